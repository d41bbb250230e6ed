"""The drop-in boundary from C and C++ (no GPU needed).

* A plain C99 program that includes only include/geon_hip.h calls every
  host-side entry point and the error paths (tests/abi/abi_c_driver.c).
* The C++ face (geon_amd/cpp/geon.hpp) evaluates the Metric trait, dot4,
  inv4, rhs and rk4_step on the host, bit for bit with the oracle's
  restatement of src/{types,metric,geodesic,integrator,camera}.rs (g⁻¹ and
  inv4 by value: structural zeros may carry the other sign of zero, and the
  cofactor inverse is checked against the oracle's to 1e-15).
* geon-render's flags map to the scene the reference's render binary
  builds (src/bin/render.rs:14-134), and its messages use Rust's float
  formatting (Display, {:.1}, {:.3e}).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "geon_amd", "lib", "geon-render")


@pytest.fixture(scope="module")
def drivers():
    import sys
    sys.path.insert(0, ROOT)
    from tests import build_drivers
    c, cpp = build_drivers.C_DRIVER, build_drivers.CPP_DRIVER
    if not (os.path.exists(c) and os.path.exists(cpp)):
        build_drivers.build_all()
    return c, cpp


def _run(args, ok=True):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, (args, r.stdout, r.stderr)
    return r


def test_c_driver_exports(drivers):
    r = _run([drivers[0], "exports"])
    assert r.stdout.startswith("OK ") and int(r.stdout.split()[1]) >= 60, r.stdout


def _args(cpp, *flags, ok=True):
    r = _run([cpp, "args", *flags], ok=ok)
    return r.returncode, json.loads(r.stdout)


def test_cli_defaults_are_the_reference_defaults(drivers):
    """render.rs:22-67 defaults: Schwarzschild::new(1.0), 1024×1024, fov 60°,
    r_obs 20 on the equator, step 0.05, max 20,000, 1 sample, stars on,
    outputs/render.png; RenderConfig::default colours (render.rs:32-41)."""
    rc, a = _args(drivers[1])
    assert a["kind"] == 1 and a["name"] == "schwarzschild" and a["p"][0] == 1.0
    assert (a["horizon_pad"], a["escape_radius"], a["fd_step"]) == (1.001, 1000.0, 1e-5)
    assert a["observer"] == [0.0, 20.0, math.pi / 2, 0.0]
    assert a["fov_y_rad"] == 60.0 * (math.pi / 180.0)
    assert (a["width"], a["height"], a["step"], a["max_steps"], a["samples_per_axis"]) == \
        (1024, 1024, 0.05, 20000, 1)
    assert a["captured_rgb"] == [0, 0, 0] and a["diverged_rgb"] == [1, 0, 1]
    assert a["show_stars"] is True and a["out"] == "outputs/render.png"
    assert a["banner"] == "rendering Schwarzschild @ 1024x1024, fov 60°, r_obs 20, step 0.05"


def test_cli_flags_build_the_reference_scenes(drivers):
    rc, a = _args(drivers[1], "--metric", "alcubierre", "--bubble-v", "0.9", "--r-obs=12",
                  "--fov", "45", "--width", "640", "--height=480", "--step", "0.02",
                  "--max-steps", "777", "--samples", "3", "--no-stars", "--out", "x/y.png")
    # Alcubierre { ship_velocity: bubble_v, ..Alcubierre::default() }, Cartesian observer
    assert a["kind"] == 2 and a["p"][:4] == [2.0, 0.5, 0.9, 0.0]
    assert a["observer"] == [0.0, 12.0, 0.0, 0.0] and a["fov_y_rad"] == 45.0 * (math.pi / 180.0)
    assert (a["width"], a["height"], a["step"], a["max_steps"], a["samples_per_axis"]) == \
        (640, 480, 0.02, 777, 3)
    assert a["show_stars"] is False and a["out"] == "x/y.png"
    assert a["banner"] == "rendering Alcubierre @ 640x480, fov 45°, r_obs 12, step 0.02"
    rc, a = _args(drivers[1], "--metric", "morris-thorne", "--r-obs", "15")
    assert a["kind"] == 3 and a["p"][0] == 1.0 and a["observer"] == [0.0, 15.0, math.pi / 2, 0.0]
    assert a["banner"].startswith("rendering MorrisThorne @ ")
    rc, a = _args(drivers[1], "--metric", "minkowski")
    assert a["kind"] == 0 and a["observer"] == [0.0, 20.0, 0.0, 0.0]
    for bad in (["--metric", "kerr"], ["--width", "-3"], ["--fov", "wide"], ["--colour"],
                ["--step"], ["--no-stars=yes"]):
        rc, a = _args(drivers[1], *bad, ok=False)
        assert rc == 2 and a == {"error": "usage"}, bad


def test_cli_binary_usage(drivers):
    r = _run([CLI, "--help"])
    assert "--max-steps" in r.stdout and "--bubble-v" in r.stdout
    r = _run([CLI, "--metric", "kerr"], ok=False)
    assert r.returncode == 2 and r.stderr.startswith("error: ")


def test_rust_float_formatting(drivers):
    """f64 Display is the shortest round-trip decimal in positional notation;
    {:.3e} has no exponent padding or '+'; {:.1} rounds half to even."""
    vals = ["60", "20", "0.05", "1e-5", "123.456", "0.25", "6.84e-13", "0", "1e21", "2.5e-7",
            "13.25", "0.1"]
    r = _run([drivers[1], "format", *vals])
    got = json.loads(r.stdout)
    assert [g[0] for g in got] == ["60", "20", "0.05", "0.00001", "123.456", "0.25",
                                   "0.000000000000684", "0", "1000000000000000000000",
                                   "0.00000025", "13.25", "0.1"]
    assert [g[1] for g in got] == ["6.000e1", "2.000e1", "5.000e-2", "1.000e-5", "1.235e2",
                                   "2.500e-1", "6.840e-13", "0.000e0", "1.000e21", "2.500e-7",
                                   "1.325e1", "1.000e-1"]
    assert got[5][2] == "0.2" and got[10][2] == "13.2" and got[0][2] == "60.0"


def test_cpp_host_api_matches_oracle(drivers):
    r = _run([drivers[1], "host"])
    h = json.loads(r.stdout)
    cases = {"schwarzschild": (O.metric(O.SCHWARZSCHILD), [0.0, 7.0, 1.1, 0.3]),
             "morris_thorne": (O.metric(O.MORRIS_THORNE), [0.0, 1.7, 0.9, 0.2]),
             "alcubierre": (O.metric(O.ALCUBIERRE, bubble_thickness=0.125, ship_velocity=0.5),
                            [0.0, 1.9, 0.4, -0.3])}
    bits = lambda a: np.asarray(a, dtype=np.float64).view(np.uint64)  # noqa: E731
    for name, (om, x) in cases.items():
        c = h[name]
        assert np.array_equal(bits(c["g"]), bits(O.g(om, x).reshape(-1))), name
        assert np.array_equal(bits(c["gamma"]), bits(O.christoffel(om, x).reshape(-1))), name
        st, gi = O.g_inv(om, x)
        assert st == 0 and np.array_equal(np.array(c["g_inv"]), gi.reshape(-1)), name
        st, inv = O.inv4(O.g(om, x))
        assert np.abs(np.array(c["inv4"]) - inv.reshape(-1)).max() <= 1e-15 * np.abs(inv).max()
        p = np.array([1.3, -0.4, 0.02, 0.05])
        assert c["dot4"] == O.dot4(O.g(om, x), p, p)
        y = np.array([0.0, *x[1:], *p])
        assert np.array_equal(bits(c["rhs"]), bits(O.rhs(om, y))), name
        assert np.array_equal(bits(c["rk4"]), bits(O.rk4_step(om, y, 0.05))), name
        assert c["terminate"] == -1 and c["fd_step"] == 1e-5
    assert h["singular_throws"] is True
    om = O.metric(O.SCHWARZSCHILD)
    cam = O.camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 16, 8)
    st, tet = O.tetrad(om, cam.observer)
    assert np.array_equal(np.array(h["ray_state"]), O.ray_state_sub(cam, tet, 3, 2))
