"""The drop-in boundary on the GPU, from outside Python.

* A plain C99 program (tests/abi/abi_c_driver.c), using only
  include/geon_hip.h, renders a Schwarzschild frame with geon_render; every
  pixel is held to the FP64 bar against the oracle.
* The C++ face (geon_amd/cpp/geon.hpp): trace_many, trace, render,
  stress_energy and nec_contraction against the oracle and the reference's
  published Morris–Thorne throat values (docs/energy_conditions.txt).
* geon-render, the reference's render binary with its flags
  (src/bin/render.rs): its messages and its PNG against the oracle's render
  of the same scene.
"""
import json
import math
import os
import subprocess
import zlib

import numpy as np
import pytest

from parity_util import compare_rays

pytestmark = pytest.mark.gpu

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


def _c1_oracle(O, w, h):
    om = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    cam = O.camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, w, h)
    return om, cam, O.camera_rays(om, cam)


def test_c_driver_render_matches_oracle(drivers, O, tmp_path):
    w, h = 64, 48
    out = tmp_path / "c_render.bin"
    print(f"case: abi_c_driver render {w} {h}", flush=True)
    r = subprocess.run([drivers[0], "render", str(w), str(h), str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = w * h
    raw = out.read_bytes()
    cls = np.frombuffer(raw, np.int32, n, 0)
    d = np.frombuffer(raw, np.float64, 3 * n, 4 * n).reshape(3, n).T
    rgb = np.frombuffer(raw, np.float32, 3 * n, 28 * n).reshape(n, 3)
    st = np.frombuffer(raw, np.uint64, 3, 40 * n)               # captured, escaped, diverged
    om, cam, y0 = _c1_oracle(O, w, h)
    ref = O.trace_rays(om, y0, 0.05, 4000)
    compare_rays(O, om, y0, dict(cls=cls, steps=ref["steps"], dir=d), ref, 0.05, 4000,
                 "C driver render", esc_r=100.0)
    rend = O.render(om, cam, O.render_cfg(step=0.05, max_steps=4000))
    assert (st[0], st[1], st[2]) == (rend["stats"].captured, rend["stats"].escaped,
                                     rend["stats"].diverged)
    # colours: every pixel except rays flung past the escape sphere at the
    # horizon, whose sky direction is rounding noise in both
    flung = (ref["cls"] == O.ESCAPED) & (ref["y"][:, 1] > 1000.0)
    same = (np.abs(rgb - rend["rgb"]).max(axis=1) == 0)[~flung]
    assert same.mean() > 0.999, (same.mean(), int(flung.sum()))


def test_cpp_api_on_gpu(drivers, O):
    print("case: api_driver gpu", flush=True)
    r = subprocess.run([drivers[1], "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = json.loads(r.stdout)
    om, cam, y0 = _c1_oracle(O, 32, 24)
    ref = O.trace_rays(om, y0, 0.05, 4000)
    got = dict(cls=np.array(g["cls"]), steps=np.array(g["steps"]), y=np.array(g["final"]))
    # TraceResult folds max-steps into Diverged, as the reference's does
    want_cls = np.where(ref["cls"] == O.MAXSTEPS, O.DIVERGED, ref["cls"])
    assert np.array_equal(got["cls"], want_cls)
    assert np.array_equal(got["steps"], ref["steps"].astype(np.int64))
    esc = (ref["cls"] == O.ESCAPED) & (ref["y"][:, 1] < 1000.0)
    rel = np.abs(got["y"][esc] - ref["y"][esc]) / np.maximum(1.0, np.abs(ref["y"][esc]))
    assert rel.max() < 1e-10
    assert g["one_steps"] == int(ref["steps"][5])
    rend = O.render(om, cam, O.render_cfg(step=0.05, max_steps=4000))
    s = g["stats"]
    assert (s["captured"], s["escaped"], s["diverged"]) == (
        rend["stats"].captured, rend["stats"].escaped, rend["stats"].diverged)
    assert s["avg_steps"] == rend["stats"].avg_steps
    px = np.array(g["pixels"], dtype=np.float32).reshape(-1, 3)
    flung = (ref["cls"] == O.ESCAPED) & (ref["y"][:, 1] > 1000.0)
    assert (np.abs(px - rend["rgb"]).max(axis=1) == 0)[~flung].mean() > 0.998
    # docs/energy_conditions.txt:37,46 — Morris–Thorne throat (b0 = 1)
    assert float(f"{g['mt_T_tt']:.4e}") == -3.9789e-2
    assert float(f"{g['mt_nec']:.4e}") == -7.9577e-2


def _rust_lower_exp(x, prec=3):
    m, e = f"{x:.{prec}e}".split("e")
    return f"{m}e{int(e)}"


def test_cli_render_matches_reference(drivers, O, tmp_path):
    """geon-render with the reference's flags: the Schwarzschild scene of
    render.rs:99-109 (Schwarzschild::new(1.0), observer on the equator at
    r_obs), 48×32 with the default step and max_steps; the stats line and
    the PNG against the oracle's render of the same scene."""
    png = tmp_path / "out" / "cli.png"
    args = [CLI, "--metric", "schwarzschild", "--width", "48", "--height", "32", "--r-obs", "20",
            "--out", str(png)]
    print("case: " + " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[0] == "rendering Schwarzschild @ 48x32, fov 60°, r_obs 20, step 0.05"
    assert lines[2] == f"wrote {png}"
    om = O.metric(O.SCHWARZSCHILD)
    cam = O.camera((0.0, 20.0, math.pi / 2, 0.0), 60.0, 48, 32)
    rend = O.render(om, cam, O.render_cfg(step=0.05, max_steps=20000))
    st = rend["stats"]
    head = (f"done. captured={st.captured}, escaped={st.escaped}, diverged={st.diverged}, "
            f"avg_steps={st.avg_steps:.1f}, max_null_residual=")
    print(lines[1], "| oracle:", head + _rust_lower_exp(st.max_null_residual), flush=True)
    assert lines[1].startswith(head)
    mres = float(lines[1].rsplit("=", 1)[1])
    assert abs(mres - st.max_null_residual) <= 5e-3 * st.max_null_residual
    # the PNG: 8-bit RGB rows top to bottom, write_png's tonemap of the render
    data = png.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, wh = 8, b"", None
    while pos < len(data):
        ln = int.from_bytes(data[pos:pos + 4], "big")
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + ln]
        assert zlib.crc32(tag + body) == int.from_bytes(data[pos + 8 + ln:pos + 12 + ln], "big")
        if tag == b"IHDR":
            wh = (int.from_bytes(body[:4], "big"), int.from_bytes(body[4:8], "big"), body[8], body[9])
        elif tag == b"IDAT":
            idat += body
        pos += 12 + ln
    assert wh == (48, 32, 8, 2)
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(32, 1 + 48 * 3)
    assert (rows[:, 0] == 0).all()
    img = rows[:, 1:].reshape(-1, 3)
    want = O.tonemap_u8(rend["rgb"].reshape(-1)).reshape(-1, 3)
    # every pixel but the rays flung out at the horizon (escape radius 1000 M)
    tr = O.trace_rays(om, O.camera_rays(om, cam), 0.05, 20000)
    flung = (tr["cls"] == O.ESCAPED) & (tr["y"][:, 1] > 10000.0)
    same = (img == want).all(axis=1)[~flung]
    print(f"PNG: {int(same.sum())} of {same.size} pixels equal, {int(flung.sum())} flung", flush=True)
    assert same.mean() > 0.998
