"""Accuracy of the kernels' sincos_cw (Cody–Waite reduction + Taylor
polynomials) against 40-digit arithmetic: within 1.5 ulp of the correctly
rounded sine and cosine over the angles a ray's θ takes, including the
neighbourhoods of the multiples of π/2 where sin θ or cos θ is tiny (the
poles), and falling back to the library beyond |x| = 2^20."""
import ctypes as C
import math
import os
import shutil
import subprocess

import mpmath
import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.fixture(scope="module")
def fn(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    so = str(tmp_path_factory.mktemp("scw") / "libscw.so")
    b = subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                        "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
                        os.path.join(ROOT, "tests", "host", "sincos_cw_driver.cpp"), "-o", so],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    L = C.CDLL(so)
    L.geon_test_sincos_cw.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]

    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        s, c = np.zeros_like(x), np.zeros_like(x)
        L.geon_test_sincos_cw(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
        return s, c
    return run


def _ulp_err(got, x, f):
    mpmath.mp.dps = 40
    out = []
    for g, v in zip(got, x):
        exact = f(mpmath.mpf(float(v)))
        ulp = math.ulp(float(exact)) if float(exact) != 0.0 else 5e-324
        out.append(abs(float((mpmath.mpf(float(g)) - exact) / ulp)))
    return np.array(out)


def test_sincos_cw_within_one_and_a_half_ulp(fn):
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-10.0, 10.0, 4000),
        rng.uniform(0.0, math.pi, 4000),
        # next to the multiples of π/2 (poles and equator), down to 1e-12 away
        np.concatenate([k * math.pi / 2 + np.sign(rng.uniform(-1, 1, 300)) *
                        10.0 ** rng.uniform(-12, -2, 300) for k in range(-2, 5)]),
        np.array([0.0, -0.0, 1e-300, math.pi / 4, 3 * math.pi / 4, 1e5, -7e5]),
    ])
    s, c = fn(x)
    es, ec = _ulp_err(s, x, mpmath.sin), _ulp_err(c, x, mpmath.cos)
    print(f"sincos_cw: max error sin {es.max():.3f} ulp, cos {ec.max():.3f} ulp over {x.size} angles")
    # the rounding of the reduced argument (up to ~0.55 ulp near |r| = π/4)
    # plus the polynomial's final sums; the library's own is within ~0.5
    assert es.max() <= 1.5 and ec.max() <= 1.5


def test_sincos_cw_falls_back_beyond_range(fn):
    x = np.array([2.0 ** 21, -3.0e7, 1e300, math.inf, math.nan])
    s, c = fn(x)
    assert np.array_equal(s[:3], np.sin(x[:3])) and np.array_equal(c[:3], np.cos(x[:3]))
    assert np.isnan(s[3:]).all() and np.isnan(c[3:]).all()
