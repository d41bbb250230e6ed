"""Accuracy of the kernels' own transcendental routines against 40-digit
arithmetic. sincos_acc_dw (nearly correctly rounded, with two-word values)
and sincos_rot_dw (stage values by rotation from them): the Schwarzschild
steps'. sincos_cw (Cody–Waite reduction + Taylor polynomials): within
1.5 ulp of the correctly rounded sine and cosine over the angles a ray's θ
takes, including the neighbourhoods of the multiples of π/2 where sin θ or
cos θ is tiny (the poles), and falling back to the library beyond
|x| = 2^20. tanh_cw (exp by Cody–Waite + Taylor, two-word denominator,
remainder-corrected quotient): within an ulp over the arguments the
Alcubierre shape function takes and far beyond, tiny, saturating and
special values included."""
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
def dmlib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    so = str(tmp_path_factory.mktemp("scw") / "libscw.so")
    b = subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                        "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
                        os.path.join(ROOT, "tests", "host", "device_math_driver.cpp"), "-o", so],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    L = C.CDLL(so)
    L.geon_test_sincos_cw.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
    L.geon_test_tanh_cw.argtypes = [C.c_void_p, C.c_long, C.c_void_p]
    L.geon_test_sincosf_cw.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
    L.geon_test_sincos_acc.argtypes = [C.c_void_p, C.c_long] + [C.c_void_p] * 4
    L.geon_test_sincos_rot_dw.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p,
                                          C.c_void_p]
    return L


@pytest.fixture(scope="module")
def fn(dmlib):
    def run(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        s, c = np.zeros_like(x), np.zeros_like(x)
        dmlib.geon_test_sincos_cw(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
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


def _sincos_acc(dmlib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = [np.zeros_like(x) for _ in range(4)]
    dmlib.geon_test_sincos_acc(x.ctypes.data, x.size, *(o.ctypes.data for o in out))
    return out


def test_sincos_acc_nearly_correctly_rounded(dmlib):
    """sincos_acc_dw (the reduced argument as an exact two-word sum, the
    rounding errors of r², of 1 − r²/2 and of the final sums carried): within
    0.65 ulp of the correctly rounded sine and cosine on sincos_cw's angle set
    (measured 0.62), the two-word values within 0.25 ulp (measured 0.23), and
    the same last bit as glibc on ≥ 98 % of θ in [0, π] (measured 98.6 %;
    sincos_cw ≈ 78 %)."""
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.uniform(-10.0, 10.0, 3000),
        rng.uniform(0.0, math.pi, 3000),
        np.concatenate([k * math.pi / 2 + np.sign(rng.uniform(-1, 1, 200)) *
                        10.0 ** rng.uniform(-12, -2, 200) for k in range(-2, 5)]),
        np.array([0.0, 1e-300, math.pi / 4, 3 * math.pi / 4, 1e5, -7e5]),
    ])
    s, sl, c, cl = _sincos_acc(dmlib, x)
    es, ec = _ulp_err(s, x, mpmath.sin), _ulp_err(c, x, mpmath.cos)
    mpmath.mp.dps = 40
    def dw_err(hi, lo, f):
        out = []
        for h, l, v in zip(hi, lo, x):
            exact = f(mpmath.mpf(float(v)))
            ulp = math.ulp(float(h)) if h != 0.0 else 5e-324
            out.append(abs(float((mpmath.mpf(float(h)) + mpmath.mpf(float(l)) - exact) / ulp)))
        return np.array(out)
    ds, dc = dw_err(s, sl, mpmath.sin), dw_err(c, cl, mpmath.cos)
    print(f"sincos_acc: max error sin {es.max():.3f} cos {ec.max():.3f} ulp; two-word "
          f"{ds.max():.3f} / {dc.max():.3f}")
    assert es.max() <= 0.65 and ec.max() <= 0.65
    assert ds.max() <= 0.25 and dc.max() <= 0.25
    u = rng.uniform(0.0, math.pi, 20000)
    su, _, cu, _ = _sincos_acc(dmlib, u)
    same = (np.mean(su == np.array([math.sin(v) for v in u])),
            np.mean(cu == np.array([math.cos(v) for v in u])))
    print(f"sincos_acc: same bits as glibc on {same[0]:.4f} (sin), {same[1]:.4f} (cos) of [0, π]")
    assert min(same) >= 0.98


def test_sincos_rot_dw_stage_values(dmlib):
    """sin and cos at an RK4 stage by rotation from the step's two-word
    values (sincos_rot_dw): the error stays within 1.5 ulp of the largest of
    |value|, |value at the step| and |d| (measured 1.03; the rotation's own
    products round at that scale, so where θ passes a zero of the function the
    error is that absolute amount), and for stage-sized increments |d| ≤ 0.02
    the values have glibc's last bit on ≥ 96 % of θ0 in [1/8, π − 1/8]
    (measured 98.2 % for sin, 97.0 % for cos)."""
    rng = np.random.default_rng(13)
    th0 = np.concatenate([rng.uniform(0.125, math.pi - 0.125, 4000), rng.uniform(-0.125, 0.125, 1000)])
    d = np.concatenate([rng.uniform(-0.0625, 0.0625, 2500),
                        np.sign(rng.uniform(-1, 1, 2500)) * 10.0 ** rng.uniform(-9, -1.3, 2500)])
    th = th0 + d
    s, c = np.zeros_like(th), np.zeros_like(th)
    dmlib.geon_test_sincos_rot_dw(th0.ctypes.data, th.ctypes.data, th.size, s.ctypes.data,
                                  c.ctypes.data)
    mpmath.mp.dps = 40
    worst = 0.0
    for g, h, v, v0 in zip(s, c, th, th0):
        X = mpmath.mpf(float(v))
        for got, ex, start in ((g, mpmath.sin(X), math.sin(v0)), (h, mpmath.cos(X), math.cos(v0))):
            scale = math.ulp(max(abs(float(ex)), abs(start), abs(v - v0)))
            worst = max(worst, abs(float((mpmath.mpf(float(got)) - ex) / scale)))
    print(f"sincos_rot_dw: max error {worst:.3f} ulp of the largest term")
    assert worst <= 1.5
    th0 = rng.uniform(0.125, math.pi - 0.125, 20000)
    th = th0 + rng.uniform(-0.02, 0.02, th0.size)
    s, c = np.zeros_like(th), np.zeros_like(th)
    dmlib.geon_test_sincos_rot_dw(th0.ctypes.data, th.ctypes.data, th.size, s.ctypes.data,
                                  c.ctypes.data)
    same = (np.mean(s == np.array([math.sin(v) for v in th])),
            np.mean(c == np.array([math.cos(v) for v in th])))
    print(f"sincos_rot_dw: same bits as glibc on {same[0]:.4f} (sin), {same[1]:.4f} (cos)")
    assert min(same) >= 0.96


def _tanh_cw(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib.geon_test_tanh_cw(x.ctypes.data, x.size, out.ctypes.data)
    return out


def test_tanh_cw_within_an_ulp(dmlib):
    """Over [0, 25] (the shape function's arguments (r_s ± R)/s and beyond
    the saturation at 19.06), log-uniform down to 1e-20, next to every
    reduction boundary k·ln2/2, and negative arguments: within 0.75 ulp of
    the correctly rounded tanh (measured 0.68; glibc's is within ~1.1 on the
    same sample)."""
    rng = np.random.default_rng(5)
    ln2 = math.log(2.0)
    x = np.concatenate([
        rng.uniform(0.0, 25.0, 6000),
        10.0 ** rng.uniform(-20, 1.4, 6000),
        np.concatenate([k * ln2 / 2 * (1 + rng.uniform(-1e-9, 1e-9, 40)) for k in range(1, 64)]),
        -rng.uniform(0.0, 25.0, 1000),
    ])
    got = _tanh_cw(dmlib, x)
    err = _ulp_err(got, x, mpmath.tanh)
    print(f"tanh_cw: max error {err.max():.3f} ulp, mean {err.mean():.3f} over {x.size} arguments")
    assert err.max() <= 0.75


def test_tanh_cw_special_values(dmlib):
    x = np.array([0.0, -0.0, math.inf, -math.inf, 22.0, 19.5, 700.0, -700.0, 1e-300, -5e-324,
                  2.0 ** -28, math.nan])
    got = _tanh_cw(dmlib, x)
    want = np.tanh(x)
    assert np.array_equal(got[:-1], want[:-1])
    assert np.array_equal(np.signbit(got[:-1]), np.signbit(want[:-1]))
    assert np.isnan(got[-1])


def test_sincosf_cw_fp32(dmlib):
    """The FP32 path's sincos (π/2 in three floats, Taylor to r^9 / r^10):
    within 2 ulp of the correctly rounded FP32 sine and cosine over the
    angles a ray's θ takes and the neighbourhoods of the multiples of π/2,
    falling back to the library beyond 2^13."""
    rng = np.random.default_rng(12)
    x = np.concatenate([
        rng.uniform(-10.0, 10.0, 3000), rng.uniform(0.0, math.pi, 3000),
        np.concatenate([k * math.pi / 2 + np.sign(rng.uniform(-1, 1, 200)) *
                        10.0 ** rng.uniform(-6, -1, 200) for k in range(-2, 5)]),
        np.array([0.0, 1e-30, 5000.0, -7000.0]),
    ]).astype(np.float32)
    s, c = np.zeros_like(x), np.zeros_like(x)
    dmlib.geon_test_sincosf_cw(x.ctypes.data, x.size, s.ctypes.data, c.ctypes.data)
    mpmath.mp.dps = 30
    def err(got, f):
        out = []
        for g, v in zip(got, x):
            exact = f(mpmath.mpf(float(v)))
            e32 = np.float32(float(exact))
            ulp = float(np.spacing(np.abs(e32))) if e32 != 0 else 1e-45
            out.append(abs(float(mpmath.mpf(float(g)) - exact)) / ulp)
        return np.array(out)
    es, ec = err(s, mpmath.sin), err(c, mpmath.cos)
    print(f"sincosf_cw: max error sin {es.max():.3f} ulp, cos {ec.max():.3f} ulp")
    assert es.max() <= 2.0 and ec.max() <= 2.0
    big = np.array([9000.0, -1e6], dtype=np.float32)
    s2, c2 = np.zeros_like(big), np.zeros_like(big)
    dmlib.geon_test_sincosf_cw(big.ctypes.data, big.size, s2.ctypes.data, c2.ctypes.data)
    assert np.allclose(s2, np.sin(big.astype(np.float64)), atol=1e-6)
