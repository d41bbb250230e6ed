"""The host runtime and the oracle built with AddressSanitizer and
UndefinedBehaviorSanitizer (host code only; no GPU code is instrumented) and
exercised on the CPU: the C ABI entry points that need no GPU, error paths
included, and oracle traces that cross the horizon, the pole and NaN."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=undefined"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")


def _run(cmd, **kw):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=600, **kw)


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    return tmp_path_factory.mktemp("san")


def test_host_runtime_under_sanitizers(tmp):
    host = [os.path.join(ROOT, "geon_amd", "csrc", "host", f)
            for f in ("geon_runtime.cpp", "geon_abi.cpp", "geon_batch.cpp", "geon_serve.cpp",
                      "geon_metric_host.cpp")]
    exe = str(tmp / "abi_driver")
    b = _run(["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
              f"-I{os.path.join(ROOT, 'include')}",
              os.path.join(ROOT, "tests", "sanitize", "abi_host_driver.cpp"), *host,
              f"-L{ROCM}/lib", "-lamdhip64", "-lpthread", f"-Wl,-rpath,{ROCM}/lib", "-o", exe])
    assert b.returncode == 0, b.stderr[-3000:]
    r = _run([exe], env=ENV)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]
    assert r.returncode == 0 and "failures 0" in out, out[-3000:]


def test_oracle_under_sanitizers(tmp):
    lib = str(tmp / "liboracle_san.so")
    b = _run(["gcc", "-std=c11", *SAN, "-ffp-contract=off", "-shared", "-fPIC",
              os.path.join(ROOT, "oracle", "geon_oracle.c"), "-lm", "-lpthread", "-o", lib])
    assert b.returncode == 0, b.stderr[-3000:]
    script = f"""
import ctypes, math, sys
import numpy as np
sys.path.insert(0, {ROOT!r})
from oracle import oracle as O
O.LIB_PATH = {lib!r}   # load the instrumented build through the usual declarations
m = O.metric(O.SCHWARZSCHILD, horizon_pad=1.005, escape_radius=100.0)
y0 = O.camera_rays(m, O.camera([0, 30, math.pi / 2, 0], 60.0, 24, 24))
y0[5, 6] = float("nan")
r = O.trace_rays(m, y0, 0.05, 4000, nthreads=4)
ra = O.render(m, O.camera([0, 12, 1.4, 0.2], 70.0, 9, 7), O.render_cfg(max_steps=2000, samples_per_axis=2), nthreads=2)
c = O.curvature(O.metric(O.ALCUBIERRE), np.c_[np.zeros(6), np.linspace(-3, 3, 18).reshape(6, 3)], O.CURV_RIEMANN)
print("oracle ok", int(r["cls"].sum()), int(ra["cls"].sum()))
"""
    # preload the sanitizer runtime so that python can load the instrumented library
    asan = _run(["gcc", "-print-file-name=libasan.so"]).stdout.strip()
    ubsan = _run(["gcc", "-print-file-name=libubsan.so"]).stdout.strip()
    env = dict(ENV, LD_PRELOAD=f"{asan}:{ubsan}")
    r = _run(["python3", "-c", script], env=env)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-3000:]
    assert r.returncode == 0 and "oracle ok" in out, out[-3000:]
