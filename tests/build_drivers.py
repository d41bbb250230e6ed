"""Build the test drivers of the C ABI and the C++ face (test infrastructure).

    python tests/build_drivers.py

  tests/abi/abi_c_driver   plain C99 client of include/geon_hip.h (gcc)
  tests/cpp/api_driver     client of geon_amd/cpp/geon.hpp and render_cli.hpp (g++)

Both link geon_amd/lib/libgeon_hip.so, which must be built first
(python -m geon_amd.build).
"""
from __future__ import annotations

import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "geon_amd", "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
C_DRIVER = os.path.join(ROOT, "tests", "abi", "abi_c_driver")
CPP_DRIVER = os.path.join(ROOT, "tests", "cpp", "api_driver")


def _link():
    return [f"-L{LIB}", "-lgeon_hip", f"-Wl,-rpath,{LIB}", f"-Wl,-rpath,{ROCM}/lib"]


def build_c_driver(out=C_DRIVER):
    cc = shutil.which("gcc") or "gcc"
    tmp = f"{out}.{os.getpid()}.tmp"
    subprocess.run([cc, "-std=c99", "-O1", "-Wall", "-Wextra", "-pedantic", out + ".c", *_link(),
                    "-lm", "-o", tmp], check=True)
    os.replace(tmp, out)
    return out


def build_cpp_driver(out=CPP_DRIVER):
    cxx = shutil.which("g++") or "g++"
    tmp = f"{out}.{os.getpid()}.tmp"
    subprocess.run([cxx, "-std=c++17", "-O1", "-Wall", "-Wno-unknown-pragmas", out + ".cpp",
                    *_link(), "-o", tmp], check=True)
    os.replace(tmp, out)
    return out


def build_all():
    return build_c_driver(), build_cpp_driver()


if __name__ == "__main__":
    build_all()
