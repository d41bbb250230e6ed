"""Build the gfx950 code object and the host runtime in-tree.

    python -m geon_amd.build

Outputs (git-ignored, shipped beside the sources):
  geon_amd/lib/geon_kernels.hsaco   hipcc --genco --offload-arch=gfx950
  geon_amd/lib/libgeon_hip.so       host runtime + C ABI (g++, links libamdhip64), plus
                                    geon_noalloc.hip (hipcc -c: a device array of its own)
  geon_amd/lib/opcount              FP64 operation counter of the device code
  geon_amd/lib/geon-render          the render command line (bin/render.rs flags)

`python -m geon_amd.build --opcount` also rewrites geon_amd/opcount/op_counts.json,
the per-ray-step and per-point operation counts used for the roofline.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# FP64 kernels keep the reference's rounding: no FMA contraction, no fast math.
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["--genco", "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-ffp-contract=off",
              "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-Wall", "-Wno-unused-result",
              "-Wno-unknown-pragmas"]

KERNEL_SRC = [os.path.join(CSRC, "kernels", "geon_kernels.hip")]
KERNEL_DEPS = KERNEL_SRC + [os.path.join(CSRC, "kernels", f)
                            for f in ("geon_device.hpp", "geon_fp32.hpp", "geon_sched.hpp",
                                      "geon_curv.hpp")] + \
    [os.path.join(CSRC, "geon_dev_types.h")]
HOST_SRC = [os.path.join(CSRC, "host", f)
            for f in ("geon_runtime.cpp", "geon_abi.cpp", "geon_batch.cpp", "geon_serve.cpp",
                      "geon_metric_host.cpp")]
# The allocation-free render keeps its scratch in a file-scope __device__
# array, so its translation unit is compiled by hipcc (host code plus the
# gfx950 image that holds the array) and linked into the same library.
NOALLOC_SRC = os.path.join(CSRC, "host", "geon_noalloc.hip")
NOALLOC_FLAGS = ["-c", "-fPIC", "--offload-arch=gfx950", "-O2", "-std=c++17", "-fvisibility=hidden",
                 "-ffp-contract=off", "-Wall", "-Xclang", "-target-feature", "-Xclang",
                 "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_kernels(out=None):
    out = out or os.path.join(LIB, "geon_kernels.hsaco")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, *KERNEL_SRC, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host(out=None):
    out = out or os.path.join(LIB, "libgeon_hip.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cxx = shutil.which("g++") or "g++"
    with tempfile.TemporaryDirectory() as td:
        obj = os.path.join(td, "geon_noalloc.o")
        _run([HIPCC, *NOALLOC_FLAGS, NOALLOC_SRC, "-o", obj])
        _run([cxx, *HOST_FLAGS, *HOST_SRC, obj, f"-L{ROCM}/lib", "-lamdhip64", "-lpthread", "-lm",
              f"-Wl,-rpath,{ROCM}/lib",
              "-o", tmp])
    os.replace(tmp, out)
    return out


CLI_SRC = os.path.join(HERE, "cpp", "render_main.cpp")


def build_cli(out=None):
    """geon-render: the reference's render binary over the C ABI (links libgeon_hip.so)."""
    out = out or os.path.join(LIB, "geon-render")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", CLI_SRC, f"-L{LIB}",
          "-lgeon_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib", "-o", tmp])
    os.replace(tmp, out)
    return out


OPCOUNT_SRC = os.path.join(HERE, "opcount", "opcount.cpp")
OPCOUNT_JSON = os.path.join(HERE, "opcount", "op_counts.json")


def build_opcount(out=None):
    """Host build of the device header with a counting scalar in place of double."""
    out = out or os.path.join(LIB, "opcount")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    tmp = f"{out}.{os.getpid()}.tmp"   # concurrent test workers each build their own
    _run([cxx, "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
          OPCOUNT_SRC, "-o", tmp])
    os.replace(tmp, out)
    return out


def run_opcount(write=False):
    text = subprocess.run([build_opcount()], check=True, capture_output=True, text=True).stdout
    if write:
        with open(OPCOUNT_JSON, "w") as f:
            f.write(text)
    return text


def build_all():
    build_kernels()
    build_host()
    build_cli()


if __name__ == "__main__":
    build_all()
    if "--opcount" in sys.argv:
        run_opcount(write=True)
    sys.exit(0)
