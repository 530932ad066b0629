"""Build librbmdx.so (HIP kernels for gfx950 + host engine) in-tree.

The kernels are compiled by hipcc for gfx950 only; the host engine and the
RBE sampler are compiled by g++ as plain C++17 (host code has no HIP kernels).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librbmdx.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", INC, "-I", CSRC,
]
CXX_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
    "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
    "-I", os.path.join(ROCM, "include"), "-I", INC, "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    jobs = [
        [hipcc] + HIPCC_FLAGS + ["-c", "-o", os.path.join(BUILD, "kernels.o"),
                                 os.path.join(CSRC, "kernels.hip")],
        [hipcc] + HIPCC_FLAGS + ["-c", "-o", os.path.join(BUILD, "batch_kernels.o"),
                                 os.path.join(CSRC, "batch.hip")],
        [hipcc] + HIPCC_FLAGS + ["-c", "-o", os.path.join(BUILD, "md_kernels.o"),
                                 os.path.join(CSRC, "md.hip")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "batch.o"),
                               os.path.join(CSRC, "batch.cpp")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "lammps_data.o"),
                               os.path.join(CSRC, "lammps_data.cpp")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "data_api.o"),
                               os.path.join(CSRC, "data_api.cpp")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "output.o"),
                               os.path.join(CSRC, "output.cpp")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "api.o"),
                               os.path.join(CSRC, "api.cpp")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "sampler.o"),
                               os.path.join(CSRC, "sampler.cpp")],
        ["g++"] + CXX_FLAGS + ["-c", "-o", os.path.join(BUILD, "sog.o"),
                               os.path.join(CSRC, "sog.cpp")],
    ]
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(_run, jobs))
    objs = [os.path.join(BUILD, f) for f in ("kernels.o", "batch_kernels.o",
                                             "md_kernels.o", "api.o",
                                             "batch.o", "lammps_data.o", "data_api.o", "output.o",
                                             "sampler.o", "sog.o")]
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
         + ["-lpthread"])
    os.replace(tmp, LIB)
    # C++ driver of the force-module layer, linked against the library
    drv = os.path.join(HERE, "rbmdx_plugin_driver")
    _run(["g++"] + CXX_FLAGS + ["-o", drv + ".tmp", os.path.join(CSRC, "plugin_driver.cpp"),
                                "-L", HERE, "-lrbmdx", "-Wl,-rpath,$ORIGIN",
                                "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
                                "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(drv + ".tmp", drv)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
