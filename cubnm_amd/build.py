"""Build the gfx950 simulation core in-tree.

Produces, next to this file:
  libcubnm_amd.so                          C ABI (include/cubnm_amd.h), HIP kernels
  core.cpython-<ver>-<platform>.so         CPython module `cubnm_amd.core`
Object files go to cubnm_amd/_build/.  Run ``python -m cubnm_amd.build``.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["session.cpp", "integrate_general.hip", "integrate_batched.hip", "fc_fcd.hip", "fc_gram.hip", "gof.hip", "adjoint.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
]

LIB = os.path.join(HERE, "libcubnm_amd.so")


def core_module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "core" + suffix)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB] + objs)
    import numpy
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          "-I", py_inc, "-I", numpy.get_include(), "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "pycore.cpp"), "-o", core_module_path(),
          "-L", HERE, "-lcubnm_amd", "-Wl,-rpath,$ORIGIN",
          "-L", os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    if verbose:
        print("built", LIB, core_module_path())
    return LIB


if __name__ == "__main__":
    build(verbose=True)
