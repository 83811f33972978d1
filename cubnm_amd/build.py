"""Build the gfx950 simulation core in-tree.

Produces, next to this file:
  libcubnm_amd.so                          C ABI (include/cubnm_amd.h), HIP kernels
  core.cpython-<ver>-<platform>.so         CPython module `cubnm_amd.core`
  .build_stamp                             hash of the sources both were built from
Object files go to cubnm_amd/_build/.

``python -m cubnm_amd.build`` (or ``build()``) always compiles.
``ensure_built()`` compiles only when a library is missing or the sources hash
differs from the stamp; the package calls it on import, so a clean checkout
builds itself on first use (hipcc cross-compiles gfx950 without a GPU).  A
file lock serialises concurrent first imports.

This module imports nothing from the package, so it can be run or loaded by
path before anything is built.
"""
import fcntl
import hashlib
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["session.cpp", "integrate_general.hip", "integrate_batched.hip", "integrate_batched_wide.hip",
               "fc_fcd.hip", "fc_gram.hip", "gof.hip", "adjoint.hip", "loss.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I", CSRC, "-I", INCLUDE,
]

LIB = os.path.join(HERE, "libcubnm_amd.so")
STAMP = os.path.join(HERE, ".build_stamp")
LOCK = os.path.join(HERE, ".build_lock")


def core_module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "core" + suffix)


def source_hash():
    """sha256 over every source and header the two libraries are built from,
    plus the compiler flags and the Python ABI tag."""
    h = hashlib.sha256()
    files = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC))
    files += sorted(os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h"))
    for f in files:
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    h.update(" ".join(HIP_SOURCES + HIP_FLAGS[:-4]).encode())
    h.update((sysconfig.get_config_var("EXT_SUFFIX") or "").encode())
    return h.hexdigest()


def is_current():
    if not (os.path.exists(LIB) and os.path.exists(core_module_path()) and os.path.exists(STAMP)):
        return False
    with open(STAMP) as fh:
        return fh.read().strip() == source_hash()


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


def _build_unlocked(verbose):
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}; cubnm_amd needs ROCm to build its gfx950 core")
    digest = source_hash()
    os.makedirs(BUILD, exist_ok=True)
    if os.path.exists(STAMP):
        os.remove(STAMP)
    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    # link to temporary names, then rename: a process that already mapped the
    # old library keeps its inode
    tmp_lib = LIB + ".tmp%d" % os.getpid()
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp_lib] + objs)
    os.replace(tmp_lib, LIB)
    import numpy
    py_inc = sysconfig.get_paths()["include"]
    tmp_core = core_module_path() + ".tmp%d" % os.getpid()
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
          "-I", py_inc, "-I", numpy.get_include(), "-I", INCLUDE,
          os.path.join(CSRC, "pycore.cpp"), "-o", tmp_core,
          "-L", HERE, "-lcubnm_amd", "-Wl,-rpath,$ORIGIN",
          "-L", os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp_core, core_module_path())
    with open(STAMP, "w") as fh:
        fh.write(digest + "\n")
    if verbose:
        print("built", LIB, core_module_path(), flush=True)
    return LIB


class _Lock:
    def __enter__(self):
        self.fh = open(LOCK, "a")
        fcntl.flock(self.fh, fcntl.LOCK_EX)
        return self

    def __exit__(self, *exc):
        fcntl.flock(self.fh, fcntl.LOCK_UN)
        self.fh.close()


def build(verbose=False):
    """Compile both libraries unconditionally."""
    with _Lock():
        return _build_unlocked(verbose)


def ensure_built(verbose=False):
    """Compile only if a library is missing or older than its sources."""
    if is_current():
        return LIB
    with _Lock():
        if is_current():
            return LIB
        sys.stderr.write("cubnm_amd: compiling the gfx950 core (hipcc)...\n")
        sys.stderr.flush()
        return _build_unlocked(verbose)


if __name__ == "__main__":
    build(verbose=True)
