"""A clean checkout builds and imports.

The files git would commit (tracked plus untracked-but-not-ignored, so built
libraries are left behind) are copied to a temporary directory.  There:

* ``__graft_entry__.build()`` compiles the gfx950 library, the CPython module
  and the oracle, and imports ``cubnm_amd.core``;
* a second copy, never built, is imported directly: the package builds itself
  on first import (``cubnm_amd.build.ensure_built``), and a second import does
  not rebuild.

hipcc cross-compiles gfx950 without a GPU and the library's ``init()`` makes no
device call, so this runs on the CPU.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _checkout(dest):
    if not os.path.isdir(os.path.join(ROOT, ".git")) or shutil.which("git") is None:
        pytest.skip("needs the git work tree")
    files = subprocess.run(["git", "ls-files", "-co", "--exclude-standard", "-z"], cwd=ROOT, check=True,
                           capture_output=True).stdout.decode().split("\0")
    n = 0
    for f in files:
        if not f or not os.path.isfile(os.path.join(ROOT, f)):
            continue
        os.makedirs(os.path.join(dest, os.path.dirname(f)), exist_ok=True)
        shutil.copy2(os.path.join(ROOT, f), os.path.join(dest, f))
        n += 1
    assert not any(f.endswith(".so") for f in files), "a built library is not ignored by git"
    return n


def _py(code, cwd):
    env = dict(os.environ, PYTHONDONTWRITEBYTECODE="1")
    return subprocess.run([sys.executable, "-c", code], cwd=cwd, env=env, capture_output=True, text=True,
                          timeout=600)


def test_clean_checkout_builds_and_imports(tmp_path):
    assert _checkout(tmp_path) > 20
    assert not os.path.exists(tmp_path / "cubnm_amd" / "libcubnm_amd.so")
    r = _py("import __graft_entry__ as g; g.build()", tmp_path)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "cubnm_amd built and imported" in r.stdout
    r = _py("import cubnm_amd.core as c; c.init(); c.set_const('k1', 2.38); print(c.__file__)", tmp_path)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.strip().startswith(str(tmp_path))
    assert "compiling" not in r.stderr


def test_first_import_builds_on_demand(tmp_path):
    _checkout(tmp_path)
    code = ("import cubnm_amd, cubnm_amd.core as c; from cubnm_amd import capi; capi.lib(); "
            "print(c.__file__)")
    r = _py(code, tmp_path)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "compiling the gfx950 core" in r.stderr
    r2 = _py(code, tmp_path)
    assert r2.returncode == 0 and "compiling" not in r2.stderr, r2.stderr[-2000:]
