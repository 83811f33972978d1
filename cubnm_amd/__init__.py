"""cubnm_amd — MI355X (gfx950) simulation core for cuBNM brain network models.

``cubnm_amd.core`` is the drop-in replacement of the reference's ``cubnm.core``
extension (run_simulations / init / set_const); ``cubnm_amd.sim`` holds
SimGroup-compatible callers and ``cubnm_amd.utils`` the goodness-of-fit and
empirical FC/FCD helpers.

Importing the package compiles the core in-tree when it is missing or older
than its sources (``build.ensure_built``); nothing here imports the compiled
module itself, so the builder stays importable on an unbuilt tree.
``run_simulations``, ``set_const`` and ``init`` are resolved on first access.
"""
from . import build as _build

_build.ensure_built()

__version__ = "0.1.0"


def __getattr__(name):
    if name in ("run_simulations", "set_const", "init"):
        from . import _core
        return getattr(_core, name)
    raise AttributeError(f"module 'cubnm_amd' has no attribute {name!r}")
