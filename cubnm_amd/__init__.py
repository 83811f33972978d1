"""cubnm_amd — MI355X (gfx950) simulation core for cuBNM brain network models.

``cubnm_amd.core`` is the drop-in replacement of the reference's ``cubnm.core``
extension (run_simulations / init / set_const); ``cubnm_amd.sim`` holds
SimGroup-compatible callers and ``cubnm_amd.utils`` the goodness-of-fit and
empirical FC/FCD helpers.
"""
from ._core import run_simulations, set_const, init  # noqa: F401

__version__ = "0.1.0"
