"""Loads the compiled core and initialises its session constants on import,
as the reference's src/cubnm/_core.py:9-17 does for ``cubnm.core``.

The package import has already compiled the core if it was missing or stale;
there is no fallback: if the extension cannot be loaded the import fails loudly.
"""
from . import build as _build

_build.ensure_built()
from . import core  # noqa: E402

run_simulations = core.run_simulations
set_const = core.set_const
init = core.init
gof_fc_corr = core.gof_fc_corr
gof_fcd_ks = core.gof_fcd_ks
last_timings = core.last_timings
force_integrator = core.force_integrator

core.init()
