"""Loads the compiled core and initialises its session constants on import,
as the reference's src/cubnm/_core.py:9-17 does for ``cubnm.core``.

There is no fallback: if the extension is missing the import fails loudly.
"""
try:
    from . import core
except ImportError as e:  # pragma: no cover - exercised only on unbuilt trees
    raise ImportError(
        "cubnm_amd.core is not built; run `python -m cubnm_amd.build` (hipcc, gfx950)"
    ) from e

run_simulations = core.run_simulations
set_const = core.set_const
init = core.init
gof_fc_corr = core.gof_fc_corr
gof_fcd_ks = core.gof_fcd_ks
last_timings = core.last_timings
force_integrator = core.force_integrator

core.init()
