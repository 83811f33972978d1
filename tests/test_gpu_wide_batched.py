"""Batched integrator above 256 nodes (integrate_batched_wide.hip) vs the oracle.

257..512 nodes: the SC fragments stream from L2 every step, the 16-simulation
groups keep one to four 16-node tiles per wave, and the regional parameters
and Balloon-Windkessel states live in a lane-ordered copy in global memory.
The node counts cover both tile deals (two or three tiles per wave, with and
without a remainder tile: 257 -> 17 tiles, 300 -> 19, 384 -> 24, 400 -> 25,
512 -> 32), a last tile that reaches past N, and a padded group (19
simulations = 16 + 3).  Tolerances as the other parity tests: BOLD norm-wise
relative error <= 1e-9 per simulation, FC absolute <= 1e-9, time-mean states
relative <= 1e-9.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _check(out, ref):
    rel = cases.bold_rel_err(out["sim_bold"], ref["sim_bold"])
    assert rel.max() <= TOL, f"BOLD rel err {rel.max():.3e}"
    e = cases.max_abs_err(out["sim_fc_trils"], ref["sim_fc_trils"])
    assert e <= TOL, f"FC abs err {e:.3e}"
    s, r = out["_sim_states"], ref["_sim_states"]
    scale = np.maximum(np.nanmax(np.abs(r), axis=(1, 2), keepdims=True), 1e-300)
    e = cases.max_abs_err(s / scale, r / scale)
    assert e <= TOL, f"states rel err {e:.3e}"


def _run(model, nodes, S, G, duration_s=3, sc_indices=None, scs=None, dt="0.1", noise_s=30, **kw):
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    sc = scs if scs is not None else cases.synthetic_sc(nodes, seed=nodes)
    do_fic = kw.pop("do_fic", False)
    gp, rp = cases.params(model, S, nodes, G=G, do_fic=do_fic, **kw)
    cfg = cases.model_config(model, TR=0.5, window_size=2, window_step=1, bold_remove_s=0, do_fcd=False,
                             do_fic=do_fic, noise_segment_length=noise_s)
    args = cases.sim_args(model, sc, gp, rp, cfg, S, duration_s=duration_s, TR_s=0.5, dt=dt,
                          sc_indices=sc_indices)
    out = _core.run_simulations(*cases.copy_args(args))
    assert _core.core.last_timings()["integrator"] == 1, "expected the batched integrator"
    ref = cases.run_oracle(cases.copy_args(args))
    return out, ref


@pytest.mark.parametrize("nodes", [257, 300, 384, 400, 512])
def test_rww_node_counts(nodes):
    S = 19
    out, ref = _run("rWW", nodes, S, np.linspace(0.05, 0.4, S), wIE=1.0)
    _check(out, ref)


def test_rww_analytical_fic():
    """Analytical FIC: wIE from the host root finder, then the batched step."""
    S = 16
    out, ref = _run("rWW", 400, S, np.linspace(0.05, 0.4, S), do_fic=True)
    _check(out, ref)
    np.testing.assert_array_equal(out["_global_bools"], ref["_global_bools"])


def test_rwwex_and_wc():
    out, ref = _run("rWWEx", 300, 17, np.linspace(0.1, 0.6, 17))
    _check(out, ref)
    out, ref = _run("WC", 270, 16, np.linspace(0.05, 0.3, 16))
    _check(out, ref)


def test_groups_per_sc():
    """Two SCs: simulations are grouped per SC, each group streams its own
    fragments."""
    N, S = 320, 20
    scs = np.stack([cases.synthetic_sc(N, seed=s) for s in range(2)])
    sci = np.array([i % 2 for i in range(S)], np.intc)
    out, ref = _run("rWW", N, S, np.linspace(0.05, 0.4, S), scs=scs, sc_indices=sci, wIE=1.0)
    _check(out, ref)


@pytest.mark.parametrize("dt", ["1.0", "0.5"])
def test_noise_repeats_and_inner_steps(dt):
    """A 1 s noise segment reused (with its node and step shuffles) three
    times over 3.5 s, and one or two model steps per BW step: the noise row
    staged one step ahead must follow the shuffled segment across BW steps
    and repeats."""
    out, ref = _run("rWW", 300, 16, np.linspace(0.05, 0.4, 16), duration_s=3.5, dt=dt, noise_s=1, wIE=1.0)
    _check(out, ref)
