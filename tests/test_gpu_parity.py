"""HIP core vs CPU oracle on identical inputs and identical host noise.

Tolerances (FP64 throughout): BOLD norm-wise relative error
max|Δ|/max|BOLD| <= 1e-9 per simulation; FC and FCD absolute error <= 1e-9;
time-mean states relative <= 1e-9; FIC-written wIE relative <= 1e-12 (the
analytical FIC is host arithmetic in both).  Kuramoto is chaotic: its phase
trajectories separate from the oracle at a rate set by the largest Lyapunov
exponent, so it is checked over a short horizon only (tolerance stated there).
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.fixture(scope="module", params=["general", "general_ks4", "batched"])
def core(request):
    """general_ks4: the per-simulation integrator with its coupling sum split
    over 4 partitions of source nodes (partial sums added in partition order,
    a different rounding order from the oracle's single sum)."""
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    _core.force_integrator(1 if request.param == "batched" else 0)
    _core.core.set_tuning("general_ksplit", 4 if request.param == "general_ks4" else 0)
    yield _core
    _core.force_integrator(-1)
    _core.core.set_tuning("general_ksplit", 0)


def _batched(core):
    return core.core.last_timings()["integrator"] == 1


def _compare(out, ref, model, tol=TOL, check_states=True):
    rel = cases.bold_rel_err(out["sim_bold"], ref["sim_bold"])
    assert rel.max() <= tol, f"BOLD rel err {rel.max():.3e}"
    if "sim_fc_trils" in ref:
        e = cases.max_abs_err(out["sim_fc_trils"], ref["sim_fc_trils"])
        assert e <= tol, f"FC abs err {e:.3e}"
    if "sim_fcd_trils" in ref:
        e = cases.max_abs_err(out["sim_fcd_trils"], ref["sim_fcd_trils"])
        assert e <= tol, f"FCD abs err {e:.3e}"
    if check_states:
        s, r = out["_sim_states"], ref["_sim_states"]
        scale = np.maximum(np.nanmax(np.abs(r), axis=(1, 2), keepdims=True), 1e-300)
        e = cases.max_abs_err(s / scale, r / scale)
        assert e <= tol, f"states rel err {e:.3e}"
    np.testing.assert_array_equal(out["_global_bools"], ref["_global_bools"])
    np.testing.assert_array_equal(out["_global_ints"], ref["_global_ints"])


def _run_both(core, args):
    a1, a2 = cases.copy_args(args), cases.copy_args(args)
    out = core.run_simulations(*a1)
    ref = cases.run_oracle(a2)
    return out, ref, a1, a2


@pytest.mark.parametrize("model", ["rWW", "JR", "WC", "rWWEx"])
def test_small_network_all_models(core, model):
    N, S = 10, 3
    sc = cases.synthetic_sc(N, seed=1)
    gp, rp = cases.params(model, S, N, G=[0.5, 1.0, 2.0])
    cfg = cases.model_config(model, TR=1.0, window_size=4, window_step=2, bold_remove_s=2)
    args = cases.sim_args(model, sc, gp, rp, cfg, S, duration_s=12, noise_out=True, force_reinit=True)
    out, ref, a1, a2 = _run_both(core, args)
    _compare(out, ref, model)
    np.testing.assert_array_equal(out["_noise"], ref["_noise"])
    np.testing.assert_array_equal(out["_shuffled_nodes"], ref["_shuffled_nodes"])
    np.testing.assert_array_equal(out["_shuffled_ts"], ref["_shuffled_ts"])
    np.testing.assert_allclose(a1[5], a2[5], rtol=1e-12, atol=0)


def test_kuramoto_short_horizon(core):
    # chaotic model: 4 s of simulated time keeps trajectories within 1e-7
    N, S = 10, 2
    sc = cases.synthetic_sc(N, seed=2)
    gp, rp = cases.params("Kuramoto", S, N, G=[0.5, 1.0])
    cfg = cases.model_config("Kuramoto", TR=0.5, window_size=1, window_step=1, bold_remove_s=1, do_fcd=False)
    args = cases.sim_args("Kuramoto", sc, gp, rp, cfg, S, duration_s=4, TR_s=0.5)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, "Kuramoto", tol=1e-7)


@pytest.mark.parametrize("fic", [0, 1, 2])
def test_rww_fic_modes(core, fic, request):
    if fic == 2 and request.node.callspec.params["core"] == "batched":
        pytest.skip("numerical-FIC restarts run on the general integrator")
    N, S = 16, 4
    sc = cases.synthetic_sc(N, seed=3) * 0.05
    gp, rp = cases.params("rWW", S, N, G=[0.5, 1.0, 1.5, 2.0], do_fic=fic > 0)
    cfg = cases.model_config("rWW", TR=1.0, window_size=4, window_step=2, bold_remove_s=10,
                             do_fic=fic > 0, max_fic_trials=5 if fic == 2 else 0)
    args = cases.sim_args("rWW", sc, gp, rp, cfg, S, duration_s=20)
    out, ref, a1, a2 = _run_both(core, args)
    _compare(out, ref, "rWW")
    np.testing.assert_allclose(a1[5], a2[5], rtol=1e-12, atol=0)


@pytest.mark.parametrize("model", ["rWW", "JR"])
def test_conduction_delay(core, model, request):
    if request.node.callspec.params["core"] == "batched":
        pytest.skip("conduction delay runs on the general integrator")
    N, S = 12, 2
    sc = cases.synthetic_sc(N, seed=4)
    dist = cases.synthetic_sc(N, seed=5) * 60.0
    gp, rp = cases.params(model, S, N, G=[0.5, 1.0])
    cfg = cases.model_config(model, TR=1.0, window_size=4, window_step=2, bold_remove_s=2)
    args = cases.sim_args(model, sc, gp, rp, cfg, S, duration_s=10, dt="1.0", sc_dist=dist, v=[0.5, 2.0])
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, model)


@pytest.mark.parametrize("model", ["rWW", "Kuramoto"])
def test_large_network_pair_mode(core, model):
    # 144 nodes: SC too large for LDS, so the per-simulation integrator sums the
    # coupling two columns per lane with 16-byte loads (pair mode)
    N, S = 144, 2
    sc = cases.synthetic_sc(N, seed=8)
    gp, rp = cases.params(model, S, N, G=[0.4, 0.8], do_fic=False)
    TR = 0.5
    cfg = cases.model_config(model, TR=TR, window_size=1, window_step=1, bold_remove_s=1, do_fcd=False,
                             do_fic=False)
    dur = 4 if model == "rWW" else 3
    args = cases.sim_args(model, sc, gp, rp, cfg, S, duration_s=dur, TR_s=TR)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, model, tol=TOL if model == "rWW" else 1e-7)


def test_states_time_series(core, request):
    if request.node.callspec.params["core"] == "batched":
        pytest.skip("state time series run on the general integrator")
    N, S = 8, 2
    sc = cases.synthetic_sc(N, seed=6)
    gp, rp = cases.params("WC", S, N, G=[0.5, 1.0])
    cfg = cases.model_config("WC", TR=0.5, window_size=2, window_step=1, bold_remove_s=1)
    args = cases.sim_args("WC", sc, gp, rp, cfg, S, duration_s=6, TR_s=0.5, states_ts=True)
    out, ref, _, _ = _run_both(core, args)
    assert out["_sim_states"].shape == (2, S, N * 12)
    _compare(out, ref, "WC")


def test_c1_plain_grid_rww(core):
    """C1: 8 sims, N=100, 60 s, G in [0.5, 3] x w_p in {0, 2}, J_N 0.15."""
    N = 100
    sc = cases.synthetic_sc(N, seed=0)
    G = np.repeat(np.linspace(0.5, 3, 4), 2)
    wp = np.tile([0.0, 2.0], 4)
    gp, rp = cases.params("rWW", 8, N, G=G, w_p=wp[:, None], J_N=0.15, do_fic=False, wIE=1.0)
    cfg = cases.model_config("rWW", do_fic=False)
    args = cases.sim_args("rWW", sc, gp, rp, cfg, 8, duration_s=60)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, "rWW")


@pytest.mark.parametrize("nodes", [1, 5, 16, 33, 68, 97, 104, 130])
def test_batched_node_counts(core, nodes):
    """Tile padding: node counts that are not multiples of 4 or 16, one tile,
    several tiles, 97..104 nodes (7 tiles of 26 fragments: the compile-time
    fragment count) and the L2-streamed SC fragments (N > 112)."""
    S = 19  # one full group of 16 plus a padded group
    sc = cases.synthetic_sc(nodes, seed=nodes)
    gp, rp = cases.params("rWW", S, nodes, G=np.linspace(0.2, 2.0, S), do_fic=False, wIE=1.0)
    cfg = cases.model_config("rWW", do_fic=False, TR=1.0, window_size=4, window_step=2, bold_remove_s=2,
                             do_fc=nodes > 1, do_fcd=nodes > 1)
    args = cases.sim_args("rWW", sc, gp, rp, cfg, S, duration_s=10)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, "rWW")


@pytest.mark.parametrize("nodes", [33, 100])
@pytest.mark.parametrize("dt", ["0.2", "1.0"])
def test_odd_inner_step_count(core, nodes, dt):
    """An odd number of integration steps per Balloon-Windkessel step (5 at
    dt 0.2 ms, 1 at dt 1 ms): the batched integrator's two X buffers then
    alternate across bw steps as well as inside them."""
    S = 16
    sc = cases.synthetic_sc(nodes, seed=nodes + 3)
    gp, rp = cases.params("rWW", S, nodes, G=np.linspace(0.2, 2.0, S), do_fic=False, wIE=1.0)
    cfg = cases.model_config("rWW", do_fic=False, TR=1.0, window_size=4, window_step=2, bold_remove_s=2)
    args = cases.sim_args("rWW", sc, gp, rp, cfg, S, duration_s=10, dt=dt)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, "rWW")


def test_rwwex_c1_shape(core):
    """rWWEx on the C1 shape (N = 100, 60 s, 8 simulations over G)."""
    N = 100
    sc = cases.synthetic_sc(N, seed=0)
    gp, rp = cases.params("rWWEx", 8, N, G=np.linspace(0.2, 2.0, 8))
    cfg = cases.model_config("rWWEx")
    args = cases.sim_args("rWWEx", sc, gp, rp, cfg, 8, duration_s=60)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, "rWWEx")


def test_batched_mixed_scs(core):
    """Simulations with different SCs are grouped per SC."""
    N, S = 20, 21
    scs = np.stack([cases.synthetic_sc(N, seed=s) for s in range(3)])
    sci = np.array([i % 3 for i in range(S)], np.intc)
    gp, rp = cases.params("WC", S, N, G=np.linspace(0.1, 1.0, S))
    cfg = cases.model_config("WC", TR=1.0, window_size=4, window_step=2, bold_remove_s=2)
    args = cases.sim_args("WC", scs, gp, rp, cfg, S, duration_s=8, sc_indices=sci)
    out, ref, _, _ = _run_both(core, args)
    _compare(out, ref, "WC")
