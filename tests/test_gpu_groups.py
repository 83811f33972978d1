"""Heterogeneous batches through one call (cubnm_run_groups / core.run_groups).

Each request of a mixed batch (different models, node counts, durations,
SCs, FIC settings and noise outputs) must give what the same request gives
alone through run_simulations: noise, shuffles and integer outputs exactly,
BOLD and states to 1e-12, FC / FCD to 1e-10 (requests on the
per-simulation integrator run in one merged launch whose kernel the compiler
contracts into FMAs a little differently: BOLD 1.2e-14 observed, amplified
to 8.7e-12 in the correlation between two 5-volume windows).  The parity tests tie
run_simulations to the oracle; one request is also checked against the
oracle directly at the usual 1e-9 tolerance.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _requests():
    reqs = []
    # G ranges keep every simulation well conditioned (one ulp of G moves its
    # BOLD by < 1e-11 in the oracle)
    specs = [("rWW", 10, 12, 3, 1, 1.5), ("JR", 17, 8, 2, 2, 1.5), ("Kuramoto", 6, 10, 4, 3, 1.5),
             ("WC", 33, 10, 5, 4, 0.7), ("rWW", 10, 20, 2, 5, 1.5), ("rWW", 70, 8, 17, 6, 1.5)]
    for model, N, dur, S, seed, g_hi in specs:
        sc = cases.synthetic_sc(N, seed=seed)
        gp, rp = cases.params(model, S, N, G=np.linspace(0.3, g_hi, S))
        cfg = cases.model_config(model, window_size=4, window_step=2, bold_remove_s=2)
        reqs.append(cases.sim_args(model, sc, gp, rp, cfg, S, duration_s=dur, noise_out=True, sim_seed=seed % 3))
    return reqs


def test_groups_equal_single_requests():
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    reqs = _requests()
    grouped = _core.core.run_groups([cases.copy_args(r) for r in reqs])
    assert len(grouped) == len(reqs)
    for r, g in zip(reqs, grouped):
        a = cases.copy_args(r)
        single = _core.run_simulations(*a)
        # grouped results also carry the integration time of their launch
        assert g.pop("integrate_ms") > 0
        assert set(g) == set(single)
        for k in single:
            if k in ("init_time", "run_time"):
                continue
            msg = f"{r[0]} N={r[15]}: {k}"
            if k == "sim_bold":
                assert cases.bold_rel_err(g[k], single[k]).max() <= 1e-12, msg
            elif k in ("sim_fc_trils", "sim_fcd_trils"):
                assert cases.max_abs_err(g[k], single[k]) <= 1e-10, msg
            elif k == "_sim_states":
                scale = np.maximum(np.abs(single[k]).max(), 1e-300)
                assert cases.max_abs_err(g[k] / scale, single[k] / scale) <= 1e-12, msg
            else:
                np.testing.assert_array_equal(g[k], single[k], err_msg=msg)


def test_groups_match_oracle_and_write_back():
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    reqs = _requests()[:2]
    a = [cases.copy_args(r) for r in reqs]
    grouped = _core.core.run_groups(a)
    ref = cases.run_oracle(cases.copy_args(reqs[0]))
    assert cases.bold_rel_err(grouped[0]["sim_bold"], ref["sim_bold"]).max() <= 1e-9
    assert cases.max_abs_err(grouped[0]["sim_fc_trils"], ref["sim_fc_trils"]) <= 1e-9
    # analytical FIC wrote wIE into the request's own regional_params
    b = cases.copy_args(reqs[0])
    cases.run_oracle(b)
    np.testing.assert_allclose(a[0][5], b[5], rtol=1e-12, atol=0)
    assert np.all(a[0][5][2] > 0)


def test_pair_mode_requests_match_oracle():
    # 100 and 120 nodes: SCs above the merged launch's 64 KB LDS budget, so the
    # one per-simulation launch sums their coupling in pair mode
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    reqs = []
    for model, N, dur, seed in [("rWW", 100, 4, 9), ("Kuramoto", 120, 2, 10)]:
        sc = cases.synthetic_sc(N, seed=seed)
        gp, rp = cases.params(model, 2, N, G=[0.4, 0.8], do_fic=False)
        cfg = cases.model_config(model, TR=0.5, window_size=1, window_step=1, bold_remove_s=1, do_fcd=False,
                                 do_fic=False)
        reqs.append(cases.sim_args(model, sc, gp, rp, cfg, 2, duration_s=dur, TR_s=0.5))
    grouped = _core.core.run_groups([cases.copy_args(r) for r in reqs])
    for r, g, tol in zip(reqs, grouped, [1e-9, 1e-7]):
        ref = cases.run_oracle(cases.copy_args(r))
        assert cases.bold_rel_err(g["sim_bold"], ref["sim_bold"]).max() <= tol, r[0]


@pytest.mark.parametrize("pair", [1, -1], ids=["pair_mode", "single_column"])
def test_c3_shaped_large_networks_match_oracle(pair):
    """The uneven batch's large members: N = 200 and 400 (TR 0.72 s; the
    Kuramoto request N = 200 at TR 0.5 s), every simulation with its own
    seeded lognormal SC (the per-simulation integrator streams each
    SC from memory), rWW, JR and Kuramoto, in one grouped call;
    the coupling sums two adjacent SC columns per lane with 16-byte loads
    (pair mode) or one column per lane (the form many one-SC-each
    simulations take).  Kuramoto is chaotic and is checked at 1e-7 over a
    short horizon, the others at 1e-9."""
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    # (model, N, duration s, simulations, seed, TR s); the oracle's Kuramoto
    # evaluates N^2 sines per step, so that request is kept short
    specs = [("rWW", 200, 5.0, 3, 21, 0.72), ("rWW", 400, 3.0, 2, 22, 0.72), ("JR", 200, 4.0, 2, 23, 0.72),
             ("Kuramoto", 200, 1.5, 2, 24, 0.5)]
    reqs = []
    for model, N, dur, S, seed, TR in specs:
        scs = np.stack([cases.synthetic_sc(N, seed=seed * 100 + i) for i in range(S)])
        gp, rp = cases.params(model, S, N, G=np.linspace(0.2, 0.8, S), do_fic=False)
        cfg = cases.model_config(model, TR=TR, window_size=2, window_step=1, bold_remove_s=0, do_fcd=False,
                                 do_fic=False)
        reqs.append(cases.sim_args(model, scs, gp, rp, cfg, S, duration_s=dur, TR_s=TR,
                                   sc_indices=np.arange(S), sim_seed=seed % 4))
    _core.core.set_tuning("general_pair", pair)
    try:
        grouped = _core.core.run_groups([cases.copy_args(r) for r in reqs])
    finally:
        _core.core.set_tuning("general_pair", 0)
    for r, g, spec in zip(reqs, grouped, specs):
        ref = cases.run_oracle(cases.copy_args(r))
        tol = 1e-7 if spec[0] == "Kuramoto" else 1e-9
        msg = f"{spec[0]} N={spec[1]}"
        assert cases.bold_rel_err(g["sim_bold"], ref["sim_bold"]).max() <= tol, msg
        assert cases.max_abs_err(g["sim_fc_trils"], ref["sim_fc_trils"]) <= tol, msg
        s, rs = g["_sim_states"], ref["_sim_states"]
        scale = np.maximum(np.nanmax(np.abs(rs), axis=(1, 2), keepdims=True), 1e-300)
        assert cases.max_abs_err(s / scale, rs / scale) <= tol, msg


def _per_sim_requests(g_shift):
    # requests that all run on the per-simulation integrator (own SCs, or a
    # delay), so the merged launch is the whole integration
    reqs = []
    for model, N, S, seed in [("rWW", 12, 3, 31), ("JR", 9, 2, 32), ("Kuramoto", 7, 2, 33)]:
        scs = np.stack([cases.synthetic_sc(N, seed=seed * 10 + i) for i in range(S)])
        gp, rp = cases.params(model, S, N, G=np.linspace(0.3, 0.9, S) + g_shift, do_fic=False)
        cfg = cases.model_config(model, window_size=4, window_step=2, bold_remove_s=2, do_fic=False)
        reqs.append(cases.sim_args(model, scs, gp, rp, cfg, S, duration_s=8, sc_indices=np.arange(S),
                                   sim_seed=seed % 3))
    return reqs


def test_group_graph_replays_match_direct_launches():
    """The merged launch is captured into a hipGraph on the first call and
    replayed on later calls with the same shapes (new parameter values each
    time); every replay equals a direct launch of the same call, with direct
    launches and synchronisations between the replays."""
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    try:
        for call, shift in enumerate([0.0, 0.05, -0.1, 0.2]):
            reqs = _per_sim_requests(shift)
            _core.core.set_tuning("group_graph", 1)
            graphed = _core.core.run_groups([cases.copy_args(r) for r in reqs])
            info = _core.core.run_groups_info()
            assert info["graph"] == (1 if call == 0 else 2), (call, info)
            _core.core.set_tuning("group_graph", 0)
            direct = _core.core.run_groups([cases.copy_args(r) for r in reqs])
            assert _core.core.run_groups_info()["graph"] == 0
            for g, d in zip(graphed, direct):
                for k in ["sim_bold", "sim_fc_trils", "sim_fcd_trils", "_sim_states"]:
                    np.testing.assert_array_equal(g[k], d[k], err_msg=f"call {call} {k}")
            _core.core.set_tuning("group_graph", 1)
    finally:
        _core.core.set_tuning("group_graph", 1)


def test_autotuner_times_shapes_and_caches(tmp_path, monkeypatch):
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("CUBNM_AMD_TUNE_CACHE", str(cache))
    reqs = _per_sim_requests(0.0)
    _core.core.set_tuning("autotune", 1)
    try:
        first = _core.core.run_groups([cases.copy_args(r) for r in reqs])
        info = _core.core.run_groups_info()
        assert info["tuned"] in (1, 2) and info["ksplit"] >= 1
        if info["tuned"] == 1:
            assert info["shapes_timed"] >= 2 and info["trial_ms"] > 0
            assert cache.exists() and "general" in cache.read_text()
        second = _core.core.run_groups([cases.copy_args(r) for r in reqs])
        assert _core.core.run_groups_info()["tuned"] == 2
    finally:
        _core.core.set_tuning("autotune", 0)
    for r, a, b in zip(reqs, first, second):
        np.testing.assert_array_equal(a["sim_bold"], b["sim_bold"])
        ref = cases.run_oracle(cases.copy_args(r))
        tol = 1e-7 if r[0] == "Kuramoto" else 1e-9
        assert cases.bold_rel_err(a["sim_bold"], ref["sim_bold"]).max() <= tol, r[0]
