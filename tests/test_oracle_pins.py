"""Pins for the CPU oracle (no GPU).

What each test pins, and the plausible mistake it would catch:
  * FC/FCD from the fixtures' own BOLD, through the oracle's window table and
    one-pass Pearson, equal the reference's FC/FCD fixtures to 1e-10: pair
    order, burn-in rows, window starts/ends (ws+1 volumes), FCD pair order.
  * Simulation semantics vs the fixtures: this image's libstdc++ does not
    reproduce the GCC 12 noise stream the fixtures were made with, so the
    simulation is pinned statistically: the node-wise time-mean BOLD of the
    fixture lies inside a 4-seed oracle ensemble (RMS z <= 2), while the same
    statistic rejects a shifted coupling G (RMS z >= 3).
  * Pearson: one-pass recurrence vs numpy.corrcoef on random data.
  * Analytical-FIC root: the Brent iterate satisfies the reference's stopping
    rule against scipy.optimize.brentq (|root - exact| <= 1e-3 * exact).
  * Balloon-Windkessel rest state: zero input from x=0, f=nu=q=1 stays at
    rest, so BOLD is exactly 0 (Kuramoto with theta = omega = sigma = G = 0).
  * Node-permutation equivariance without noise; identical simulations give
    identical outputs; loop counts and table sizes for the named configs.
"""
import numpy as np
import pytest
import scipy.optimize

import cases


def _fixture_keys(model):
    d = np.load(f"{cases.GOLDEN}/sim_{model}.npz")
    return d, sorted({k.split("|")[0] for k in d.files})


@pytest.mark.parametrize("model", ["rWW", "JR", "Kuramoto", "WC", "rWWEx"])
def test_fc_fcd_from_fixture_bold(oracle_lib, model):
    O = oracle_lib
    d, keys = _fixture_keys(model)
    dims = O.dims(model, cases.model_config(model, window_size=10, window_step=2), 100, 60000, 1000, 1000, 0.1, 1.0)
    ws, we = dims["window_starts"], dims["window_ends"]
    pi, pj = np.tril_indices(100, -1)
    for key in keys:
        bold = d[f"{key}|sim_bold"][0]
        z = bold[30:]
        fc = np.array([O.corr(z[:, i], z[:, j]) for i, j in zip(pi, pj)])
        np.testing.assert_allclose(fc, d[f"{key}|sim_fc_trils"][0], rtol=0, atol=1e-10)
        wfc = np.array([[O.corr(bold[s:e + 1, i], bold[s:e + 1, j]) for i, j in zip(pi, pj)] for s, e in zip(ws, we)])
        fcd = np.array([O.corr(wfc[a], wfc[b]) for a, b in zip(*np.tril_indices(len(ws), -1))])
        np.testing.assert_allclose(fcd, d[f"{key}|sim_fcd_trils"][0], rtol=0, atol=1e-10)


def _oracle_bold(model, seed, G, do_fic):
    sc, dist = cases.hcp_sc(), cases.hcp_sc_length()
    gp, rp = cases.params(model, 1, 100, G=[G], sim_seed=0, do_fic=do_fic)
    cfg = cases.model_config(model, do_fic=do_fic)
    args = cases.sim_args(model, sc, gp, rp, cfg, 1, duration_s=60, dt="1.0", sc_dist=dist, v=0.5, sim_seed=seed)
    return cases.run_oracle(args, n_threads=1)["sim_bold"].reshape(60, 100)


@pytest.mark.parametrize("model,fic,g_bad", [("JR", 0, 0.8), ("WC", 0, 0.6), ("rWW", 0, 0.6), ("rWW", 1, 0.8)])
def test_fixture_statistical_consistency(model, fic, g_bad):
    """Reference fixture (do_delay:1, CPU) vs an oracle seed ensemble."""
    d, _ = _fixture_keys(model)
    key = "force_cpu:1,do_delay:1" + (f",do_fic:{fic}" if model == "rWW" else "")
    fixture = d[f"{key}|sim_bold"][0]
    tmean = lambda b: b[30:].mean(axis=0)
    ens = np.array([tmean(_oracle_bold(model, s, 0.5, fic > 0)) for s in range(4)])
    mu, sd = ens.mean(axis=0), ens.std(axis=0, ddof=1) + 1e-12
    rms_z = lambda b: float(np.sqrt(np.mean(((tmean(b) - mu) / sd) ** 2)))
    assert rms_z(fixture) <= 2.0
    assert rms_z(_oracle_bold(model, 7, g_bad, fic > 0)) >= 3.0


def test_corr_matches_numpy(oracle_lib):
    rng = np.random.default_rng(0)
    for n in [2, 3, 11, 30, 4950]:
        x, y = rng.standard_normal(n), rng.standard_normal(n) + 0.3
        assert abs(oracle_lib.corr(x, y) - np.corrcoef(x, y)[0, 1]) < 1e-13


def test_fic_root_meets_bracket_rule(oracle_lib):
    a_I, b_I, d_I = 615.0, 177.0, 0.087
    I0_I, S_E_ss = 0.382 * 0.7, 0.1647572075

    def f(x, w):
        return I0_I + w * S_E_ss - 0.01 * ((a_I * x - b_I) / (1 - np.exp(-1 * d_I * (a_I * x - b_I)))) - x

    for w in np.linspace(0.001, 0.5, 60):
        r = oracle_lib.fic_root(w)
        exact = scipy.optimize.brentq(f, 0, 2, args=(w,), xtol=1e-16, rtol=8.9e-16)
        assert abs(r - exact) <= 1e-3 * exact
    # J_N = 0.15 (reference default): converged well beyond the stopping rule
    r = oracle_lib.fic_root(0.15)
    assert abs(r - scipy.optimize.brentq(f, 0, 2, args=(0.15,), xtol=1e-16, rtol=8.9e-16)) < 1e-9


def test_bw_rest_state_gives_zero_bold():
    N, S = 6, 2
    gp, rp = cases.params("Kuramoto", S, N, G=[0.0, 0.0], init_theta=0.0, omega=0.0, sigma=0.0)
    cfg = cases.model_config("Kuramoto", window_size=2, window_step=1, bold_remove_s=2)
    out = cases.run_oracle(cases.sim_args("Kuramoto", cases.synthetic_sc(N), gp, rp, cfg, S, duration_s=6))
    assert np.all(out["sim_bold"] == 0.0)


def test_permutation_equivariance_without_noise():
    N, S = 12, 2
    rng = np.random.default_rng(3)
    sc = cases.synthetic_sc(N, seed=9)
    wp = rng.uniform(0.5, 1.5, N)
    perm = rng.permutation(N)
    cfg = cases.model_config("rWW", window_size=4, window_step=2, bold_remove_s=2, do_fic=False)
    gp, rp = cases.params("rWW", S, N, G=[0.8, 1.6], w_p=wp, sigma=0.0, do_fic=False, wIE=1.0)
    gp2, rp2 = cases.params("rWW", S, N, G=[0.8, 1.6], w_p=wp[perm], sigma=0.0, do_fic=False, wIE=1.0)
    a = cases.run_oracle(cases.sim_args("rWW", sc, gp, rp, cfg, S, duration_s=10))
    b = cases.run_oracle(cases.sim_args("rWW", sc[np.ix_(perm, perm)], gp2, rp2, cfg, S, duration_s=10))
    ba, bb = a["sim_bold"].reshape(S, -1, N), b["sim_bold"].reshape(S, -1, N)
    np.testing.assert_allclose(bb, ba[:, :, perm], rtol=1e-11, atol=1e-14)


def test_identical_sims_identical():
    N, S = 10, 2
    gp, rp = cases.params("WC", S, N, G=[0.7, 0.7])
    cfg = cases.model_config("WC", window_size=4, window_step=2, bold_remove_s=2)
    out = cases.run_oracle(cases.sim_args("WC", cases.synthetic_sc(N), gp, rp, cfg, S, duration_s=8))
    np.testing.assert_array_equal(out["sim_bold"][0], out["sim_bold"][1])
    np.testing.assert_array_equal(out["sim_fc_trils"][0], out["sim_fc_trils"][1])


def test_loop_counts_and_tables(oracle_lib):
    O = oracle_lib
    c1 = O.dims("rWW", cases.model_config("rWW"), 100, 60000, 1000, 1000, 0.1, 1.0)
    assert (c1["bw_it"], c1["inner_it"], c1["bold_len"], c1["n_pairs"]) == (60000, 10, 60, 4950)
    assert (c1["n_windows"], c1["n_window_pairs"], c1["n_vols_remove"], c1["corr_len"]) == (10, 45, 30, 30)
    assert list(c1["window_starts"]) == list(range(30, 50, 2))
    assert c1["noise_bw_it"] == 30000 and c1["noise_repeats"] == 2 and c1["noise_size"] == 100 * 30000 * 10 * 2
    c4 = O.dims("rWW", cases.model_config("rWW", TR=0.72), 100, 120000, 720, 720, 0.1, 1.0)
    assert c4["bold_len"] == 166
    # N=400, 300 s, TR 0.72 s, window 30 TRs step 1, no burn-in: reference window rule
    cfg5 = cases.model_config("rWW", TR=0.72, bold_remove_s=0)
    cfg5.update({"window_size": "30", "window_step": "1"})
    c5 = O.dims("rWW", cfg5, 400, 300000, 720, 720, 0.1, 1.0)
    assert (c5["bold_len"], c5["n_pairs"], c5["n_windows"]) == (416, 79800, 386)
