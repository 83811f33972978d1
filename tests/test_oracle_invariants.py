"""Known-answer and invariant pins for the CPU oracle (no GPU).

Each test fixes something the fixtures leave open (their noise stream is not
reproducible with this image's libstdc++), chosen so that a plausible slip
fails it:
  * noise indexing: Kuramoto with G = 0 and omega = 0 is a pure random walk,
    theta_t = theta_0 + sqrt(dt) * sigma * (sum of the noise it drew), so the
    sampled phases reveal which noise element every (step, node) read.  The
    expected draw order restates the reference's own test_get_noise loop
    (tests/sim/test.py:188-258) over the returned _noise / _shuffled_* arrays.
    A wrong segment, repeat, shuffle or inner-step index fails it.
  * analytical FIC: without noise, every node's firing rate must settle at
    the FIC target r_E = 3.0773 Hz (rww.cpp:47-52) for any G, on an
    asymmetric SC; and wIE solved for SC must miss the target when the
    simulation couples through SC transposed (the FIC K_EE and the
    simulation's coupling use the same orientation).
  * G = 0 decouples nodes: outputs do not depend on SC.
  * Kuramoto coupling, two nodes without noise, SC[0,1] = a != SC[1,0] = b:
    a*theta_0 + b*theta_1 - (a + b)*omega*t is conserved (the coupling terms
    cancel; the transposed SC conserves b*theta_0 + a*theta_1 instead), and
    the phase difference follows tan(D/2) = tan(D0/2) exp(-G (a + b) t)
    (a flipped sine argument makes it grow, a lost or doubled G misses it).
  * FC/FCD of a 4-node simulation recomputed by brute force from its BOLD
    with numpy (burn-in rows, window bounds, pair order,
    inter-hemispheric exclusion).
  * hand-enumerated pair and window tables for a small case.
"""
import numpy as np
import pytest

import cases


def _drawn_noise(out, bw_it, inner_it, noise_bw_it, nodes, n_noise, which=0):
    """Noise element read at (bw step, inner step, node), following the
    reference's test_get_noise reconstruction loop."""
    noise = out["_noise"]
    shn = out["_shuffled_nodes"].ravel()
    sht = out["_shuffled_ts"].ravel()
    xi = np.empty((bw_it, inner_it, nodes))
    rep = 0
    for bw in range(bw_it):
        ts = sht[bw % noise_bw_it + rep * noise_bw_it]
        base = (ts * inner_it + np.arange(inner_it)[:, None]) * nodes * n_noise
        xi[bw] = noise[base + shn[rep * nodes:(rep + 1) * nodes][None, :] * n_noise + which]
        if (bw + 1) % noise_bw_it == 0 and bw + 1 < bw_it:
            rep += 1
    return xi


def test_noise_indexing_kuramoto_random_walk():
    N, dur_s, seg_s, sigma = 6, 3, 1, 1e-3
    gp, rp = cases.params("Kuramoto", 1, N, G=[0.0], init_theta=np.pi, omega=0.0, sigma=sigma)
    cfg = cases.model_config("Kuramoto", do_fc=False, do_fcd=False, bold_remove_s=0,
                             noise_segment_length=seg_s, random_init_theta=False)
    args = cases.sim_args("Kuramoto", cases.synthetic_sc(N), gp, rp, cfg, 1, duration_s=dur_s,
                          states_ts=True, noise_out=True, states_sampling_s=0.001)
    out = cases.run_oracle(args)
    assert out["_shuffled_nodes"].shape == (3, N) and out["_shuffled_ts"].shape == (3, 1000)
    theta = out["_sim_states"][0, 0].reshape(dur_s * 1000, N)     # sampled every BW step
    xi = _drawn_noise(out, dur_s * 1000, 10, seg_s * 1000, N, 1)
    expected = np.pi + np.sqrt(0.1) * sigma * np.cumsum(xi.sum(axis=1), axis=0)
    assert np.abs(theta - np.pi).max() > 1e-2                     # the walk moved
    np.testing.assert_allclose(theta, expected, rtol=0, atol=1e-12)


def test_kuramoto_two_node_closed_form():
    a, b, G, omega = 0.6, 0.2, 0.004, 0.01                # omega in rad/ms, dt = 0.1 ms
    sc = np.array([[0.0, a], [b, 0.0]])
    th0 = np.array([0.3, 2.9])
    gp, rp = cases.params("Kuramoto", 1, 2, G=[G], init_theta=th0, omega=omega, sigma=0.0)
    cfg = cases.model_config("Kuramoto", do_fc=False, do_fcd=False, bold_remove_s=0, random_init_theta=False)
    args = cases.sim_args("Kuramoto", sc, gp, rp, cfg, 1, duration_s=2, states_ts=True, states_sampling_s=0.001)
    out = cases.run_oracle(args)
    theta = np.unwrap(out["_sim_states"][0, 0].reshape(2000, 2), axis=0)   # phases are kept mod 2 pi
    t = np.arange(1, 2001, dtype=float)                                   # sample i is after i + 1 ms
    conserved = a * theta[:, 0] + b * theta[:, 1] - (a + b) * omega * t
    np.testing.assert_allclose(conserved, a * th0[0] + b * th0[1], rtol=0, atol=1e-12)
    transposed = b * theta[:, 0] + a * theta[:, 1] - (a + b) * omega * t
    assert np.abs(transposed - (b * th0[0] + a * th0[1])).max() > 0.5
    d = theta[:, 1] - theta[:, 0]
    expected = np.tan((th0[1] - th0[0]) / 2) * np.exp(-G * (a + b) * t)
    assert d[-1] < 0.02                                                   # synchronised
    np.testing.assert_allclose(np.tan(d / 2), expected, rtol=2e-3)        # Euler error ~6e-4


def _asym_case():
    N = 8
    rng = np.random.default_rng(4)
    sc = rng.uniform(0, 1, (N, N))
    np.fill_diagonal(sc, 0)
    sc /= sc.max() * 4
    return N, sc, rng.uniform(0.5, 1.5, N), rng.uniform(0.1, 0.3, N)


def test_fic_fixed_point_and_coupling_orientation():
    N, sc, wp, jn = _asym_case()
    Gs = [0.0, 0.1, 0.2, 0.4]
    S = len(Gs)
    r_target = 3.0773259   # rww.cpp:47-52 (S_E_ss-consistent r_E)
    gp, rp = cases.params("rWW", S, N, G=Gs, w_p=wp, J_N=jn, sigma=0.0, do_fic=True)
    cfg = cases.model_config("rWW", do_fc=False, do_fcd=False, bold_remove_s=10, do_fic=True)
    out = cases.run_oracle(cases.sim_args("rWW", sc, gp, rp, cfg, S, duration_s=20))
    assert np.abs(out["_sim_states"][2] - r_target).max() < 1e-4
    assert not out["_global_bools"].any()
    wie = rp[2].copy()   # FIC wrote wIE in place
    assert np.all(wie > 0)
    cfg0 = cases.model_config("rWW", do_fc=False, do_fcd=False, bold_remove_s=10, do_fic=False)
    for s, hit in [(sc, True), (sc.T.copy(), False)]:
        gp2, rp2 = cases.params("rWW", S, N, G=Gs, w_p=wp, J_N=jn, sigma=0.0, do_fic=False, wIE=1.0)
        rp2[2] = wie
        dev = np.abs(cases.run_oracle(cases.sim_args("rWW", s, gp2, rp2, cfg0, S, duration_s=20))
                     ["_sim_states"][2] - r_target).max(axis=1)
        assert dev[0] < 1e-4          # G = 0 does not see SC
        if hit:
            assert dev.max() < 1e-4
        else:
            assert dev[1:].min() > 1e-2


@pytest.mark.parametrize("model", ["rWW", "JR", "WC"])
def test_zero_coupling_ignores_sc(model):
    N, S = 6, 2
    gp, rp = cases.params(model, S, N, G=[0.0, 0.0])
    cfg = cases.model_config(model, window_size=4, window_step=2, bold_remove_s=2)
    a = cases.run_oracle(cases.sim_args(model, cases.synthetic_sc(N, seed=1), gp, rp.copy(), cfg, S, duration_s=8))
    b = cases.run_oracle(cases.sim_args(model, cases.synthetic_sc(N, seed=2), gp, rp.copy(), cfg, S, duration_s=8))
    np.testing.assert_array_equal(a["sim_bold"], b["sim_bold"])
    gp1, _ = cases.params(model, S, N, G=[0.5, 0.5])
    c = cases.run_oracle(cases.sim_args(model, cases.synthetic_sc(N, seed=2), gp1, rp.copy(), cfg, S, duration_s=8))
    assert np.abs(c["sim_bold"] - b["sim_bold"]).max() > 0


def _pearson(x, y):
    x, y = x - x.mean(), y - y.mean()
    return (x * y).sum() / np.sqrt((x * x).sum() * (y * y).sum())


@pytest.mark.parametrize("exc_inter", [False, True])
def test_small_fc_fcd_brute_force(exc_inter):
    N, S, TR, nvr_s, ws, step = 4, 2, 1.0, 3, 4, 3
    gp, rp = cases.params("WC", S, N, G=[0.4, 0.9])
    cfg = cases.model_config("WC", window_size=ws, window_step=step, bold_remove_s=nvr_s,
                             exc_interhemispheric=exc_inter)
    out = cases.run_oracle(cases.sim_args("WC", cases.synthetic_sc(N, seed=3), gp, rp, cfg, S, duration_s=14))
    bold = out["sim_bold"].reshape(S, 14, N)
    pairs = [(i, j) for i in range(N) for j in range(i) if not exc_inter or (i < N // 2) == (j < N // 2)]
    # windows centred every `step` volumes from ws/2 + nvr, each [c - ws/2, c + ws/2]
    centres = list(range(ws // 2 + nvr_s, 14 - 1 - ws // 2 + 1, step))
    for s in range(S):
        fc = [_pearson(bold[s, nvr_s:, i], bold[s, nvr_s:, j]) for i, j in pairs]
        # one-pass (oracle, as GSL) vs two-pass Pearson on 5-11 samples: ~1e-12 apart
        np.testing.assert_allclose(out["sim_fc_trils"][s], fc, rtol=0, atol=1e-10)
        wfc = np.array([[_pearson(bold[s, c - ws // 2:c + ws // 2 + 1, i], bold[s, c - ws // 2:c + ws // 2 + 1, j])
                         for i, j in pairs] for c in centres])
        fcd = [_pearson(wfc[a], wfc[b]) for a in range(len(centres)) for b in range(a)]
        np.testing.assert_allclose(out["sim_fcd_trils"][s], fcd, rtol=0, atol=1e-10)


def test_small_tables_hand_enumerated(oracle_lib):
    cfg = cases.model_config("rWW", window_size=4, window_step=3, bold_remove_s=2)
    d = oracle_lib.dims("rWW", cfg, 4, 12000, 1000, 1000, 0.1, 1.0)
    # 12 volumes, 2 removed -> 10 used; centres 4 and 7 (last allowed 9)
    assert (d["bold_len"], d["n_vols_remove"], d["corr_len"], d["n_pairs"]) == (12, 2, 10, 6)
    assert list(d["window_starts"]) == [2, 5] and list(d["window_ends"]) == [6, 9]
    assert d["n_windows"] == 2 and d["n_window_pairs"] == 1
    cfg["exc_interhemispheric"] = "1"
    d = oracle_lib.dims("rWW", cfg, 4, 12000, 1000, 1000, 0.1, 1.0)
    assert d["n_pairs"] == 2   # (1,0) and (3,2)
