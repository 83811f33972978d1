"""SimGroup-compatible layer (cubnm_amd.sim) on the CPU.

``run_simulations`` is swapped for the oracle here (test infrastructure), so
the layer's marshalling, output shaping, noise reconstruction, scoring and
multi-group merge are checked without a GPU:
  * run() builds exactly the argument tuple of the reference's
    SimGroup.run (sim/base.py:705-750) for the C1 grid, and the reinit rule;
  * get_noise() equals the reference's own test_get_noise reconstruction
    loop (tests/sim/test.py:188-258) on the same settings;
  * outputs have the reference's shapes (sim/base.py:805-818) and rWW FIC
    writes wIE back into param_lists (rww.py:170-178);
  * score(): -fc_diff / -fc_normec / +gof arithmetic, SciPy terms, NaN ->
    worst score, FIC penalty (rww.py:180-234);
  * MultiSimGroup merges duplicate SCs and splits outputs so each child sees
    what it would see alone;
  * save() writes the outputs and parameter lists (base.py:1209-1239) and
    the file reads back with numpy's non-pickle loader.
"""
from decimal import Decimal

import numpy as np
import pytest
import scipy.stats

import cases


@pytest.fixture
def oracle_core(monkeypatch):
    from cubnm_amd import _core
    calls = []

    def fake(*args):
        calls.append(args)
        return cases.run_oracle(args)

    monkeypatch.setattr(_core, "run_simulations", fake)
    return calls


def _c1_group(**kw):
    from cubnm_amd.sim import rWWSimGroup
    sg = rWWSimGroup(duration=60, TR=1, sc=cases.synthetic_sc(100), window_size=10, window_step=2,
                     do_fic=False, **kw)
    G, wp = np.meshgrid(np.linspace(0.5, 3, 4), [0.0, 2.0], indexing="ij")
    sg.N = 8
    sg.param_lists["G"] = G.ravel()
    sg.param_lists["w_p"] = np.repeat(wp.ravel()[:, None], 100, axis=1)
    sg._set_default_params()
    return sg, G.ravel(), wp.ravel()


def test_run_arguments_match_reference_marshalling():
    sg, G, wp = _c1_group()
    args, _ = sg._run_args()
    gp, rp = cases.params("rWW", 8, 100, G=G, w_p=np.repeat(wp[:, None], 100, axis=1), do_fic=False, wIE=1.0)
    cfg = cases.model_config("rWW", do_fic=False)
    ref = cases.sim_args("rWW", cases.synthetic_sc(100), gp, rp, cfg, 8, duration_s=60, force_reinit=True)
    assert len(args) == len(ref) == 22
    for a, b in zip(args, ref):
        if isinstance(a, np.ndarray):
            assert a.dtype == b.dtype and a.shape == b.shape
            np.testing.assert_array_equal(a, b)
        else:
            assert a == b and type(a) is type(b), (a, b)


def test_reinit_rule(oracle_core):
    from cubnm_amd.sim import rWWSimGroup
    sg = rWWSimGroup(duration=4, TR=1, sc=cases.synthetic_sc(6), window_size=2, window_step=1,
                     bold_remove_s=1, do_fic=False)
    sg.N = 2
    sg.param_lists["G"] = np.array([0.5, 1.0])
    sg._set_default_params()
    sg.run()
    sg.run()
    sg.TR = 0.5
    sg.run()
    sg.N = 3
    sg.param_lists["G"] = np.array([0.5, 1.0, 1.5])
    sg._set_default_params()
    sg.run()
    assert [c[12] for c in oracle_core] == [True, False, True, True]   # force_reinit flag
    assert sg.it == 4


def test_get_noise_matches_core_loop(oracle_core):
    from cubnm_amd.sim import rWWSimGroup
    nodes = 10
    sc = np.ones((nodes, nodes))
    np.fill_diagonal(sc, 0)
    sg = rWWSimGroup(duration=8, noise_segment_length=3, TR=0.1, dt="0.2", bw_dt="0.8", sc=sc,
                     do_fc=False, do_fcd=False, gof_terms=[], noise_out=True)
    sg.N = 1
    sg.param_lists["G"] = np.array([0.5])
    sg._set_default_params(missing=True)
    sg.run()
    recon = sg.get_noise()
    assert recon.shape == (2, nodes, sg.bw_it, sg.inner_it) == (2, 10, 10000, 4)
    bw_dt_s = sg.bw_dt / 1000
    noise_bw_it = int((Decimal(sg.noise_segment_length * 1000) / 1000) / bw_dt_s)
    repeats = int(np.ceil(sg.bw_it / noise_bw_it))
    assert repeats == sg._shuffled_nodes.shape[0] == 3
    sts, snd = sg._shuffled_ts.ravel(), sg._shuffled_nodes.ravel()
    mimic = np.empty_like(recon)
    r = 0
    for bw in range(sg.bw_it):
        ts = sts[bw % noise_bw_it + r * noise_bw_it]
        idx = ((ts * sg.inner_it + np.arange(sg.inner_it)[None, :]) * nodes * sg.n_noise
               + snd[r * nodes:(r + 1) * nodes][:, None] * sg.n_noise)
        mimic[0, :, bw, :] = sg._noise[idx]
        mimic[1, :, bw, :] = sg._noise[idx + 1]
        if (bw + 1) % noise_bw_it == 0 and bw + 1 < sg.bw_it:
            r += 1
    np.testing.assert_array_equal(recon, mimic)


def test_outputs_shapes_and_fic_writeback(oracle_core):
    from cubnm_amd.sim import rWWSimGroup
    sg = rWWSimGroup(duration=12, TR=1, sc=cases.synthetic_sc(10), window_size=4, window_step=2,
                     bold_remove_s=2, states_ts=True)
    sg.N = 3
    assert np.all(sg.param_lists["wIE"] == 0)
    sg.param_lists["G"] = np.array([0.5, 1.0, 2.0])
    sg._set_default_params()
    sg.run()
    assert sg.sim_bold.shape == (3, 12, 10)
    assert sg.sim_fc_trils.shape == (3, 45)
    n_w = len(range(2 + 2, 12 - 1 - 2 + 1, 2))
    assert sg.sim_fcd_trils.shape == (3, n_w * (n_w - 1) // 2)
    assert set(sg.sim_states) == set(sg.state_names)
    assert sg.sim_states["r_E"].shape == (3, 12, 10)
    assert sg.fic_unstable.shape == (3,) and sg.fic_ntrials.shape == (3,)
    assert np.all(sg.param_lists["wIE"] > 0) and sg.param_lists["wIE"].shape == (3, 10)
    fc = sg.get_sim_fc(1)
    assert fc.shape == (10, 10) and np.allclose(fc, fc.T) and np.all(np.diag(fc) == 1)
    assert fc[3, 1] == sg.sim_fc_trils[1][np.where((np.tril_indices(10, -1)[0] == 3) & (np.tril_indices(10, -1)[1] == 1))[0][0]]
    fcd = sg.get_sim_fcd(2)
    assert fcd.shape == (n_w, n_w)
    avg = sg.get_state_averages()
    np.testing.assert_allclose(avg["S_E"], sg.sim_states["S_E"][:, 2:, :].mean(axis=(1, 2)))
    one = sg.slice(1)
    assert one.N == 1 and one.sim_bold.shape == (1, 12, 10)
    sg.clear()
    assert not hasattr(sg, "sim_bold")


def test_score_terms(oracle_core):
    from cubnm_amd.sim import rWWSimGroup, WORST_SCORES
    sg = rWWSimGroup(duration=12, TR=1, sc=cases.synthetic_sc(10), window_size=4, window_step=2,
                     bold_remove_s=2, gof_terms=["+fc_corr", "-fcd_ks", "-fc_diff", "-fc_normec"])
    sg.N = 3
    sg.param_lists["G"] = np.array([0.5, 1.0, 2.0])
    sg._set_default_params()
    sg.run()
    rng = np.random.default_rng(0)
    emp_fc = np.tanh(rng.standard_normal(45))
    emp_fcd = rng.uniform(-0.5, 1, 20)
    sg.sim_fc_trils[2, 5] = np.nan
    s = sg.score(emp_fc, emp_fcd, force_cpu=True)
    for i in range(2):
        assert s.loc[i, "+fc_corr"] == pytest.approx(scipy.stats.pearsonr(sg.sim_fc_trils[i], emp_fc)[0], abs=1e-15)
        assert s.loc[i, "-fcd_ks"] == -scipy.stats.ks_2samp(sg.sim_fcd_trils[i], emp_fcd).statistic
        assert s.loc[i, "-fc_diff"] == pytest.approx(-abs(sg.sim_fc_trils[i].mean() - emp_fc.mean()), abs=1e-15)
        assert s.loc[i, "-fc_normec"] == pytest.approx(
            -np.sqrt(((sg.sim_fc_trils[i] - emp_fc) ** 2).sum()) / (2 * np.sqrt(45)), abs=1e-15)
    assert s.loc[2, "+fc_corr"] == WORST_SCORES["+fc_corr"]
    assert s.loc[2, "-fc_diff"] == WORST_SCORES["-fc_diff"]
    np.testing.assert_allclose(s["+gof"], s[["+fc_corr", "-fcd_ks", "-fc_diff", "-fc_normec"]].sum(axis=1))
    # FIC penalty: only deviations beyond 1 Hz from 3 Hz count
    r = np.full((3, 10), 3.2)
    r[1, :3] = [5.0, 0.5, 3.9]
    sg.sim_states["r_E"] = r
    s = sg.score(emp_fc, emp_fcd, force_cpu=True)
    assert s.loc[0, "-fic_penalty"] == 0.0
    want = -((1 - np.exp(-0.05 * 1.0)) + (1 - np.exp(-0.05 * 1.5))) * 0.5 / 10
    assert s.loc[1, "-fic_penalty"] == pytest.approx(want, rel=1e-14)


def test_multi_sim_group_split(oracle_core):
    from cubnm_amd.sim import MultiSimGroup, WCSimGroup
    kw = dict(duration=8, TR=1, window_size=4, window_step=2, bold_remove_s=2)
    sc_a, sc_b = cases.synthetic_sc(6, seed=1), cases.synthetic_sc(6, seed=2)
    groups = []
    for sc, G in [(sc_a, [0.3, 0.6]), (sc_b, [0.9]), (sc_a, [1.2, 1.5, 1.8])]:
        g = WCSimGroup(sc=sc, **kw)
        g.N = len(G)
        g.param_lists["G"] = np.array(G)
        g._set_default_params()
        groups.append(g)
    multi = MultiSimGroup(groups)
    multi.run()
    merged = oracle_core[-1]
    assert merged[1].shape == (2, 36) and list(merged[2]) == [0, 0, 1, 0, 0, 0]
    for g in groups:
        bold = g.sim_bold.copy()
        fc = g.sim_fc_trils.copy()
        g.run()
        np.testing.assert_array_equal(g.sim_bold, bold)
        np.testing.assert_array_equal(g.sim_fc_trils, fc)
    # the class factory of the reference's batch optimisation
    from cubnm_amd.sim import JRSimGroup, create_multi_sim_group
    MultiWC = create_multi_sim_group(WCSimGroup)
    assert MultiWC.__name__ == "MultiWCSimGroup" and issubclass(MultiWC, MultiSimGroup)
    m2 = MultiWC(groups[:2])
    m2.run()
    assert oracle_core[-1][14] == 3                               # N_SIMS of the merged call
    with pytest.raises(TypeError):
        MultiWC([JRSimGroup(sc=sc_a, **kw)])
    with pytest.raises(TypeError):
        create_multi_sim_group(int)


def test_avail_gpus_counts_gpu_topology_nodes(tmp_path, monkeypatch):
    """utils.avail_gpus counts kernel-driver topology nodes with SIMDs: a CPU
    node (simd_count 0) and an unreadable entry are not GPUs."""
    from cubnm_amd import utils
    for i, simds in enumerate([0, 1024, 1024]):
        d = tmp_path / str(i)
        d.mkdir()
        (d / "properties").write_text(f"cpu_cores_count {0 if simds else 64}\nsimd_count {simds}\nmax_waves_per_simd 8\n")
    (tmp_path / "3").mkdir()
    (tmp_path / "3" / "properties").write_text("simd_count x\n")
    monkeypatch.setattr(utils, "_KFD_TOPOLOGY", str(tmp_path))
    assert utils.avail_gpus() == 2
    monkeypatch.setattr(utils, "_KFD_TOPOLOGY", str(tmp_path / "absent"))
    assert utils.avail_gpus() == 0
    assert utils.is_jupyter() is False


def test_save_writes_outputs_and_parameters(oracle_core, tmp_path):
    from cubnm_amd.sim import rWWSimGroup
    sg = rWWSimGroup(duration=12, TR=1, sc=cases.synthetic_sc(10), window_size=4, window_step=2,
                     bold_remove_s=2, out_dir=str(tmp_path / "sims"))
    sg.N = 2
    sg.param_lists["G"] = np.array([0.5, 1.5])
    sg._set_default_params()
    sg.run()
    sg.save()
    with np.load(tmp_path / "sims" / "sim_data.npz") as d:
        expected = {"sim_bold", "sim_fc_trils", "sim_fcd_trils", *sg.state_names,
                    *[k for k, v in sg.param_lists.items() if v is not None]}
        assert set(d.files) == expected
        np.testing.assert_array_equal(d["sim_bold"], sg.sim_bold)
        np.testing.assert_array_equal(d["sim_fcd_trils"], sg.sim_fcd_trils)
        np.testing.assert_array_equal(d["wIE"], sg.param_lists["wIE"])     # after the FIC write-back
        np.testing.assert_array_equal(d["S_E"], sg.sim_states["S_E"])
    sg.out_dir = None
    with pytest.raises(ValueError, match="out_dir"):
        sg.save()
