"""Optimisation layer (cubnm_amd.optimize) on the CPU.

The parameter mapping of BNMProblem is host arithmetic, checked against
what it must satisfy rather than against a restated formula:
  * the unit cube maps linearly onto the bounds and back (round trip);
  * fixed parameters are constant, free global and bias parameters follow
    their column, group parameters are shared exactly by their group's
    members ('sym': node i and node i + N/2; 'node': one per node; a fixed
    group value stays fixed);
  * map coefficients: the automatic range keeps every node's value
    non-negative at both ends for a z-scored map and is a multiple of 0.1;
    (0, 1) for a positive map; the regional value is bias x (1 + sum of
    coefficient x map), which at zero coefficients is the bias;
  * range normalisation maps the values into the range with an affine,
    order-preserving map, leaves them unchanged when inside, and keeps an
    in-range end where it is;
  * the reference's argument errors.
With run_simulations swapped for the oracle (test infrastructure), the grid
and CMA-ES drivers evaluate the problem end to end: costs are -(+gof), the
FIC penalty enters the cost, and the grid's rows equal a direct SimGroup run.
batch_optimize (one simulation call per generation for several optimisers)
gives exactly the histories of the same optimisers run alone, on one problem
and on two problems with different SCs, and raises the reference's argument
errors.  save() writes the reference's run-directory layout.  NSGA-II: its
non-dominated sorting against the definition, crowding distance on a
hand-worked front, convergence to the Pareto set of ZDT1, and the
multi-objective BNMProblem path.
"""
import os

import numpy as np
import pytest

import cases

N = 6


@pytest.fixture
def oracle_core(monkeypatch):
    """Simulations on the oracle and the two GPU GOF reductions as SciPy."""
    import scipy.stats
    from cubnm_amd import _core, utils
    monkeypatch.setattr(_core, "run_simulations", lambda *a: cases.run_oracle(a))
    monkeypatch.setattr(utils, "fc_corr_device",
                        lambda sim, emp, usable_mem=None: np.array([scipy.stats.pearsonr(x, emp)[0] for x in sim]))
    monkeypatch.setattr(utils, "fcd_ks_device",
                        lambda sim, emp, usable_mem=None: np.array([scipy.stats.ks_2samp(x, emp).statistic
                                                                    for x in sim]))


def _problem(params, **kw):
    from cubnm_amd.optimize import BNMProblem
    base = dict(model="rWW", params=params, duration=12, TR=1, sc=cases.synthetic_sc(N, seed=1), window_size=4,
                window_step=2, bold_remove_s=2, do_fic=False)
    base.update(kw)
    if "emp_fc_tril" not in kw and "emp_bold" not in kw:
        rng = np.random.default_rng(0)
        base["emp_fc_tril"] = np.tanh(rng.standard_normal(N * (N - 1) // 2))
        base["emp_fcd_tril"] = np.tanh(rng.standard_normal(6))
    return BNMProblem(**base)


def test_homogeneous_mapping_and_round_trip():
    p = _problem({"G": (0.5, 3.0), "w_p": (0.0, 2.0), "J_N": 0.15})
    assert p.free_params == ["G", "w_p"] and set(p.fixed_params) == {"J_N"} and p.ndim == 2
    X = np.array([[0.0, 0.0], [1.0, 1.0], [0.5, 0.25]])
    np.testing.assert_allclose(p._get_X(p._get_Xt(X)), X, atol=1e-15)
    sp = p._get_sim_params(X)
    np.testing.assert_allclose(sp["G"], [0.5, 3.0, 1.75])
    assert sp["w_p"].shape == (3, N)
    np.testing.assert_allclose(sp["w_p"], np.repeat([[0.0], [2.0], [0.5]], N, axis=1))
    np.testing.assert_array_equal(sp["J_N"], np.full((3, N), 0.15))
    assert p.xl.tolist() == [0, 0] and p.xu.tolist() == [1, 1] and p.n_obj == 1


def test_node_groups_sym_node_and_fixed_group():
    p = _problem({"G": (0.5, 3.0), "w_p": (0.0, 2.0)}, het_params=["w_p"], node_grouping="sym")
    assert p.free_params == ["G"] + [f"w_p{g}" for g in range(N // 2)]
    X = np.random.default_rng(1).uniform(0, 1, (4, p.ndim))
    wp = p._get_sim_params(X)["w_p"]
    np.testing.assert_array_equal(wp[:, :N // 2], wp[:, N // 2:])
    np.testing.assert_allclose(wp[:, :N // 2], 2.0 * X[:, 1:])
    q = _problem({"G": (0.5, 3.0), "w_p": (0.0, 2.0), "w_p3": 0.7}, het_params=["w_p"], node_grouping="node")
    assert "w_p3" not in q.free_params and len(q.free_params) == 1 + N - 1
    wq = q._get_sim_params(np.full((2, q.ndim), 0.5))["w_p"]
    np.testing.assert_array_equal(wq[:, 3], [0.7, 0.7])
    np.testing.assert_allclose(np.delete(wq, 3, axis=1), 1.0)
    groups = np.array([0, 0, 1, 1, 1, 2])
    r = _problem({"G": (0.5, 3.0), "J_N": (0.1, 0.3)}, het_params=["J_N"], node_grouping=groups)
    jn = r._get_sim_params(np.array([[0.0, 0.0, 0.5, 1.0]]))["J_N"][0]
    np.testing.assert_allclose(jn, [0.1, 0.1, 0.2, 0.2, 0.2, 0.3])


def test_map_coefficients_and_ranges():
    rng = np.random.default_rng(2)
    z = rng.standard_normal(N)
    z = (z - z.mean()) / z.std()
    pos = rng.uniform(0.1, 1.0, N)
    maps = np.vstack([z, pos])
    p = _problem({"G": (0.5, 3.0), "w_p": (0.5, 2.0)}, het_params=["w_p"], maps=maps, het_params_range=None)
    assert p.free_params == ["G", "w_p", "w_pscale0", "w_pscale1"]
    lo, hi = p.lb[2], p.ub[2]
    assert lo < 0 < hi
    for v in (lo, hi):
        assert abs(round(v * 10) - v * 10) < 1e-9
        assert np.all(1 + v * z >= -1e-12)            # the bias is never flipped in sign
    assert (p.lb[3], p.ub[3]) == (0.0, 1.0)
    # zero coefficients: every node at the bias
    Xz = np.array([[0.5, 0.5, p._get_X([0, 0, 0, 0])[2], 0.0]])
    np.testing.assert_allclose(p._get_sim_params(Xz)["w_p"], np.full((1, N), 1.25))
    # bias x (1 + c0 z + c1 pos)
    X = np.array([[0.5, 1.0, 1.0, 1.0]])
    np.testing.assert_allclose(p._get_sim_params(X)["w_p"][0], 2.0 * (1 + hi * z + 1.0 * pos))


def test_range_normalisation():
    z = np.array([-2.0, -1.0, 0.0, 0.5, 1.0, 1.5])
    p = _problem({"G": (0.5, 3.0), "w_p": (0.0, 2.0)}, het_params=["w_p"], maps=z[None, :],
                 maps_coef_range=(-1.0, 1.0))
    raw = _problem({"G": (0.5, 3.0), "w_p": (0.0, 2.0)}, het_params=["w_p"], maps=z[None, :],
                   maps_coef_range=(-1.0, 1.0), het_params_range=None)
    for X in ([0.5, 1.0, 1.0], [0.5, 0.5, 0.6], [0.5, 0.25, 0.5]):
        X = np.array([X])
        v, r = p._get_sim_params(X)["w_p"][0], raw._get_sim_params(X)["w_p"][0]
        assert v.min() >= -1e-12 and v.max() <= 2.0 + 1e-12
        if r.min() >= 0 and r.max() <= 2:
            np.testing.assert_allclose(v, r)          # inside the range: unchanged
        else:
            np.testing.assert_array_equal(np.argsort(v, kind="stable"), np.argsort(r, kind="stable"))
            a = np.polyfit(r, v, 1)
            np.testing.assert_allclose(np.polyval(a, r), v, atol=1e-12)   # affine
            if 0 <= r.max() <= 2:
                assert abs(v.max() - r.max()) < 1e-12   # the in-range end stays
            if 0 <= r.min() <= 2:
                assert abs(v.min() - r.min()) < 1e-12


def test_argument_errors():
    with pytest.raises(ValueError, match="node_grouping"):
        _problem({"G": (0.5, 3.0), "w_p": (0, 2)}, het_params=["w_p"], maps=np.ones((1, N)), node_grouping="node")
    with pytest.raises(ValueError, match="wIE"):
        _problem({"G": (0.5, 3.0), "wIE": (0.5, 2)}, het_params=["wIE"], node_grouping="node", do_fic=True)
    with pytest.raises(ValueError, match="cannot be fixed"):
        _problem({"G": (0.5, 3.0), "w_p": 1.0}, het_params=["w_p"], maps=np.ones((1, N)))
    with pytest.raises(ValueError, match="shape"):
        _problem({"G": (0.5, 3.0), "w_p": (0, 2)}, het_params=["w_p"], node_grouping=np.zeros(N + 1))


def test_multiobjective_names():
    p = _problem({"G": (0.5, 3.0)}, multiobj=True, do_fic=True)
    assert p.obj_names == ["-fc_corr", "+fcd_ks", "-fic_penalty"] and p.n_obj == 3


def test_grid_optimizer_end_to_end(oracle_core):
    from cubnm_amd.optimize import GridOptimizer
    p = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5), "J_N": 0.15})
    go = GridOptimizer()
    opt = go.optimize(p, grid_shape={"G": 2, "w_p": 2})
    h = go.history
    assert len(h) == 4 and list(h.columns[:2]) == ["G", "w_p"]
    np.testing.assert_allclose(h["G"], [0.5, 0.5, 3.0, 3.0])
    np.testing.assert_allclose(h["w_p"], [0.5, 1.5, 0.5, 1.5])
    np.testing.assert_allclose(h["cost"], -h["+gof"])
    assert opt["cost"] == h["cost"].min()
    # the same simulations through the SimGroup directly
    sg = p.sim_group
    sg.N = 4
    sg.param_lists["G"] = h["G"].values
    sg.param_lists["w_p"] = np.repeat(h["w_p"].values[:, None], N, axis=1)
    sg.param_lists["J_N"] = np.full((4, N), 0.15)
    sg._set_default_params()
    sg.run()
    s = sg.score(p.emp_fc_tril, p.emp_fcd_tril)
    np.testing.assert_allclose(s["+gof"].values, h["+gof"].values, rtol=0, atol=1e-12)


def test_fic_penalty_enters_the_cost(oracle_core):
    p = _problem({"G": (0.5, 3.0)}, do_fic=True)
    F, scores = p.evaluate(np.array([[0.2], [0.8]]))
    assert "-fic_penalty" in scores.columns
    np.testing.assert_allclose(F, -scores["+gof"].values - scores["-fic_penalty"].values)


def test_cmaes_optimizer_end_to_end(oracle_core):
    from cubnm_amd.optimize import CMAESOptimizer
    p = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5)})
    cm = CMAESOptimizer(popsize=4, n_iter=2, seed=3)
    cm.setup_problem(p)
    opt = cm.optimize()
    h = cm.history
    assert len(h) == 12 and sorted(h["gen"].unique().tolist()) == [0, 1, 2]
    assert ((h["G"] >= 0.5) & (h["G"] <= 3.0) & (h["w_p"] >= 0.5) & (h["w_p"] <= 1.5)).all()
    assert opt["cost"] == h["cost"].min() and len(cm.opt_history) == 3


def _history_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.values, b.values)


def test_batch_optimize_matches_independent_runs(oracle_core):
    """Reference tests/optimize/test.py:119-155: four optimisers (two seeds,
    each twice) on one problem in one batch; equal seeds give equal
    histories, different seeds different ones, and each equals the same
    optimiser run alone."""
    from cubnm_amd.optimize import CMAESOptimizer, batch_optimize
    p = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5)})
    o1, o2 = CMAESOptimizer(popsize=4, n_iter=2, seed=1), CMAESOptimizer(popsize=4, n_iter=2, seed=2)
    b = batch_optimize([o1, o1, o2, o2], p, save=False)
    assert len(b) == 4 and all(o.is_fit for o in b)
    assert _history_equal(b[0].history, b[1].history) and _history_equal(b[2].history, b[3].history)
    assert not np.allclose(b[0].history[["G", "w_p"]].values, b[2].history[["G", "w_p"]].values)
    import copy
    for o, k in ((o1, 0), (o2, 2)):
        o.setup_problem(copy.deepcopy(p))
        o.optimize()
        assert _history_equal(b[k].history, o.history)
        assert b[k].opt["cost"] == o.opt["cost"]


def test_batch_optimize_variable_sc(oracle_core):
    """Reference tests/optimize/test.py:157-193: one optimiser on two
    problems that differ in SC (the second reversed) runs as one call with
    two SCs; each history equals the independent run on its problem."""
    from cubnm_amd.optimize import CMAESOptimizer, batch_optimize
    p1 = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5)})
    p2 = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5)}, sc=cases.synthetic_sc(N, seed=1)[::-1, ::-1].copy())
    o = CMAESOptimizer(popsize=4, n_iter=2, seed=1)
    b = batch_optimize(o, [p1, p2], save=False)
    assert not np.allclose(b[0].history["cost"].values, b[1].history["cost"].values)
    for k, p in enumerate((p1, p2)):
        alone = CMAESOptimizer(popsize=4, n_iter=2, seed=1)
        alone.setup_problem(p)
        alone.optimize()
        assert _history_equal(b[k].history, alone.history)


def test_batch_optimize_argument_errors():
    from cubnm_amd.optimize import BNMProblem, CMAESOptimizer, GridOptimizer, batch_optimize
    p = _problem({"G": (0.5, 3.0)})
    o = CMAESOptimizer(popsize=4, n_iter=1)
    with pytest.raises(ValueError, match="must be a list"):
        batch_optimize(o, p)
    with pytest.raises(ValueError, match="Same number"):
        batch_optimize([o, o], [p, p, p])
    with pytest.raises(TypeError, match="population-based"):
        batch_optimize([GridOptimizer()], [p])
    with pytest.raises(ValueError, match="same type of model"):
        batch_optimize(o, [p, BNMProblem(model="JR", params={"G": (0.5, 3.0)}, duration=12, TR=1,
                                         sc=cases.synthetic_sc(N, seed=1), window_size=4, window_step=2,
                                         bold_remove_s=2, emp_fc_tril=p.emp_fc_tril, emp_fcd_tril=p.emp_fcd_tril)])
    with pytest.raises(ValueError, match="out_dir"):
        batch_optimize([o], p, save=True)


def test_optimizer_save_layout(oracle_core, tmp_path):
    """Optimizer.save (optimize.py:636-731, 1235-1254): run directories
    numbered in turn, the history tables, problem.json with arrays moved to
    .txt files, optimizer.json, and the optimum's simulation in opt_sim/
    (sliced out of the grid run; rerun for CMA-ES, whose scores.csv row
    equals the best history row)."""
    import json
    import pandas as pd
    from cubnm_amd.optimize import CMAESOptimizer, GridOptimizer, batch_optimize
    p = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5)}, out_dir=str(tmp_path))
    go = GridOptimizer()
    go.optimize(p, grid_shape=2)
    d = go.save()
    assert d == str(tmp_path / "grid_run-0")
    assert {"history.csv", "opt.csv", "problem.json", "optimizer.json", "opt_sim", "state_averages.csv",
            "sc.txt"} <= set(os.listdir(d))
    assert "opt_history.csv" not in os.listdir(d)
    cfg = json.load(open(os.path.join(d, "problem.json")))
    np.testing.assert_array_equal(np.loadtxt(cfg["sc"]), p.sim_group.sc)
    assert json.load(open(os.path.join(d, "optimizer.json"))) == {"grid_shape": {"G": 2, "w_p": 2}}
    with np.load(os.path.join(d, "opt_sim", "sim_data.npz")) as z:
        k = int(go.opt.name)
        np.testing.assert_array_equal(z["G"], [go.history["G"][k]])
        assert z["sim_bold"].shape[0] == 1
    assert go.save(save_opt=False) == str(tmp_path / "grid_run-1")
    cm = CMAESOptimizer(popsize=4, n_iter=1, seed=2)
    cm.setup_problem(p)
    cm.optimize()
    d = cm.save()
    assert d == str(tmp_path / "cmaes_run-0") and "opt_history.csv" in os.listdir(d)
    assert json.load(open(os.path.join(d, "optimizer.json"))) == {"n_iter": 1, "popsize": 4, "seed": 2}
    sc = pd.read_csv(os.path.join(d, "opt_sim", "scores.csv"))
    np.testing.assert_allclose(sc["+gof"].values, [cm.opt["+gof"]], rtol=0, atol=1e-12)
    with pytest.raises(RuntimeError, match="fitted"):
        CMAESOptimizer(popsize=4).save()
    b = batch_optimize(cm, [p, p])
    assert os.path.isdir(tmp_path / "cmaes_run-1") and os.path.isdir(tmp_path / "cmaes_run-2") and len(b) == 2


def test_nondominated_ranks_and_crowding():
    """Fast non-dominated sorting against the definition (a row's front is one
    more than the highest front among the rows dominating it, 0 if none),
    on random sets with ties; crowding distance on a hand-worked front."""
    from cubnm_amd.optimize import _crowding_distance, _nondominated_ranks
    rng = np.random.default_rng(5)
    for n, m in ((1, 2), (7, 2), (30, 2), (25, 3)):
        F = rng.integers(0, 5, (n, m)).astype(float)          # many ties and duplicates
        rank = _nondominated_ranks(F)
        for i in range(n):
            dominators = [j for j in range(n) if np.all(F[j] <= F[i]) and np.any(F[j] < F[i])]
            assert rank[i] == (1 + max(rank[j] for j in dominators) if dominators else 0)
    d = _crowding_distance(np.array([[0.0, 4.0], [1.0, 2.0], [2.0, 1.0], [4.0, 0.0]]))
    assert np.isinf(d[0]) and np.isinf(d[3])
    np.testing.assert_allclose(d[1:3], [2 / 4 + 3 / 4, 3 / 4 + 2 / 4])


class _Zdt1:
    """Two-objective test problem (Zitzler, Deb and Thiele 2000, ZDT1 with two
    variables): f1 = x0, f2 = g (1 - sqrt(x0 / g)), g = 1 + 9 x1; the Pareto
    front is x1 = 0."""
    n_obj, ndim, multiobj = 2, 2, True
    obj_names = ["f1", "f2"]
    free_params = ["x0", "x1"]

    def _get_Xt(self, X):
        return np.asarray(X)

    def evaluate(self, X):
        g = 1 + 9 * X[:, 1]
        return np.stack([X[:, 0], g * (1 - np.sqrt(X[:, 0] / g))], axis=1), None


def test_nsga2_converges_on_zdt1():
    from cubnm_amd.optimize import NSGA2Optimizer, _nondominated_ranks
    o = NSGA2Optimizer(popsize=24, n_iter=40, seed=3)
    o.setup_problem(_Zdt1())
    assert o.optimize() is None                               # several objectives: no single optimum
    h = o.history
    assert len(h) == 24 * 41 and list(h.columns[1:]) == ["x0", "x1", "f1", "f2", "gen"]
    assert ((h[["x0", "x1"]] >= 0) & (h[["x0", "x1"]] <= 1)).all().all()
    assert h.loc[h["gen"] == 0, "x1"].mean() > 0.3
    front = o._pop_X[_nondominated_ranks(o._pop_F) == 0]
    assert len(front) >= 12 and front[:, 1].max() < 0.05      # the population reached the Pareto set
    assert np.ptp(front[:, 0]) > 0.5                           # and spread along it
    with pytest.raises(ValueError, match="at most 3"):
        p4 = _Zdt1()
        p4.n_obj = 4
        o.setup_problem(p4)


def test_nsga2_multiobjective_problem_and_batch(oracle_core, tmp_path):
    """NSGA-II on a multi-objective BNMProblem: objective columns are the
    problem's obj_names (-(+fc_corr) and -(-fcd_ks)); a batch of two seeds equals
    the seeds run alone; save reruns the final non-dominated set."""
    import copy
    from cubnm_amd.optimize import NSGA2Optimizer, batch_optimize, _nondominated_ranks
    p = _problem({"G": (0.5, 3.0), "w_p": (0.5, 1.5)}, multiobj=True, out_dir=str(tmp_path))
    o1, o2 = NSGA2Optimizer(popsize=4, n_iter=1, seed=1), NSGA2Optimizer(popsize=4, n_iter=1, seed=2)
    b = batch_optimize([o1, o2], p, save=False)
    for o, k in ((o1, 0), (o2, 1)):
        o.setup_problem(copy.deepcopy(p))
        o.optimize()
        assert _history_equal(b[k].history, o.history)
    assert list(o1.history.columns[3:5]) == p.obj_names
    d = o1.save()
    with np.load(os.path.join(d, "opt_sim", "sim_data.npz")) as z:
        assert z["sim_bold"].shape[0] == int((_nondominated_ranks(o1._pop_F) == 0).sum())


def test_population_optimizer_arguments_and_history_sims(oracle_core):
    """Constructor arguments in the reference's order and meaning
    (optimize.py:1267-1330, 1386-1410, 1487-1504); without save_history_sim
    each generation's simulation outputs are cleared once scored."""
    from cubnm_amd.optimize import CMAESOptimizer, NSGA2Optimizer
    o = CMAESOptimizer(4, None, 0.3, n_iter=1, seed=1)
    assert (o.popsize, o.x0, o.sigma, o.n_iter, o.seed, o.save_history_sim) == (4, None, 0.3, 1, 1, False)
    for bad in (dict(use_bound_penalty=True), dict(algorithm_kws={"CMA_elitist": True}), dict(termination=object())):
        with pytest.raises(NotImplementedError):
            CMAESOptimizer(4, **bad)
    n = NSGA2Optimizer(6, {"eta_c": 5.0})
    assert (n.eta_c, n.prob_c, n.eta_m) == (5.0, 0.9, 20.0)
    with pytest.raises(NotImplementedError, match="mutation"):
        NSGA2Optimizer(6, {"mutation": None})
    p = _problem({"G": (0.5, 3.0)})
    o.setup_problem(p)
    o.optimize()
    assert not hasattr(p.sim_group, "sim_bold")
    q = _problem({"G": (0.5, 3.0)})
    k = CMAESOptimizer(4, None, 0.3, n_iter=1, seed=1, save_history_sim=True)
    k.setup_problem(q)
    k.optimize()
    assert q.sim_group.sim_bold.shape[0] == 4
    np.testing.assert_array_equal(o.history.values, k.history.values)
