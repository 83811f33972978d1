"""Optimisation drivers on the GPU core (cubnm_amd.optimize).

* GridOptimizer over G x w_p (2 x 2, rWW, 10 nodes, 12 s): its history rows
  equal a direct SimGroup run of the same four parameter sets (same core,
  same noise: identical scores) and its costs are -(+gof).
* CMAESOptimizer (popsize 4, 2 generations after the uniform first one)
  keeps every candidate inside the bounds and reports the best row.
* batch_optimize of one CMA-ES over two problems with different SCs: one
  simulation call per generation, histories equal to independent runs.
* NSGA-II on the two GOF objectives: its recorded objectives equal a direct
  evaluation of the same candidates.
The arithmetic of both against the oracle is in tests/test_optimize.py.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

N = 10


def _problem(**kw):
    import torch  # noqa: F401
    from cubnm_amd.optimize import BNMProblem
    rng = np.random.default_rng(0)
    sc = kw.pop("sc", cases.synthetic_sc(N, seed=1))
    return BNMProblem(model="rWW", params={"G": (0.5, 3.0), "w_p": (0.5, 1.5), "J_N": 0.15}, duration=12, TR=1,
                      sc=sc, window_size=4, window_step=2, bold_remove_s=2,
                      do_fic=False, emp_fc_tril=np.tanh(rng.standard_normal(N * (N - 1) // 2)),
                      emp_fcd_tril=np.tanh(rng.standard_normal(6)), **kw)


def test_grid_optimizer_matches_direct_run():
    from cubnm_amd.optimize import GridOptimizer
    p = _problem()
    go = GridOptimizer()
    go.optimize(p, grid_shape=2)
    h = go.history
    np.testing.assert_allclose(h["cost"], -h["+gof"])
    sg = p.sim_group
    sg.N = 4
    sg.param_lists["G"] = h["G"].values
    sg.param_lists["w_p"] = np.repeat(h["w_p"].values[:, None], N, axis=1)
    sg.param_lists["J_N"] = np.full((4, N), 0.15)
    sg._set_default_params()
    sg.run()
    s = sg.score(p.emp_fc_tril, p.emp_fcd_tril)
    np.testing.assert_array_equal(s["+gof"].values, h["+gof"].values)


def test_cmaes_optimizer_runs():
    from cubnm_amd.optimize import CMAESOptimizer
    p = _problem()
    cm = CMAESOptimizer(popsize=4, n_iter=2, seed=5)
    cm.setup_problem(p)
    opt = cm.optimize()
    h = cm.history
    assert len(h) == 12
    assert ((h["G"] >= 0.5) & (h["G"] <= 3.0)).all() and ((h["w_p"] >= 0.5) & (h["w_p"] <= 1.5)).all()
    assert opt["cost"] == h["cost"].min()


def test_batch_optimize_variable_sc_matches_independent_runs():
    """Reference tests/optimize/test.py:157-193 on the GPU core: one CMA-ES
    on two problems with different SCs (the second reversed) in one call per
    generation; each history equals the independent run to 1e-6 (the
    reference's tolerance), and the two differ."""
    from cubnm_amd.optimize import CMAESOptimizer, batch_optimize
    p1 = _problem()
    p2 = _problem(sc=cases.synthetic_sc(N, seed=1)[::-1, ::-1].copy())
    o = CMAESOptimizer(popsize=4, n_iter=2, seed=1)
    b = batch_optimize(o, [p1, p2], save=False)
    assert not np.allclose(b[0].history["cost"].values, b[1].history["cost"].values, atol=1e-6)
    for k, p in enumerate((p1, p2)):
        alone = CMAESOptimizer(popsize=4, n_iter=2, seed=1)
        alone.setup_problem(p)
        alone.optimize()
        assert b[k].history.shape == alone.history.shape
        np.testing.assert_allclose(b[k].history.values.astype(float), alone.history.values.astype(float),
                                   rtol=0, atol=1e-6)


def test_nsga2_multiobjective_runs():
    """NSGA-II on the two GOF objectives: objective columns are the problem's
    obj_names, candidates stay in bounds, the values equal a direct score of
    the same candidates."""
    from cubnm_amd.optimize import NSGA2Optimizer
    p = _problem(multiobj=True)
    o = NSGA2Optimizer(popsize=4, n_iter=1, seed=2)
    o.setup_problem(p)
    o.optimize()
    h = o.history
    assert len(h) == 8 and list(h.columns[3:5]) == p.obj_names
    assert ((h["G"] >= 0.5) & (h["G"] <= 3.0)).all()
    X = p._get_X(h.loc[h["gen"] == 1, ["G", "w_p"]].values)
    F, _ = p.evaluate(X)
    np.testing.assert_allclose(F, h.loc[h["gen"] == 1, p.obj_names].values, rtol=0, atol=1e-12)
