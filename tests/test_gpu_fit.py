"""Fitting harness on the GPU: the loss gradient (adjoint kernel + torch
autograd through FC correlation and FCD quantiles) matches central
differences of the loss; CMA-ES and Adam run and do not increase the best
loss."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _fit(n_sims):
    import torch  # noqa: F401
    from cubnm_amd import fit, grad
    N = 6
    prob = grad.RwwProblem(cases.synthetic_sc(N, seed=2), np.ones((n_sims, N)), 0.01, 20.0, 0.72, seed=1,
                           noise_segment_s=5, checkpoint_bw_steps=500)
    emp = np.random.default_rng(0).standard_normal((N, prob.bold_len))
    bounds = {"G": (0.001, 10.0), "w_p": (0.0, 2.0), "J_N": (0.001, 0.5)}
    return fit.RegionalFit(prob, emp, bounds, 6, 2, 2)


def test_loss_gradient_matches_finite_differences():
    import torch
    f = _fit(2)
    U = torch.tensor(np.random.default_rng(1).uniform(0.2, 0.6, (2, 13)), dtype=torch.float64, device="cuda")
    loss, g = f.loss_and_grad(U)
    for s, k in [(0, 0), (1, 0), (0, 3), (1, 9)]:
        h = 1e-6
        vals = []
        for sgn in (1, -1):
            V = U.clone()
            V[s, k] += sgn * h
            vals.append(float(f.evaluate(V)[s]))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - float(g[s, k])) <= 1e-4 * max(1e-3, float(g.abs().max())), (s, k, fd, float(g[s, k]))


def test_cmaes_and_adam_run():
    from cubnm_amd import fit
    f = _fit(8)
    h = fit.cmaes_fit(f, 3, seed=0)
    assert len(h) == 4 and all(b2 <= b1 for (_, b1), (_, b2) in zip(h, h[1:]))
    h2 = fit.adam_fit(f, 3, seed=0)
    assert len(h2) == 3 and np.isfinite(h2[-1][1])
