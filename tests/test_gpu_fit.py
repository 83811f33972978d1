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


def _loss_case(S=3, T=40, N=7, nvr=6, seed=0):
    import torch
    from cubnm_amd import fit, grad
    rng = np.random.default_rng(seed)
    bold = torch.tensor(rng.standard_normal((S, T, N)).cumsum(axis=1) * 0.1 + rng.standard_normal((S, 1, N)),
                        dtype=torch.float64, device="cuda")
    windows = fit.windows_for(T, 6, 3, nvr)
    emp_bold = torch.tensor(rng.standard_normal((1, T, N)).cumsum(axis=1), dtype=torch.float64, device="cuda")
    emp_fc = grad.fc_tril(emp_bold, nvr)[0]
    qs = torch.linspace(0, 1, 64, dtype=torch.float64, device="cuda")
    emp_q = torch.quantile(fit.fcd_tril(emp_bold, windows)[0], qs)
    return bold, windows, emp_fc, emp_q, nvr


def test_loss_kernel_matches_torch_autograd():
    """The hand-written loss and its adjoint (cubnm_fc_fcd_loss) against the
    same loss written in torch operations (fit.fit_loss) and its autograd,
    float64 on the CPU: loss 1e-12, gradient 1e-10 of its largest entry."""
    import torch
    from cubnm_amd import fit
    bold, windows, emp_fc, emp_q, nvr = _loss_case()
    S, T, N = bold.shape
    lossobj = fit.FcFcdLoss(emp_fc, emp_q, windows, nvr, S, T, N)
    loss, dbold = lossobj.value_and_grad(bold)
    b = bold.cpu().clone().requires_grad_(True)
    ref = fit.fit_loss(b, emp_fc.cpu(), emp_q.cpu(), windows, nvr)
    ref.sum().backward()
    np.testing.assert_allclose(loss.cpu().numpy(), ref.detach().numpy(), rtol=0, atol=1e-12)
    scale = float(b.grad.abs().max())
    assert float((dbold.cpu() - b.grad).abs().max()) <= 1e-10 * scale
    np.testing.assert_array_equal(lossobj.value(bold).cpu().numpy(), loss.cpu().numpy())
    # through autograd: the backward multiplies by the incoming gradient
    bv = bold.clone().requires_grad_(True)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    (lossobj(bv) * w).sum().backward()
    np.testing.assert_allclose(bv.grad.cpu().numpy(), (dbold * w[:, None, None]).cpu().numpy(), rtol=0, atol=0)


def test_loss_kernel_nan_simulation():
    from cubnm_amd import fit
    bold, windows, emp_fc, emp_q, nvr = _loss_case(seed=1)
    bold[1, :, 2] = 0.25  # a constant node: undefined correlations
    S, T, N = bold.shape
    loss, dbold = fit.FcFcdLoss(emp_fc, emp_q, windows, nvr, S, T, N).value_and_grad(bold)
    loss = loss.cpu().numpy()
    assert np.isnan(loss[1]) and np.isfinite(loss[[0, 2]]).all()
    assert float(dbold[1].abs().max()) == 0.0 and float(dbold[0].abs().max()) > 0.0
