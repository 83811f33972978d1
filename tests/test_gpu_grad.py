"""Gradients through the rWW simulation (adjoint kernel) on the GPU.

* The differentiable forward pass reproduces run_simulations' BOLD
  (feature-complete integrator, same noise, FIC off, wIE given) to 1e-12
  relative: the same operations, contracted into FMAs differently by the
  compiler in the two kernels (observed 1.5e-14).
* The adjoint gradient of L = sum(W * BOLD) with respect to G, w_p and J_N
  matches central finite differences of the same forward pass
  (h = 1e-6 relative), with checkpoint segments shorter than the run so the
  recompute path is exercised: relative error <= 1e-5 of the largest
  gradient entry of each kind.
* Through torch autograd, the gradient of 1 - corr(FC_sim, FC_emp) with
  respect to G matches central differences of that loss.
* The same two checks at 60 nodes, where the kernels split the coupling sum
  (and its transpose) over two partitions of source nodes.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

N, S, DUR, TR = 6, 3, 4.0, 0.5


def _problem(ck=3, n=N):
    import torch  # noqa: F401
    from cubnm_amd import grad
    wIE = np.full((S, n), 1.2)
    sig = np.full((S, n), 0.01)
    return grad.RwwProblem(cases.synthetic_sc(n, seed=3), wIE, sig, DUR, TR, seed=4, noise_segment_s=1,
                           checkpoint_bw_steps=ck)


def _params(torch, n=N):
    rng = np.random.default_rng(0)
    G = torch.tensor([0.4, 0.9, 1.3], dtype=torch.float64, device="cuda")
    wp = torch.tensor(rng.uniform(0.8, 1.6, (S, n)), dtype=torch.float64, device="cuda")
    JN = torch.tensor(rng.uniform(0.1, 0.3, (S, n)), dtype=torch.float64, device="cuda")
    return G, wp, JN


@pytest.mark.parametrize("n", [N, 60])
def test_forward_matches_run_simulations(n):
    import torch
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    N = n
    prob = _problem(n=n)
    G, wp, JN = _params(torch, n)
    bold = prob.forward(G, wp, JN).cpu().numpy()
    gp, rp = cases.params("rWW", S, N, G=G.cpu().numpy(), w_p=wp.cpu().numpy(), J_N=JN.cpu().numpy(),
                          wIE=1.2, sigma=0.01, do_fic=False)
    cfg = cases.model_config("rWW", TR=TR, do_fc=False, do_fcd=False, bold_remove_s=0, do_fic=False,
                             noise_segment_length=1)
    args = cases.sim_args("rWW", cases.synthetic_sc(N, seed=3), gp, rp, cfg, S, duration_s=DUR, TR_s=TR, sim_seed=4)
    _core.force_integrator(0)
    try:
        out = _core.run_simulations(*args)
    finally:
        _core.force_integrator(-1)
    assert cases.bold_rel_err(bold.reshape(S, -1), out["sim_bold"]).max() <= 1e-12


@pytest.mark.parametrize("n", [N, 60])
def test_adjoint_matches_finite_differences(n):
    import torch
    N = n
    prob = _problem(ck=3, n=n)
    G, wp, JN = _params(torch, n)
    W = torch.tensor(np.random.default_rng(1).standard_normal((S, prob.bold_len, N)), dtype=torch.float64,
                     device="cuda")
    L = lambda g, w, j: float((prob.forward(g, w, j) * W).sum())
    prob.forward(G, wp, JN)
    gG, gwp, gJN = [t.cpu().numpy() for t in prob.backward(G, wp, JN, W)]

    def fd(which, idx):
        base = {"G": G, "wp": wp, "JN": JN}
        x = base[which]
        h = 1e-6 * max(1.0, abs(float(x.reshape(-1)[idx])))
        vals = []
        for sgn in (1, -1):
            y = x.clone().reshape(-1)
            y[idx] += sgn * h
            args = dict(base)
            args[which] = y.reshape(x.shape)
            vals.append(L(args["G"], args["wp"], args["JN"]))
        return (vals[0] - vals[1]) / (2 * h)

    for s in range(S):
        assert abs(fd("G", s) - gG[s]) <= 1e-5 * np.abs(gG).max(), (s, fd("G", s), gG[s])
    for flat in [0, 5, 7, 13, 17]:
        assert abs(fd("wp", flat) - gwp.reshape(-1)[flat]) <= 1e-5 * np.abs(gwp).max()
        assert abs(fd("JN", flat) - gJN.reshape(-1)[flat]) <= 1e-5 * np.abs(gJN).max()


def test_fc_loss_gradient_through_autograd():
    import torch
    from cubnm_amd import grad
    prob = _problem(ck=2)
    G, wp, JN = _params(torch)
    emp = torch.tensor(np.tanh(np.random.default_rng(2).standard_normal(N * (N - 1) // 2)), dtype=torch.float64,
                       device="cuda")
    Gv = G.clone().requires_grad_(True)
    loss = grad.fc_corr_loss(grad.rww_bold(prob, Gv, wp, JN), emp).sum()
    loss.backward()
    for s in range(S):
        h = 1e-6
        vals = []
        for sgn in (1, -1):
            Gs = G.clone()
            Gs[s] += sgn * h
            vals.append(float(grad.fc_corr_loss(prob.forward(Gs, wp, JN), emp).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - float(Gv.grad[s])) <= 1e-5 * float(Gv.grad.abs().max()), (s, fd, float(Gv.grad[s]))
