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


def _torch_rww_bold(prob, G, wp, JN):
    """An independent restatement of the differentiable forward model in
    float64 torch operations on the CPU (rww.cpp:187-225 with the noise
    indexing of bnm.cpp:261-284, bw.cpp:24-34, BOLD bnm.cpp:318), for
    autograd: the reference gradient of the adjoint kernel."""
    import torch
    N = prob.nodes
    dt, sqrt_dt = prob.dt, np.sqrt(prob.dt)
    inner = int(round(prob.bw_dt / prob.dt))
    bw_it = prob.time_steps // int(round(prob.bw_dt))
    tr_it = prob.TR // int(round(prob.bw_dt))
    nbw = prob.noise_ts // int(round(prob.bw_dt))
    sc = prob.sc.cpu()
    noise, shn, sht = prob.noise.cpu(), prob.shn.cpu().long(), prob.sht.cpu().long()
    wIE, sigma = prob.wIE.cpu(), prob.sigma.cpu()
    a_E, b_E, d_E, a_I, b_I, d_I = 310.0, 125.0, 0.16, 615.0, 177.0, 0.087
    gE, gI, tE, tI, I0, w_E, w_I, w_II = 0.641e-3, 1e-3, 100.0, 10.0, 0.382, 1.0, 0.7, 1.0
    rho, alpha, tau, gam, kap = 0.34, 0.32, 0.98, 1 / 0.41, 1 / 0.65
    k1, k2, k3, V0 = 7 * rho, 2.0, 2 * rho - 0.2, 2.0
    bdt = prob.bw_dt / 1000.0
    H = lambda x, a, b, d: (a * x - b) / (1 - torch.exp(-d * (a * x - b)))  # noqa: E731
    S = G.shape[0]
    SE = torch.full((S, N), 0.001, dtype=torch.float64)
    SI = torch.full((S, N), 0.001, dtype=torch.float64)
    x = torch.zeros((S, N), dtype=torch.float64)
    f, nu, q = torch.ones_like(x), torch.ones_like(x), torch.ones_like(x)
    out = []
    for bw_i in range(bw_it):
        rep = bw_i // nbw
        sh_t = int(sht[rep * nbw + bw_i % nbw])
        sh_j = shn[rep * N:(rep + 1) * N]
        for inner_i in range(inner):
            g = SE @ sc  # g_j = sum_k SC[k][j] S_E(k)
            base = ((sh_t * inner + inner_i) * N + sh_j) * 2
            n0, n1 = noise[base], noise[base + 1]
            IE = w_E * I0 + wp * JN * SE + g * G[:, None] * JN - wIE * SI
            II = w_I * I0 + JN * SE - w_II * SI
            SE_n = SE + n0 * sigma * sqrt_dt + dt * gE * (1 - SE) * H(IE, a_E, b_E, d_E) - dt / tE * SE
            SI_n = SI + n1 * sigma * sqrt_dt + dt * gI * H(II, a_I, b_I, d_I) - dt / tI * SI
            SE, SI = SE_n.clamp(0.0, 1.0), SI_n.clamp(0.0, 1.0)
        x_n = x + bdt * (SE - kap * x - gam * (f - 1))
        nu_n = nu + bdt / tau * (f - nu ** (1 / alpha))
        q = q + bdt / tau * (f * (1 - (1 - rho) ** (1 / f)) / rho - nu_n ** (1 / alpha) * q / nu_n)
        f = f + bdt * x_n
        x, nu = x_n, nu_n
        if (bw_i + 1) % tr_it == 0:
            out.append(V0 * (k1 * (1 - q) + k2 * (1 - q / nu) + k3 * (1 - nu)))
    return torch.stack(out, dim=1)


def test_adjoint_matches_torch_autograd_reference():
    """The adjoint kernel's gradient of L = sum(W * BOLD) against float64
    autograd of the restatement above, on the same noise (N = 4, 0.6 s at
    dt 0.1 ms, 6 volumes at TR 0.1 s, checkpoint segments of 20 BW steps):
    relative 1e-8 of the largest entry of each gradient."""
    import torch
    from cubnm_amd import grad
    n, s = 4, 2
    prob = grad.RwwProblem(cases.synthetic_sc(n, seed=7), np.full((s, n), 1.1), np.full((s, n), 0.01), 0.6, 0.1,
                           seed=3, noise_segment_s=0.25, checkpoint_bw_steps=20)
    rng = np.random.default_rng(5)
    G = torch.tensor([0.7, 1.4], dtype=torch.float64, device="cuda")
    wp = torch.tensor(rng.uniform(0.8, 1.6, (s, n)), dtype=torch.float64, device="cuda")
    JN = torch.tensor(rng.uniform(0.1, 0.3, (s, n)), dtype=torch.float64, device="cuda")
    bold = prob.forward(G, wp, JN)
    W = torch.tensor(rng.standard_normal(tuple(bold.shape)), dtype=torch.float64, device="cuda")
    gG, gwp, gJN = [t.cpu() for t in prob.backward(G, wp, JN, W)]
    Gc, wpc, JNc = [t.cpu().clone().requires_grad_(True) for t in (G, wp, JN)]
    ref = _torch_rww_bold(prob, Gc, wpc, JNc)
    np.testing.assert_allclose(ref.detach().numpy(), bold.cpu().numpy(), rtol=1e-10, atol=1e-14)
    (ref * W.cpu()).sum().backward()
    for got, want in [(gG, Gc.grad), (gwp, wpc.grad), (gJN, JNc.grad)]:
        scale = float(want.abs().max())
        assert float((got - want).abs().max()) <= 1e-8 * scale, (got, want)
