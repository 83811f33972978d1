"""Fitting regional rWW parameters: gradient descent through the adjoint
kernel versus CMA-ES, on the same forward model, noise and loss.

* Loss (per parameter set): (1 - Pearson(FC_sim, FC_emp)) plus the mean
  absolute difference of 64 quantiles of the FCD lower triangles of the
  simulated BOLD (``cubnm_amd.grad.rww_bold``).  ``FcFcdLoss`` evaluates it
  and its gradient with respect to BOLD in one hand-written kernel
  (cubnm_fc_fcd_loss); ``fit_loss`` is the same loss in torch operations,
  kept as the reference the kernel is tested against.
* Parameters live in the unit cube, mapped linearly to their bounds, as the
  reference's optimisers do (optimize.py:432-537 maps [0, 1] to each
  parameter's range).
* ``CMAES``: the (mu/mu_w, lambda)-CMA-ES with cumulative step-size
  adaptation, rank-one and rank-mu covariance updates and the default
  strategy parameters of Hansen, "The CMA Evolution Strategy: A Tutorial"
  (arXiv:1604.00772), with sampled points clipped to the unit cube.  Like
  the reference's CMAESOptimizer with x0=None (optimize.py:1386-1484), the
  first generation is a uniform sample and its best point is the initial
  mean; sigma = 0.5.
* ``adam_fit``: Adam on a batch of starting points (one simulation each),
  clipped to the unit cube.
"""
import time

import numpy as np
import torch

from . import grad


def window_fc(bold, windows, n_vols_remove=0):
    """FC lower triangles of each window: (sims, windows, pairs)."""
    out = []
    for s, e in windows:
        out.append(grad.fc_tril(bold[:, s:e + 1, :]))
    return torch.stack(out, dim=1)


def fcd_tril(bold, windows):
    """FCD lower triangle (sims, window pairs): Pearson r between window-FC vectors."""
    w = window_fc(bold, windows)
    w = w - w.mean(dim=2, keepdim=True)
    w = w / torch.sqrt((w * w).sum(dim=2, keepdim=True))
    fcd = torch.bmm(w, w.transpose(1, 2).contiguous())
    i, j = torch.tril_indices(len(windows), len(windows), -1, device=bold.device)
    return fcd[:, i, j]


def fit_loss(bold, emp_fc, emp_fcd_q, windows, n_vols_remove, n_quantiles=64):
    """(1 - corr(FC_sim, FC_emp)) + mean |quantiles(FCD_sim) - quantiles(FCD_emp)|, per simulation."""
    l_fc = grad.fc_corr_loss(bold, emp_fc, n_vols_remove)
    qs = torch.linspace(0, 1, n_quantiles, dtype=bold.dtype, device=bold.device)
    q = torch.quantile(fcd_tril(bold, windows), qs, dim=1).T
    return l_fc + (q - emp_fcd_q[None, :]).abs().mean(dim=1)


class FcFcdLoss:
    """The loss of ``fit_loss`` on device BOLD (n_sims, bold_len, nodes),
    evaluated by the library's kernel; ``value(bold)`` gives the loss per
    simulation, ``value_and_grad(bold)`` also dL/dBOLD."""

    def __init__(self, emp_fc, emp_q, windows, n_vols_remove, n_sims, bold_len, nodes, n_quantiles=64):
        from . import capi
        self.capi = capi
        dev = emp_fc.device
        if any(s < 0 or e >= bold_len or e <= s for s, e in windows) or len(windows) < 2:
            raise ValueError("windows must lie within the BOLD volumes, at least two, each two volumes or more")
        self.S, self.T, self.N, self.nvr = n_sims, bold_len, nodes, n_vols_remove
        self.W = len(windows)
        self.Lmax = max(e - s + 1 for s, e in windows)
        self.ws = torch.tensor([w[0] for w in windows], dtype=torch.int32, device=dev)
        self.we = torch.tensor([w[1] for w in windows], dtype=torch.int32, device=dev)
        self.emp_fc = emp_fc.contiguous()
        self.emp_q = emp_q.contiguous()
        self.qs = torch.linspace(0, 1, n_quantiles, dtype=torch.float64, device=dev)
        self.nq = n_quantiles
        per = capi.fc_fcd_loss_scratch(bold_len, nodes, n_vols_remove, self.W, self.Lmax)
        self.scratch = torch.empty(n_sims * per, dtype=torch.float64, device=dev)

    def _run(self, bold, with_grad):
        bold = bold.contiguous()
        assert bold.shape == (self.S, self.T, self.N) and bold.dtype == torch.float64
        loss = torch.empty(self.S, dtype=torch.float64, device=bold.device)
        dbold = torch.empty_like(bold) if with_grad else None
        torch.cuda.synchronize()
        self.capi.fc_fcd_loss(bold.data_ptr(), self.S, self.T, self.N, self.nvr, self.ws.data_ptr(),
                              self.we.data_ptr(), self.W, self.Lmax, self.emp_fc.data_ptr(), self.emp_q.data_ptr(),
                              self.qs.data_ptr(), self.nq, loss.data_ptr(),
                              dbold.data_ptr() if with_grad else None, self.scratch.data_ptr())
        torch.cuda.synchronize()
        return loss, dbold

    def value(self, bold):
        return self._run(bold, False)[0]

    def value_and_grad(self, bold):
        return self._run(bold, True)

    def __call__(self, bold):
        """Differentiable loss: backward multiplies the kernel's dL/dBOLD."""
        return _FcFcdLossFn.apply(bold, self)


class _FcFcdLossFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, bold, lossobj):
        loss, dbold = lossobj.value_and_grad(bold.detach())
        ctx.save_for_backward(dbold)
        return loss

    @staticmethod
    def backward(ctx, gl):
        (dbold,) = ctx.saved_tensors
        return dbold * gl[:, None, None], None


def windows_for(bold_len, window, step, n_vols_remove):
    """Window bounds [c - window/2, c + window/2] for centres every `step`
    volumes after burn-in, dropping edge windows (utils.cpp:21-72)."""
    corr_len = bold_len - n_vols_remove
    first, last = window // 2 + n_vols_remove, corr_len - 1 - window // 2 + n_vols_remove
    return [(c - window // 2, c + window // 2) for c in range(first, last + 1, step)]


class CMAES:
    def __init__(self, x0, sigma, popsize, seed=0):
        self.rng = np.random.default_rng(seed)
        n = len(x0)
        self.n, self.lam = n, popsize
        self.mu = popsize // 2
        w = np.log(self.mu + 0.5) - np.log(np.arange(1, self.mu + 1))
        self.w = w / w.sum()
        self.mueff = 1.0 / np.sum(self.w ** 2)
        self.cc = (4 + self.mueff / n) / (n + 4 + 2 * self.mueff / n)
        self.cs = (self.mueff + 2) / (n + self.mueff + 5)
        self.c1 = 2 / ((n + 1.3) ** 2 + self.mueff)
        self.cmu = min(1 - self.c1, 2 * (self.mueff - 2 + 1 / self.mueff) / ((n + 2) ** 2 + self.mueff))
        self.damps = 1 + 2 * max(0.0, np.sqrt((self.mueff - 1) / (n + 1)) - 1) + self.cs
        self.chiN = np.sqrt(n) * (1 - 1 / (4 * n) + 1 / (21 * n * n))
        self.mean = np.array(x0, float)
        self.sigma = sigma
        self.C = np.eye(n)
        self.B, self.D = np.eye(n), np.ones(n)
        self.pc, self.ps = np.zeros(n), np.zeros(n)
        self.gen = 0

    def ask(self):
        z = self.rng.standard_normal((self.lam, self.n))
        y = (z * self.D) @ self.B.T
        self._y = y
        return np.clip(self.mean + self.sigma * y, 0.0, 1.0)

    def tell(self, x, f):
        order = np.argsort(f)
        xs = x[order[:self.mu]]
        old = self.mean
        self.mean = self.w @ xs
        ymean = (self.mean - old) / self.sigma
        Cinvsqrt = (self.B / self.D) @ self.B.T
        self.ps = (1 - self.cs) * self.ps + np.sqrt(self.cs * (2 - self.cs) * self.mueff) * (Cinvsqrt @ ymean)
        self.gen += 1
        hsig = (np.linalg.norm(self.ps) / np.sqrt(1 - (1 - self.cs) ** (2 * self.gen)) / self.chiN
                < 1.4 + 2 / (self.n + 1))
        self.pc = (1 - self.cc) * self.pc + hsig * np.sqrt(self.cc * (2 - self.cc) * self.mueff) * ymean
        yk = (xs - old) / self.sigma
        self.C = ((1 - self.c1 - self.cmu) * self.C
                  + self.c1 * (np.outer(self.pc, self.pc) + (1 - hsig) * self.cc * (2 - self.cc) * self.C)
                  + self.cmu * (yk.T * self.w) @ yk)
        self.sigma *= np.exp((self.cs / self.damps) * (np.linalg.norm(self.ps) / self.chiN - 1))
        self.C = (self.C + self.C.T) / 2
        d, B = np.linalg.eigh(self.C)
        self.D, self.B = np.sqrt(np.maximum(d, 1e-30)), B


class RegionalFit:
    """G (1) + w_p (N) + J_N (N) in the unit cube for one SC and empirical
    BOLD (nodes x volumes, as the reference's emp_bold); ``evaluate(U)``
    returns the loss of each row of U (a batch the size of the problem),
    ``loss_and_grad(U)`` also its gradient."""

    def __init__(self, problem, emp_bold, bounds, window, step, n_vols_remove):
        self.p = problem
        self.N = problem.nodes
        dev = problem.device
        self.lo = torch.tensor(np.concatenate([[bounds["G"][0]], np.full(self.N, bounds["w_p"][0]),
                                               np.full(self.N, bounds["J_N"][0])]), dtype=torch.float64, device=dev)
        self.hi = torch.tensor(np.concatenate([[bounds["G"][1]], np.full(self.N, bounds["w_p"][1]),
                                               np.full(self.N, bounds["J_N"][1])]), dtype=torch.float64, device=dev)
        self.nvr = n_vols_remove
        self.windows = windows_for(problem.bold_len, window, step, n_vols_remove)
        eb = torch.tensor(np.asarray(emp_bold, float).T.copy(), dtype=torch.float64, device=dev)[None]
        self.emp_fc = grad.fc_tril(eb, n_vols_remove)[0]
        qs = torch.linspace(0, 1, 64, dtype=torch.float64, device=dev)
        self.emp_q = torch.quantile(fcd_tril(eb, self.windows)[0], qs)
        self.loss = FcFcdLoss(self.emp_fc, self.emp_q, self.windows, n_vols_remove, problem.n_sims,
                              problem.bold_len, self.N)

    def _split(self, theta):
        return theta[:, 0].contiguous(), theta[:, 1:1 + self.N].contiguous(), theta[:, 1 + self.N:].contiguous()

    def evaluate(self, U):
        with torch.no_grad():
            theta = self.lo + (self.hi - self.lo) * U
            bold = self.p.forward(*self._split(theta))
            return self.loss.value(bold)

    def loss_and_grad(self, U):
        loss, finish = self.loss_then_grad(U)
        return loss, finish()

    def loss_then_grad(self, U):
        """The loss of each row of U (forward pass only) and a function that
        runs the backward pass and returns the gradient, so that a caller
        that stops on the loss skips the backward pass of its last point."""
        Uv = U.clone().requires_grad_(True)
        theta = self.lo + (self.hi - self.lo) * Uv
        G, wp, JN = self._split(theta)
        loss = self.loss(grad.rww_bold(self.p, G, wp, JN))

        def finish():
            loss.sum().backward()
            return Uv.grad.detach()
        return loss.detach(), finish


def cmaes_fit(fit, n_gen, seed=0, sigma=0.5, log=None):
    """CMA-ES in the unit cube, popsize = the problem's batch; returns the
    (wall seconds, best loss so far) history."""
    rng = np.random.default_rng(seed)
    dim = 1 + 2 * fit.N
    lam = fit.p.n_sims
    t0 = time.time()
    hist = []
    X = rng.uniform(0, 1, (lam, dim))
    f = fit.evaluate(torch.tensor(X, device=fit.p.device)).cpu().numpy()
    best = float(np.nanmin(f))
    hist.append((time.time() - t0, best))
    es = CMAES(X[int(np.nanargmin(f))], sigma, lam, seed=seed + 1)
    for g in range(n_gen):
        X = es.ask()
        f = fit.evaluate(torch.tensor(X, device=fit.p.device)).cpu().numpy()
        f = np.where(np.isnan(f), np.inf, f)
        es.tell(X, f)
        best = min(best, float(f.min()))
        hist.append((time.time() - t0, best))
        if log:
            log(f"cmaes gen {g + 1}: best {best:.5f} at {hist[-1][0]:.1f} s")
    return hist


def adam_fit(fit, n_steps, lr=0.02, seed=0, target=None, log=None):
    """Adam from a uniform batch of starts; stops early at `target` (checked
    on each point's loss before its gradient is computed); returns the (wall
    seconds, best loss so far) history."""
    rng = np.random.default_rng(seed)
    dim = 1 + 2 * fit.N
    U = torch.tensor(rng.uniform(0, 1, (fit.p.n_sims, dim)), dtype=torch.float64, device=fit.p.device)
    m = torch.zeros_like(U)
    v = torch.zeros_like(U)
    b1, b2, eps = 0.9, 0.999, 1e-8
    t0 = time.time()
    hist = []
    best = float("inf")
    for k in range(1, n_steps + 1):
        loss, finish = fit.loss_then_grad(U)
        lf = loss.cpu().numpy()
        best = min(best, float(np.nanmin(np.where(np.isnan(lf), np.inf, lf))))
        hist.append((time.time() - t0, best))
        if log:
            log(f"adam step {k}: best {best:.5f} at {hist[-1][0]:.1f} s")
        # the gradient is needed only to take another step
        if (target is not None and best <= target) or k == n_steps:
            break
        g = torch.nan_to_num(finish())
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        U = (U - lr * (m / (1 - b1 ** k)) / (torch.sqrt(v / (1 - b2 ** k)) + eps)).clamp(0.0, 1.0)
    return hist
