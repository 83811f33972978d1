"""Differentiable rWW simulations for gradient-based fitting of regional
parameters (no counterpart in the reference, which fits with grid search and
CMA-ES; the forward model is the reference's rWW + Balloon-Windkessel BOLD,
src/ext/models/rww.cpp:187-225, bw.cpp:24-34).

``RwwBold(problem)(G, w_p, J_N)`` returns the simulated BOLD of a batch of
parameter sets as a torch tensor whose backward pass is the hand-written
adjoint kernel (cubnm_rww_grad_backward): the noise is the reference's
host-generated segment for the problem's seed, held fixed, so the gradient
is that of the deterministic map parameters -> BOLD.  Losses on the BOLD
(FC correlation, FCD quantiles) are written in torch on top of it.

Device tensors; FP64 throughout.  torch must be imported before the library
is loaded (it owns the HIP runtime of the process).
"""
import numpy as np
import torch

from . import capi


class RwwProblem:
    """Fixed part of a differentiable rWW batch: SC (nodes x nodes, shared),
    wIE and sigma (n_sims x nodes), timing, noise seed and the checkpoint
    interval (BW steps between stored states)."""

    def __init__(self, sc, wIE, sigma, duration_s, TR_s, dt="0.1", bw_dt="1.0", seed=0,
                 noise_segment_s=30, checkpoint_bw_steps=1000, device="cuda"):
        sc = np.ascontiguousarray(sc, dtype=float)
        self.nodes = sc.shape[0]
        wIE = np.asarray(wIE, float)
        self.n_sims = wIE.shape[0]
        self.time_steps = int(round(duration_s * 1000))
        self.TR = int(round(TR_s * 1000))
        self.dt, self.bw_dt = float(dt), float(bw_dt)
        self.ck = int(checkpoint_bw_steps)
        self.noise_ts = int(round(noise_segment_s * 1000))
        self.device = device
        # the reference's noise for this seed and shape, without running a simulation
        cfg = {"do_fc": "0", "do_fcd": "0", "bold_remove_s": "0", "exc_interhemispheric": "0",
               "window_size": "2", "window_step": "1", "drop_edges": "1", "noise_time_steps": str(self.noise_ts),
               "verbose": "0", "progress_interval": "500", "do_fic": "0"}
        d = capi.prepare("rWW", cfg, 1, self.nodes, self.time_steps, self.TR, self.TR, seed, self.dt, self.bw_dt,
                         noise_out=True)
        noise = np.empty(d.noise_size)
        shn = np.empty(d.noise_repeats * self.nodes, np.intc)
        sht = np.empty(d.noise_repeats * d.noise_bw_it, np.intc)
        capi.copy_noise(noise, shn, sht)
        self.bold_len = d.bold_len
        t = lambda a, dt=torch.float64: torch.as_tensor(np.array(a, copy=True), dtype=dt, device=device)
        self.sc, self.wIE, self.sigma = t(sc), t(wIE), t(np.broadcast_to(np.asarray(sigma, float), wIE.shape))
        self.noise, self.shn, self.sht = t(noise), t(shn, torch.int32), t(sht, torch.int32)
        nb, nc, ns = capi.rww_grad_sizes(self.n_sims, self.nodes, self.time_steps, self.TR, self.dt, self.bw_dt, self.ck)
        self.ckpt = torch.empty(nc, dtype=torch.float64, device=device)
        self.scratch = torch.empty(ns, dtype=torch.float64, device=device)

    def _common(self):
        return (self.n_sims, self.nodes, self.time_steps, self.TR, self.dt, self.bw_dt, self.ck, self.noise_ts)

    def forward(self, G, w_p, J_N):
        bold = torch.empty(self.n_sims, self.bold_len, self.nodes, dtype=torch.float64, device=self.device)
        torch.cuda.synchronize()
        capi.rww_grad_forward(*self._common(), self.sc.data_ptr(), G.data_ptr(), w_p.data_ptr(), J_N.data_ptr(),
                              self.wIE.data_ptr(), self.sigma.data_ptr(), self.noise.data_ptr(), self.shn.data_ptr(),
                              self.sht.data_ptr(), bold.data_ptr(), self.ckpt.data_ptr())
        return bold

    def backward(self, G, w_p, J_N, dbold):
        gG = torch.empty_like(G)
        gwp = torch.empty_like(w_p)
        gJN = torch.empty_like(J_N)
        dbold = dbold.contiguous()
        torch.cuda.synchronize()
        capi.rww_grad_backward(*self._common(), self.sc.data_ptr(), G.data_ptr(), w_p.data_ptr(), J_N.data_ptr(),
                               self.wIE.data_ptr(), self.sigma.data_ptr(), self.noise.data_ptr(), self.shn.data_ptr(),
                               self.sht.data_ptr(), self.ckpt.data_ptr(), dbold.data_ptr(), self.scratch.data_ptr(),
                               gG.data_ptr(), gwp.data_ptr(), gJN.data_ptr())
        return gG, gwp, gJN


class _RwwBoldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, G, w_p, J_N, problem):
        G, w_p, J_N = G.contiguous(), w_p.contiguous(), J_N.contiguous()
        ctx.problem = problem
        ctx.save_for_backward(G, w_p, J_N)
        return problem.forward(G, w_p, J_N)

    @staticmethod
    def backward(ctx, dbold):
        G, w_p, J_N = ctx.saved_tensors
        # the checkpoints of the last forward belong to these parameters
        gG, gwp, gJN = ctx.problem.backward(G, w_p, J_N, dbold)
        return gG, gwp, gJN, None


def rww_bold(problem, G, w_p, J_N):
    """Differentiable BOLD (n_sims, bold_len, nodes) of the problem's batch."""
    return _RwwBoldFn.apply(G, w_p, J_N, problem)


def fc_tril(bold, n_vols_remove=0):
    """Pearson FC lower triangle (i > j, row-major) per simulation, in torch."""
    x = bold[:, n_vols_remove:, :]
    x = x - x.mean(dim=1, keepdim=True)
    x = x / torch.sqrt((x * x).sum(dim=1, keepdim=True))
    fc = torch.bmm(x.transpose(1, 2).contiguous(), x.contiguous())
    i, j = torch.tril_indices(bold.shape[2], bold.shape[2], -1, device=bold.device)
    return fc[:, i, j]


def fc_corr_loss(bold, emp_fc_tril, n_vols_remove=0):
    """1 - Pearson(FC_sim, FC_emp) per simulation."""
    sim = fc_tril(bold, n_vols_remove)
    sim = sim - sim.mean(dim=1, keepdim=True)
    emp = emp_fc_tril - emp_fc_tril.mean()
    r = (sim * emp).sum(dim=1) / torch.sqrt((sim * sim).sum(dim=1) * (emp * emp).sum())
    return 1.0 - r
