"""Goodness-of-fit and empirical-target helpers around the simulation core.

Mirrors the reference's src/cubnm/utils.py for the functions on the hot
path's caller side:
  get_bw_params   utils.py:130-160  (Balloon-Windkessel k1..k3 sets)
  calculate_fc    utils.py:162-201  (empirical FC tril)
  calculate_fcd   utils.py:204-304  (empirical FCD tril)
  fc_corr_device  utils.py:416-507  (Pearson r of each simulated FC vs empirical)
  fcd_ks_device   utils.py:306-414  (KS distance of each simulated FCD vs empirical)
  avail_gpus      utils.py:68-111   (GPUs on the machine)
  is_jupyter      utils.py:113-128  (running under a Jupyter kernel)
The two *_device functions run on the GPU through the C ABI
(cubnm_gof_fc_corr / cubnm_gof_fcd_ks); the empirical helpers are the
reference's NumPy restated (they run once per problem).
"""
import glob
import sys

import numpy as np
import scipy.stats

from . import _core


_KFD_TOPOLOGY = "/sys/class/kfd/kfd/topology/nodes"


def avail_gpus():
    """Number of GPUs on the machine, counted from the kernel driver's
    topology (nodes with SIMDs are GPUs; CPU nodes report 0), without
    initialising a GPU runtime.  0 where the driver is absent."""
    n = 0
    for props in glob.glob(f"{_KFD_TOPOLOGY}/*/properties"):
        try:
            with open(props) as f:
                for line in f:
                    key, _, val = line.partition(" ")
                    if key == "simd_count" and int(val) > 0:
                        n += 1
                        break
        except (OSError, ValueError):
            continue
    return n


def is_jupyter():
    """True when running under a Jupyter (IPython kernel) session."""
    ipython = sys.modules.get("IPython")
    shell = ipython.get_ipython() if ipython is not None and hasattr(ipython, "get_ipython") else None
    return shell is not None and "IPKernelApp" in getattr(shell, "config", {})


def get_bw_params(src):
    """Balloon-Windkessel k1, k2, k3 for 'friston2003' or 'heinzle2016-3T'."""
    if src == "friston2003":
        rho = 0.34
        return {"k1": 7 * rho, "k2": 2.0, "k3": 2 * rho - 0.2}
    elif src == "heinzle2016-3T":
        return {"k1": 3.72, "k2": 0.527, "k3": 0.53}
    raise ValueError(f"Unknown BW parameters source: {src}")


def _zscore_nonoutlier(bold):
    """z-score each node over non-outlier volumes (all-zero columns are
    motion outliers and stay zero).  Works on a copy: the reference
    z-scores the caller's array in place (utils.py:189-190)."""
    bold = np.array(bold, dtype=float, copy=True)
    outlier = bold.sum(axis=0) == 0
    bold[:, ~outlier] = scipy.stats.zscore(bold[:, ~outlier], axis=1)
    return bold


def calculate_fc(bold, exc_interhemispheric=False, return_tril=True):
    """Empirical FC of BOLD (nodes, volumes); lower triangle by default."""
    bold = _zscore_nonoutlier(bold)
    fc = np.corrcoef(bold)
    if exc_interhemispheric:
        rh = bold.shape[0] // 2
        fc[:rh, rh:] = np.nan
        fc[rh:, :rh] = np.nan
    if return_tril:
        fc = fc[np.tril_indices(fc.shape[0], -1)]
        fc = fc[~np.isnan(fc)]
    return fc


def calculate_fcd(bold, window_size, window_step, drop_edges=True, outlier_threshold=0.5,
                  exc_interhemispheric=False, return_tril=True, return_dfc=False):
    """Empirical FCD of BOLD (nodes, volumes) with windows of window_size+1
    volumes every window_step volumes; windows with more than
    outlier_threshold outlier volumes are dropped."""
    assert window_size % 2 == 0, "Window size must be even"
    bold = _zscore_nonoutlier(bold)
    nodes, n_vols = bold.shape
    half = window_size // 2
    first, last = (half, n_vols - 1 - half) if drop_edges else (0, n_vols - 1)
    trils, fcs = [], []
    c = first
    while c <= last:
        s0, s1 = max(c - half, 0), min(c + half, n_vols - 1)
        c += window_step
        wb = bold[:, s0:s1 + 1]
        if (wb.sum(axis=0) == 0).mean() > outlier_threshold:
            continue
        wfc = np.corrcoef(wb)
        if exc_interhemispheric:
            rh = nodes // 2
            wfc[:rh, rh:] = np.nan
            wfc[rh:, :rh] = np.nan
        fcs.append(wfc[:, :, None])
        t = wfc[np.tril_indices(nodes, -1)]
        trils.append(t[~np.isnan(t)][:, None])
    trils = np.concatenate(trils, axis=1)
    fcd = np.corrcoef(trils.T)
    if return_tril:
        fcd = fcd[np.tril_indices(fcd.shape[0], -1)]
    if return_dfc:
        return fcd, np.concatenate(fcs, axis=2)
    return fcd


def fc_corr_device(sim_fc_trils, emp_fc_tril, usable_mem=None):
    """Pearson r between every simulated FC tril (N, n_pairs) and the
    empirical FC tril, on the GPU.  ``usable_mem`` is accepted for
    signature compatibility; batching is not needed at 288 GB of HBM."""
    return _core.gof_fc_corr(np.ascontiguousarray(sim_fc_trils, float), np.ascontiguousarray(emp_fc_tril, float))


def fcd_ks_device(sim_fcd_trils, emp_fcd_tril, usable_mem=None):
    """Two-sample KS statistic between every simulated FCD tril (N, m1) and
    the empirical FCD tril (m2,), on the GPU."""
    return _core.gof_fcd_ks(np.ascontiguousarray(sim_fcd_trils, float), np.ascontiguousarray(emp_fcd_tril, float))
