"""Empirical FC / FCD targets (cubnm_amd.utils.calculate_fc / calculate_fcd),
the reference's utils.py:162-304, pinned without the reference:

* FC is Pearson r of every node pair i > j in np.tril_indices order: checked
  pair by pair against scipy.stats.pearsonr on BOLD without outliers (the
  z-scoring is a per-node affine map and leaves r unchanged).
* Motion outliers (all-zero volumes) are excluded from each node's z-scoring
  but stay in the correlation as zeros: checked against the statistic written
  out by hand (mean and ddof-0 standard deviation over the kept volumes,
  then Pearson over all volumes).  Since the z-scored kept volumes have mean
  0, the zero volumes add nothing to any sum, so FC equals Pearson r over the
  kept volumes alone: a closed form that also pins the result.
* Inter-hemispheric exclusion keeps only pairs within the first or the
  second half of the nodes.
* FCD windows: centres from window_size/2 to n_vols-1-window_size/2 every
  window_step volumes (drop_edges), window_size+1 volumes each, windows with
  more than half outlier volumes dropped; checked against a hand-enumerated
  window list and brute-force Pearson of the window FC vectors.
* The caller's array is not modified (the reference z-scores it in place;
  a documented difference that cannot change the returned values).
"""
import numpy as np
import pytest
import scipy.stats

from cubnm_amd import utils


def _bold(nodes, vols, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((nodes, vols)).cumsum(axis=1) * 0.3 + rng.standard_normal((nodes, 1))


def test_fc_pairs_and_values():
    b = _bold(7, 40, 0)
    fc = utils.calculate_fc(b)
    i, j = np.tril_indices(7, -1)
    assert fc.shape == (21,)
    assert all(ii > jj for ii, jj in zip(i, j))
    ref = np.array([scipy.stats.pearsonr(b[a], b[c])[0] for a, c in zip(i, j)])
    np.testing.assert_allclose(fc, ref, rtol=0, atol=1e-13)
    full = utils.calculate_fc(b, return_tril=False)
    np.testing.assert_allclose(np.diag(full), 1.0, atol=1e-14)
    np.testing.assert_allclose(full[i, j], fc, atol=0)


def _hand_zscore(b):
    keep = b.sum(axis=0) != 0
    z = np.zeros_like(b)
    for n in range(b.shape[0]):
        x = b[n, keep]
        z[n, keep] = (x - x.mean()) / np.sqrt(np.mean((x - x.mean()) ** 2))
    return z


def _pearson(x, y):
    dx, dy = x - x.mean(), y - y.mean()
    return float((dx * dy).sum() / np.sqrt((dx * dx).sum() * (dy * dy).sum()))


def test_fc_with_outlier_volumes():
    b = _bold(5, 30, 1)
    b[:, [3, 17, 18]] = 0.0
    orig = b.copy()
    fc = utils.calculate_fc(b)
    np.testing.assert_array_equal(b, orig)
    z = _hand_zscore(orig)
    i, j = np.tril_indices(5, -1)
    ref = np.array([_pearson(z[a], z[c]) for a, c in zip(i, j)])
    np.testing.assert_allclose(fc, ref, rtol=0, atol=1e-13)
    keep = orig.sum(axis=0) != 0
    kept_only = np.array([_pearson(orig[a, keep], orig[c, keep]) for a, c in zip(i, j)])
    np.testing.assert_allclose(fc, kept_only, rtol=0, atol=1e-13)
    # whereas correlating the raw series, zeros included, gives other values
    raw = np.array([_pearson(orig[a], orig[c]) for a, c in zip(i, j)])
    assert np.abs(raw - fc).max() > 1e-3


def test_fc_exc_interhemispheric():
    b = _bold(6, 25, 2)
    fc = utils.calculate_fc(b, exc_interhemispheric=True)
    pairs = [(a, c) for a, c in zip(*np.tril_indices(6, -1)) if (a < 3) == (c < 3)]
    assert fc.shape == (len(pairs),) == (6,)
    ref = np.array([scipy.stats.pearsonr(b[a], b[c])[0] for a, c in pairs])
    np.testing.assert_allclose(fc, ref, rtol=0, atol=1e-13)


def test_fcd_windows_and_values():
    b = _bold(6, 20, 3)
    ws, step = 4, 3
    fcd, dfc = utils.calculate_fcd(b, ws, step, return_dfc=True)
    # centres 2, 5, 8, 11, 14, 17 (<= 20 - 1 - 2); windows [c-2, c+2]
    windows = [(c - 2, c + 2) for c in range(2, 18, 3)]
    assert dfc.shape == (6, 6, len(windows))
    i, j = np.tril_indices(6, -1)
    z = _hand_zscore(b)
    trils = []
    for w, (s0, s1) in enumerate(windows):
        t = np.array([_pearson(z[a, s0:s1 + 1], z[c, s0:s1 + 1]) for a, c in zip(i, j)])
        np.testing.assert_allclose(dfc[:, :, w][i, j], t, rtol=0, atol=1e-13)
        trils.append(t)
    wi, wj = np.tril_indices(len(windows), -1)
    ref = np.array([_pearson(trils[a], trils[c]) for a, c in zip(wi, wj)])
    assert fcd.shape == (len(windows) * (len(windows) - 1) // 2,)
    np.testing.assert_allclose(fcd, ref, rtol=0, atol=1e-13)


def test_fcd_keeps_edges_without_drop():
    b = _bold(4, 12, 4)
    _, dfc = utils.calculate_fcd(b, 4, 2, drop_edges=False, return_dfc=True)
    # centres 0, 2, ..., 10: edge windows are truncated to the series
    assert dfc.shape[2] == 6


def test_fcd_drops_outlier_windows():
    b = _bold(5, 30, 5)
    ws, step = 4, 1
    b[:, 10:13] = 0.0  # volumes 10, 11, 12
    _, dfc = utils.calculate_fcd(b.copy(), ws, step, return_dfc=True)
    centres = list(range(2, 28))
    kept = [c for c in centres if np.mean([(v in (10, 11, 12)) for v in range(c - 2, c + 3)]) <= 0.5]
    # windows centred on 10, 11 and 12 hold 3 of 5 outlier volumes and are
    # dropped; those centred on 9 and 13 hold 2 and are kept
    assert not {10, 11, 12} & set(kept) and {9, 13} <= set(kept)
    assert dfc.shape[2] == len(kept)
    with pytest.raises(AssertionError):
        utils.calculate_fcd(b, 5, 1)
