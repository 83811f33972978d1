"""The allocation-free KS entry point (cubnm_gof_fcd_ks_device_static)
against the allocating one (cubnm_gof_fcd_ks_device), through the C ABI.

Both run the same kernels on the same inputs, and the statistic is a ratio
of exact integer counts, so the outputs must be identical to the bit.  Two
calls in a row, the second with a smaller input, show that the fixed
workspace's leftovers from a larger call do not leak into a smaller one; the
first call uses more simulations than fit the workspace at once (batches).
Also: sizes beyond the workspace are rejected with an error, and results
match SciPy.
"""
import numpy as np
import pytest
import scipy.stats

pytestmark = pytest.mark.gpu


def _case(n_sims, n, m, seed):
    rng = np.random.default_rng(seed)
    sim = np.tanh(rng.standard_normal((n_sims, n)) * 0.5 + rng.uniform(0, 0.5, (n_sims, 1)))
    emp = np.sort(np.tanh(rng.standard_normal(m) * 0.5 + 0.2))
    sim[1, 5] = np.nan
    return sim, emp


def _both(sim, emp):
    import torch
    from cubnm_amd import capi
    n_sims, n = sim.shape
    m = emp.size
    d_sim = torch.from_numpy(np.ascontiguousarray(sim)).cuda()
    d_emp = torch.from_numpy(np.ascontiguousarray(emp)).cuda()
    out_a = torch.full((n_sims,), -1.0, dtype=torch.float64, device="cuda")
    out_b = torch.full((n_sims,), -2.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    capi.gof_fcd_ks_device(d_sim.data_ptr(), n_sims, n, d_emp.data_ptr(), m, out_a.data_ptr())
    capi.gof_fcd_ks_device_static(d_sim.data_ptr(), n_sims, n, d_emp.data_ptr(), m, out_b.data_ptr(), None)
    torch.cuda.synchronize()
    return out_a.cpu().numpy(), out_b.cpu().numpy()


def test_static_matches_allocating_twice():
    # 40 x 74,691 values: 23.9 MB of sorted copies, more than the 16 MiB workspace
    sim, emp = _case(40, 74691, 300, 1)
    a, b = _both(sim, emp)
    np.testing.assert_array_equal(a, b)
    assert np.isnan(b[1]) and not np.isnan(b[0])
    ref = np.array([scipy.stats.ks_2samp(sim[i], emp).statistic for i in (0, 2, 39)])
    np.testing.assert_allclose(b[[0, 2, 39]], ref, rtol=0, atol=1e-15)
    # second call, smaller input
    sim2, emp2 = _case(7, 45, 300, 2)
    a2, b2 = _both(sim2, emp2)
    np.testing.assert_array_equal(a2, b2)
    ref2 = np.array([scipy.stats.ks_2samp(s, emp2).statistic for i, s in enumerate(sim2) if i != 1])
    np.testing.assert_allclose(np.delete(b2, 1), ref2, rtol=0, atol=1e-15)
    assert np.isnan(b2[1])


def test_static_rejects_oversized_rows():
    import torch
    from cubnm_amd import capi
    n = 2_097_151  # one more value than the 16 MiB workspace holds
    d_sim = torch.zeros((1, n), dtype=torch.float64, device="cuda")
    d_emp = torch.zeros(4, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="fixed workspace"):
        capi.gof_fcd_ks_device_static(d_sim.data_ptr(), 1, n, d_emp.data_ptr(), 4, out.data_ptr(), None)
