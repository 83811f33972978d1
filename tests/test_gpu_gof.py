"""GOF reductions on the GPU vs SciPy (the reference scores with
scipy.stats.pearsonr and scipy.stats.ks_2samp, sim/base.py:1176-1195, or
with its device helpers utils.py:306-507, which follow the same formulas).

Tolerances: Pearson r within 1e-12 absolute (two-pass FP64 in both); the KS
statistic is a difference of two counts over sample sizes, so it must be
equal to SciPy's to the last bit (1e-15 allows for its final division
order).  Ties within and across samples, non-power-of-two sizes, NaN rows
and sizes past one LDS sort chunk (8,192 values: C5's 74,305 FCD values per
simulation, and a NaN in a later chunk) are covered; NaN gives NaN (callers
map it to the worst score, sim/base.py:37-42).
"""
import numpy as np
import pytest
import scipy.stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def core():
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    return _core


def _pearson_ref(sim, emp):
    return np.array([scipy.stats.pearsonr(s, emp)[0] for s in sim])


def _ks_ref(sim, emp):
    return np.array([scipy.stats.ks_2samp(s, emp).statistic for s in sim])


@pytest.mark.parametrize("n_sims,n_pairs", [(1, 3), (37, 4950), (5, 79800)])
def test_fc_corr_matches_scipy(core, n_sims, n_pairs):
    rng = np.random.default_rng(n_pairs)
    emp = np.tanh(rng.standard_normal(n_pairs))
    sim = np.tanh(0.6 * emp + rng.standard_normal((n_sims, n_pairs)) + rng.uniform(-1, 1, (n_sims, 1)))
    np.testing.assert_allclose(core.gof_fc_corr(sim, emp), _pearson_ref(sim, emp), rtol=0, atol=1e-12)


def test_fc_corr_nan_row(core):
    rng = np.random.default_rng(1)
    emp = rng.standard_normal(45)
    sim = rng.standard_normal((3, 45))
    sim[1, 7] = np.nan
    r = core.gof_fc_corr(sim, emp)
    assert np.isnan(r[1]) and not np.isnan(r[0]) and not np.isnan(r[2])


@pytest.mark.parametrize("n_sim_values,n_emp", [(1, 1), (45, 300), (46, 300), (1000, 77), (4095, 4950)])
def test_fcd_ks_matches_scipy(core, n_sim_values, n_emp):
    rng = np.random.default_rng(n_sim_values + n_emp)
    emp = rng.uniform(-0.2, 1.0, n_emp)
    sim = rng.uniform(-0.5, 1.0, (9, n_sim_values)) + rng.uniform(-0.3, 0.3, (9, 1))
    np.testing.assert_allclose(core.gof_fcd_ks(sim, emp), _ks_ref(sim, emp), rtol=0, atol=1e-15)


@pytest.mark.parametrize("n_sim_values,n_emp", [(8192, 300), (8193, 1000), (74305, 74305), (74691, 300)])
def test_fcd_ks_large_matches_scipy(core, n_sim_values, n_emp):
    """Several sorted chunks per simulation (C5: 386-387 windows give
    74,305-74,691 FCD values per simulation)."""
    rng = np.random.default_rng(n_sim_values)
    emp = np.tanh(rng.standard_normal(n_emp) * 0.5 + 0.3)
    sim = np.tanh(rng.standard_normal((3, n_sim_values)) * 0.5 + rng.uniform(0, 0.6, (3, 1)))
    sim[2, ::7] = np.round(sim[2, ::7], 2)  # ties across chunks
    np.testing.assert_allclose(core.gof_fcd_ks(sim, emp), _ks_ref(sim, emp), rtol=0, atol=1e-15)


def test_fcd_ks_large_nan_in_late_chunk(core):
    rng = np.random.default_rng(11)
    emp = rng.standard_normal(500)
    sim = rng.standard_normal((2, 20000))
    sim[1, 19999] = np.nan
    ks = core.gof_fcd_ks(sim, emp)
    assert not np.isnan(ks[0]) and np.isnan(ks[1])
    assert ks[0] == pytest.approx(scipy.stats.ks_2samp(sim[0], emp).statistic, abs=1e-15)


def test_fcd_ks_with_ties(core):
    rng = np.random.default_rng(5)
    emp = np.round(rng.uniform(0, 1, 300), 1)          # 11 distinct values
    sim = np.round(rng.uniform(0, 1, (16, 45)), 1)     # ties within and across samples
    sim[3] = emp[:45]
    sim[4] = 0.5                                       # constant row
    np.testing.assert_allclose(core.gof_fcd_ks(sim, emp), _ks_ref(sim, emp), rtol=0, atol=1e-15)


def test_fcd_ks_nan_and_identical(core):
    rng = np.random.default_rng(6)
    emp = rng.standard_normal(45)
    sim = np.stack([emp, emp[::-1], rng.standard_normal(45)])
    sim[2, 0] = np.nan
    ks = core.gof_fcd_ks(sim, emp)
    assert ks[0] == 0.0 and ks[1] == 0.0 and np.isnan(ks[2])


def test_device_forms_match_host(core):
    import torch
    from cubnm_amd import utils
    rng = np.random.default_rng(7)
    emp_fc, emp_fcd = rng.standard_normal(4950), np.sort(rng.standard_normal(300))
    sim_fc, sim_fcd = rng.standard_normal((64, 4950)), rng.standard_normal((64, 45))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    out_r = torch.empty(64, dtype=torch.float64, device="cuda")
    out_k = torch.empty(64, dtype=torch.float64, device="cuda")
    from cubnm_amd import capi
    a, b, c, d = dev(sim_fc), dev(emp_fc), dev(sim_fcd), dev(emp_fcd)
    torch.cuda.synchronize()
    capi.gof_fc_corr_device(a.data_ptr(), 64, 4950, b.data_ptr(), out_r.data_ptr())
    capi.gof_fcd_ks_device(c.data_ptr(), 64, 45, d.data_ptr(), 300, out_k.data_ptr())
    np.testing.assert_array_equal(out_r.cpu().numpy(), utils.fc_corr_device(sim_fc, emp_fc))
    np.testing.assert_array_equal(out_k.cpu().numpy(), utils.fcd_ks_device(sim_fcd, emp_fcd))
    np.testing.assert_allclose(out_k.cpu().numpy(), _ks_ref(sim_fcd, emp_fcd), rtol=0, atol=1e-15)
