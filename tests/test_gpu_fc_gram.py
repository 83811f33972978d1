"""FC/FCD as Gram products on the matrix cores (fc_mode 1: FP64 MFMA,
fc_mode 2: split-BF16 MFMA) against FP64 NumPy on the same simulated BOLD.

The expected values restate the reference's definitions (fc.cpp:1-91):
Pearson r over the post-burn-in volumes for every pair (i > j, row-major,
optionally same-hemisphere only), Pearson r inside each window, and Pearson
r between window-FC vectors; the window table is the oracle's (pinned by the
reference fixtures).  Tolerances: FP64 Gram 1e-10 absolute (summation order
only: NumPy and the MFMA K loop add the 79,800 pair products of an N=400 FCD
entry in different orders; observed 4e-12); split-BF16 Gram (static FC and
FCD in 3-term BF16, window FCs in FP64), stated per result and asserted
here: FC 1e-4, FCD 1e-4 absolute.  Operands carry ~2^-16 relative error and
each entry averages it over K products; measured on MI355X with 40 volumes
at TR 1 s: FC 6e-6 - 1e-5, FCD 3e-7 (N=100) - 4e-6 (N=8); errors are
printed.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _expected(bold, nvr, ws, we, exc_inter):
    T, N = bold.shape
    i, j = np.tril_indices(N, -1)
    if exc_inter:
        keep = (i < N // 2) == (j < N // 2)
        i, j = i[keep], j[keep]
    fc = np.corrcoef(bold[nvr:].T)[i, j]
    wfc = np.array([np.corrcoef(bold[s:e + 1].T)[i, j] for s, e in zip(ws, we)])
    a, b = np.tril_indices(len(ws), -1)
    fcd = np.corrcoef(wfc)[a, b]
    return fc, fcd


def _run(core, nodes, n_sims, duration, TR, window, step, nvr_s, exc_inter, fc_mode):
    from oracle import oracle as O
    gp, rp = cases.params("rWW", n_sims, nodes, G=np.linspace(0.5, 1.5, n_sims), do_fic=False, wIE=1.0)
    cfg = cases.model_config("rWW", TR=TR, bold_remove_s=nvr_s, exc_interhemispheric=exc_inter, do_fic=False)
    cfg.update({"window_size": str(window), "window_step": str(step)})
    args = cases.sim_args("rWW", cases.synthetic_sc(nodes, seed=11), gp, rp, cfg, n_sims, duration_s=duration,
                          TR_s=TR)
    core.core.set_tuning("fc_mode", fc_mode)
    try:
        out = core.run_simulations(*args)
        path = core.core.last_timings()["fc_path"]
    finally:
        core.core.set_tuning("fc_mode", 0)
    d = O.dims("rWW", cfg, nodes, args[16], args[17], args[18], 0.1, 1.0)
    return out, path, d


@pytest.fixture(scope="module")
def core():
    import cubnm_amd  # noqa: F401
    from cubnm_amd import _core
    return _core


@pytest.mark.parametrize("nodes,exc_inter", [(24, False), (130, True), (400, False)])
def test_fp64_gram_matches_numpy(core, nodes, exc_inter):
    duration, TR = (30, 0.72) if nodes == 400 else (20, 1.0)
    out, path, d = _run(core, nodes, 3, duration, TR, 10, 1 if nodes == 400 else 2, 0 if nodes == 400 else 2,
                        exc_inter, 1)
    assert path == 1
    bold = out["sim_bold"].reshape(3, d["bold_len"], nodes)
    for s in range(3):
        fc, fcd = _expected(bold[s], d["n_vols_remove"], d["window_starts"], d["window_ends"], exc_inter)
        np.testing.assert_allclose(out["sim_fc_trils"][s], fc, rtol=0, atol=1e-10)
        np.testing.assert_allclose(out["sim_fcd_trils"][s], fcd, rtol=0, atol=1e-10)


def test_fp64_gram_equals_lds_kernel(core):
    a, pa, _ = _run(core, 100, 20, 40, 1.0, 10, 2, 10, False, 0)
    b, pb, _ = _run(core, 100, 20, 40, 1.0, 10, 2, 10, False, 1)
    assert (pa, pb) == (0, 1)
    np.testing.assert_array_equal(a["sim_bold"], b["sim_bold"])
    np.testing.assert_allclose(b["sim_fc_trils"], a["sim_fc_trils"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(b["sim_fcd_trils"], a["sim_fcd_trils"], rtol=0, atol=1e-13)


@pytest.mark.parametrize("nodes", [24, 100, 400])
def test_split_bf16_gram_error(core, nodes):
    # the FCD product runs in one K slice at 24 nodes (276 pairs), in 4 K
    # slices at 100 and in 16 at 400 (partial sums added in slice order)
    duration, TR = (40, 0.72) if nodes == 400 else (40, 1.0)
    out, path, d = _run(core, nodes, 2, duration, TR, 10, 1, 0, False, 2)
    assert path == 2
    bold = out["sim_bold"].reshape(2, d["bold_len"], nodes)
    e_fc = e_fcd = 0.0
    for s in range(2):
        fc, fcd = _expected(bold[s], d["n_vols_remove"], d["window_starts"], d["window_ends"], False)
        e_fc = max(e_fc, np.abs(out["sim_fc_trils"][s] - fc).max())
        e_fcd = max(e_fcd, np.abs(out["sim_fcd_trils"][s] - fcd).max())
    print(f"split-BF16 N={nodes}: max |FC err| {e_fc:.2e}, max |FCD err| {e_fcd:.2e}")
    assert e_fc <= 1e-4 and e_fcd <= 1e-4
    assert e_fc > 0   # the BF16 path really ran
