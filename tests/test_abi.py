"""C-ABI and CPython boundary checks that need no GPU.

* libcubnm_amd.so loads and exports every function include/cubnm_amd.h
  declares.
* cubnm_amd.core exposes the reference's three functions (core.cpp:147-192, 495-626)
  and raises (instead of returning None) on the reference's error paths that
  are decided on the host before any device work: use_cpu, unknown model,
  malformed arguments, too few BOLD volumes after burn-in, zero FCD windows,
  odd node count with inter-hemispheric exclusion (core.cpp:226-274,
  bnm.cpp:638-640, utils.cpp:9-14, 50-52).
"""
import ctypes
import os
import re

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "cubnm_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cubnm_\w+)\s*\(", src)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(os.path.join(ROOT, "cubnm_amd", "libcubnm_amd.so"))
    names = _declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_core_module_surface():
    import cubnm_amd.core as core
    for fn in ["run_simulations", "init", "set_const"]:
        assert callable(getattr(core, fn))
    core.init()
    core.set_const("k1", 3.72)
    core.set_const("unknown_key", 1.0)  # ignored, as upstream
    core.set_const("k1", 7 * 0.34)


def _args(model="rWW", nodes=10, duration=20, TR=1.0, **cfg_kw):
    sc = cases.synthetic_sc(nodes)
    gp, rp = cases.params(model, 2, nodes, G=[0.5, 1.0])
    cfg = cases.model_config(model, TR=TR, **{"window_size": 4, "window_step": 2, "bold_remove_s": 2, **cfg_kw})
    return list(cases.sim_args(model, sc, gp, rp, cfg, 2, duration_s=duration, TR_s=TR))


def test_use_cpu_is_rejected():
    import cubnm_amd.core as core
    a = _args()
    a[13] = True
    with pytest.raises(RuntimeError, match="use_cpu"):
        core.run_simulations(*a)


def test_unknown_model_raises():
    import cubnm_amd.core as core
    a = _args()
    a[0] = "NoSuchModel"
    with pytest.raises(RuntimeError, match="not found"):
        core.run_simulations(*a)


def test_bad_argument_tuple_raises():
    import cubnm_amd.core as core
    with pytest.raises(TypeError):
        core.run_simulations("rWW", 1, 2)
    a = _args()
    a[7] = "not a dict"
    with pytest.raises(TypeError):
        core.run_simulations(*a)


def test_too_few_volumes_raises():
    import cubnm_amd.core as core
    a = _args(duration=3, bold_remove_s=2)
    with pytest.raises(RuntimeError, match="too low for FC"):
        core.run_simulations(*a)


def test_zero_fcd_windows_raises():
    import cubnm_amd.core as core
    a = _args(duration=6, bold_remove_s=2)  # 4 volumes after burn-in, window of 5
    with pytest.raises(RuntimeError, match="windows is 0"):
        core.run_simulations(*a)


def test_odd_nodes_with_interhemispheric_exclusion_raises():
    import cubnm_amd.core as core
    a = _args(nodes=9, exc_interhemispheric=True)
    with pytest.raises(RuntimeError, match="not even"):
        core.run_simulations(*a)


def test_too_many_nodes_raises():
    import cubnm_amd.core as core
    a = _args(nodes=8193)
    with pytest.raises(RuntimeError, match="8192"):
        core.run_simulations(*a)


def test_short_parameter_arrays_raise():
    import cubnm_amd.core as core
    a = _args()
    a[5] = a[5][:2].copy()  # rWW needs 4 regional parameter rows
    with pytest.raises((ValueError, RuntimeError)):
        core.run_simulations(*a)


def test_sc_indices_beyond_scs_raise():
    import cubnm_amd.core as core
    a = _args()
    a[2] = np.array([0, 3], np.intc)
    with pytest.raises(ValueError, match="max\\(SC_indices\\)"):
        core.run_simulations(*a)


def test_run_groups_argument_errors():
    import cubnm_amd.core as core
    assert core.run_groups([]) == []
    with pytest.raises(TypeError):
        core.run_groups(5)
    with pytest.raises(TypeError):
        core.run_groups([[1, 2]])
    a = _args()
    a[13] = True
    with pytest.raises(RuntimeError, match="use_cpu"):
        core.run_groups([tuple(a)])


def test_tuning_keys_validated():
    import cubnm_amd.core as core
    for key, bad in [("batched_tpw", 3), ("general_ksplit", 9), ("general_ksplit", -1), ("fc_mode", 3)]:
        with pytest.raises(ValueError):
            core.set_tuning(key, bad)
    with pytest.raises(ValueError, match="unknown tuning key"):
        core.set_tuning("no_such_key", 0)
    for key in ["batched_tpw", "general_ksplit", "fc_mode"]:
        core.set_tuning(key, 0)
