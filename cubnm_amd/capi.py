"""ctypes binding of the C ABI declared in include/cubnm_amd.h.

The CPython module ``cubnm_amd.core`` is the drop-in for the reference's
``cubnm.core``; this binding exposes the same library to callers that keep
their arrays in device memory (``run_device``), e.g. torch CUDA tensors.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Dims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "bold_len", "n_pairs", "n_windows", "n_window_pairs", "states_len", "n_state_vars",
        "n_global_params", "n_regional_params", "n_global_out_bool", "n_global_out_int",
        "noise_bw_it", "noise_repeats", "bw_it", "inner_it", "n_vols_remove", "corr_len",
        "do_fc", "do_fcd", "modifies_params", "N_SIMS", "nodes")] + [
        ("noise_size", ctypes.c_longlong), ("init_time", ctypes.c_double)]


class Outputs(ctypes.Structure):
    _fields_ = [("sim_bold", ctypes.c_void_p), ("sim_fc_trils", ctypes.c_void_p),
                ("sim_fcd_trils", ctypes.c_void_p), ("sim_states", ctypes.c_void_p),
                ("global_bools", ctypes.c_void_p), ("global_ints", ctypes.c_void_p),
                ("noise", ctypes.c_void_p), ("shuffled_nodes", ctypes.c_void_p),
                ("shuffled_ts", ctypes.c_void_p), ("run_time", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    from . import build as _build
    path = _build.ensure_built()
    L = ctypes.CDLL(path)
    I, D, P, C = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
    CP = ctypes.POINTER(ctypes.c_char_p)
    L.cubnm_last_error.restype = C
    L.cubnm_init.restype = I
    L.cubnm_set_const.argtypes = [C, D]
    L.cubnm_prepare.argtypes = [C, I, CP, CP] + [I] * 13 + [D, D, ctypes.POINTER(Dims)]
    L.cubnm_run.argtypes = [P] * 6 + [ctypes.POINTER(Outputs)]
    L.cubnm_run_device.argtypes = [P] * 6 + [ctypes.POINTER(Outputs)]
    L.cubnm_last_timings.argtypes = [ctypes.POINTER(ctypes.c_double * 6)]
    L.cubnm_force_integrator.argtypes = [I]
    L.cubnm_last_integrator.restype = I
    L.cubnm_last_launch_shape.argtypes = [ctypes.POINTER(ctypes.c_int * 5)]
    L.cubnm_set_tuning.argtypes = [C, I]
    L.cubnm_gof_fc_corr.argtypes = [P, I, I, P, P]
    L.cubnm_gof_fcd_ks.argtypes = [P, I, I, P, I, P]
    L.cubnm_gof_fc_corr_device.argtypes = [P, I, I, P, P]
    L.cubnm_gof_fcd_ks_device.argtypes = [P, I, I, P, I, P]
    L.cubnm_gof_fcd_ks_device_static.argtypes = [P, I, I, P, I, P, P]
    LL = ctypes.POINTER(ctypes.c_longlong)
    L.cubnm_copy_noise.argtypes = [P, P, P]
    L.cubnm_rww_grad_sizes.argtypes = [I, I, I, I, D, D, I, LL, LL, LL]
    L.cubnm_rww_grad_forward.argtypes = [I, I, I, I, D, D, I, I] + [P] * 11
    L.cubnm_rww_grad_backward.argtypes = [I, I, I, I, D, D, I, I] + [P] * 15
    L.cubnm_fc_fcd_loss_scratch.argtypes = [I, I, I, I, I]
    L.cubnm_fc_fcd_loss_scratch.restype = ctypes.c_longlong
    L.cubnm_fc_fcd_loss.argtypes = [P, I, I, I, I, P, P, I, I, P, P, P, I, P, P, P, P]
    L.cubnm_init()
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().cubnm_last_error().decode())


def prepare(model, config, N_SIMS, nodes, time_steps, BOLD_TR, states_sampling, sim_seed, dt, bw_dt,
            N_SCs=1, ext_out=True, states_ts=False, noise_out=False, do_delay=False, force_reinit=False):
    keys = (ctypes.c_char_p * len(config))(*[k.encode() for k in config])
    vals = (ctypes.c_char_p * len(config))(*[str(v).encode() for v in config.values()])
    d = Dims()
    _check(lib().cubnm_prepare(model.encode(), len(config), keys, vals, int(ext_out), int(states_ts),
                               int(noise_out), int(do_delay), int(force_reinit), 0, N_SIMS, nodes, N_SCs,
                               time_steps, BOLD_TR, states_sampling, sim_seed, float(dt), float(bw_dt),
                               ctypes.byref(d)))
    return d


def run_device(SC, SC_indices, SC_dist, global_params, regional_params, v_list, bold, fc=None, fcd=None,
               states=None, global_bools=None, global_ints=None):
    """All arguments are device addresses (ints), e.g. tensor.data_ptr()."""
    o = Outputs(bold, fc, fcd, states, global_bools, global_ints, None, None, None, 0.0)
    _check(lib().cubnm_run_device(SC, SC_indices, SC_dist, global_params, regional_params, v_list, ctypes.byref(o)))
    return o.run_time


def last_timings():
    t = (ctypes.c_double * 6)()
    lib().cubnm_last_timings(ctypes.byref(t))
    keys = ["prep_ms", "h2d_ms", "integrate_ms", "fc_fcd_ms", "d2h_ms", "integrate_launches"]
    return dict(zip(keys, list(t)))


def set_tuning(key, value):
    _check(lib().cubnm_set_tuning(key.encode(), int(value)))


def last_launch_shape():
    """Launch shape of the last run: integrator, tiles per wave, SC fragments
    in LDS, k partitions, chosen by (0 rule, 1 autotuner trials, 2 cache)."""
    o = (ctypes.c_int * 5)()
    lib().cubnm_last_launch_shape(ctypes.byref(o))
    keys = ["integrator", "tiles_per_wave", "sc_in_lds", "ksplit", "chosen_by"]
    return dict(zip(keys, list(o)))


def force_integrator(k):
    lib().cubnm_force_integrator(int(k))


def gof_fc_corr_device(sim_fc, n_sims, n_pairs, emp_fc, out):
    _check(lib().cubnm_gof_fc_corr_device(sim_fc, n_sims, n_pairs, emp_fc, out))


def gof_fcd_ks_device(sim_fcd, n_sims, n, emp_sorted, m, out):
    _check(lib().cubnm_gof_fcd_ks_device(sim_fcd, n_sims, n, emp_sorted, m, out))


def gof_fcd_ks_device_static(sim_fcd, n_sims, n, emp_sorted, m, out, stream=None):
    """Allocation-free KS (fixed per-GPU workspace); queued on ``stream``
    (a hipStream_t address, None = null stream) without waiting."""
    _check(lib().cubnm_gof_fcd_ks_device_static(sim_fcd, n_sims, n, emp_sorted, m, out, stream))


def copy_noise(noise, shuffled_nodes, shuffled_ts):
    """Host copies (numpy arrays of the sizes in the prepared Dims) of the
    noise segment and shuffles of the last prepare()."""
    _check(lib().cubnm_copy_noise(noise.ctypes.data, shuffled_nodes.ctypes.data, shuffled_ts.ctypes.data))


def rww_grad_sizes(n_sims, nodes, time_steps, BOLD_TR, dt, bw_dt, ck):
    b, c, s = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    _check(lib().cubnm_rww_grad_sizes(n_sims, nodes, time_steps, BOLD_TR, float(dt), float(bw_dt), ck,
                                      ctypes.byref(b), ctypes.byref(c), ctypes.byref(s)))
    return b.value, c.value, s.value


def rww_grad_forward(n_sims, nodes, time_steps, BOLD_TR, dt, bw_dt, ck, noise_time_steps, *ptrs):
    """ptrs: SC, G, w_p, J_N, wIE, sigma, noise, shuffled_nodes, shuffled_ts, bold, ckpt (device addresses)."""
    _check(lib().cubnm_rww_grad_forward(n_sims, nodes, time_steps, BOLD_TR, float(dt), float(bw_dt), ck,
                                        noise_time_steps, *ptrs))


def rww_grad_backward(n_sims, nodes, time_steps, BOLD_TR, dt, bw_dt, ck, noise_time_steps, *ptrs):
    """ptrs: SC, G, w_p, J_N, wIE, sigma, noise, shuffled_nodes, shuffled_ts, ckpt, dbold, scratch,
    grad_G, grad_w_p, grad_J_N (device addresses)."""
    _check(lib().cubnm_rww_grad_backward(n_sims, nodes, time_steps, BOLD_TR, float(dt), float(bw_dt), ck,
                                         noise_time_steps, *ptrs))


def fc_fcd_loss_scratch(bold_len, nodes, n_vols_remove, n_windows, max_window_len):
    """Scratch doubles per simulation of fc_fcd_loss."""
    return lib().cubnm_fc_fcd_loss_scratch(bold_len, nodes, n_vols_remove, n_windows, max_window_len)


def fc_fcd_loss(bold, n_sims, bold_len, nodes, n_vols_remove, win_start, win_end, n_windows, max_window_len,
                emp_fc, emp_q, qs, n_quantiles, loss, dbold, scratch, stream=None):
    """Regional-fit loss (and, with dbold not None, its gradient) of device
    BOLD; all array arguments are device addresses."""
    _check(lib().cubnm_fc_fcd_loss(bold, n_sims, bold_len, nodes, n_vols_remove, win_start, win_end, n_windows,
                                   max_window_len, emp_fc, emp_q, qs, n_quantiles, loss, dbold, scratch, stream))
