"""ctypes front end of the CPU oracle (oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the cubnm_amd package.

``run`` mirrors the argument meaning and returned dict of the reference's
``cubnm.core.run_simulations`` (src/ext/core.cpp:194-493) for the CPU path, so
tests can compare the two dicts key by key.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FRISTON2003 = (7 * 0.34, 2.0, 2 * 0.34 - 0.2, 0.02 * 100)


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-C", _HERE, "-s", "liboracle.so"], check=True)


def _bind(L):
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    L.oracle_last_error.restype = ctypes.c_char_p
    L.oracle_dims.argtypes = [ctypes.c_char_p, ctypes.c_char_p, I, I, I, I, D, D, P, P, P, P]
    L.oracle_dims.restype = I
    L.oracle_run.argtypes = ([ctypes.c_char_p, ctypes.c_char_p, I, I, I, P, P, P, P, P, P,
                              I, I, I, I, I, I, D, D, P, I] + [P] * 10)
    L.oracle_run.restype = I
    L.oracle_fic_root.argtypes = [D]
    L.oracle_fic_root.restype = D
    L.oracle_corr.argtypes = [P, P, ctypes.c_long]
    L.oracle_corr.restype = D
    return L


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _bind(ctypes.CDLL(path))
    return _LIB


def build_variant(source_text, out_dir, name="liboracle_variant.so"):
    """Compile a modified copy of oracle.cpp (same flags as the Makefile) into
    out_dir and return it bound like lib().  Tests use this to plant a mistake
    in the oracle and show that a pin rejects it."""
    import subprocess
    src = os.path.join(out_dir, "oracle_variant.cpp")
    with open(src, "w") as fh:
        fh.write(source_text)
    so = os.path.join(out_dir, name)
    subprocess.run(["g++", "-std=c++11", "-O3", "-m64", "-fopenmp", "-fPIC", "-DNOISE_SEGMENT", "-shared",
                    "-o", so, src], check=True)
    return _bind(ctypes.CDLL(so))


def source():
    with open(os.path.join(_HERE, "oracle.cpp")) as fh:
        return fh.read()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def config_string(config):
    return ";".join(f"{k}={v}" for k, v in config.items())


def dims(model, config, nodes, time_steps, BOLD_TR, states_sampling, dt, bw_dt, lib=lib):
    out = np.zeros(16, np.int32)
    ns = np.zeros(1, np.int64)
    rc = lib().oracle_dims(model.encode(), config_string(config).encode(), nodes, time_steps, BOLD_TR,
                           states_sampling, dt, bw_dt, _ptr(out), _ptr(ns), None, None)
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    keys = ["bold_len", "n_pairs", "n_windows", "n_window_pairs", "states_len", "n_state_vars",
            "n_global", "n_regional", "n_gout_bool", "n_gout_int", "noise_bw_it", "noise_repeats",
            "bw_it", "inner_it", "n_vols_remove", "corr_len"]
    d = {k: int(v) for k, v in zip(keys, out)}
    d["noise_size"] = int(ns[0])
    if d["n_windows"] > 0:
        ws = np.zeros(d["n_windows"], np.int32)
        we = np.zeros(d["n_windows"], np.int32)
        lib().oracle_dims(model.encode(), config_string(config).encode(), nodes, time_steps, BOLD_TR,
                          states_sampling, dt, bw_dt, _ptr(out), _ptr(ns), _ptr(ws), _ptr(we))
        d["window_starts"], d["window_ends"] = ws, we
    return d


def run(model, SC, SC_indices, SC_dist, global_params, regional_params, v_list, config,
        states_ts, noise_out, do_delay, N_SIMS, nodes, time_steps, BOLD_TR, states_sampling,
        sim_seed, dt, bw_dt, bw_consts=FRISTON2003, n_threads=0, lib=lib):
    """Run simulations on the CPU oracle.

    ``global_params`` and ``regional_params`` are float64 C-contiguous arrays
    modified in place (analytical/numerical FIC writes wIE), as the reference
    does when the model modifies parameters (core.cpp:437-450).
    """
    if isinstance(lib, ctypes.CDLL):
        _variant = lib
        lib = lambda: _variant  # noqa: E731
    d = dims(model, config, nodes, time_steps, BOLD_TR, states_sampling, dt, bw_dt, lib=lib)
    SC = np.ascontiguousarray(SC, np.float64).reshape(-1, nodes * nodes)
    SC_indices = np.ascontiguousarray(SC_indices, np.int32)
    SC_dist = np.ascontiguousarray(SC_dist, np.float64).reshape(-1)
    assert global_params.dtype == np.float64 and global_params.flags.c_contiguous
    assert regional_params.dtype == np.float64 and regional_params.flags.c_contiguous
    v_list = np.ascontiguousarray(v_list, np.float64)
    bold = np.zeros((N_SIMS, d["bold_len"] * nodes))
    fc = np.zeros((N_SIMS, d["n_pairs"]))
    fcd = np.zeros((N_SIMS, d["n_window_pairs"]))
    st_len = nodes * (d["states_len"] if states_ts else 1)
    states = np.zeros((d["n_state_vars"], N_SIMS, st_len))
    gb = np.zeros((d["n_gout_bool"], N_SIMS), np.uint8)
    gi = np.zeros((d["n_gout_int"], N_SIMS), np.int32)
    noise = shn = sht = None
    if noise_out:
        noise = np.zeros(d["noise_size"])
        shn = np.zeros((d["noise_repeats"], nodes), np.int32)
        sht = np.zeros((d["noise_repeats"], d["noise_bw_it"]), np.int32)
    times = np.zeros(2)
    bwc = np.asarray(bw_consts, np.float64)
    rc = lib().oracle_run(model.encode(), config_string(config).encode(), N_SIMS, nodes, SC.shape[0],
                          _ptr(SC), _ptr(SC_indices), _ptr(SC_dist), _ptr(global_params),
                          _ptr(regional_params), _ptr(v_list), int(states_ts), int(do_delay),
                          time_steps, BOLD_TR, states_sampling, sim_seed, dt, bw_dt, _ptr(bwc),
                          n_threads, _ptr(bold), _ptr(fc), _ptr(fcd), _ptr(states), _ptr(gb),
                          _ptr(gi), _ptr(noise), _ptr(shn), _ptr(sht), _ptr(times))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    out = {"init_time": float(times[0]), "run_time": float(times[1]), "sim_bold": bold,
           "_sim_states": states, "_global_bools": gb.astype(bool), "_global_ints": gi}
    cfg_fc = config.get("do_fc", "1") != "0"
    cfg_fcd = cfg_fc and config.get("do_fcd", "1") != "0"
    if cfg_fc:
        out["sim_fc_trils"] = fc
        if cfg_fcd:
            out["sim_fcd_trils"] = fcd
    if noise_out:
        out["_noise"], out["_shuffled_nodes"], out["_shuffled_ts"] = noise, shn, sht
    return out


def fic_root(w_EI):
    return lib().oracle_fic_root(w_EI)


def corr(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().oracle_corr(_ptr(x), _ptr(y), x.size)
