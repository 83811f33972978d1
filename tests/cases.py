"""Input builders shared by the tests, the smoke check and the benchmark.

``sim_args`` assembles the 22 positional arguments of run_simulations exactly
as the reference's SimGroup.run does (src/cubnm/sim/base.py:673-750) and the
model config strings as SimGroup._model_config / rWWSimGroup._model_config do
(base.py:635-657, rww.py:114-128).
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

MODELS = {
    # name: (global names, regional names, defaults, state names, n_noise)
    "rWW": (["G"], ["w_p", "J_N", "wIE", "sigma"],
            {"w_p": 1.4, "J_N": 0.15, "wIE": 1.0, "sigma": 0.01},
            ["I_E", "I_I", "r_E", "r_I", "S_E", "S_I"], 2),
    "JR": (["G"], ["J", "a_1", "a_2", "a_3", "a_4", "sigma"],
           {"J": 135.0, "a_1": 1.0, "a_2": 0.8, "a_3": 0.25, "a_4": 0.25, "sigma": 0.0001},
           ["y0", "y1", "y2", "y3", "y4", "y5", "y1_y2", "s_y1_y2"], 1),
    "Kuramoto": (["G"], ["init_theta", "omega", "sigma"],
                 {"init_theta": None, "omega": np.pi, "sigma": 0.17}, ["theta"], 1),
    "WC": (["G"], ["c_EE", "c_IE", "c_EI", "c_II", "P_E", "P_I", "sigma_E", "sigma_I"],
           {"c_EE": 16, "c_IE": 12, "c_EI": 15, "c_II": 3, "P_E": 0.5, "P_I": 0.0,
            "sigma_E": 5e-05, "sigma_I": 5e-05}, ["E", "I"], 2),
    "rWWEx": (["G"], ["w", "I0", "sigma"], {"w": 0.9, "I0": 0.3, "sigma": 0.001}, ["x", "r", "S"], 1),
}

FRISTON2003 = (7 * 0.34, 2.0, 2 * 0.34 - 0.2, 0.02 * 100)


def hcp_sc():
    """HCP group-train706 Schaefer-100 SC strength, normalised to mean 0.01
    (datasets.py:73-75), as used by the reference's regression fixtures."""
    m = np.load(os.path.join(GOLDEN, "hcp_sc_strength.npz"))["mat"].copy()
    m /= m.mean() * 100
    return m


def hcp_sc_length():
    return np.load(os.path.join(GOLDEN, "hcp_sc_length.npz"))["mat"].copy()


def synthetic_sc(nodes, seed=0):
    """Seeded dense SC: lognormal(0,1) weights, symmetrised, zero diagonal,
    scaled to max 1."""
    rng = np.random.default_rng(seed)
    w = rng.lognormal(0.0, 1.0, size=(nodes, nodes))
    w = (w + w.T) / 2
    np.fill_diagonal(w, 0.0)
    m = w.max()
    return w / m if m > 0 else w


def model_config(model, TR=1.0, window_size=10, window_step=2, bold_remove_s=30,
                 exc_interhemispheric=False, do_fc=True, do_fcd=True, drop_edges=True,
                 noise_segment_length=30, do_fic=True, max_fic_trials=0,
                 random_init_theta=True):
    ws_trs = int(np.round(window_size / (TR * 2))) * 2
    st_trs = int(np.round(window_step / TR))
    cfg = {
        "do_fc": str(int(do_fc)), "do_fcd": str(int(do_fcd)), "bold_remove_s": str(bold_remove_s),
        "exc_interhemispheric": str(int(exc_interhemispheric)), "window_size": str(ws_trs),
        "window_step": str(st_trs), "drop_edges": str(int(drop_edges)),
        "noise_time_steps": str(int(noise_segment_length * 1000)), "verbose": "0",
        "progress_interval": "500",
    }
    if model == "rWW":
        cfg.update({"do_fic": str(int(do_fic)), "max_fic_trials": str(max_fic_trials),
                    "I_SAMPLING_START": "1000", "I_SAMPLING_END": "10000", "init_delta": "0.02",
                    "fic_penalty_scale": "0.5"})
    if model == "Kuramoto":
        cfg["random_init_theta"] = str(int(random_init_theta))
    return cfg


def params(model, n_sims, nodes, G, sim_seed=0, do_fic=True, **regional):
    """Global (n_g, N) and regional (n_r, N*nodes) arrays with the SimGroup
    defaults (sim/base.py:494-545, rww.py:164-168, kuramoto.py:126-137)."""
    gnames, rnames, defaults, _, _ = MODELS[model]
    gp = np.ascontiguousarray(np.broadcast_to(np.asarray(G, float), (n_sims,)))[None, :].copy()
    rows = []
    for r in rnames:
        if r in regional:
            v = np.broadcast_to(np.asarray(regional[r], float), (n_sims, nodes))
        elif model == "rWW" and r == "wIE" and do_fic:
            v = np.zeros((n_sims, nodes))
        elif model == "Kuramoto" and r == "init_theta":
            th = np.random.default_rng(sim_seed).uniform(0, 2 * np.pi, nodes)
            v = np.tile(th, (n_sims, 1))
        else:
            v = np.full((n_sims, nodes), defaults[r], float)
        rows.append(np.ascontiguousarray(v, float).reshape(-1))
    return np.ascontiguousarray(gp, float), np.ascontiguousarray(np.vstack(rows), float)


def sim_args(model, sc, gp, rp, config, n_sims, duration_s=60, TR_s=1.0, dt="0.1", bw_dt="1.0",
             sc_dist=None, v=None, sim_seed=0, states_ts=False, noise_out=False,
             force_reinit=False, sc_indices=None, states_sampling_s=None):
    """The positional argument tuple of run_simulations (sim/base.py:727-750)."""
    sc = np.asarray(sc, float)
    nodes = sc.shape[-1]
    if sc.ndim == 2:
        sc2 = np.ascontiguousarray(sc.reshape(1, -1))
        sci = np.zeros(n_sims, np.intc)
    else:
        sc2 = np.ascontiguousarray(sc.reshape(sc.shape[0], -1))
        sci = np.ascontiguousarray(np.asarray(sc_indices, np.intc))
    do_delay = sc_dist is not None
    dist = np.zeros((nodes, nodes)) if sc_dist is None else np.asarray(sc_dist, float)
    vv = np.zeros(n_sims) if v is None else np.ascontiguousarray(np.broadcast_to(np.asarray(v, float), (n_sims,)))
    ss = TR_s if states_sampling_s is None else states_sampling_s
    return (model, sc2, sci, np.ascontiguousarray(dist.reshape(-1)), gp, rp, vv, config,
            True, states_ts, noise_out, do_delay, force_reinit, False, n_sims, nodes,
            int(round(duration_s * 1000)), int(round(TR_s * 1000)), int(round(ss * 1000)),
            sim_seed, float(dt), float(bw_dt))


def run_oracle(args, bw_consts=FRISTON2003, n_threads=0, lib=None):
    """Run the same argument tuple through the CPU oracle (test infrastructure)."""
    from oracle import oracle as O
    (model, sc2, sci, dist, gp, rp, vv, config, ext_out, states_ts, noise_out, do_delay, _fr, _cpu,
     n_sims, nodes, time_steps, TR, ss, seed, dt, bw_dt) = args
    return O.run(model, sc2, sci, dist.reshape(nodes, nodes), gp, rp, vv, config, states_ts, noise_out,
                 do_delay, n_sims, nodes, time_steps, TR, ss, seed, dt, bw_dt, bw_consts=bw_consts,
                 n_threads=n_threads, lib=O.lib if lib is None else lib)


def copy_args(args):
    """Fresh copies of the in/out parameter arrays (FIC writes them in place)."""
    a = list(args)
    a[4] = a[4].copy()
    a[5] = a[5].copy()
    return tuple(a)


def bold_rel_err(a, b):
    """Norm-wise relative BOLD error per simulation: max|a-b| / max|b|.
    NaN in both at the same position counts as equal; NaN in one is inf."""
    a = np.asarray(a).reshape(a.shape[0], -1)
    b = np.asarray(b).reshape(b.shape[0], -1)
    d = np.abs(a - b)
    both = np.isnan(a) & np.isnan(b)
    d[both] = 0.0
    d[np.isnan(a) ^ np.isnan(b)] = np.inf
    return d.max(axis=1) / np.nanmax(np.abs(b), axis=1)


def max_abs_err(a, b):
    """max|a-b| with NaN-at-the-same-place treated as equal."""
    d = np.abs(np.asarray(a) - np.asarray(b))
    d[np.isnan(a) & np.isnan(b)] = 0.0
    d[np.isnan(a) ^ np.isnan(b)] = np.inf
    return d.max() if d.size else 0.0
