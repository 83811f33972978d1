"""SimGroup-compatible caller of the simulation core.

A SimGroup holds the settings of a group of simulations (timing, SC, FC/FCD
options, noise) and their parameters, runs them through
``cubnm_amd.core.run_simulations`` in one call and exposes the outputs with
the reference's attribute names and shapes, so code written against the
reference's ``cubnm.sim`` classes runs unchanged:

  run marshalling and reinit rule   reference sim/base.py:673-766
  output attributes and shapes      sim/base.py:768-818
  get_noise                         sim/base.py:845-887
  get_sim_fc / get_sim_fcd          sim/base.py:889-960
  score (GOF terms, NaN -> worst)   sim/base.py:37-42, 1084-1207
  save (outputs and parameters)     sim/base.py:1209-1239
  multi-group merge and split       sim/base.py:1303-1438

Differences: the core has no CPU engine, so ``force_cpu=True`` only affects
``score`` (SciPy instead of the GPU reductions) and ``run`` raises; nodes are
limited to 1024 (no co-launch).
"""
import copy
import gc
import logging
import os
from decimal import Decimal

import numpy as np

from .. import _core, utils

logger = logging.getLogger(__name__)

# NaN scores are replaced by the worst value of their term
WORST_SCORES = {"+fc_corr": -1.0, "-fc_diff": -2.0, "-fc_normec": -1.0, "-fcd_ks": -1.0}

MAX_NODES = 1024


def _load_matrix(m):
    if isinstance(m, (str, os.PathLike)):
        path = os.fspath(m)
        if path.endswith(".txt"):
            return np.loadtxt(path)
        if path.endswith(".npy"):
            return np.load(path)
        raise ValueError(f"unsupported matrix file {path!r} (.txt or .npy)")
    return np.asarray(m, dtype=float)


class SimGroup:
    """Group of simulations of one model run together on the GPU.

    Arguments and defaults are the reference's (sim/base.py:45-71); times are
    in seconds except ``dt``/``bw_dt`` (milliseconds, exact decimals).
    Subclasses set ``model_name``, ``global_param_names``,
    ``regional_param_names``, ``state_names``, ``sel_state_var``, ``n_noise``
    and ``default_params``.
    """
    model_name = None
    global_param_names = []
    regional_param_names = []
    state_names = []
    sel_state_var = None
    n_noise = 1
    default_params = {}

    def __init__(self, duration, TR, sc, sc_dist=None, out_dir=None, dt="0.1", bw_dt="1.0",
                 states_ts=False, states_sampling=None, noise_out=False, do_fc=True, do_fcd=True,
                 window_size=30, window_step=5, sim_seed=0, exc_interhemispheric=False,
                 force_cpu=False, force_gpu=False, gof_terms=("+fc_corr", "-fcd_ks"),
                 bw_params="friston2003", bold_remove_s=30, fcd_drop_edges=True,
                 noise_segment_length=30, sim_verbose=False, progress_interval=500):
        self.input_sc = sc
        self.sc = _load_matrix(sc)
        if self.sc.ndim == 2 and np.any(np.diag(self.sc)):
            logger.warning("The diagonal of the SC matrix is not 0; self-connections are simulated.")
        self.nodes = self.sc.shape[-1]
        if self.nodes > MAX_NODES:
            raise NotImplementedError(f"at most {MAX_NODES} nodes are supported")
        self.input_sc_dist = sc_dist
        self.do_delay = sc_dist is not None
        self.sc_dist = np.zeros((self.nodes, self.nodes)) if sc_dist is None else _load_matrix(sc_dist)
        self.duration = duration
        self.TR = TR
        self._dt, self._bw_dt = Decimal(dt), Decimal(bw_dt)
        self._check_dt()
        self.states_ts = states_ts
        self.states_sampling = states_sampling
        self.noise_out = noise_out
        self.gof_terms = list(gof_terms)
        self._do_fc, self._do_fcd = bool(do_fc), bool(do_fcd)
        self._check_fc_flags()
        self.window_size, self.window_step = window_size, window_step
        self.sim_seed = sim_seed
        self.exc_interhemispheric = exc_interhemispheric
        self.force_cpu, self.force_gpu = force_cpu, force_gpu
        self.bold_remove_s = bold_remove_s
        self.fcd_drop_edges = fcd_drop_edges
        self.noise_segment_length = noise_segment_length
        self.sim_verbose, self.progress_interval = sim_verbose, progress_interval
        self.avail_gpus, self.avail_cpus = utils.avail_gpus(), os.cpu_count()   # base.py:246-248
        self.bw_params = bw_params
        self.out_dir = self.input_out_dir = out_dir
        self.param_lists = {k: None for k in self.global_param_names + self.regional_param_names + ["v"]}
        self._N = 0
        self.last_N = 0
        self.last_config = self.get_config(for_reinit=True)
        self.it = 0
        self.post_init()

    def post_init(self):
        """Hook run at the end of construction; subclasses may extend it
        (base.py:286-294)."""

    # ---- timing -----------------------------------------------------------
    @property
    def duration(self):
        return self._duration

    @duration.setter
    def duration(self, duration):
        ms = Decimal(str(duration)) * 1000
        if ms % 1 != 0:
            raise ValueError(f"duration {duration} s is not a whole number of milliseconds")
        self._duration = duration
        self.duration_msec = int(ms)

    @property
    def TR(self):
        return self._TR

    @TR.setter
    def TR(self, TR):
        self._TR = TR
        self.TR_msec = int(TR * 1000)

    @property
    def states_sampling(self):
        return self.TR if self._states_sampling is None else self._states_sampling

    @states_sampling.setter
    def states_sampling(self, v):
        if v is not None and v < 0.001:
            raise ValueError("states_sampling cannot be lower than 0.001 s")
        self._states_sampling = v

    @property
    def states_sampling_msec(self):
        return int(self.states_sampling * 1000)

    @property
    def n_states_samples_remove(self):
        return int(self.bold_remove_s / self.states_sampling)

    @property
    def noise_segment_length(self):
        return self.duration if self._noise_segment_length is None else self._noise_segment_length

    @noise_segment_length.setter
    def noise_segment_length(self, v):
        self._noise_segment_length = v

    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, v):
        self._dt = Decimal(v)
        self._check_dt()

    @property
    def bw_dt(self):
        return self._bw_dt

    @bw_dt.setter
    def bw_dt(self, v):
        self._bw_dt = Decimal(v)
        self._check_dt()

    @property
    def bw_it(self):
        return int(self.duration_msec / self.bw_dt)

    @property
    def inner_it(self):
        return int(self.bw_dt / self.dt)

    def _check_dt(self):
        if not self.bw_dt >= self.dt:
            raise ValueError("Balloon-Windkessel dt must be >= model dt")
        if self.bw_dt % self.dt != 0:
            raise ValueError("Balloon-Windkessel dt must be divisible by model dt")
        if self.duration_msec % self.bw_dt != 0:
            raise ValueError("duration must be divisible by Balloon-Windkessel dt")

    @property
    def window_size_TRs(self):
        """Window size in volumes: nearest even number of TRs (+1 centre volume in use)."""
        return int(np.round(self.window_size / (self.TR * 2))) * 2

    @property
    def window_step_TRs(self):
        return int(np.round(self.window_step / self.TR))

    # ---- FC / FCD flags -----------------------------------------------------
    def _check_fc_flags(self):
        if self._do_fcd and not self._do_fc:
            raise ValueError("Cannot calculate FCD without FC")
        if not self._do_fc and {"+fc_corr", "-fc_diff", "-fc_normec"} & set(self.gof_terms):
            raise ValueError("FC goodness-of-fit terms need do_fc=True")
        if not self._do_fcd and "-fcd_ks" in self.gof_terms:
            raise ValueError("FCD goodness-of-fit terms need do_fcd=True")

    @property
    def do_fc(self):
        return self._do_fc

    @do_fc.setter
    def do_fc(self, v):
        old = self._do_fc
        self._do_fc = bool(v)
        try:
            self._check_fc_flags()
        except ValueError:
            self._do_fc = old
            raise

    @property
    def do_fcd(self):
        return self._do_fcd

    @do_fcd.setter
    def do_fcd(self, v):
        old = self._do_fcd
        self._do_fcd = bool(v)
        try:
            self._check_fc_flags()
        except ValueError:
            self._do_fcd = old
            raise

    # ---- Balloon-Windkessel constants ---------------------------------------
    @property
    def bw_params(self):
        return self._bw_params

    @bw_params.setter
    def bw_params(self, src):
        self._bw_params = src
        p = src if isinstance(src, dict) else utils.get_bw_params(src)
        for k in ("k1", "k2", "k3"):
            _core.set_const(k, float(p[k]))

    # ---- parameters ----------------------------------------------------------
    @property
    def N(self):
        return self._N

    @N.setter
    def N(self, N):
        self._N = int(N)
        if not self.do_delay:
            self.param_lists["v"] = np.zeros(self._N)

    def _set_default_params(self, missing=True):
        """Fill parameters that are unset (or of the wrong length) with the
        model defaults; with missing=False overwrite all of them."""
        defaults = dict(self.default_params)
        if self.do_delay and "v" not in defaults:
            defaults["v"] = 0.5
        for name, v in defaults.items():
            cur = self.param_lists.get(name)
            if missing and cur is not None and np.shape(cur)[0] == self.N:
                continue
            if v is None:
                raise ValueError(f"Parameter {name} has no default and must be set")
            if name in self.global_param_names or name == "v":
                self.param_lists[name] = np.repeat(float(v), self.N)
            elif not self._skip_default(name):
                self.param_lists[name] = np.full((self.N, self.nodes), float(v))

    def _skip_default(self, name):
        return False

    # ---- configuration -------------------------------------------------------
    def get_config(self, include_N=False, for_reinit=False):
        config = {
            "duration": self.duration, "TR": self.TR, "nodes": self.nodes, "states_ts": self.states_ts,
            "states_sampling": self.states_sampling, "do_fc": self.do_fc, "do_fcd": self.do_fcd,
            "window_size": self.window_size, "window_step": self.window_step,
            "window_size_TRs": self.window_size_TRs, "window_step_TRs": self.window_step_TRs,
            "sim_seed": self.sim_seed, "exc_interhemispheric": self.exc_interhemispheric,
            "force_cpu": self.force_cpu, "force_gpu": self.force_gpu, "hardware_used": "gpu",
            "bold_remove_s": self.bold_remove_s, "fcd_drop_edges": self.fcd_drop_edges,
            "noise_segment_length": self.noise_segment_length,
        }
        if not for_reinit:
            config.update(out_dir=self.input_out_dir, gof_terms=self.gof_terms, sc=self.input_sc,
                          sc_dist=self.input_sc_dist, bw_params=self.bw_params)
            if include_N:
                config["N"] = self.N
        return config

    @property
    def _model_config(self):
        """String config passed to the core (keys of base.hpp:271-297)."""
        return {
            "do_fc": str(int(self.do_fc)), "do_fcd": str(int(self.do_fcd)),
            "bold_remove_s": str(self.bold_remove_s),
            "exc_interhemispheric": str(int(self.exc_interhemispheric)),
            "window_size": str(self.window_size_TRs), "window_step": str(self.window_step_TRs),
            "drop_edges": str(int(self.fcd_drop_edges)),
            "noise_time_steps": str(int(self.noise_segment_length * 1000)),
            "verbose": str(int(self.sim_verbose)), "progress_interval": str(int(self.progress_interval)),
        }

    @property
    def use_cpu(self):
        return False

    # ---- running -------------------------------------------------------------
    def _sc_arrays(self):
        if self.sc.ndim == 2 or self.sc.shape[0] == 1:
            sc = np.ascontiguousarray(self.sc.reshape(1, -1), dtype=float)
            sci = np.zeros(self.N, np.intc)
        else:
            sc = np.ascontiguousarray(self.sc.reshape(self.sc.shape[0], -1), dtype=float)
            sci = np.ascontiguousarray(np.asarray(self.sc_indices), dtype=np.intc)
        used = np.unique(sci)
        if sci.size != self.N or sc.shape[0] != used.size or not np.array_equal(used, np.arange(used.size)):
            raise ValueError("sc_indices must have N entries and use every SC 0..n_SCs-1")
        return sc, sci

    def _run_args(self, force_reinit=False):
        """The positional arguments of run_simulations for the current state."""
        if self.force_cpu and not self.force_gpu:
            raise RuntimeError("force_cpu: the MI355X core has no CPU engine")
        curr = self.get_config(for_reinit=True)
        force_reinit = bool(force_reinit or self.N != self.last_N or curr != self.last_config)
        self._global_params = np.vstack([np.ascontiguousarray(self.param_lists[p], float)[None, :]
                                         for p in self.global_param_names])
        self._regional_params = np.vstack([np.ascontiguousarray(self.param_lists[p], float).reshape(-1)
                                           for p in self.regional_param_names])
        sc, sci = self._sc_arrays()
        args = (self.model_name, sc, sci, np.ascontiguousarray(self.sc_dist, float).reshape(-1),
                self._global_params, self._regional_params,
                np.ascontiguousarray(self.param_lists["v"], float), self._model_config,
                True, self.states_ts, self.noise_out, self.do_delay, force_reinit, False,
                self.N, self.nodes, self.duration_msec, self.TR_msec, self.states_sampling_msec,
                self.sim_seed, float(self.dt), float(self.bw_dt))
        return args, curr

    def run(self, force_reinit=False):
        """Run all N simulations in one call of the core."""
        args, curr = self._run_args(force_reinit)
        out = _core.run_simulations(*args)
        self.last_N, self.last_config = self.N, curr
        self.it += 1
        self._process_out(out)

    def _process_out(self, out):
        """Outputs as attributes: sim_bold (N, volumes, nodes), sim_states
        {name: (N, nodes) or (N, samples, nodes)}, sim_fc_trils (N, pairs),
        sim_fcd_trils (N, window pairs), plus the core's raw keys."""
        self.__dict__.update(out)
        self.sim_bold = self.sim_bold.reshape(self.N, -1, self.nodes)
        self.sim_states = {}
        for i, name in enumerate(self.state_names):
            s = self._sim_states[i]
            self.sim_states[name] = s.reshape(self.N, -1, self.nodes) if self.states_ts else s
        if self.do_fc:
            self.sim_fc_trils = self.sim_fc_trils.reshape(self.N, -1)
            if self.do_fcd:
                self.sim_fcd_trils = self.sim_fcd_trils.reshape(self.N, -1)

    # ---- outputs -------------------------------------------------------------
    def _get_save_data(self):
        """Outputs and parameters written by ``save`` (base.py:1209-1227):
        sim_bold, one array per state variable, the FC / FCD trils when
        computed, and every parameter list."""
        data = {"sim_bold": self.sim_bold}
        data.update(self.sim_states)
        if self.do_fc:
            data["sim_fc_trils"] = self.sim_fc_trils
            if self.do_fcd:
                data["sim_fcd_trils"] = self.sim_fcd_trils
        data.update({k: v for k, v in self.param_lists.items() if v is not None})
        return data

    def save(self):
        """Write the outputs of the last run to ``out_dir/sim_data.npz``
        (compressed; base.py:1229-1239)."""
        if self.out_dir is None:
            raise ValueError("Cannot save the simulations when `.out_dir` is not set")
        os.makedirs(self.out_dir, exist_ok=True)
        np.savez_compressed(os.path.join(self.out_dir, "sim_data.npz"), **self._get_save_data())

    def get_state_averages(self):
        import pandas as pd
        avg = {}
        for name in self.state_names:
            s = self.sim_states[name]
            avg[name] = (s[:, self.n_states_samples_remove:, :].mean(axis=(1, 2)) if self.states_ts
                         else s.mean(axis=1))
        return pd.DataFrame(avg)

    def get_noise(self):
        """Noise each (node, step) received, shape (n_noise, nodes, bw_it,
        inner_it): the segment, re-indexed by each repeat's shuffled time
        points and nodes, concatenated and cut to the duration."""
        if not self.noise_out:
            raise ValueError("Noise was not returned; set noise_out=True and run again")
        if not hasattr(self, "_noise"):
            raise ValueError("Run the simulations first")
        seg_it = int(int(self.noise_segment_length * 1000) / self.bw_dt)
        seg = self._noise.reshape(seg_it, self.inner_it, self.nodes, self.n_noise)
        parts = [seg[ts][:, :, nodes] for ts, nodes in zip(self._shuffled_ts, self._shuffled_nodes)]
        full = np.concatenate(parts, axis=0)[:self.bw_it]
        return full.transpose(3, 2, 0, 1)

    def _pair_index(self):
        if self.exc_interhemispheric:
            h = self.nodes // 2
            i, j = np.tril_indices(self.nodes, -1)
            keep = (i < h) == (j < h)
            return i[keep], j[keep]
        return np.tril_indices(self.nodes, -1)

    def get_sim_fc(self, idx):
        """FC of simulation idx as a symmetric (nodes, nodes) matrix; NaN
        between hemispheres when they are excluded."""
        if not self.do_fc:
            raise ValueError("FC is not calculated in this simulation group")
        if idx >= self.N:
            raise IndexError("Index out of range")
        fc = np.full((self.nodes, self.nodes), np.nan if self.exc_interhemispheric else 0.0)
        i, j = self._pair_index()
        fc[i, j] = self.sim_fc_trils[idx]
        fc[j, i] = self.sim_fc_trils[idx]
        np.fill_diagonal(fc, 1.0)
        return fc

    def get_sim_fcd(self, idx):
        if not self.do_fcd:
            raise ValueError("FCD is not calculated in this simulation group")
        if idx >= self.N:
            raise IndexError("Index out of range")
        tril = self.sim_fcd_trils[idx]
        w = int(round((1 + np.sqrt(1 + 8 * tril.size)) / 2))
        if w * (w - 1) // 2 != tril.size:
            raise ValueError("FCD tril length is not a triangular number")
        fcd = np.zeros((w, w))
        fcd[np.tril_indices(w, -1)] = tril
        fcd += fcd.T
        np.fill_diagonal(fcd, 1.0)
        return fcd

    def slice(self, key, inplace=False):
        if not isinstance(key, int):
            raise ValueError("Only integer indexing is supported")
        if key >= self.N:
            raise IndexError("Index out of range")
        obj = self if inplace else copy.copy(self)
        obj.param_lists = {k: (None if v is None else v[key][None, ...]) for k, v in self.param_lists.items()}
        obj.sim_bold = self.sim_bold[key][None, ...]
        obj.sim_states = {k: v[key][None, ...] for k, v in self.sim_states.items()}
        if obj.do_fc:
            obj.sim_fc_trils = self.sim_fc_trils[key][None, ...]
            if obj.do_fcd:
                obj.sim_fcd_trils = self.sim_fcd_trils[key][None, ...]
        obj._N = 1
        return obj

    @property
    def _to_clear(self):
        attrs = ["sim_bold", "sim_fc_trils", "sim_fcd_trils", "sim_states", "_sim_states", "_noise",
                 "_shuffled_nodes", "_shuffled_ts"]
        return [a for a in attrs if hasattr(self, a)]

    def clear(self):
        for a in self._to_clear:
            delattr(self, a)
        gc.collect()

    # ---- goodness of fit ------------------------------------------------------
    # ---- hooks of optimize.BNMProblem (model-specific costs) --------------
    def _problem_init(self, problem):
        """Model-specific checks and objectives of a BNMProblem (none here)."""

    def _problem_evaluate(self, problem, X, out, *args, **kwargs):
        """Model-specific additions to a BNMProblem's costs (none here)."""

    def score(self, emp_fc_tril=None, emp_fcd_tril=None, emp_bold=None, force_cpu=False, usable_mem=None):
        """GOF terms per simulation (pandas DataFrame, one row per
        simulation): +fc_corr (Pearson r), -fc_diff (-|mean difference|),
        -fc_normec (-||diff||_2 / (2 sqrt(pairs))), -fcd_ks (-KS statistic)
        and +gof, the sum of ``gof_terms``.  With ``emp_bold`` (nodes,
        volumes) the empirical FC/FCD are computed from it.  NaN terms take
        their worst value.  +fc_corr and -fcd_ks run on the GPU unless
        force_cpu (SciPy)."""
        import pandas as pd
        import scipy.stats
        if emp_bold is not None:
            if self.do_fc:
                emp_fc_tril = utils.calculate_fc(emp_bold, self.exc_interhemispheric, return_tril=True)
            if self.do_fcd:
                emp_fcd_tril = utils.calculate_fcd(emp_bold, self.window_size_TRs, self.window_step_TRs,
                                                   drop_edges=self.fcd_drop_edges,
                                                   exc_interhemispheric=self.exc_interhemispheric)
        cols = []
        if self.do_fc and emp_fc_tril is not None:
            emp_fc_tril = np.asarray(emp_fc_tril, float)
            if emp_fc_tril.size != self.sim_fc_trils.shape[1]:
                raise ValueError("Empirical FC lower triangle size does not match the simulated FC size")
            cols += [c for c in ("+fc_corr", "-fc_diff", "-fc_normec") if c in self.gof_terms]
        if self.do_fcd and emp_fcd_tril is not None:
            emp_fcd_tril = np.asarray(emp_fcd_tril, float)
            cols += [c for c in ("-fcd_ks",) if c in self.gof_terms]
        scores = pd.DataFrame(index=range(self.N), columns=cols + ["+gof"], dtype=float)
        for c in cols:
            if c == "-fc_diff":
                v = -np.abs(self.sim_fc_trils.mean(axis=1) - emp_fc_tril.mean())
            elif c == "-fc_normec":
                v = -np.linalg.norm(self.sim_fc_trils - emp_fc_tril[None, :], axis=1) / (2 * np.sqrt(emp_fc_tril.size))
            elif c == "+fc_corr":
                v = (np.array([scipy.stats.pearsonr(x, emp_fc_tril)[0] for x in self.sim_fc_trils]) if force_cpu
                     else utils.fc_corr_device(self.sim_fc_trils, emp_fc_tril, usable_mem))
            else:
                v = -(np.array([scipy.stats.ks_2samp(x, emp_fcd_tril).statistic for x in self.sim_fcd_trils])
                      if force_cpu else utils.fcd_ks_device(self.sim_fcd_trils, emp_fcd_tril, usable_mem))
            scores[c] = np.where(np.isnan(v), WORST_SCORES.get(c, np.nan), v)
        scores["+gof"] = scores[[c for c in self.gof_terms if c in scores.columns]].sum(axis=1)
        return scores


class MultiSimGroup:
    """Several SimGroups of one model (e.g. parallel optimisation runs) run
    as one call: SCs are merged (duplicates shared), parameters
    concatenated, and the outputs split back to each child
    (reference sim/base.py:1303-1438)."""

    def __init__(self, sim_groups):
        if not sim_groups:
            raise ValueError("need at least one SimGroup")
        self.children = list(sim_groups)
        if len({type(c) for c in self.children}) != 1:
            raise ValueError("all SimGroups must be of the same model class")
        self._proto = copy.copy(self.children[0])
        self._proto.param_lists = dict(self._proto.param_lists)

    @property
    def N(self):
        return sum(c.N for c in self.children)

    def _merged(self):
        scs, idx = [], []
        for c in self.children:
            csc = c.sc if c.sc.ndim == 3 else c.sc[None]
            cidx = np.zeros(c.N, int) if csc.shape[0] == 1 else np.asarray(c.sc_indices, int)
            remap = np.empty(csc.shape[0], int)
            for k, m in enumerate(csc):
                hit = [i for i, s in enumerate(scs) if np.array_equal(s, m)]
                if hit:
                    remap[k] = hit[0]
                else:
                    scs.append(m)
                    remap[k] = len(scs) - 1
            idx.append(remap[cidx])
        return np.stack(scs), np.concatenate(idx)

    def run(self, force_reinit=False):
        g = self._proto
        sc, sci = self._merged()
        g.sc = sc if sc.shape[0] > 1 else sc[0]
        g.sc_indices = sci
        g.N = self.N
        for p in g.param_lists:
            vals = [c.param_lists[p] for c in self.children]
            g.param_lists[p] = None if any(v is None for v in vals) else np.concatenate(vals)
        args, curr = g._run_args(force_reinit)
        out = _core.run_simulations(*args)
        g.last_N, g.last_config = g.N, curr
        start = 0
        nodes = g.nodes
        for c in self.children:
            end = start + c.N
            c._global_params = g._global_params[:, start:end]
            c._regional_params = g._regional_params[:, start * nodes:end * nodes]
            part = {}
            for k, v in out.items():
                if k in ("init_time", "run_time", "_noise", "_shuffled_nodes", "_shuffled_ts"):
                    part[k] = v
                elif k in ("sim_bold", "sim_fc_trils", "sim_fcd_trils"):
                    part[k] = v[start:end]
                else:
                    part[k] = v[:, start:end]
            c.it += 1
            c._process_out(part)
            start = end


def create_multi_sim_group(sim_group_cls):
    """A MultiSimGroup class for one SimGroup class (sim/base.py:1440-1462):
    its instances take a list of ``sim_group_cls`` instances and run them as
    one call."""
    if not (isinstance(sim_group_cls, type) and issubclass(sim_group_cls, SimGroup)):
        raise TypeError("sim_group_cls must be a SimGroup subclass")

    class _Multi(MultiSimGroup):
        def __init__(self, sim_groups):
            if not all(isinstance(g, sim_group_cls) for g in sim_groups):
                raise TypeError(f"all simulation groups must be {sim_group_cls.__name__} instances")
            super().__init__(sim_groups)

    _Multi.__name__ = _Multi.__qualname__ = f"Multi{sim_group_cls.__name__}"
    return _Multi
