"""Optimisation problems over a SimGroup's parameters, and the grid and
CMA-ES drivers that evaluate them (the caller side of the simulation core).

``BNMProblem`` follows the reference's ``cubnm.optimize.BNMProblem``
(optimize.py:40-619) without its pymoo base class: the same constructor
arguments and attributes (``free_params``, ``fixed_params``, ``lb``, ``ub``,
``ndim``, ``obj_names``, ``n_obj``, ``xl``, ``xu``), the same mapping from
normalised candidates X in [0, 1]^ndim to simulation parameters
(``_get_Xt``, ``_get_X``, ``_get_sim_params``: global parameters,
regional biases, node-group heterogeneity and map-based heterogeneity with
the optional range normalisation), the same ``eval`` (run + score) and the
same ``_evaluate(X, out, scores=[...])`` calling convention, including the
rWW FIC penalty hook.  ``GridOptimizer`` (optimize.py:1168-1233) and
``CMAESOptimizer`` (optimize.py:1386-1486; here driven by this package's
own CMA-ES, ``fit.CMAES``, in the unit cube) produce ``history``,
``opt_history`` and ``opt`` tables with the reference's columns.
``NSGA2Optimizer`` (optimize.py:1487-1504; this package's NSGA-II with
pymoo's default operators) handles up to three objectives, and ``save``
(optimize.py:636-731) writes the reference's run directory.
``batch_optimize`` (optimize.py:1506-1715) runs several CMA-ES or NSGA-II
optimisations with one simulation call per generation.
"""
import itertools
import logging
import os

import numpy as np
import pandas as pd

from . import sim, utils

logger = logging.getLogger(__name__)


class BNMProblem:
    def __init__(self, model, params, emp_fc_tril=None, emp_fcd_tril=None, emp_bold=None, het_params=(),
                 het_params_range="same", maps=None, maps_coef_range="auto", node_grouping=None, multiobj=False,
                 **kwargs):
        """Free parameters are the entries of ``params`` given as (min, max)
        tuples, fixed ones are given as values.  ``het_params`` (regional
        parameters) vary across nodes through ``node_grouping`` ('node',
        'sym', a path or an array of group labels: one free parameter per
        group, named f"{param}{group}") or through ``maps`` (n_maps x nodes:
        one free coefficient per map, f"{param}scale{map_idx}", the regional
        value being bias x (1 + sum_m coef_m map_m), optionally renormalised
        into ``het_params_range``).  Other keyword arguments construct the
        model's SimGroup."""
        self.model = model
        self.params = params
        self.emp_fc_tril = emp_fc_tril
        self.emp_fcd_tril = emp_fcd_tril
        self.emp_bold = emp_bold
        self.het_params = list(het_params)
        self.input_maps = maps
        self.maps_coef_range = maps_coef_range
        self.het_params_range = het_params_range
        self.node_grouping = node_grouping
        self.multiobj = multiobj
        self.sim_group = getattr(sim, f"{model}SimGroup")(**kwargs)
        sg = self.sim_group
        # empirical targets (optimize.py:143-187)
        if emp_bold is not None:
            if emp_fc_tril is not None or emp_fcd_tril is not None:
                logger.warning("Both empirical BOLD and empirical FC/FCD are provided; "
                               "FC/FCD are computed from the BOLD")
            if sg.do_fc:
                self.emp_fc_tril = utils.calculate_fc(emp_bold, sg.exc_interhemispheric, return_tril=True)
            if sg.do_fcd:
                self.emp_fcd_tril = utils.calculate_fcd(emp_bold, sg.window_size_TRs, sg.window_step_TRs,
                                                        drop_edges=sg.fcd_drop_edges,
                                                        exc_interhemispheric=sg.exc_interhemispheric)
        else:
            # flags from the targets given, in the reference's order (the
            # SimGroup setters reject FCD without FC and GOF terms without
            # their measure)
            if self.emp_fc_tril is not None:
                sg.do_fc = True
                if self.emp_fcd_tril is not None:
                    sg.do_fcd = True
            else:
                sg.do_fc = False
            if self.emp_fcd_tril is not None:
                # the core computes FC and FCD together
                sg.do_fc = True
                sg.do_fcd = True
            else:
                sg.do_fcd = False
        if node_grouping is not None and maps is not None:
            raise ValueError("Both `node_grouping` and `maps` cannot be used")
        self.global_params = list(sg.global_param_names) + ["v"]
        self.regional_params = list(sg.regional_param_names)
        self.free_params, lb, ub = [], [], []
        # node groups (optimize.py:205-236)
        self.is_node_based = node_grouping is not None
        if self.is_node_based:
            if isinstance(node_grouping, (str, os.PathLike)) and node_grouping == "node":
                self.memberships = np.arange(sg.nodes)
            elif isinstance(node_grouping, (str, os.PathLike)) and node_grouping == "sym":
                if sg.nodes % 2:
                    raise ValueError("Number of nodes must be even")
                self.memberships = np.tile(np.arange(sg.nodes // 2), 2)
            elif isinstance(node_grouping, (str, os.PathLike)):
                self.memberships = np.loadtxt(node_grouping).astype(int)
            elif isinstance(node_grouping, np.ndarray):
                self.memberships = node_grouping.astype(int)
            else:
                raise ValueError("Invalid node_grouping provided.")
            if self.memberships.shape != (sg.nodes,):
                raise ValueError(f"Invalid node_grouping shape: {self.memberships.shape}.")
            self.node_groups = np.unique(self.memberships)
        # global, bias and group parameters (optimize.py:238-261)
        for p, v in params.items():
            if p not in self.global_params + self.regional_params or not isinstance(v, tuple):
                continue
            if self.is_node_based and p in self.regional_params and p in self.het_params:
                for group in self.node_groups:
                    name = f"{p}{group}"
                    if name not in params:
                        self.free_params.append(name)
                        lb.append(v[0])
                        ub.append(v[1])
            else:
                self.free_params.append(p)
                lb.append(v[0])
                ub.append(v[1])
        self.fixed_params = [p for p in params if p not in self.free_params]
        # map coefficients (optimize.py:263-320)
        self.is_map_based = maps is not None
        if self.is_map_based:
            self.maps = np.loadtxt(maps) if isinstance(maps, (str, os.PathLike)) else np.asarray(maps, float)
            if self.maps.ndim != 2 or self.maps.shape[1] != sg.nodes:
                raise ValueError(f"Maps second dimension {self.maps.shape[-1]} != nodes {sg.nodes}")
            for p in self.het_params:
                for m in range(self.maps.shape[0]):
                    name = f"{p}scale{m}"
                    if name in params:
                        continue
                    if maps_coef_range == "auto":
                        mx, mn = self.maps[m].max(), self.maps[m].min()
                        if mn < 0 < mx:
                            # e.g. z-scored: the map can scale the bias down to 0 at either end
                            lo = np.ceil((-1 / mx) / 0.1) * 0.1
                            hi = np.floor((-1 / mn) / 0.1) * 0.1
                        else:
                            lo, hi = 0.0, 1.0
                    elif isinstance(maps_coef_range, tuple):
                        lo, hi = maps_coef_range
                    elif isinstance(maps_coef_range, list):
                        lo, hi = maps_coef_range[m]
                    else:
                        raise ValueError("Invalid maps_coef_range provided")
                    self.free_params.append(name)
                    lb.append(lo)
                    ub.append(hi)
            if het_params_range == "same":
                self._het_params_range = {}
                for p in self.het_params:
                    if p not in params:
                        raise ValueError(f"Parameter {p} is not defined in params")
                    if not isinstance(params[p], tuple):
                        raise ValueError(f"Parameter {p} baseline cannot be fixed when map-based heterogeneity "
                                         "is applied and het_params_range is set to 'same'")
                    self._het_params_range[p] = params[p]
            elif isinstance(het_params_range, dict) or het_params_range is None:
                self._het_params_range = het_params_range
            else:
                raise ValueError("Invalid het_params_range provided")
        self.lb = np.array(lb, dtype=float)
        self.ub = np.array(ub, dtype=float)
        self.ndim = len(self.free_params)
        if multiobj:
            # costs are minimised: +term becomes -term and vice versa
            self.obj_names = [("+" + t[1:]) if t.startswith("-") else ("-" + t[1:]) for t in sg.gof_terms]
        else:
            self.obj_names = ["cost"]
        self.n_obj = len(self.obj_names)
        sg._problem_init(self)
        self.n_var = self.ndim
        self.xl = np.zeros(self.ndim)
        self.xu = np.ones(self.ndim)

    def get_config(self, include_sim_group=True, include_N=False):
        config = {
            "model": self.model, "params": self.params, "het_params": self.het_params, "maps": self.input_maps,
            "node_grouping": self.node_grouping, "maps_coef_range": self.maps_coef_range,
            "het_params_range": self.het_params_range, "emp_fc_tril": self.emp_fc_tril,
            "emp_fcd_tril": self.emp_fcd_tril, "emp_bold": self.emp_bold,
        }
        if include_N:
            config["N"] = self.sim_group.N
        if include_sim_group:
            config.update(self.sim_group.get_config())
        return config

    def _get_Xt(self, X):
        """Normalised candidates (N, ndim) in [0, 1] -> parameter values."""
        return np.asarray(X, float) * (self.ub - self.lb) + self.lb

    def _get_X(self, Xt):
        """Parameter values -> normalised candidates."""
        return (np.asarray(Xt, float) - self.lb) / (self.ub - self.lb)

    def _get_sim_params(self, X):
        """Global (N,) and regional (N, nodes) parameter arrays of the
        SimGroup for candidates X (optimize.py:432-537)."""
        X = np.atleast_2d(np.asarray(X, float))
        N, nodes = X.shape[0], self.sim_group.nodes
        Xt = self._get_Xt(X)
        col = {p: i for i, p in enumerate(self.free_params)}
        out = {}
        for p in self.fixed_params:
            if isinstance(self.params[p], tuple):
                continue  # a group-heterogeneous parameter's range: set per group below
            if p in self.global_params:
                out[p] = np.repeat(float(self.params[p]), N)
            elif p in self.regional_params:
                out[p] = np.broadcast_to(np.asarray(self.params[p], float), (N, nodes)).copy()
        for p in self.free_params:
            if p in self.global_params:
                out[p] = Xt[:, col[p]].copy()
            elif p in self.regional_params:
                out[p] = np.repeat(Xt[:, col[p]][:, None], nodes, axis=1)
        if self.is_map_based:
            for p in self.het_params:
                for s in range(N):
                    scal = np.ones(nodes)
                    for m in range(self.maps.shape[0]):
                        name = f"{p}scale{m}"
                        coef = self.params[name] if name in self.fixed_params else Xt[s, col[name]]
                        scal = scal + self.maps[m] * coef
                    out[p][s] *= scal
                    if self._het_params_range is not None:
                        # min-max renormalise into the range when an end is
                        # outside it; an end that is inside stays where it is
                        lo_lim, hi_lim = self._het_params_range[p]
                        umin, umax = out[p][s].min(), out[p][s].max()
                        tmin = min(max(lo_lim, umin), hi_lim)
                        tmax = max(min(hi_lim, umax), lo_lim)
                        if umin < tmin or umax > tmax:
                            out[p][s] = (out[p][s] - umin) / (umax - umin) * (tmax - tmin) + tmin
        if self.is_node_based:
            for p in self.het_params:
                v = np.zeros((N, nodes))
                for group in self.node_groups:
                    name = f"{p}{group}"
                    sel = self.memberships == group
                    if name in self.fixed_params:
                        v[:, sel] = self.params[name]
                    else:
                        v[:, sel] = Xt[:, col[name]][:, None]
                out[p] = v
        return out

    def _set_sim_params(self, X):
        self.sim_group.param_lists.update(self._get_sim_params(X))
        self.sim_group._set_default_params(missing=True)

    def eval(self, X, skip_run=False):
        """Run the candidates' simulations and score them (pandas DataFrame,
        one row per candidate)."""
        self.sim_group.N = np.atleast_2d(X).shape[0]
        self._set_sim_params(X)
        if not skip_run:
            self.sim_group.run()
        return self.sim_group.score(self.emp_fc_tril, self.emp_fcd_tril)

    def _evaluate(self, X, out, *args, **kwargs):
        """pymoo calling convention: out['F'] = costs (N, n_obj); the scores
        table is appended to kwargs['scores']."""
        skip_run = kwargs.pop("skip_run", False)
        scores = self.eval(X, skip_run=skip_run)
        kwargs.setdefault("scores", []).append(scores)
        if self.multiobj:
            out["F"] = -scores.loc[:, self.sim_group.gof_terms].values
        else:
            out["F"] = -scores.loc[:, "+gof"].values
        self.sim_group._problem_evaluate(self, X, out, *args, **kwargs)

    def evaluate(self, X):
        """Costs (N,) or (N, n_obj) and the scores table of candidates X."""
        out, scores = {}, []
        self._evaluate(X, out, scores=scores)
        return out["F"], scores[0]


def _json_value(v):
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, (set, tuple)):
        return list(v)
    return str(v)


class _Saving:
    """``save`` of the reference's Optimizer (optimize.py:636-731): a new
    ``<name>_run-<i>`` directory under the problem's ``out_dir`` with
    history.csv, opt.csv and opt_history.csv (single objective),
    problem.json (arrays as .txt files beside it, existing paths made
    absolute), optimizer.json and, with ``save_opt``, the optimal
    simulation(s) in opt_sim/ (sim_data.npz; scores.csv when rerun).
    ``save_obj`` (a pickle of the whole optimiser) is not provided, and the
    problem's SimGroup keeps its own ``out_dir`` after the optimum is
    written (the reference leaves it pointing at opt_sim/)."""

    def get_config(self):
        return {"n_iter": self.n_iter, "popsize": self.popsize, "seed": self.seed}

    def save(self, save_opt=True, save_obj=False):
        import json
        if not getattr(self, "is_fit", False):
            raise RuntimeError("Optimizer must be fitted before saving. Call `optimize()` method to fit the optimizer.")
        if save_obj:
            raise NotImplementedError("save_obj: pickling the optimizer is not provided")
        out_dir = self.problem.sim_group.out_dir
        if out_dir is None:
            raise ValueError("Cannot save the optimizer when the problem's `out_dir` is not set")
        name = type(self).__name__.lower().replace("optimizer", "")
        run_idx = 0
        while os.path.exists(os.path.join(out_dir, f"{name}_run-{run_idx}")):
            run_idx += 1
        opt_dir = os.path.join(out_dir, f"{name}_run-{run_idx}")
        os.makedirs(opt_dir)
        self.history.to_csv(os.path.join(opt_dir, "history.csv"))
        if self.problem.n_obj == 1:
            self.opt.to_csv(os.path.join(opt_dir, "opt.csv"))
            if not isinstance(self, GridOptimizer):
                self.opt_history.to_csv(os.path.join(opt_dir, "opt_history.csv"))
        config = self.problem.get_config(include_sim_group=True, include_N=True)
        for k, v in config.items():
            if isinstance(v, np.ndarray):
                path = os.path.join(opt_dir, k + ".txt")
                np.savetxt(path, v.reshape(v.shape[0], -1) if v.ndim > 2 else v)
                config[k] = path
            if isinstance(config[k], (str, os.PathLike)) and os.path.exists(config[k]):
                config[k] = os.path.abspath(config[k])
        with open(os.path.join(opt_dir, "problem.json"), "w") as f:
            json.dump(config, f, indent=4, default=_json_value)
        with open(os.path.join(opt_dir, "optimizer.json"), "w") as f:
            json.dump(self.get_config(), f, indent=4, default=_json_value)
        if save_opt:
            opt_sim_dir = os.path.join(opt_dir, "opt_sim")
            os.makedirs(opt_sim_dir, exist_ok=True)
            sg = self._opt_sim_group(opt_sim_dir)
            prev, sg.out_dir = sg.out_dir, opt_sim_dir
            sg.save()
            if sg is self.problem.sim_group:   # the problem keeps saving under its own out_dir
                sg.out_dir = prev
            self.problem.opt_sim_group = sg
        return opt_dir

    def _opt_X(self):
        return self.problem._get_X(np.atleast_2d(self.opt[self.problem.free_params].values.astype(float)))

    def _opt_sim_group(self, opt_sim_dir):
        """Rerun the optimum (the best candidate, or the final population's
        non-dominated set with several objectives) and write its scores
        (optimize.py:717-727)."""
        X = self._opt_X()
        Xt = pd.DataFrame(self.problem._get_Xt(X), columns=self.problem.free_params)
        pd.concat([Xt, self.problem.eval(X)], axis=1).to_csv(os.path.join(opt_sim_dir, "scores.csv"))
        return self.problem.sim_group


class GridOptimizer(_Saving):
    """Every point of a grid over the free parameters, evaluated in one
    batch (optimize.py:1168-1233)."""

    def get_config(self):
        return {"grid_shape": self.grid_shape}

    def save(self, save_avg_states=True, **kwargs):
        """As the base ``save``, plus state_averages.csv of every grid point
        (optimize.py:1235-1254); the optimal simulation is sliced out of the
        grid run instead of rerun."""
        opt_dir = super().save(**kwargs)
        if save_avg_states:
            self.problem.sim_group.get_state_averages().to_csv(os.path.join(opt_dir, "state_averages.csv"))
        return opt_dir

    def _opt_sim_group(self, opt_sim_dir):
        return self.problem.sim_group.slice(int(self.opt.name), inplace=False)

    def optimize(self, problem, grid_shape):
        self.problem = problem
        if isinstance(grid_shape, int):
            grid_shape = {p: grid_shape for p in problem.free_params}
        self.grid_shape = grid_shape
        axes = [np.linspace(0, 1, grid_shape[p]) for p in problem.free_params]
        X = np.array(list(itertools.product(*axes)))
        self.popsize = X.shape[0]
        out, scores = {}, []
        problem._evaluate(X, out, scores=scores)
        scores = scores[0]
        scores["cost"] = out["F"]
        Xt = pd.DataFrame(problem._get_Xt(X), columns=problem.free_params)
        self.history = pd.concat([Xt, scores], axis=1)
        self.opt = self.history.loc[self.history["cost"].idxmin()]
        self.is_fit = True
        return self.opt


class _PopulationOptimizer(_Saving):
    """Generation-at-a-time driver shared by the population optimisers, so
    that batch_optimize can run the generations of several optimisers as
    one simulation call.  A subclass provides _start_algorithm, _ask and
    _tell_algorithm; history rows are the candidates' parameters, their
    objective values (column 'cost' for one objective, the problem's
    obj_names otherwise), with one objective the score columns, and 'gen'.
    Keyword arguments of the reference's PymooOptimizer (optimize.py:1267-1330):
    n_iter, seed, print_history, and save_history_sim (False: each
    generation's simulation outputs are cleared once scored); a pymoo
    ``termination`` object is not provided."""

    def __init__(self, popsize, n_iter=2, seed=0, print_history=False, save_history_sim=False, termination=None):
        if termination is not None:
            raise NotImplementedError("termination: pymoo termination objects are not provided; use n_iter")
        self.popsize, self.n_iter, self.seed = popsize, n_iter, seed
        self.print_history, self.save_history_sim = print_history, save_history_sim

    def setup_problem(self, problem):
        if problem.n_obj > self.max_obj:
            raise ValueError(f"{type(self).__name__} optimises at most {self.max_obj} objective(s)")
        self.problem = problem

    def _record(self, X, F, scores, gen):
        Xt = pd.DataFrame(self.problem._get_Xt(X), columns=self.problem.free_params)
        Xt.index.name = "sim_idx"
        single = self.problem.n_obj == 1
        Fd = pd.DataFrame(np.reshape(F, (len(X), -1)), columns=["cost"] if single else self.problem.obj_names)
        res = pd.concat([Xt, Fd, scores] if single else [Xt, Fd], axis=1)
        res["gen"] = gen
        if self.print_history:
            print(res.to_string())
        self.history.append(res)
        if single:
            self.opt_history.append(res.loc[res["cost"].argmin()])

    def _start(self):
        self._rng = np.random.default_rng(self.seed)
        self.history, self.opt_history = [], []
        self._gen = 0
        self._start_algorithm()

    def _has_next(self):
        return self._gen <= self.n_iter

    def _tell(self, X, F, scores):
        F = np.where(np.isnan(F), np.inf, F)
        self._tell_algorithm(X, F)
        self._record(X, F, scores, self._gen)
        self._gen += 1
        sg = getattr(self.problem, "sim_group", None)
        if sg is not None and not self.save_history_sim:   # the generation's outputs are not kept
            sg.clear()

    def _finish(self):
        self.history = pd.concat(self.history, axis=0).reset_index(drop=False)
        self.is_fit = True
        if self.problem.n_obj == 1:
            self.opt_history = pd.DataFrame(self.opt_history).reset_index(drop=True)
            self.opt = self.history.loc[self.history["cost"].argmin()]
            return self.opt
        return None

    def optimize(self):
        self._start()
        while self._has_next():
            X = self._ask()
            F, scores = self.problem.evaluate(X)
            self._tell(X, F, scores)
        return self._finish()


class CMAESOptimizer(_PopulationOptimizer):
    """CMA-ES in the unit cube over a BNMProblem's free parameters
    (optimize.py:1386-1486): with x0=None the first generation is popsize
    uniform samples whose best point becomes the initial mean; sigma 0.5;
    popsize candidates per generation for n_iter generations after it.
    Uses this package's CMA-ES (fit.CMAES, Hansen's tutorial defaults,
    samples clipped to the cube) instead of pymoo's wrapper of `cma`, so the
    trajectories are not the reference's; hence ``use_bound_penalty`` and
    ``algorithm_kws`` (options of `cma`) raise NotImplementedError."""
    max_obj = 1

    def __init__(self, popsize, x0=None, sigma=0.5, use_bound_penalty=False, algorithm_kws=None, **kwargs):
        if use_bound_penalty:
            raise NotImplementedError("use_bound_penalty: candidates are clipped to the bounds instead")
        if algorithm_kws:
            raise NotImplementedError("algorithm_kws: options of the `cma` package are not provided")
        super().__init__(popsize, **kwargs)
        self.sigma, self.x0 = sigma, x0

    def setup_problem(self, problem):
        if problem.n_obj != 1:
            raise ValueError("CMA-ES optimises a single objective")
        self.problem = problem

    def _start_algorithm(self):
        from .fit import CMAES
        self._es = None
        if self.x0 is not None:
            x0 = self.problem._get_X(np.asarray(self.x0, float))
            self._es = CMAES(x0, self.sigma, self.popsize, seed=self.seed + 1)
            self._gen = 1

    def _ask(self):
        if self._es is None:      # generation 0: uniform samples in the cube
            return self._rng.uniform(0, 1, (self.popsize, self.problem.ndim))
        return self._es.ask()

    def _tell_algorithm(self, X, F):
        from .fit import CMAES
        if self._es is None:
            self._es = CMAES(X[int(np.argmin(F))], self.sigma, self.popsize, seed=self.seed + 1)
        else:
            self._es.tell(X, F)


def _nondominated_ranks(F):
    """Front index of every row of F (minimised objectives): front 0 is the
    rows no other row dominates, front r+1 those dominated only by rows of
    fronts <= r (Deb et al. 2002, fast non-dominated sorting)."""
    F = np.asarray(F, float).reshape(len(F), -1)
    le = np.all(F[:, None, :] <= F[None, :, :], axis=2)
    lt = np.any(F[:, None, :] < F[None, :, :], axis=2)
    dom = le & lt                                  # dom[i, j]: i dominates j
    count = dom.sum(axis=0)
    rank = np.full(len(F), -1)
    left = np.ones(len(F), bool)
    r = 0
    while left.any():
        front = np.where(left & (count == 0))[0]
        rank[front] = r
        left[front] = False
        count = count - dom[front].sum(axis=0)
        r += 1
    return rank


def _crowding_distance(F):
    """Crowding distance of the rows of one front: per objective, the gap
    between the two neighbours normalised by the front's range, summed;
    the extreme rows of every objective are infinite."""
    F = np.asarray(F, float).reshape(len(F), -1)
    n = len(F)
    if n <= 2:
        return np.full(n, np.inf)
    d = np.zeros(n)
    for k in range(F.shape[1]):
        o = np.argsort(F[:, k], kind="stable")
        f = F[o, k]
        d[o[0]] = d[o[-1]] = np.inf
        span = f[-1] - f[0]
        if span > 0 and np.isfinite(span):
            d[o[1:-1]] += (f[2:] - f[:-2]) / span
    return d


def _rank_and_crowding(F):
    rank = _nondominated_ranks(F)
    crowd = np.empty(len(rank))
    for r in np.unique(rank):
        m = rank == r
        crowd[m] = _crowding_distance(np.asarray(F).reshape(len(rank), -1)[m])
    return rank, crowd


class NSGA2Optimizer(_PopulationOptimizer):
    """NSGA-II (Deb et al. 2002) in the unit cube over a BNMProblem's free
    parameters, for up to three objectives (optimize.py:1487-1504):
    generation 0 is popsize uniform samples; each later generation makes
    popsize offspring by binary tournaments on (front, crowding distance),
    simulated binary crossover (eta 15, probability 0.9 per pair, 0.5 per
    variable) and polynomial mutation (eta 20, probability 1/ndim per
    variable), and keeps the best popsize of parents and offspring by front
    then crowding distance.  These are pymoo's NSGA2 defaults, but the
    random streams are this package's, so trajectories differ from the
    reference's.  ``algorithm_kws`` may set eta_c, prob_c and eta_m."""
    max_obj = 3

    def __init__(self, popsize, algorithm_kws=None, **kwargs):
        kws = dict(algorithm_kws or {})
        self.eta_c, self.prob_c, self.eta_m = kws.pop("eta_c", 15.0), kws.pop("prob_c", 0.9), kws.pop("eta_m", 20.0)
        if kws:
            raise NotImplementedError(f"algorithm_kws {sorted(kws)}: only eta_c, prob_c and eta_m are provided")
        super().__init__(popsize, **kwargs)

    def _start_algorithm(self):
        self._pop_X = self._pop_F = None

    def _tournament(self, rank, crowd, n):
        a = self._rng.integers(0, len(rank), n)
        b = self._rng.integers(0, len(rank), n)
        better_a = (rank[a] < rank[b]) | ((rank[a] == rank[b]) & (crowd[a] > crowd[b]))
        tie = (rank[a] == rank[b]) & (crowd[a] == crowd[b])
        coin = self._rng.random(n) < 0.5
        return np.where(better_a | (tie & coin), a, b)

    def _sbx(self, p1, p2):
        """Simulated binary crossover bounded to [0, 1] (Deb and Agrawal 1995;
        Deb's bounded form)."""
        c1, c2 = p1.copy(), p2.copy()
        eta = self.eta_c
        for i in range(len(p1)):
            if self._rng.random() >= self.prob_c:
                continue
            for k in range(p1.shape[1]):
                if self._rng.random() > 0.5 or abs(p1[i, k] - p2[i, k]) <= 1e-14:
                    continue
                y1, y2 = min(p1[i, k], p2[i, k]), max(p1[i, k], p2[i, k])
                u = self._rng.random()
                out = []
                for beta in (1 + 2 * y1 / (y2 - y1), 1 + 2 * (1 - y2) / (y2 - y1)):
                    alpha = 2 - beta ** -(eta + 1)
                    bq = (u * alpha) ** (1 / (eta + 1)) if u <= 1 / alpha else (1 / (2 - u * alpha)) ** (1 / (eta + 1))
                    out.append(bq)
                v1 = min(max(0.5 * (y1 + y2 - out[0] * (y2 - y1)), 0.0), 1.0)
                v2 = min(max(0.5 * (y1 + y2 + out[1] * (y2 - y1)), 0.0), 1.0)
                if self._rng.random() < 0.5:
                    v1, v2 = v2, v1
                c1[i, k], c2[i, k] = v1, v2
        return c1, c2

    def _mutate(self, X):
        """Polynomial mutation bounded to [0, 1] (Deb and Goyal 1996)."""
        X = X.copy()
        eta, pm = self.eta_m, 1.0 / X.shape[1]
        for i in range(X.shape[0]):
            for k in range(X.shape[1]):
                if self._rng.random() >= pm:
                    continue
                y, u = X[i, k], self._rng.random()
                if u < 0.5:
                    val = 2 * u + (1 - 2 * u) * (1 - y) ** (eta + 1)
                    dq = val ** (1 / (eta + 1)) - 1
                else:
                    val = 2 * (1 - u) + 2 * (u - 0.5) * y ** (eta + 1)
                    dq = 1 - val ** (1 / (eta + 1))
                X[i, k] = min(max(y + dq, 0.0), 1.0)
        return X

    def _ask(self):
        if self._pop_X is None:
            return self._rng.uniform(0, 1, (self.popsize, self.problem.ndim))
        rank, crowd = _rank_and_crowding(self._pop_F)
        n_pairs = (self.popsize + 1) // 2
        p1 = self._pop_X[self._tournament(rank, crowd, n_pairs)]
        p2 = self._pop_X[self._tournament(rank, crowd, n_pairs)]
        c1, c2 = self._sbx(p1, p2)
        return self._mutate(np.vstack([c1, c2])[:self.popsize])

    def _tell_algorithm(self, X, F):
        F = np.reshape(F, (len(X), -1))
        if self._pop_X is None:
            self._pop_X, self._pop_F = X.copy(), F.copy()
            return
        X_all, F_all = np.vstack([self._pop_X, X]), np.vstack([self._pop_F, F])
        rank, crowd = _rank_and_crowding(F_all)
        keep = np.lexsort((-crowd, rank))[:self.popsize]
        self._pop_X, self._pop_F = X_all[keep], F_all[keep]

    def _opt_X(self):
        if self.problem.n_obj == 1:
            return super()._opt_X()
        return self._pop_X[_nondominated_ranks(self._pop_F) == 0]


def batch_optimize(optimizers, problems, save=True):
    """Run several CMA-ES / NSGA-II optimisations together (optimize.py:1506-1715):
    each generation, the candidates of every optimisation still running are
    simulated in one call through a MultiSimGroup (SCs merged, duplicates
    shared), then each optimiser scores its own part and takes its next
    step.  One optimiser for a list of problems, a list of optimisers for
    one problem, or two lists of the same length.  Optimisers and problems
    are deep-copied, so the same object may appear several times; every
    copy draws from its own seeded generators, so each history equals that
    of the same optimiser run alone on its problem.  Returns the list of
    fitted optimisers.  With ``save`` (the default) each optimiser is saved
    as by its ``save()`` (the optima rerun one problem at a time), which
    needs the problems' ``out_dir``."""
    import copy
    if not isinstance(optimizers, list) and not isinstance(problems, list):
        raise ValueError("At least one of optimizers or problems must be a list")
    if not isinstance(optimizers, list):
        optimizers = [optimizers] * len(problems)
    if not isinstance(problems, list):
        problems = [problems] * len(optimizers)
    if len(optimizers) != len(problems):
        raise ValueError("Same number of optimizers and problems must be provided.")
    if not all(isinstance(o, _PopulationOptimizer) for o in optimizers):
        raise TypeError("Batch optimization is only supported for population-based optimizers "
                        "(CMAESOptimizer, NSGA2Optimizer)")
    if len({type(p.sim_group) for p in problems}) != 1:
        raise ValueError("All optimizations must be done on the same type of model.")
    if save and any(p.sim_group.out_dir is None for p in problems):
        raise ValueError("save=True needs every problem's `out_dir` (or pass save=False)")
    optimizers = [copy.deepcopy(o) for o in optimizers]
    problems = [copy.deepcopy(p) for p in problems]
    for o, p in zip(optimizers, problems):
        o.setup_problem(p)
        o._start()
    while any(o._has_next() for o in optimizers):
        ongoing = [o for o in optimizers if o._has_next()]
        Xs = []
        for o in ongoing:
            X = o._ask()
            o.problem.sim_group.N = X.shape[0]
            o.problem._set_sim_params(X)
            Xs.append(X)
        sim.MultiSimGroup([o.problem.sim_group for o in ongoing]).run()
        for o, X in zip(ongoing, Xs):
            out, scores = {}, []
            o.problem._evaluate(X, out, scores=scores, skip_run=True)
            o._tell(X, out["F"], scores[0])
    for o in optimizers:
        o._finish()
    if save:
        for o in optimizers:
            o.save()
    return optimizers
