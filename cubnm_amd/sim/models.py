"""Model SimGroups: parameter names, state names, noise width and defaults
of each model (reference sim/rww.py, jr.py, kuramoto.py, wc.py, rwwex.py), plus the
rWW feedback-inhibition-control (FIC) options, outputs and penalty
(rww.py:19-234) and the Kuramoto initial phases (kuramoto.py:28-137)."""
import logging

import numpy as np

from .base import SimGroup

logger = logging.getLogger(__name__)


class rWWSimGroup(SimGroup):
    """Reduced Wong-Wang excitatory-inhibitory model.  With ``do_fic`` the
    core solves wIE analytically (and numerically when max_fic_trials > 0)
    and writes it back to ``param_lists['wIE']``."""
    model_name = "rWW"
    global_param_names = ["G"]
    regional_param_names = ["w_p", "J_N", "wIE", "sigma"]
    state_names = ["I_E", "I_I", "r_E", "r_I", "S_E", "S_I"]
    sel_state_var = "r_E"
    n_noise = 2
    default_params = {"G": None, "w_p": 1.4, "J_N": 0.15, "wIE": 1.0, "sigma": 0.01}

    def __init__(self, *args, do_fic=True, max_fic_trials=0, I_SAMPLING_START=1000, I_SAMPLING_END=10000,
                 init_delta=0.02, fic_penalty_scale=0.5, **kwargs):
        self.do_fic = do_fic
        self.max_fic_trials = max_fic_trials
        self.I_SAMPLING_START = I_SAMPLING_START
        self.I_SAMPLING_END = I_SAMPLING_END
        self.init_delta = init_delta
        self.fic_penalty_scale = fic_penalty_scale
        super().__init__(*args, **kwargs)

    def get_config(self, *args, **kwargs):
        config = super().get_config(*args, **kwargs)
        config.update(do_fic=self.do_fic, max_fic_trials=self.max_fic_trials,
                      I_SAMPLING_START=self.I_SAMPLING_START, I_SAMPLING_END=self.I_SAMPLING_END,
                      init_delta=self.init_delta, fic_penalty_scale=self.fic_penalty_scale)
        return config

    @property
    def _model_config(self):
        config = super()._model_config
        config.update({
            "do_fic": str(int(self.do_fic)), "max_fic_trials": str(self.max_fic_trials),
            "I_SAMPLING_START": str(self.I_SAMPLING_START), "I_SAMPLING_END": str(self.I_SAMPLING_END),
            "init_delta": str(self.init_delta), "fic_penalty_scale": str(self.fic_penalty_scale),
        })
        return config

    @SimGroup.N.setter
    def N(self, N):
        SimGroup.N.fset(self, N)
        if self.do_fic:
            self.param_lists["wIE"] = np.zeros((self._N, self.nodes))

    def _skip_default(self, name):
        return name == "wIE" and self.do_fic

    @property
    def _to_clear(self):
        return super()._to_clear + [a for a in ("fic_unstable", "fic_failed", "fic_ntrials") if hasattr(self, a)]

    def _process_out(self, out):
        super()._process_out(out)
        if self.do_fic:
            self.fic_unstable = self._global_bools[0]
            self.fic_failed = self._global_bools[1]
            self.fic_ntrials = self._global_ints[0]
            self.param_lists["wIE"] = self._regional_params[2].reshape(self.N, -1)

    def _problem_init(self, problem):
        """wIE cannot be heterogeneous when FIC sets it; with several
        objectives the FIC penalty is one of them (rww.py:254-271)."""
        if self.do_fic and "wIE" in problem.het_params:
            raise ValueError("In rWW wIE should not be specified as a heterogeneous parameter when FIC is done")
        if problem.multiobj and self.do_fic and self.fic_penalty_scale > 0:
            problem.obj_names.append("-fic_penalty")
            problem.n_obj += 1

    def _problem_evaluate(self, problem, X, out, *args, **kwargs):
        """The FIC penalty enters the cost (single objective: subtracted,
        being a negative score) or is appended as an objective
        (rww.py:273-297)."""
        scores = kwargs["scores"][-1]
        if self.do_fic and self.fic_penalty_scale > 0:
            if problem.multiobj:
                out["F"] = np.concatenate([out["F"], -scores.loc[:, ["-fic_penalty"]].values], axis=1)
            else:
                out["F"] = out["F"] - scores.loc[:, "-fic_penalty"].values

    def score(self, emp_fc_tril=None, emp_fcd_tril=None, emp_bold=None, **kwargs):
        """GOF terms plus ``-fic_penalty`` (with FIC): per node, a mean r_E
        more than 1 Hz from 3 Hz costs 1 - exp(-0.05 (|dev| - 1)); the sum
        is scaled by fic_penalty_scale / nodes (rww.py:180-234).  The
        penalty is not part of +gof."""
        scores = super().score(emp_fc_tril, emp_fcd_tril, emp_bold, **kwargs)
        if self.do_fic and self.fic_penalty_scale > 0:
            r_E = self.sim_states["r_E"]
            if self.states_ts:
                r_E = r_E[:, self.n_states_samples_remove:].mean(axis=1)
            dev = np.abs(r_E - 3.0)
            per_node = np.where(dev > 1, 1 - np.exp(-0.05 * (dev - 1)), 0.0)
            scores["-fic_penalty"] = -per_node.sum(axis=1) * self.fic_penalty_scale / self.nodes
        return scores


class JRSimGroup(SimGroup):
    """Jansen-Rit neural mass model."""
    model_name = "JR"
    global_param_names = ["G"]
    regional_param_names = ["J", "a_1", "a_2", "a_3", "a_4", "sigma"]
    state_names = ["y0", "y1", "y2", "y3", "y4", "y5", "y1_y2", "s_y1_y2"]
    sel_state_var = "y1_y2"
    n_noise = 1
    default_params = {"G": None, "J": 135.0, "a_1": 1.0, "a_2": 0.8, "a_3": 0.25, "a_4": 0.25, "sigma": 0.0001}


class KuramotoSimGroup(SimGroup):
    """Kuramoto phase oscillators.  With ``random_init_theta`` the initial
    phases are drawn once per node from U(0, 2 pi) with the simulation seed
    (numpy default_rng) and shared by all simulations; otherwise they start
    at 0 unless set in ``param_lists['init_theta']``."""
    model_name = "Kuramoto"
    global_param_names = ["G"]
    regional_param_names = ["init_theta", "omega", "sigma"]
    state_names = ["theta"]
    sel_state_var = "theta"
    n_noise = 1
    default_params = {"G": None, "init_theta": None, "omega": np.pi, "sigma": 0.17}

    def __init__(self, *args, random_init_theta=True, **kwargs):
        self.random_init_theta = random_init_theta
        super().__init__(*args, **kwargs)

    def get_config(self, *args, **kwargs):
        config = super().get_config(*args, **kwargs)
        config["random_init_theta"] = self.random_init_theta
        return config

    @property
    def _model_config(self):
        config = super()._model_config
        config["random_init_theta"] = str(int(self.random_init_theta))
        return config

    @SimGroup.N.setter
    def N(self, N):
        SimGroup.N.fset(self, N)
        if self.random_init_theta:
            theta = np.random.default_rng(self.sim_seed).uniform(0, 2 * np.pi, self.nodes)
            self.param_lists["init_theta"] = np.tile(theta, (self._N, 1))
        else:
            self.param_lists["init_theta"] = np.zeros((self._N, self.nodes))


class WCSimGroup(SimGroup):
    """Wilson-Cowan excitatory-inhibitory model."""
    model_name = "WC"
    global_param_names = ["G"]
    regional_param_names = ["c_EE", "c_IE", "c_EI", "c_II", "P_E", "P_I", "sigma_E", "sigma_I"]
    state_names = ["E", "I"]
    sel_state_var = "E"
    n_noise = 2
    default_params = {"G": None, "c_EE": 16, "c_IE": 12, "c_EI": 15, "c_II": 3, "P_E": 0.5, "P_I": 0.0,
                      "sigma_E": 5e-05, "sigma_I": 5e-05}


class rWWExSimGroup(SimGroup):
    """Reduced Wong-Wang model, excitatory population only (Deco et al. 2013):
    x = w J_N S + G J_N sum_j C_ij S_j + I0, r = (a x - b) / (1 - exp(-d (a x - b))),
    dS/dt = -S/tau + (1 - S) gamma r + sigma noise (sim/rwwex.py:14-66)."""
    model_name = "rWWEx"
    global_param_names = ["G"]
    regional_param_names = ["w", "I0", "sigma"]
    state_names = ["x", "r", "S"]
    sel_state_var = "r"
    n_noise = 1
    default_params = {"G": None, "w": 0.9, "I0": 0.3, "sigma": 0.001}
