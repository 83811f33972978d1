"""Balloon-Windkessel pins for the oracle, from the mathematics of the model.

The BW input is the BOLD state variable.  Kuramoto with omega = G = sigma = 0
keeps every phase at its initial value (theta' = 0, fmod leaves |theta| < 2 pi
unchanged), so each node's BW integrator sees a constant input s = theta_j
from the rest state x = 0, f = nu = q = 1 (bw.cpp:24-34, bnm.cpp:318).

1. Steady state.  Any Euler map's fixed point is the ODE's equilibrium:
   x = 0, f = 1 + s/gamma, nu = f^alpha, q = nu (1 - (1-rho)^(1/f)) / rho and
   BOLD = V0 [k1 (1-q) + k2 (1 - q/nu) + k3 (1 - nu)].  After 200 s every mode
   (slowest decay ~0.77/s) has settled to machine precision.  A flipped
   gamma (f-1) sign, nu^alpha for nu^(1/alpha), a wrong q balance or a wrong
   k/V0 term changes the fixed point; the k's are also checked under the
   heinzle2016-3T constants.
2. Small-signal response.  For s = eps the deviations from rest obey the
   linearised balloon model z' = J z + b s, z = (x, f-1, nu-1, q-1), whose
   Jacobian is derived below by differentiating the model's right-hand sides
   at rest.  (a) The oracle's BOLD(t)/eps approaches the continuous linear
   response (matrix exponential) to first order in the BW step (checked at
   1 ms and 0.5 ms: error ~ h, ratio ~ 2).  (b) The reference updates the BW
   variables one after another in the order x, nu, q, f, each from the
   latest values (bw.cpp:30-34: f is advanced from the new x, nu and q from
   the old f, q from the new nu).  The same sequential Euler sweep applied to
   J reproduces the oracle's response to 1e-6 relative; any other order (f
   from the old x, q from the old nu) differs by O(h) ~ 1e-3.
"""
import numpy as np
import pytest
import scipy.linalg

import cases

RHO, ALPHA, TAU, GAMMA, KAPPA = 0.34, 0.32, 0.98, 1 / 0.41, 1 / 0.65
HEINZLE = (3.72, 0.527, 0.53, 2.0)


def closed_form_bold(s, k1, k2, k3, V0):
    f = 1 + s / GAMMA
    nu = f ** ALPHA
    q = nu * (1 - (1 - RHO) ** (1 / f)) / RHO
    return V0 * (k1 * (1 - q) + k2 * (1 - q / nu) + k3 * (1 - nu))


def constant_input_args(theta, duration_s, dt="1.0", bw_dt="1.0", TR=1.0):
    N = len(theta)
    gp, rp = cases.params("Kuramoto", 1, N, G=[0.0], init_theta=np.asarray(theta)[None, :], omega=0.0,
                          sigma=0.0)
    cfg = cases.model_config("Kuramoto", TR=TR, window_size=2 * TR, window_step=TR, bold_remove_s=2)
    return cases.sim_args("Kuramoto", cases.synthetic_sc(N), gp, rp, cfg, 1, duration_s=duration_s, TR_s=TR,
                          dt=dt, bw_dt=bw_dt)


THETA = np.array([0.05, 0.3, 0.7, 1.5, 3.0, -0.2])


@pytest.mark.parametrize("consts", [cases.FRISTON2003, HEINZLE], ids=["friston2003", "heinzle2016"])
def test_bw_steady_state_closed_form(consts):
    out = cases.run_oracle(constant_input_args(THETA, 200), bw_consts=consts)
    bold = out["sim_bold"].reshape(-1, len(THETA))
    expect = closed_form_bold(THETA, *consts)
    np.testing.assert_allclose(bold[-1], expect, rtol=1e-11, atol=1e-14)
    assert np.abs(bold[-1] - bold[-2]).max() < 1e-13


def linear_model():
    """Jacobian of the balloon model at rest and the input vector, by
    differentiating its right-hand sides:
      x'  = s - kappa x - gamma (f - 1)
      f'  = x
      nu' = (f - nu^(1/alpha)) / tau
      q'  = (f (1 - (1-rho)^(1/f)) / rho - nu^(1/alpha - 1) q) / tau
    d/df [f (1 - (1-rho)^(1/f)) / rho] at f = 1 is (rho + (1-rho) ln(1-rho)) / rho;
    d/dnu [nu^(1/alpha - 1) q] = 1/alpha - 1 and d/dq = 1 at nu = q = 1."""
    cf = (RHO + (1 - RHO) * np.log(1 - RHO)) / RHO
    J = np.array([
        [-KAPPA, -GAMMA, 0.0, 0.0],
        [1.0, 0.0, 0.0, 0.0],
        [0.0, 1 / TAU, -1 / (ALPHA * TAU), 0.0],
        [0.0, cf / TAU, -(1 / ALPHA - 1) / TAU, -1 / TAU],
    ])
    b = np.array([1.0, 0.0, 0.0, 0.0])
    return J, b


def linear_bold_weights(k1, k2, k3, V0):
    # BOLD = V0 [k1 (1-q) + k2 (1 - q/nu) + k3 (1 - nu)] linearised at nu = q = 1
    return V0 * np.array([0.0, 0.0, k2 - k3, -(k1 + k2)])


def continuous_step_response(t, consts):
    J, b = linear_model()
    w = linear_bold_weights(*consts)
    # z(t) = J^-1 (e^{Jt} - I) b for a unit step from z = 0
    Jinv = np.linalg.inv(J)
    return np.array([w @ (Jinv @ (scipy.linalg.expm(J * ti) - np.eye(4)) @ b) for ti in t])


def sequential_euler_response(h, n_steps, sample_every, consts, order):
    """Unit-step response of z' = J z + b under an Euler sweep that updates the
    components one at a time in ``order``, each from the latest values;
    ``order=None`` is the explicit Euler step (every component from the old
    values)."""
    J, b = linear_model()
    w = linear_bold_weights(*consts)
    z = np.zeros(4)
    out = []
    for n in range(1, n_steps + 1):
        if order is None:
            z = z + h * (J @ z + b)
        else:
            for i in order:
                z[i] = z[i] + h * (J[i] @ z + b[i])
        if n % sample_every == 0:
            out.append(w @ z)
    return np.array(out)


X, F, NU, Q = 0, 1, 2, 3
EPS = np.array([1e-7, -2e-7, 5e-8])


@pytest.mark.parametrize("h_ms", [1.0, 0.5])
def test_bw_small_signal_vs_continuous_model(h_ms):
    dur = 40
    out = cases.run_oracle(constant_input_args(EPS, dur, dt=str(h_ms), bw_dt=str(h_ms)))
    resp = out["sim_bold"].reshape(-1, len(EPS)) / EPS
    t = np.arange(1, dur + 1, dtype=float)
    cont = continuous_step_response(t, cases.FRISTON2003)
    err = np.abs(resp - cont[:, None]).max() / np.abs(cont).max()
    # first-order scheme: error ~ C h with C of order the model's rates (~1/s)
    assert 1e-4 * h_ms < err < 5e-3 * h_ms, err


def test_bw_first_order_convergence():
    dur = 40
    t = np.arange(1, dur + 1, dtype=float)
    cont = continuous_step_response(t, cases.FRISTON2003)
    errs = []
    for h in (1.0, 0.5):
        out = cases.run_oracle(constant_input_args(EPS[:2], dur, dt=str(h), bw_dt=str(h)))
        errs.append(np.abs(out["sim_bold"].reshape(-1, 2)[:, 0] / EPS[0] - cont).max())
    assert 1.8 < errs[0] / errs[1] < 2.2, errs


def small_signal(lib=None, eps=1e-4, dur=40):
    """BOLD(t)/eps by a central difference of inputs +eps and -eps (cancels the
    quadratic term; eps large enough that the rounding of 1 - q stays below
    1e-8 relative)."""
    out = cases.run_oracle(constant_input_args(np.array([eps, -eps]), dur), lib=lib)
    b = out["sim_bold"].reshape(-1, 2)
    return (b[:, 0] - b[:, 1]) / (2 * eps)


def test_bw_update_order_small_signal():
    dur = 40
    resp = small_signal(dur=dur)
    ref = sequential_euler_response(1e-3, dur * 1000, 1000, cases.FRISTON2003, order=(X, NU, Q, F))
    scale = np.abs(ref).max()
    assert np.abs(resp - ref).max() / scale < 1e-7
    # the pin separates the update schemes: explicit Euler (f from the old x),
    # q from the old nu, q from the new f
    for wrong in [None, (X, Q, NU, F), (X, NU, F, Q)]:
        alt = sequential_euler_response(1e-3, dur * 1000, 1000, cases.FRISTON2003, order=wrong)
        assert np.abs(resp - alt).max() / scale > 1e-5


PLANTS = {
    "gamma_sign": ("- bw.y * (bf[j] - 1.0)", "+ bw.y * (bf[j] - 1.0)"),
    "nu_exponent": ("bn[j] = bn[j] + bw.dt_itau * (bf[j] - std::pow(bn[j], bw.ialpha));",
                    "bn[j] = bn[j] + bw.dt_itau * (bf[j] - std::pow(bn[j], bw.alpha));"),
    "f_from_old_x": ("double tmp_f = bf[j] + bw.dt * bx[j];\n", ""),
    "k2_term": ("bw.k2 * (1 - bq[j] / bn[j])", "bw.k2 * (1 - bq[j] * bn[j])"),
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    from oracle import oracle as O
    src = O.source()
    libs = {}
    for name, (a, b) in PLANTS.items():
        assert src.count(a) == 1, name
        s = src.replace(a, b)
        if name == "f_from_old_x":
            s = s.replace("bx[j] = bx[j] + bw.dt *", "double tmp_f = bf[j] + bw.dt * bx[j];\n            bx[j] = bx[j] + bw.dt *")
        libs[name] = O.build_variant(s, str(tmp_path_factory.mktemp(name)))
    return libs


@pytest.mark.parametrize("plant", list(PLANTS))
def test_bw_pins_reject_planted_mistakes(planted, plant):
    """Each planted mistake fails the steady-state or the update-order pin."""
    L = planted[plant]
    bad = False
    with np.errstate(all="ignore"):
        out = cases.run_oracle(constant_input_args(THETA, 200), lib=L)
        last = out["sim_bold"].reshape(-1, len(THETA))[-1]
        bad |= not np.allclose(last, closed_form_bold(THETA, *cases.FRISTON2003), rtol=1e-11, atol=1e-14)
        resp = small_signal(lib=L)
        ref = sequential_euler_response(1e-3, 40000, 1000, cases.FRISTON2003, order=(X, NU, Q, F))
        bad |= not (np.abs(resp - ref).max() / np.abs(ref).max() < 1e-7)
    assert bad
