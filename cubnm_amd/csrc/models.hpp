// Neural-mass model definitions shared by the host (constants, analytical
// FIC) and the gfx950 integrator kernels (per-node Euler-Maruyama steps).
//
// Counts, state layouts and update expressions follow the reference models:
//   rWW       include/cubnm/models/rww.hpp:88-109, src/ext/models/rww.cpp:16-53, 149-313
//   JR        include/cubnm/models/jr.hpp:55-75,   src/ext/models/jr.cpp:14-133
//   Kuramoto  include/cubnm/models/kuramoto.hpp,   src/ext/models/kuramoto.cpp:14-61
//   WC        include/cubnm/models/wc.hpp:50-70,   src/ext/models/wc.cpp:14-80
//   rWWEx     include/cubnm/models/rwwex.hpp,      src/ext/models/rwwex.cpp:14-66
// The expressions keep the reference's operation order (left-to-right
// products such as (w_p*J_N)*S_E) so that FP64 results agree with the CPU
// path to rounding; noise enters as already-gathered values.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define CUBNM_HD __host__ __device__ inline
// Division in the split steps (pre/post) used by the batched integrator:
// the reciprocal-based sequence on the device, IEEE division on the host.
// Exponential in the split steps: exp_cw (below) on the device, the C
// library's exp on the host.
#ifdef __HIP_DEVICE_COMPILE__
#define DIV(a, b) fast_div((a), (b))
#define EXP(x) exp_cw(x)
#else
#define DIV(a, b) ((a) / (b))
#define EXP(x) exp(x)
#endif

namespace cubnm {

enum ModelId { MODEL_RWW = 0, MODEL_JR = 1, MODEL_KURAMOTO = 2, MODEL_WC = 3, MODEL_RWWEX = 4, MODEL_COUNT = 5 };

// Balloon-Windkessel constants (src/ext/models/bw.cpp:4-22; core.cpp:157-192, 291-292)
struct BWConsts {
    double dt, rho, alpha, tau, y, kappa, V_0, k1, k2, k3;
    double ialpha, itau, oneminrho, dt_itau;
    double log_oneminrho;  // log(1 - rho), for the exp/log form of the BW step
    double irho;           // 1 / rho
};

// a / b from the hardware reciprocal: one Newton step and one residual
// correction (6 instructions instead of the 11 of the IEEE division
// sequence); within one ulp of a / b for normal operands.
__device__ inline double fast_div(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

// exp(x) by Cody-Waite reduction and a Taylor polynomial: x = k ln2 + r with
// k = rint(x log2 e) and |r| <= ln2/2 (ln2 split into a head with trailing
// zero bits, so k * head is exact, and a tail), exp(r) by its Taylor series
// to degree 13 in Horner form (truncation r^14/14! < 5e-18 relative), then
// 2^k by ldexp.  Within 1 ulp of the C library's exp over [-700, 700]
// (checked on 2e7 points); underflows to 0 and overflows to inf through
// ldexp.  Unlike the device library's exp it has no branches or range
// checks (19 instructions instead of about 25), and exp(+-inf) is NaN: the
// integrators never pass an infinite argument (the model inputs are bounded
// by clipped states and finite parameters).
__device__ inline double exp_cw(double x) {
    const double k = __builtin_rint(x * 1.4426950408889634074);
    double r = fma(-k, 6.93147180369123816490e-01, x);
    r = fma(-k, 1.90821492927058770002e-10, r);
    double p = 1.0 / 6227020800.0;
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)k);
}

// The exponential of the split model steps (pre/post) as a functor, so that
// an integrator can pass its own; this one is EXP.
struct ExpPoly {
    CUBNM_HD double operator()(double x) const { return EXP(x); }
};

// rWW ---------------------------------------------------------------------
struct RWW {
    static constexpr int n_state = 6, n_noise = 2, n_global = 1, n_regional = 4;
    static constexpr int conn_idx = 4, bold_idx = 4, n_gout_bool = 2, n_gout_int = 1;
    static constexpr bool is_osc = false;
    struct C {
        double dt, sqrt_dt, a_E, b_E, d_E, a_I, b_I, d_I, w_II;
        double w_E__I_0, w_I__I_0, b_a_ratio_E, dt_itau_E, dt_gamma_E, dt_itau_I, dt_gamma_I;
        // numerical FIC config (rww.hpp:61-69)
        int do_fic, max_fic_trials, I_SAMPLING_START, I_SAMPLING_END, I_SAMPLING_DURATION;
        double init_delta;
        // analytical FIC constants
        double tau_I_s, gamma_I_s, I_0, w_E, w_I, I_ext, I_E_ss, S_E_ss;
    };
    static void init_constants(C& c, double dt) {
        c.dt = dt; c.sqrt_dt = std::sqrt(dt);
        c.a_E = 310; c.b_E = 125; c.d_E = 0.16; c.a_I = 615; c.b_I = 177; c.d_I = 0.087;
        const double gamma_E = (double)0.641 / (double)1000.0, gamma_I = (double)1.0 / (double)1000.0;
        const double tau_E = 100, tau_I = 10;
        c.I_0 = 0.382; c.w_E = 1.0; c.w_I = 0.7; c.w_II = 1.0; c.I_ext = 0.0;
        c.w_E__I_0 = c.w_E * c.I_0; c.w_I__I_0 = c.w_I * c.I_0; c.b_a_ratio_E = c.b_E / c.a_E;
        c.dt_itau_E = dt * (1.0 / tau_E); c.dt_gamma_E = dt * gamma_E;
        c.dt_itau_I = dt * (1.0 / tau_I); c.dt_gamma_I = dt * gamma_I;
        c.tau_I_s = 0.01; c.gamma_I_s = 1.0;
        c.I_E_ss = 0.3773805650; c.S_E_ss = 0.1647572075;
        c.do_fic = 1; c.max_fic_trials = 0; c.I_SAMPLING_START = 1000; c.I_SAMPLING_END = 10000;
        c.I_SAMPLING_DURATION = 10000 - 1000 + 1; c.init_delta = 0.02;
    }
    CUBNM_HD static void init(const C&, double* s, const double*) { s[4] = 0.001; s[5] = 0.001; }
    // rww.cpp:187-225
    CUBNM_HD static void step(const C& c, double* s, const double* gp, const double* rp, double g,
                              double n0, double n1) {
        s[0] = c.w_E__I_0 + rp[0] * rp[1] * s[4] + g * gp[0] * rp[1] - rp[2] * s[5];
        s[1] = c.w_I__I_0 + rp[1] * s[4] - c.w_II * s[5];
        const double aE = c.a_E * s[0] - c.b_E;
        const double aI = c.a_I * s[1] - c.b_I;
        s[2] = aE / (1 - exp(-c.d_E * aE));
        s[3] = aI / (1 - exp(-c.d_I * aI));
        const double dSE = n0 * rp[3] * c.sqrt_dt + c.dt_gamma_E * ((1 - s[4]) * s[2]) - c.dt_itau_E * s[4];
        const double dSI = n1 * rp[3] * c.sqrt_dt + c.dt_gamma_I * s[3] - c.dt_itau_I * s[5];
        s[4] += dSE;
        s[5] += dSI;
        s[4] = fmax(0.0, fmin(1.0, s[4]));
        s[5] = fmax(0.0, fmin(1.0, s[5]));
    }
    // The same step in two parts for the batched integrator: `pre` is
    // everything that reads neither this step's coupling input g nor its
    // noise (I_I, r_I, the deterministic part of dS_I and the partial sums
    // of I_E), so it can run while the coupling product is on the matrix
    // cores and the noise loads are in flight; `post` finishes I_E, the
    // excitatory population and S_I.  The operations are those of `step`;
    // dS_I adds its noise term last (step adds it first), a difference in
    // rounding only.  The split steps take the noise scale sigma*sqrt(dt) in
    // rp[3] (scale_rp, applied once when the parameters are loaded): one
    // multiplication per noise term instead of two, (sigma sqrt(dt)) n
    // instead of (n sigma) sqrt(dt), a difference in rounding only.
    static constexpr int n_tmp = 3;
    CUBNM_HD static void scale_rp(const C& c, double* rp) { rp[3] *= c.sqrt_dt; }
    template <class E = ExpPoly>
    CUBNM_HD static void pre(const C& c, double* s, const double* gp, const double* rp, double n0, double n1,
                             double* t, const E& ex = E()) {
        (void)gp; (void)n0; (void)n1;
        t[0] = c.w_E__I_0 + rp[0] * rp[1] * s[4];
        t[1] = rp[2] * s[5];
        s[1] = c.w_I__I_0 + rp[1] * s[4] - c.w_II * s[5];
        const double aI = c.a_I * s[1] - c.b_I;
        s[3] = DIV(aI, 1 - ex(-c.d_I * aI));
        t[2] = c.dt_gamma_I * s[3] - c.dt_itau_I * s[5];
    }
    template <class E = ExpPoly>
    CUBNM_HD static void post(const C& c, double* s, const double* gp, const double* rp, double g, double n0,
                              double n1, const double* t, const E& ex = E()) {
        s[0] = t[0] + g * gp[0] * rp[1] - t[1];
        const double aE = c.a_E * s[0] - c.b_E;
        s[2] = DIV(aE, 1 - ex(-c.d_E * aE));
        const double dSE = n0 * rp[3] + c.dt_gamma_E * ((1 - s[4]) * s[2]) - c.dt_itau_E * s[4];
        s[4] += dSE;
        s[4] = fmax(0.0, fmin(1.0, s[4]));
        s[5] += n1 * rp[3] + t[2];
        s[5] = fmax(0.0, fmin(1.0, s[5]));
    }
};

// Jansen-Rit ----------------------------------------------------------------
struct JR {
    static constexpr int n_state = 8, n_noise = 1, n_global = 1, n_regional = 6;
    static constexpr int conn_idx = 7, bold_idx = 6, n_gout_bool = 0, n_gout_int = 0;
    static constexpr bool is_osc = false;
    struct C { double dt, sqrt_dt, A, B, a, b, v0, nu_max, r, mu, cmin, cmax, midpoint, a2, b2; };
    static void init_constants(C& c, double dt) {
        c.dt = dt; c.sqrt_dt = std::sqrt(dt);
        c.A = 3.25; c.B = 22.0; c.a = 0.1; c.b = 0.05; c.v0 = 6.0; c.nu_max = 0.0025; c.r = 0.56;
        c.mu = 0.22; c.cmin = 0.0; c.cmax = 0.005; c.midpoint = 6.0;
        c.a2 = std::pow(c.a, 2); c.b2 = std::pow(c.b, 2);  // the reference evaluates pow(a, 2) per step
    }
    CUBNM_HD static void init(const C&, double* s, const double*) { for (int i = 0; i < 8; i++) s[i] = 0.0; }
    // jr.cpp:80-133
    CUBNM_HD static void step(const C& c, double* s, const double* gp, const double* rp, double g,
                              double n0, double) {
        const double sg0 = 2.0 * c.nu_max / (1.0 + exp(c.r * (c.v0 - (s[1] - s[2]))));
        const double sg1 = 2.0 * c.nu_max / (1.0 + exp(c.r * (c.v0 - (rp[1] * rp[0] * s[0]))));
        const double sg2 = 2.0 * c.nu_max / (1.0 + exp(c.r * (c.v0 - (rp[3] * rp[0] * s[0]))));
        const double d0 = s[3], d1 = s[4], d2 = s[5];
        const double d3 = c.A * c.a * sg0 - 2.0 * c.a * s[3] - c.a2 * s[0];
        const double d4 = c.A * c.a * (c.mu + rp[2] * rp[0] * sg1 + gp[0] * g) - 2.0 * c.a * s[4] - c.a2 * s[1];
        const double d5 = c.B * c.b * (rp[4] * rp[0] * sg2) - 2.0 * c.b * s[5] - c.b2 * s[2];
        s[0] += c.dt * d0;
        s[1] += c.dt * d1;
        s[2] += c.dt * d2;
        s[3] += c.dt * d3 + c.sqrt_dt * n0 * rp[5];
        s[4] += c.dt * d4;
        s[5] += c.dt * d5;
        s[6] = s[1] - s[2];
        s[7] = c.cmin + (c.cmax - c.cmin) / (1.0 + exp(c.r * (c.midpoint - (s[6]))));
    }
    // `step` split around the coupling input (see RWW::pre): only y4 reads g.
    static constexpr int n_tmp = 2;
    CUBNM_HD static void scale_rp(const C&, double*) {}
    template <class E = ExpPoly>
    CUBNM_HD static void pre(const C& c, double* s, const double* gp, const double* rp, double n0, double,
                             double* t, const E& ex = E()) {
        (void)gp;
        const double sg0 = DIV(2.0 * c.nu_max, 1.0 + ex(c.r * (c.v0 - (s[1] - s[2]))));
        const double sg1 = DIV(2.0 * c.nu_max, 1.0 + ex(c.r * (c.v0 - (rp[1] * rp[0] * s[0]))));
        const double sg2 = DIV(2.0 * c.nu_max, 1.0 + ex(c.r * (c.v0 - (rp[3] * rp[0] * s[0]))));
        const double d0 = s[3], d1 = s[4], d2 = s[5];
        const double d3 = c.A * c.a * sg0 - 2.0 * c.a * s[3] - c.a2 * s[0];
        t[0] = c.mu + rp[2] * rp[0] * sg1;
        t[1] = c.a2 * s[1];
        const double d5 = c.B * c.b * (rp[4] * rp[0] * sg2) - 2.0 * c.b * s[5] - c.b2 * s[2];
        s[0] += c.dt * d0;
        s[1] += c.dt * d1;
        s[2] += c.dt * d2;
        s[3] += c.dt * d3 + c.sqrt_dt * n0 * rp[5];
        s[5] += c.dt * d5;
        s[6] = s[1] - s[2];
        s[7] = c.cmin + DIV(c.cmax - c.cmin, 1.0 + ex(c.r * (c.midpoint - (s[6]))));
    }
    template <class E = ExpPoly>
    CUBNM_HD static void post(const C& c, double* s, const double* gp, const double*, double g, double, double,
                              const double* t, const E& ex = E()) {
        const double d4 = c.A * c.a * (t[0] + gp[0] * g) - 2.0 * c.a * s[4] - t[1];
        s[4] += c.dt * d4;
    }
};

// Kuramoto ------------------------------------------------------------------
struct Kuramoto {
    static constexpr int n_state = 1, n_noise = 1, n_global = 1, n_regional = 3;
    static constexpr int conn_idx = 0, bold_idx = 0, n_gout_bool = 0, n_gout_int = 0;
    static constexpr bool is_osc = true;
    struct C { double dt, sqrt_dt, twopi; };
    static void init_constants(C& c, double dt) { c.dt = dt; c.sqrt_dt = std::sqrt(dt); c.twopi = 2.0 * M_PI; }
    CUBNM_HD static void init(const C&, double* s, const double* rp) { s[0] = rp[0]; }
    // kuramoto.cpp:48-61
    CUBNM_HD static void step(const C& c, double* s, const double* gp, const double* rp, double g,
                              double n0, double) {
        const double td = rp[1] + gp[0] * g;
        s[0] = s[0] + c.dt * td + c.sqrt_dt * n0 * rp[2];
        s[0] = fmod(s[0], c.twopi);
    }
    // the whole step reads the coupling input
    static constexpr int n_tmp = 1;
    CUBNM_HD static void scale_rp(const C&, double*) {}
    template <class E = ExpPoly>
    CUBNM_HD static void pre(const C&, double*, const double*, const double*, double, double, double*, const E& = E()) {}
    template <class E = ExpPoly>
    CUBNM_HD static void post(const C& c, double* s, const double* gp, const double* rp, double g, double n0,
                              double n1, const double*, const E& = E()) {
        step(c, s, gp, rp, g, n0, n1);
    }
};

// Wilson-Cowan ----------------------------------------------------------------
struct WC {
    static constexpr int n_state = 2, n_noise = 2, n_global = 1, n_regional = 8;
    static constexpr int conn_idx = 0, bold_idx = 0, n_gout_bool = 0, n_gout_int = 0;
    static constexpr bool is_osc = false;
    struct C { double dt, sqrt_dt, itau_E, itau_I, a_E, a_I, mu_E, mu_I; };
    static void init_constants(C& c, double dt) {
        c.dt = dt; c.sqrt_dt = std::sqrt(dt);
        c.itau_E = 1.0 / 2.5; c.itau_I = 1.0 / 3.75; c.a_E = 1.5; c.a_I = 1.5; c.mu_E = 3.0; c.mu_I = 3.0;
    }
    CUBNM_HD static void init(const C&, double* s, const double*) { s[0] = 0.001; s[1] = 0.001; }
    // wc.cpp:51-80
    CUBNM_HD static void step(const C& c, double* s, const double* gp, const double* rp, double g,
                              double n0, double n1) {
        const double SE = 1.0 / (1.0 + exp(-c.a_E * (rp[0] * s[0] - rp[1] * s[1] + gp[0] * g + rp[4] - c.mu_E)));
        const double SI = 1.0 / (1.0 + exp(-c.a_I * (rp[2] * s[0] - rp[3] * s[1] + rp[5] - c.mu_I)));
        const double dE = (-s[0] + (1.0 - s[0]) * SE) * c.itau_E;
        const double dI = (-s[1] + (1.0 - s[1]) * SI) * c.itau_I;
        s[0] += c.dt * dE + c.sqrt_dt * n0 * rp[6];
        s[1] += c.dt * dI + c.sqrt_dt * n1 * rp[7];
        s[0] = fmax(0.0, fmin(1.0, s[0]));
        s[1] = fmax(0.0, fmin(1.0, s[1]));
    }
    // `step` split around the coupling input (see RWW::pre): the inhibitory
    // population does not read g.
    static constexpr int n_tmp = 1;
    CUBNM_HD static void scale_rp(const C&, double*) {}
    template <class E = ExpPoly>
    CUBNM_HD static void pre(const C& c, double* s, const double* gp, const double* rp, double n0, double n1,
                             double* t, const E& ex = E()) {
        (void)gp; (void)n0;
        t[0] = rp[0] * s[0] - rp[1] * s[1];
        const double SI = DIV(1.0, 1.0 + ex(-c.a_I * (rp[2] * s[0] - rp[3] * s[1] + rp[5] - c.mu_I)));
        const double dI = (-s[1] + (1.0 - s[1]) * SI) * c.itau_I;
        s[1] += c.dt * dI + c.sqrt_dt * n1 * rp[7];
        s[1] = fmax(0.0, fmin(1.0, s[1]));
    }
    template <class E = ExpPoly>
    CUBNM_HD static void post(const C& c, double* s, const double* gp, const double* rp, double g, double n0,
                              double, const double* t, const E& ex = E()) {
        const double SE = DIV(1.0, 1.0 + ex(-c.a_E * (t[0] + gp[0] * g + rp[4] - c.mu_E)));
        const double dE = (-s[0] + (1.0 - s[0]) * SE) * c.itau_E;
        s[0] += c.dt * dE + c.sqrt_dt * n0 * rp[6];
        s[0] = fmax(0.0, fmin(1.0, s[0]));
    }
};

// reduced Wong-Wang, excitatory population only ------------------------------
// State [x, r, S]; regional [w, I0, sigma]; global [G]; one noise per node.
struct RWWEx {
    static constexpr int n_state = 3, n_noise = 1, n_global = 1, n_regional = 3;
    static constexpr int conn_idx = 2, bold_idx = 2, n_gout_bool = 0, n_gout_int = 0;
    static constexpr bool is_osc = false;
    struct C { double dt, sqrt_dt, J_N, a, b, d, dt_itau, dt_gamma; };
    static void init_constants(C& c, double dt) {
        c.dt = dt; c.sqrt_dt = std::sqrt(dt);
        c.J_N = 0.2609; c.a = 270; c.b = 108; c.d = 0.154;
        const double gamma = (double)0.641 / (double)1000.0, tau = 100;
        c.dt_itau = dt * (1.0 / tau); c.dt_gamma = dt * gamma;
    }
    CUBNM_HD static void init(const C&, double* s, const double*) { s[2] = 0.001; }
    // rwwex.cpp:45-66
    CUBNM_HD static void step(const C& c, double* s, const double* gp, const double* rp, double g,
                              double n0, double) {
        s[0] = rp[0] * c.J_N * s[2] + gp[0] * c.J_N * g + rp[1];
        const double axb = c.a * s[0] - c.b;
        s[1] = axb / (1 - exp(-c.d * axb));
        const double dS = c.dt_gamma * ((1 - s[2]) * s[1]) - c.dt_itau * s[2] + n0 * c.sqrt_dt * rp[2];
        s[2] += dS;
        s[2] = fmax(0.0, fmin(1.0, s[2]));
    }
    // `step` split around the coupling input (see RWW::pre): only w J_N S
    // precedes it.
    static constexpr int n_tmp = 1;
    CUBNM_HD static void scale_rp(const C&, double*) {}
    template <class E = ExpPoly>
    CUBNM_HD static void pre(const C& c, double* s, const double*, const double* rp, double, double, double* t, const E& ex = E()) {
        t[0] = rp[0] * c.J_N * s[2];
    }
    template <class E = ExpPoly>
    CUBNM_HD static void post(const C& c, double* s, const double* gp, const double* rp, double g, double n0,
                              double, const double* t, const E& ex = E()) {
        s[0] = t[0] + gp[0] * c.J_N * g + rp[1];
        const double axb = c.a * s[0] - c.b;
        s[1] = DIV(axb, 1 - ex(-c.d * axb));
        const double dS = c.dt_gamma * ((1 - s[2]) * s[1]) - c.dt_itau * s[2] + n0 * c.sqrt_dt * rp[2];
        s[2] += dS;
        s[2] = fmax(0.0, fmin(1.0, s[2]));
    }
};

// One Balloon-Windkessel Euler step (bw.cpp:24-34): nu and q use the updated
// nu, both use the old f, and f advances with the updated x.
CUBNM_HD void bw_step(const BWConsts& b, double s, double& x, double& f, double& nu, double& q) {
    x = x + b.dt * (s - b.kappa * x - b.y * (f - 1.0));
    const double tmp_f = f + b.dt * x;
    nu = nu + b.dt_itau * (f - pow(nu, b.ialpha));
    q = q + b.dt_itau * (f * (1.0 - pow(b.oneminrho, (1.0 / f))) / b.rho - pow(nu, b.ialpha) * q / nu);
    f = tmp_f;
}

// The same BW step with the powers evaluated as exp(y * log(x)):
// nu^(1/alpha) = exp(ialpha * log(nu)) and (1-rho)^(1/f) = exp(log(1-rho) / f).
// This matches pow to a few ulp (the BW system is contracting, so the
// difference stays at rounding level) and costs a third of the device pow.
// Special values follow pow: log(0) = -inf -> 0, log(x<0) = NaN -> NaN.
// nu^(1/alpha) with the reference's alpha = 0.32 (bw.cpp:4-22; no entry point
// changes it, and 1/alpha = 3.125 exactly in FP64): nu^3 * nu^(1/8).  The
// eighth root comes from y = nu^(-1/8): three hardware reciprocal square
// roots give it to about 2^-22 (each is good to 2^-23), two Newton steps
// y <- y (1 + (1 - nu y^8)/8) (quadratic convergence) to rounding, and
// nu^(1/8) = 1/y by the refined reciprocal: within 4 ulp of pow(nu, 3.125)
// for nu in [2e-9, 5e8] (checked on 2e7 points with a starting error of up to
// 1e-6), 23 instructions instead of the 63 of three IEEE square roots.  nu <= 0 gives NaN (pow gives 0 at nu = 0, where the BW q update
// divides 0 by 0 one half-step later).  There is no general-alpha path: its
// log/exp coefficients, kept in registers across the integrators' loops,
// cost more registers than the rest of the BW step.
__device__ inline double bw_nu_pow(const BWConsts& b, double nu) {
    (void)b;
    double y = __builtin_amdgcn_rsq(__builtin_amdgcn_rsq(__builtin_amdgcn_rsq(nu)));
#pragma unroll
    for (int it = 0; it < 2; it++) {
        const double y2 = y * y, y4 = y2 * y2, y8 = y4 * y4;
        const double e = fma(-nu, y8, 1.0);
        y = fma(y, e * 0.125, y);
    }
    double r = __builtin_amdgcn_rcp(y);
    r = fma(fma(-y, r, 1.0), r, r);
    r = fma(fma(-y, r, 1.0), r, r);
    return (nu * nu) * nu * r;
}

// The BW step of bw_step_explog with fewer instructions: powers of nu as
// above, the divisions by f and nu through fast_div, 1/rho as a product.  The
// exponential is a functor (the C library's exp unless the caller passes its
// own).
struct ExpLib {
    __device__ double operator()(double x) const { return exp(x); }
};
template <class E = ExpLib>
__device__ inline void bw_step_fast(const BWConsts& b, double s, double& x, double& f, double& nu, double& q,
                                    const E& ex = E()) {
    x = x + b.dt * (s - b.kappa * x - b.y * (f - 1.0));
    const double tmp_f = f + b.dt * x;
    nu = nu + b.dt_itau * (f - bw_nu_pow(b, nu));
    const double nu_pow = bw_nu_pow(b, nu);
    q = q + b.dt_itau * (f * (1.0 - ex(fast_div(b.log_oneminrho, f))) * b.irho - fast_div(nu_pow * q, nu));
    f = tmp_f;
}

// bw_step_fast with nu^(1/alpha) carried between steps: the power of the old
// nu that the nu update reads is the power of the new nu that the previous
// step's q update computed, so one bw_nu_pow per step instead of two (the
// same values bit for bit).  nup enters as bw_nu_pow(nu) and leaves as the
// power of the updated nu.
template <class E = ExpLib>
__device__ inline void bw_step_carry(const BWConsts& b, double s, double& x, double& f, double& nu, double& q,
                                     double& nup, const E& ex = E()) {
    x = x + b.dt * (s - b.kappa * x - b.y * (f - 1.0));
    const double tmp_f = f + b.dt * x;
    nu = nu + b.dt_itau * (f - nup);
    nup = bw_nu_pow(b, nu);
    q = q + b.dt_itau * (f * (1.0 - ex(fast_div(b.log_oneminrho, f))) * b.irho - fast_div(nup * q, nu));
    f = tmp_f;
}

__device__ inline void bw_step_explog(const BWConsts& b, double s, double& x, double& f, double& nu, double& q) {
    x = x + b.dt * (s - b.kappa * x - b.y * (f - 1.0));
    const double tmp_f = f + b.dt * x;
    nu = nu + b.dt_itau * (f - exp(b.ialpha * log(nu)));
    const double nu_pow = exp(b.ialpha * log(nu));
    q = q + b.dt_itau * (f * (1.0 - exp(b.log_oneminrho / f)) / b.rho - nu_pow * q / nu);
    f = tmp_f;
}

// BOLD observation in the CPU engine's form (bnm.cpp:318).
CUBNM_HD double bold_signal(const BWConsts& b, double nu, double q) {
    return b.V_0 * (b.k1 * (1 - q) + b.k2 * (1 - q / nu) + b.k3 * (1 - nu));
}

}  // namespace cubnm
