// Oracle: FP64 CPU restatement of the cuBNM simulation core.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the timed CPU
// baseline.  Nothing in the cubnm_amd package links, loads or calls it; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// It restates, in this project's own code, the algorithm of the reference CPU
// engine (built upstream with -std=c++11 -O3 -m64 -fopenmp -DNOISE_SEGMENT):
//   * loop structure, Jacobi coupling update, BW cadence, BOLD write, state
//     means, numerical-FIC restart ........ src/ext/bnm.cpp:69-522
//   * OpenMP over simulations ............. src/ext/bnm.cpp:524-583
//   * session init, noise + shuffles ...... src/ext/bnm.cpp:585-715, src/ext/utils.cpp:74-94
//   * timing / loop-count arithmetic ...... include/cubnm/models/base.hpp:51-68
//   * base config keys .................... include/cubnm/models/base.hpp:111-124, 271-297
//   * BW constants and step ............... src/ext/models/bw.cpp:4-34
//   * rWW constants/config/step/FIC ....... src/ext/models/rww.cpp:16-313
//   * analytical FIC ...................... src/ext/models/rww_fic.cpp:105-231
//   * JR / Kuramoto / WC steps ............ src/ext/models/jr.cpp:14-133, kuramoto.cpp:14-61, wc.cpp:14-80
//   * rWWEx step .......................... src/ext/models/rwwex.cpp:14-66
//   * FC / FCD trils ...................... src/ext/fc.cpp:1-91
//   * pair count and FCD windows .......... src/ext/utils.cpp:6-72
//
// Third-party arithmetic the reference takes from GSL 2.7 (setup.py:227-240),
// restated here from the published algorithms:
//   * gsl_stats_correlation: one-pass (Welford) mean/cross-product recurrence.
//   * gsl_root_fsolver_brent + gsl_root_test_interval: Brent's (1973) zero
//     finder, one interpolation/bisection step per iterate, bracket test
//     |hi-lo| < epsabs + epsrel*min(|lo|,|hi|) (same-sign bracket), max 100 iterates.
//   * gsl_blas_ddot: plain sequential dot product.
// Deliberate deviation: a NaN correlation yields NaN in the output instead of
// the reference's NULL return and crash (fc.cpp:86-89, bnm.cpp:472-473).
//
// Pinning (tests/test_oracle_pins.py, tests/test_oracle_invariants.py):
// reference golden fixtures (tests/expected/sim/*.pkl.gz, decoded into
// tests/golden/sim_*.npz) for FC/FCD exactly and for the simulations
// statistically; numpy.corrcoef for Pearson; scipy.optimize.brentq for the FIC
// root; closed-form BW rest state; the Kuramoto random walk for noise
// indexing; the two-node Kuramoto conservation law and synchronisation
// closed form; the FIC target rate; node-permutation equivariance.
// Parity unpinned at the bit level: the simulations, because the fixtures'
// GCC 12 normal_distribution/shuffle stream is not reproducible with this
// image's libstdc++ (pinned statistically and by the invariants above), and
// the exact Brent iterate sequence (pinned to the bracket rule only).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// ---------------------------------------------------------------- constants
struct BW {
    double dt, rho, alpha, tau, y, kappa, V_0, k1, k2, k3;
    double ialpha, itau, oneminrho, dt_itau;
    void init(double bw_dt_s, double k1_, double k2_, double k3_, double V0_) {
        dt = bw_dt_s; rho = 0.34; alpha = 0.32; tau = 0.98;
        y = 1.0 / 0.41; kappa = 1.0 / 0.65;
        V_0 = V0_; k1 = k1_; k2 = k2_; k3 = k3_;
        ialpha = 1.0 / alpha; itau = 1.0 / tau; oneminrho = 1.0 - rho;
        dt_itau = dt * itau;
    }
};

struct Timing {
    int nodes, time_steps, BOLD_TR, states_sampling;
    double dt, bw_dt;  // dt in ms, bw_dt in s
    int bold_len, bold_size, BOLD_TR_iters, states_len, states_sampling_iters;
    int bw_it, inner_it;
    void set(int nodes_, int time_steps_, int BOLD_TR_, int states_sampling_, double dt_, double bw_dt_ms) {
        nodes = nodes_; time_steps = time_steps_; BOLD_TR = BOLD_TR_;
        states_sampling = states_sampling_; dt = dt_; bw_dt = bw_dt_ms / 1000.0;
        bold_len = time_steps / BOLD_TR;
        bold_size = bold_len * nodes;
        BOLD_TR_iters = (int)(((double)BOLD_TR / 1000.0) / bw_dt);
        states_len = time_steps / states_sampling;
        states_sampling_iters = (int)(((double)states_sampling / 1000.0) / bw_dt);
        bw_it = (int)(((double)time_steps / 1000.0) / bw_dt);
        inner_it = (int)((bw_dt * 1000) / dt);
    }
};

struct BaseConf {
    bool do_fc = true, do_fcd = true;
    int bold_remove_s = 30;
    bool exc_interhemispheric = true;
    int window_size = 10, window_step = 2;
    bool drop_edges = true;
    int noise_time_steps = 30000;
};

struct RWWConf {
    bool do_fic = true;
    int max_fic_trials = 0;
    int I_SAMPLING_START = 1000, I_SAMPLING_END = 10000;
    int I_SAMPLING_DURATION = 10000 - 1000 + 1;
    double init_delta = 0.02;
};

struct Conf {
    BaseConf base;
    RWWConf rww;
    bool modifies_params = false;
};

typedef std::map<std::string, std::string> KV;

static KV parse_config(const char* cfg) {
    KV m;
    if (!cfg) return m;
    std::string s(cfg);
    size_t pos = 0;
    while (pos < s.size()) {
        size_t end = s.find(';', pos);
        if (end == std::string::npos) end = s.size();
        std::string item = s.substr(pos, end - pos);
        size_t eq = item.find('=');
        if (eq != std::string::npos) m[item.substr(0, eq)] = item.substr(eq + 1);
        pos = end + 1;
    }
    return m;
}

static void apply_config(const KV& m, Conf& c, bool is_rww, double bw_dt_s) {
    for (KV::const_iterator it = m.begin(); it != m.end(); ++it) {
        const std::string& k = it->first;
        const std::string& v = it->second;
        if (k == "do_fc") c.base.do_fc = (bool)std::stoi(v);
        else if (k == "do_fcd") c.base.do_fcd = (bool)std::stoi(v);
        else if (k == "bold_remove_s") c.base.bold_remove_s = std::stoi(v);
        else if (k == "exc_interhemispheric") c.base.exc_interhemispheric = (bool)std::stoi(v);
        else if (k == "window_size") c.base.window_size = std::stoi(v);
        else if (k == "window_step") c.base.window_step = std::stoi(v);
        else if (k == "drop_edges") c.base.drop_edges = (bool)std::stoi(v);
        else if (k == "noise_time_steps") c.base.noise_time_steps = std::stoi(v);
    }
    if (!is_rww) return;
    // rww.cpp:54-71; the duration is refreshed after every key, as upstream.
    for (KV::const_iterator it = m.begin(); it != m.end(); ++it) {
        const std::string& k = it->first;
        const std::string& v = it->second;
        if (k == "do_fic") { c.rww.do_fic = (bool)std::stoi(v); c.modifies_params = true; }
        else if (k == "max_fic_trials") c.rww.max_fic_trials = std::stoi(v);
        else if (k == "I_SAMPLING_START") c.rww.I_SAMPLING_START = (int)((std::stoi(v) / 1000) / bw_dt_s);
        else if (k == "I_SAMPLING_END") c.rww.I_SAMPLING_END = (int)((std::stoi(v) / 1000) / bw_dt_s);
        else if (k == "init_delta") c.rww.init_delta = std::stod(v);
        c.rww.I_SAMPLING_DURATION = c.rww.I_SAMPLING_END - c.rww.I_SAMPLING_START + 1;
    }
}

// ---------------------------------------------------------------- models
// Each model: counts, init(), step(); rWW adds the numerical-FIC hooks.
struct RWW {
    static const int n_state = 6, n_inter = 7, n_noise = 2, n_global = 1, n_regional = 4;
    static const int conn_idx = 4, bold_idx = 4, n_gout_bool = 2, n_gout_int = 1;
    static const bool is_osc = false;
    double dt, sqrt_dt, a_E, b_E, d_E, a_I, b_I, d_I, gamma_E, gamma_I, tau_E, tau_I, I_0, w_E, w_I, w_II, I_ext;
    double w_E__I_0, w_I__I_0, b_a_ratio_E, itau_E, itau_I, dt_itau_E, dt_gamma_E, dt_itau_I, dt_gamma_I;
    double tau_E_s, tau_I_s, gamma_E_s, gamma_I_s, r_I_ss, r_E_ss, I_I_ss, I_E_ss, S_I_ss, S_E_ss;
    RWWConf conf;
    void init_constants(double dt_) {
        dt = dt_; sqrt_dt = std::sqrt(dt);
        a_E = 310; b_E = 125; d_E = 0.16; a_I = 615; b_I = 177; d_I = 0.087;
        gamma_E = (double)0.641 / (double)1000.0; gamma_I = (double)1.0 / (double)1000.0;
        tau_E = 100; tau_I = 10; I_0 = 0.382; w_E = 1.0; w_I = 0.7; w_II = 1.0; I_ext = 0.0;
        w_E__I_0 = w_E * I_0; w_I__I_0 = w_I * I_0; b_a_ratio_E = b_E / a_E;
        itau_E = 1.0 / tau_E; itau_I = 1.0 / tau_I;
        dt_itau_E = dt * itau_E; dt_gamma_E = dt * gamma_E;
        dt_itau_I = dt * itau_I; dt_gamma_I = dt * gamma_I;
        tau_E_s = 0.1; tau_I_s = 0.01; gamma_E_s = 0.641; gamma_I_s = 1.0;
        r_I_ss = 3.9218448633; r_E_ss = 3.0773270642; I_I_ss = 0.2528951325;
        I_E_ss = 0.3773805650; S_I_ss = 0.0392184486; S_E_ss = 0.1647572075;
    }
    // rww.cpp:149-170
    void init(double* s, double* iv, int* ext_int_sh, bool* ext_bool_sh) const {
        s[4] = 0.001; s[5] = 0.001;
        iv[4] = 0.0; iv[5] = conf.init_delta;
        ext_int_sh[0] = 0;
        ext_bool_sh[0] = conf.do_fic & (conf.max_fic_trials > 0);
        ext_bool_sh[1] = false;
    }
    // rww.cpp:172-185
    void restart(double* s, double* iv) const { s[4] = 0.001; s[5] = 0.001; iv[4] = 0.0; }
    // rww.cpp:187-225
    void step(double* s, double* iv, const double* gp, const double* rp, double g,
              const double* noise, long idx) const {
        s[0] = w_E__I_0 + rp[0] * rp[1] * s[4] + g * gp[0] * rp[1] - rp[2] * s[5];
        s[1] = w_I__I_0 + rp[1] * s[4] - w_II * s[5];
        iv[0] = a_E * s[0] - b_E;
        iv[1] = a_I * s[1] - b_I;
        s[2] = iv[0] / (1 - std::exp(-d_E * iv[0]));
        s[3] = iv[1] / (1 - std::exp(-d_I * iv[1]));
        iv[2] = noise[idx] * rp[3] * sqrt_dt + dt_gamma_E * ((1 - s[4]) * s[2]) - dt_itau_E * s[4];
        iv[3] = noise[idx + 1] * rp[3] * sqrt_dt + dt_gamma_I * s[3] - dt_itau_I * s[5];
        s[4] += iv[2];
        s[5] += iv[3];
        s[4] = std::max(0.0, std::min(1.0, s[4]));
        s[5] = std::max(0.0, std::min(1.0, s[5]));
    }
    // rww.cpp:226-257 (one node); returns that node's restart vote
    bool post_bw_node(double* s, double* iv, const int* ext_int_sh, const bool* ext_bool_sh,
                      double* rp, int bw_i) const {
        bool r = false;
        if (ext_bool_sh[0]) {
            if (((bw_i + 1) >= conf.I_SAMPLING_START) & ((bw_i + 1) <= conf.I_SAMPLING_END)) iv[4] += s[0];
            if ((bw_i + 1) == conf.I_SAMPLING_END) {
                iv[4] /= conf.I_SAMPLING_DURATION;
                iv[6] = iv[4] - b_a_ratio_E;
                if (std::fabs(iv[6] + 0.026) > 0.005) {
                    r = true;
                    if (ext_int_sh[0] < conf.max_fic_trials) {
                        if (iv[6] < -0.026) {
                            rp[2] -= iv[5];
                            iv[5] -= 0.001;
                            iv[5] = std::fmax(iv[5], 0.001);
                        } else {
                            rp[2] += iv[5];
                        }
                    }
                }
            }
        }
        return r;
    }
};

struct JR {
    static const int n_state = 8, n_inter = 9, n_noise = 1, n_global = 1, n_regional = 6;
    static const int conn_idx = 7, bold_idx = 6, n_gout_bool = 0, n_gout_int = 0;
    static const bool is_osc = false;
    double dt, sqrt_dt, A, B, a, b, v0, nu_max, r, p_min, p_max, mu, cmin, cmax, midpoint;
    void init_constants(double dt_) {
        dt = dt_; sqrt_dt = std::sqrt(dt);
        A = 3.25; B = 22.0; a = 0.1; b = 0.05; v0 = 6.0; nu_max = 0.0025; r = 0.56;
        p_min = 0.12; p_max = 0.32; mu = 0.22; cmin = 0.0; cmax = 0.005; midpoint = 6.0;
    }
    void init(double* s) const { for (int i = 0; i < 8; i++) s[i] = 0.0; }
    // jr.cpp:80-133
    void step(double* s, double* iv, const double* gp, const double* rp, double g,
              const double* noise, long idx) const {
        iv[0] = 2.0 * nu_max / (1.0 + std::exp(r * (v0 - (s[1] - s[2]))));
        iv[1] = 2.0 * nu_max / (1.0 + std::exp(r * (v0 - (rp[1] * rp[0] * s[0]))));
        iv[2] = 2.0 * nu_max / (1.0 + std::exp(r * (v0 - (rp[3] * rp[0] * s[0]))));
        iv[3] = s[3];
        iv[4] = s[4];
        iv[5] = s[5];
        iv[6] = A * a * iv[0] - 2.0 * a * s[3] - (std::pow(a, 2)) * s[0];
        iv[7] = A * a * (mu + rp[2] * rp[0] * iv[1] + gp[0] * g) - 2.0 * a * s[4] - (std::pow(a, 2)) * s[1];
        iv[8] = B * b * (rp[4] * rp[0] * iv[2]) - 2.0 * b * s[5] - (std::pow(b, 2)) * s[2];
        s[0] += dt * iv[3];
        s[1] += dt * iv[4];
        s[2] += dt * iv[5];
        s[3] += dt * iv[6] + sqrt_dt * noise[idx] * rp[5];
        s[4] += dt * iv[7];
        s[5] += dt * iv[8];
        s[6] = s[1] - s[2];
        s[7] = cmin + (cmax - cmin) / (1.0 + std::exp(r * (midpoint - (s[6]))));
    }
};

struct Kuramoto {
    static const int n_state = 1, n_inter = 1, n_noise = 1, n_global = 1, n_regional = 3;
    static const int conn_idx = 0, bold_idx = 0, n_gout_bool = 0, n_gout_int = 0;
    static const bool is_osc = true;
    double dt, sqrt_dt, twopi;
    void init_constants(double dt_) { dt = dt_; sqrt_dt = std::sqrt(dt); twopi = 2.0 * M_PI; }
    void init(double* s, const double* rp) const { s[0] = rp[0]; }
    // kuramoto.cpp:48-61
    void step(double* s, double* iv, const double* gp, const double* rp, double g,
              const double* noise, long idx) const {
        iv[0] = rp[1] + gp[0] * g;
        s[0] = s[0] + dt * iv[0] + sqrt_dt * noise[idx] * rp[2];
        s[0] = std::fmod(s[0], twopi);
    }
};

struct WC {
    static const int n_state = 2, n_inter = 4, n_noise = 2, n_global = 1, n_regional = 8;
    static const int conn_idx = 0, bold_idx = 0, n_gout_bool = 0, n_gout_int = 0;
    static const bool is_osc = false;
    double dt, sqrt_dt, tau_E, tau_I, itau_E, itau_I, a_E, a_I, mu_E, mu_I;
    void init_constants(double dt_) {
        dt = dt_; sqrt_dt = std::sqrt(dt);
        tau_E = 2.5; tau_I = 3.75; itau_E = 1.0 / tau_E; itau_I = 1.0 / tau_I;
        a_E = 1.5; a_I = 1.5; mu_E = 3.0; mu_I = 3.0;
    }
    void init(double* s) const { s[0] = 0.001; s[1] = 0.001; }
    // wc.cpp:51-80
    void step(double* s, double* iv, const double* gp, const double* rp, double g,
              const double* noise, long idx) const {
        iv[0] = 1.0 / (1.0 + std::exp(-a_E * (rp[0] * s[0] - rp[1] * s[1] + gp[0] * g + rp[4] - mu_E)));
        iv[1] = 1.0 / (1.0 + std::exp(-a_I * (rp[2] * s[0] - rp[3] * s[1] + rp[5] - mu_I)));
        iv[2] = (-s[0] + (1.0 - s[0]) * iv[0]) * itau_E;
        iv[3] = (-s[1] + (1.0 - s[1]) * iv[1]) * itau_I;
        s[0] += dt * iv[2] + sqrt_dt * noise[idx] * rp[6];
        s[1] += dt * iv[3] + sqrt_dt * noise[idx + 1] * rp[7];
        s[0] = std::max(0.0, std::min(1.0, s[0]));
        s[1] = std::max(0.0, std::min(1.0, s[1]));
    }
};

struct RWWEx {
    static const int n_state = 3, n_inter = 2, n_noise = 1, n_global = 1, n_regional = 3;
    static const int conn_idx = 2, bold_idx = 2, n_gout_bool = 0, n_gout_int = 0;
    static const bool is_osc = false;
    double dt, sqrt_dt, J_N, a, b, d, gamma, tau, itau, dt_itau, dt_gamma;
    void init_constants(double dt_) {
        dt = dt_; sqrt_dt = std::sqrt(dt);
        J_N = 0.2609; a = 270; b = 108; d = 0.154;
        gamma = (double)0.641 / (double)1000.0; tau = 100; itau = 1.0 / tau;
        dt_itau = dt * itau; dt_gamma = dt * gamma;
    }
    void init(double* s) const { s[0] = 0.0; s[1] = 0.0; s[2] = 0.001; }
    // rwwex.cpp:45-66: state [x, r, S], regional [w, I0, sigma]
    void step(double* s, double* iv, const double* gp, const double* rp, double g,
              const double* noise, long idx) const {
        s[0] = rp[0] * J_N * s[2] + gp[0] * J_N * g + rp[1];
        iv[0] = a * s[0] - b;
        s[1] = iv[0] / (1 - std::exp(-d * iv[0]));
        iv[1] = dt_gamma * ((1 - s[2]) * s[1]) - dt_itau * s[2] + noise[idx] * sqrt_dt * rp[2];
        s[2] += iv[1];
        s[2] = std::max(0.0, std::min(1.0, s[2]));
    }
};

// ---------------------------------------------------------------- statistics
// Pearson correlation by the one-pass mean / cross-product recurrence
// (the published algorithm of gsl_stats_correlation).
static double corr_onepass(const double* x, long sx, const double* y, long sy, long n) {
    double sum_xsq = 0.0, sum_ysq = 0.0, sum_cross = 0.0;
    double mean_x = x[0], mean_y = y[0];
    for (long i = 1; i < n; ++i) {
        double ratio = i / (i + 1.0);
        double dx = x[i * sx] - mean_x;
        double dy = y[i * sy] - mean_y;
        sum_xsq += dx * dx * ratio;
        sum_ysq += dy * dy * ratio;
        sum_cross += dx * dy * ratio;
        mean_x += dx / (i + 1.0);
        mean_y += dy / (i + 1.0);
    }
    return sum_cross / (std::sqrt(sum_xsq) * std::sqrt(sum_ysq));
}

// utils.cpp:6-19
static int fc_n_pairs(int nodes, bool exc_inter) {
    int n = (nodes * (nodes - 1)) / 2;
    if (exc_inter) {
        if (nodes % 2 != 0) throw std::runtime_error("exc_interhemispheric is set but number of nodes is not even");
        int rh = nodes / 2;
        n -= rh * rh;
    }
    return n;
}

// utils.cpp:21-72
static int dfc_windows(std::vector<int>& starts, std::vector<int>& ends, int corr_len, int bold_len,
                       int n_vols_remove, int step, int size, bool drop_edges) {
    int first, last;
    if (drop_edges) { first = size / 2; last = corr_len - 1 - (size / 2); }
    else { first = 0; last = corr_len - 1; }
    first += n_vols_remove; last += n_vols_remove;
    int n = 0;
    for (int c = first; c <= last; c += step) {
        int s = c - (size / 2); if (s < 0) s = 0;
        int e = c + (size / 2); if (e >= bold_len) e = bold_len - 1;
        starts.push_back(s); ends.push_back(e); n++;
    }
    if (n == 0) throw std::runtime_error("Number of dynamic FC windows is 0");
    return n;
}

// FC tril over rows [r0, r1) of a (rows x nodes) row-major BOLD matrix (fc.cpp:1-39)
static void fc_tril(const double* bold, int nodes, int r0, int r1, bool exc_inter, double* out) {
    int rh = nodes / 2, idx = 0;
    for (int i = 0; i < nodes; i++)
        for (int j = 0; j < i; j++) {
            if (exc_inter && ((i < rh) ^ (j < rh))) continue;
            out[idx++] = corr_onepass(bold + (long)r0 * nodes + i, nodes, bold + (long)r0 * nodes + j, nodes, r1 - r0);
        }
}

// ---------------------------------------------------------------- FIC
struct InhParams { double I0_I, w_EI, S_E_ss, w_II, gamma_I_s, tau_I_s; const RWW* m; };

static double phi_I(const RWW& m, double II) {
    return ((m.a_I * II) - m.b_I) / (1 - std::exp(-1 * m.d_I * ((m.a_I * II) - m.b_I)));
}
static double inh_curr_fixed_pts(double x, const InhParams& p) {
    return p.I0_I + p.w_EI * p.S_E_ss - p.w_II * p.gamma_I_s * p.tau_I_s * phi_I(*p.m, x) - x;
}

// Brent (1973) zero finder, driven like the reference's gsl_fsolve
// (rww_fic.cpp:62-102): set on [lo, hi], iterate, stop when the bracket
// passes |hi-lo| < 1e-3*min(|lo|,|hi|) or after 100 iterates; -1 on failure.
static double brent_root(const InhParams& P, double x_lo, double x_hi) {
    const double EPS = 2.2204460492503131e-16;
    double fl = inh_curr_fixed_pts(x_lo, P), fu = inh_curr_fixed_pts(x_hi, P);
    if ((fl < 0.0 && fu < 0.0) || (fl > 0.0 && fu > 0.0)) return -1;
    double a = x_lo, fa = fl, b = x_hi, fb = fu, c = x_hi, fc = fu;
    double d = x_hi - x_lo, e = x_hi - x_lo;
    double root = 0.5 * (x_lo + x_hi);
    for (int iter = 1; iter <= 100; iter++) {
        // ---- one Brent iterate
        bool ac_equal = false;
        double lo, hi;
        if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { ac_equal = true; c = a; fc = fa; d = b - a; e = b - a; }
        if (std::fabs(fc) < std::fabs(fb)) { ac_equal = true; a = b; b = c; c = a; fa = fb; fb = fc; fc = fa; }
        double tol = 0.5 * EPS * std::fabs(b);
        double m = 0.5 * (c - b);
        if (fb == 0) { root = b; lo = b; hi = b; }
        else if (std::fabs(m) <= tol) { root = b; if (b < c) { lo = b; hi = c; } else { lo = c; hi = b; } }
        else {
            if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) { d = m; e = m; }
            else {
                double p, q, r, s = fb / fa;
                if (ac_equal) { p = 2 * m * s; q = 1 - s; }
                else {
                    q = fa / fc; r = fb / fc;
                    p = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
                    q = (q - 1) * (r - 1) * (s - 1);
                }
                if (p > 0) q = -q; else p = -p;
                if (2 * p < std::min(3 * m * q - std::fabs(tol * q), std::fabs(e * q))) { e = d; d = p / q; }
                else { d = m; e = m; }
            }
            a = b; fa = fb;
            if (std::fabs(d) > tol) b += d; else b += (m > 0 ? +tol : -tol);
            fb = inh_curr_fixed_pts(b, P);
            root = b;
            double cc = c;
            if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) cc = a;
            if (b < cc) { lo = b; hi = cc; } else { lo = cc; hi = b; }
        }
        // ---- interval test (epsabs 0, epsrel 1e-3)
        double min_abs = ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) ? std::min(std::fabs(lo), std::fabs(hi)) : 0.0;
        double tolerance = 0.0 + 0.001 * min_abs;
        if (std::fabs(hi - lo) < tolerance) return root;
    }
    return -1;
}

// rww_fic.cpp:138-231 for one simulation. SC is (source k, target j) row-major.
// Returns true when unstable; w_IE_out is written only on success.
static bool analytical_fic(const RWW& m, const double* SC, int nodes, double G,
                           const double* w_EE, const double* w_EI, double* w_IE_out) {
    std::vector<double> wIE(nodes, 0.0), krow(nodes);
    const double I0_E = m.I_0 * m.w_E, I0_I = m.I_0 * m.w_I;
    for (int j = 0; j < nodes; j++) {
        InhParams P = {I0_I, w_EI[j], m.S_E_ss, m.w_II, m.gamma_I_s, m.tau_I_s, &m};
        double II = brent_root(P, 0.0, 2.0);
        if (II == -1) return true;
        double rI = phi_I(m, II);
        double SI = rI * m.tau_I_s * m.gamma_I_s;
        // K_EE row j: (SC^T[j][k] * G) * w_EI[j], plus w_EE[j] on the diagonal
        double dot = 0.0;
        for (int k = 0; k < nodes; k++) {
            double v = (SC[(long)k * nodes + j] * G) * w_EI[j];
            if (k == j) v += w_EE[j];
            dot += v * m.S_E_ss;
        }
        double w = (-1 / SI) * (m.I_E_ss - m.I_ext - I0_E - dot);
        if (w < 0) return true;
        wIE[j] = w;
    }
    std::copy(wIE.begin(), wIE.end(), w_IE_out);
    return false;
}

// ---------------------------------------------------------------- session
struct Dims {
    int bold_len, n_pairs, n_windows, n_window_pairs, states_len, n_vols_remove, corr_len;
    int noise_bw_it, noise_repeats, bw_it, inner_it, n_states_samples_remove;
    long long noise_size;
    std::vector<int> w_starts, w_ends;
};

template <class M> struct Traits {};

struct Inputs {
    int N_SIMS, nodes;
    const double* SC; const int* SC_indices; const double* SC_dist;
    double* gp; double* rp; const double* v_list;
    bool states_ts, do_delay;
};

struct Outputs {
    double *bold, *fc, *fcd, *states;
    unsigned char* gbool; int* gint;
};

// Noise segment and shuffles: bnm.cpp:659-707, utils.cpp:74-94.
static void make_noise(int seed, long long noise_size, int nodes, int noise_bw_it, int noise_repeats,
                       std::vector<double>& noise, std::vector<int>& sh_nodes, std::vector<int>& sh_ts) {
    std::mt19937 gen(seed);
    std::normal_distribution<double> nd(0, 1);
    noise.resize(noise_size);
    for (long long i = 0; i < noise_size; i++) noise[i] = nd(gen);
    sh_nodes.resize((size_t)noise_repeats * nodes);
    sh_ts.resize((size_t)noise_repeats * noise_bw_it);
    std::vector<int> ni(nodes);
    std::iota(ni.begin(), ni.end(), 0);
    for (int r = 0; r < noise_repeats; r++) {
        std::shuffle(ni.begin(), ni.end(), gen);
        std::copy(ni.begin(), ni.end(), sh_nodes.begin() + (size_t)r * nodes);
    }
    std::vector<int> ti(noise_bw_it);
    std::iota(ti.begin(), ti.end(), 0);
    for (int r = 0; r < noise_repeats; r++) {
        std::shuffle(ti.begin(), ti.end(), gen);
        std::copy(ti.begin(), ti.end(), sh_ts.begin() + (size_t)r * noise_bw_it);
    }
}

static void compute_dims(const Timing& t, const Conf& c, int n_noise, Dims& d) {
    d.bold_len = t.bold_len;
    d.states_len = t.states_len;
    d.bw_it = t.bw_it; d.inner_it = t.inner_it;
    d.n_states_samples_remove = c.base.bold_remove_s * 1000 / t.states_sampling;
    d.n_pairs = d.n_windows = d.n_window_pairs = 0;
    d.n_vols_remove = d.corr_len = 0;
    if (c.base.do_fc) {
        d.n_vols_remove = c.base.bold_remove_s * 1000 / t.BOLD_TR;
        d.corr_len = t.bold_len - d.n_vols_remove;
        if (d.corr_len < 2) throw std::runtime_error("Number of BOLD volumes (after removing initial volumes) is too low for FC calculations");
        d.n_pairs = fc_n_pairs(t.nodes, c.base.exc_interhemispheric);
        if (c.base.do_fcd) {
            d.n_windows = dfc_windows(d.w_starts, d.w_ends, d.corr_len, t.bold_len, d.n_vols_remove,
                                      c.base.window_step, c.base.window_size, c.base.drop_edges);
            d.n_window_pairs = (d.n_windows * (d.n_windows - 1)) / 2;
        }
    }
    d.noise_bw_it = (int)(((double)(c.base.noise_time_steps) / 1000.0) / t.bw_dt);
    d.noise_size = (long long)t.nodes * d.noise_bw_it * t.inner_it * n_noise;
    d.noise_repeats = (int)std::ceil((double)(t.bw_it) / (double)(d.noise_bw_it));
}

// One simulation (bnm.cpp:69-522).
template <class M>
static void simulate(const M& model, const Timing& t, const Conf& c, const Dims& d, const BW& bw,
                     const Inputs& in, const Outputs& out, int sim,
                     const double* noise, const int* sh_nodes, const int* sh_ts) {
    const int N = t.nodes;
    const int NS = M::n_state, NI = M::n_inter, NR = M::n_regional, NG = M::n_global;
    const double* SC = in.SC + (long)in.SC_indices[sim] * N * N;
    std::vector<double> gp(NG), rp((size_t)N * NR), st((size_t)N * NS, 0.0), iv((size_t)N * NI, 0.0);
    for (int i = 0; i < NG; i++) gp[i] = in.gp[(long)i * in.N_SIMS + sim];
    for (int j = 0; j < N; j++)
        for (int i = 0; i < NR; i++) rp[(size_t)j * NR + i] = in.rp[(long)i * in.N_SIMS * N + (long)sim * N + j];
    int ext_int_sh[2] = {0, 0};
    bool ext_bool_sh[2] = {false, false};
    long states_stride = in.states_ts ? (long)t.states_len * N : N;
    double* st_out = out.states;  // layout (n_state, N_SIMS, states_stride)
    if (!in.states_ts)
        for (int j = 0; j < N; j++)
            for (int i = 0; i < NS; i++) st_out[(long)i * in.N_SIMS * states_stride + (long)sim * states_stride + j] = 0;
    for (int j = 0; j < N; j++) init_node(model, &st[(size_t)j * NS], &iv[(size_t)j * NI], &rp[(size_t)j * NR], ext_int_sh, ext_bool_sh);
    std::vector<double> bx(N, 0.0), bf(N, 1.0), bn(N, 1.0), bq(N, 1.0);

    // delay (bnm.cpp:146-216)
    std::vector<int> delay;
    int max_delay = 0;
    double velocity = in.v_list[sim] * t.dt;
    if (in.do_delay) {
        delay.assign((size_t)N * N, 0);
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) {
                double len = in.SC_dist[(long)i * N + j];
                if (i > j) {
                    int cd = (int)rintf((float)(len / velocity));
                    cd = std::max(cd, 1);
                    delay[(size_t)i * N + j] = cd; delay[(size_t)j * N + i] = cd;
                    if (cd > max_delay) max_delay = cd;
                } else if (i == j) {
                    delay[(size_t)i * N + j] = 1;
                }
            }
    }
    const bool has_delay = max_delay > 0;
    int buff_idx = max_delay - 1;
    std::vector<double> hist(has_delay ? (size_t)N * max_delay : (size_t)N);
    for (int j = 0; j < N; j++) {
        double v = st[(size_t)j * NS + M::conn_idx];
        if (has_delay) for (int bi = 0; bi < max_delay; bi++) hist[(size_t)bi * N + j] = v;
        else hist[j] = v;
    }
    std::vector<double> bold_local;
    if (c.base.do_fc) bold_local.assign((size_t)t.bold_len * N, 0.0);
    std::vector<double> g(N);
    double* BOLD = out.bold + (long)sim * t.bold_size;

    bool restart = false;
    int bw_i = 0, bold_i = 0, states_i = 0, rep = 0;
    while (bw_i < t.bw_it) {
        int sh_t = sh_ts[(bw_i % d.noise_bw_it) + rep * d.noise_bw_it];
        for (int inner_i = 0; inner_i < t.inner_it; inner_i++) {
            for (int j = 0; j < N; j++) {
                double acc = 0;
                if (has_delay) {
                    for (int k = 0; k < N; k++) {
                        int kb = (buff_idx + delay[(size_t)k * N + j]) % max_delay;
                        if (M::is_osc) acc += SC[(long)k * N + j] * std::sin(hist[(size_t)kb * N + k] - hist[(size_t)buff_idx * N + j]);
                        else acc += SC[(long)k * N + j] * hist[(size_t)kb * N + k];
                    }
                } else {
                    for (int k = 0; k < N; k++) {
                        if (M::is_osc) acc += SC[(long)k * N + j] * std::sin(hist[k] - hist[j]);
                        else acc += SC[(long)k * N + j] * hist[k];
                    }
                }
                g[j] = acc;
            }
            for (int j = 0; j < N; j++) {
                int sh_j = sh_nodes[rep * N + j];
                long nidx = (((long)(sh_t * t.inner_it + inner_i) * N * M::n_noise) + ((long)sh_j * M::n_noise));
                model.step(&st[(size_t)j * NS], &iv[(size_t)j * NI], gp.data(), &rp[(size_t)j * NR], g[j], noise, nidx);
            }
            if (has_delay) {
                for (int j = 0; j < N; j++) hist[(size_t)buff_idx * N + j] = st[(size_t)j * NS + M::conn_idx];
                buff_idx = (buff_idx + max_delay - 1) % max_delay;
            } else {
                for (int j = 0; j < N; j++) hist[j] = st[(size_t)j * NS + M::conn_idx];
            }
        }
        // BW (bw.cpp:24-34)
        for (int j = 0; j < N; j++) {
            double s = st[(size_t)j * NS + M::bold_idx];
            bx[j] = bx[j] + bw.dt * (s - bw.kappa * bx[j] - bw.y * (bf[j] - 1.0));
            double tmp_f = bf[j] + bw.dt * bx[j];
            bn[j] = bn[j] + bw.dt_itau * (bf[j] - std::pow(bn[j], bw.ialpha));
            bq[j] = bq[j] + bw.dt_itau * (bf[j] * (1.0 - std::pow(bw.oneminrho, (1.0 / bf[j]))) / bw.rho - std::pow(bn[j], bw.ialpha) * bq[j] / bn[j]);
            bf[j] = tmp_f;
        }
        if ((bw_i + 1) % t.BOLD_TR_iters == 0) {
            for (int j = 0; j < N; j++) {
                double b = bw.V_0 * (bw.k1 * (1 - bq[j]) + bw.k2 * (1 - bq[j] / bn[j]) + bw.k3 * (1 - bn[j]));
                BOLD[(long)bold_i * N + j] = b;
                if (c.base.do_fc) bold_local[(size_t)bold_i * N + j] = b;
            }
            bold_i++;
        }
        if ((bw_i + 1) % t.states_sampling_iters == 0) {
            for (int j = 0; j < N; j++) {
                for (int i = 0; i < NS; i++) {
                    double* o = st_out + (long)i * in.N_SIMS * states_stride + (long)sim * states_stride;
                    if (in.states_ts) o[(long)states_i * N + j] = st[(size_t)j * NS + i];
                    else if (states_i >= d.n_states_samples_remove) o[j] += st[(size_t)j * NS + i];
                }
            }
            states_i++;
        }
        if ((bw_i + 1) % d.noise_bw_it == 0) {
            if (bw_i + 1 < t.bw_it) rep++;
        }
        post_bw(model, st, iv, rp, ext_int_sh, ext_bool_sh, restart, bw_i, N);
        bw_i++;
        if (restart) {
            for (int j = 0; j < N; j++) restart_node(model, &st[(size_t)j * NS], &iv[(size_t)j * NI], &rp[(size_t)j * NR]);
            for (int j = 0; j < N; j++) {
                bx[j] = 0.0; bf[j] = 1.0; bn[j] = 1.0; bq[j] = 1.0;
                double v = st[(size_t)j * NS + M::conn_idx];
                if (has_delay) for (int bi = 0; bi < max_delay; bi++) hist[(size_t)bi * N + j] = v;
                else hist[j] = v;
            }
            bold_i = 0; states_i = 0; bw_i = 0;
            buff_idx = max_delay - 1;
            rep = 0;
            restart = false;
        }
    }
    post_integration(model, in, out, rp, ext_int_sh, ext_bool_sh, sim, N);
    if (!in.states_ts) {
        int tp = states_i - d.n_states_samples_remove;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < NS; i++) st_out[(long)i * in.N_SIMS * states_stride + (long)sim * states_stride + j] /= tp;
    }
    if (c.base.do_fc) {
        fc_tril(bold_local.data(), N, d.n_vols_remove, t.bold_len, c.base.exc_interhemispheric, out.fc + (long)sim * d.n_pairs);
        if (c.base.do_fcd) {
            std::vector<double> wfc((size_t)d.n_windows * d.n_pairs);
            for (int w = 0; w < d.n_windows; w++)
                fc_tril(bold_local.data(), N, d.w_starts[w], d.w_ends[w] + 1, c.base.exc_interhemispheric, &wfc[(size_t)w * d.n_pairs]);
            double* fcd = out.fcd + (long)sim * d.n_window_pairs;
            int idx = 0;
            for (int wi = 0; wi < d.n_windows; wi++)
                for (int wj = 0; wj < wi; wj++)
                    fcd[idx++] = corr_onepass(&wfc[(size_t)wi * d.n_pairs], 1, &wfc[(size_t)wj * d.n_pairs], 1, d.n_pairs);
        }
    }
}

// per-model hook dispatch
static void init_node(const RWW& m, double* s, double* iv, double*, int* ei, bool* eb) { m.init(s, iv, ei, eb); }
static void init_node(const JR& m, double* s, double*, double*, int*, bool*) { m.init(s); }
static void init_node(const Kuramoto& m, double* s, double*, double* rp, int*, bool*) { m.init(s, rp); }
static void init_node(const WC& m, double* s, double*, double*, int*, bool*) { m.init(s); }
static void init_node(const RWWEx& m, double* s, double*, double*, int*, bool*) { m.init(s); }
static void restart_node(const RWW& m, double* s, double* iv, double*) { m.restart(s, iv); }
static void restart_node(const JR& m, double* s, double*, double*) { m.init(s); }
static void restart_node(const Kuramoto& m, double* s, double*, double* rp) { m.init(s, rp); }
static void restart_node(const WC& m, double* s, double*, double*) { m.init(s); }
static void restart_node(const RWWEx& m, double* s, double*, double*) { m.init(s); }

// rww.cpp:259-291
static void post_bw(const RWW& m, std::vector<double>& st, std::vector<double>& iv, std::vector<double>& rp,
                    int* ei, bool* eb, bool& restart, int bw_i, int N) {
    for (int j = 0; j < N; j++) {
        bool r = m.post_bw_node(&st[(size_t)j * RWW::n_state], &iv[(size_t)j * RWW::n_inter], ei, eb,
                                &rp[(size_t)j * RWW::n_regional], bw_i);
        restart = restart || r;
    }
    if (eb[0] && ((bw_i + 1) == m.conf.I_SAMPLING_END)) {
        if (restart) {
            if (ei[0] < m.conf.max_fic_trials) ei[0]++;
            else { restart = false; eb[0] = false; eb[1] = true; }
        } else {
            eb[0] = false;
        }
    }
}
template <class M>
static void post_bw(const M&, std::vector<double>&, std::vector<double>&, std::vector<double>&,
                    int*, bool*, bool&, int, int) {}

// rww.cpp:293-313
static void post_integration(const RWW& m, const Inputs& in, const Outputs& out, std::vector<double>& rp,
                             int* ei, bool* eb, int sim, int N) {
    if (!m.conf.do_fic) return;
    for (int j = 0; j < N; j++) in.rp[2L * in.N_SIMS * N + (long)sim * N + j] = rp[(size_t)j * RWW::n_regional + 2];
    out.gint[sim] = ei[0];
    out.gbool[1L * in.N_SIMS + sim] = eb[1];
}
template <class M>
static void post_integration(const M&, const Inputs&, const Outputs&, std::vector<double>&, int*, bool*, int, int) {}

// rww.cpp:73-146 (prep_params); other models do nothing.
static void prep_params(const RWW& m, const Inputs& in, const Outputs& out, int N_SCs) {
    if (!m.conf.do_fic) return;
    const int N = in.nodes;
    std::vector<double> wEE(N), wEI(N), wIE(N);
    for (int sim = 0; sim < in.N_SIMS; sim++) {
        for (int j = 0; j < N; j++) {
            wEI[j] = in.rp[1L * in.N_SIMS * N + (long)sim * N + j];
            wEE[j] = in.rp[0L * in.N_SIMS * N + (long)sim * N + j] * wEI[j];
        }
        out.gbool[sim] = 0;
        const double* SC = in.SC + (long)in.SC_indices[sim] * N * N;
        bool unstable = analytical_fic(m, SC, N, in.gp[sim], wEE.data(), wEI.data(), wIE.data());
        out.gbool[sim] = unstable;
        for (int j = 0; j < N; j++)
            in.rp[2L * in.N_SIMS * N + (long)sim * N + j] = unstable ? 1.0 : wIE[j];
    }
    (void)N_SCs;
}
template <class M>
static void prep_params(const M&, const Inputs&, const Outputs&, int) {}

template <class M>
static void run_model(M& model, const Timing& t, const Conf& c, const Dims& d, const BW& bw,
                      const Inputs& in, const Outputs& out, int N_SCs, int sim_seed, int n_threads,
                      double* noise_out, int* shn_out, int* sht_out, double* times) {
    model.init_constants(t.dt);
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::vector<double> noise;
    std::vector<int> shn, sht;
    make_noise(sim_seed, d.noise_size, t.nodes, d.noise_bw_it, d.noise_repeats, noise, shn, sht);
    std::chrono::steady_clock::time_point t1 = std::chrono::steady_clock::now();
    if (noise_out) std::copy(noise.begin(), noise.end(), noise_out);
    if (shn_out) std::copy(shn.begin(), shn.end(), shn_out);
    if (sht_out) std::copy(sht.begin(), sht.end(), sht_out);
    const int nb = M::n_gout_bool, ni = M::n_gout_int;
    for (long i = 0; i < (long)nb * in.N_SIMS; i++) out.gbool[i] = 0;
    for (long i = 0; i < (long)ni * in.N_SIMS; i++) out.gint[i] = 0;
    std::chrono::steady_clock::time_point t2 = std::chrono::steady_clock::now();
    prep_params(model, in, out, N_SCs);
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(static)
#endif
    for (int sim = 0; sim < in.N_SIMS; sim++)
        simulate(model, t, c, d, bw, in, out, sim, noise.data(), shn.data(), sht.data());
    std::chrono::steady_clock::time_point t3 = std::chrono::steady_clock::now();
    if (times) {
        times[0] = std::chrono::duration<double>(t1 - t0).count();
        times[1] = std::chrono::duration<double>(t3 - t2).count();
    }
}

static std::string g_err;

static int model_counts(const std::string& name, int* n_state, int* n_global, int* n_regional, int* n_noise,
                        int* n_gb, int* n_gi) {
#define OR_COUNTS(M) { *n_state = M::n_state; *n_global = M::n_global; *n_regional = M::n_regional; \
    *n_noise = M::n_noise; *n_gb = M::n_gout_bool; *n_gi = M::n_gout_int; return 0; }
    if (name == "rWW") OR_COUNTS(RWW)
    if (name == "JR") OR_COUNTS(JR)
    if (name == "Kuramoto") OR_COUNTS(Kuramoto)
    if (name == "WC") OR_COUNTS(WC)
    if (name == "rWWEx") OR_COUNTS(RWWEx)
#undef OR_COUNTS
    return -1;
}

}  // namespace oracle

using namespace oracle;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// dims_out: [bold_len, n_pairs, n_windows, n_window_pairs, states_len, n_state_vars,
//            n_global, n_regional, n_gout_bool, n_gout_int, noise_bw_it, noise_repeats,
//            bw_it, inner_it, n_vols_remove, corr_len]
int oracle_dims(const char* model, const char* config, int nodes, int time_steps, int BOLD_TR,
                int states_sampling, double dt, double bw_dt, int* dims_out, long long* noise_size,
                int* window_starts, int* window_ends) {
    try {
        int ns, ng, nr, nn, nb, ni;
        if (model_counts(model, &ns, &ng, &nr, &nn, &nb, &ni) != 0) { g_err = "unknown model"; return -1; }
        Timing t; t.set(nodes, time_steps, BOLD_TR, states_sampling, dt, bw_dt);
        Conf c; apply_config(parse_config(config), c, std::string(model) == "rWW", t.bw_dt);
        Dims d; compute_dims(t, c, nn, d);
        int v[16] = {d.bold_len, d.n_pairs, d.n_windows, d.n_window_pairs, d.states_len, ns, ng, nr, nb, ni,
                     d.noise_bw_it, d.noise_repeats, d.bw_it, d.inner_it, d.n_vols_remove, d.corr_len};
        std::memcpy(dims_out, v, sizeof(v));
        *noise_size = d.noise_size;
        if (window_starts) std::copy(d.w_starts.begin(), d.w_starts.end(), window_starts);
        if (window_ends) std::copy(d.w_ends.begin(), d.w_ends.end(), window_ends);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

int oracle_run(const char* model, const char* config, int N_SIMS, int nodes, int N_SCs,
               const double* SC, const int* SC_indices, const double* SC_dist,
               double* global_params, double* regional_params, const double* v_list,
               int states_ts, int do_delay, int time_steps, int BOLD_TR, int states_sampling,
               int sim_seed, double dt, double bw_dt, const double* bw_consts, int n_threads,
               double* bold, double* fc, double* fcd, double* states, unsigned char* gbool, int* gint,
               double* noise_out, int* shn_out, int* sht_out, double* times) {
    try {
        std::string name(model);
        Timing t; t.set(nodes, time_steps, BOLD_TR, states_sampling, dt, bw_dt);
        Conf c; apply_config(parse_config(config), c, name == "rWW", t.bw_dt);
        BW bw; bw.init(t.bw_dt, bw_consts[0], bw_consts[1], bw_consts[2], bw_consts[3]);
        Inputs in = {N_SIMS, nodes, SC, SC_indices, SC_dist, global_params, regional_params, v_list,
                     states_ts != 0, do_delay != 0};
        Outputs out = {bold, fc, fcd, states, gbool, gint};
        Dims d;
        if (name == "rWW") {
            RWW m; m.conf = c.rww; compute_dims(t, c, RWW::n_noise, d);
            run_model(m, t, c, d, bw, in, out, N_SCs, sim_seed, n_threads, noise_out, shn_out, sht_out, times);
        } else if (name == "JR") {
            JR m; compute_dims(t, c, JR::n_noise, d);
            run_model(m, t, c, d, bw, in, out, N_SCs, sim_seed, n_threads, noise_out, shn_out, sht_out, times);
        } else if (name == "Kuramoto") {
            Kuramoto m; compute_dims(t, c, Kuramoto::n_noise, d);
            run_model(m, t, c, d, bw, in, out, N_SCs, sim_seed, n_threads, noise_out, shn_out, sht_out, times);
        } else if (name == "WC") {
            WC m; compute_dims(t, c, WC::n_noise, d);
            run_model(m, t, c, d, bw, in, out, N_SCs, sim_seed, n_threads, noise_out, shn_out, sht_out, times);
        } else if (name == "rWWEx") {
            RWWEx m; compute_dims(t, c, RWWEx::n_noise, d);
            run_model(m, t, c, d, bw, in, out, N_SCs, sim_seed, n_threads, noise_out, shn_out, sht_out, times);
        } else {
            g_err = "unknown model " + name;
            return -1;
        }
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// Single-root probe of the FIC solver for tests: returns I_I root for given J_N.
double oracle_fic_root(double w_EI) {
    RWW m; m.init_constants(0.1);
    InhParams P = {m.I_0 * m.w_I, w_EI, m.S_E_ss, m.w_II, m.gamma_I_s, m.tau_I_s, &m};
    return brent_root(P, 0.0, 2.0);
}

// Pearson by the one-pass recurrence, exposed for tests.
double oracle_corr(const double* x, const double* y, long n) { return corr_onepass(x, 1, y, 1, n); }

}  // extern "C"
