// Host side of the MI355X simulation core: the process-global session behind
// the C ABI in include/cubnm_amd.h.
//
// Mirrors the reference boundary src/ext/core.cpp:147-504 and the host parts
// of its GPU engine (src/ext/bnm.cu:507-1265): model/session selection,
// string config, loop-count arithmetic (include/cubnm/models/base.hpp:51-68),
// FC pair and FCD window tables (src/ext/utils.cpp:6-72), host noise
// generation with std::mt19937 + std::normal_distribution + std::shuffle
// (bnm.cu:1207-1263, utils.cpp:74-94), rWW analytical FIC
// (src/ext/models/rww.cpp:73-146, rww_fic.cpp:105-231), per-simulation
// conduction delays (bnm.cpp:146-216), then the HIP kernels.
#include "engine.hpp"
#include "../../include/cubnm_amd.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <map>
#include <set>
#include <sys/stat.h>
#include <memory>
#include <tuple>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace cubnm {

hipError_t launch_integrate_batched(int model, const IntegrateArgs& a, const void* consts, hipStream_t s,
                                    int* launches, const int* d_group_sims, int n_groups, const double* d_afrag,
                                    int want_tpw, int alds, double* d_lanes);
bool batched_afrag_fits_lds(int model, int nodes, int tpw);
bool batched_supported(int model, int nodes);
int batched_wide_grid(int n_groups);
size_t batched_wide_lane_doubles(int model, int nodes, int grid);
void batched_plan(int nodes, int want_tpw, int* tpw, int* nt_pad);
size_t batched_afrag_doubles(int nodes, int nt_pad, int n_scs);
hipError_t build_afrag(const double* d_SC, int n_scs, int nodes, int nt_pad, double* d_afrag, hipStream_t s);
extern int g_batched_tpw;
extern int g_general_pair;

namespace {

thread_local std::string g_err;

struct HipError : std::runtime_error {
    explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- BW constants
BWConsts g_bw;
bool g_bw_ready = false;

void bw_defaults(BWConsts& b) {
    b.dt = 0.001; b.rho = 0.34; b.alpha = 0.32; b.tau = 0.98; b.y = 1.0 / 0.41; b.kappa = 1.0 / 0.65;
    b.V_0 = 0.02 * 100; b.k1 = 7 * b.rho; b.k2 = 2.0; b.k3 = 2 * b.rho - 0.2;
    b.ialpha = 1.0 / b.alpha; b.itau = 1.0 / b.tau; b.oneminrho = 1.0 - b.rho; b.dt_itau = b.dt * b.itau;
}

// ---------------------------------------------------------------- device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n) {
        if (n <= bytes && p) return;
        if (p) check(hipFree(p), "hipFree");
        p = nullptr; bytes = 0;
        if (n == 0) return;
        check(hipMalloc(&p, n), "hipMalloc");
        bytes = n;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------- model table
struct ModelInfo {
    const char* name; int id, n_state, n_global, n_regional, n_noise, n_gb, n_gi;
};
const ModelInfo kModels[] = {
    {"rWW", MODEL_RWW, RWW::n_state, RWW::n_global, RWW::n_regional, RWW::n_noise, RWW::n_gout_bool, RWW::n_gout_int},
    {"JR", MODEL_JR, JR::n_state, JR::n_global, JR::n_regional, JR::n_noise, JR::n_gout_bool, JR::n_gout_int},
    {"Kuramoto", MODEL_KURAMOTO, Kuramoto::n_state, Kuramoto::n_global, Kuramoto::n_regional, Kuramoto::n_noise, 0, 0},
    {"WC", MODEL_WC, WC::n_state, WC::n_global, WC::n_regional, WC::n_noise, 0, 0},
    {"rWWEx", MODEL_RWWEX, RWWEx::n_state, RWWEx::n_global, RWWEx::n_regional, RWWEx::n_noise, 0, 0},
};

const ModelInfo* find_model(const char* name) {
    for (const auto& m : kModels) if (std::strcmp(m.name, name) == 0) return &m;
    return nullptr;
}

// ---------------------------------------------------------------- config
struct BaseConf {
    bool do_fc = true, do_fcd = true;
    int bold_remove_s = 30;
    bool exc_interhemispheric = true;
    int window_size = 10, window_step = 2;
    bool drop_edges = true;
    int noise_time_steps = 30000;
    bool ext_out = true, states_ts = false;
    bool verbose = false;
};

// ---------------------------------------------------------------- noise cache
// The noise segment depends only on (seed, nodes, noise width, segment
// length, inner steps per BW step); sessions that agree on these share one
// host and one device copy.  The generator state right after the normals is
// kept so each session can draw its own shuffles (whose count depends on the
// duration) exactly as one uninterrupted stream would (bnm.cu:1207-1263).
struct NoiseKey {
    int seed, nodes, n_noise, noise_bw_it, inner_it;
    bool operator<(const NoiseKey& o) const {
        return std::tie(seed, nodes, n_noise, noise_bw_it, inner_it) <
               std::tie(o.seed, o.nodes, o.n_noise, o.noise_bw_it, o.inner_it);
    }
    bool operator==(const NoiseKey& o) const { return !(*this < o) && !(o < *this); }
};

struct NoiseEntry {
    std::vector<double> noise;
    std::mt19937 after;
    DevBuf d_noise;
};

std::map<NoiseKey, std::weak_ptr<NoiseEntry>> g_noise_cache;

// Run the kernels on device-resident inputs already in s.d_* (or the caller's
// device pointers), writing device outputs.
struct DevIO {
    const double* SC; const int* SCi; double* gp; double* rp;
    double *bold, *fc, *fcd, *states; unsigned char* gb; int* gi;
};

// ---------------------------------------------------------------- session
struct Session {
    const ModelInfo* model = nullptr;
    bool initialized = false;
    // run shape
    int N_SIMS = 0, nodes = 0, N_SCs = 0, BOLD_TR = 0, states_sampling = 0, time_steps = 0, sim_seed = 0;
    bool do_delay = false, noise_out = false;
    double dt = 0.1, bw_dt = 0.001;  // ms, s
    BaseConf conf;
    bool modifies_params = false;
    RWW::C rww; JR::C jr; Kuramoto::C kur; WC::C wc; RWWEx::C rwwex;
    // derived
    int bold_len = 0, bold_size = 0, BOLD_TR_iters = 0, states_len = 0, states_sampling_iters = 0;
    int bw_it = 0, inner_it = 0, n_states_samples_remove = 0, n_vols_remove = 0, corr_len = 0;
    int n_pairs = 0, n_windows = 0, n_window_pairs = 0;
    std::vector<int> w_starts, w_ends, pairs_i, pairs_j, wpairs_i, wpairs_j;
    // noise: the segment is shared through the process-wide cache; the
    // per-repeat shuffles belong to the session (their count depends on the
    // duration)
    int noise_bw_it = 0, noise_repeats = 0;
    long long noise_size = 0;
    std::shared_ptr<NoiseEntry> noise_e;
    NoiseKey noise_key{};
    int shuffled_repeats = -1;
    std::vector<int> sh_nodes, sh_ts;
    int tpw = 0;              // batched tiles per wave for this session (0 = global knob)
    int tune_tpw = 0;         // batched tiles per wave chosen by the autotuner for the last request (0 = none)
    int alds = -1;            // batched SC fragments: -1 LDS when they fit, 0 streamed from L2
    int shape[5] = {-1, 0, 0, 0, 0};  // last launch: integrator, tiles per wave, fragments in LDS, k partitions, chosen by
    // device buffers
    DevBuf d_shn, d_sht, d_pairs, d_wins, d_wpairs;
    DevBuf d_SC, d_SCi, d_gp, d_rp, d_v, d_bold, d_fc, d_fcd, d_states, d_gb, d_gi, d_delays, d_maxd, d_hist, d_wfc;
    DevBuf d_groups, d_afrag, d_pair_of, d_gram, d_lanes;
    std::vector<int> groups_h;
    int last_fc_path = -1;    // 0 LDS kernel, 1 FP64 Gram, 2 split-BF16 Gram
    std::vector<unsigned char> gb_host;   // staged FIC flags (alive until the run completes)
    std::vector<int> delays_h, maxd_h;    // staged delays (alive until the run completes)
    double t_start = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    double timings[6] = {0, 0, 0, 0, 0, 0};
    int last_integrator = -1;
    int n_groups = 0;           // batched groups of the last request
    IntegrateArgs iargs{};      // integrator arguments of a staged request
    DevIO io{};                 // device buffers of a staged request
    double* h_gp = nullptr;     // caller's parameter arrays (write-back target)
    double* h_rp = nullptr;
    double merged_ms = -1;      // integrator time when run in a merged launch
};

Session g_s;
std::vector<std::unique_ptr<Session>> g_pool;   // sessions of the grouped entry point
int g_last_groups = 0;                           // group count of the last cubnm_run_groups call
int g_force_integrator = -1;  // -1 auto, 0 general, 1 batched
int g_fc_mode = 0;            // 0 auto (LDS kernel when a simulation fits), 1 FP64 Gram, 2 split-BF16 Gram
int g_general_ksplit = 0;     // partitions of the general kernel's coupling sum (0 auto)

int parse_int(const std::string& v) { return std::stoi(v); }

void apply_config(Session& s, const std::vector<std::pair<std::string, std::string>>& kv) {
    s.conf = BaseConf();
    for (const auto& p : kv) {
        const std::string& k = p.first; const std::string& v = p.second;
        if (k == "do_fc") s.conf.do_fc = (bool)parse_int(v);
        else if (k == "do_fcd") s.conf.do_fcd = (bool)parse_int(v);
        else if (k == "bold_remove_s") s.conf.bold_remove_s = parse_int(v);
        else if (k == "exc_interhemispheric") s.conf.exc_interhemispheric = (bool)parse_int(v);
        else if (k == "window_size") s.conf.window_size = parse_int(v);
        else if (k == "window_step") s.conf.window_step = parse_int(v);
        else if (k == "drop_edges") s.conf.drop_edges = (bool)parse_int(v);
        else if (k == "noise_time_steps") s.conf.noise_time_steps = parse_int(v);
        else if (k == "verbose") s.conf.verbose = (bool)parse_int(v);
        // progress_interval is accepted and ignored (the integrators are not polled)
    }
    s.modifies_params = false;
    if (s.model->id == MODEL_RWW) {
        // rww.cpp:54-71; the duration is recomputed after every key, as upstream
        for (const auto& p : kv) {
            const std::string& k = p.first; const std::string& v = p.second;
            if (k == "do_fic") { s.rww.do_fic = (bool)parse_int(v); s.modifies_params = true; }
            else if (k == "max_fic_trials") s.rww.max_fic_trials = parse_int(v);
            else if (k == "I_SAMPLING_START") s.rww.I_SAMPLING_START = (int)((parse_int(v) / 1000) / s.bw_dt);
            else if (k == "I_SAMPLING_END") s.rww.I_SAMPLING_END = (int)((parse_int(v) / 1000) / s.bw_dt);
            else if (k == "init_delta") s.rww.init_delta = std::stod(v);
            s.rww.I_SAMPLING_DURATION = s.rww.I_SAMPLING_END - s.rww.I_SAMPLING_START + 1;
        }
    }
}

// base.hpp:51-68 and bnm.cpp:631-657: double division, then truncation
void derive_shape(Session& s) {
    s.bold_len = s.time_steps / s.BOLD_TR;
    s.bold_size = s.bold_len * s.nodes;
    s.BOLD_TR_iters = (int)(((double)s.BOLD_TR / 1000.0) / s.bw_dt);
    s.states_len = s.time_steps / s.states_sampling;
    s.states_sampling_iters = (int)(((double)s.states_sampling / 1000.0) / s.bw_dt);
    s.bw_it = (int)(((double)s.time_steps / 1000.0) / s.bw_dt);
    s.inner_it = (int)((s.bw_dt * 1000) / s.dt);
    if (s.bw_it <= 0 || s.inner_it <= 0 || s.BOLD_TR_iters <= 0 || s.states_sampling_iters <= 0)
        throw std::runtime_error("invalid timing: duration, TR, states_sampling, dt and bw_dt give a zero loop count");
    s.n_states_samples_remove = s.conf.bold_remove_s * 1000 / s.states_sampling;
    s.n_pairs = s.n_windows = s.n_window_pairs = 0;
    s.pairs_i.clear(); s.pairs_j.clear(); s.w_starts.clear(); s.w_ends.clear();
    s.wpairs_i.clear(); s.wpairs_j.clear();
    if (!s.conf.do_fc) return;
    s.n_vols_remove = s.conf.bold_remove_s * 1000 / s.BOLD_TR;
    s.corr_len = s.bold_len - s.n_vols_remove;
    if (s.corr_len < 2)
        throw std::runtime_error("Number of BOLD volumes (after removing initial volumes) is too low for FC calculations");
    // utils.cpp:6-19 and bnm.cu:1093-1122
    const int N = s.nodes;
    int rh = N / 2;
    if (s.conf.exc_interhemispheric && (N % 2) != 0)
        throw std::runtime_error("exc_interhemispheric is set but number of nodes is not even");
    for (int i = 0; i < N; i++)
        for (int j = 0; j < i; j++) {
            if (s.conf.exc_interhemispheric && ((i < rh) ^ (j < rh))) continue;
            s.pairs_i.push_back(i); s.pairs_j.push_back(j);
        }
    s.n_pairs = (int)s.pairs_i.size();
    if (!s.conf.do_fcd) return;
    // utils.cpp:21-72
    int first, last;
    if (s.conf.drop_edges) { first = s.conf.window_size / 2; last = s.corr_len - 1 - (s.conf.window_size / 2); }
    else { first = 0; last = s.corr_len - 1; }
    first += s.n_vols_remove; last += s.n_vols_remove;
    if (s.conf.window_step <= 0) throw std::runtime_error("window_step must be positive");
    for (int c = first; c <= last; c += s.conf.window_step) {
        int ws = std::max(c - s.conf.window_size / 2, 0);
        int we = std::min(c + s.conf.window_size / 2, s.bold_len - 1);
        s.w_starts.push_back(ws); s.w_ends.push_back(we);
    }
    s.n_windows = (int)s.w_starts.size();
    if (s.n_windows == 0) throw std::runtime_error("Number of dynamic FC windows is 0");
    for (int i = 0; i < s.n_windows; i++)
        for (int j = 0; j < i; j++) { s.wpairs_i.push_back(i); s.wpairs_j.push_back(j); }
    s.n_window_pairs = (int)s.wpairs_i.size();
}

void init_model_constants(Session& s) {
    // Config values survive this call: keep them aside and restore.
    RWW::C keep = s.rww;
    RWW::init_constants(s.rww, s.dt);
    s.rww.do_fic = keep.do_fic; s.rww.max_fic_trials = keep.max_fic_trials;
    s.rww.I_SAMPLING_START = keep.I_SAMPLING_START; s.rww.I_SAMPLING_END = keep.I_SAMPLING_END;
    s.rww.I_SAMPLING_DURATION = keep.I_SAMPLING_DURATION; s.rww.init_delta = keep.init_delta;
    JR::init_constants(s.jr, s.dt);
    Kuramoto::init_constants(s.kur, s.dt);
    WC::init_constants(s.wc, s.dt);
    RWWEx::init_constants(s.rwwex, s.dt);
}

const void* model_consts(const Session& s) {
    switch (s.model->id) {
        case MODEL_RWW: return &s.rww;
        case MODEL_JR: return &s.jr;
        case MODEL_KURAMOTO: return &s.kur;
        case MODEL_RWWEX: return &s.rwwex;
        default: return &s.wc;
    }
}

// While set (cubnm_prepare_groups), ensure_noise records the keys it would
// generate instead of generating them, so that distinct segments can be
// generated in parallel; each is still one uninterrupted std::mt19937 stream.
bool g_noise_deferred = false;
std::vector<NoiseKey> g_deferred_keys;

// Host part of a noise entry: the segment's normals and the generator state
// after them (bnm.cu:1207-1263).
void generate_noise(const NoiseKey& key, NoiseEntry& e) {
    std::mt19937 gen(key.seed);
    std::normal_distribution<double> nd(0, 1);
    const long long size = (long long)key.nodes * key.noise_bw_it * key.inner_it * key.n_noise;
    e.noise.resize(size);
    for (long long i = 0; i < size; i++) e.noise[i] = nd(gen);
    e.after = gen;
}

// Noise segment and per-repeat shuffles from one std::mt19937 stream.
void ensure_noise(Session& s, bool force) {
    const int n_noise = s.model->n_noise;
    s.noise_bw_it = (int)(((double)(s.conf.noise_time_steps) / 1000.0) / s.bw_dt);
    if (s.noise_bw_it <= 0) throw std::runtime_error("noise segment shorter than one BW step");
    s.noise_size = (long long)s.nodes * s.noise_bw_it * s.inner_it * n_noise;
    s.noise_repeats = (int)std::ceil((double)(s.bw_it) / (double)(s.noise_bw_it));
    const NoiseKey key{s.sim_seed, s.nodes, n_noise, s.noise_bw_it, s.inner_it};
    if (force || !s.noise_e || !(s.noise_key == key)) {
        std::shared_ptr<NoiseEntry> e;
        auto it = g_noise_cache.find(key);
        if (!force && it != g_noise_cache.end()) e = it->second.lock();
        if (!e && g_noise_deferred && !force) {
            g_deferred_keys.push_back(key);
            s.noise_e.reset();
            return;
        }
        if (!e) {
            e = std::make_shared<NoiseEntry>();
            generate_noise(key, *e);
            e->d_noise.ensure(sizeof(double) * s.noise_size);
            check(hipMemcpy(e->d_noise.p, e->noise.data(), sizeof(double) * s.noise_size, hipMemcpyHostToDevice),
                  "noise H2D");
            g_noise_cache[key] = e;
        }
        s.noise_e = e;
        s.noise_key = key;
        s.shuffled_repeats = -1;
    }
    if (s.shuffled_repeats == s.noise_repeats) return;
    std::mt19937 gen = s.noise_e->after;
    s.sh_nodes.assign((size_t)s.noise_repeats * s.nodes, 0);
    s.sh_ts.assign((size_t)s.noise_repeats * s.noise_bw_it, 0);
    std::vector<int> idx(s.nodes);
    std::iota(idx.begin(), idx.end(), 0);
    for (int r = 0; r < s.noise_repeats; r++) {
        std::shuffle(idx.begin(), idx.end(), gen);
        std::copy(idx.begin(), idx.end(), s.sh_nodes.begin() + (size_t)r * s.nodes);
    }
    std::vector<int> tix(s.noise_bw_it);
    std::iota(tix.begin(), tix.end(), 0);
    for (int r = 0; r < s.noise_repeats; r++) {
        std::shuffle(tix.begin(), tix.end(), gen);
        std::copy(tix.begin(), tix.end(), s.sh_ts.begin() + (size_t)r * s.noise_bw_it);
    }
    s.d_shn.ensure(sizeof(int) * s.sh_nodes.size());
    check(hipMemcpy(s.d_shn.p, s.sh_nodes.data(), sizeof(int) * s.sh_nodes.size(), hipMemcpyHostToDevice), "shuffle H2D");
    s.d_sht.ensure(sizeof(int) * s.sh_ts.size());
    check(hipMemcpy(s.d_sht.p, s.sh_ts.data(), sizeof(int) * s.sh_ts.size(), hipMemcpyHostToDevice), "shuffle H2D");
    s.shuffled_repeats = s.noise_repeats;
}

void upload_tables(Session& s) {
    auto up = [](DevBuf& b, const std::vector<int>& a, const std::vector<int>& c) {
        b.ensure(sizeof(int) * std::max<size_t>(1, a.size() * 2));
        if (!a.empty()) {
            check(hipMemcpy(b.p, a.data(), sizeof(int) * a.size(), hipMemcpyHostToDevice), "table H2D");
            check(hipMemcpy(b.as<int>() + a.size(), c.data(), sizeof(int) * c.size(), hipMemcpyHostToDevice), "table H2D");
        }
    };
    up(s.d_pairs, s.pairs_i, s.pairs_j);
    up(s.d_wins, s.w_starts, s.w_ends);
    up(s.d_wpairs, s.wpairs_i, s.wpairs_j);
}

// ---------------------------------------------------------------- analytical FIC
// Brent's (1973) bracketing zero finder, iterated the way the reference
// drives gsl_root_fsolver_brent (rww_fic.cpp:62-102): one step per iterate,
// stop once the bracket satisfies |hi-lo| < 1e-3*min(|lo|,|hi|) (same sign),
// at most 100 iterates; -1 when the bracket does not straddle a root or the
// iteration does not converge.
double phi_I(const RWW::C& c, double II) {
    return ((c.a_I * II) - c.b_I) / (1 - std::exp(-1 * c.d_I * ((c.a_I * II) - c.b_I)));
}

double inh_fixed_point(const RWW::C& c, double w_EI, double x) {
    const double I0_I = c.I_0 * c.w_I;
    return I0_I + w_EI * c.S_E_ss - c.w_II * c.gamma_I_s * c.tau_I_s * phi_I(c, x) - x;
}

double fic_root(const RWW::C& c, double w_EI) {
    const double eps = 2.2204460492503131e-16;
    double lo0 = 0.0, hi0 = 2.0;
    double fa = inh_fixed_point(c, w_EI, lo0), fb = inh_fixed_point(c, w_EI, hi0);
    if ((fa < 0.0 && fb < 0.0) || (fa > 0.0 && fb > 0.0)) return -1;
    double a = lo0, b = hi0, cc = hi0, fc = fb, d = hi0 - lo0, e = hi0 - lo0;
    for (int it = 0; it < 100; it++) {
        bool interp3 = true;  // false when the third point coincides with a (secant step)
        if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) { interp3 = false; cc = a; fc = fa; d = b - a; e = b - a; }
        if (std::fabs(fc) < std::fabs(fb)) { interp3 = false; a = b; b = cc; cc = a; fa = fb; fb = fc; fc = fa; }
        const double tol = 0.5 * eps * std::fabs(b);
        const double m = 0.5 * (cc - b);
        double root, lo, hi;
        if (fb == 0) { root = lo = hi = b; }
        else if (std::fabs(m) <= tol) { root = b; lo = std::min(b, cc); hi = std::max(b, cc); if (b == cc) lo = hi = b; }
        else {
            if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) { d = m; e = m; }
            else {
                const double sr = fb / fa;
                double p, q;
                if (!interp3) { p = 2 * m * sr; q = 1 - sr; }
                else {
                    const double qa = fa / fc, rb = fb / fc;
                    p = sr * (2 * m * qa * (qa - rb) - (b - a) * (rb - 1));
                    q = (qa - 1) * (rb - 1) * (sr - 1);
                }
                if (p > 0) q = -q; else p = -p;
                if (2 * p < std::min(3 * m * q - std::fabs(tol * q), std::fabs(e * q))) { e = d; d = p / q; }
                else { d = m; e = m; }
            }
            a = b; fa = fb;
            b += (std::fabs(d) > tol) ? d : (m > 0 ? +tol : -tol);
            fb = inh_fixed_point(c, w_EI, b);
            root = b;
            const double other = ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) ? a : cc;
            lo = std::min(b, other); hi = std::max(b, other);
        }
        const bool same_sign = (lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0);
        const double min_abs = same_sign ? std::min(std::fabs(lo), std::fabs(hi)) : 0.0;
        if (std::fabs(hi - lo) < 0.001 * min_abs) return root;
    }
    return -1;
}

// rww.cpp:73-146 + rww_fic.cpp:138-231 for all simulations, threaded over sims.
// The root depends only on J_N, so roots are memoised per distinct J_N value.
void analytical_fic(const Session& s, const double* SC, const int* SCi, const double* gp, double* rp,
                    unsigned char* fic_unstable) {
    const int N = s.nodes, S = s.N_SIMS;
    const RWW::C& c = s.rww;
    std::unordered_map<uint64_t, double> roots;
    for (long long i = 0; i < (long long)S * N; i++) {
        const double jn = rp[1LL * S * N + i];
        uint64_t key; std::memcpy(&key, &jn, 8);
        if (!roots.count(key)) roots[key] = fic_root(c, jn);
    }
    const double I0_E = c.I_0 * c.w_E;
    auto work = [&](int s0, int s1) {
        std::vector<double> wIE(N);
        for (int sim = s0; sim < s1; sim++) {
            const double G = gp[sim];
            const double* sc = SC + (long long)SCi[sim] * N * N;
            bool unstable = false;
            for (int j = 0; j < N && !unstable; j++) {
                const double jn = rp[1LL * S * N + (long long)sim * N + j];
                const double wEE = rp[(long long)sim * N + j] * jn;
                uint64_t key; std::memcpy(&key, &jn, 8);
                const double II = roots.at(key);
                if (II == -1) { unstable = true; break; }
                const double SI = phi_I(c, II) * c.tau_I_s * c.gamma_I_s;
                double dot = 0.0;
                for (int k = 0; k < N; k++) {
                    double v = (sc[(long long)k * N + j] * G) * jn;
                    if (k == j) v += wEE;
                    dot += v * c.S_E_ss;
                }
                const double w = (-1 / SI) * (c.I_E_ss - c.I_ext - I0_E - dot);
                if (w < 0) { unstable = true; break; }
                wIE[j] = w;
            }
            fic_unstable[sim] = unstable;
            for (int j = 0; j < N; j++) rp[2LL * S * N + (long long)sim * N + j] = unstable ? 1.0 : wIE[j];
        }
    };
    int nt = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 64u));
    nt = std::min(nt, std::max(1, S / 8));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(work, (int)((long long)S * t / nt), (int)((long long)S * (t + 1) / nt));
    for (auto& t : th) t.join();
}

// bnm.cpp:146-216: symmetric integer delays from the lower triangle of
// SC_dist, diagonal 1, max over pairs; 0 when there is no pair.
void build_delays(const Session& s, const double* SC_dist, const double* v_list, std::vector<int>& delays,
                  std::vector<int>& maxd) {
    const int N = s.nodes;
    delays.assign((size_t)s.N_SIMS * N * N, 0);
    maxd.assign(s.N_SIMS, 0);
    for (int sim = 0; sim < s.N_SIMS; sim++) {
        const double vel = v_list[sim] * s.dt;
        int* d = &delays[(size_t)sim * N * N];
        int md = 0;
        for (int i = 0; i < N; i++)
            for (int j = 0; j <= i; j++) {
                if (i == j) { d[(size_t)i * N + j] = 1; continue; }
                int cd = (int)rintf((float)(SC_dist[(long long)i * N + j] / vel));
                cd = std::max(cd, 1);
                d[(size_t)i * N + j] = cd; d[(size_t)j * N + i] = cd;
                md = std::max(md, cd);
            }
        maxd[sim] = md;
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.f;
    return ms;
}

void ensure_stream(Session& s) {
    if (s.stream) return;
    check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
    for (auto& e : s.ev) check(hipEventCreate(&e), "hipEventCreate");
}


// Groups of 16 simulations sharing one SC for the batched integrator: sims
// are ordered by SC index (stable) and each SC's run is padded to a multiple
// of 16 with -1 entries.
int build_groups(Session& s, const int* SCi_host) {
    std::vector<int> order(s.N_SIMS);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return SCi_host[x] < SCi_host[y]; });
    s.groups_h.clear();
    for (size_t i = 0; i < order.size();) {
        size_t e = i;
        while (e < order.size() && SCi_host[order[e]] == SCi_host[order[i]]) e++;
        for (size_t b = i; b < e; b += 16)
            for (size_t k = 0; k < 16; k++) s.groups_h.push_back(b + k < e ? order[b + k] : -1);
        i = e;
    }
    s.d_groups.ensure(sizeof(int) * s.groups_h.size());
    check(hipMemcpyAsync(s.d_groups.p, s.groups_h.data(), sizeof(int) * s.groups_h.size(), hipMemcpyHostToDevice,
                         s.stream), "groups H2D");
    return (int)(s.groups_h.size() / 16);
}

// Integrator arguments for the session's request; uploads delays, decides
// batched vs per-simulation integrator and builds the batched groups and SC
// fragments (queued on the session's stream).
// Launch-shape cache shared by the autotuners of single requests and of the
// merged per-simulation launch: key -> choice, in memory and in a text file
// ($CUBNM_AMD_TUNE_CACHE, else ~/.cache/cubnm_amd/tune.txt) so later
// processes on the same machine reuse it.  The two integers of a choice
// depend on the key's kind (see tune_single and cubnm_run_groups).
struct TuneChoice { int ks = 0; int pair = 0; double trial_ms = 0; };
std::map<std::string, TuneChoice> g_tune_cache;
bool g_tune_loaded = false;
int g_autotune = 0;   // -1 never, 0 large launches, 1 always
int g_group_graph = 1;  // capture and replay the merged launch

std::string tune_cache_path() {
    if (const char* p = std::getenv("CUBNM_AMD_TUNE_CACHE")) return p;
    const char* h = std::getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/cubnm_amd/tune.txt";
}

void tune_cache_load() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    std::ifstream f(tune_cache_path());
    std::string key;
    TuneChoice c;
    while (f >> key >> c.ks >> c.pair >> c.trial_ms) g_tune_cache[key] = c;
}

void tune_cache_store(const std::string& key, const TuneChoice& c) {
    g_tune_cache[key] = c;
    try {
        const std::string path = tune_cache_path();
        const size_t slash = path.rfind('/');
        if (slash != std::string::npos) {
            std::string dir = path.substr(0, slash), cur;
            for (size_t i = 1; i <= dir.size(); i++)
                if (i == dir.size() || dir[i] == '/') { cur = dir.substr(0, i); mkdir(cur.c_str(), 0755); }
        }
        std::ofstream f(path, std::ios::app);
        f << key << " " << c.ks << " " << c.pair << " " << c.trial_ms << "\n";
    } catch (...) {
    }
}


// Autotuner of a single request that the batched integrator can run
// (no delay, no numerical-FIC restarts, up to 256 nodes): times, on the
// request's own launch cut to 50 BW steps (the real launch overwrites every
// output), each candidate of
//   simulations per workgroup: 16 (batched) or 1 (per-simulation);
//   nodes per lane (batched): 4 or 8 (one or two 16-node tiles per wave);
//   LDS layout (batched): SC fragments in LDS (when they fit) or streamed
//     from L2; (per-simulation) 1, 2 or 4 k partitions of the coupling sum,
//     SC staging and pair mode by the integrator's rule;
// and keeps the fastest, cached per (model, nodes, simulations rounded to a
// power of two, simulations per SC rounded down to a power of two).  Cache
// entry: ks = tiles per wave (0: per-simulation integrator), pair = SC
// fragments in LDS (batched) or k partitions (per-simulation).  Runs for
// launches of at least 2e10 node-steps (autotune 0) or always (autotune 1);
// the batched_tpw, general_ksplit and integrator knobs switch it off.
// Returns 1 when trials ran, 2 for a cached choice; sets the session's tiles
// per wave and fragment layout, or switches it to the per-simulation
// integrator with the chosen k partitions in `a`.
int tune_single(Session& s, IntegrateArgs& a, const DevIO& io) {
    const int N = s.nodes, mid = s.model->id;
    tune_cache_load();
    int sims_p2 = 1;
    while (sims_p2 < s.N_SIMS) sims_p2 <<= 1;
    int share = 1;
    while (share * 2 <= s.N_SIMS / s.N_SCs) share <<= 1;
    const std::string key = "single," + std::to_string(mid) + "|N" + std::to_string(N) + "|S" +
                            std::to_string(sims_p2) + "|per_sc" + std::to_string(share);
    TuneChoice best;
    int by = 2;
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end() && it->second.ks >= 0 && it->second.ks <= 2) {
        best = it->second;
    } else {
        by = 1;
        IntegrateArgs t = a;
        t.bw_it = std::min(a.bw_it, 50);
        int tp, nt_max;
        batched_plan(N, 2, &tp, &nt_max);
        s.d_afrag.ensure(sizeof(double) * batched_afrag_doubles(N, nt_max, s.N_SCs));
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "event");
        check(hipEventCreate(&e1), "event");
        best.trial_ms = 1e300;
        auto timed = [&](const char* what, auto&& launch) {
            check(hipEventRecord(e0, s.stream), "event");
            check(launch(), what);
            check(hipEventRecord(e1, s.stream), "event");
            check(hipEventSynchronize(e1), "autotune trial");
            return ev_ms(e0, e1);
        };
        for (int tpw : {1, 2}) {
            int tt, nt_pad;
            batched_plan(N, tpw, &tt, &nt_pad);
            check(build_afrag(io.SC, s.N_SCs, N, nt_pad, s.d_afrag.as<double>(), s.stream), "build_afrag launch");
            for (int alds : {1, 0}) {
                if (alds && !batched_afrag_fits_lds(mid, N, tpw)) continue;
                int launches = 1;
                const double ms = timed("autotune batched trial", [&] {
                    return launch_integrate_batched(mid, t, model_consts(s), s.stream, &launches, s.d_groups.as<int>(),
                                                    s.n_groups, s.d_afrag.as<double>(), tpw, alds,
                                                    s.d_lanes.as<double>());
                });
                if (ms < best.trial_ms) { best.trial_ms = ms; best.ks = tpw; best.pair = alds; }
            }
        }
        for (int ks : {1, 2, 4}) {
            if (ks * general_np(N) > 1024) continue;
            t.ksplit = ks;
            const double ms = timed("autotune per-simulation trial", [&] {
                return launch_integrate_general(mid, t, model_consts(s), s.stream,
                                                s.N_SIMS <= 512 || s.N_SIMS >= 4 * s.N_SCs);
            });
            if (ms < best.trial_ms) { best.trial_ms = ms; best.ks = 0; best.pair = ks; }
        }
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        tune_cache_store(key, best);
    }
    if (best.ks == 0) {
        s.last_integrator = 0;
        a.ksplit = best.pair;
    } else {
        s.tune_tpw = best.ks;
        s.alds = best.pair;
    }
    return by;
}

// Tiles per wave of the session's batched launches: the request's own
// setting, else the autotuner's choice, else the global knob / rule.
int eff_tpw(const Session& s) { return s.tpw != 0 ? s.tpw : s.tune_tpw; }

IntegrateArgs setup_integrate(Session& s, const DevIO& io, const std::vector<int>* delays,
                              const std::vector<int>* maxd, const int* SCi_host) {
    const int N = s.nodes;
    IntegrateArgs a{};
    a.N_SIMS = s.N_SIMS; a.nodes = N;
    a.bw_it = s.bw_it; a.inner_it = s.inner_it; a.BOLD_TR_iters = s.BOLD_TR_iters;
    a.states_sampling_iters = s.states_sampling_iters; a.n_states_samples_remove = s.n_states_samples_remove;
    a.noise_bw_it = s.noise_bw_it; a.states_ts = s.conf.states_ts; a.bold_len = s.bold_len;
    a.states_len = s.states_len; a.noise_repeats = s.noise_repeats;
    a.do_delay = s.do_delay; a.dt = s.dt;
    a.SC = io.SC; a.SC_indices = io.SCi; a.gp = io.gp; a.rp = io.rp;
    a.noise = s.noise_e->d_noise.as<double>(); a.sh_nodes = s.d_shn.as<int>(); a.sh_ts = s.d_sht.as<int>();
    a.bold = io.bold; a.states = io.states; a.gbool = io.gb; a.gint = io.gi;
    a.bw = g_bw;
    a.bw.dt = s.bw_dt; a.bw.dt_itau = a.bw.dt * a.bw.itau;  // core.cpp:291-292
    a.bw.log_oneminrho = std::log(a.bw.oneminrho);
    a.bw.irho = 1.0 / a.bw.rho;
    bool any_delay = false;
    if (s.do_delay) {
        int mmax = 0;
        for (int v : *maxd) mmax = std::max(mmax, v);
        any_delay = mmax > 0;
        s.d_delays.ensure(sizeof(int) * delays->size());
        check(hipMemcpyAsync(s.d_delays.p, delays->data(), sizeof(int) * delays->size(), hipMemcpyHostToDevice, s.stream), "delay H2D");
        s.d_maxd.ensure(sizeof(int) * maxd->size());
        check(hipMemcpyAsync(s.d_maxd.p, maxd->data(), sizeof(int) * maxd->size(), hipMemcpyHostToDevice, s.stream), "delay H2D");
        a.hist_stride = (long long)std::max(mmax, 1) * N;
        s.d_hist.ensure(sizeof(double) * a.hist_stride * s.N_SIMS);
        a.delays = s.d_delays.as<int>(); a.max_delays = s.d_maxd.as<int>(); a.hist = s.d_hist.as<double>();
    }
    a.do_delay = any_delay;
    a.ksplit = general_ksplit(N, g_general_ksplit);
    const bool fic_restarts = s.model->id == MODEL_RWW && s.rww.do_fic && s.rww.max_fic_trials > 0;
    bool use_batched = !any_delay && !fic_restarts && !s.conf.states_ts && batched_supported(s.model->id, N);
    if (use_batched && g_force_integrator != 1) {
        // groups of 16 share one SC: with many SCs and few simulations per SC
        // most lanes would be padding; past 2x padding the per-simulation
        // integrator does less work
        std::vector<int> per_sc(s.N_SCs, 0);
        for (int i = 0; i < s.N_SIMS; i++) per_sc[SCi_host[i]]++;
        long long slots = 0;
        for (int c : per_sc) slots += 16LL * ((c + 15) / 16);
        if (slots > 2LL * s.N_SIMS) use_batched = false;
    }
    if (g_force_integrator == 0) use_batched = false;
    if (g_force_integrator == 1 && !use_batched)
        throw std::runtime_error("batched integrator does not support this configuration (delay, numerical FIC, states_ts or node count)");
    if (N > 1024) {
        // the many-node integrator (integrate_wide_kernel)
        if (any_delay) throw std::runtime_error("conduction delay is not supported with more than 1024 nodes");
        if (fic_restarts)
            throw std::runtime_error("numerical FIC (max_fic_trials > 0) is not supported with more than 1024 nodes");
        use_batched = false;
    }
    s.last_integrator = use_batched ? 1 : (N > 1024 ? 2 : 0);
    s.n_groups = 0;
    s.tune_tpw = 0;
    s.alds = -1;
    int chosen_by = 0;
    if (use_batched) {
        s.n_groups = build_groups(s, SCi_host);
        const long long work = (long long)s.N_SIMS * s.bw_it * s.inner_it * N;
        if (N <= 256 && g_force_integrator < 0 && g_batched_tpw == 0 && s.tpw == 0 && g_general_ksplit == 0 &&
            (g_autotune > 0 || (g_autotune == 0 && work >= 20000000000LL)))
            chosen_by = tune_single(s, a, io);
    }
    if (s.last_integrator == 1) {
        int tpw, nt_pad;
        batched_plan(N, eff_tpw(s), &tpw, &nt_pad);
        s.d_afrag.ensure(sizeof(double) * batched_afrag_doubles(N, nt_pad, s.N_SCs));
        check(build_afrag(io.SC, s.N_SCs, N, nt_pad, s.d_afrag.as<double>(), s.stream), "build_afrag launch");
        if (N > 256)  // lane-ordered parameters and BW states of the wide integrator
            s.d_lanes.ensure(sizeof(double) *
                             batched_wide_lane_doubles(s.model->id, N, batched_wide_grid(s.n_groups)));
    }
    // launch shape for cubnm_last_launch_shape
    s.shape[0] = s.last_integrator;
    s.shape[1] = 0; s.shape[2] = 0; s.shape[3] = 0;
    if (s.last_integrator == 1 && N <= 256) {
        int tpw, nt_pad;
        batched_plan(N, eff_tpw(s), &tpw, &nt_pad);
        s.shape[1] = tpw;
        s.shape[2] = s.alds != 0 && batched_afrag_fits_lds(s.model->id, N, tpw);
    } else if (s.last_integrator == 0) {
        s.shape[3] = a.ksplit;
    }
    s.shape[4] = chosen_by;
    return a;
}

void launch_integrator(Session& s, const IntegrateArgs& a) {
    int launches = 1;
    check(hipEventRecord(s.ev[2], s.stream), "event");
    if (s.last_integrator == 2)
        check(launch_integrate_wide(s.model->id, a, model_consts(s), s.stream), "integrate_wide launch");
    else if (s.last_integrator == 1)
        check(launch_integrate_batched(s.model->id, a, model_consts(s), s.stream, &launches, s.d_groups.as<int>(),
                                       s.n_groups, s.d_afrag.as<double>(), eff_tpw(s), s.alds, s.d_lanes.as<double>()),
              "integrate_batched launch");
    else check(launch_integrate_general(s.model->id, a, model_consts(s), s.stream,
                                        s.N_SIMS <= 512 || s.N_SIMS >= 4 * s.N_SCs),
               "integrate_general launch");
    check(hipEventRecord(s.ev[3], s.stream), "event");
    s.timings[5] = launches;
}

// FC / FCD of the session's BOLD (queued on its stream).
void launch_fc(Session& s, const DevIO& io) {
    const int N = s.nodes;
    if (s.conf.do_fc) {
        FcArgs f{};
        f.N_SIMS = s.N_SIMS; f.nodes = N; f.bold_len = s.bold_len; f.n_vols_remove = s.n_vols_remove;
        f.n_pairs = s.n_pairs; f.n_windows = s.conf.do_fcd ? s.n_windows : 0;
        f.n_window_pairs = s.conf.do_fcd ? s.n_window_pairs : 0;
        f.exc_inter = s.conf.exc_interhemispheric;
        f.bold = io.bold;
        f.pairs_i = s.d_pairs.as<int>(); f.pairs_j = s.d_pairs.as<int>() + s.n_pairs;
        f.w_starts = s.d_wins.as<int>(); f.w_ends = s.d_wins.as<int>() + s.n_windows;
        f.wpairs_i = s.d_wpairs.as<int>(); f.wpairs_j = s.d_wpairs.as<int>() + s.n_window_pairs;
        f.fc = io.fc; f.fcd = io.fcd;
        int path = g_fc_mode;
        if (path == 0) path = fc_fcd_fits_lds(f) ? 0 : 1;
        s.last_fc_path = path;
        if (path == 0) {
            f.n_slots = std::min(s.N_SIMS, 2048);
            s.d_wfc.ensure(sizeof(double) * std::max<size_t>(1, (size_t)f.n_slots * f.n_windows * f.n_pairs));
            f.wfc_scratch = s.d_wfc.as<double>();
            check(launch_fc_fcd(f, s.stream), "fc_fcd launch");
        } else {
            int max_win = 1;
            for (int w = 0; w < s.n_windows; w++) max_win = std::max(max_win, s.w_ends[w] - s.w_starts[w] + 1);
            const int* pair_of = nullptr;
            if (s.conf.exc_interhemispheric) {
                std::vector<int> map((size_t)N * N, -1);
                for (int p = 0; p < s.n_pairs; p++) map[(size_t)s.pairs_i[p] * N + s.pairs_j[p]] = p;
                s.d_pair_of.ensure(sizeof(int) * map.size());
                check(hipMemcpyAsync(s.d_pair_of.p, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice,
                                     s.stream), "pair map upload");
                check(hipStreamSynchronize(s.stream), "pair map upload");
                pair_of = s.d_pair_of.as<int>();
            }
            const size_t per_sim = fc_gram_workspace_bytes(f, max_win, path == 2, 1);
            const size_t cap = std::max(per_sim, (size_t)4 << 30);  // scratch cap: 4 GiB or one simulation
            const size_t want = std::min(cap, per_sim * (size_t)s.N_SIMS);
            s.d_gram.ensure(want);
            check(launch_fc_gram(f, max_win, path == 2, pair_of, s.d_gram.p, s.d_gram.bytes, s.stream),
                  "fc_gram launch");
        }
    }
    check(hipEventRecord(s.ev[4], s.stream), "event");
}

void run_kernels(Session& s, const DevIO& io, const std::vector<int>* delays, const std::vector<int>* maxd,
                 const int* SCi_host) {
    const IntegrateArgs a = setup_integrate(s, io, delays, maxd, SCi_host);
    launch_integrator(s, a);
    launch_fc(s, io);
}

void require_session() {
    if (!g_s.initialized || !g_s.model) throw std::runtime_error("cubnm_prepare must be called before cubnm_run");
}

// core.cpp:194-386: model selection, config, loop counts, tables, noise.
void prepare_session(Session& s, const char* model_name, int n_config, const char* const* keys,
                     const char* const* values, int ext_out, int states_ts, int noise_out, int do_delay,
                     int force_reinit, int use_cpu, int N_SIMS, int nodes, int N_SCs, int time_steps, int BOLD_TR,
                     int states_sampling, int sim_seed, double dt, double bw_dt, cubnm_dims* dims) {
    if (!g_bw_ready) { bw_defaults(g_bw); g_bw_ready = true; }
    if (use_cpu)
        throw std::runtime_error("use_cpu is not supported: this library is the GPU simulation core "
                                 "(the CPU restatement is test infrastructure only)");
    const ModelInfo* m = find_model(model_name);
    if (!m) throw std::runtime_error(std::string("Model ") + model_name + " not found");
    if (N_SIMS <= 0 || nodes <= 0 || N_SCs <= 0) throw std::runtime_error("N_SIMS, nodes and N_SCs must be positive");
    if (nodes > 8192) throw std::runtime_error("nodes > 8192 is not supported");
    if (dt <= 0 || bw_dt <= 0 || BOLD_TR <= 0 || states_sampling <= 0 || time_steps <= 0)
        throw std::runtime_error("durations and steps must be positive");
    const double t0 = now_s();
    s.model = m;
    s.N_SIMS = N_SIMS; s.nodes = nodes; s.N_SCs = N_SCs; s.BOLD_TR = BOLD_TR;
    s.states_sampling = states_sampling; s.time_steps = time_steps; s.do_delay = do_delay != 0;
    s.sim_seed = sim_seed; s.dt = dt; s.bw_dt = bw_dt / 1000.0; s.noise_out = noise_out != 0;
    std::vector<std::pair<std::string, std::string>> kv;
    for (int i = 0; i < n_config; i++) kv.emplace_back(keys[i], values[i]);
    // rWW config defaults are reset on every call (rww.hpp:61-69)
    RWW::init_constants(s.rww, dt);
    apply_config(s, kv);
    s.conf.ext_out = ext_out != 0;
    s.conf.states_ts = states_ts != 0;
    init_model_constants(s);
    derive_shape(s);
    ensure_stream(s);
    ensure_noise(s, force_reinit != 0);
    upload_tables(s);
    s.initialized = true;
    dims->bold_len = s.bold_len; dims->n_pairs = s.n_pairs; dims->n_windows = s.n_windows;
    dims->n_window_pairs = s.n_window_pairs; dims->states_len = s.states_len;
    dims->n_state_vars = m->n_state; dims->n_global_params = m->n_global; dims->n_regional_params = m->n_regional;
    dims->n_global_out_bool = m->n_gb; dims->n_global_out_int = m->n_gi;
    dims->noise_bw_it = s.noise_bw_it; dims->noise_repeats = s.noise_repeats; dims->bw_it = s.bw_it;
    dims->inner_it = s.inner_it; dims->n_vols_remove = s.n_vols_remove; dims->corr_len = s.corr_len;
    dims->do_fc = s.conf.do_fc; dims->do_fcd = s.conf.do_fc && s.conf.do_fcd;
    dims->modifies_params = s.modifies_params; dims->N_SIMS = N_SIMS; dims->nodes = nodes;
    dims->noise_size = s.noise_size;
    dims->init_time = now_s() - t0;
}

// core.cpp:388-450 on host buffers, first part: host parameter preparation,
// device buffers, uploads and integrator set-up queued on the session's stream.
void stage_run(Session& s, const double* SC, const int* SC_indices, const double* SC_dist, double* gp, double* rp,
               const double* v_list) {
    s.t_start = now_s();
    const int N = s.nodes, S = s.N_SIMS;
    const ModelInfo* m = s.model;
    for (int i = 0; i < S; i++)
        if (SC_indices[i] < 0 || SC_indices[i] >= s.N_SCs) throw std::runtime_error("SC_indices out of range");
    // host parameter preparation (rWW analytical FIC)
    s.gb_host.assign((size_t)std::max(1, m->n_gb) * S, 0);
    const double tp0 = now_s();
    if (m->id == MODEL_RWW && s.rww.do_fic) analytical_fic(s, SC, SC_indices, gp, rp, s.gb_host.data());
    s.delays_h.clear(); s.maxd_h.clear();
    if (s.do_delay) build_delays(s, SC_dist, v_list, s.delays_h, s.maxd_h);
    s.timings[0] = (now_s() - tp0) * 1e3;
    const size_t scb = sizeof(double) * (size_t)s.N_SCs * N * N;
    const size_t gpb = sizeof(double) * (size_t)m->n_global * S;
    const size_t rpb = sizeof(double) * (size_t)m->n_regional * S * N;
    const long long st_stride = s.conf.states_ts ? (long long)s.states_len * N : N;
    const size_t stb = sizeof(double) * (size_t)m->n_state * S * st_stride;
    s.d_SC.ensure(scb); s.d_SCi.ensure(sizeof(int) * S); s.d_gp.ensure(gpb); s.d_rp.ensure(rpb);
    s.d_bold.ensure(sizeof(double) * (size_t)S * s.bold_size);
    s.d_fc.ensure(sizeof(double) * std::max<size_t>(1, (size_t)S * s.n_pairs));
    s.d_fcd.ensure(sizeof(double) * std::max<size_t>(1, (size_t)S * s.n_window_pairs));
    s.d_states.ensure(stb);
    s.d_gb.ensure(std::max<size_t>(1, s.gb_host.size()));
    s.d_gi.ensure(sizeof(int) * std::max<size_t>(1, (size_t)m->n_gi * S));
    check(hipEventRecord(s.ev[0], s.stream), "event");
    check(hipMemcpyAsync(s.d_SC.p, SC, scb, hipMemcpyHostToDevice, s.stream), "SC H2D");
    check(hipMemcpyAsync(s.d_SCi.p, SC_indices, sizeof(int) * S, hipMemcpyHostToDevice, s.stream), "SC_indices H2D");
    check(hipMemcpyAsync(s.d_gp.p, gp, gpb, hipMemcpyHostToDevice, s.stream), "params H2D");
    check(hipMemcpyAsync(s.d_rp.p, rp, rpb, hipMemcpyHostToDevice, s.stream), "params H2D");
    check(hipMemcpyAsync(s.d_gb.p, s.gb_host.data(), s.gb_host.size(), hipMemcpyHostToDevice, s.stream), "bools H2D");
    check(hipMemsetAsync(s.d_gi.p, 0, sizeof(int) * std::max<size_t>(1, (size_t)m->n_gi * S), s.stream), "memset");
    check(hipEventRecord(s.ev[1], s.stream), "event");
    s.io = DevIO{s.d_SC.as<double>(), s.d_SCi.as<int>(), s.d_gp.as<double>(), s.d_rp.as<double>(),
                 s.d_bold.as<double>(), s.d_fc.as<double>(), s.d_fcd.as<double>(), s.d_states.as<double>(),
                 s.d_gb.as<unsigned char>(), s.d_gi.as<int>()};
    s.iargs = setup_integrate(s, s.io, s.do_delay ? &s.delays_h : nullptr, s.do_delay ? &s.maxd_h : nullptr,
                              SC_indices);
    s.h_gp = gp; s.h_rp = rp;
}

// Last part: FC/FCD and downloads queued after the integrator.
void queue_outputs(Session& s, cubnm_outputs* out) {
    const int N = s.nodes, S = s.N_SIMS;
    const ModelInfo* m = s.model;
    launch_fc(s, s.io);
    const long long st_stride = s.conf.states_ts ? (long long)s.states_len * N : N;
    const size_t stb = sizeof(double) * (size_t)m->n_state * S * st_stride;
    const size_t gpb = sizeof(double) * (size_t)m->n_global * S;
    const size_t rpb = sizeof(double) * (size_t)m->n_regional * S * N;
    check(hipMemcpyAsync(out->sim_bold, s.d_bold.p, sizeof(double) * (size_t)S * s.bold_size, hipMemcpyDeviceToHost, s.stream), "BOLD D2H");
    if (s.conf.do_fc && out->sim_fc_trils)
        check(hipMemcpyAsync(out->sim_fc_trils, s.d_fc.p, sizeof(double) * (size_t)S * s.n_pairs, hipMemcpyDeviceToHost, s.stream), "FC D2H");
    if (s.conf.do_fc && s.conf.do_fcd && out->sim_fcd_trils)
        check(hipMemcpyAsync(out->sim_fcd_trils, s.d_fcd.p, sizeof(double) * (size_t)S * s.n_window_pairs, hipMemcpyDeviceToHost, s.stream), "FCD D2H");
    if (out->sim_states) check(hipMemcpyAsync(out->sim_states, s.d_states.p, stb, hipMemcpyDeviceToHost, s.stream), "states D2H");
    if (out->global_bools && m->n_gb > 0)
        check(hipMemcpyAsync(out->global_bools, s.d_gb.p, (size_t)m->n_gb * S, hipMemcpyDeviceToHost, s.stream), "bools D2H");
    if (out->global_ints && m->n_gi > 0)
        check(hipMemcpyAsync(out->global_ints, s.d_gi.p, sizeof(int) * (size_t)m->n_gi * S, hipMemcpyDeviceToHost, s.stream), "ints D2H");
    if (s.modifies_params) {
        check(hipMemcpyAsync(s.h_gp, s.d_gp.p, gpb, hipMemcpyDeviceToHost, s.stream), "params D2H");
        check(hipMemcpyAsync(s.h_rp, s.d_rp.p, rpb, hipMemcpyDeviceToHost, s.stream), "params D2H");
    }
    check(hipEventRecord(s.ev[5], s.stream), "event");
}

void launch_run(Session& s, const double* SC, const int* SC_indices, const double* SC_dist, double* gp, double* rp,
                const double* v_list, cubnm_outputs* out) {
    stage_run(s, SC, SC_indices, SC_dist, gp, rp, v_list);
    launch_integrator(s, s.iargs);
    queue_outputs(s, out);
}

// Second half: wait for the stream, timings, host-side noise outputs.
void finish_run(Session& s, cubnm_outputs* out) {
    check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    s.timings[1] = ev_ms(s.ev[0], s.ev[1]);
    s.timings[2] = s.merged_ms >= 0 ? s.merged_ms : ev_ms(s.ev[2], s.ev[3]);
    s.timings[3] = ev_ms(s.ev[3], s.ev[4]);
    s.timings[4] = ev_ms(s.ev[4], s.ev[5]);
    if (s.noise_out) {
        if (out->noise) std::copy(s.noise_e->noise.begin(), s.noise_e->noise.end(), out->noise);
        if (out->shuffled_nodes) std::copy(s.sh_nodes.begin(), s.sh_nodes.end(), out->shuffled_nodes);
        if (out->shuffled_ts) std::copy(s.sh_ts.begin(), s.sh_ts.end(), out->shuffled_ts);
    }
    out->run_time = now_s() - s.t_start;
}

// Relative cost of a group for launch ordering: node-steps times the
// coupling width.
double group_cost(const cubnm_group& g) {
    return (double)g.N_SIMS * g.nodes * ((double)g.time_steps / std::max(1e-9, g.dt)) * (2.0 * g.nodes + 42.0);
}

}  // namespace

void set_last_error(const std::string& e) { g_err = e; }

// Balloon-Windkessel constants as set by cubnm_init / cubnm_set_const, with
// the derived fields of the device steps (the caller sets dt and dt_itau).
BWConsts current_bw() {
    if (!g_bw_ready) { bw_defaults(g_bw); g_bw_ready = true; }
    BWConsts b = g_bw;
    b.log_oneminrho = std::log(b.oneminrho);
    b.irho = 1.0 / b.rho;
    return b;
}

}  // namespace cubnm

using namespace cubnm;

// ---------------------------------------------------------------------------
// Launch shape of the merged per-simulation launch: the autotuner and the
// captured graph.
//
// Autotuner: for a large merged launch (or always, with the "autotune" knob
// at 1) it times short trials of every candidate shape - k partitions of
// the coupling sum in {1, 2, 4} (as fit 1024 lanes) x pair mode {as the rule
// allows, off} - on the real launch with every request cut to a few BW
// steps, and keeps the fastest.  Choices are cached per (models, largest
// node count, simulations rounded to a power of two, SC-sharing) in memory
// and in a text file ($CUBNM_AMD_TUNE_CACHE, else ~/.cache/cubnm_amd/tune.txt)
// so later processes on the same machine reuse them.  Requests with
// numerical-FIC restarts are never cut into trials (their restarts write wIE
// back), so launches holding them keep the rule-based shape.
//
// Graph: the launch tables' uploads and the merged launch are captured into
// a hipGraph the first time a launch signature (every request's arguments,
// tables, block shape and LDS size) is seen and replayed while the next
// call's signature is identical (for example the generations of an
// optimiser, whose parameters change but whose buffers do not).  The tables
// live in pinned host memory owned by the cached graph.
struct GroupGraph {
    std::vector<char> sig;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    void* pinned[4] = {nullptr, nullptr, nullptr, nullptr};
    std::unique_ptr<DevBuf[]> bufs;
    int replays = 0;
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void*& p : pinned) if (p) { (void)hipHostFree(p); p = nullptr; }
        exec = nullptr; graph = nullptr;
        bufs.reset();
        sig.clear();
    }
};
GroupGraph g_group_graph_cache;
// last call: [0] graph (0 none, 1 captured, 2 replayed), [1] k partitions,
// [2] pair mode (1 allowed by the rule, 0 off), [3] tuned (0 rule, 1 trials
// run, 2 cached choice), [4] trial time ms, [5] candidates timed
double g_group_info[6] = {0, 0, 0, 0, 0, 0};

extern "C" {

const char* cubnm_last_error(void) { return g_err.c_str(); }

int cubnm_init(void) {
    bw_defaults(g_bw);
    g_bw_ready = true;
    return 0;
}

int cubnm_set_const(const char* key, double value) {
    if (!g_bw_ready) cubnm_init();
    if (std::strcmp(key, "k1") == 0) g_bw.k1 = value;
    else if (std::strcmp(key, "k2") == 0) g_bw.k2 = value;
    else if (std::strcmp(key, "k3") == 0) g_bw.k3 = value;
    else if (std::strcmp(key, "V_0") == 0) g_bw.V_0 = value;
    return 0;
}

int cubnm_force_integrator(int force) { g_force_integrator = force; return 0; }

int cubnm_set_tuning(const char* key, int value) {
    if (std::strcmp(key, "batched_tpw") == 0) {
        if (value != 0 && value != 1 && value != 2) { g_err = "batched_tpw must be 0 (auto), 1 or 2"; return -1; }
        g_batched_tpw = value;
        return 0;
    }
    if (std::strcmp(key, "general_ksplit") == 0) {
        if (value < 0 || value > 8) { g_err = "general_ksplit must be 0 (auto) or 1..8"; return -1; }
        g_general_ksplit = value;
        return 0;
    }
    if (std::strcmp(key, "general_pair") == 0) {
        if (value < -1 || value > 1) { g_err = "general_pair must be -1 (never), 0 (auto) or 1 (whenever eligible)"; return -1; }
        g_general_pair = value;
        return 0;
    }
    if (std::strcmp(key, "autotune") == 0) {
        if (value < -1 || value > 1) { g_err = "autotune must be -1 (never), 0 (large launches) or 1 (always)"; return -1; }
        g_autotune = value;
        return 0;
    }
    if (std::strcmp(key, "group_graph") == 0) {
        if (value < 0 || value > 1) { g_err = "group_graph must be 0 (off) or 1 (capture and replay)"; return -1; }
        g_group_graph = value;
        return 0;
    }
    if (std::strcmp(key, "fc_mode") == 0) {
        if (value < 0 || value > 2) { g_err = "fc_mode must be 0 (auto), 1 (FP64 Gram) or 2 (split-BF16 Gram)"; return -1; }
        g_fc_mode = value;
        return 0;
    }
    g_err = std::string("unknown tuning key ") + key;
    return -1;
}
int cubnm_copy_noise(double* noise, int* shuffled_nodes, int* shuffled_ts) {
    try {
        require_session();
        const Session& s = g_s;
        if (noise) std::copy(s.noise_e->noise.begin(), s.noise_e->noise.end(), noise);
        if (shuffled_nodes) std::copy(s.sh_nodes.begin(), s.sh_nodes.end(), shuffled_nodes);
        if (shuffled_ts) std::copy(s.sh_ts.begin(), s.sh_ts.end(), shuffled_ts);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int cubnm_last_integrator(void) { return g_s.last_integrator; }
int cubnm_last_launch_shape(int* out) {
    for (int i = 0; i < 5; i++) out[i] = g_s.shape[i];
    return 0;
}
int cubnm_last_fc_path(void) { return g_s.last_fc_path; }

int cubnm_last_timings(double* out6) {
    for (int i = 0; i < 6; i++) out6[i] = g_s.timings[i];
    return 0;
}

int cubnm_group_timings(int group, double* out6) {
    if (group < 0 || group >= g_last_groups || group >= (int)g_pool.size() || !g_pool[group]) {
        g_err = "no such group in the last cubnm_run_groups call";
        return -1;
    }
    for (int i = 0; i < 6; i++) out6[i] = g_pool[group]->timings[i];
    return 0;
}

int cubnm_prepare(const char* model_name, int n_config, const char* const* keys, const char* const* values,
                  int ext_out, int states_ts, int noise_out, int do_delay, int force_reinit, int use_cpu,
                  int N_SIMS, int nodes, int N_SCs, int time_steps, int BOLD_TR, int states_sampling,
                  int sim_seed, double dt, double bw_dt, cubnm_dims* dims) {
    try {
        prepare_session(g_s, model_name, n_config, keys, values, ext_out, states_ts, noise_out, do_delay,
                        force_reinit, use_cpu, N_SIMS, nodes, N_SCs, time_steps, BOLD_TR, states_sampling, sim_seed,
                        dt, bw_dt, dims);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

int cubnm_run(const double* SC, const int* SC_indices, const double* SC_dist, double* gp, double* rp,
              const double* v_list, cubnm_outputs* out) {
    try {
        require_session();
        launch_run(g_s, SC, SC_indices, SC_dist, gp, rp, v_list, out);
        finish_run(g_s, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        if (g_s.stream) hipStreamSynchronize(g_s.stream);
        return -1;
    }
}

int cubnm_prepare_groups(int n_groups, cubnm_group* groups) {
    g_deferred_keys.clear();
    try {
        if (n_groups <= 0) throw std::runtime_error("need at least one group");
        while ((int)g_pool.size() < n_groups) g_pool.emplace_back(new Session());
        // 1. sessions, with the generation of missing noise segments deferred
        g_noise_deferred = true;
        for (int i = 0; i < n_groups; i++) {
            cubnm_group& g = groups[i];
            try {
                prepare_session(*g_pool[i], g.model_name, g.n_config, g.config_keys, g.config_values, g.ext_out,
                                g.states_ts, g.noise_out, g.do_delay, g.force_reinit, 0, g.N_SIMS, g.nodes, g.N_SCs,
                                g.time_steps, g.BOLD_TR, g.states_sampling, g.sim_seed, g.dt, g.bw_dt, &g.dims);
                g_pool[i]->tpw = g.batched_tpw;
            } catch (const std::exception& e) {
                g_noise_deferred = false;
                throw std::runtime_error("group " + std::to_string(i) + ": " + e.what());
            }
        }
        g_noise_deferred = false;
        // 2. distinct missing segments generated in parallel (each segment one
        //    uninterrupted stream, segments taken in turn by at most 64
        //    threads), uploaded in order
        std::sort(g_deferred_keys.begin(), g_deferred_keys.end());
        g_deferred_keys.erase(std::unique(g_deferred_keys.begin(), g_deferred_keys.end()), g_deferred_keys.end());
        const double t0 = now_s();
        std::vector<std::shared_ptr<NoiseEntry>> fresh(g_deferred_keys.size());
        for (auto& e : fresh) e = std::make_shared<NoiseEntry>();
        {
            const size_t n_keys = fresh.size();
            const size_t n_workers =
                std::min(n_keys, (size_t)std::max(1u, std::min(std::thread::hardware_concurrency(), 64u)));
            std::atomic<size_t> next{0};
            std::vector<std::exception_ptr> failed(n_workers);
            std::vector<std::thread> workers;
            for (size_t w = 0; w < n_workers; w++)
                workers.emplace_back([&, w]() {
                    try {
                        for (size_t k = next.fetch_add(1); k < n_keys; k = next.fetch_add(1))
                            generate_noise(g_deferred_keys[k], *fresh[k]);
                    } catch (...) {
                        failed[w] = std::current_exception();
                    }
                });
            for (auto& w : workers) w.join();
            for (auto& f : failed)
                if (f) std::rethrow_exception(f);
        }
        for (size_t k = 0; k < fresh.size(); k++) {
            const size_t bytes = sizeof(double) * fresh[k]->noise.size();
            fresh[k]->d_noise.ensure(bytes);
            check(hipMemcpy(fresh[k]->d_noise.p, fresh[k]->noise.data(), bytes, hipMemcpyHostToDevice), "noise H2D");
            g_noise_cache[g_deferred_keys[k]] = fresh[k];
        }
        groups[0].dims.init_time += now_s() - t0;  // the shared generation is booked on the first group
        // 3. every session takes its segment from the cache and draws its shuffles
        for (int i = 0; i < n_groups; i++) {
            const double t1 = now_s();
            ensure_noise(*g_pool[i], false);
            groups[i].dims.init_time += now_s() - t1;
        }
        g_deferred_keys.clear();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}


int cubnm_run_groups(int n_groups, cubnm_group* groups) {
    int staged = 0;
    std::vector<int> order(std::max(0, n_groups));
    g_last_groups = 0;
    try {
        if (n_groups <= 0 || (int)g_pool.size() < n_groups) throw std::runtime_error("cubnm_prepare_groups first");
        for (int i = 0; i < n_groups; i++)
            if (!g_pool[i]->initialized || g_pool[i]->N_SIMS != groups[i].N_SIMS || g_pool[i]->nodes != groups[i].nodes)
                throw std::runtime_error("group " + std::to_string(i) + " does not match its prepared session");
        bool verbose = false;
        for (int i = 0; i < n_groups; i++) verbose = verbose || g_pool[i]->conf.verbose;
        const double t0 = now_s();
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return group_cost(groups[a]) > group_cost(groups[b]); });
        // 1. host preparation, uploads and integrator set-up of every group
        for (; staged < n_groups; staged++) {
            const int i = order[staged];
            cubnm_group& g = groups[i];
            try {
                stage_run(*g_pool[i], g.SC, g.SC_indices, g.SC_dist, g.global_params, g.regional_params, g.v_list);
            } catch (const std::exception& e) {
                throw std::runtime_error("group " + std::to_string(i) + ": " + e.what());
            }
            g_pool[i]->merged_ms = -1;
        }
        for (int i = 0; i < n_groups; i++) check(hipStreamSynchronize(g_pool[i]->stream), "staging");
        // 2. integrators: batched groups on their own streams; all groups on
        //    the per-simulation integrator, whatever their model, merged into
        //    one launch (block shaped for the largest node count; smaller
        //    requests leave the extra waves idle), longest simulations first
        struct Bucket {
            int np = 64, ks = 1, threads = 64;
            size_t lds = 0;
            std::vector<int> members;
            hipEvent_t start = nullptr, stop = nullptr;
            hipStream_t stream = nullptr;
            bool done = false;
            int n_sims = 0, reported_vols = 0;
            long long n_vols = 0;
            double last_report = 0;
        };
        // host-mapped progress counters of the merged launch: BOLD volumes
        // written and simulations completed
        static int* progress_host = nullptr;
        static int* progress_dev = nullptr;
        if (!progress_host) {
            check(hipHostMalloc((void**)&progress_host, sizeof(int) * 8, hipHostMallocMapped | hipHostMallocCoherent),
                  "hipHostMalloc");
            check(hipHostGetDevicePointer((void**)&progress_dev, progress_host, 0), "hipHostGetDevicePointer");
        }
        for (int q = 0; q < 8; q++) progress_host[q] = 0;
        std::map<int, Bucket> buckets;
        std::vector<int> batched;
        for (int k = 0; k < n_groups; k++) {
            const int i = order[k];
            Session& s = *g_pool[i];
            if (s.last_integrator == 1 || s.last_integrator == 2) {
                launch_integrator(s, s.iargs);
                batched.push_back(i);
                continue;
            }
            Bucket& b = buckets[0];
            b.np = std::max(b.np, general_np(s.nodes));
            b.n_vols += (long long)s.bold_len * s.N_SIMS;
            b.members.push_back(i);
        }
        // shape of a bucket's launch for k partitions ks and pair policy
        // (1: the rule's latency-bound condition, 0: off)
        auto shape_bucket = [&](Bucket& b, int ks, int pair_policy) {
            int sims = 0;
            for (int i : b.members) sims += g_pool[i]->N_SIMS;
            b.ks = ks;
            b.threads = b.ks * b.np;
            b.lds = 0;
            for (int i : b.members) {
                Session& s = *g_pool[i];
                // small SCs staged; larger ones read through the caches, in
                // pair mode where the request allows it
                const size_t with_sc = general_lds_bytes(s.model->id, s.nodes, b.np, b.ks, true, false, b.threads);
                s.iargs.sc_lds = with_sc <= 64 * 1024;
                s.iargs.ksplit = b.ks;
                s.iargs.pair = pair_policy && general_pair_ok(s.iargs, b.threads,
                                                              sims <= 512 || s.N_SIMS >= 4 * s.N_SCs);
                b.lds = std::max(b.lds, general_lds_bytes(s.model->id, s.nodes, b.np, b.ks, s.iargs.sc_lds,
                                                          s.iargs.pair != 0, b.threads));
            }
        };
        for (int q = 0; q < 6; q++) g_group_info[q] = 0;
        for (auto& kv : buckets) {
            Bucket& b = kv.second;
            int nmax = 1;
            for (int i : b.members) nmax = std::max(nmax, g_pool[i]->nodes);
            shape_bucket(b, general_ksplit(nmax, g_general_ksplit), 1);
        }
        // host copies of the launch tables stay alive until the next call
        static std::vector<std::vector<char>> host_tables;
        static std::vector<std::unique_ptr<DevBuf[]>> multi_bufs;
        host_tables.clear();
        multi_bufs.clear();
        for (auto& kv : buckets) {
            Bucket& b = kv.second;
            const size_t cstride = model_consts_stride();
            std::vector<char> args, consts, models;
            std::vector<std::pair<double, std::pair<int, int>>> blk;  // (cost, (request, sim))
            for (size_t r = 0; r < b.members.size(); r++) {
                Session& s = *g_pool[b.members[r]];
                const char* ap = reinterpret_cast<const char*>(&s.iargs);
                args.insert(args.end(), ap, ap + sizeof(IntegrateArgs));
                const char* c = static_cast<const char*>(model_consts(s));
                const size_t cb = model_consts_bytes(s.model->id);
                consts.insert(consts.end(), c, c + cb);
                consts.resize(consts.size() + (cstride - cb), 0);
                const int mid = s.model->id;
                const char* mp = reinterpret_cast<const char*>(&mid);
                models.insert(models.end(), mp, mp + sizeof(int));
                const double cost = (double)s.bw_it * s.inner_it * s.nodes * (s.nodes + 40.0);
                for (int sim = 0; sim < s.N_SIMS; sim++) blk.push_back({cost, {(int)r, sim}});
            }
            std::stable_sort(blk.begin(), blk.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
            std::vector<char> blocks(sizeof(int2) * blk.size());
            for (size_t q = 0; q < blk.size(); q++) {
                const int2 v = make_int2(blk[q].second.first, blk[q].second.second);
                std::memcpy(blocks.data() + q * sizeof(int2), &v, sizeof(int2));
            }
            b.stream = g_pool[b.members[0]]->stream;
            // argument table of the bucket's members as shaped now, with every
            // request cut to `trial_bw` BW steps when trial_bw > 0
            auto args_table = [&](int trial_bw) {
                std::vector<char> t;
                for (int i : b.members) {
                    IntegrateArgs a = g_pool[i]->iargs;
                    if (trial_bw > 0) a.bw_it = std::min(a.bw_it, trial_bw);
                    const char* ap = reinterpret_cast<const char*>(&a);
                    t.insert(t.end(), ap, ap + sizeof(IntegrateArgs));
                }
                return t;
            };
            // ---- autotuner
            {
                long long work = 0;
                bool restarts = false;
                std::set<int> mids;
                int nmax = 1, sims = 0;
                bool shared = true;
                for (int i : b.members) {
                    const Session& s = *g_pool[i];
                    work += (long long)s.N_SIMS * s.bw_it * s.inner_it * s.nodes;
                    restarts = restarts || (s.model->id == MODEL_RWW && s.rww.do_fic && s.rww.max_fic_trials > 0);
                    mids.insert(s.model->id);
                    nmax = std::max(nmax, s.nodes);
                    sims += s.N_SIMS;
                    shared = shared && s.N_SIMS >= 4 * s.N_SCs;
                }
                const bool want = !restarts && g_general_ksplit == 0 && g_general_pair == 0 &&
                                  (g_autotune > 0 || (g_autotune == 0 && work >= 20000000000LL));
                if (want) {
                    tune_cache_load();
                    int sims_p2 = 1;
                    while (sims_p2 < sims) sims_p2 <<= 1;
                    std::string key = "general";
                    for (int m : mids) key += "," + std::to_string(m);
                    key += "|N" + std::to_string(nmax) + "|S" + std::to_string(sims_p2) + (shared ? "|shared" : "|own");
                    auto it = g_tune_cache.find(key);
                    TuneChoice best;
                    if (it != g_tune_cache.end() && it->second.ks * b.np <= 1024) {
                        best = it->second;
                        g_group_info[3] = 2;
                        g_group_info[4] = best.trial_ms;
                    } else {
                        // trials: every request cut to 20 BW steps (200 steps at the usual dt)
                        std::unique_ptr<DevBuf[]> tb(new DevBuf[4]);
                        tb[1].ensure(consts.size()); tb[2].ensure(blocks.size()); tb[3].ensure(models.size());
                        check(hipMemcpyAsync(tb[1].p, consts.data(), consts.size(), hipMemcpyHostToDevice, b.stream), "H2D");
                        check(hipMemcpyAsync(tb[2].p, blocks.data(), blocks.size(), hipMemcpyHostToDevice, b.stream), "H2D");
                        check(hipMemcpyAsync(tb[3].p, models.data(), models.size(), hipMemcpyHostToDevice, b.stream), "H2D");
                        hipEvent_t e0, e1;
                        check(hipEventCreate(&e0), "event");
                        check(hipEventCreate(&e1), "event");
                        best.trial_ms = 1e300;
                        int timed = 0;
                        for (int ks : {1, 2, 4}) {
                            if (ks * b.np > 1024) continue;
                            for (int pr : {1, 0}) {
                                shape_bucket(b, ks, pr);
                                std::vector<char> targs = args_table(20);
                                tb[0].ensure(targs.size());
                                check(hipMemcpyAsync(tb[0].p, targs.data(), targs.size(), hipMemcpyHostToDevice, b.stream),
                                      "H2D");
                                check(hipEventRecord(e0, b.stream), "event");
                                check(launch_integrate_general_multi(tb[0].as<IntegrateArgs>(), tb[1].p, tb[3].as<int>(),
                                                                     tb[2].as<int2>(), (int)blk.size(), b.threads, b.lds,
                                                                     nullptr, b.stream),
                                      "autotune trial launch");
                                check(hipEventRecord(e1, b.stream), "event");
                                check(hipEventSynchronize(e1), "autotune trial");
                                const double ms = ev_ms(e0, e1);
                                timed++;
                                if (ms < best.trial_ms) { best.trial_ms = ms; best.ks = ks; best.pair = pr; }
                                bool any_pair = false;
                                for (int i : b.members) any_pair = any_pair || g_pool[i]->iargs.pair;
                                if (!any_pair) break;  // pair mode changes nothing for this launch
                            }
                        }
                        hipEventDestroy(e0); hipEventDestroy(e1);
                        tune_cache_store(key, best);
                        g_group_info[3] = 1;
                        g_group_info[4] = best.trial_ms;
                        g_group_info[5] = timed;
                        if (verbose)
                            std::fprintf(stderr, "cubnm_run_groups: autotuner %s: %d k partitions, pair mode %s "
                                         "(%d shapes timed, best trial %.2f ms)\n", key.c_str(), best.ks,
                                         best.pair ? "allowed" : "off", timed, best.trial_ms);
                    }
                    shape_bucket(b, best.ks, best.pair);
                }
                g_group_info[1] = b.ks;
                g_group_info[2] = 0;
                for (int i : b.members) g_group_info[2] = g_group_info[2] || g_pool[i]->iargs.pair;
            }
            args = args_table(0);
            for (int q = 0; q < 8; q++) progress_host[q] = 0;
            // ---- graph: capture the table uploads and the launch, or replay
            std::vector<char> sig;
            auto put = [&](const void* p, size_t n) {
                const char* c = static_cast<const char*>(p);
                sig.insert(sig.end(), c, c + n);
            };
            put(args.data(), args.size()); put(consts.data(), consts.size());
            put(blocks.data(), blocks.size()); put(models.data(), models.size());
            put(&b.threads, sizeof(b.threads)); put(&b.lds, sizeof(b.lds)); put(&b.stream, sizeof(b.stream));
            GroupGraph& gg = g_group_graph_cache;
            bool replay = g_group_graph && buckets.size() == 1 && gg.exec && gg.sig == sig;
            std::unique_ptr<DevBuf[]> bufs;
            check(hipEventCreate(&b.start), "event");
            check(hipEventCreate(&b.stop), "event");
            b.n_sims = (int)blk.size();
            if (replay) {
                check(hipEventRecord(b.start, b.stream), "event");
                check(hipGraphLaunch(gg.exec, b.stream), "hipGraphLaunch");
                check(hipEventRecord(b.stop, b.stream), "event");
                gg.replays++;
                g_group_info[0] = 2;
            } else {
                if (g_group_graph) gg.release();  // a new capture replaces the cached graph
                bufs.reset(new DevBuf[4]);
                bufs[0].ensure(args.size());
                bufs[1].ensure(consts.size());
                bufs[2].ensure(blocks.size());
                bufs[3].ensure(models.size());
                bool captured = false;
                if (g_group_graph && buckets.size() == 1) {
                    // pinned copies of the tables: the captured copies read them on every replay
                    const std::vector<char>* src[4] = {&args, &consts, &blocks, &models};
                    bool ok = true;
                    for (int q = 0; q < 4 && ok; q++) {
                        ok = hipHostMalloc(&gg.pinned[q], std::max<size_t>(1, src[q]->size()), hipHostMallocDefault) ==
                             hipSuccess;
                        if (ok) std::memcpy(gg.pinned[q], src[q]->data(), src[q]->size());
                    }
                    if (ok && hipStreamBeginCapture(b.stream, hipStreamCaptureModeRelaxed) == hipSuccess) {
                        for (int q = 0; q < 4; q++)
                            (void)hipMemcpyAsync(bufs[q].p, gg.pinned[q], src[q]->size(), hipMemcpyHostToDevice,
                                                 b.stream);
                        const hipError_t le = launch_integrate_general_multi(
                            bufs[0].as<IntegrateArgs>(), bufs[1].p, bufs[3].as<int>(), bufs[2].as<int2>(),
                            (int)blk.size(), b.threads, b.lds, progress_dev, b.stream);
                        hipGraph_t graph = nullptr;
                        const hipError_t ee = hipStreamEndCapture(b.stream, &graph);
                        if (le == hipSuccess && ee == hipSuccess && graph &&
                            hipGraphInstantiate(&gg.exec, graph, nullptr, nullptr, 0) == hipSuccess) {
                            gg.graph = graph;
                            captured = true;
                        } else if (graph) {
                            (void)hipGraphDestroy(graph);
                        }
                        (void)hipGetLastError();
                    }
                }
                check(hipEventRecord(b.start, b.stream), "event");
                if (captured) {
                    check(hipGraphLaunch(gg.exec, b.stream), "hipGraphLaunch");
                    gg.sig = sig;
                    gg.bufs = std::move(bufs);
                    g_group_info[0] = 1;
                } else {
                    if (g_group_graph) gg.release();
                    check(hipMemcpyAsync(bufs[0].p, args.data(), args.size(), hipMemcpyHostToDevice, b.stream), "args H2D");
                    check(hipMemcpyAsync(bufs[1].p, consts.data(), consts.size(), hipMemcpyHostToDevice, b.stream), "consts H2D");
                    check(hipMemcpyAsync(bufs[2].p, blocks.data(), blocks.size(), hipMemcpyHostToDevice, b.stream), "blocks H2D");
                    check(hipMemcpyAsync(bufs[3].p, models.data(), models.size(), hipMemcpyHostToDevice, b.stream), "models H2D");
                    check(launch_integrate_general_multi(bufs[0].as<IntegrateArgs>(), bufs[1].p, bufs[3].as<int>(),
                                                         bufs[2].as<int2>(), (int)blk.size(), b.threads, b.lds,
                                                         progress_dev, b.stream),
                          "integrate_general_multi launch");
                }
                check(hipEventRecord(b.stop, b.stream), "event");
            }
            host_tables.push_back(std::move(args));
            host_tables.push_back(std::move(consts));
            host_tables.push_back(std::move(blocks));
            host_tables.push_back(std::move(models));
            if (bufs) multi_bufs.push_back(std::move(bufs));
            if (verbose)
                std::fprintf(stderr, "cubnm_run_groups: %zu requests, %zu simulations in one per-simulation launch "
                             "(%d threads in %d k partitions, %zu B LDS)\n",
                             b.members.size(), blk.size(), b.threads, b.ks, b.lds);
        }
        // 3. wait for the integrators, reporting each launch as it completes
        for (size_t left = buckets.size() + batched.size(); left > 0;) {
            bool progressed = false;
            for (auto& kv : buckets) {
                Bucket& b = kv.second;
                if (!b.done && verbose) {
                    const int v = __atomic_load_n(progress_host, __ATOMIC_RELAXED);
                    const int c = __atomic_load_n(progress_host + 1, __ATOMIC_RELAXED);
                    if (v != b.reported_vols && (now_s() - b.last_report > 10.0 || c == b.n_sims)) {
                        std::fprintf(stderr, "cubnm_run_groups: per-simulation launch: %d of %d simulations done, "
                                     "%.1f%% of volumes at %.1f s\n", c, b.n_sims,
                                     100.0 * v / std::max(1LL, b.n_vols), now_s() - t0);
                        b.reported_vols = v;
                        b.last_report = now_s();
                    }
                }
                if (b.done || hipEventQuery(b.stop) == hipErrorNotReady) continue;
                b.done = true; left--; progressed = true;
                const float ms = ev_ms(b.start, b.stop);
                for (int i : b.members) g_pool[i]->merged_ms = ms;
                if (verbose)
                    std::fprintf(stderr, "cubnm_run_groups: per-simulation launch of %zu requests done in %.1f s "
                                 "(at %.1f s)\n", b.members.size(), ms / 1000.0, now_s() - t0);
            }
            for (int& i : batched) {
                if (i < 0 || hipEventQuery(g_pool[i]->ev[3]) == hipErrorNotReady) continue;
                if (verbose)
                    std::fprintf(stderr, "cubnm_run_groups: group %d (batched, %d x %d nodes) integrated at %.1f s\n",
                                 i, groups[i].N_SIMS, groups[i].nodes, now_s() - t0);
                i = -1; left--; progressed = true;
            }
            if (!progressed && left > 0) std::this_thread::sleep_for(std::chrono::milliseconds(20));
        }
        for (auto& kv : buckets) { hipEventDestroy(kv.second.start); hipEventDestroy(kv.second.stop); }
        // 4. FC/FCD and downloads, then collect
        for (int k = 0; k < n_groups; k++) {
            const int i = order[k];
            Session& s = *g_pool[i];
            if (s.merged_ms >= 0) {
                check(hipEventRecord(s.ev[2], s.stream), "event");
                check(hipEventRecord(s.ev[3], s.stream), "event");
            }
            queue_outputs(s, &groups[i].out);
        }
        for (int k = 0; k < n_groups; k++) {
            const int i = order[k];
            try {
                finish_run(*g_pool[i], &groups[i].out);
            } catch (const std::exception& e) {
                throw std::runtime_error("group " + std::to_string(i) + ": " + e.what());
            }
        }
        if (verbose) std::fprintf(stderr, "cubnm_run_groups: all %d groups done at %.1f s\n", n_groups, now_s() - t0);
        g_last_groups = n_groups;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        hipDeviceSynchronize();
        return -1;
    }
}

int cubnm_run_device(const double* d_SC, const int* d_SCi, const double* d_SC_dist, double* d_gp, double* d_rp,
                     const double* d_v, cubnm_outputs* o) {
    try {
        require_session();
        Session& s = g_s;
        const double t_start = now_s();
        const ModelInfo* m = s.model;
        const int S = s.N_SIMS, N = s.nodes;
        std::vector<int> delays, maxd;
        s.timings[0] = 0;
        const long long st_stride = s.conf.states_ts ? (long long)s.states_len * N : N;
        if (!o->sim_states) s.d_states.ensure(sizeof(double) * (size_t)m->n_state * S * st_stride);
        if (!o->global_bools) s.d_gb.ensure(std::max<size_t>(1, (size_t)m->n_gb * S));
        if (!o->global_ints) s.d_gi.ensure(sizeof(int) * std::max<size_t>(1, (size_t)m->n_gi * S));
        if (!o->sim_fc_trils) s.d_fc.ensure(sizeof(double) * std::max<size_t>(1, (size_t)S * s.n_pairs));
        if (!o->sim_fcd_trils) s.d_fcd.ensure(sizeof(double) * std::max<size_t>(1, (size_t)S * s.n_window_pairs));
        unsigned char* d_gb = o->global_bools ? o->global_bools : s.d_gb.as<unsigned char>();
        int* d_gi = o->global_ints ? o->global_ints : s.d_gi.as<int>();
        if (m->n_gb > 0) check(hipMemsetAsync(d_gb, 0, (size_t)m->n_gb * S, s.stream), "memset");
        if (m->n_gi > 0) check(hipMemsetAsync(d_gi, 0, sizeof(int) * (size_t)m->n_gi * S, s.stream), "memset");
        auto d2h = [&](void* dst, const void* src, size_t n) {
            check(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s.stream), "D2H");
        };
        if (m->id == MODEL_RWW && s.rww.do_fic) {
            // FIC is host arithmetic: stage the parameters through the host.
            const double tp0 = now_s();
            std::vector<double> SC((size_t)s.N_SCs * N * N), gp((size_t)m->n_global * S), rp((size_t)m->n_regional * S * N);
            std::vector<int> SCi(S);
            d2h(SC.data(), d_SC, sizeof(double) * SC.size());
            d2h(SCi.data(), d_SCi, sizeof(int) * S);
            d2h(gp.data(), d_gp, sizeof(double) * gp.size());
            d2h(rp.data(), d_rp, sizeof(double) * rp.size());
            check(hipStreamSynchronize(s.stream), "sync");
            std::vector<unsigned char> gb((size_t)m->n_gb * S, 0);
            analytical_fic(s, SC.data(), SCi.data(), gp.data(), rp.data(), gb.data());
            check(hipMemcpyAsync(d_rp, rp.data(), sizeof(double) * rp.size(), hipMemcpyHostToDevice, s.stream), "H2D");
            check(hipMemcpyAsync(d_gb, gb.data(), gb.size(), hipMemcpyHostToDevice, s.stream), "H2D");
            check(hipStreamSynchronize(s.stream), "sync");
            s.timings[0] = (now_s() - tp0) * 1e3;
        }
        if (s.do_delay) {
            std::vector<double> dist((size_t)N * N), v(S);
            d2h(dist.data(), d_SC_dist, sizeof(double) * dist.size());
            d2h(v.data(), d_v, sizeof(double) * S);
            check(hipStreamSynchronize(s.stream), "sync");
            build_delays(s, dist.data(), v.data(), delays, maxd);
        }
        DevIO io{d_SC, d_SCi, d_gp, d_rp, o->sim_bold,
                 o->sim_fc_trils ? o->sim_fc_trils : s.d_fc.as<double>(),
                 o->sim_fcd_trils ? o->sim_fcd_trils : s.d_fcd.as<double>(),
                 o->sim_states ? o->sim_states : s.d_states.as<double>(),
                 d_gb, d_gi};
        check(hipEventRecord(s.ev[0], s.stream), "event");
        check(hipEventRecord(s.ev[1], s.stream), "event");
        std::vector<int> sci_h(S);
        d2h(sci_h.data(), d_SCi, sizeof(int) * S);
        check(hipStreamSynchronize(s.stream), "sync");
        for (int i = 0; i < S; i++)
            if (sci_h[i] < 0 || sci_h[i] >= s.N_SCs) throw std::runtime_error("SC_indices out of range");
        run_kernels(s, io, s.do_delay ? &delays : nullptr, s.do_delay ? &maxd : nullptr, sci_h.data());
        check(hipEventRecord(s.ev[5], s.stream), "event");
        check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
        s.timings[1] = 0;
        s.timings[2] = ev_ms(s.ev[2], s.ev[3]);
        s.timings[3] = ev_ms(s.ev[3], s.ev[4]);
        s.timings[4] = 0;
        o->run_time = now_s() - t_start;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

}  // extern "C"

extern "C" int cubnm_run_groups_info(double* out6) {
    for (int i = 0; i < 6; i++) out6[i] = g_group_info[i];
    return 0;
}
