// General FP64 integrator: one workgroup per simulation, one lane per node.
//
// This kernel covers every feature of the reference simulation loop
// (src/ext/bnm.cu:117-505, semantics of src/ext/bnm.cpp:69-522): conduction
// delays with a circular history, oscillator coupling, rWW numerical-FIC
// restarts, state time series, any node count up to 1024.  The batched
// MFMA kernel (integrate_batched.hip) is the throughput path for the common
// case (no delay, shared SC); this one is the feature-complete path.
//
// Per inner step: barrier, coupling g_j = sum_k SC[k][j] * s_k (Jacobi: all
// lanes read the previous step's state from LDS), model step, barrier,
// publish s_j.  BW runs once per bw step, BOLD every TR.
#include "engine.hpp"

namespace cubnm {

namespace {

template <class M>
struct FicState {  // rWW numerical FIC per node and per simulation
    double mean_I_E, delta;
    int trial;
    bool adjust, failed;
};

template <class M>
__device__ __forceinline__ void model_restart(const typename M::C& mc, double* s, const double* rp) {
    M::init(mc, s, rp);
}

}  // namespace

template <class M>
__device__ __forceinline__ void general_body(const IntegrateArgs& a, const typename M::C& mc, const int sim) {
    const int j = threadIdx.x;
    const int N = a.nodes;
    const bool act = j < N;
    const int jj = act ? j : 0;  // inactive lanes mirror node 0 but never publish
    // waves with no node (a launch shaped for larger requests) skip the work
    // but keep meeting the barriers
    const bool wact = (int)(threadIdx.x & ~63u) < N;
    const int kN = wact ? N : 0;
    extern __shared__ double conn[];  // [N] previous-step coupling state (no delay), then SC when staged
    const double* SCg = a.SC + (long long)a.SC_indices[sim] * N * N;
    // SC in LDS (rows padded by one double) when it fits: the coupling loop
    // then reads LDS instead of L1/L2 every step; summation order unchanged
    double* scs = conn + N;
    const int ldsc = a.sc_lds ? N + 1 : N;
    if (a.sc_lds)
        for (int i = threadIdx.x; i < N * N; i += blockDim.x) scs[(i / N) * ldsc + i % N] = SCg[i];
    const double* SC = a.sc_lds ? scs : SCg;
    double gp[M::n_global], rp[M::n_regional];
    for (int i = 0; i < M::n_global; i++) gp[i] = a.gp[(long long)i * a.N_SIMS + sim];
    for (int i = 0; i < M::n_regional; i++) rp[i] = a.rp[(long long)i * a.N_SIMS * N + (long long)sim * N + jj];

    double s[M::n_state];
    for (int i = 0; i < M::n_state; i++) s[i] = 0.0;
    M::init(mc, s, rp);

    // rWW numerical FIC (rww.cpp:149-170, 226-291)
    double fic_mean = 0.0, fic_delta = 0.0;
    int fic_trial = 0;
    bool fic_adjust = false, fic_failed = false;
    if constexpr (M::n_gout_int > 0) {
        fic_delta = mc.init_delta;
        fic_adjust = mc.do_fic && (mc.max_fic_trials > 0);
    }

    double bx = 0.0, bf = 1.0, bn = 1.0, bq = 1.0;

    const int max_delay = a.do_delay ? a.max_delays[sim] : 0;
    const bool has_delay = max_delay > 0;
    double* hist = has_delay ? a.hist + (long long)sim * a.hist_stride : nullptr;
    const int* delay = has_delay ? a.delays + (long long)sim * N * N : nullptr;
    int buff_idx = max_delay - 1;
    auto publish_all = [&]() {
        if (!act) return;
        if (has_delay) for (int bi = 0; bi < max_delay; bi++) hist[(long long)bi * N + j] = s[M::conn_idx];
        else conn[j] = s[M::conn_idx];
    };
    publish_all();

    const long long st_stride = a.states_ts ? (long long)a.states_len * N : N;
    double st_sum[M::n_state];
    for (int i = 0; i < M::n_state; i++) st_sum[i] = 0.0;

    double* bold = a.bold + (long long)sim * a.bold_len * N;
    int bw_i = 0, bold_i = 0, states_i = 0, rep = 0;
    while (bw_i < a.bw_it) {
        const int sh_t = a.sh_ts[(bw_i % a.noise_bw_it) + rep * a.noise_bw_it];
        const int sh_j = a.sh_nodes[rep * N + jj];
        for (int inner_i = 0; inner_i < a.inner_it; inner_i++) {
            __syncthreads();
            double g = 0.0;
            if (has_delay) {
                const double sj = hist[(long long)buff_idx * N + jj];
                for (int k = 0; k < kN; k++) {
                    const int kb = (buff_idx + delay[(long long)k * N + jj]) % max_delay;
                    const double hk = hist[(long long)kb * N + k];
                    if constexpr (M::is_osc) g += SC[(long long)k * ldsc + jj] * sin(hk - sj);
                    else g += SC[(long long)k * ldsc + jj] * hk;
                }
            } else {
                const double sj = conn[jj];
#pragma unroll 8
                for (int k = 0; k < kN; k++) {
                    if constexpr (M::is_osc) g += SC[(long long)k * ldsc + jj] * sin(conn[k] - sj);
                    else g += SC[(long long)k * ldsc + jj] * conn[k];
                }
            }
            const long long nidx = ((long long)(sh_t * a.inner_it + inner_i) * N + sh_j) * M::n_noise;
            const double n0 = a.noise[nidx];
            const double n1 = (M::n_noise > 1) ? a.noise[nidx + 1] : 0.0;
            if (wact) M::step(mc, s, gp, rp, g, n0, n1);
            __syncthreads();
            if (has_delay) {
                if (act) hist[(long long)buff_idx * N + j] = s[M::conn_idx];
                buff_idx = (buff_idx + max_delay - 1) % max_delay;
            } else if (act) {
                conn[j] = s[M::conn_idx];
            }
        }
        if (wact) bw_step(a.bw, s[M::bold_idx], bx, bf, bn, bq);
        if ((bw_i + 1) % a.BOLD_TR_iters == 0) {
            if (act && bold_i < a.bold_len) bold[(long long)bold_i * N + j] = bold_signal(a.bw, bn, bq);
            bold_i++;
        }
        if ((bw_i + 1) % a.states_sampling_iters == 0) {
            if (a.states_ts) {
                if (act && states_i < a.states_len)
                    for (int i = 0; i < M::n_state; i++)
                        a.states[(long long)i * a.N_SIMS * st_stride + (long long)sim * st_stride + (long long)states_i * N + j] = s[i];
            } else if (states_i >= a.n_states_samples_remove) {
                for (int i = 0; i < M::n_state; i++) st_sum[i] += s[i];
            }
            states_i++;
        }
        if ((bw_i + 1) % a.noise_bw_it == 0) {
            if (bw_i + 1 < a.bw_it) rep++;
        }
        bool restart = false;
        if constexpr (M::n_gout_int > 0) {
            // block-uniform: fic_adjust and bw_i are identical in every lane
            if (fic_adjust) {
                const int b1 = bw_i + 1;
                if (b1 >= mc.I_SAMPLING_START && b1 <= mc.I_SAMPLING_END) fic_mean += s[0];
                if (b1 == mc.I_SAMPLING_END) {
                    fic_mean /= mc.I_SAMPLING_DURATION;
                    const double diff = fic_mean - mc.b_a_ratio_E;
                    bool vote = false;
                    if (fabs(diff + 0.026) > 0.005) {
                        vote = true;
                        if (fic_trial < mc.max_fic_trials) {
                            if (diff < -0.026) {
                                rp[2] -= fic_delta;
                                fic_delta -= 0.001;
                                fic_delta = fmax(fic_delta, 0.001);
                            } else {
                                rp[2] += fic_delta;
                            }
                        }
                    }
                    restart = __syncthreads_or(act && vote);
                    if (restart) {
                        if (fic_trial < mc.max_fic_trials) fic_trial++;
                        else { restart = false; fic_adjust = false; fic_failed = true; }
                    } else {
                        fic_adjust = false;
                    }
                }
            }
        }
        bw_i++;
        if (restart) {
            __syncthreads();
            model_restart<M>(mc, s, rp);
            if constexpr (M::n_gout_int > 0) fic_mean = 0.0;
            bx = 0.0; bf = 1.0; bn = 1.0; bq = 1.0;
            buff_idx = max_delay - 1;
            publish_all();
            bold_i = 0; states_i = 0; bw_i = 0; rep = 0;
            __syncthreads();
        }
    }
    if (!act) return;
    if constexpr (M::n_gout_int > 0) {
        if (mc.do_fic) {
            a.rp[2LL * a.N_SIMS * N + (long long)sim * N + j] = rp[2];
            if (j == 0) {
                a.gint[sim] = fic_trial;
                a.gbool[1LL * a.N_SIMS + sim] = fic_failed;
            }
        }
    }
    if (!a.states_ts) {
        const int tp = states_i - a.n_states_samples_remove;
        for (int i = 0; i < M::n_state; i++)
            a.states[(long long)i * a.N_SIMS * st_stride + (long long)sim * st_stride + j] = st_sum[i] / tp;
    }
}

template <class M>
__global__ void __launch_bounds__(1024) integrate_general_kernel(IntegrateArgs a, typename M::C mc) {
    general_body<M>(a, mc, blockIdx.x);
}

// Several requests of one model in one launch: block b integrates simulation
// blocks[b].y of request blocks[b].x, with that request's arguments and
// model constants (durations, SCs, outputs and configs may all differ; the
// launch shares the block size and the LDS layout).
template <class M>
__global__ void __launch_bounds__(1024) integrate_general_multi_kernel(const IntegrateArgs* args,
                                                                       const typename M::C* consts,
                                                                       const int2* blocks, int* progress) {
    const int2 b = blocks[blockIdx.x];
    general_body<M>(args[b.x], consts[b.x], b.y);
    // completed simulations, readable by the host while the launch runs
    if (threadIdx.x == 0 && progress)
        __hip_atomic_fetch_add(progress, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

size_t general_lds(const IntegrateArgs& a, int* sc_lds);

template <class M>
static hipError_t launch_general(const IntegrateArgs& a, const void* consts, hipStream_t st) {
    const typename M::C mc = *static_cast<const typename M::C*>(consts);
    int threads = ((a.nodes + 63) / 64) * 64;
    IntegrateArgs b = a;
    const size_t lds = general_lds(a, &b.sc_lds);
    hipLaunchKernelGGL(integrate_general_kernel<M>, dim3(a.N_SIMS), dim3(threads), lds, st, b, mc);
    return hipGetLastError();
}

// LDS bytes and SC staging of a request on the general integrator.
size_t general_lds(const IntegrateArgs& a, int* sc_lds) {
    const size_t lds = sizeof(double) * (size_t)a.nodes;
    const size_t with_sc = lds + sizeof(double) * (size_t)a.nodes * (a.nodes + 1);
    // staged when it leaves room for two workgroups per CU, or when there are
    // too few simulations to fill the CUs anyway (latency-bound)
    *sc_lds = with_sc <= 64 * 1024 || (with_sc <= 150 * 1024 && a.N_SIMS <= 512);
    return *sc_lds ? with_sc : lds;
}

template <class M>
static hipError_t launch_general_multi(const IntegrateArgs* d_args, const void* d_consts, const int2* d_blocks,
                                       int n_blocks, int threads, size_t lds, int* progress, hipStream_t st) {
    hipLaunchKernelGGL(integrate_general_multi_kernel<M>, dim3(n_blocks), dim3(threads), lds, st, d_args,
                       static_cast<const typename M::C*>(d_consts), d_blocks, progress);
    return hipGetLastError();
}

size_t model_consts_bytes(int model) {
    switch (model) {
        case MODEL_RWW: return sizeof(RWW::C);
        case MODEL_JR: return sizeof(JR::C);
        case MODEL_KURAMOTO: return sizeof(Kuramoto::C);
        default: return sizeof(WC::C);
    }
}

// d_args: [n_req] arguments (sc_lds set consistently with lds); d_consts:
// [n_req] model constants of `model`; d_blocks: [n_blocks] (request, sim);
// progress: null, or a host-mapped counter of completed simulations.
hipError_t launch_integrate_general_multi(int model, const IntegrateArgs* d_args, const void* d_consts,
                                          const int2* d_blocks, int n_blocks, int threads, size_t lds,
                                          int* progress, hipStream_t s) {
    if (threads < 64 || threads > 1024 || n_blocks < 1) return hipErrorInvalidValue;
    switch (model) {
        case MODEL_RWW: return launch_general_multi<RWW>(d_args, d_consts, d_blocks, n_blocks, threads, lds, progress, s);
        case MODEL_JR: return launch_general_multi<JR>(d_args, d_consts, d_blocks, n_blocks, threads, lds, progress, s);
        case MODEL_KURAMOTO: return launch_general_multi<Kuramoto>(d_args, d_consts, d_blocks, n_blocks, threads, lds, progress, s);
        case MODEL_WC: return launch_general_multi<WC>(d_args, d_consts, d_blocks, n_blocks, threads, lds, progress, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_integrate_general(int model, const IntegrateArgs& a, const void* consts, hipStream_t s) {
    if (a.nodes < 1 || a.nodes > 1024) return hipErrorInvalidValue;
    switch (model) {
        case MODEL_RWW: return launch_general<RWW>(a, consts, s);
        case MODEL_JR: return launch_general<JR>(a, consts, s);
        case MODEL_KURAMOTO: return launch_general<Kuramoto>(a, consts, s);
        case MODEL_WC: return launch_general<WC>(a, consts, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cubnm
