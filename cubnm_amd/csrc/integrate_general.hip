// General FP64 integrator: one workgroup per simulation, one lane per node.
//
// This kernel covers every feature of the reference simulation loop
// (src/ext/bnm.cu:117-505, semantics of src/ext/bnm.cpp:69-522): conduction
// delays with a circular history, oscillator coupling, rWW numerical-FIC
// restarts, state time series, any node count up to 1024 (and, in
// integrate_wide_kernel below, up to 8192 nodes).  The batched
// MFMA kernel (integrate_batched.hip) is the throughput path for the common
// case (no delay, shared SC); this one is the feature-complete path.
//
// Per inner step: barrier, coupling g_j = sum_k SC[k][j] * s_k (Jacobi: all
// lanes read the previous step's state from LDS), model step, barrier,
// publish s_j.  BW runs once per bw step, BOLD every TR.
//
// The coupling sum is the hot loop: with one SC per simulation it streams
// nodes^2 doubles per step from L2 / MALL / HBM.  It is split over `ksplit`
// partitions of source nodes (more waves, more loads in flight; partial sums
// added in partition order), its loads are typed as global (generic loads
// would wait for each other) and issued eight at a time.  Oscillators
// without delay use sin(x_k - x_j) = sin x_k cos x_j - cos x_k sin x_j: two
// sums over published sin/cos instead of nodes^2 sine evaluations per step.
#include "engine.hpp"
#include <cstdint>

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

// Global-memory view of a pointer whose address space the compiler cannot
// infer (read from a table in memory, or chosen at run time between LDS and
// global): generic (flat) loads are counted against both the vector-memory
// and the LDS counters, so every load would wait for the previous one.
template <class T>
using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem<T>* as_global(T* p) { return (gmem<T>*)p; }

// sum_{k0 <= k < k1} SC[k][j] * x_k (and the same with the second half of x
// for oscillators: x = sin then cos), in ascending k.  SC is an LDS or a
// global pointer (typed, so loads are issued back to back); x is in LDS.
template <bool OSC, class P>
__device__ __forceinline__ void coupling_sum(P sc, const int ldsc, const double* x, const int N, const int k0,
                                             const int k1, double& g, double& g2) {
    // batches of U loads issued back to back, then consumed in order (a
    // software-pipelined form, loads of batch b+1 before batch b is consumed,
    // spills at this kernel's 128-VGPR budget and is scheduled with a full
    // wait at the loop head)
    constexpr int U = 8;
    int k = k0;
    for (; k + U <= k1; k += U) {
        double w[U];
#pragma unroll
        for (int u = 0; u < U; u++) w[u] = sc[(long long)u * ldsc];
        sc += (long long)U * ldsc;
#pragma unroll
        for (int u = 0; u < U; u++) {
            g += w[u] * x[k + u];
            if constexpr (OSC) g2 += w[u] * x[N + k + u];
        }
    }
    for (; k < k1; k++, sc += ldsc) {
        const double w = *sc;
        g += w * x[k];
        if constexpr (OSC) g2 += w * x[N + k];
    }
}

typedef double dbl2 __attribute__((ext_vector_type(2)));

// Pair mode: this lane's two adjacent columns (c, c+1) of
// sum_{k0 <= k < k1} SC[k][c] * x_k, with 16-byte loads (ld2: row stride in
// column pairs), in ascending k.  Halves the load instructions of the
// single-column form, which bound it.
template <bool OSC>
__device__ __forceinline__ void pair_sum(const gmem<const double>* sc, const int ld2, const double* x, const int N,
                                         const int k0, const int k1, double& ga, double& gb, double& ga2,
                                         double& gb2) {
    const gmem<const dbl2>* p = (const gmem<const dbl2>*)sc;
    constexpr int U = 4;
    int k = k0;
    for (; k + U <= k1; k += U) {
        dbl2 w[U];
#pragma unroll
        for (int u = 0; u < U; u++) w[u] = p[(long long)u * ld2];
        p += (long long)U * ld2;
#pragma unroll
        for (int u = 0; u < U; u++) {
            ga += w[u].x * x[k + u];
            gb += w[u].y * x[k + u];
            if constexpr (OSC) {
                ga2 += w[u].x * x[N + k + u];
                gb2 += w[u].y * x[N + k + u];
            }
        }
    }
    for (; k < k1; k++, p += ld2) {
        const dbl2 w = *p;
        ga += w.x * x[k];
        gb += w.y * x[k];
        if constexpr (OSC) {
            ga2 += w.x * x[N + k];
            gb2 += w.y * x[N + k];
        }
    }
}

}  // namespace

template <class M>
__device__ __forceinline__ void general_body(const IntegrateArgs& a, const typename M::C& mc, const int sim,
                                             int* progress) {
    // the block is ksplit partitions of np lanes (np a multiple of 64, so no
    // wave straddles two partitions); every partition sums the coupling over
    // its own range of source nodes k, partition 0 owns the node state and
    // adds the other partitions' partial sums in partition order
    const int ks = a.ksplit;
    const int np = blockDim.x / ks;
    const int part = threadIdx.x / np;
    const int j = threadIdx.x - part * np;
    const int N = a.nodes;
    const bool own = part == 0;
    const bool act = own && j < N;
    const int jj = j < N ? j : 0;  // inactive lanes mirror node 0 but never publish
    // waves with no node (a launch shaped for larger requests) skip the work
    // but keep meeting the barriers
    const bool wnode = (j & ~63) < N;
    const bool wact = own && wnode;
    const int kc = (N + ks - 1) / ks;
    const int k0 = wnode ? min(N, part * kc) : 0;
    const int k1 = wnode ? min(N, k0 + kc) : 0;
    // pair mode (SC in global memory, no delay, even N): the coupling is
    // summed by P partitions of npc lanes, lane t of partition p taking
    // columns 2t and 2t+1 over partition p's source rows; node lanes then add
    // the P partial sums from LDS in partition order
    const bool pair = a.pair != 0;
    const int npc = general_npc(N);
    const int P = pair ? (int)blockDim.x / npc : 1;
    const int pp = threadIdx.x / npc;
    const int pc = 2 * (threadIdx.x - pp * npc);
    const bool pact = pair && pp < P && pc < N;
    const int pkc = (N + P - 1) / P;
    const int pk0 = pact ? min(N, pp * pkc) : 0;
    const int pk1 = pact ? min(N, pk0 + pkc) : 0;
    extern __shared__ double lds[];
    // partial sums: k partitions 1..ks-1 ([ks-1][2][np]) or pair partitions ([2][P][N])
    double* red = lds;
    double* conn = lds + general_red_doubles(M::is_osc, N, np, ks, pair, blockDim.x);
    const int nc = M::is_osc ? 2 * N : N;
    const gmem<const double>* SCg = as_global(a.SC) + (long long)a.SC_indices[sim] * N * N;
    // SC in LDS (rows padded by one double) when it fits: the coupling loop
    // then reads LDS instead of L1/L2 every step; summation order unchanged
    double* scs = conn + nc;
    const int ldsc = a.sc_lds ? N + 1 : N;
    if (a.sc_lds)
        for (int i = threadIdx.x; i < N * N; i += blockDim.x) scs[(i / N) * ldsc + i % N] = SCg[i];
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
    gmem<double>* hist = has_delay ? as_global(a.hist) + (long long)sim * a.hist_stride : nullptr;
    const gmem<const int>* delay = has_delay ? as_global(a.delays) + (long long)sim * N * N : nullptr;
    int buff_idx = max_delay - 1;
    // oscillators without delay publish sin and cos of the phase:
    // sum_k SC[k][j] sin(x_k - x_j) = cos x_j (SC^T sin x)_j - sin x_j (SC^T cos x)_j
    auto publish = [&](double v) {
        if constexpr (M::is_osc) sincos(v, &conn[j], &conn[N + j]);
        else conn[j] = v;
    };
    auto publish_all = [&]() {
        if (!act) return;
        if (has_delay) for (int bi = 0; bi < max_delay; bi++) hist[(long long)bi * N + j] = s[M::conn_idx];
        else publish(s[M::conn_idx]);
    };
    publish_all();

    const long long st_stride = a.states_ts ? (long long)a.states_len * N : N;
    double st_sum[M::n_state];
    for (int i = 0; i < M::n_state; i++) st_sum[i] = 0.0;

    gmem<double>* bold = as_global(a.bold) + (long long)sim * a.bold_len * N;
    const gmem<const double>* noise = as_global(a.noise);
    int bw_i = 0, bold_i = 0, states_i = 0, rep = 0;
    while (bw_i < a.bw_it) {
        const int sh_t = as_global(a.sh_ts)[(bw_i % a.noise_bw_it) + rep * a.noise_bw_it];
        const int sh_j = as_global(a.sh_nodes)[rep * N + jj];
        for (int inner_i = 0; inner_i < a.inner_it; inner_i++) {
            // this step's noise, issued before the barrier so that its latency
            // overlaps the coupling sum; without delay the barriers order LDS
            // only (with delay, lanes exchange state through the history in
            // global memory and need the full barrier)
            const long long nidx = ((long long)(sh_t * a.inner_it + inner_i) * N + sh_j) * M::n_noise;
            const double n0 = wact ? noise[nidx] : 0.0;
            const double n1 = (M::n_noise > 1 && wact) ? noise[nidx + 1] : 0.0;
            if (has_delay) __syncthreads();
            else lds_barrier();
            double g = 0.0, g2 = 0.0;
            if (has_delay) {
                const double sj = hist[(long long)buff_idx * N + jj];
                for (int k = k0; k < k1; k++) {
                    const int kb = (buff_idx + delay[(long long)k * N + jj]) % max_delay;
                    const double hk = hist[(long long)kb * N + k];
                    const double w = a.sc_lds ? scs[(long long)k * ldsc + jj] : SCg[(long long)k * ldsc + jj];
                    if constexpr (M::is_osc) g += w * sin(hk - sj);
                    else g += w * hk;
                }
            } else if (a.sc_lds) {
                coupling_sum<M::is_osc>(scs + (long long)k0 * ldsc + jj, ldsc, conn, N, k0, k1, g, g2);
            } else if (pair) {
                double ga = 0.0, gb = 0.0, ga2 = 0.0, gb2 = 0.0;
                pair_sum<M::is_osc>(SCg + (long long)pk0 * N + (pact ? pc : 0), N / 2, conn, N, pk0, pk1, ga, gb,
                                    ga2, gb2);
                if (pact) {
                    red[(long long)pp * N + pc] = ga;
                    red[(long long)pp * N + pc + 1] = gb;
                    if constexpr (M::is_osc) {
                        red[(long long)(P + pp) * N + pc] = ga2;
                        red[(long long)(P + pp) * N + pc + 1] = gb2;
                    }
                }
                lds_barrier();
                if (act) {
                    g = red[jj];
                    for (int p = 1; p < P; p++) g += red[(long long)p * N + jj];
                    if constexpr (M::is_osc) {
                        g2 = red[(long long)P * N + jj];
                        for (int p = 1; p < P; p++) g2 += red[(long long)(P + p) * N + jj];
                    }
                }
            } else {
                coupling_sum<M::is_osc>(SCg + (long long)k0 * ldsc + jj, ldsc, conn, N, k0, k1, g, g2);
            }
            if (ks > 1 && !pair) {
                if (part > 0) {
                    red[(long long)(part - 1) * 2 * np + j] = g;
                    red[(long long)((part - 1) * 2 + 1) * np + j] = g2;
                }
                lds_barrier();
                if (own)
                    for (int p = 1; p < ks; p++) {
                        g += red[(long long)(p - 1) * 2 * np + j];
                        g2 += red[(long long)((p - 1) * 2 + 1) * np + j];
                    }
            }
            if constexpr (M::is_osc) {
                if (!has_delay) g = conn[N + jj] * g - conn[jj] * g2;
            }
            if (wact) M::step(mc, s, gp, rp, g, n0, n1);
            if (has_delay) __syncthreads();
            else lds_barrier();
            if (has_delay) {
                if (act) hist[(long long)buff_idx * N + j] = s[M::conn_idx];
                buff_idx = (buff_idx + max_delay - 1) % max_delay;
            } else if (act) {
                publish(s[M::conn_idx]);
            }
        }
        if (wact) bw_step(a.bw, s[M::bold_idx], bx, bf, bn, bq);
        if ((bw_i + 1) % a.BOLD_TR_iters == 0) {
            if (act && bold_i < a.bold_len) bold[(long long)bold_i * N + j] = bold_signal(a.bw, bn, bq);
            bold_i++;
            // volumes written, readable by the host while the launch runs (bnm.cu:362-370)
            if (progress && threadIdx.x == 0)
                __hip_atomic_fetch_add(progress, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
            if (progress && threadIdx.x == 0)  // the restarted volumes are written again (bnm.cu:443-446)
                __hip_atomic_fetch_add(progress, -bold_i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
    general_body<M>(a, mc, blockIdx.x, nullptr);
}

// More than 1024 nodes: one workgroup of 1024 lanes per simulation, lane t
// owning nodes t, t + 1024, ..., t + (NPL - 1) * 1024; the coupling state
// (N doubles, sin and cos for oscillators) in LDS and SC read from memory.
// This is the reference's many-node regime (bnm.cu:702-743 splits such a
// simulation over co-launched blocks with a grid barrier per step); here one
// CU integrates one simulation and simulations run side by side, so no
// grid-wide barrier is needed.  Without conduction delay and numerical-FIC
// restarts (the host rejects both for N > 1024); per node the arithmetic and
// summation order are those of integrate_general_kernel with one partition.
template <class M, int NPL>
__global__ void __launch_bounds__(1024) integrate_wide_kernel(IntegrateArgs a, typename M::C mc) {
    const int sim = blockIdx.x;
    const int N = a.nodes, L = blockDim.x;
    extern __shared__ double conn[];  // [N], oscillators [2N]
    const gmem<const double>* SCg = as_global(a.SC) + (long long)a.SC_indices[sim] * N * N;
    double gp[M::n_global];
    for (int i = 0; i < M::n_global; i++) gp[i] = a.gp[(long long)i * a.N_SIMS + sim];
    double rp[NPL][M::n_regional], s[NPL][M::n_state], st_sum[NPL][M::n_state];
    double bx[NPL], bf[NPL], bn[NPL], bq[NPL];
    int jj[NPL];
    bool act[NPL];
#pragma unroll
    for (int v = 0; v < NPL; v++) {
        const int j = threadIdx.x + v * L;
        act[v] = j < N;
        jj[v] = act[v] ? j : 0;
        for (int i = 0; i < M::n_regional; i++)
            rp[v][i] = a.rp[(long long)i * a.N_SIMS * N + (long long)sim * N + jj[v]];
        for (int i = 0; i < M::n_state; i++) { s[v][i] = 0.0; st_sum[v][i] = 0.0; }
        M::init(mc, s[v], rp[v]);
        bx[v] = 0.0; bf[v] = 1.0; bn[v] = 1.0; bq[v] = 1.0;
    }
    auto publish = [&](int j, double val) {
        if constexpr (M::is_osc) sincos(val, &conn[j], &conn[N + j]);
        else conn[j] = val;
    };
#pragma unroll
    for (int v = 0; v < NPL; v++)
        if (act[v]) publish(jj[v], s[v][M::conn_idx]);
    const long long st_stride = a.states_ts ? (long long)a.states_len * N : N;
    gmem<double>* bold = as_global(a.bold) + (long long)sim * a.bold_len * N;
    const gmem<const double>* noise = as_global(a.noise);
    int bw_i = 0, bold_i = 0, states_i = 0, rep = 0;
    while (bw_i < a.bw_it) {
        const int sh_t = as_global(a.sh_ts)[(bw_i % a.noise_bw_it) + rep * a.noise_bw_it];
        int sh_j[NPL];
#pragma unroll
        for (int v = 0; v < NPL; v++) sh_j[v] = as_global(a.sh_nodes)[rep * N + jj[v]];
        for (int inner_i = 0; inner_i < a.inner_it; inner_i++) {
            double n0[NPL], n1[NPL];
#pragma unroll
            for (int v = 0; v < NPL; v++) {
                const long long nidx = ((long long)(sh_t * a.inner_it + inner_i) * N + sh_j[v]) * M::n_noise;
                n0[v] = act[v] ? noise[nidx] : 0.0;
                n1[v] = (M::n_noise > 1 && act[v]) ? noise[nidx + 1] : 0.0;
            }
            lds_barrier();
#pragma unroll
            for (int v = 0; v < NPL; v++) {
                if (!act[v]) continue;
                double g = 0.0, g2 = 0.0;
                coupling_sum<M::is_osc>(SCg + jj[v], N, conn, N, 0, N, g, g2);
                if constexpr (M::is_osc) g = conn[N + jj[v]] * g - conn[jj[v]] * g2;
                M::step(mc, s[v], gp, rp[v], g, n0[v], n1[v]);
            }
            lds_barrier();
#pragma unroll
            for (int v = 0; v < NPL; v++)
                if (act[v]) publish(jj[v], s[v][M::conn_idx]);
        }
#pragma unroll
        for (int v = 0; v < NPL; v++)
            if (act[v]) bw_step(a.bw, s[v][M::bold_idx], bx[v], bf[v], bn[v], bq[v]);
        if ((bw_i + 1) % a.BOLD_TR_iters == 0) {
            if (bold_i < a.bold_len)
#pragma unroll
                for (int v = 0; v < NPL; v++)
                    if (act[v]) bold[(long long)bold_i * N + jj[v]] = bold_signal(a.bw, bn[v], bq[v]);
            bold_i++;
        }
        if ((bw_i + 1) % a.states_sampling_iters == 0) {
#pragma unroll
            for (int v = 0; v < NPL; v++) {
                if (!act[v]) continue;
                if (a.states_ts) {
                    if (states_i < a.states_len)
                        for (int i = 0; i < M::n_state; i++)
                            a.states[(long long)i * a.N_SIMS * st_stride + (long long)sim * st_stride +
                                     (long long)states_i * N + jj[v]] = s[v][i];
                } else if (states_i >= a.n_states_samples_remove) {
                    for (int i = 0; i < M::n_state; i++) st_sum[v][i] += s[v][i];
                }
            }
            states_i++;
        }
        if ((bw_i + 1) % a.noise_bw_it == 0 && bw_i + 1 < a.bw_it) rep++;
        bw_i++;
    }
#pragma unroll
    for (int v = 0; v < NPL; v++) {
        if (!act[v]) continue;
        const int j = jj[v];
        if constexpr (M::n_gout_int > 0) {
            // analytical FIC only (restarts are rejected for this kernel):
            // wIE is unchanged, no trials, no failure
            if (mc.do_fic && j == 0) { a.gint[sim] = 0; a.gbool[1LL * a.N_SIMS + sim] = false; }
        }
        if (!a.states_ts) {
            const int tp = states_i - a.n_states_samples_remove;
            for (int i = 0; i < M::n_state; i++)
                a.states[(long long)i * a.N_SIMS * st_stride + (long long)sim * st_stride + j] = st_sum[v][i] / tp;
        }
    }
}

// Every per-simulation request of a heterogeneous batch in one launch: block b
// integrates simulation blocks[b].y of request blocks[b].x, with that
// request's model (models[r]), arguments and model constants (consts + r *
// cstride).  Durations, models, node counts, SCs, outputs and configs may all
// differ; the launch shares the block size, the k partitions and the LDS
// size.  progress: null, or two host-mapped counters: [0] BOLD volumes
// written, [1] simulations completed.
__global__ void __launch_bounds__(1024) integrate_general_multi_kernel(const IntegrateArgs* args, const char* consts,
                                                                       int cstride, const int* models,
                                                                       const int2* blocks, int* progress) {
    const int2 b = blocks[blockIdx.x];
    const char* c = consts + (long long)b.x * cstride;
    switch (models[b.x]) {
        case MODEL_RWW: general_body<RWW>(args[b.x], *reinterpret_cast<const RWW::C*>(c), b.y, progress); break;
        case MODEL_JR: general_body<JR>(args[b.x], *reinterpret_cast<const JR::C*>(c), b.y, progress); break;
        case MODEL_KURAMOTO:
            general_body<Kuramoto>(args[b.x], *reinterpret_cast<const Kuramoto::C*>(c), b.y, progress);
            break;
        case MODEL_RWWEX: general_body<RWWEx>(args[b.x], *reinterpret_cast<const RWWEx::C*>(c), b.y, progress); break;
        default: general_body<WC>(args[b.x], *reinterpret_cast<const WC::C*>(c), b.y, progress); break;
    }
    if (threadIdx.x == 0 && progress)
        __hip_atomic_fetch_add(progress + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// LDS bytes of one general-kernel workgroup: partial sums of k partitions
// 1..ksplit-1 (two per lane), the previous-step coupling state (sin and cos
// for oscillators) and, when staged, SC with rows padded by one double.
size_t general_lds_bytes(int model, int nodes, int np, int ksplit, bool sc_staged, bool pair, int threads) {
    size_t d = (size_t)general_red_doubles(model == MODEL_KURAMOTO, nodes, np, ksplit, pair, threads) +
               (size_t)(model == MODEL_KURAMOTO ? 2 : 1) * nodes;
    if (sc_staged) d += (size_t)nodes * (nodes + 1);
    return d * sizeof(double);
}

// Lanes per k partition (nodes rounded up to whole waves) and the number of
// partitions for `want` (0 = auto), at most what fits 1024 lanes.  Auto: about
// 25 source nodes per partition, at most 4 (a partition's loads hide the
// others' latency; each extra partition costs a barrier and an LDS exchange
// per step).  Measured on MI355X, rWW with one SC per simulation, per step:
// N=68 x 1024 sims (SC in LDS) 3.0 / 2.8 / 3.7 us with 1 / 2 / 4 partitions;
// N=100 x 1024 10.6 / 10.6 / 8.2 us; N=100 x 1 3.6 / 2.3 / 1.7 us; N=200 and
// N=400 within 5% of each other (bandwidth-bound: 7.2-7.6 TB/s of SC reads).
int general_np(int nodes) { return ((nodes + 63) / 64) * 64; }
int general_ksplit(int nodes, int want) {
    const int kmax = std::max(1, std::min(8, 1024 / general_np(nodes)));
    if (want > 0) return std::min(want, kmax);
    return std::max(1, std::min({4, kmax, nodes / 25}));
}

// Pair mode for a request on a launch of `threads` lanes: SC read from global
// memory (not staged, no delay), an even node count of at least 96, the SC
// 16-byte aligned, room for two pair partitions, and a launch that is bound
// by load instructions rather than HBM (`latency_bound`: a shared SC or few
// simulations).  Measured on MI355X, rWW per step: N=400 one simulation 21 ->
// 12.7 us, 256 simulations sharing one SC 21.4 -> 13.1 us, N=200 one
// simulation 6.6 -> 4.3 us; but 256 simulations with one SC each (HBM-bound)
// 43-45 -> 48 us, N=200 x 512 21.5 -> 23 us.
int g_general_pair = 0;  // 0 = auto (latency-bound launches), 1 = whenever eligible, -1 = never

bool general_pair_ok(const IntegrateArgs& a, int threads, bool latency_bound) {
    if (g_general_pair < 0) return false;
    return (latency_bound || g_general_pair > 0) && !a.do_delay && !a.sc_lds && a.nodes % 2 == 0 && a.nodes >= 96 &&
           reinterpret_cast<uintptr_t>(a.SC) % 16 == 0 && 2 * general_npc(a.nodes) <= threads;
}

// SC staging, pair mode and LDS bytes of a request on the general integrator:
// SC staged when it leaves room for two workgroups per CU, or when there are
// too few simulations to fill the CUs anyway (latency-bound).
static size_t general_lds(int model, IntegrateArgs& b, int threads, bool latency_bound) {
    const int np = general_np(b.nodes);
    const size_t with_sc = general_lds_bytes(model, b.nodes, np, b.ksplit, true, false, threads);
    b.sc_lds = with_sc <= 64 * 1024 || (with_sc <= 150 * 1024 && b.N_SIMS <= 512);
    b.pair = general_pair_ok(b, threads, latency_bound);
    return general_lds_bytes(model, b.nodes, np, b.ksplit, b.sc_lds, b.pair != 0, threads);
}

template <class M>
static hipError_t launch_general(int model, const IntegrateArgs& a, const void* consts, hipStream_t st,
                                 bool latency_bound) {
    const typename M::C mc = *static_cast<const typename M::C*>(consts);
    IntegrateArgs b = a;
    if (b.ksplit < 1 || b.ksplit * general_np(a.nodes) > 1024) return hipErrorInvalidValue;
    const int threads = b.ksplit * general_np(a.nodes);
    const size_t lds = general_lds(model, b, threads, latency_bound);
    hipLaunchKernelGGL(integrate_general_kernel<M>, dim3(a.N_SIMS), dim3(threads), lds, st, b, mc);
    return hipGetLastError();
}

size_t model_consts_bytes(int model) {
    switch (model) {
        case MODEL_RWW: return sizeof(RWW::C);
        case MODEL_JR: return sizeof(JR::C);
        case MODEL_KURAMOTO: return sizeof(Kuramoto::C);
        case MODEL_RWWEX: return sizeof(RWWEx::C);
        default: return sizeof(WC::C);
    }
}

// Stride of the model-constants table of a merged launch: the largest model's
// constants, rounded up to 16 bytes.
size_t model_consts_stride() {
    size_t m = 0;
    for (int k : {MODEL_RWW, MODEL_JR, MODEL_KURAMOTO, MODEL_WC, MODEL_RWWEX}) m = std::max(m, model_consts_bytes(k));
    return (m + 15) & ~(size_t)15;
}

// d_args: [n_req] arguments (sc_lds and ksplit set consistently with threads
// and lds: threads = ksplit * general_np(largest node count)); d_consts:
// [n_req] model constants at model_consts_stride(); d_models: [n_req] model
// ids; d_blocks: [n_blocks] (request, sim); progress: null, or two
// host-mapped counters (volumes written, simulations completed).
hipError_t launch_integrate_general_multi(const IntegrateArgs* d_args, const void* d_consts, const int* d_models,
                                          const int2* d_blocks, int n_blocks, int threads, size_t lds,
                                          int* progress, hipStream_t s) {
    if (threads < 64 || threads > 1024 || n_blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(integrate_general_multi_kernel, dim3(n_blocks), dim3(threads), lds, s, d_args,
                       static_cast<const char*>(d_consts), (int)model_consts_stride(), d_models, d_blocks, progress);
    return hipGetLastError();
}

template <class M>
static hipError_t launch_wide(const IntegrateArgs& a, const void* consts, hipStream_t st) {
    const typename M::C mc = *static_cast<const typename M::C*>(consts);
    if (a.do_delay) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)(M::is_osc ? 2 : 1) * a.nodes;
    const int npl = (a.nodes + 1023) / 1024;
    if (npl <= 2) hipLaunchKernelGGL((integrate_wide_kernel<M, 2>), dim3(a.N_SIMS), dim3(1024), lds, st, a, mc);
    else if (npl <= 4) hipLaunchKernelGGL((integrate_wide_kernel<M, 4>), dim3(a.N_SIMS), dim3(1024), lds, st, a, mc);
    else if (npl <= 8) hipLaunchKernelGGL((integrate_wide_kernel<M, 8>), dim3(a.N_SIMS), dim3(1024), lds, st, a, mc);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_integrate_wide(int model, const IntegrateArgs& a, const void* consts, hipStream_t s) {
    if (a.nodes <= 1024 || a.nodes > 8192) return hipErrorInvalidValue;
    switch (model) {
        case MODEL_RWW: return launch_wide<RWW>(a, consts, s);
        case MODEL_JR: return launch_wide<JR>(a, consts, s);
        case MODEL_KURAMOTO: return launch_wide<Kuramoto>(a, consts, s);
        case MODEL_WC: return launch_wide<WC>(a, consts, s);
        case MODEL_RWWEX: return launch_wide<RWWEx>(a, consts, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_integrate_general(int model, const IntegrateArgs& a, const void* consts, hipStream_t s,
                                    bool latency_bound) {
    if (a.nodes < 1 || a.nodes > 1024) return hipErrorInvalidValue;
    switch (model) {
        case MODEL_RWW: return launch_general<RWW>(model, a, consts, s, latency_bound);
        case MODEL_JR: return launch_general<JR>(model, a, consts, s, latency_bound);
        case MODEL_KURAMOTO: return launch_general<Kuramoto>(model, a, consts, s, latency_bound);
        case MODEL_WC: return launch_general<WC>(model, a, consts, s, latency_bound);
        case MODEL_RWWEX: return launch_general<RWWEx>(model, a, consts, s, latency_bound);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cubnm
