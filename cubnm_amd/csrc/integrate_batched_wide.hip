// Batched FP64 integrator for 257..512 nodes: 16 simulations per workgroup,
// coupling on MFMA with the SC fragments streamed from L2.
//
// Same simulation semantics as the reference loop (src/ext/bnm.cu:117-505,
// src/ext/bnm.cpp:69-522) for the no-delay, no-restart case, and the same
// arithmetic as the batched integrator for up to 256 nodes
// (integrate_batched.hip): the coupling input of the group's 16 simulations
// for one step is the product G^T[j][c] = sum_k SC[k][j] x_c[k] on
// v_mfma_f64_16x16x4f64, with the D layout (lane l, register r: node
// 4r + l/16 of the tile, simulation l%16) as the node-to-lane map of the model
// step.  What changes with the node count:
//
//   * An SC of 400 nodes is 1.28 MB: the fragments cannot stay in LDS, so
//     every step streams them from L2 (all workgroups of a launch that share
//     an SC read the same 1.28 MB, which stays L2-resident), two fragment
//     pairs ahead of the MFMAs consuming them.
//   * 16 simulations x 512 nodes are 8192 node slots, 16 per lane on eight
//     waves: the registers hold only each slot's dynamic state (S_E and S_I
//     for rWW) and its noise index.  The regional parameters (read every
//     step) and the Balloon-Windkessel state (read and written every 10th
//     step) live in a per-workgroup lane-ordered copy in global memory, so
//     each of their loads is one contiguous 512-byte wave access.
//   * Tiles are dealt to the eight waves round-robin (tile t of wave w is
//     16 (w + 8t)), so a SIMD's two waves together hold at most
//     ceil(NT/4) tiles; a wave's last tile is skipped on a wave-uniform
//     branch where the tile count does not divide evenly.
//   * MFMA and VALU work of different waves do not overlap on a CDNA4 SIMD,
//     so the whole model step runs after the product; the second wave on the
//     SIMD fills the latency of the step's parameter and noise loads.
#include "engine.hpp"
#include "exp_table.hpp"

namespace cubnm {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

constexpr int WB_WAVES = 8;

struct WideBatchedArgs {
    IntegrateArgs base;
    const int* group_sims;  // [n_groups][16], -1 = padding lane
    const double* afrag;    // [N_SCs][NT][KT][64] SC^T in MFMA A-fragment order
    double* lanes;          // [grid][WB_WAVES][4 (TMIN+1) slots][n_regional+4][64]
    int n_groups, KT, NT;
};

// Wave-uniform value: keeps per-wave branches scalar.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The same pointer, opaque to the optimiser: addresses derived from it are
// computed where they are used instead of being hoisted out of loops.
template <class T>
__device__ __forceinline__ T* fresh(T* p) {
    auto g = (__attribute__((address_space(1))) T*)p;  // global memory: keeps global_* loads and stores
    asm volatile("" : "+v"(g));
    return (T*)g;
}
__device__ __forceinline__ int fresh_int(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

}  // namespace

// TMIN: tiles every wave owns (2 for 257..384 nodes, 3 for 385..512); waves
// with a remainder tile own TMIN + 1.
//
// Register budget: 16 node slots per lane leave no room for one address per
// slot, which the compiler would otherwise hoist out of the time loop (one
// 64-bit address per slot and variable).  Every per-slot access therefore
// starts from a per-lane base passed through fresh(), which the compiler
// cannot hoist, plus a compile-time offset.
//
// LDS: [2][KN][16] coupling-state buffers, [2][N n_noise] noise rows (the
// step's noise row is shared by the 16 simulations: it is loaded once per
// step by the whole workgroup, one step ahead), 64 doubles for the writes of
// lanes whose node is padding.
template <class M, int TMIN>
__global__ void __launch_bounds__(64 * WB_WAVES) integrate_batched_wide_kernel(WideBatchedArgs wa, typename M::C mc) {
    constexpr int TM = TMIN + 1;
    constexpr int NV = M::n_regional + 4;        // lane-copy variables: regional params, then x, f, nu, q
    constexpr int NNZ = M::n_noise;
    const IntegrateArgs& a = wa.base;
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const int c = lane & 15, q = lane >> 4;
    const int N = a.nodes, KT = wa.KT, KN = 4 * KT, NT = wa.NT;
    const int row = N * NNZ;                     // doubles per noise row
    const bool extra = uniform(wave + WB_WAVES * TMIN < NT);  // owns tile TMIN
    const long long NSN = (long long)a.N_SIMS * N;
    extern __shared__ double lds[];
    double* xs = lds;                            // [2][KN][16]
    double* nzs = lds + 2 * KN * 16;             // [2][row]
    double* trash = nzs + 2 * row;               // [64]
    double* etab = trash + 64;                   // [kExpTableSize] for ExpTable
    load_exp_table(etab);
    const ExpTable ex{etab};
    // lane copy of this wave: [slot = 4t + r][NV][64 lanes]
    double* const lw = wa.lanes + (long long)(blockIdx.x * WB_WAVES + wave) * TM * 4 * NV * 64 + lane;
    // node of slot (t, r) is j0 + 128 t + 4 r
    const int j0 = 16 * wave + q;
    // tile t holds nodes 16 (wave + 8t) .. +15; only a tile reaching past N
    // needs a per-node check
    auto tile_full = [&](int t) { return 16 * (wave + WB_WAVES * t) + 16 <= N; };
    // noise row of step (bw step b, inner step i) in noise repeat rp
    auto noise_row = [&](int b, int i, int rp) {
        const int sh_t = a.sh_ts[(b % a.noise_bw_it) + rp * a.noise_bw_it];
        return a.noise + ((long long)sh_t * a.inner_it + i) * row;
    };

    for (int g = blockIdx.x; g < wa.n_groups; g += gridDim.x) {
        const int* gs = wa.group_sims + 16 * g;
        const int sim0 = gs[0];
        const int sim = gs[c];
        const bool sim_ok = sim >= 0;
        const int simr = sim_ok ? sim : sim0;
        const int sc_idx = a.SC_indices[sim0];
        const double* afrag = wa.afrag + (long long)sc_idx * NT * KT * 64;
        double* const st_lane = a.states + (long long)sim * N + j0;             // + i NSN + 128 t + 4 r
        double* const bold_lane = a.bold + (long long)simr * a.bold_len * N + j0;  // + bold_i N + 128 t + 4 r

        double gp[M::n_global];
#pragma unroll
        for (int i = 0; i < M::n_global; i++) gp[i] = a.gp[(long long)i * a.N_SIMS + simr];

        double s[TM][4][M::n_state];
        int shj[TM][4];  // noise offset of the slot's node within a row (shuffled node x n_noise)
        // zero both X buffers (padding rows stay finite: A is zero there);
        // the first step's noise row
        for (int i = threadIdx.x; i < 2 * KN * 16; i += blockDim.x) xs[i] = 0.0;
        {
            const double* nr = noise_row(0, 0, 0);
            for (int i = threadIdx.x; i < row; i += blockDim.x) nzs[i] = nr[i];
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TM; t++) {
            if (t == TMIN && !extra) break;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = j0 + 128 * t + 4 * r;
                const int jr = j < N ? j : 0;
                double* ps = fresh(lw) + (4 * t + r) * NV * 64;
                double rp[M::n_regional];
#pragma unroll
                for (int i = 0; i < M::n_regional; i++) rp[i] = a.rp[(long long)i * NSN + (long long)simr * N + jr];
#pragma unroll
                for (int i = 0; i < M::n_state; i++) s[t][r][i] = 0.0;
                M::init(mc, s[t][r], rp);
                M::scale_rp(mc, rp);  // the split steps' form of the parameters
#pragma unroll
                for (int i = 0; i < M::n_regional; i++) ps[i * 64] = rp[i];
                ps[(M::n_regional + 0) * 64] = 0.0;  // x
                ps[(M::n_regional + 1) * 64] = 1.0;  // f
                ps[(M::n_regional + 2) * 64] = 1.0;  // nu
                ps[(M::n_regional + 3) * 64] = 1.0;  // q
                shj[t][r] = a.sh_nodes[jr] * NNZ;
                if (j < N) xs[j * 16 + c] = s[t][r][M::conn_idx];
                if (j < N && sim_ok)
#pragma unroll
                    for (int i = 0; i < M::n_state; i++) st_lane[i * NSN + 128 * t + 4 * r] = 0.0;
            }
        }
        __syncthreads();

        const double* ab[TM];
#pragma unroll
        for (int t = 0; t < TM; t++) ab[t] = afrag + (long long)(wave + WB_WAVES * (t < TMIN || extra ? t : 0)) * KT * 64 + lane;
        int buf = 0, bw_i = 0, bold_i = 0, states_i = 0, rep = 0;
        while (bw_i < a.bw_it) {
            // states are summed from the last inner step of a sampling bw step
            const bool sample = ((bw_i + 1) % a.states_sampling_iters == 0) &&
                                (states_i >= a.n_states_samples_remove);
            for (int inner_i = 0; inner_i < a.inner_it; inner_i++) {
                // next step's noise row: loaded now, stored to LDS after the
                // model step (the loads complete under the MFMAs)
                int nb = inner_i + 1, nbw = bw_i, nrep = rep;
                if (nb == a.inner_it) {
                    nb = 0; nbw = bw_i + 1;
                    if (nbw % a.noise_bw_it == 0) nrep++;
                }
                const bool has_next = nbw < a.bw_it;
                double nx0 = 0.0, nx1 = 0.0;
                if (has_next) {
                    const double* nr = noise_row(nbw, nb, nrep);
                    if ((int)threadIdx.x < row) nx0 = nr[threadIdx.x];
                    if ((int)threadIdx.x + 64 * WB_WAVES < row) nx1 = nr[threadIdx.x + 64 * WB_WAVES];
                }
                const double* bb = xs + buf * KN * 16 + lane;
                v4d acc[TM];
#pragma unroll
                for (int t = 0; t < TM; t++) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
                // two register sets of two k-steps each: the loads of one
                // set are in flight while the MFMAs consume the other
                double a0[TM], a1[TM], a2[TM], a3[TM];
                double b0, b1, b2, b3;
                auto load_pair = [&](double* x0, double* x1, double& y0, double& y1, int k) {
#pragma unroll
                    for (int t = 0; t < TMIN; t++) { x0[t] = ab[t][k * 64]; x1[t] = ab[t][(k + 1) * 64]; }
                    x0[TMIN] = x1[TMIN] = 0.0;
                    if (extra) { x0[TMIN] = ab[TMIN][k * 64]; x1[TMIN] = ab[TMIN][(k + 1) * 64]; }
                    y0 = bb[k * 64]; y1 = bb[(k + 1) * 64];
                };
                auto mfma_pair = [&](const double* x0, const double* x1, double y0, double y1) {
#pragma unroll
                    for (int t = 0; t < TMIN; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0[t], y0, acc[t], 0, 0, 0);
                    if (extra) acc[TMIN] = __builtin_amdgcn_mfma_f64_16x16x4f64(x0[TMIN], y0, acc[TMIN], 0, 0, 0);
#pragma unroll
                    for (int t = 0; t < TMIN; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1[t], y1, acc[t], 0, 0, 0);
                    if (extra) acc[TMIN] = __builtin_amdgcn_mfma_f64_16x16x4f64(x1[TMIN], y1, acc[TMIN], 0, 0, 0);
                };
                load_pair(a0, a1, b0, b1, 0);
                for (int kt = 0; kt < KT; kt += 4) {
                    // KT is a multiple of 4; the last iteration's prefetch
                    // re-reads fragments 0..1 (in bounds, unused)
                    const int kn = kt + 4 < KT ? kt + 4 : 0;
                    load_pair(a2, a3, b2, b3, kt + 2);
                    __builtin_amdgcn_sched_barrier(0);
                    mfma_pair(a0, a1, b0, b1);
                    __builtin_amdgcn_sched_barrier(0);
                    load_pair(a0, a1, b0, b1, kn);
                    __builtin_amdgcn_sched_barrier(0);
                    mfma_pair(a2, a3, b2, b3);
                    __builtin_amdgcn_sched_barrier(0);
                }
                double* xn = xs + (buf ^ 1) * KN * 16 + j0 * 16 + c;  // + (128 t + 4 r) 16
                const double* nzl = nzs + buf * row;   // this step's row (the noise buffers alternate with the X buffers)
                const bool last = sample && (inner_i == a.inner_it - 1);
#pragma unroll
                for (int t = 0; t < TM; t++) {
                    if (t == TMIN && !extra) break;
                    const bool full = tile_full(t);
                    double n0[4], n1[4], rp[4][M::n_regional];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        n0[r] = nzl[shj[t][r]];
                        n1[r] = (NNZ > 1) ? nzl[shj[t][r] + 1] : 0.0;
                        const double* ps = fresh(lw) + (4 * t + r) * NV * 64;
#pragma unroll
                        for (int i = 0; i < M::n_regional; i++) rp[r][i] = ps[i * 64];
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        double tp[M::n_tmp];
                        M::pre(mc, s[t][r], gp, rp[r], n0[r], n1[r], tp, ex);
                        M::post(mc, s[t][r], gp, rp[r], acc[t][r], n0[r], n1[r], tp, ex);
                        const bool valid = full || fresh_int(j0) + 128 * t + 4 * r < N;
                        *(valid ? xn + (128 * t + 4 * r) * 16 : trash + lane) = s[t][r][M::conn_idx];
                        if (last && valid && sim_ok) {
                            double* o = fresh(st_lane) + 128 * t + 4 * r;
#pragma unroll
                            for (int i = 0; i < M::n_state; i++) o[i * NSN] += s[t][r][i];
                        }
                    }
                }
                if (has_next) {
                    double* nzn = nzs + (buf ^ 1) * row;
                    if ((int)threadIdx.x < row) nzn[threadIdx.x] = nx0;
                    if ((int)threadIdx.x + 64 * WB_WAVES < row) nzn[threadIdx.x + 64 * WB_WAVES] = nx1;
                }
                buf ^= 1;
                lds_barrier();
            }
            const bool out = (bw_i + 1) % a.BOLD_TR_iters == 0 && sim_ok && bold_i < a.bold_len;
#pragma unroll
            for (int t = 0; t < TM; t++) {
                if (t == TMIN && !extra) break;
                const bool full = tile_full(t);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    double* ps = fresh(lw) + ((4 * t + r) * NV + M::n_regional) * 64;
                    double x = ps[0], f = ps[64], nu = ps[128], qq = ps[192];
                    bw_step_fast(a.bw, s[t][r][M::bold_idx], x, f, nu, qq, ex);
                    ps[0] = x; ps[64] = f; ps[128] = nu; ps[192] = qq;
                    if (out && (full || fresh_int(j0) + 128 * t + 4 * r < N))
                        fresh(bold_lane)[(long long)bold_i * N + 128 * t + 4 * r] = bold_signal(a.bw, nu, qq);
                }
            }
            if ((bw_i + 1) % a.BOLD_TR_iters == 0) bold_i++;
            if ((bw_i + 1) % a.states_sampling_iters == 0) states_i++;
            if ((bw_i + 1) % a.noise_bw_it == 0 && bw_i + 1 < a.bw_it) {
                rep++;
#pragma unroll
                for (int t = 0; t < TM; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int j = j0 + 128 * t + 4 * r;
                        shj[t][r] = a.sh_nodes[rep * N + (j < N ? j : 0)] * NNZ;
                    }
            }
            bw_i++;
        }
        if (sim_ok) {
            const int tp = states_i - a.n_states_samples_remove;
#pragma unroll
            for (int t = 0; t < TM; t++) {
                if (t == TMIN && !extra) break;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (j0 + 128 * t + 4 * r < N)
#pragma unroll
                        for (int i = 0; i < M::n_state; i++) {
                            double* o = st_lane + i * NSN + 128 * t + 4 * r;
                            *o = *o / tp;
                        }
                }
            }
        }
        __syncthreads();
    }
}

namespace {

size_t batched_wide_lds_bytes(int KT, int nodes, int n_noise) {
    return sizeof(double) * (2 * 4 * (size_t)KT * 16 + 2 * (size_t)nodes * n_noise + 64 + kExpTableSize);
}

template <class M, int TMIN>
hipError_t launch_wide_batched_t(const WideBatchedArgs& wa, const void* consts, hipStream_t st, int grid) {
    const typename M::C mc = *static_cast<const typename M::C*>(consts);
    const size_t lds = batched_wide_lds_bytes(wa.KT, wa.base.nodes, M::n_noise);
    hipLaunchKernelGGL((integrate_batched_wide_kernel<M, TMIN>), dim3(grid), dim3(64 * WB_WAVES), lds, st, wa, mc);
    return hipGetLastError();
}

template <class M>
hipError_t dispatch_wide(const WideBatchedArgs& wa, const void* consts, hipStream_t st, int grid) {
    if (wa.NT > 3 * WB_WAVES) return launch_wide_batched_t<M, 3>(wa, consts, st, grid);
    return launch_wide_batched_t<M, 2>(wa, consts, st, grid);
}

}  // namespace

bool batched_wide_supported(int model, int nodes) {
    return nodes > 256 && nodes <= 512 && (model == MODEL_RWW || model == MODEL_RWWEX || model == MODEL_WC);
}

// Workgroups of the wide launch: one per CU (each takes 8 waves and up to
// 128 KB of LDS), never more than there are groups.
int batched_wide_grid(int n_groups) {
    int dev = 0, n_cu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    return std::max(1, std::min(n_groups, n_cu));
}

// Doubles of the lane-ordered parameter and BW-state copy for `grid` workgroups.
size_t batched_wide_lane_doubles(int model, int nodes, int grid) {
    int nreg = 0;
    switch (model) {
        case MODEL_RWW: nreg = RWW::n_regional; break;
        case MODEL_RWWEX: nreg = RWWEx::n_regional; break;
        case MODEL_WC: nreg = WC::n_regional; break;
        default: return 0;
    }
    const int nt = (nodes + 15) / 16;
    const int tm = (nt > 3 * WB_WAVES ? 3 : 2) + 1;
    return (size_t)grid * WB_WAVES * tm * (nreg + 4) * 4 * 64;
}

hipError_t launch_integrate_batched_wide(int model, const IntegrateArgs& a, const void* consts, hipStream_t s,
                                         const int* d_group_sims, int n_groups, const double* d_afrag, int KT,
                                         double* d_lanes, int grid) {
    WideBatchedArgs wa;
    wa.base = a;
    wa.group_sims = d_group_sims;
    wa.afrag = d_afrag;
    wa.lanes = d_lanes;
    wa.n_groups = n_groups;
    wa.KT = KT;
    wa.NT = (a.nodes + 15) / 16;
    if (!batched_wide_supported(model, a.nodes) || KT < 4 || (KT & 3) || 4 * KT < a.nodes) return hipErrorInvalidValue;
    if (grid < 1 || grid > n_groups) return hipErrorInvalidConfiguration;
    switch (model) {
        case MODEL_RWW: return dispatch_wide<RWW>(wa, consts, s, grid);
        case MODEL_RWWEX: return dispatch_wide<RWWEx>(wa, consts, s, grid);
        case MODEL_WC: return dispatch_wide<WC>(wa, consts, s, grid);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cubnm
