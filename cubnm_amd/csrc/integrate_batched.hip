// Batched FP64 integrator: 16 simulations per workgroup, coupling on MFMA.
//
// Same simulation semantics as the reference loop (src/ext/bnm.cu:117-505,
// src/ext/bnm.cpp:69-522) for the no-delay, no-restart case.  The MI355X
// design replaces the reference's "one block per simulation, one thread per
// node, N LDS broadcasts per node per step" with:
//
//   * a workgroup integrates a group of 16 simulations that share one SC;
//   * the coupling input of all 16 simulations for one step is the product
//       G^T[j][c] = sum_k SC[k][j] * x_c[k]    (c = simulation in the group)
//     computed with v_mfma_f64_16x16x4f64: A = SC^T tiles (16 nodes x 4
//     sources) stored once per SC in MFMA fragment order (in LDS when they
//     fit, else streamed from L2), B = the group's coupling state (4 sources x
//     16 simulations) read from a double-buffered LDS tile, D = 16 nodes x 16
//     simulations;
//   * the f64 16x16x4 D layout (lane l, register r: row 4*r + l/16, column
//     l%16; A[i][k] in lane i + 16k, B[k][j] in lane j + 16k) is the
//     node-to-lane map of the model step: lane l updates nodes 4r + l/16 of
//     simulation l%16 in every 16-node tile, without moving the coupling result;
//   * each wave owns TPW node tiles; one workgroup barrier per inner step
//     (double buffering makes a second barrier unnecessary);
//   * oscillator coupling sum_k SC[k][j] sin(x_k - x_j) is evaluated as
//     cos(x_j) * (SC^T sin x)_j - sin(x_j) * (SC^T cos x)_j, i.e. two products.
//
// The noise element of (step, node) is shared by all simulations of the
// group (one seed), so its loads are broadcasts.  BOLD, states and FIC
// outputs use the reference layouts.
#include <type_traits>
#include "engine.hpp"
#include "exp_table.hpp"

namespace cubnm {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef int v2i __attribute__((ext_vector_type(2)));
typedef int v4i __attribute__((ext_vector_type(4)));

// word 3 of a raw buffer resource on gfx9-family GPUs (32-bit data format,
// no swizzle); out-of-range loads return zero
constexpr int kBufferConfig = 0x00020000;

struct BatchedArgs {
    IntegrateArgs base;
    const int* group_sims;  // [n_groups][16], -1 = padding lane
    const double* afrag;    // [N_SCs][NT][KT][64] SC^T in MFMA A-fragment order
    int n_groups, KT, NT;   // KT: fragments per tile (even), NT: tiles (multiple of TPW)
};

int batched_kt(int nodes);

namespace {

template <class M>
__device__ __forceinline__ void publish(double* xs, int kn, int j, int c, double v, double* sn, double* cs) {
    if constexpr (M::is_osc) {
        double sv, cv;
        sincos(v, &sv, &cv);
        *sn = sv; *cs = cv;
        xs[j * 16 + c] = sv;
        xs[kn * 16 + j * 16 + c] = cv;
    } else {
        xs[j * 16 + c] = v;
    }
}

}  // namespace

// A_LDS: A fragments staged in LDS (small N) or read from global memory.
// KTC: compile-time fragment count per tile (0 = runtime ba.KT); KTU: the
// fragments the product uses (ceil(N/4)), at most KTC (the layout keeps KTC
// even; the last fragment is all zero when N <= 4 KTC - 4).
template <class M, int TPW, int MAXW, bool A_LDS, int KTC, int KTU>
__global__ void __launch_bounds__(64 * MAXW) integrate_batched_kernel(BatchedArgs ba, typename M::C mc) {
    constexpr int NS = 4 * TPW;                 // node slots per lane
    constexpr int NX = M::is_osc ? 2 : 1;        // B operands per step
    const IntegrateArgs& a = ba.base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, q = lane >> 4;
    const int N = a.nodes, KT = KTC > 0 ? KTC : ba.KT, KN = 4 * KT, NT = ba.NT;
    // a plane holds the KN source rows and one spare row (KN) that takes
    // the writes of padding rows; no product reads it
    const int KNP = KN + 1;
    const int buf_stride = NX * KNP * 16;       // doubles per X buffer
    extern __shared__ double lds[];
    double* etab = lds;                          // [kExpTableSize] for ExpTable
    double* xs = lds + kExpTableSize;            // [2][NX][KNP][16]
    double* as = xs + 2 * buf_stride;            // [NT][KT][64] when A_LDS
    load_exp_table(etab);
    const ExpTable ex{etab};
    int loaded_sc = -1;

    for (int g = blockIdx.x; g < ba.n_groups; g += gridDim.x) {
        const int* gs = ba.group_sims + 16 * g;
        const int sim0 = gs[0];
        const int sim = gs[c];
        const bool sim_ok = sim >= 0;
        const int simr = sim_ok ? sim : sim0;
        const int sc_idx = a.SC_indices[sim0];
        const double* afrag = ba.afrag + (long long)sc_idx * NT * KT * 64;
        if (A_LDS && sc_idx != loaded_sc) {
            for (int i = threadIdx.x; i < NT * KT * 64; i += blockDim.x) as[i] = afrag[i];
            loaded_sc = sc_idx;
        }
        const double* abase = A_LDS ? as : afrag;

        double gp[M::n_global];
#pragma unroll
        for (int i = 0; i < M::n_global; i++) gp[i] = a.gp[(long long)i * a.N_SIMS + simr];

        double s[NS][M::n_state], rp[NS][M::n_regional];
        double bx[NS], bf[NS], bn[NS], bq[NS], bnp[NS], sn[NS], cs[NS];
        unsigned shb[NS];  // byte offset of the node's noise in a step's row
        // zero both X buffers (padding rows stay finite: A is zero there)
        for (int i = threadIdx.x; i < 2 * buf_stride; i += blockDim.x) xs[i] = 0.0;
        __syncthreads();
#pragma unroll
        for (int sl = 0; sl < NS; sl++) {
            const int j = 16 * (wave * TPW + sl / 4) + 4 * (sl & 3) + q;
            const int jr = j < N ? j : 0;
#pragma unroll
            for (int i = 0; i < M::n_regional; i++)
                rp[sl][i] = a.rp[(long long)i * a.N_SIMS * N + (long long)simr * N + jr];
#pragma unroll
            for (int i = 0; i < M::n_state; i++) s[sl][i] = 0.0;
            M::init(mc, s[sl], rp[sl]);
            M::scale_rp(mc, rp[sl]);  // the split steps' form of the parameters
            bx[sl] = 0.0; bf[sl] = 1.0; bn[sl] = 1.0; bq[sl] = 1.0;
            bnp[sl] = bw_nu_pow(a.bw, 1.0);  // nu^(1/alpha) of the initial nu
            sn[sl] = 0.0; cs[sl] = 1.0;
            shb[sl] = (unsigned)a.sh_nodes[jr] * (unsigned)(sizeof(double) * M::n_noise);
            if (j < N) publish<M>(xs, KNP, j, c, s[sl][M::conn_idx], &sn[sl], &cs[sl]);
            if (j < N && sim_ok)
#pragma unroll
                for (int i = 0; i < M::n_state; i++) a.states[(long long)i * a.N_SIMS * N + (long long)sim * N + j] = 0.0;
        }
        __syncthreads();

        double* bold = a.bold + (long long)simr * a.bold_len * N;
        int buf = 0, bw_i = 0, bold_i = 0, states_i = 0, rep = 0;
        while (bw_i < a.bw_it) {
            // (readfirstlane: the value is the same in every lane; it keeps the
            // noise buffer below in scalar registers)
            const int sh_t = __builtin_amdgcn_readfirstlane(a.sh_ts[(bw_i % a.noise_bw_it) + rep * a.noise_bw_it]);
            // states are summed from the last inner step of a sampling bw step
            const bool sample = ((bw_i + 1) % a.states_sampling_iters == 0) &&
                                (states_i >= a.n_states_samples_remove) && sim_ok;
            // this bw step's noise rows as one buffer: the loads below take
            // the row offset in a scalar register and the node's offset in a
            // vector one, with no vector address arithmetic
            const int row_bytes = N * M::n_noise * (int)sizeof(double);
            const __amdgpu_buffer_rsrc_t nbuf = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<double*>(a.noise) + (long long)sh_t * a.inner_it * N * M::n_noise, 0,
                a.inner_it * row_bytes, kBufferConfig);
            // one inner step; B, the X buffer it reads, is a compile-time
            // constant, so the LDS addresses of both buffers are immediate offsets
            auto inner_step = [&](auto BUF, const int inner_i) __attribute__((always_inline)) {
                constexpr int B = decltype(BUF)::value;
                // this step's noise: issued before the coupling product so the
                // loads complete under the MFMAs
                double n0[NS], n1[NS];
#pragma unroll
                for (int sl = 0; sl < NS; sl++) {
                    if constexpr (M::n_noise > 1) {
                        const v4i v = __builtin_amdgcn_raw_buffer_load_b128(nbuf, (int)shb[sl], inner_i * row_bytes, 0);
                        n0[sl] = __builtin_bit_cast(double, (v2i){v[0], v[1]});
                        n1[sl] = __builtin_bit_cast(double, (v2i){v[2], v[3]});
                    } else {
                        const v2i v = __builtin_amdgcn_raw_buffer_load_b64(nbuf, (int)shb[sl], inner_i * row_bytes, 0);
                        n0[sl] = __builtin_bit_cast(double, v);
                        n1[sl] = 0.0;
                    }
                }
                // the part of the model step that does not read this step's
                // coupling input, placed before the product so that the
                // scheduler can issue it between the MFMAs
                double tp[NS][M::n_tmp];
#pragma unroll
                for (int sl = 0; sl < NS; sl++) M::pre(mc, s[sl], gp, rp[sl], n0[sl], n1[sl], tp[sl], ex);
                const double* xb = xs + B * buf_stride;
                // two accumulators per product (even and odd source
                // fragments), added at the end: two independent MFMA chains
                // instead of one dependent one
                v4d acc[TPW][NX], acc2[TPW][NX];
#pragma unroll
                for (int t = 0; t < TPW; t++)
#pragma unroll
                    for (int x = 0; x < NX; x++) {
                        acc[t][x] = v4d{0.0, 0.0, 0.0, 0.0};
                        acc2[t][x] = v4d{0.0, 0.0, 0.0, 0.0};
                    }
                const double* ab = abase + (long long)(wave * TPW) * KT * 64 + lane;
                const double* bb = xb + lane;
                if constexpr (KTC > 0) {
                    // compile-time fragment count: fully unrolled, immediate LDS offsets
#pragma unroll
                    for (int kt = 0; kt < KTU; kt++) {
#pragma unroll
                        for (int x = 0; x < NX; x++) {
                            const double b = bb[x * KNP * 16 + kt * 64];
#pragma unroll
                            for (int t = 0; t < TPW; t++) {
                                if (kt & 1)
                                    acc2[t][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(ab[(t * KT + kt) * 64], b,
                                                                                      acc2[t][x], 0, 0, 0);
                                else
                                    acc[t][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(ab[(t * KT + kt) * 64], b,
                                                                                     acc[t][x], 0, 0, 0);
                            }
                        }
                    }
                } else {
                    // runtime fragment count: software-pipelined two k-steps ahead
                    double a0[TPW], a1[TPW], b0[NX], b1[NX];
#pragma unroll
                    for (int t = 0; t < TPW; t++) { a0[t] = ab[t * KT * 64]; a1[t] = ab[(t * KT + 1) * 64]; }
#pragma unroll
                    for (int x = 0; x < NX; x++) { b0[x] = bb[x * KNP * 16]; b1[x] = bb[x * KNP * 16 + 64]; }
                    for (int kt = 0; kt < KT; kt += 2) {
                        const int k2 = min(kt + 2, KT - 2), k3 = k2 + 1;
                        double a2[TPW], a3[TPW], b2[NX], b3[NX];
#pragma unroll
                        for (int t = 0; t < TPW; t++) { a2[t] = ab[(t * KT + k2) * 64]; a3[t] = ab[(t * KT + k3) * 64]; }
#pragma unroll
                        for (int x = 0; x < NX; x++) { b2[x] = bb[x * KNP * 16 + k2 * 64]; b3[x] = bb[x * KNP * 16 + k3 * 64]; }
                        // keep the prefetch ahead of the MFMAs consuming the previous pair
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int x = 0; x < NX; x++)
#pragma unroll
                            for (int t = 0; t < TPW; t++) {
                                acc[t][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], b0[x], acc[t][x], 0, 0, 0);
                                acc2[t][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], b1[x], acc2[t][x], 0, 0, 0);
                            }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int t = 0; t < TPW; t++) { a0[t] = a2[t]; a1[t] = a3[t]; }
#pragma unroll
                        for (int x = 0; x < NX; x++) { b0[x] = b2[x]; b1[x] = b3[x]; }
                    }
                }
#pragma unroll
                for (int t = 0; t < TPW; t++)
#pragma unroll
                    for (int x = 0; x < NX; x++) acc[t][x] += acc2[t][x];
                double* xn = xs + (B ^ 1) * buf_stride;
#pragma unroll
                for (int t = 0; t < TPW; t++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int sl = 4 * t + r;
                        double gin;
                        if constexpr (M::is_osc) gin = cs[sl] * acc[t][0][r] - sn[sl] * acc[t][1][r];
                        else gin = acc[t][0][r];
                        M::post(mc, s[sl], gp, rp[sl], gin, n0[sl], n1[sl], tp[sl], ex);
                    }
                // every slot writes its coupling state, padding rows into the
                // spare row: no branch between the slots' steps, so they form
                // one block whose dependent chains the scheduler interleaves
#pragma unroll
                for (int sl = 0; sl < NS; sl++) {
                    const int j = 16 * (wave * TPW + sl / 4) + 4 * (sl & 3) + q;
                    publish<M>(xn, KNP, j < N ? j : KN, c, s[sl][M::conn_idx], &sn[sl], &cs[sl]);
                }
                if (sample && inner_i == a.inner_it - 1) {
#pragma unroll
                    for (int sl = 0; sl < NS; sl++) {
                        const int j = 16 * (wave * TPW + sl / 4) + 4 * (sl & 3) + q;
                        if (j < N)
#pragma unroll
                            for (int i = 0; i < M::n_state; i++)
                                a.states[(long long)i * a.N_SIMS * N + (long long)sim * N + j] += s[sl][i];
                    }
                }
                __syncthreads();
            };
            using B0 = std::integral_constant<int, 0>;
            using B1 = std::integral_constant<int, 1>;
            int inner_i = 0;
            if (buf == 1) inner_step(B1{}, inner_i++);
            for (; inner_i + 1 < a.inner_it; inner_i += 2) {
                inner_step(B0{}, inner_i);
                inner_step(B1{}, inner_i + 1);
            }
            buf = 0;
            if (inner_i < a.inner_it) {
                inner_step(B0{}, inner_i);
                buf = 1;
            }
#pragma unroll
            for (int sl = 0; sl < NS; sl++) {
                // carrying nu^(1/alpha) costs a register pair per slot: only
                // where the slots leave room (one tile per wave, and not the
                // JR and WC models with their 6 and 8 regional parameters)
                if constexpr (TPW == 1 && M::n_state + 2 * M::n_regional <= 14)
                    bw_step_carry(a.bw, s[sl][M::bold_idx], bx[sl], bf[sl], bn[sl], bq[sl], bnp[sl], ex);
                else
                    bw_step_fast(a.bw, s[sl][M::bold_idx], bx[sl], bf[sl], bn[sl], bq[sl], ex);
            }
            if ((bw_i + 1) % a.BOLD_TR_iters == 0) {
                if (sim_ok && bold_i < a.bold_len)
#pragma unroll
                    for (int sl = 0; sl < NS; sl++) {
                        const int j = 16 * (wave * TPW + sl / 4) + 4 * (sl & 3) + q;
                        if (j < N) bold[(long long)bold_i * N + j] = bold_signal(a.bw, bn[sl], bq[sl]);
                    }
                bold_i++;
            }
            if ((bw_i + 1) % a.states_sampling_iters == 0) states_i++;
            if ((bw_i + 1) % a.noise_bw_it == 0 && bw_i + 1 < a.bw_it) {
                rep++;
#pragma unroll
                for (int sl = 0; sl < NS; sl++) {
                    const int j = 16 * (wave * TPW + sl / 4) + 4 * (sl & 3) + q;
                    shb[sl] = (unsigned)a.sh_nodes[rep * N + (j < N ? j : 0)] * (unsigned)(sizeof(double) * M::n_noise);
                }
            }
            bw_i++;
        }
        if (sim_ok) {
            const int tp = states_i - a.n_states_samples_remove;
#pragma unroll
            for (int sl = 0; sl < NS; sl++) {
                const int j = 16 * (wave * TPW + sl / 4) + 4 * (sl & 3) + q;
                if (j < N)
#pragma unroll
                    for (int i = 0; i < M::n_state; i++) {
                        double* o = a.states + (long long)i * a.N_SIMS * N + (long long)sim * N + j;
                        *o = *o / tp;
                    }
            }
        }
        __syncthreads();
    }
}

namespace {

// SC^T in A-fragment order: frag[(jt*KT + kt)*64 + lane] = SC[(4kt + lane/16)][16jt + lane%16]
__global__ void build_afrag_kernel(const double* SC, int n_scs, int N, int NT, int KT, double* afrag) {
    const long long total = (long long)n_scs * NT * KT * 64;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const int lane = (int)(i % 64);
        long long r = i / 64;
        const int kt = (int)(r % KT); r /= KT;
        const int jt = (int)(r % NT);
        const int sc = (int)(r / NT);
        const int k = 4 * kt + lane / 16, j = 16 * jt + lane % 16;
        afrag[i] = (k < N && j < N) ? SC[(long long)sc * N * N + (long long)k * N + j] : 0.0;
    }
}

template <class M, int TPW, int MAXW, bool A_LDS, int KTC = 0, int KTU = KTC>
hipError_t launch_one(const BatchedArgs& ba, const void* consts, int waves, hipStream_t st) {
    if (waves > MAXW) return hipErrorInvalidConfiguration;
    const typename M::C mc = *static_cast<const typename M::C*>(consts);
    const int nx = M::is_osc ? 2 : 1;
    size_t lds = sizeof(double) * (kExpTableSize + 2 * nx * (4 * ba.KT + 1) * 16);
    if (A_LDS) lds += sizeof(double) * (size_t)ba.NT * ba.KT * 64;
    if (KTC > 0 && (ba.KT != KTC || 4 * KTU < ba.base.nodes)) return hipErrorInvalidConfiguration;
    auto kern = integrate_batched_kernel<M, TPW, MAXW, A_LDS, KTC, KTU>;
    const int threads = 64 * waves;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    int dev = 0, n_cu = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    const int grid = std::min(ba.n_groups, std::max(1, per_cu * n_cu));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, st, ba, mc);
    return hipGetLastError();
}

// alds: SC fragments in LDS when they fit (1, or -1 = automatic) or
// streamed from L2 (0).
template <class M>
hipError_t dispatch(const BatchedArgs& ba, const void* consts, int tpw, int alds, hipStream_t st) {
    const int waves = ba.NT / tpw;  // NT is padded to a multiple of tpw
    const int nx = M::is_osc ? 2 : 1;
    const size_t lds_a = sizeof(double) * ((size_t)ba.NT * ba.KT * 64 + kExpTableSize + 2 * nx * (4 * ba.KT + 1) * 16);
    const bool a_lds = alds != 0 && lds_a <= 150 * 1024;
    if (tpw == 1) {
        if (waves <= 4) return a_lds ? launch_one<M, 1, 4, true>(ba, consts, waves, st) : launch_one<M, 1, 4, false>(ba, consts, waves, st);
        if (waves <= 8 && a_lds && ba.KT == 26) {
            if (ba.base.nodes <= 100) return launch_one<M, 1, 8, true, 26, 25>(ba, consts, waves, st);  // N = 97..100
            return launch_one<M, 1, 8, true, 26, 26>(ba, consts, waves, st);                            // N = 101..104
        }
        if (waves <= 8) return a_lds ? launch_one<M, 1, 8, true>(ba, consts, waves, st) : launch_one<M, 1, 8, false>(ba, consts, waves, st);
        return hipErrorInvalidValue;
    }
    if (waves <= 4) return a_lds ? launch_one<M, 2, 4, true>(ba, consts, waves, st) : launch_one<M, 2, 4, false>(ba, consts, waves, st);
    if (waves <= 8) return launch_one<M, 2, 8, false>(ba, consts, waves, st);
    return hipErrorInvalidValue;
}

}  // namespace

int g_batched_tpw = 0;  // 0 = heuristic; set by the tuning entry point

bool batched_wide_supported(int model, int nodes);
int batched_wide_grid(int n_groups);
hipError_t launch_integrate_batched_wide(int model, const IntegrateArgs& a, const void* consts, hipStream_t s,
                                         const int* d_group_sims, int n_groups, const double* d_afrag, int KT,
                                         double* d_lanes, int grid);

// Up to 256 nodes every model (SC fragments in LDS or streamed, tiles per
// wave 1 or 2); 257..512 nodes the conductance models with at most two
// dynamic states per node (integrate_batched_wide.hip).
bool batched_supported(int model, int nodes) {
    return (nodes >= 1 && nodes <= 256) || batched_wide_supported(model, nodes);
}

void batched_plan(int nodes, int want_tpw, int* tpw, int* nt_pad);

// Whether the SC fragments of a launch fit in LDS beside the coupling-state
// buffers (what dispatch does with alds = -1).
bool batched_afrag_fits_lds(int model, int nodes, int tpw) {
    if (nodes > 256) return false;
    int t, nt_pad;
    batched_plan(nodes, tpw, &t, &nt_pad);
    const int KT = batched_kt(nodes), nx = model == MODEL_KURAMOTO ? 2 : 1;
    return sizeof(double) * ((size_t)nt_pad * KT * 64 + kExpTableSize + 2 * nx * (4 * KT + 1) * 16) <= 150 * 1024;
}

// Tiles per wave and the padded tile count (waves * TPW) the fragment array
// and the kernel use; every wave then owns exactly TPW allocated tiles.
void batched_plan(int nodes, int want_tpw, int* tpw, int* nt_pad) {
    const int nt = (nodes + 15) / 16;
    if (nodes > 256) {  // wide integrator: tiles dealt round-robin to 8 waves, no padding tiles
        *tpw = 0;
        *nt_pad = nt;
        return;
    }
    int t = want_tpw != 0 ? want_tpw : g_batched_tpw;
    if (t != 1 && t != 2) t = (nt > 8) ? 2 : 1;
    const int waves = (nt + t - 1) / t;
    *tpw = t;
    *nt_pad = waves * t;
}

// Source-node fragments per tile: ceil(nodes/4) rounded up to even (the
// coupling loop consumes fragments in pairs), to a multiple of 4 above 256
// nodes (the wide integrator's loop consumes them in fours); padded fragments
// are zero.
int batched_kt(int nodes) {
    const int kt = (nodes + 3) / 4;
    return nodes > 256 ? (kt + 3) & ~3 : (kt + 1) & ~1;
}

size_t batched_afrag_doubles(int nodes, int nt_pad, int n_scs) {
    const int KT = batched_kt(nodes);
    return (size_t)n_scs * nt_pad * KT * 64;
}

hipError_t build_afrag(const double* d_SC, int n_scs, int nodes, int nt_pad, double* d_afrag, hipStream_t s) {
    const int KT = batched_kt(nodes), NT = nt_pad;
    const long long total = (long long)n_scs * NT * KT * 64;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(build_afrag_kernel, dim3(blocks), dim3(256), 0, s, d_SC, n_scs, nodes, NT, KT, d_afrag);
    return hipGetLastError();
}

hipError_t launch_integrate_batched(int model, const IntegrateArgs& a, const void* consts, hipStream_t s,
                                    int* launches, const int* d_group_sims, int n_groups, const double* d_afrag,
                                    int want_tpw, int alds, double* d_lanes) {
    if (a.nodes > 256) {
        *launches = 1;
        return launch_integrate_batched_wide(model, a, consts, s, d_group_sims, n_groups, d_afrag, batched_kt(a.nodes),
                                             d_lanes, batched_wide_grid(n_groups));
    }
    BatchedArgs ba;
    ba.base = a;
    ba.group_sims = d_group_sims;
    ba.afrag = d_afrag;
    ba.n_groups = n_groups;
    ba.KT = batched_kt(a.nodes);
    int tpw, nt_pad;
    batched_plan(a.nodes, want_tpw, &tpw, &nt_pad);
    ba.NT = nt_pad;
    *launches = 1;
    switch (model) {
        case MODEL_RWW: return dispatch<RWW>(ba, consts, tpw, alds, s);
        case MODEL_JR: return dispatch<JR>(ba, consts, tpw, alds, s);
        case MODEL_KURAMOTO: return dispatch<Kuramoto>(ba, consts, tpw, alds, s);
        case MODEL_WC: return dispatch<WC>(ba, consts, tpw, alds, s);
        case MODEL_RWWEX: return dispatch<RWWEx>(ba, consts, tpw, alds, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cubnm
