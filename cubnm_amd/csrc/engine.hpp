// Launch-argument structures shared by the host session and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "models.hpp"

namespace cubnm {

// Workgroup barrier that orders LDS accesses only.  __syncthreads() is a
// workgroup fence on every address space, which on gfx9 waits for all
// outstanding global loads and stores (vmcnt(0)) before the barrier; this
// form waits only for LDS (lgkmcnt), so global loads issued before it (the
// next step's noise, a prefetch) stay in flight across it.  Use it only where
// the barrier protects LDS data: lanes that exchange data through global
// memory need __syncthreads().
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Everything the integrator kernels read or write for one batch of
// simulations. Layouts follow the reference boundary (src/ext/core.cpp:336-386):
//   gp    [n_global][N_SIMS]            rp    [n_regional][N_SIMS*nodes]
//   SC    [N_SCs][source k][target j]   bold  [N_SIMS][bold_len][nodes]
//   states[n_state][N_SIMS][nodes or states_len*nodes]
//   gbool [n_gout_bool][N_SIMS]         gint  [n_gout_int][N_SIMS]
struct IntegrateArgs {
    int N_SIMS, nodes;
    int bw_it, inner_it, BOLD_TR_iters, states_sampling_iters, n_states_samples_remove;
    int noise_bw_it, states_ts, bold_len, states_len, noise_repeats;
    int do_delay;
    double dt;
    const double* SC;
    const int* SC_indices;
    const double* gp;
    double* rp;
    const double* noise;      // noise segment [noise_bw_it*inner_it][nodes][n_noise]
    const int* sh_nodes;      // [noise_repeats][nodes]
    const int* sh_ts;         // [noise_repeats][noise_bw_it]
    double* bold;
    double* states;
    unsigned char* gbool;
    int* gint;
    // conduction delay (general kernel only)
    const int* delays;        // [N_SIMS][nodes][nodes] (k, j), or null
    const int* max_delays;    // [N_SIMS]
    double* hist;             // circular history, hist_stride doubles per sim
    long long hist_stride;
    BWConsts bw;
    int sc_lds;               // general kernel: SC staged in LDS (rows padded by one)
    int ksplit;               // general kernel: partitions of the coupling sum over k (>= 1)
    int pair;                 // general kernel: coupling lanes sum two adjacent columns (16-byte loads)
};

struct FcArgs {
    int N_SIMS, nodes, bold_len, n_vols_remove, n_pairs, n_windows, n_window_pairs;
    int exc_inter;
    const double* bold;        // [N_SIMS][bold_len][nodes]
    const int* pairs_i;        // [n_pairs]
    const int* pairs_j;
    const int* w_starts;       // [n_windows], inclusive
    const int* w_ends;         // [n_windows], inclusive
    const int* wpairs_i;       // [n_window_pairs]
    const int* wpairs_j;
    double* fc;                // [N_SIMS][n_pairs]
    double* fcd;               // [N_SIMS][n_window_pairs]
    double* wfc_scratch;       // [n_slots][n_windows][n_pairs]
    int n_slots;
};

// Host-side launchers (defined in the .hip files).
// latency_bound: the request's simulations share SCs or are few (pair mode
// allowed; see general_pair_ok)
hipError_t launch_integrate_general(int model, const IntegrateArgs& a, const void* model_consts, hipStream_t s,
                                    bool latency_bound);
// more than 1024 nodes (up to 8192): one 1024-lane workgroup per simulation,
// several nodes per lane; no delay, no numerical-FIC restarts
hipError_t launch_integrate_wide(int model, const IntegrateArgs& a, const void* model_consts, hipStream_t s);
// general kernel geometry: lanes per k partition, partitions (want 0 = auto),
// LDS bytes of a workgroup; several requests of one model in one launch
int general_np(int nodes);
int general_ksplit(int nodes, int want);
size_t general_lds_bytes(int model, int nodes, int np, int ksplit, bool sc_staged, bool pair, int threads);
// pair mode of the general kernel: lanes per partition (two adjacent source
// columns per lane), its eligibility, and the doubles of the partial-sum
// region (the k partitions' or the pair partitions', whichever is larger)
__host__ __device__ inline int general_npc(int nodes) { return 64 * ((nodes / 2 + 63) / 64); }
bool general_pair_ok(const IntegrateArgs& a, int threads, bool latency_bound);
__host__ __device__ inline int general_red_doubles(bool osc, int nodes, int np, int ks, bool pair, int threads) {
    int r = (ks - 1) * 2 * np;
    if (pair) {
        const int pr = (threads / general_npc(nodes)) * nodes * (osc ? 2 : 1);
        r = r > pr ? r : pr;
    }
    return r;
}
size_t model_consts_bytes(int model);
size_t model_consts_stride();
hipError_t launch_integrate_general_multi(const IntegrateArgs* d_args, const void* d_consts, const int* d_models,
                                          const int2* d_blocks, int n_blocks, int threads, size_t lds,
                                          int* progress, hipStream_t s);
hipError_t launch_fc_fcd(const FcArgs& a, hipStream_t s);
// true when the per-simulation LDS kernel can hold a simulation's BOLD
bool fc_fcd_fits_lds(const FcArgs& a);
// FC/FCD as Gram products on MFMA (FP64, or split-BF16 when split_bf16 != 0);
// pair_of: [nodes*nodes] tril index of pair (i > j) or -1, or null for the
// plain tril order; work: device scratch of work_bytes (at least one
// simulation's fc_gram_workspace_bytes).
size_t fc_gram_workspace_bytes(const FcArgs& a, int max_win_len, int split_bf16, int batch);
hipError_t launch_fc_gram(const FcArgs& a, int max_win_len, int split_bf16, const int* pair_of, void* work,
                          size_t work_bytes, hipStream_t s);

}  // namespace cubnm
