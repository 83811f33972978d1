// FC and FCD from simulated BOLD, FP64.
//
// Semantics follow the reference (src/ext/fc.cu:4-240 on the GPU,
// src/ext/fc.cpp:1-91 on the CPU):
//   FC[p]   = Pearson(BOLD[nvr:, i], BOLD[nvr:, j]) for pair p=(i>j) in
//             row-major lower-triangle order, optionally skipping
//             inter-hemispheric pairs;
//   wFC_w   = the same over each window's rows [start_w, end_w];
//   FCD[q]  = Pearson over pairs of wFC_wi and wFC_wj for q=(wi>wj).
// Each workgroup handles one simulation at a time (grid-stride over sims):
// the column means and root-sum-of-squares are computed once per row range,
// the columns are z-scored into LDS so that every correlation is a plain
// dot product, window FCs go to a per-workgroup scratch slice and the FCD
// pass runs two-pass Pearson over that slice.
#include "engine.hpp"

namespace cubnm {

namespace {
constexpr int FC_THREADS = 256;

__device__ double block_sum(double v, double* red) {
    // wave64 butterfly, then across waves through LDS
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    return t;
}

// z-score columns of rows [r0, r1) into z[(r - r0) * N + i]; returns nothing.
__device__ void zscore_rows(const double* bold, int N, int r0, int r1, double* z) {
    const int len = r1 - r0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double m = 0.0;
        for (int r = r0; r < r1; r++) m += bold[(long long)r * N + i];
        m /= len;
        double ss = 0.0;
        for (int r = r0; r < r1; r++) { const double d = bold[(long long)r * N + i] - m; ss += d * d; }
        const double inv = 1.0 / sqrt(ss);
        for (int r = r0; r < r1; r++) z[(r - r0) * N + i] = (bold[(long long)r * N + i] - m) * inv;
    }
}

__device__ void pair_corrs(const double* z, int N, int len, int n_pairs, const int* pi, const int* pj, double* out) {
    for (int p = threadIdx.x; p < n_pairs; p += blockDim.x) {
        const int i = pi[p], j = pj[p];
        double acc = 0.0;
        for (int r = 0; r < len; r++) acc += z[r * N + i] * z[r * N + j];
        out[p] = acc;
    }
}
}  // namespace

__global__ void __launch_bounds__(FC_THREADS) fc_fcd_kernel(FcArgs a) {
    extern __shared__ double z[];   // [max rows][N]
    __shared__ double red[FC_THREADS / 64];
    __shared__ double wstat[2 * 1024];  // per-window mean and 1/ssd (n_windows <= 1024)
    const int N = a.nodes;
    double* wfc = a.wfc_scratch + (long long)blockIdx.x * a.n_windows * a.n_pairs;
    for (int sim = blockIdx.x; sim < a.N_SIMS; sim += gridDim.x) {
        const double* bold = a.bold + (long long)sim * a.bold_len * N;
        __syncthreads();
        zscore_rows(bold, N, a.n_vols_remove, a.bold_len, z);
        __syncthreads();
        pair_corrs(z, N, a.bold_len - a.n_vols_remove, a.n_pairs, a.pairs_i, a.pairs_j,
                   a.fc + (long long)sim * a.n_pairs);
        if (a.n_windows == 0) continue;
        for (int w = 0; w < a.n_windows; w++) {
            const int r0 = a.w_starts[w], r1 = a.w_ends[w] + 1;
            __syncthreads();
            zscore_rows(bold, N, r0, r1, z);
            __syncthreads();
            pair_corrs(z, N, r1 - r0, a.n_pairs, a.pairs_i, a.pairs_j, wfc + (long long)w * a.n_pairs);
        }
        __syncthreads();
        // per-window mean and root-sum-of-squares over pairs
        for (int w = 0; w < a.n_windows; w++) {
            const double* v = wfc + (long long)w * a.n_pairs;
            double s = 0.0;
            for (int p = threadIdx.x; p < a.n_pairs; p += blockDim.x) s += v[p];
            const double m = block_sum(s, red) / a.n_pairs;
            double ss = 0.0;
            for (int p = threadIdx.x; p < a.n_pairs; p += blockDim.x) { const double d = v[p] - m; ss += d * d; }
            ss = block_sum(ss, red);
            if (threadIdx.x == 0) { wstat[2 * w] = m; wstat[2 * w + 1] = 1.0 / sqrt(ss); }
        }
        __syncthreads();
        double* fcd = a.fcd + (long long)sim * a.n_window_pairs;
        for (int q = 0; q < a.n_window_pairs; q++) {
            const int wi = a.wpairs_i[q], wj = a.wpairs_j[q];
            const double* x = wfc + (long long)wi * a.n_pairs;
            const double* y = wfc + (long long)wj * a.n_pairs;
            const double mx = wstat[2 * wi], my = wstat[2 * wj];
            double acc = 0.0;
            for (int p = threadIdx.x; p < a.n_pairs; p += blockDim.x) acc += (x[p] - mx) * (y[p] - my);
            acc = block_sum(acc, red);
            if (threadIdx.x == 0) fcd[q] = acc * wstat[2 * wi + 1] * wstat[2 * wj + 1];
        }
    }
}

bool fc_fcd_fits_lds(const FcArgs& a) {
    return a.n_windows <= 1024 && sizeof(double) * (size_t)a.bold_len * a.nodes <= 150 * 1024;
}

hipError_t launch_fc_fcd(const FcArgs& a, hipStream_t s) {
    if (!fc_fcd_fits_lds(a)) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)a.bold_len * a.nodes;
    hipLaunchKernelGGL(fc_fcd_kernel, dim3(a.n_slots), dim3(FC_THREADS), lds, s, a);
    return hipGetLastError();
}

}  // namespace cubnm
