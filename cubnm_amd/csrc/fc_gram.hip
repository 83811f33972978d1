// FC and FCD as correlation GEMMs (Gram products) on the matrix cores.
//
// Same results as the reference's FC/FCD (src/ext/fc.cu:4-240,
// src/ext/fc.cpp:1-91): Pearson r of every node pair over the post-burn-in
// volumes, of every node pair inside every window, and of every pair of
// window-FC vectors.  Written as Gram products so the work lands on MFMA:
//
//   1. zscore_kernel   per simulation and row range (the FC range and each
//      window), every node's column is centred and scaled to unit norm in
//      FP64 and stored node-major, X[node][k], zero padded to a multiple of
//      the GEMM K step;
//   2. gram            X X^T lower triangle = Pearson r of every pair,
//      scattered to the tril order (row-major, i > j; optional
//      inter-hemispheric exclusion through a pair map);
//   3. center_kernel   every window-FC vector is centred and scaled to unit
//      norm over its pairs in FP64 (two-pass);
//   4. gram            U U^T lower triangle over windows = FCD.
//
// Two arithmetic forms of the Gram product:
//   * FP64: v_mfma_f64_16x16x4f64 with FP64 accumulation (default; parity
//     with the FP64 oracle at the 1e-13 level);
//   * split-BF16 (for the static FC and the FCD, the large-K products):
//     each FP64 operand x is split into hi = bf16(x) and lo = bf16(x - hi);
//     x.y ~ hi.hi + hi.lo + lo.hi on v_mfma_f32_16x16x32_bf16 with FP32
//     accumulation, flushed to FP64 every 512 K.  The normalisations stay
//     FP64, and so do the window FCs (K = one window, a small share of the
//     work): the FCD is a Pearson r between window-FC vectors whose spread
//     across pairs can be small, so their errors are amplified by
//     1 / (their standard deviation) and must stay at FP64 level.  Operand
//     error is about 2^-16 relative; measured errors are stated per result
//     by the tests.
//
// Tiles: 64 x 64 outputs per workgroup of 4 waves (each 32 x 32 = 2 x 2
// MFMA blocks); only tiles on or below the diagonal are launched.  The
// split-BF16 FCD product (long K, few tiles) is further split over K slices
// whose partial sums are added in slice order.
#include "engine.hpp"
#include <algorithm>
#include <cmath>
#include <vector>

namespace cubnm {

namespace {
constexpr int GT = 64;        // output tile edge
constexpr int KC64 = 16;      // K chunk staged per step, FP64 form
constexpr int KCB = 32;       // K chunk staged per step, split-BF16 form
constexpr int FLUSH = 16;     // BF16 chunks accumulated in FP32 before an FP64 flush
constexpr int ZT = 256;       // threads of the z-score / centring kernels
constexpr int MAX_SPLIT = 16; // K slices of the split-BF16 FCD product

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short bf16_bits(float x) {
    // round to nearest even; operands are finite (z-scores, correlations)
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf16_float(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

__device__ inline void split_store(double v, unsigned short* hi, unsigned short* lo, long long idx) {
    const float f = (float)v;
    const unsigned short h = bf16_bits(f);
    hi[idx] = h;
    lo[idx] = bf16_bits((float)(v - (double)bf16_float(h)));
}

__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    return t;
}

__device__ inline void tri_tile(int t, int* ti, int* tj) {
    int i = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) i++;
    while (i * (i + 1) / 2 > t) i--;
    *ti = i;
    *tj = t - i * (i + 1) / 2;
}

__device__ inline long long pair_index(int i, int j, const int* pair_of, int rows) {
    return pair_of ? (long long)pair_of[(long long)i * rows + j] : (long long)i * (i - 1) / 2 + j;
}
}  // namespace

// ---------------------------------------------------------------- 1. z-scores
// blockIdx.x = simulation within the batch, blockIdx.y = row range (0 = FC,
// 1 + w = window w).  Output row stride is k_fc or k_w.
template <bool FC_BF16>
__global__ void __launch_bounds__(ZT) zscore_kernel(const double* bold, int bold_len, int N, int r0_fc,
                                                    const int* w_starts, const int* w_ends, int n_windows,
                                                    int k_fc, int k_w, double* x_fc, double* x_w,
                                                    unsigned short* h_fc, unsigned short* l_fc) {
    const int b = blockIdx.x, m = blockIdx.y;
    const double* B = bold + (long long)b * bold_len * N;
    int r0, r1, K;
    long long base;
    if (m == 0) { r0 = r0_fc; r1 = bold_len; K = k_fc; base = (long long)b * N * k_fc; }
    else {
        r0 = w_starts[m - 1]; r1 = w_ends[m - 1] + 1; K = k_w;
        base = ((long long)b * n_windows + (m - 1)) * N * k_w;
    }
    const int len = r1 - r0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double s = 0.0;
        for (int r = r0; r < r1; r++) s += B[(long long)r * N + i];
        const double mean = s / len;
        double ss = 0.0;
        for (int r = r0; r < r1; r++) { const double d = B[(long long)r * N + i] - mean; ss += d * d; }
        const double inv = 1.0 / sqrt(ss);
        const long long row = base + (long long)i * K;
        for (int k = 0; k < K; k++) {
            const double z = k < len ? (B[(long long)(r0 + k) * N + i] - mean) * inv : 0.0;
            if (FC_BF16 && m == 0) split_store(z, h_fc, l_fc, row + k);
            else (m == 0 ? x_fc : x_w)[row + k] = z;
        }
    }
}

// ---------------------------------------------------------------- 3. centring
// One workgroup per window-FC vector: u = (v - mean) / ||v - mean|| over the
// first P entries (the padding stays 0).
template <bool BF16>
__global__ void __launch_bounds__(ZT) center_kernel(double* U, int P, int ld, unsigned short* hi, unsigned short* lo) {
    __shared__ double red[ZT / 64];
    double* v = U + (long long)blockIdx.x * ld;
    double s = 0.0;
    for (int p = threadIdx.x; p < P; p += blockDim.x) s += v[p];
    const double mean = block_sum(s, red) / P;
    double ss = 0.0;
    for (int p = threadIdx.x; p < P; p += blockDim.x) { const double d = v[p] - mean; ss += d * d; }
    const double inv = 1.0 / sqrt(block_sum(ss, red));
    const long long row = (long long)blockIdx.x * ld;
    for (int p = threadIdx.x; p < ld; p += blockDim.x) {
        const double u = p < P ? (v[p] - mean) * inv : 0.0;
        if (BF16) split_store(u, hi, lo, row + p);
        else v[p] = u;
    }
}

// ---------------------------------------------------------------- 2/4. Gram, FP64
// X: matrices of `rows` rows of K (multiple of KC64) doubles, row stride K,
// matrix stride xs.  out: tril of each matrix, matrix stride os.
__global__ void __launch_bounds__(256) gram_f64_kernel(const double* X, long long xs, int rows, int K,
                                                       double* out, long long os, const int* pair_of, int n_tiles) {
    __shared__ double As[GT][KC64 + 1];
    __shared__ double Bs[GT][KC64 + 1];
    const long long mat = blockIdx.x / n_tiles;
    int ti, tj;
    tri_tile((int)(blockIdx.x % n_tiles), &ti, &tj);
    const double* Xm = X + mat * xs;
    const int i0 = ti * GT, j0 = tj * GT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    doublex4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++) acc[a][c] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KC64) {
        __syncthreads();
        for (int q = threadIdx.x; q < GT * KC64; q += 256) {
            const int r = q / KC64, k = q % KC64;
            As[r][k] = i0 + r < rows ? Xm[(long long)(i0 + r) * K + k0 + k] : 0.0;
            Bs[r][k] = j0 + r < rows ? Xm[(long long)(j0 + r) * K + k0 + k] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC64; kk += 4) {
            double a[2], b[2];
            for (int x = 0; x < 2; x++) {
                a[x] = As[wr + 16 * x + li][kk + lk];
                b[x] = Bs[wc + 16 * x + li][kk + lk];
            }
            for (int x = 0; x < 2; x++)
                for (int y = 0; y < 2; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
        }
    }
    double* o = out + mat * os;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int r = 0; r < 4; r++) {
                // f64 16x16 D layout: register r of lane l holds row 4r + l/16, column l%16
                const int i = i0 + wr + 16 * x + 4 * r + lk, j = j0 + wc + 16 * y + li;
                if (i < rows && j < i) {
                    const long long p = pair_index(i, j, pair_of, rows);
                    if (p >= 0) o[p] = acc[x][y][r];
                }
            }
}

// ---------------------------------------------------------------- 2/4. Gram, split-BF16
// Grid: (matrix, tile, K slice), the slice innermost.  With n_split == 1 the
// tile's pairs are written to out; otherwise each slice writes its partial
// sums to part (slice stride ps, matrix stride os, every pair of every
// matrix written by exactly one tile per slice) and reduce_slices adds the
// slices in order.  The next K chunk is loaded into registers while the
// current one is multiplied; the barriers order LDS only, so those loads
// stay in flight across them.
__global__ void __launch_bounds__(256) gram_bf16x3_kernel(const unsigned short* Xh, const unsigned short* Xl,
                                                          long long xs, int rows, int K, double* out, long long os,
                                                          const int* pair_of, int n_tiles, int n_split,
                                                          int k_per_split, double* part, long long ps) {
    // rows padded to 40 halfwords (80 B): 16-byte aligned 8-element reads
    __shared__ __attribute__((aligned(16))) unsigned short Ah[GT][KCB + 8], Al[GT][KCB + 8];
    __shared__ __attribute__((aligned(16))) unsigned short Bh[GT][KCB + 8], Bl[GT][KCB + 8];
    const int split = (int)(blockIdx.x % n_split);
    const long long rest = blockIdx.x / n_split;
    const long long mat = rest / n_tiles;
    int ti, tj;
    tri_tile((int)(rest % n_tiles), &ti, &tj);
    const long long mo = mat * xs;
    const int i0 = ti * GT, j0 = tj * GT;
    const int kb = split * k_per_split, ke = min(K, kb + k_per_split);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    floatx4 f[2][2];
    doublex4 d[2][2];
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++) { f[a][c] = floatx4{0.f, 0.f, 0.f, 0.f}; d[a][c] = doublex4{0.0, 0.0, 0.0, 0.0}; }
    // 64 rows x 32 halfwords = 256 loads of 8 halfwords (16 B) per array
    const int r = threadIdx.x >> 2, c8 = (threadIdx.x & 3) * 8;
    const bool ra = i0 + r < rows, rb = j0 + r < rows;
    const uint4 z = {0u, 0u, 0u, 0u};
    const long long ga0 = mo + (long long)(i0 + r) * K + c8;
    const long long gb0 = mo + (long long)(j0 + r) * K + c8;
    uint4 pah = z, pal = z, pbh = z, pbl = z;
    if (kb < ke) {
        if (ra) { pah = *(const uint4*)(Xh + ga0 + kb); pal = *(const uint4*)(Xl + ga0 + kb); }
        if (rb) { pbh = *(const uint4*)(Xh + gb0 + kb); pbl = *(const uint4*)(Xl + gb0 + kb); }
    }
    int chunk = 0;
    for (int k0 = kb; k0 < ke; k0 += KCB, chunk++) {
        lds_barrier();  // the previous chunk's LDS reads are done
        *(uint4*)&Ah[r][c8] = pah;
        *(uint4*)&Al[r][c8] = pal;
        *(uint4*)&Bh[r][c8] = pbh;
        *(uint4*)&Bl[r][c8] = pbl;
        lds_barrier();
        if (k0 + KCB < ke) {
            const int kn = k0 + KCB;
            if (ra) { pah = *(const uint4*)(Xh + ga0 + kn); pal = *(const uint4*)(Xl + ga0 + kn); }
            if (rb) { pbh = *(const uint4*)(Xh + gb0 + kn); pbl = *(const uint4*)(Xl + gb0 + kn); }
        }
        bf16x8 ah[2], al[2], bh[2], bl[2];
        for (int x = 0; x < 2; x++) {
            // 16x16x32 bf16 operands: lane l holds row (or column) l%16, k = 8*(l/16) .. +7
            ah[x] = *(const bf16x8*)&Ah[wr + 16 * x + li][8 * lk];
            al[x] = *(const bf16x8*)&Al[wr + 16 * x + li][8 * lk];
            bh[x] = *(const bf16x8*)&Bh[wc + 16 * x + li][8 * lk];
            bl[x] = *(const bf16x8*)&Bl[wc + 16 * x + li][8 * lk];
        }
        for (int x = 0; x < 2; x++)
            for (int y = 0; y < 2; y++) {
                f[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], f[x][y], 0, 0, 0);
                f[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], f[x][y], 0, 0, 0);
                f[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], f[x][y], 0, 0, 0);
            }
        if ((chunk + 1) % FLUSH == 0) {
            for (int x = 0; x < 2; x++)
                for (int y = 0; y < 2; y++)
                    for (int q = 0; q < 4; q++) { d[x][y][q] += (double)f[x][y][q]; f[x][y][q] = 0.f; }
        }
    }
    double* o = part ? part + split * ps + mat * os : out + mat * os;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int q = 0; q < 4; q++) {
                // f32 16x16 D layout: register q of lane l holds row 4*(l/16) + q, column l%16
                const int i = i0 + wr + 16 * x + 4 * lk + q, j = j0 + wc + 16 * y + li;
                if (i < rows && j < i) {
                    const long long p = pair_index(i, j, pair_of, rows);
                    if (p >= 0) o[p] = d[x][y][q] + (double)f[x][y][q];
                }
            }
}

// out[i] = sum over slices s (in order) of part[s * ps + i], i < n.
__global__ void __launch_bounds__(256) reduce_slices(const double* part, long long ps, int n_split, long long n,
                                                     double* out) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        double v = part[i];
        for (int q = 1; q < n_split; q++) v += part[q * ps + i];
        out[i] = v;
    }
}

// ---------------------------------------------------------------- host pipeline
namespace {
inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
inline int tri_tiles(int rows) { const int t = (rows + GT - 1) / GT; return t * (t + 1) / 2; }

struct GramLayout {
    int k_fc, k_w, p_ld;
    size_t per_sim_bytes;
};

GramLayout gram_layout(const FcArgs& a, int max_win_len, bool bf16) {
    GramLayout g{};
    const int step = bf16 ? KCB : KC64;
    g.k_fc = round_up(std::max(1, a.bold_len - a.n_vols_remove), step);
    g.k_w = round_up(std::max(1, max_win_len), KC64);       // window FCs are FP64 in both forms
    g.p_ld = round_up(std::max(1, a.n_pairs), step);
    const size_t e = bf16 ? 4 : 8;  // bytes per FC operand element (hi + lo, or one double)
    size_t per = e * (size_t)a.nodes * g.k_fc + 8 * (size_t)a.n_windows * a.nodes * g.k_w;
    per += 8 * (size_t)a.n_windows * g.p_ld;                 // FP64 window-FC vectors
    if (bf16) per += 4 * (size_t)a.n_windows * g.p_ld;       // their hi/lo split
    if (bf16) per += 8 * (size_t)MAX_SPLIT * a.n_window_pairs;  // FCD K-slice partial sums
    g.per_sim_bytes = per;
    return g;
}
}  // namespace

size_t fc_gram_workspace_bytes(const FcArgs& a, int max_win_len, int split_bf16, int batch) {
    return gram_layout(a, max_win_len, split_bf16 != 0).per_sim_bytes * (size_t)std::max(1, batch);
}

hipError_t launch_fc_gram(const FcArgs& a, int max_win_len, int split_bf16, const int* pair_of, void* work,
                          size_t work_bytes, hipStream_t s) {
    const bool bf16 = split_bf16 != 0;
    const GramLayout g = gram_layout(a, max_win_len, bf16);
    const int N = a.nodes, W = a.n_windows;
    const int batch = (int)std::min<size_t>((size_t)a.N_SIMS, work_bytes / std::max<size_t>(1, g.per_sim_bytes));
    if (batch < 1) return hipErrorOutOfMemory;
    if (a.n_windows > 65535) return hipErrorInvalidValue;
    char* w = static_cast<char*>(work);
    const size_t e = bf16 ? 2 : 8;
    // carve: [x_fc | x_w] (FP64, or hi then lo), U (FP64), [U hi | U lo]
    const size_t n_fc = (size_t)batch * N * g.k_fc, n_w = (size_t)batch * W * N * g.k_w;
    const size_t n_u = (size_t)batch * W * g.p_ld;
    double *x_fc = nullptr, *x_w = nullptr;
    unsigned short *h_fc = nullptr, *l_fc = nullptr, *uh = nullptr, *ul = nullptr;
    if (bf16) {
        h_fc = (unsigned short*)w; l_fc = h_fc + n_fc;
        w += 2 * e * n_fc;
    } else {
        x_fc = (double*)w;
        w += e * n_fc;
    }
    x_w = (double*)w;
    w += 8 * n_w;
    double* U = (double*)w;
    w += 8 * n_u;
    double* part = nullptr;
    if (bf16) {
        uh = (unsigned short*)w; ul = uh + n_u;
        w += 4 * n_u;
        part = (double*)w;  // [MAX_SPLIT][batch][n_window_pairs]
    }
    const int mats = 1 + (a.n_window_pairs > 0 ? W : 0);
    if (W > 0 && a.n_window_pairs > 0) {
        hipError_t e0 = hipMemsetAsync(U, 0, 8 * n_u, s);   // pair padding stays 0
        if (e0 != hipSuccess) return e0;
    }
    for (int s0 = 0; s0 < a.N_SIMS; s0 += batch) {
        const int nb = std::min(batch, a.N_SIMS - s0);
        const double* bold = a.bold + (long long)s0 * a.bold_len * N;
        if (bf16)
            hipLaunchKernelGGL(zscore_kernel<true>, dim3(nb, mats), dim3(ZT), 0, s, bold, a.bold_len, N,
                               a.n_vols_remove, a.w_starts, a.w_ends, W, g.k_fc, g.k_w, nullptr, x_w, h_fc, l_fc);
        else
            hipLaunchKernelGGL(zscore_kernel<false>, dim3(nb, mats), dim3(ZT), 0, s, bold, a.bold_len, N,
                               a.n_vols_remove, a.w_starts, a.w_ends, W, g.k_fc, g.k_w, x_fc, x_w,
                               nullptr, nullptr);
        // static FC straight into the output
        double* fc = a.fc + (long long)s0 * a.n_pairs;
        const int tn = tri_tiles(N), tw = tri_tiles(W);
        if (bf16)
            hipLaunchKernelGGL(gram_bf16x3_kernel, dim3(tn * nb), dim3(256), 0, s, h_fc, l_fc,
                               (long long)N * g.k_fc, N, g.k_fc, fc, (long long)a.n_pairs, pair_of, tn, 1, g.k_fc,
                               (double*)nullptr, 0LL);
        else
            hipLaunchKernelGGL(gram_f64_kernel, dim3(tn * nb), dim3(256), 0, s, x_fc,
                               (long long)N * g.k_fc, N, g.k_fc, fc, (long long)a.n_pairs, pair_of, tn);
        if (mats > 1) {
            hipLaunchKernelGGL(gram_f64_kernel, dim3(tn * nb * W), dim3(256), 0, s, x_w,
                               (long long)N * g.k_w, N, g.k_w, U, (long long)g.p_ld, pair_of, tn);
            double* fcd = a.fcd + (long long)s0 * a.n_window_pairs;
            if (bf16) {
                hipLaunchKernelGGL(center_kernel<true>, dim3(nb * W), dim3(ZT), 0, s, U, a.n_pairs, g.p_ld, uh, ul);
                // K slices: about 2048 workgroups (8 per CU) over the batch,
                // each slice at least 32 chunks long
                const int chunks = g.p_ld / KCB;
                int n_split = std::max(1, std::min({MAX_SPLIT, chunks / 32, (2048 + tw * nb - 1) / (tw * nb)}));
                const int kps = ((chunks + n_split - 1) / n_split) * KCB;
                n_split = (g.p_ld + kps - 1) / kps;
                const long long ps = (long long)nb * a.n_window_pairs;
                hipLaunchKernelGGL(gram_bf16x3_kernel, dim3(tw * nb * n_split), dim3(256), 0, s, uh, ul,
                                   (long long)W * g.p_ld, W, g.p_ld, fcd, (long long)a.n_window_pairs,
                                   (const int*)nullptr, tw, n_split, kps, n_split > 1 ? part : (double*)nullptr, ps);
                if (n_split > 1)
                    hipLaunchKernelGGL(reduce_slices, dim3((int)std::min<long long>((ps + 255) / 256, 4096)),
                                       dim3(256), 0, s, part, ps, n_split, ps, fcd);
            } else {
                hipLaunchKernelGGL(center_kernel<false>, dim3(nb * W), dim3(ZT), 0, s, U, a.n_pairs, g.p_ld,
                                   nullptr, nullptr);
                hipLaunchKernelGGL(gram_f64_kernel, dim3(tw * nb), dim3(256), 0, s, U,
                                   (long long)W * g.p_ld, W, g.p_ld, fcd, (long long)a.n_window_pairs,
                                   (const int*)nullptr, tw);
            }
        }
        hipError_t e1 = hipGetLastError();
        if (e1 != hipSuccess) return e1;
    }
    return hipSuccess;
}

}  // namespace cubnm
