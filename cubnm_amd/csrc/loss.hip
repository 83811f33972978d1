// The regional-fit loss of simulated BOLD and its hand-written adjoint, FP64
// (no counterpart in the reference, which scores with FC correlation and
// FCD KS after the fact, sim/base.py:1084-1207; this is the differentiable
// objective of gradient fitting):
//
//   L = (1 - r(FC_sim, FC_emp)) + (1/nq) sum_q |Q_q(FCD_sim) - E_q|
//
// * FC_sim: Pearson r of every node pair (i > j, row-major) over the volumes
//   after the burn-in; r(.,.) the Pearson correlation of the two vectors.
// * FCD_sim: Pearson r between the window-FC vectors of every window pair
//   (k > l), each window's FC over its volumes [start_k, end_k].
// * Q_q: quantile at position q_q (n-1) of the sorted FCD values with linear
//   interpolation (the definition of torch.quantile's default, with its
//   lerp); E_q the empirical FCD quantiles.
//
// One workgroup per simulation.  Forward: z-scored columns, FC, window FCs,
// FCD, a stable rank sort, quantiles.  Backward (when dbold is given):
//   dL/dQ_q = sign(Q_q - E_q)/nq, shared by the two sorted values the lerp
//     reads (1 - w, w), then sent back to the FCD entries through the ranks;
//   FCD_kl = u_k . u_l with u = centred, normalised window FC: dL/du_k =
//     sum_l H_kl u_l, dL/dw_k = (dL/du_k - u_k (u_k . dL/du_k)) / sigma_k;
//   a Pearson r_ij of columns x_i, x_j over rows T: dr_ij/dx_ti =
//     (z_tj - r_ij z_ti) / s_i with z the centred columns over their norms s,
//     so dL/dx_ti = (sum_j G_ij z_tj - z_ti sum_j G_ij r_ij) / s_i for the
//     symmetric G = dL/dr (static FC: G from the FC correlation term,
//     windows: G = dL/dw_k);
//   dL/dFC from the correlation with the empirical FC: -(b/(|a||b|) -
//     r a/|a|^2) with a, b the centred vectors.
// A NaN anywhere in a simulation's FC or FCD gives a NaN loss and a zero
// gradient for that simulation.
#include "engine.hpp"
#include <string>

namespace cubnm {

void set_last_error(const std::string& e);

namespace {

constexpr int LT = 256;

struct LossArgs {
    const double* bold;         // [S][T][N]
    int n_sims, T, N, nvr;
    const int *w_start, *w_end;  // [W], inclusive volume bounds
    int W, Lmax;
    const double *emp_fc, *emp_q, *qs;  // [P], [nq], [nq]
    int nq;
    double* loss;    // [S]
    double* dbold;   // [S][T][N] or null
    double* scratch; // [S][stride]
    long long stride;
};

__device__ inline long long pair_index(int i, int j) {  // i != j
    const int a = i > j ? i : j, b = i > j ? j : i;
    return (long long)a * (a - 1) / 2 + b;
}

__device__ inline void pair_of(long long p, int* i, int* j) {
    int a = (int)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
    while ((long long)a * (a - 1) / 2 > p) a--;
    while ((long long)(a + 1) * a / 2 <= p) a++;
    *i = a;
    *j = (int)(p - (long long)a * (a - 1) / 2);
}

__device__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < LT / 64; w++) t += red[w];
    return t;
}

// torch's lerp: start + w (end - start) for |w| < 0.5, else end - (end - start)(1 - w)
__device__ inline double lerp_t(double a, double b, double w) {
    const double d = b - a;
    return fabs(w) < 0.5 ? a + w * d : b - d * (1.0 - w);
}

// z-scores columns of rows [r0, r1] of x (T x N) into z ((r1-r0+1) x N) and
// their norms into s (centred column / s; r over these rows = z^T z)
__device__ void zscore(const double* x, int N, int r0, int r1, double* z, double* s) {
    const int L = r1 - r0 + 1;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double m = 0.0;
        for (int t = r0; t <= r1; t++) m += x[(long long)t * N + i];
        m /= L;
        double ss = 0.0;
        for (int t = r0; t <= r1; t++) { const double d = x[(long long)t * N + i] - m; ss += d * d; }
        const double sd = sqrt(ss);
        s[i] = sd;
        for (int t = r0; t <= r1; t++) z[(long long)(t - r0) * N + i] = (x[(long long)t * N + i] - m) / sd;
    }
}

// r_p = z_i . z_j over L rows for every pair p
__device__ void pair_dots(const double* z, int L, int N, long long P, double* out) {
    for (long long p = threadIdx.x; p < P; p += blockDim.x) {
        int i, j;
        pair_of(p, &i, &j);
        double acc = 0.0;
        for (int t = 0; t < L; t++) acc += z[(long long)t * N + i] * z[(long long)t * N + j];
        out[p] = acc;
    }
}

// dx[r0 + t][i] += (sum_j G_ij z_tj - z_ti c_i) / s_i, c_i = sum_j G_ij r_ij,
// for G given as pair values g (symmetric, zero diagonal)
__device__ void pearson_backward(const double* z, const double* s, int L, int N, const double* g, const double* r,
                                 double* c, double* dx, int r0) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < N; j++)
            if (j != i) { const long long p = pair_index(i, j); acc += g[p] * r[p]; }
        c[i] = acc;
    }
    __syncthreads();
    for (long long e = threadIdx.x; e < (long long)L * N; e += blockDim.x) {
        const int t = (int)(e / N), i = (int)(e % N);
        double acc = 0.0;
        for (int j = 0; j < N; j++)
            if (j != i) acc += g[pair_index(i, j)] * z[(long long)t * N + j];
        dx[(long long)(r0 + t) * N + i] += (acc - z[(long long)t * N + i] * c[i]) / s[i];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(LT) fc_fcd_loss_kernel(LossArgs a) {
    __shared__ double red[LT / 64];
    __shared__ int bad;
    const int sim = blockIdx.x, N = a.N, T = a.T, W = a.W;
    const int Tp = T - a.nvr;
    const long long P = (long long)N * (N - 1) / 2;
    const long long Q = (long long)W * (W - 1) / 2;
    const double* x = a.bold + (long long)sim * T * N;
    double* sc = a.scratch + (long long)sim * a.stride;
    double* zs = sc;                          // Tp*N
    double* zw = zs + (long long)Tp * N;      // Lmax*N
    double* fc = zw + (long long)a.Lmax * N;  // P
    double* gfc = fc + P;                     // P
    double* wfc = gfc + P;                    // W*P
    double* dwfc = wfc + W * P;               // W*P
    double* wmean = dwfc + W * P;             // W
    double* wsig = wmean + W;                 // W
    double* fcd = wsig + W;                   // Q
    double* rnk = fcd + Q;                    // Q
    double* srt = rnk + Q;                    // Q
    double* gsrt = srt + Q;                   // Q
    double* H = gsrt + Q;                     // Q
    double* sst = H + Q;                      // N
    double* cst = sst + N;                    // N
    double* sw = cst + N;                     // N
    double* cw = sw + N;                      // N
    if (threadIdx.x == 0) bad = 0;

    // static FC and its correlation with the empirical FC
    zscore(x, N, a.nvr, T - 1, zs, sst);
    __syncthreads();
    pair_dots(zs, Tp, N, P, fc);
    __syncthreads();
    double ps = 0.0, pe = 0.0;
    for (long long p = threadIdx.x; p < P; p += blockDim.x) { ps += fc[p]; pe += a.emp_fc[p]; }
    const double mf = block_sum(ps, red) / P, me = block_sum(pe, red) / P;
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    for (long long p = threadIdx.x; p < P; p += blockDim.x) {
        const double da = fc[p] - mf, db = a.emp_fc[p] - me;
        sab += da * db; saa += da * da; sbb += db * db;
    }
    const double Sab = block_sum(sab, red), Saa = block_sum(saa, red), Sbb = block_sum(sbb, red);
    const double r_fc = Sab / sqrt(Saa * Sbb);

    // window FCs, centred and normalised
    for (int k = 0; k < W; k++) {
        const int r0 = a.w_start[k], r1 = a.w_end[k];
        zscore(x, N, r0, r1, zw, sw);
        __syncthreads();
        pair_dots(zw, r1 - r0 + 1, N, P, wfc + (long long)k * P);
        __syncthreads();
        double s = 0.0;
        for (long long p = threadIdx.x; p < P; p += blockDim.x) s += wfc[(long long)k * P + p];
        const double m = block_sum(s, red) / P;
        double ss = 0.0;
        for (long long p = threadIdx.x; p < P; p += blockDim.x) {
            const double d = wfc[(long long)k * P + p] - m;
            ss += d * d;
        }
        const double sg = sqrt(block_sum(ss, red));
        if (threadIdx.x == 0) { wmean[k] = m; wsig[k] = sg; }
        __syncthreads();
    }
    // FCD lower triangle, (k, l) with k > l row-major
    for (long long q = threadIdx.x; q < Q; q += blockDim.x) {
        int k, l;
        pair_of(q, &k, &l);
        const double *wk = wfc + (long long)k * P, *wl = wfc + (long long)l * P;
        const double mk = wmean[k], ml = wmean[l];
        double acc = 0.0;
        for (long long p = 0; p < P; p++) acc += (wk[p] - mk) * (wl[p] - ml);
        fcd[q] = acc / (wsig[k] * wsig[l]);
    }
    __syncthreads();
    if (isnan(r_fc) && threadIdx.x == 0) bad = 1;
    for (long long q = threadIdx.x; q < Q; q += blockDim.x)
        if (isnan(fcd[q])) bad = 1;
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) a.loss[sim] = NAN;
        if (a.dbold)
            for (long long e = threadIdx.x; e < (long long)T * N; e += blockDim.x) a.dbold[(long long)sim * T * N + e] = 0.0;
        return;
    }
    // stable ranks and the sorted values
    for (long long q = threadIdx.x; q < Q; q += blockDim.x) {
        const double v = fcd[q];
        long long r = 0;
        for (long long o = 0; o < Q; o++) {
            const double u = fcd[o];
            r += (u < v) || (u == v && o < q);
        }
        rnk[q] = (double)r;
        srt[r] = v;
    }
    __syncthreads();
    double lq = 0.0;
    for (int i = threadIdx.x; i < a.nq; i += blockDim.x) {
        const double h = a.qs[i] * (double)(Q - 1);
        const double lo = floor(h), hi = ceil(h);
        lq += fabs(lerp_t(srt[(long long)lo], srt[(long long)hi], h - lo) - a.emp_q[i]);
    }
    lq = block_sum(lq, red) / a.nq;
    if (threadIdx.x == 0) a.loss[sim] = (1.0 - r_fc) + lq;
    if (!a.dbold) return;

    // ---- backward
    for (long long q = threadIdx.x; q < Q; q += blockDim.x) gsrt[q] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < a.nq; i++) {
            const double h = a.qs[i] * (double)(Q - 1);
            const double lo = floor(h), hi = ceil(h), w = h - lo;
            const double d = lerp_t(srt[(long long)lo], srt[(long long)hi], w) - a.emp_q[i];
            const double g = (d > 0.0 ? 1.0 : (d < 0.0 ? -1.0 : 0.0)) / a.nq;
            gsrt[(long long)lo] += g * (1.0 - w);
            gsrt[(long long)hi] += g * w;
        }
    __syncthreads();
    for (long long q = threadIdx.x; q < Q; q += blockDim.x) H[q] = gsrt[(long long)rnk[q]];
    __syncthreads();
    // dL/dw_k for every window: dU_k = sum_l H_kl u_l, then remove the u_k
    // component and divide by sigma_k (dwfc first holds dU)
    for (long long e = threadIdx.x; e < (long long)W * P; e += blockDim.x) {
        const int k = (int)(e / P);
        const long long p = e % P;
        double acc = 0.0;
        for (int l = 0; l < W; l++)
            if (l != k) acc += H[pair_index(k, l)] * (wfc[(long long)l * P + p] - wmean[l]) / wsig[l];
        dwfc[e] = acc;
    }
    __syncthreads();
    for (int k = 0; k < W; k++) {
        double s = 0.0;
        for (long long p = threadIdx.x; p < P; p += blockDim.x)
            s += (wfc[(long long)k * P + p] - wmean[k]) / wsig[k] * dwfc[(long long)k * P + p];
        const double ck = block_sum(s, red);
        for (long long p = threadIdx.x; p < P; p += blockDim.x) {
            const double u = (wfc[(long long)k * P + p] - wmean[k]) / wsig[k];
            dwfc[(long long)k * P + p] = (dwfc[(long long)k * P + p] - u * ck) / wsig[k];
        }
        __syncthreads();
    }
    // dL/dFC (static): minus the gradient of the correlation with the empirical FC
    const double nab = sqrt(Saa * Sbb);
    for (long long p = threadIdx.x; p < P; p += blockDim.x) {
        const double da = fc[p] - mf, db = a.emp_fc[p] - me;
        gfc[p] = -(db / nab - r_fc * da / Saa);
    }
    double* dx = a.dbold + (long long)sim * T * N;
    for (long long e = threadIdx.x; e < (long long)T * N; e += blockDim.x) dx[e] = 0.0;
    __syncthreads();
    pearson_backward(zs, sst, Tp, N, gfc, fc, cst, dx, a.nvr);
    for (int k = 0; k < W; k++) {
        const int r0 = a.w_start[k], r1 = a.w_end[k];
        zscore(x, N, r0, r1, zw, sw);
        __syncthreads();
        pearson_backward(zw, sw, r1 - r0 + 1, N, dwfc + (long long)k * P, wfc + (long long)k * P, cw, dx, r0);
    }
}

}  // namespace

}  // namespace cubnm

using namespace cubnm;

extern "C" {

long long cubnm_fc_fcd_loss_scratch(int bold_len, int nodes, int n_vols_remove, int n_windows, int max_window_len) {
    const long long N = nodes, P = N * (N - 1) / 2, W = n_windows, Q = W * (W - 1) / 2;
    return (long long)(bold_len - n_vols_remove) * N + (long long)max_window_len * N + 2 * P + 2 * W * P + 2 * W +
           5 * Q + 4 * N;
}

int cubnm_fc_fcd_loss(const double* d_bold, int n_sims, int bold_len, int nodes, int n_vols_remove,
                      const int* d_win_start, const int* d_win_end, int n_windows, int max_window_len,
                      const double* d_emp_fc, const double* d_emp_q, const double* d_qs, int n_quantiles,
                      double* d_loss, double* d_dbold, double* d_scratch, void* stream) {
    if (n_sims < 1 || nodes < 3 || bold_len - n_vols_remove < 2 || n_windows < 2 || n_quantiles < 1 ||
        max_window_len < 2) {
        set_last_error("cubnm_fc_fcd_loss: need n_sims >= 1, nodes >= 3, >= 2 volumes after burn-in, >= 2 windows "
                       "of >= 2 volumes and >= 1 quantile");
        return -1;
    }
    LossArgs a{d_bold, n_sims, bold_len, nodes, n_vols_remove, d_win_start, d_win_end, n_windows, max_window_len,
               d_emp_fc, d_emp_q, d_qs, n_quantiles, d_loss, d_dbold, d_scratch,
               cubnm_fc_fcd_loss_scratch(bold_len, nodes, n_vols_remove, n_windows, max_window_len)};
    hipLaunchKernelGGL(fc_fcd_loss_kernel, dim3(n_sims), dim3(LT), 0, static_cast<hipStream_t>(stream), a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { set_last_error(std::string("cubnm_fc_fcd_loss: ") + hipGetErrorString(e)); return -1; }
    return 0;
}

}  // extern "C"
