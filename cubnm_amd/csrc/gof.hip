// Goodness-of-fit reductions on the GPU, FP64.
//
// fc_corr: Pearson r between each simulated FC tril and the empirical FC
//   tril (reference: sim/base.py:1176-1189 with scipy.stats.pearsonr on the
//   CPU, utils.py:416-507 fc_corr_device on the GPU).  Two-pass: means, then
//   centred cross and square sums.
// fcd_ks: two-sample Kolmogorov-Smirnov statistic D = max_x |F_sim(x) -
//   F_emp(x)| over the pooled sample points, with right-continuous ECDFs
//   (reference: scipy.stats.ks_2samp(...).statistic, and utils.py:306-414
//   fcd_ks_device with its right-searchsorted cdf_2d_2d_right kernel).
//   Each workgroup sorts one simulation's FCD values in LDS (bitonic), then
//   every pooled point evaluates both ECDFs by binary search; the empirical
//   values arrive sorted from the host.  NaN anywhere in a row gives NaN.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace cubnm {

namespace {
constexpr int GOF_THREADS = 256;

__device__ double wg_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
    return t;
}

__device__ double wg_max(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) t = fmax(t, red[w]);
    return t;
}

// number of elements <= x in sorted a[0..n)
__device__ int upper_bound(const double* a, int n, double x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
}  // namespace

__global__ void __launch_bounds__(GOF_THREADS) fc_corr_kernel(const double* sim, int n_sims, int n, const double* emp,
                                                              double* out) {
    __shared__ double red[GOF_THREADS / 64];
    double e_acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) e_acc += emp[i];
    const double emp_mean = wg_sum(e_acc, red) / n;
    double e_ss = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { const double d = emp[i] - emp_mean; e_ss += d * d; }
    const double emp_ss = wg_sum(e_ss, red);
    for (int s = blockIdx.x; s < n_sims; s += gridDim.x) {
        const double* x = sim + (long long)s * n;
        double acc = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) acc += x[i];
        const double mx = wg_sum(acc, red) / n;
        double sxy = 0.0, sxx = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double dx = x[i] - mx, dy = emp[i] - emp_mean;
            sxy += dx * dy; sxx += dx * dx;
        }
        sxy = wg_sum(sxy, red);
        sxx = wg_sum(sxx, red);
        if (threadIdx.x == 0) {
            double r = sxy / sqrt(sxx * emp_ss);
            if (!isnan(r)) r = fmax(-1.0, fmin(1.0, r));
            out[s] = r;
        }
    }
}

__global__ void __launch_bounds__(GOF_THREADS) fcd_ks_kernel(const double* sim, int n_sims, int n, int n_pow2,
                                                             const double* emp_sorted, int m, double* out) {
    extern __shared__ double v[];  // [n_pow2]
    __shared__ double red[GOF_THREADS / 64];
    for (int s = blockIdx.x; s < n_sims; s += gridDim.x) {
        const double* x = sim + (long long)s * n;
        double nan_seen = 0.0;
        for (int i = threadIdx.x; i < n_pow2; i += blockDim.x) {
            double val = i < n ? x[i] : INFINITY;
            if (isnan(val)) { nan_seen = 1.0; val = INFINITY; }
            v[i] = val;
        }
        nan_seen = wg_max(nan_seen, red);
        // bitonic sort, ascending
        for (int k = 2; k <= n_pow2; k <<= 1) {
            for (int jb = k >> 1; jb > 0; jb >>= 1) {
                __syncthreads();
                for (int i = threadIdx.x; i < n_pow2; i += blockDim.x) {
                    const int p = i ^ jb;
                    if (p > i) {
                        const double a = v[i], b = v[p];
                        const bool up = (i & k) == 0;
                        if ((a > b) == up) { v[i] = b; v[p] = a; }
                    }
                }
            }
        }
        __syncthreads();
        // |F_sim - F_emp| = |c_s m - c_e n| / (n m): the integer numerator is
        // exact, so equal ECDFs give exactly 0 and the one division rounds once
        long long dmax = 0;
        for (int i = threadIdx.x; i < n + m; i += blockDim.x) {
            const double z = i < n ? v[i] : emp_sorted[i - n];
            const long long cs = upper_bound(v, n, z), ce = upper_bound(emp_sorted, m, z);
            const long long dd = cs * m - ce * n;
            dmax = max(dmax, dd < 0 ? -dd : dd);
        }
        const double d = wg_max((double)dmax, red);   // integers below 2^53: exact in double
        if (threadIdx.x == 0) out[s] = nan_seen > 0.0 ? NAN : d / ((double)n * (double)m);
        __syncthreads();
    }
}

}  // namespace cubnm

namespace cubnm { void set_last_error(const std::string& e); }

namespace {
thread_local std::string g_gof_err;

template <class F>
int guarded(F f) {
    hipError_t e = f();
    if (e != hipSuccess) { g_gof_err = hipGetErrorString(e); cubnm::set_last_error(g_gof_err); return -1; }
    return 0;
}
}  // namespace

namespace {
template <class F>
int with_device_copies(const double* sim, size_t sim_n, const double* emp, size_t emp_n, double* out, size_t out_n, F f) {
    double *d_sim = nullptr, *d_emp = nullptr, *d_out = nullptr;
    int rc = -1;
    if (hipMalloc(&d_sim, sizeof(double) * sim_n) == hipSuccess && hipMalloc(&d_emp, sizeof(double) * emp_n) == hipSuccess &&
        hipMalloc(&d_out, sizeof(double) * out_n) == hipSuccess &&
        hipMemcpy(d_sim, sim, sizeof(double) * sim_n, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(d_emp, emp, sizeof(double) * emp_n, hipMemcpyHostToDevice) == hipSuccess) {
        rc = f(d_sim, d_emp, d_out);
        if (rc == 0 && hipMemcpy(out, d_out, sizeof(double) * out_n, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    } else {
        g_gof_err = "device allocation or copy failed";
        cubnm::set_last_error(g_gof_err);
    }
    if (d_sim) hipFree(d_sim);
    if (d_emp) hipFree(d_emp);
    if (d_out) hipFree(d_out);
    return rc;
}
}  // namespace

extern "C" {

int cubnm_gof_fc_corr_device(const double* d_sim, int n_sims, int n_pairs, const double* d_emp, double* d_out);
int cubnm_gof_fcd_ks_device(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m, double* d_out);

const char* cubnm_gof_last_error(void) { return g_gof_err.c_str(); }

int cubnm_gof_fc_corr_device(const double* d_sim, int n_sims, int n_pairs, const double* d_emp, double* d_out) {
    if (n_sims <= 0 || n_pairs <= 1) { g_gof_err = "need n_sims > 0 and n_pairs > 1"; return -1; }
    return guarded([&]() -> hipError_t {
        hipLaunchKernelGGL(cubnm::fc_corr_kernel, dim3(std::min(n_sims, 4096)), dim3(cubnm::GOF_THREADS), 0, 0,
                           d_sim, n_sims, n_pairs, d_emp, d_out);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? hipDeviceSynchronize() : e;
    });
}

int cubnm_gof_fcd_ks_device(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m, double* d_out) {
    if (n_sims <= 0 || n <= 0 || m <= 0) { g_gof_err = "need n_sims, n and m > 0"; return -1; }
    int n_pow2 = 1;
    while (n_pow2 < n) n_pow2 <<= 1;
    if ((size_t)n_pow2 * sizeof(double) > 128 * 1024) { g_gof_err = "too many FCD values per simulation for the LDS sort"; return -1; }
    return guarded([&]() -> hipError_t {
        hipLaunchKernelGGL(cubnm::fcd_ks_kernel, dim3(std::min(n_sims, 4096)), dim3(cubnm::GOF_THREADS),
                           sizeof(double) * n_pow2, 0, d_sim, n_sims, n, n_pow2, d_emp_sorted, m, d_out);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? hipDeviceSynchronize() : e;
    });
}


int cubnm_gof_fc_corr(const double* sim_fc, int n_sims, int n_pairs, const double* emp_fc, double* out) {
    if (n_sims <= 0 || n_pairs <= 1) { g_gof_err = "need n_sims > 0 and n_pairs > 1"; return -1; }
    return with_device_copies(sim_fc, (size_t)n_sims * n_pairs, emp_fc, n_pairs, out, n_sims,
                              [&](double* ds, double* de, double* dout) {
                                  return cubnm_gof_fc_corr_device(ds, n_sims, n_pairs, de, dout);
                              });
}

int cubnm_gof_fcd_ks(const double* sim_fcd, int n_sims, int n, const double* emp_fcd, int m, double* out) {
    if (n_sims <= 0 || n <= 0 || m <= 0) { g_gof_err = "need n_sims, n and m > 0"; return -1; }
    std::vector<double> emp(emp_fcd, emp_fcd + m);
    for (double x : emp) if (std::isnan(x)) { for (int i = 0; i < n_sims; i++) out[i] = NAN; return 0; }
    std::sort(emp.begin(), emp.end());
    return with_device_copies(sim_fcd, (size_t)n_sims * n, emp.data(), m, out, n_sims,
                              [&](double* ds, double* de, double* dout) {
                                  return cubnm_gof_fcd_ks_device(ds, n_sims, n, de, m, dout);
                              });
}

}  // extern "C"
