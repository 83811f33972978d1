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
//   Any number of values per simulation: each simulation's FCD values are
//   cut into chunks of at most 8,192, each chunk bitonic-sorted in LDS and
//   written to a device workspace; then every pooled point (simulation and
//   empirical values) counts the simulation values <= it by a binary search
//   in each sorted chunk and the empirical values by one in the sorted
//   empirical array, and the largest |c_sim m - c_emp n| over the points is
//   reduced with an integer atomic max.  The empirical values arrive sorted.
//   NaN anywhere in a row gives NaN.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <limits>
#include <mutex>
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

constexpr int KS_CHUNK = 8192;  // values per LDS sort (64 KB)

// Sort chunk blockIdx.x of simulation blockIdx.y (of the batch starting at
// sim0) into ws[s][chunk]; NaN values become +inf and mark the simulation.
__global__ void __launch_bounds__(GOF_THREADS) ks_sort_chunks_kernel(const double* sim, int sim0, int n, int chunk_pow2,
                                                                     double* ws, int* nan_flag) {
    extern __shared__ double v[];  // [chunk_pow2]
    const int s = blockIdx.y, c0 = blockIdx.x * chunk_pow2;
    const int len = min(chunk_pow2, n - c0);
    const double* x = sim + (long long)(sim0 + s) * n + c0;
    bool nan_seen = false;
    for (int i = threadIdx.x; i < chunk_pow2; i += blockDim.x) {
        double val = i < len ? x[i] : INFINITY;
        if (isnan(val)) { nan_seen = true; val = INFINITY; }
        v[i] = val;
    }
    if (nan_seen) atomicOr(nan_flag + s, 1);
    for (int k = 2; k <= chunk_pow2; k <<= 1) {
        for (int jb = k >> 1; jb > 0; jb >>= 1) {
            __syncthreads();
            for (int i = threadIdx.x; i < chunk_pow2; i += blockDim.x) {
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
    double* out = ws + (long long)s * n + c0;
    for (int i = threadIdx.x; i < len; i += blockDim.x) out[i] = v[i];
}

// Pooled points [blockIdx.x * per_block, ...) of simulation blockIdx.y:
// dmax[s] = max |c_sim(z) m - c_emp(z) n| (exact integers).
__global__ void __launch_bounds__(GOF_THREADS) ks_eval_kernel(const double* ws, int n, int chunk_pow2,
                                                              const double* emp_sorted, int m, int per_block,
                                                              unsigned long long* dmax) {
    __shared__ double red[GOF_THREADS / 64];
    const int s = blockIdx.y;
    const double* w = ws + (long long)s * n;
    const int n_chunks = (n + chunk_pow2 - 1) / chunk_pow2;
    const int p0 = blockIdx.x * per_block, p1 = min(n + m, p0 + per_block);
    long long best = 0;
    for (int i = p0 + threadIdx.x; i < p1; i += blockDim.x) {
        const double z = i < n ? w[i] : emp_sorted[i - n];
        long long cs = 0;
        for (int c = 0; c < n_chunks; c++) {
            const int c0 = c * chunk_pow2;
            cs += upper_bound(w + c0, min(chunk_pow2, n - c0), z);
        }
        const long long ce = upper_bound(emp_sorted, m, z);
        const long long dd = cs * m - ce * n;
        best = max(best, dd < 0 ? -dd : dd);
    }
    const double b = wg_max((double)best, red);  // integers below 2^53: exact in double
    if (threadIdx.x == 0) atomicMax(dmax + s, (unsigned long long)b);
}

__global__ void ks_finish_kernel(int sim0, int n_batch, int n, int m, const unsigned long long* dmax,
                                 const int* nan_flag, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_batch) out[sim0 + s] = nan_flag[s] ? NAN : (double)dmax[s] / ((double)n * (double)m);
}

__global__ void ks_clear_kernel(int n_batch, unsigned long long* dmax, int* nan_flag) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n_batch) { dmax[s] = 0; nan_flag[s] = 0; }
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

// Device buffers the GOF entry points reuse across calls, grown when a call
// needs more and held until the process ends (the runtime may be gone by the
// time static destructors run, so they are not freed then).  Calls are
// serialised by a mutex; they run on the null stream and wait for it (not
// for the whole device).
struct GofBuffers {
    std::mutex mu;
    void* p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t bytes[5] = {0, 0, 0, 0, 0};
    hipError_t ensure(int i, size_t b) {
        if (bytes[i] >= b) return hipSuccess;
        if (p[i]) { (void)hipFree(p[i]); p[i] = nullptr; bytes[i] = 0; }
        hipError_t e = hipMalloc(&p[i], b);
        if (e == hipSuccess) bytes[i] = b;
        return e;
    }
};
GofBuffers& bufs() {
    static GofBuffers b;
    return b;
}
enum { BUF_SIM = 0, BUF_EMP = 1, BUF_OUT = 2, BUF_KS_WS = 3, BUF_KS_RED = 4 };

hipError_t fc_corr_launch(const double* d_sim, int n_sims, int n_pairs, const double* d_emp, double* d_out) {
    hipLaunchKernelGGL(cubnm::fc_corr_kernel, dim3(std::min(n_sims, 4096)), dim3(cubnm::GOF_THREADS), 0, 0,
                       d_sim, n_sims, n_pairs, d_emp, d_out);
    return hipGetLastError();
}

// KS over all simulations in batches of at most `batch`, with the sorted
// copies in ws (batch * n doubles) and the per-simulation reductions in dmax
// and nan_flag (batch each), on stream st; launches only, no allocation and
// no synchronisation.
hipError_t fcd_ks_run(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m, double* d_out,
                      double* ws, unsigned long long* dmax, int* nan_flag, int batch, hipStream_t st) {
    using namespace cubnm;
    int chunk = 1;
    while (chunk < std::min(n, KS_CHUNK)) chunk <<= 1;
    const int n_chunks = (n + chunk - 1) / chunk;
    const int per_block = GOF_THREADS * 16;
    const int eval_blocks = (n + m + per_block - 1) / per_block;
    for (int s0 = 0; s0 < n_sims; s0 += batch) {
        const int nb = std::min(batch, n_sims - s0);
        hipLaunchKernelGGL(ks_clear_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, nb, dmax, nan_flag);
        hipLaunchKernelGGL(ks_sort_chunks_kernel, dim3(n_chunks, nb), dim3(GOF_THREADS), sizeof(double) * chunk, st,
                           d_sim, s0, n, chunk, ws, nan_flag);
        hipLaunchKernelGGL(ks_eval_kernel, dim3(eval_blocks, nb), dim3(GOF_THREADS), 0, st, ws, n, chunk,
                           d_emp_sorted, m, per_block, dmax);
        hipLaunchKernelGGL(ks_finish_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, s0, nb, n, m, dmax, nan_flag,
                           d_out);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Simulations per batch for a workspace of `bytes`: n doubles of sorted
// copies plus 12 bytes of reductions each; at most 65,535 (grid.y).
size_t ks_batch_for(size_t bytes, int n) {
    return std::min<size_t>(bytes / (sizeof(double) * (size_t)n + sizeof(unsigned long long) + sizeof(int)), 65535);
}

// The same over the cached (grown on demand) device buffers.
hipError_t fcd_ks_launch(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m, double* d_out) {
    GofBuffers& B = bufs();
    const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)n_sims, ks_batch_for((size_t)256 << 20, n)));
    hipError_t e;
    if ((e = B.ensure(BUF_KS_WS, sizeof(double) * (size_t)n * batch)) != hipSuccess) return e;
    if ((e = B.ensure(BUF_KS_RED, (sizeof(unsigned long long) + sizeof(int)) * batch)) != hipSuccess) return e;
    unsigned long long* dmax = static_cast<unsigned long long*>(B.p[BUF_KS_RED]);
    return fcd_ks_run(d_sim, n_sims, n, d_emp_sorted, m, d_out, static_cast<double*>(B.p[BUF_KS_WS]), dmax,
                      reinterpret_cast<int*>(dmax + batch), batch, 0);
}

// Fixed-size workspace of the allocation-free KS entry point, one per GPU.
// Host code must not use this name as an address: it is resolved per device
// with hipGetSymbolAddress (ks_static_scratch) and the result passed on.
constexpr size_t KS_STATIC_BYTES = (size_t)16 << 20;
__device__ __align__(256) unsigned char scratch[KS_STATIC_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// Device address of this GPU's scratch array, resolved on its first use on
// the current device.
hipError_t ks_static_scratch(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

int cubnm_gof_fc_corr_device(const double* d_sim, int n_sims, int n_pairs, const double* d_emp, double* d_out);
int cubnm_gof_fcd_ks_device(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m, double* d_out);
int cubnm_gof_fcd_ks_device_static(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m,
                                   double* d_out, void* stream);

const char* cubnm_gof_last_error(void) { return g_gof_err.c_str(); }

int cubnm_gof_fc_corr_device(const double* d_sim, int n_sims, int n_pairs, const double* d_emp, double* d_out) {
    if (n_sims <= 0 || n_pairs <= 1) { g_gof_err = "need n_sims > 0 and n_pairs > 1"; return -1; }
    return guarded([&]() -> hipError_t {
        hipError_t e = fc_corr_launch(d_sim, n_sims, n_pairs, d_emp, d_out);
        return e == hipSuccess ? hipStreamSynchronize(0) : e;
    });
}

int cubnm_gof_fcd_ks_device(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m, double* d_out) {
    if (n_sims <= 0 || n <= 0 || m <= 0) { g_gof_err = "need n_sims, n and m > 0"; return -1; }
    std::lock_guard<std::mutex> lock(bufs().mu);
    return guarded([&]() -> hipError_t {
        hipError_t e = fcd_ks_launch(d_sim, n_sims, n, d_emp_sorted, m, d_out);
        return e == hipSuccess ? hipStreamSynchronize(0) : e;
    });
}

int cubnm_gof_fcd_ks_device_static(const double* d_sim, int n_sims, int n, const double* d_emp_sorted, int m,
                                   double* d_out, void* stream) {
    if (n_sims <= 0 || n <= 0 || m <= 0) { g_gof_err = "need n_sims, n and m > 0"; return -1; }
    const size_t batch = ks_batch_for(KS_STATIC_BYTES, n);
    if (batch < 1) {
        g_gof_err = "cubnm_gof_fcd_ks_device_static: " + std::to_string(n) +
                    " values per simulation exceed its fixed workspace of " + std::to_string(KS_STATIC_BYTES) +
                    " bytes (at most " + std::to_string((KS_STATIC_BYTES - 12) / 8) + " values)";
        cubnm::set_last_error(g_gof_err);
        return -1;
    }
    return guarded([&]() -> hipError_t {
        void* p = nullptr;
        hipError_t e = ks_static_scratch(&p);
        if (e != hipSuccess) return e;
        const int nb = (int)std::min<size_t>(batch, (size_t)n_sims);
        double* ws = static_cast<double*>(p);
        unsigned long long* dmax = reinterpret_cast<unsigned long long*>(ws + (size_t)n * nb);
        int* nan_flag = reinterpret_cast<int*>(dmax + nb);
        return fcd_ks_run(d_sim, n_sims, n, d_emp_sorted, m, d_out, ws, dmax, nan_flag, nb,
                          static_cast<hipStream_t>(stream));
    });
}

int cubnm_gof_fc_corr(const double* sim_fc, int n_sims, int n_pairs, const double* emp_fc, double* out) {
    if (n_sims <= 0 || n_pairs <= 1) { g_gof_err = "need n_sims > 0 and n_pairs > 1"; return -1; }
    GofBuffers& B = bufs();
    std::lock_guard<std::mutex> lock(B.mu);
    return guarded([&]() -> hipError_t {
        hipError_t e;
        const size_t sb = sizeof(double) * (size_t)n_sims * n_pairs;
        if ((e = B.ensure(BUF_SIM, sb)) != hipSuccess) return e;
        if ((e = B.ensure(BUF_EMP, sizeof(double) * n_pairs)) != hipSuccess) return e;
        if ((e = B.ensure(BUF_OUT, sizeof(double) * n_sims)) != hipSuccess) return e;
        double *ds = static_cast<double*>(B.p[BUF_SIM]), *de = static_cast<double*>(B.p[BUF_EMP]);
        double* dout = static_cast<double*>(B.p[BUF_OUT]);
        if ((e = hipMemcpyAsync(ds, sim_fc, sb, hipMemcpyHostToDevice, 0)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(de, emp_fc, sizeof(double) * n_pairs, hipMemcpyHostToDevice, 0)) != hipSuccess) return e;
        if ((e = fc_corr_launch(ds, n_sims, n_pairs, de, dout)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(out, dout, sizeof(double) * n_sims, hipMemcpyDeviceToHost, 0)) != hipSuccess) return e;
        return hipStreamSynchronize(0);
    });
}

int cubnm_gof_fcd_ks(const double* sim_fcd, int n_sims, int n, const double* emp_fcd, int m, double* out) {
    if (n_sims <= 0 || n <= 0 || m <= 0) { g_gof_err = "need n_sims, n and m > 0"; return -1; }
    std::vector<double> emp(emp_fcd, emp_fcd + m);
    for (double x : emp) if (std::isnan(x)) { for (int i = 0; i < n_sims; i++) out[i] = NAN; return 0; }
    std::sort(emp.begin(), emp.end());
    GofBuffers& B = bufs();
    std::lock_guard<std::mutex> lock(B.mu);
    return guarded([&]() -> hipError_t {
        hipError_t e;
        const size_t sb = sizeof(double) * (size_t)n_sims * n;
        if ((e = B.ensure(BUF_SIM, sb)) != hipSuccess) return e;
        if ((e = B.ensure(BUF_EMP, sizeof(double) * m)) != hipSuccess) return e;
        if ((e = B.ensure(BUF_OUT, sizeof(double) * n_sims)) != hipSuccess) return e;
        double *ds = static_cast<double*>(B.p[BUF_SIM]), *de = static_cast<double*>(B.p[BUF_EMP]);
        double* dout = static_cast<double*>(B.p[BUF_OUT]);
        if ((e = hipMemcpyAsync(ds, sim_fcd, sb, hipMemcpyHostToDevice, 0)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(de, emp.data(), sizeof(double) * m, hipMemcpyHostToDevice, 0)) != hipSuccess) return e;
        if ((e = fcd_ks_launch(ds, n_sims, n, de, m, dout)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(out, dout, sizeof(double) * n_sims, hipMemcpyDeviceToHost, 0)) != hipSuccess) return e;
        return hipStreamSynchronize(0);
    });
}

}  // extern "C"
