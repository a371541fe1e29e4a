// Multiple-histogram (WHAM / Ferrenberg-Swendsen) reweighting on the device, FP64.
//
// Restates analysis/scripts/reweighting.py:45-108 (wham_reweight): with K
// simulations at beta_k holding N_k energy samples (concatenated, N in total),
// iterate
//   log D_n  = logsumexp_k( log N_k + f_k - beta_k E_n )
//   f_k      = -logsumexp_n( -beta_k E_n - log D_n ),   then f -= f_0,
// until max |f_new - f| < tol, and reweight every observable to a target beta
// with weights w_n ~ exp(-beta E_n - log D_n).  Every log-sum-exp is computed
// in one pass with the running (max, scaled sum) pair and merged in a fixed
// order, so results are deterministic.
#include "kernels.h"

namespace ising {

constexpr int kWT = 256;
constexpr int kWMaxK = 1024;
constexpr int kWMaxObs = 8;

// log D_n for every sample.
__global__ __launch_bounds__(kWT) void wham_denom_kernel(const double* __restrict__ E, int64_t n, int K,
                                                       const double* __restrict__ logN, const double* __restrict__ f,
                                                       const double* __restrict__ beta, double* __restrict__ logD) {
    __shared__ double s_a[kWMaxK], s_b[kWMaxK];
    for (int k = threadIdx.x; k < K; k += kWT) {
        s_a[k] = logN[k] + f[k];
        s_b[k] = beta[k];
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kWT + threadIdx.x;
    if (i >= n) return;
    const double e = E[i];
    double m = -__builtin_huge_val(), s = 0.0;
    for (int k = 0; k < K; ++k) {
        const double v = s_a[k] - s_b[k] * e;
        if (v > m) {
            s = s * exp(m - v) + 1.0;
            m = v;
        } else {
            s += exp(v - m);
        }
    }
    logD[i] = m + log(s);
}

// Partial log-sum-exp of  -b_t E_n - log D_n  (and weighted observable sums)
// over one chunk of samples, for target t = blockIdx.y.
// part layout per (t, chunk): [m, s, sum_j for each observable].
__global__ __launch_bounds__(kWT) void wham_partial_kernel(const double* __restrict__ E, int64_t n,
                                                         const double* __restrict__ logD,
                                                         const double* __restrict__ tb, int n_obs,
                                                         const double* __restrict__ obs, double* __restrict__ part) {
    __shared__ double s_m[kWT], s_s[kWT], s_o[kWMaxObs][kWT];
    const int t = blockIdx.y;
    const double b = tb[t];
    double m = -__builtin_huge_val(), s = 0.0, a[kWMaxObs];
    for (int j = 0; j < kWMaxObs; ++j) a[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kWT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWT) {
        const double v = -b * E[i] - logD[i];
        double scale_old = 1.0, w = 1.0;
        if (v > m) {
            scale_old = exp(m - v);
            m = v;
        } else {
            w = exp(v - m);
        }
        s = s * scale_old + w;
        for (int j = 0; j < n_obs; ++j) a[j] = a[j] * scale_old + w * obs[(int64_t)j * n + i];
    }
    s_m[threadIdx.x] = m;
    s_s[threadIdx.x] = s;
    for (int j = 0; j < n_obs; ++j) s_o[j][threadIdx.x] = a[j];
    __syncthreads();
    for (int w = kWT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const int u = threadIdx.x + w;
            const double m1 = s_m[threadIdx.x], m2 = s_m[u];
            const double mm = fmax(m1, m2);
            const double c1 = m1 == -__builtin_huge_val() ? 0.0 : exp(m1 - mm);
            const double c2 = m2 == -__builtin_huge_val() ? 0.0 : exp(m2 - mm);
            s_m[threadIdx.x] = mm;
            s_s[threadIdx.x] = s_s[threadIdx.x] * c1 + s_s[u] * c2;
            for (int j = 0; j < n_obs; ++j) s_o[j][threadIdx.x] = s_o[j][threadIdx.x] * c1 + s_o[j][u] * c2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* p = part + ((int64_t)t * gridDim.x + blockIdx.x) * (2 + kWMaxObs);
        p[0] = s_m[0];
        p[1] = s_s[0];
        for (int j = 0; j < n_obs; ++j) p[2 + j] = s_o[j][0];
    }
}

// Merge the chunk partials of each target in chunk order: out[t] = (logsumexp, <obs_j>).
__global__ void wham_merge_kernel(const double* __restrict__ part, int nchunk, int n_targets, int n_obs,
                                  double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_targets) return;
    double m = -__builtin_huge_val(), s = 0.0, a[kWMaxObs];
    for (int j = 0; j < kWMaxObs; ++j) a[j] = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        const double* p = part + ((int64_t)t * nchunk + c) * (2 + kWMaxObs);
        if (p[0] == -__builtin_huge_val()) continue;
        const double mm = fmax(m, p[0]);
        const double c1 = m == -__builtin_huge_val() ? 0.0 : exp(m - mm), c2 = exp(p[0] - mm);
        s = s * c1 + p[1] * c2;
        for (int j = 0; j < n_obs; ++j) a[j] = a[j] * c1 + p[2 + j] * c2;
        m = mm;
    }
    double* o = out + (int64_t)t * (1 + n_obs);
    o[0] = m + log(s);
    for (int j = 0; j < n_obs; ++j) o[1 + j] = a[j] / s;
}

hipError_t launch_wham_denom(const double* E, int64_t n, int K, const double* logN, const double* f,
                             const double* beta, double* logD, hipStream_t st) {
    hipLaunchKernelGGL(wham_denom_kernel, dim3((unsigned)((n + kWT - 1) / kWT)), dim3(kWT), 0, st, E, n, K, logN, f,
                       beta, logD);
    return hipGetLastError();
}

hipError_t launch_wham_reduce(const double* E, int64_t n, const double* logD, const double* tb, int n_targets,
                              int n_obs, const double* obs, double* part, int nchunk, double* out, hipStream_t st) {
    hipLaunchKernelGGL(wham_partial_kernel, dim3(nchunk, n_targets), dim3(kWT), 0, st, E, n, logD, tb, n_obs, obs,
                       part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(wham_merge_kernel, dim3((n_targets + 63) / 64), dim3(64), 0, st, part, nchunk, n_targets, n_obs,
                       out);
    return hipGetLastError();
}

}  // namespace ising
