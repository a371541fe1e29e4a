// Reductions of a caller's device lattice in the reference's layout, gfx950.
//
// Replaces reduce_mag_ising / reduce_energy_ising / reduce_mag_continuous /
// reduce_energy_continuous (src/cuda/reduce_kernel.cu:7-151) and the host sums
// of src/cuda/reduce_gpu.rs:23-163.  Layout: z-major, site (z*n + y)*n + x;
// Ising int8 +-1; continuous spins component-interleaved FP32 (x, y[, z]).
//
// Two launches: one partial per 4096-site tile (256 threads x 16 sites, each
// step of the block reading 256 consecutive sites), written to a buffer the
// caller provides (`part`), then one block adds the partials in a fixed order
// into `out`.  Ising sums are exact int64 (bond sum B over the three forward
// neighbours, E = -J B; the reference adds FP32 partials, exact while a
// block's sum stays below 2^24).  Continuous sums are FP64 in a fixed order,
// so a call gives the same bits every time.
#include "kernels.h"

namespace ising {

constexpr int kRedT = 256;     // threads per block
constexpr int kRedSites = 16;  // sites per thread
static_assert(kRedT * kRedSites == kReduceTile, "host tiling must match the kernels");

// Fixed-order block sum of Q values per thread; result valid in thread 0.
template <typename T, int Q>
__device__ __forceinline__ void block_sum_fixed(T v[Q]) {
    __shared__ T s[Q][kRedT / 64];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = wave_sum(v[q]);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < Q; ++q) s[q][wid] = v[q];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            T a = s[q][0];
            for (int w = 1; w < kRedT / 64; ++w) a += s[q][w];
            v[q] = a;
        }
}

__device__ __forceinline__ void site_xyz(uint32_t i, uint32_t n, uint32_t* x, uint32_t* y, uint32_t* z) {
    const uint32_t nn = n * n;
    *z = i / nn;
    const uint32_t r = i - *z * nn;
    *y = r / n;
    *x = r - *y * n;
}

// part[2 * blk] = sum over the tile of s_i (s_{x+1} + s_{y+1} + s_{z+1}), part[2 * blk + 1] = sum of s_i.
__global__ __launch_bounds__(kRedT) void red_ising_part_kernel(const int8_t* s, uint32_t n, int64_t* part) {
    const uint32_t n3 = n * n * n, nn = n * n;
    const uint32_t base = blockIdx.x * (uint32_t)kReduceTile;
    int b = 0, m = 0;  // at most 16 sites of one thread: |b| <= 48
#pragma unroll 4
    for (int k = 0; k < kRedSites; ++k) {
        const uint32_t i = base + (uint32_t)k * kRedT + threadIdx.x;
        if (i >= n3) break;
        uint32_t x, y, z;
        site_xyz(i, n, &x, &y, &z);
        const uint32_t xp = x + 1 == n ? 0 : x + 1, yp = y + 1 == n ? 0 : y + 1, zp = z + 1 == n ? 0 : z + 1;
        const int si = s[i];
        b += si * ((int)s[z * nn + y * n + xp] + (int)s[z * nn + yp * n + x] + (int)s[zp * nn + y * n + x]);
        m += si;
    }
    int64_t v[2] = {b, m};
    block_sum_fixed<int64_t, 2>(v);
    if (threadIdx.x == 0) {
        part[2 * (int64_t)blockIdx.x] = v[0];
        part[2 * (int64_t)blockIdx.x + 1] = v[1];
    }
}

// part[4 * blk + {0, 1, 2, 3}] = sum over the tile of (S_i . (S_{x+1} + S_{y+1} + S_{z+1}), Sx, Sy, Sz);
// the bond term only when n > 0 (n_sites = n^3), else 0.
template <int NC>
__global__ __launch_bounds__(kRedT) void red_cont_part_kernel(const float* s, uint32_t n_sites, uint32_t n,
                                                              double* part) {
    const uint32_t base = blockIdx.x * (uint32_t)kReduceTile;
    const uint32_t nn = n * n;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < kRedSites; ++k) {
        const uint32_t i = base + (uint32_t)k * kRedT + threadIdx.x;
        if (i >= n_sites) break;
        double si[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            si[c] = (double)s[(int64_t)i * NC + c];
            v[1 + c] += si[c];
        }
        if (n > 0) {
            uint32_t x, y, z;
            site_xyz(i, n, &x, &y, &z);
            const uint32_t xp = x + 1 == n ? 0 : x + 1, yp = y + 1 == n ? 0 : y + 1, zp = z + 1 == n ? 0 : z + 1;
            const uint32_t f[3] = {z * nn + y * n + xp, z * nn + yp * n + x, zp * nn + y * n + x};
            double e = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int c = 0; c < NC; ++c) e += si[c] * (double)s[(int64_t)f[d] * NC + c];
            v[0] += e;
        }
    }
    block_sum_fixed<double, 4>(v);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) part[4 * (int64_t)blockIdx.x + q] = v[q];
}

// out[q] = sum over blocks of part[Q * blk + q], one block, fixed order.
template <typename T, int Q>
__global__ __launch_bounds__(kRedT) void red_final_kernel(const T* part, int32_t nblk, T* out) {
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = T(0);
    for (int b = threadIdx.x; b < nblk; b += kRedT)
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q] += part[(int64_t)Q * b + q];
    block_sum_fixed<T, Q>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < Q; ++q) out[q] = v[q];
}

hipError_t launch_reduce_ising(const int8_t* spins, int32_t n, int64_t* part, int64_t* out, hipStream_t st) {
    const int nblk = reduce_blocks((int64_t)n * n * n);
    hipLaunchKernelGGL(red_ising_part_kernel, dim3(nblk), dim3(kRedT), 0, st, spins, (uint32_t)n, part);
    hipLaunchKernelGGL((red_final_kernel<int64_t, 2>), dim3(1), dim3(kRedT), 0, st, (const int64_t*)part, nblk, out);
    return hipGetLastError();
}

hipError_t launch_reduce_cont(const float* spins, int64_t n_sites, int32_t n, int32_t ncomp, double* part,
                              double* out, hipStream_t st) {
    const int nblk = reduce_blocks(n_sites);
    if (ncomp == 2)
        hipLaunchKernelGGL(red_cont_part_kernel<2>, dim3(nblk), dim3(kRedT), 0, st, spins, (uint32_t)n_sites,
                           (uint32_t)n, part);
    else
        hipLaunchKernelGGL(red_cont_part_kernel<3>, dim3(nblk), dim3(kRedT), 0, st, spins, (uint32_t)n_sites,
                           (uint32_t)n, part);
    hipLaunchKernelGGL((red_final_kernel<double, 4>), dim3(1), dim3(kRedT), 0, st, (const double*)part, nblk, out);
    return hipGetLastError();
}

}  // namespace ising
