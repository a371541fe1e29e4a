// XY (n=2) and Heisenberg (n=3) checkerboard Metropolis + overrelaxation, gfx950.
//
// Replaces continuous_metropolis_kernel / continuous_overrelax_kernel /
// init_continuous_rng_kernel (src/cuda/continuous_spin_kernel.cu:9-152) and
// reduce_mag_continuous / reduce_energy_continuous (src/cuda/reduce_kernel.cu:70-151).
//
// Storage: FP32 unit vectors, structure of arrays per colour half-lattice:
// component k of colour c of replica r lives at off + ((2r + c) * n + k) * stride.
// A pass over colour c reads n floats of c and of 1-c and writes n floats of c:
// 3 * 4n bytes per update (XY 24 B, Heisenberg 36 B).
//
// Proposal (symmetric, unlike the reference GPU's cube perturbation):
//   Heisenberg: cap around the current spin, cos(theta) uniform in [cos d, 1],
//               phi uniform in [0, 2 pi)   (heisenberg/metropolis.rs:54-92)
//   XY:         rotation by alpha uniform in [-d, d)
// One Philox call per site and pass: words 0,1 -> proposal, word 2 -> acceptance.
// Overrelaxation: S' = 2 (S.h) h / |h|^2 - S, skipped when |h|^2 < 1e-30.
//
// Measurement is fused into the Metropolis passes and summed in FP64: block
// partials are written to a buffer and the closing launch after the colour-1
// pass adds them per replica in a fixed order (one wave per replica, spread over
// the launch's blocks), so every run with the same seed gives the same bits; the
// block of that launch that finishes last writes the records, decides the
// exchanges and advances the sweep counter.
//
// Threads: one site per thread, except XY's Metropolis (cont_metro_pair_kernel)
// and overrelaxation passes, which take two sites of the block's 256-site tile
// per thread.
#include "cont_pass.h"

namespace ising {

// Metropolis pass over colour `color` of every replica (O32: 32-bit offsets,
// see comp_at).
template <int N, bool O32>
__global__ __launch_bounds__(kCT) void cont_metro_kernel(ContArgs a, int nblocks, int color, int closes) {
    const int b = xcd_block();
    if (b >= nblocks) return;
    cont_metro_block<N, O32>(a, b, color, closes);
}

// The same pass for XY with two sites per thread (see cont_metro_pair_block).
template <int N, bool O32>
__global__ __launch_bounds__(kCT / 2) void cont_metro_pair_kernel(ContArgs a, int nblocks, int color, int closes) {
    const int b = xcd_block();
    if (b >= nblocks) return;
    cont_metro_pair_block<N, O32>(a, b, color, closes, true, 0, (int)threadIdx.x);
}

// Closes a continuous sweep; one launch of kCT-thread blocks (see cont_close_block).
__global__ __launch_bounds__(kCT) void cont_close_kernel(ContArgs a) { cont_close_block(a, (int)blockIdx.x, (int)gridDim.x); }

// Overrelaxation pass over colour `color`, SPT sites per thread (see cont_or_block).
template <int N, int SPT, bool O32>
__global__ __launch_bounds__(kCT / SPT) void cont_or_kernel(ContArgs a, int nblocks, int color) {
    const int b = xcd_block();
    if (b >= nblocks) return;
    cont_or_block<N, SPT, O32>(a, b, color);
}

// Random start: Heisenberg z = 2u-1, phi = 2 pi u' (uniform on S^2); XY theta = 2 pi u.
template <int N>
__global__ __launch_bounds__(kCT) void cont_init_kernel(const DevGroup* groups, int ngroups, float* pool, int model,
                                                      uint32_t k0, uint32_t k1) {
    const int g = find_group_c(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, (int)blockIdx.x - G.blk0, &r, &tile);
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    if (h >= G.n_half) return;
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);
    for (int c = 0; c < 2; ++c) {
        float* C = pool + G.off + (int64_t)(2 * r + c) * N * G.stride;
        const int64_t row = h / G.Lh;
        const int xh = (int)(h - row * G.Lh);
        const int z = (int)(row / G.L), y = (int)(row - (int64_t)z * G.L);
        const int x = 2 * xh + ((y + z + c) & 1);
        const uint32_t site = (uint32_t)(((int64_t)z * G.L + y) * G.L + x);
        const u32x4 w = philox(site, 0u, stream, tag(P_INIT, (uint32_t)model), k0, k1);
        if (N == 2) {
            float sn, cs;
            sincosf(6.28318530717958647692f * u24f(w.x), &sn, &cs);
            C[h] = cs;
            C[G.stride + h] = sn;
        } else {
            const float zz = 2.0f * u24f(w.x) - 1.0f;
            float sn, cs;
            sincosf(6.28318530717958647692f * u24f(w.y), &sn, &cs);
            const float rr = sqrtf(fmaxf(0.0f, 1.0f - zz * zz));
            C[h] = rr * cs;
            C[G.stride + h] = rr * sn;
            C[2 * G.stride + h] = zz;
        }
    }
}

// Standalone FP64 E/M of every replica: part[4*blk + k] partials (E, Mx, My, Mz).
template <int N>
__global__ __launch_bounds__(kCT) void cont_measure_kernel(const DevGroup* groups, int ngroups, const float* pool,
                                                         double J, double* part) {
    __shared__ double s_red[4][kCT / 64];
    const int g = find_group_c(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, (int)blockIdx.x - G.blk0, &r, &tile);
    const float* C1 = pool + G.off + (int64_t)(2 * r + 1) * N * G.stride;
    const float* C0 = pool + G.off + (int64_t)(2 * r) * N * G.stride;
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (h < G.n_half) {
        uint32_t nb[6];
        neighbours(G, h, 1, nb);
        double e = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double f = 0.0;
            for (int q = 0; q < 6; ++q) f += (double)C0[k * G.stride + nb[q]];
            const double s1 = (double)C1[k * G.stride + h];
            e += s1 * f;
            v[1 + k] = s1 + (double)C0[k * G.stride + h];
        }
        v[0] = -J * e;
    }
    block_sum4(v, s_red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * (int64_t)blockIdx.x + k] = v[k];
}

// XY spin stiffness sums of every replica: over the forward bonds (i, i + d),
// d = x, y, z, C_d = sum cos(theta_i - theta_j) = S_i . S_j and
// J_d = sum sin(theta_i - theta_j) = S_i^y S_j^x - S_i^x S_j^y, as
// part[6 * blk + {C_x, C_y, C_z, J_x, J_y, J_z}] block partials (FP64).
// The helicity modulus along d is (<C_d> - beta <J_d^2>) / N.
__global__ __launch_bounds__(kCT) void cont_helicity_kernel(const DevGroup* groups, int ngroups, const float* pool,
                                                          double* part) {
    __shared__ double s_red[4][kCT / 64];
    const int g = find_group_c(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, (int)blockIdx.x - G.blk0, &r, &tile);
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (h < G.n_half) {
        for (int c = 0; c < 2; ++c) {
            const float* Ci = pool + G.off + (int64_t)(2 * r + c) * 2 * G.stride;
            const float* Cj = pool + G.off + (int64_t)(2 * r + (c ^ 1)) * 2 * G.stride;
            uint32_t nb[6];
            neighbours(G, h, c, nb);
            const int64_t row = h / G.Lh;
            const int z = (int)(row / G.L), y = (int)(row - (int64_t)z * G.L);
            const int p = (y + z + c) & 1;
            const int64_t fwd[3] = {p ? nb[1] : nb[0], nb[3], nb[5]};  // +x, +y, +z
            const double xi = (double)Ci[h], yi = (double)Ci[G.stride + h];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double xj = (double)Cj[fwd[d]], yj = (double)Cj[G.stride + fwd[d]];
                v[d] += xi * xj + yi * yj;
                v[3 + d] += yi * xj - xi * yj;
            }
        }
    }
    block_sum4(v, s_red);
    __syncthreads();
    block_sum4(v + 4, s_red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * (int64_t)blockIdx.x + k] = v[k];
}

hipError_t launch_cont_helicity(const DevGroup* groups, int ngroups, const float* pool, double* part, int nblocks,
                                hipStream_t st) {
    hipLaunchKernelGGL(cont_helicity_kernel, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, part);
    return hipGetLastError();
}

static inline int xcd_grid(int nblocks) { return (nblocks + 7) & ~7; }

template <bool O32>
static void launch_metro(const ContArgs& a, int nblocks, int color, int closes, hipStream_t st) {
    const dim3 grid(xcd_grid(nblocks));
    if (a.ncomp == 2)
        hipLaunchKernelGGL((cont_metro_pair_kernel<2, O32>), grid, dim3(kCT / 2), 0, st, a, nblocks, color, closes);
    else hipLaunchKernelGGL((cont_metro_kernel<3, O32>), grid, dim3(kCT), 0, st, a, nblocks, color, closes);
}

hipError_t launch_cont_metro(const ContArgs& a, int nblocks, int color, int closes, hipStream_t st) {
    if (a.off32) launch_metro<true>(a, nblocks, color, closes, st);
    else launch_metro<false>(a, nblocks, color, closes, st);
    return hipGetLastError();
}

hipError_t launch_cont_close(const ContArgs& a, hipStream_t st) {
    static_assert(kContClosePerBlock == kCT / 64, "one wave per replica");
    const int nb = (a.total_reps + kContClosePerBlock - 1) / kContClosePerBlock;
    hipLaunchKernelGGL(cont_close_kernel, dim3(nb), dim3(kCT), 0, st, a);
    return hipGetLastError();
}

template <bool O32>
static void launch_or(const ContArgs& a, int nblocks, int color, hipStream_t st) {
    const dim3 grid(xcd_grid(nblocks));
    const int spt = a.or_spt ? a.or_spt : (a.ncomp == 2 ? 2 : 1);
    if (a.ncomp == 2 && spt == 2)
        hipLaunchKernelGGL((cont_or_kernel<2, 2, O32>), grid, dim3(kCT / 2), 0, st, a, nblocks, color);
    else if (a.ncomp == 2) hipLaunchKernelGGL((cont_or_kernel<2, 1, O32>), grid, dim3(kCT), 0, st, a, nblocks, color);
    else if (spt == 2) hipLaunchKernelGGL((cont_or_kernel<3, 2, O32>), grid, dim3(kCT / 2), 0, st, a, nblocks, color);
    else hipLaunchKernelGGL((cont_or_kernel<3, 1, O32>), grid, dim3(kCT), 0, st, a, nblocks, color);
}

hipError_t launch_cont_or(const ContArgs& a, int nblocks, int color, hipStream_t st) {
    if (a.off32) launch_or<true>(a, nblocks, color, st);
    else launch_or<false>(a, nblocks, color, st);
    return hipGetLastError();
}

hipError_t launch_cont_init(const DevGroup* groups, int ngroups, float* pool, int ncomp, int model, int nblocks,
                            uint32_t k0, uint32_t k1, hipStream_t st) {
    if (ncomp == 2)
        hipLaunchKernelGGL(cont_init_kernel<2>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, model, k0, k1);
    else
        hipLaunchKernelGGL(cont_init_kernel<3>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, model, k0, k1);
    return hipGetLastError();
}

hipError_t launch_cont_measure(const DevGroup* groups, int ngroups, const float* pool, int ncomp, double J,
                               double* part, int nblocks, hipStream_t st) {
    if (ncomp == 2)
        hipLaunchKernelGGL(cont_measure_kernel<2>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, J, part);
    else
        hipLaunchKernelGGL(cont_measure_kernel<3>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, J, part);
    return hipGetLastError();
}

}  // namespace ising
