// Ising checkerboard Metropolis on red/black half-lattices, gfx950.
//
// Replaces metropolis_sweep_kernel / init_rng_kernel (src/cuda/kernels.cu:9-62)
// and reduce_mag_ising / reduce_energy_ising (src/cuda/reduce_kernel.cu:7-66).
//
// Storage: one byte per site, 0 = down, 1 = up, in colour half-lattices.  A
// half-sweep of colour c reads c and 1-c and writes c: 3 bytes per update.
//
// Acceptance: accept iff u < T(k), k = b*7 + n_up, with u a 32-bit uniform
// and T = floor(2^32 min(1, exp(-beta dE))) a host-built 33-bit threshold
// split as T = hi * 2^25 + lo (hi in 0..128).  The top 7 bits r7 of u are the
// low 7 bits of the site's byte of one Philox call shared by 16 consecutive
// sites (counter c0 = h >> 4, site h takes byte h & 15); only when r7 == hi
// (p = 2^-7) is a second Philox call, keyed by the site, drawn for the other
// 25 bits.  This is exactly the 32-bit comparison, at 1/16 of a Philox call
// per update.  A wave queues its tied sites in LDS and decides them 64 at a
// time, so the second call costs one lane per tie, not a whole wave.
//
// The 16 sites of a chunk are handled four per 32-bit word: neighbour counts
// by byte-wise adds, the threshold byte of each site by two byte permutes of
// the packed table (one per spin value), and r7 < hi by one subtraction of
// (r | 0x80) - hi whose per-byte sign bit is the comparison.
//
// Measurement is fused: the colour-0 pass sums its new spins, the colour-1 pass
// sums its new spins and s_i * (sum of its 6 neighbours), which counts every
// bond exactly once (each bond joins the two colours).  Block partials go to
// int64 accumulators by atomics.  Sweep t is closed at the start of the
// colour-0 launch of sweep t + 1 (sweepctl.h): its blocks decide the replica
// exchanges from the committed sums, one block writes the records and the next
// sweep's control fields.  A one-block closing kernel closes the last sweep of
// each host call, so the host always reads complete state.
#include "ising_pass.h"

namespace ising {

// Multispin coding: bit plane j (MSB first) of the class-k threshold of every
// replica of a 32-replica word, bit r of the plane belonging to replica r,
// for the permutation of parity `par`.  All threads of one block.
__device__ void msc_planes(const IsingArgs& a, int par) {
    const int32_t* r2t = a.r2t + (int64_t)par * a.total_reps;
    for (int g = 0; g < a.ngroups; ++g) {
        const DevGroup G = a.groups[g];
        const int nwords = (G.R + 31) / 32;
        for (int e = threadIdx.x; e < nwords * 96; e += blockDim.x) {
            const int wi = e / 96, kj = e - wi * 96, kk = kj >> 5, j = kj & 31;
            const int nrep = min(32, G.R - 32 * wi);
            uint32_t v = 0;
            for (int bit = 0; bit < nrep; ++bit) {
                const int ti = r2t[G.rep0 + 32 * wi + bit];
                v |= ((a.mthr[3 * (G.rep0 + ti) + kk] >> (31 - j)) & 1u) << bit;
            }
            a.mplanes[(int64_t)(G.rep0 + 32 * wi) * 96 + kj] = v;
        }
    }
}

// One colour half-sweep of every replica of every group in the ensemble.
// Grid: one block per (group, replica, tile of G.tsz half-sites), padded to
// a multiple of 8 blocks.  Blocks are dealt to the 8 XCDs round-robin (block
// i runs on XCD i % 8); the tiles are renumbered so each XCD walks one
// contiguous range of them (the z-neighbour planes of a tile then stay in
// that XCD's L2).
__global__ __launch_bounds__(kThreads) void ising_half_kernel(IsingArgs a, int color) {
    const int per = (int)(gridDim.x >> 3);
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (b >= a.nblocks) return;
    DevGroup G;
    if (a.gk_on) {
        int g = 0;
        while (g + 1 < a.ngroups && b >= a.gk[g + 1].blk0) ++g;
        G = a.gk[g];
    } else {
        G = a.groups[find_group(a.groups, a.ngroups, b)];
    }
    ising_half_block(a, b, color, G);
}

// Closes the last sweep of a host call (records, replica exchange), which no
// colour-0 launch will close.  One block, after the colour-1 launch.
__global__ __launch_bounds__(kThreads) void ising_close_kernel(IsingArgs a) {
    DevCtl* ctl = a.ctl;
    const int64_t tl = ctl->t - 1;
    if (ctl->closed >= tl) return;
    const int32_t pfl = ctl->flp;
    const int64_t round = ctl->round;
    const int64_t* acc = acc_of(a, tl);
    if (pfl & F_MEAS) write_records(a, acc, a.r2t + (round & 1) * a.total_reps);
    if (pfl & F_EXCH) {
        exch_write_all(a, round, acc);
        __syncthreads();
        if (a.mplanes) msc_planes(a, (int)((round + 1) & 1));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (pfl & F_EXCH) ctl->round = round + 1;
        ctl->closed = tl;
    }
}

// Hot start: i.i.d. +-1 with p = 1/2 from Philox(site, 0, stream, init tag).
__global__ __launch_bounds__(kThreads) void ising_init_kernel(const DevGroup* groups, int ngroups, uint8_t* pool,
                                                            uint32_t k0, uint32_t k1) {
    const int g = find_group(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, blockIdx.x - G.blk0, &r, &tile);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);
    for (int c = 0; c < 2; ++c) {
        uint8_t* C = pool + G.off + (int64_t)(2 * r + c) * G.stride;
        for (int j = 0; j < G.tsz / kThreads; ++j) {
            const int64_t h = (int64_t)tile * G.tsz + (int64_t)j * kThreads + threadIdx.x;
            if (h >= G.n_half) break;
            const int64_t row = h / G.Lh;
            const int xh = (int)(h - row * G.Lh);
            const int z = (int)(row / G.L), y = (int)(row - (int64_t)z * G.L);
            const int x = 2 * xh + ((y + z + c) & 1);
            const uint32_t site = (uint32_t)(((int64_t)z * G.L + y) * G.L + x);
            C[h] = (uint8_t)(philox(site, 0u, stream, tag(P_INIT, MODEL_ISING), k0, k1).x >> 31);
        }
    }
}

// Standalone E/M of every replica (for callers that measure outside a sweep):
// acc[2*grep] += B, acc[2*grep+1] += M.
__global__ __launch_bounds__(kThreads) void ising_measure_kernel(const DevGroup* groups, int ngroups,
                                                               const uint8_t* pool, int64_t* acc) {
    __shared__ int64_t s_red[2][kThreads / 64];
    const int g = find_group(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, blockIdx.x - G.blk0, &r, &tile);
    const int L = G.L, Lh = G.Lh;
    const uint8_t* C1 = pool + G.off + (int64_t)(2 * r + 1) * G.stride;
    const uint8_t* C0 = pool + G.off + (int64_t)(2 * r) * G.stride;
    int bsum = 0, msum = 0;
    for (int j = 0; j < G.tsz / kThreads; ++j) {
        const int64_t h = (int64_t)tile * G.tsz + (int64_t)j * kThreads + threadIdx.x;
        if (h >= G.n_half) break;
        const int64_t row = h / Lh;
        const int xh = (int)(h - row * Lh);
        const int z = (int)(row / L), y = (int)(row - (int64_t)z * L);
        const int p = (y + z + 1) & 1;
        const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
        const int zm = z == 0 ? L - 1 : z - 1, zp = z == L - 1 ? 0 : z + 1;
        const int xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
        int n = C0[row * Lh + xh] + C0[row * Lh + xo] + C0[((int64_t)z * L + ym) * Lh + xh] +
                C0[((int64_t)z * L + yp) * Lh + xh];
        if (G.dims == 3) n += C0[((int64_t)zm * L + y) * Lh + xh] + C0[((int64_t)zp * L + y) * Lh + xh];
        const int s1 = 2 * (int)C1[h] - 1;
        bsum += s1 * (2 * n - 2 * G.dims);
        msum += s1 + 2 * (int)C0[h] - 1;
    }
    const int wb = wave_sum(bsum), wm = wave_sum(msum);
    if ((threadIdx.x & 63) == 0) {
        s_red[0][threadIdx.x >> 6] = wb;
        s_red[1][threadIdx.x >> 6] = wm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t B = 0, M = 0;
        for (int i = 0; i < kThreads / 64; ++i) {
            B += s_red[0][i];
            M += s_red[1][i];
        }
        atomicAdd((unsigned long long*)&acc[2 * (G.rep0 + r)], (unsigned long long)B);
        atomicAdd((unsigned long long*)&acc[2 * (G.rep0 + r) + 1], (unsigned long long)M);
    }
}

}  // namespace ising

// ----------------------------------------------------------------------------
// Launchers used by the host API (same translation unit as the kernels).
// ----------------------------------------------------------------------------
namespace ising {

hipError_t launch_ising_half(const IsingArgs& a, int nblocks, int color, hipStream_t st) {
    const int grid = (nblocks + 7) / 8 * 8;
    hipLaunchKernelGGL(ising_half_kernel, dim3(grid), dim3(kThreads), 0, st, a, color);
    return hipGetLastError();
}

hipError_t launch_ising_close(const IsingArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(ising_close_kernel, dim3(1), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_ising_init(const DevGroup* groups, int ngroups, uint8_t* pool, int nblocks, uint32_t k0,
                             uint32_t k1, hipStream_t st) {
    hipLaunchKernelGGL(ising_init_kernel, dim3(nblocks), dim3(kThreads), 0, st, groups, ngroups, pool, k0, k1);
    return hipGetLastError();
}

hipError_t launch_ising_measure(const DevGroup* groups, int ngroups, const uint8_t* pool, int64_t* acc,
                                int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(ising_measure_kernel, dim3(nblocks), dim3(kThreads), 0, st, groups, ngroups, pool, acc);
    return hipGetLastError();
}

}  // namespace ising
