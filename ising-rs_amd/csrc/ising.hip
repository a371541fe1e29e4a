// Ising checkerboard Metropolis on red/black half-lattices, gfx950.
//
// Replaces metropolis_sweep_kernel / init_rng_kernel (src/cuda/kernels.cu:9-62)
// and reduce_mag_ising / reduce_energy_ising (src/cuda/reduce_kernel.cu:7-66).
//
// Storage: one byte per site, 0 = down, 1 = up, in colour half-lattices.  A
// half-sweep of colour c reads c and 1-c and writes c: 3 bytes per update.
//
// Acceptance: accept iff dE <= 0 or u < exp(-beta dE) with u a 32-bit uniform
// compared against a host-built 33-bit threshold T(k), k = b*7 + n_up.  The
// high 16 bits of u come from one Philox call shared by 8 consecutive sites
// (counter c0 = h >> 3); only when they equal T's high half (p = 2^-16) is a
// second Philox call (keyed by the site) drawn for the low 16 bits.  This is
// exactly a 32-bit comparison, at 1/8 of a Philox call per update.
//
// Measurement is fused: the colour-0 pass sums its new spins, the colour-1 pass
// sums its new spins and s_i * (sum of its 6 neighbours), which counts every
// bond exactly once (each bond joins the two colours).  Block partials go to
// int64 accumulators by atomics; after the colour-1 launch a closing kernel writes
// the records, performs the replica exchange and advances the sweep counter
// in a one-block closing kernel.
#include "kernels.h"

namespace ising {

constexpr int kThreads = 256;
constexpr int kSitesPerThread = 8;
constexpr int kTile = kThreads * kSitesPerThread;  // half-sites per block

__device__ __forceinline__ int find_group(const DevGroup* groups, int ngroups, int b) {
    int g = 0;
    while (g + 1 < ngroups && b >= groups[g + 1].blk0) ++g;
    return g;
}

__device__ __forceinline__ uint32_t tie_bits(uint32_t h, uint32_t hs, uint32_t stream, uint32_t k0, uint32_t k1) {
    return philox(h, hs, stream, tag(P_TIE, MODEL_ISING), k0, k1).x & 0xffffu;
}

__device__ __forceinline__ uint32_t word_of(const u32x4& w, int j) {
    return j < 2 ? w.x : (j < 4 ? w.y : (j < 6 ? w.z : w.w));
}

// Accumulated quantity q (0: bond sum, 1: magnetisation) of global replica gr,
// summed over the accumulator copies.
__device__ __forceinline__ int64_t acc_total(const IsingArgs& a, int gr, int q) {
    int64_t v = 0;
    for (int c = 0; c < a.acc_copies; ++c)
        v += __hip_atomic_load(&a.acc[2 * ((int64_t)c * a.total_reps + gr) + q], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

// Closing step of a sweep (one-block kernel after the colour-1 launch).
__device__ void close_sweep(const IsingArgs& a, int64_t t, bool meas, bool exch, int64_t round) {
    DevCtl* ctl = a.ctl;
    const int tid = threadIdx.x;
    const bool need = meas || exch;
    const int nrec = ctl->n_rec;
    const bool room = nrec < ctl->rec_cap;
    if (need) {
        for (int gr = tid; gr < a.total_reps; gr += blockDim.x) {
            const int64_t B = acc_total(a, gr, 0);
            const int64_t M = acc_total(a, gr, 1);
            if (meas && room) {
                IsingRec r;
                r.bond = B;
                r.mag = M;
                r.tidx = a.r2t[gr];
                r.pad = 0;
                a.rec[(int64_t)nrec * a.total_reps + gr] = r;
            }
        }
    }
    __syncthreads();
    if (exch) {
        for (int g = 0; g < a.ngroups; ++g) {
            const DevGroup G = a.groups[g];
            const int npairs = (G.R - (int)(round & 1)) / 2;
            for (int q = tid; q < npairs; q += blockDim.x) {
                const int i = (int)(round & 1) + 2 * q;
                const int ri = a.t2r[G.rep0 + i], rj = a.t2r[G.rep0 + i + 1];
                const int64_t Bi = acc_total(a, G.rep0 + ri, 0);
                const int64_t Bj = acc_total(a, G.rep0 + rj, 0);
                const double Ei = -a.J * (double)Bi;
                const double Ej = -a.J * (double)Bj;
                const double delta = (a.betas[G.rep0 + i] - a.betas[G.rep0 + i + 1]) * (Ej - Ei);
                const u32x4 w = philox((uint32_t)i, (uint32_t)round, stream_id((uint32_t)G.L, 0u),
                                       tag(P_EXCH, MODEL_ISING), a.k0, a.k1);
                const double u = u01(w.x, w.y);
                if (delta <= 0.0 || u < det_exp(-delta)) {
                    a.t2r[G.rep0 + i] = rj;
                    a.t2r[G.rep0 + i + 1] = ri;
                    a.r2t[G.rep0 + ri] = i + 1;
                    a.r2t[G.rep0 + rj] = i;
                    atomicAdd((unsigned long long*)&ctl->exch_accepts, 1ull);
                }
            }
            if (tid == 0) atomicAdd((unsigned long long*)&ctl->exch_tries, (unsigned long long)npairs);
        }
    }
    __syncthreads();
    if (need)
        for (int64_t i = tid; i < 2 * (int64_t)a.total_reps * a.acc_copies; i += blockDim.x) a.acc[i] = 0;
    if (tid == 0) {
        if (meas) {
            if (room) ctl->n_rec = nrec + 1;
            else ctl->overflow = 1;
        }
        ctl->t = t + 1;
        advance_flags(ctl);
    }
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Remap so each XCD walks one contiguous range of tiles: the z-neighbour
// planes a tile reads are then mostly in the same XCD's L2.
__device__ __forceinline__ int xcd_block(int nblocks) {
    const int per = (int)(gridDim.x >> 3);
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// One colour half-sweep of every replica of every group in the ensemble.
// Grid: one block per (group, replica, tile of 2048 half-sites), padded to a
// multiple of 8 blocks.
__global__ __launch_bounds__(kThreads) void ising_half_kernel(IsingArgs a, int color) {
    __shared__ uint32_t s_thr[28];
    __shared__ int64_t s_red[2][kThreads / 64];
    const int b = xcd_block(a.nblocks);
    if (b >= a.nblocks) return;
    const int g = find_group(a.groups, a.ngroups, b);
    const DevGroup G = a.groups[g];
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int tidx = a.r2t[grep];
    if (threadIdx.x < 28)
        s_thr[threadIdx.x] = a.inline_thr ? a.thr_inline[threadIdx.x] : a.thr[(int64_t)(G.rep0 + tidx) * 28 + threadIdx.x];
    const int64_t t = a.ctl->t;
    const bool need = (a.ctl->flags & (F_MEAS | F_EXCH)) != 0;
    __syncthreads();

    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);
    const int L = G.L, Lh = G.Lh;
    uint8_t* C = a.pool + G.off + (int64_t)(2 * r + color) * G.stride;
    const uint8_t* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * G.stride;

    int bsum = 0, msum = 0;
    for (int part = 0; part < G.spt; ++part) {
    const int64_t h0 = ((int64_t)tile * G.spt + part) * kTile + (int64_t)threadIdx.x * kSitesPerThread;
    if (h0 < G.n_half) {
        const u32x4 w = philox((uint32_t)(h0 >> 3), hs, stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
        if (G.fast) {
            // 8 sites in one row (Lh % 8 == 0): six 8-byte neighbour loads and one
            // edge byte for the shifted x-neighbour.  All index math in 32 bits
            // (L^3/2 < 2^32 for every L the library accepts).
            const uint32_t hh = (uint32_t)h0;
            const uint32_t row = fastdiv(hh, G.mag_Lh, G.sh_Lh);
            const uint32_t xh0 = hh - row * (uint32_t)Lh;
            const uint32_t z = fastdiv(row, G.mag_L, G.sh_L), y = row - z * (uint32_t)L;
            const uint32_t p = (y + z + (uint32_t)color) & 1u;
            const uint32_t ym = y == 0 ? L - 1 : y - 1, yp = y == (uint32_t)L - 1 ? 0 : y + 1;
            const uint32_t zm = z == 0 ? L - 1 : z - 1, zp = z == (uint32_t)L - 1 ? 0 : z + 1;
            const uint32_t rb = row * (uint32_t)Lh;
            const uint32_t xe = p ? (xh0 + 8 == (uint32_t)Lh ? 0u : xh0 + 8) : (xh0 == 0 ? (uint32_t)Lh - 1 : xh0 - 1);
            const uint64_t c8 = *(const uint64_t*)(C + hh);
            const uint64_t xr = *(const uint64_t*)(O + hh);
            const uint64_t e = O[rb + xe];
            const uint64_t nym = *(const uint64_t*)(O + (z * (uint32_t)L + ym) * (uint32_t)Lh + xh0);
            const uint64_t nyp = *(const uint64_t*)(O + (z * (uint32_t)L + yp) * (uint32_t)Lh + xh0);
            const uint64_t nzm = *(const uint64_t*)(O + (zm * (uint32_t)L + y) * (uint32_t)Lh + xh0);
            const uint64_t nzp = *(const uint64_t*)(O + (zp * (uint32_t)L + y) * (uint32_t)Lh + xh0);
            const uint64_t xs = p ? ((xr >> 8) | (e << 56)) : ((xr << 8) | e);
            const uint64_t ns = xr + xs + nym + nyp + nzm + nzp;  // per-byte up-neighbour counts 0..6
            // per-byte byte offset of the table entry k = 7 b + n (at most 52: no carries)
            const uint64_t K4 = (c8 * 7ull + ns) << 2;
            uint32_t flo = 0, fhi = 0;
            bool anytie = false;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const uint32_t kk = (uint32_t)(K4 >> (8 * j)) & 0xffu;
                const uint32_t thr = *(const uint32_t*)((const char*)s_thr + kk);
                const uint32_t r16 = (word_of(w, j) >> (16 * (j & 1))) & 0xffffu;
                const uint32_t acc = (r16 - thr) >> 31;  // r16 < thr (both below 2^17)
                if (j < 4) flo |= acc << (8 * j);
                else fhi |= acc << (8 * (j - 4));
                anytie |= (r16 == thr);
            }
            if (anytie) {
                // rare (p = 2^-16 per site): decide tied sites with 16 more bits
#pragma unroll 1
                for (int j = 0; j < 8; ++j) {
                    const uint32_t kk = (uint32_t)(K4 >> (8 * j)) & 0xffu;
                    const uint32_t thr = *(const uint32_t*)((const char*)s_thr + kk);
                    const uint32_t r16 = (word_of(w, j) >> (16 * (j & 1))) & 0xffffu;
                    if (r16 == thr) {
                        const uint32_t lo = *(const uint32_t*)((const char*)s_thr + 56 + kk);
                        const uint32_t acc = tie_bits(hh + (uint32_t)j, hs, stream, a.k0, a.k1) < lo ? 1u : 0u;
                        if (j < 4) flo = (flo & ~(1u << (8 * j))) | (acc << (8 * j));
                        else fhi = (fhi & ~(1u << (8 * (j - 4)))) | (acc << (8 * (j - 4)));
                    }
                }
            }
            const uint64_t nc8 = c8 ^ (flo | ((uint64_t)fhi << 32));
            *(uint64_t*)(C + hh) = nc8;
            if (need) {
                const int sb = __builtin_popcountll(nc8);
                msum += 2 * sb - 8;
                if (color) {
                    // sum_j (2 b'_j - 1)(2 n_j - 6) = 4 sum b'n - 12 sum b' - 2 sum n + 48
                    const uint64_t bn = (nc8 * 0xffull) & ns;
                    const int sbn = (int)((bn * 0x0101010101010101ull) >> 56);
                    const int sn = (int)((ns * 0x0101010101010101ull) >> 56);
                    bsum += 4 * sbn - 12 * sb - 2 * sn + 48;
                }
            }
        } else {
            // Generic even L: per-site addressing, same random-number assignment.
#pragma unroll 1
            for (int j = 0; j < 8; ++j) {
                const int64_t h = h0 + j;
                if (h >= G.n_half) break;
                const int64_t row = h / Lh;
                const int xh = (int)(h - row * Lh);
                const int z = (int)(row / L), y = (int)(row - (int64_t)z * L);
                const int p = (y + z + color) & 1;
                const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
                const int zm = z == 0 ? L - 1 : z - 1, zp = z == L - 1 ? 0 : z + 1;
                const int xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
                uint32_t n = (uint32_t)O[row * Lh + xh] + O[row * Lh + xo] + O[((int64_t)z * L + ym) * Lh + xh] +
                             O[((int64_t)z * L + yp) * Lh + xh];
                if (G.dims == 3) n += (uint32_t)O[((int64_t)zm * L + y) * Lh + xh] + O[((int64_t)zp * L + y) * Lh + xh];
                const uint32_t b = C[h] & 1u;
                const uint32_t k = b * (uint32_t)(2 * G.dims + 1) + n;  // table row: b * (z + 1) + n_up
                const uint32_t thi = s_thr[k];
                const uint32_t r16 = (word_of(w, j) >> (16 * (j & 1))) & 0xffffu;
                bool acc = r16 < thi;
                if (r16 == thi) acc = tie_bits((uint32_t)h, hs, stream, a.k0, a.k1) < s_thr[14 + k];
                const uint32_t nb = b ^ (uint32_t)acc;
                C[h] = (uint8_t)nb;
                if (need) {
                    const int sn = 2 * (int)nb - 1;
                    msum += sn;
                    if (color) bsum += sn * (2 * (int)n - 2 * G.dims);
                }
            }
        }
    }
    }

    // Fused measurement: wave -> block -> one int64 atomic per block and quantity.
    if (need) {
        const int wb = wave_sum(bsum), wm = wave_sum(msum);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) {
            s_red[0][wid] = wb;
            s_red[1][wid] = wm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t B = 0, M = 0;
            for (int i = 0; i < kThreads / 64; ++i) {
                B += s_red[0][i];
                M += s_red[1][i];
            }
            if (color) atomicAdd((unsigned long long*)&a.acc[2 * grep], (unsigned long long)B);
            atomicAdd((unsigned long long*)&a.acc[2 * grep + 1], (unsigned long long)M);
        }
    }

}

// Closes a sweep: records, replica exchange, sweep counter.  One block,
// launched after the colour-1 half-sweep (a kernel boundary orders it after
// every block's accumulator update).
__global__ __launch_bounds__(kThreads) void ising_close_kernel(IsingArgs a) {
    const int64_t t = a.ctl->t;
    const int32_t fl = a.ctl->flags;
    close_sweep(a, t, (fl & F_MEAS) != 0, (fl & F_EXCH) != 0, a.ctl->round);
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
        for (int j = 0; j < kSitesPerThread * G.spt; ++j) {
            const int64_t h = (int64_t)tile * kTile * G.spt + (int64_t)j * kThreads + threadIdx.x;
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
    for (int j = 0; j < kSitesPerThread * G.spt; ++j) {
        const int64_t h = (int64_t)tile * kTile * G.spt + (int64_t)j * kThreads + threadIdx.x;
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
