// The int8 Ising colour half-sweep of one block tile (the body of a colour
// launch), shared by the per-ensemble launch and the work-list launch of
// several ensembles.  Storage, random numbers and acceptance are described
// in ising.hip.
#pragma once
#include "sweepctl.h"

namespace ising {

constexpr int kThreads = 256;
constexpr int kSitesPerThread = 16;                // sites per Philox call (and per generic thread step)
static_assert(kThreads * kSitesPerThread == kIsingTile, "host tiling must match the kernel");

__device__ __forceinline__ int find_group(const DevGroup* groups, int ngroups, int b) {
    int g = 0;
    while (g + 1 < ngroups && b >= groups[g + 1].blk0) ++g;
    return g;
}

// The 25 low bits of a tied site's uniform: top 25 bits of its own Philox call.
__device__ __forceinline__ uint32_t tie_bits(uint32_t h, uint32_t hs, uint32_t stream, uint32_t k0, uint32_t k1) {
    return philox(h, hs, stream, tag(P_TIE, MODEL_ISING), k0, k1).x >> 7;
}

__device__ __forceinline__ uint32_t word_of(const u32x4& w, int d) {
    return d == 0 ? w.x : (d == 1 ? w.y : (d == 2 ? w.z : w.w));
}

// 24-bit multiply (one full-rate instruction); both operands are below 2^24
// for every lattice the library accepts (L <= 2046, so L^2 < 2^22).
__device__ __forceinline__ uint32_t mul24(uint32_t x, uint32_t y) { return (x & 0xffffffu) * (y & 0xffffffu); }

// Spin bytes c[] and up-neighbour counts ns[] of the ND words (4 * ND sites)
// starting at half index hh, all in one row: one ND-word load per neighbour
// row plus one edge byte for the x-neighbour shifted across words.  Offsets
// are 32-bit (L^3/2 < 2^32), added to the uniform base pointers.
template <int ND>
__device__ __forceinline__ void gather_row(const DevGroup& G, const uint8_t* __restrict__ C,
                                           const uint8_t* __restrict__ O, uint32_t hh, int color, uint32_t* c,
                                           uint32_t* ns) {
    typedef uint32_t vec __attribute__((ext_vector_type(ND)));
    const uint32_t L = (uint32_t)G.L, Lh = (uint32_t)G.Lh;
    const uint32_t row = fastdiv(hh, G.mag_Lh, G.sh_Lh);
    const uint32_t xh0 = hh - mul24(row, Lh);
    const uint32_t z = fastdiv(row, G.mag_L, G.sh_L), y = row - mul24(z, L);
    const uint32_t p = (y + z + (uint32_t)color) & 1u;
    const uint32_t ym = (y == 0 ? L : y) - 1, yp = y + 1 == L ? 0 : y + 1;
    const uint32_t zm = (z == 0 ? L : z) - 1, zp = z + 1 == L ? 0 : z + 1;
    const uint32_t zL = mul24(z, L);
    uint32_t xf = xh0 + 4 * ND;
    xf = xf == Lh ? 0u : xf;
    const uint32_t xb = (xh0 == 0 ? Lh : xh0) - 1;
    const uint32_t xe = p ? xf : xb;
    const vec vc = *(const vec*)(C + hh);
    const vec vx = *(const vec*)(O + hh);
    const vec v1 = *(const vec*)(O + (mul24(zL + ym, Lh) + xh0));
    const vec v2 = *(const vec*)(O + (mul24(zL + yp, Lh) + xh0));
    const vec v3 = *(const vec*)(O + (mul24(mul24(zm, L) + y, Lh) + xh0));
    const vec v4 = *(const vec*)(O + (mul24(mul24(zp, L) + y, Lh) + xh0));
    const uint32_t e = O[(hh - xh0) + xe];
    // x-neighbour bytes: shifted down one byte (p = 1, neighbour at xh + 1) or
    // up one byte (p = 0, neighbour at xh - 1), the edge byte filling the gap;
    // both shifts are formed and the row parity picks one, without branches
    const uint32_t pm = 0u - p;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const uint32_t up = __builtin_amdgcn_alignbyte(d + 1 < ND ? vx[d + 1 < ND ? d + 1 : 0] : e, vx[d], 1u);
        const uint32_t dn = __builtin_amdgcn_alignbyte(vx[d], d > 0 ? vx[d > 0 ? d - 1 : 0] : e << 24, 3u);
        const uint32_t xs = (up & pm) | (dn & ~pm);
        c[d] = vc[d];
        ns[d] = vx[d] + xs + v1[d] + v2[d] + v3[d] + v4[d];  // per byte 0..6: no carries
    }
}

// Decides one tied site of a 3D chunk layout (r7 == hi): recounts its up
// neighbours, draws the 25 low bits and flips it when u < T.  The chunk's
// provisional value (no flip) is what C[h] holds.
__device__ __forceinline__ void resolve_tie(const IsingArgs& a, const DevGroup& G, uint8_t* C, const uint8_t* O,
                                            const uint32_t* s_thr, uint32_t h, int color, uint32_t hs,
                                            uint32_t stream, bool need, int* msum, int* bsum) {
    const uint32_t L = (uint32_t)G.L, Lh = (uint32_t)G.Lh;
    const uint32_t row = fastdiv(h, G.mag_Lh, G.sh_Lh);
    const uint32_t xh = h - mul24(row, Lh);
    const uint32_t z = fastdiv(row, G.mag_L, G.sh_L), y = row - mul24(z, L);
    const uint32_t p = (y + z + (uint32_t)color) & 1u;
    const uint32_t ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
    const uint32_t zm = z == 0 ? L - 1 : z - 1, zp = z == L - 1 ? 0 : z + 1;
    const uint32_t xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
    const uint32_t rb = h - xh;
    const uint32_t n = (uint32_t)O[rb + xh] + O[rb + xo] + O[mul24(mul24(z, L) + ym, Lh) + xh] +
                       O[mul24(mul24(z, L) + yp, Lh) + xh] + O[mul24(mul24(zm, L) + y, Lh) + xh] +
                       O[mul24(mul24(zp, L) + y, Lh) + xh];
    const uint32_t b = C[h];
    const uint32_t k = b * 7u + n;
    if (tie_bits(h, hs, stream, a.k0, a.k1) < s_thr[14 + k]) {
        C[h] = (uint8_t)(b ^ 1u);
        if (need) {
            const int ds = b ? -2 : 2;
            *msum += ds;
            if (color) *bsum += ds * (2 * (int)n - 6);
        }
    }
}

// Block queue of tied thread steps (LDS): step base and tie mask.  Bit
// 8j + 7 - d marks byte j of word d of the step; word d lies in run d / RW at
// offset 4 (d % RW).  A lane queues at most one entry per step, so the queue
// holds kThreads * (steps per tile) entries and never overflows.
constexpr int kTieCap = kThreads * 4;

// Decides every queued tie of the block (all threads; after the tile).
template <int RW, uint32_t RS>
__device__ __forceinline__ void flush_ties(const IsingArgs& a, const DevGroup& G, uint8_t* C, const uint8_t* O,
                                           const uint32_t* s_thr, int color, uint32_t hs, uint32_t stream, bool need,
                                           const uint32_t* qh, const uint32_t* qm, int n, int* msum, int* bsum) {
    for (int e = threadIdx.x; e < n; e += kThreads) {
        const uint32_t h0 = qh[e];
        uint32_t m = qm[e];
        while (m) {
            const uint32_t pb = (uint32_t)__builtin_ctz(m);
            m &= m - 1u;
            const uint32_t d = 7u - (pb & 7u);
            const uint32_t site = h0 + (d / RW) * RS + 4u * (d % RW) + (pb >> 3);
            resolve_tie(a, G, C, O, s_thr, site, color, hs, stream, need, msum, bsum);
        }
    }
}

// Thread steps of one block tile for a 3D lattice with Lh % 8 == 0.  A step
// is NR runs of RW 4-site words, each run within one row; runs are RS
// half-sites apart.  Layouts: (RW 2, NR 2, RS 8) two adjacent 8-site runs;
// (4, 1) 16 sites; (4, 2, RS 256 * 16) two 16-site runs a block row apart;
// (8, 1) 32 sites.  Per 4-site word: thresholds by byte permutes of the
// packed table, r7 < hi by the sign bit of (r | 0x80) - hi, and ties (byte
// exactly 0x80) queued for the block.
template <int RW, int NR, uint32_t RS>
__device__ __forceinline__ int fast_tile(const IsingArgs& a, const DevGroup& G, int tile, uint8_t* C,
                                         const uint8_t* O, const uint32_t* s_thr, int color, uint32_t hs,
                                         uint32_t stream, bool need, uint32_t* qh, uint32_t* qm, int* qn, int* msum,
                                         int* bsum) {
    constexpr int NW = RW * NR;                          // 32-bit words per thread step (<= 8)
    constexpr uint32_t TS = RS == 8u ? 16u : 4u * RW;   // half-sites a thread step advances per thread
    constexpr uint32_t SPAN = kThreads * TS * (RS == 8u ? 1u : NR);  // half-sites per block step
    static_assert(NW <= 8, "tie masks hold 8 words");
    const uint32_t T0l = s_thr[28], T0h = s_thr[29], T1l = s_thr[30], T1h = s_thr[31];
    for (int part = 0; part < G.spt; ++part) {
        const int64_t h0 = (int64_t)tile * G.tsz + (int64_t)part * SPAN + (int64_t)threadIdx.x * TS;
        uint32_t tie = 0;
        if (h0 < G.n_half) {
            const uint32_t hh = (uint32_t)h0;  // L^3/2 < 2^32 for every accepted L
            uint32_t c[NW], ns[NW], rw[NW];
#pragma unroll
            for (int k = 0; k < NR; ++k) gather_row<RW>(G, C, O, hh + (uint32_t)k * RS, color, c + k * RW, ns + k * RW);
#pragma unroll
            for (int k = 0; k < NW / 4; ++k) {
                // words 4k..4k+3 are the 16-site group of their first site
                const uint32_t site = hh + (uint32_t)((4 * k) / RW) * RS + 4u * (uint32_t)((4 * k) % RW);
                const u32x4 w = philox(site >> 4, hs, stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
                rw[4 * k] = w.x;
                rw[4 * k + 1] = w.y;
                rw[4 * k + 2] = w.z;
                rw[4 * k + 3] = w.w;
            }
            uint32_t nc[NW];
#pragma unroll
            for (int d = 0; d < NW; ++d) {
                const uint32_t m = __builtin_amdgcn_perm(0u, 0x0000ff00u, c[d]);  // 0xff in up-spin bytes
                const uint32_t t1 = __builtin_amdgcn_perm(T1h, T1l, ns[d]);
                const uint32_t t0 = __builtin_amdgcn_perm(T0h, T0l, ns[d]);
                const uint32_t thr = (m & t1) | (~m & t0);
                // per byte (0x80 + r7) - hi >= 0: sign bit clear iff r7 < hi
                const uint32_t zz = (rw[d] | 0x80808080u) - thr;
                nc[d] = c[d] ^ (((zz >> 7) & 0x01010101u) ^ 0x01010101u);
                // tie iff the byte is exactly 0x80 (r7 == hi)
                const uint32_t u = (zz & 0x7f7f7f7fu) + 0x7f7f7f7fu;
                tie |= (zz & ~u & 0x80808080u) >> d;
            }
            typedef uint32_t vecr __attribute__((ext_vector_type(RW)));
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                vecr o;
#pragma unroll
                for (int d = 0; d < RW; ++d) o[d] = nc[k * RW + d];
                *(vecr*)(C + hh + (uint32_t)k * RS) = o;
            }
            if (need) {
                int sb = 0;
#pragma unroll
                for (int d = 0; d < NW; ++d) sb += __builtin_popcount(nc[d]);
                *msum += 2 * sb - 4 * NW;
                if (color) {
                    // sum_j (2 b'_j - 1)(2 n_j - 6) = 4 sum b'n - 12 sum b' - 2 sum n + 6 (4 NW)
                    uint32_t bn = 0, sn = 0;  // byte sums stay below 256 (at most 8 words of 6)
#pragma unroll
                    for (int d = 0; d < NW; ++d) {
                        bn += __builtin_amdgcn_perm(0u, 0x0000ff00u, nc[d]) & ns[d];
                        sn += ns[d];
                    }
                    const int sbn = (int)__builtin_amdgcn_sad_u8(bn, 0u, 0u);
                    const int snn = (int)__builtin_amdgcn_sad_u8(sn, 0u, 0u);
                    *bsum += 4 * sbn - 12 * sb - 2 * snn + 24 * NW;
                }
            }
        }
        // queue this lane's tied step: one LDS add per wave, then lane slots
        const uint64_t bal = __ballot(tie != 0u);
        if (bal) {
            int base = 0;
            if ((threadIdx.x & 63) == 0) base = atomicAdd(qn, __popcll(bal));
            base = __shfl(base, 0);
            if (tie) {
                const uint32_t pos = (uint32_t)base + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                qh[pos] = (uint32_t)h0;
                qm[pos] = tie;
            }
        }
    }
    __threadfence_block();  // this block's step stores are done before tied bytes are rewritten
    __syncthreads();
    const int n = *qn;
    flush_ties<RW, RS>(a, G, C, O, s_thr, color, hs, stream, need, qh, qm, n, msum, bsum);
    return n;
}

// One colour half-sweep of work block b (already in the XCD-aware order,
// b < a.nblocks) of size G: the tile of one (size, replica), closing the
// previous sweep in colour 0 (sweepctl.h).  All threads of a 256-thread block.
// The caller finds G: a dynamic index into the kernel arguments' descriptors
// stays a scalar load only in the kernel itself.
__device__ __forceinline__ void ising_half_block(const IsingArgs& a, const int b, const int color, const DevGroup G) {
    __shared__ uint32_t s_thr[kIsingTab];
    __shared__ int s_red[2];  // block sums of the bond sum and magnetisation
    __shared__ int s_waves;    // waves done with the block sums
    __shared__ __align__(8) uint32_t s_qh[kTieCap];
    __shared__ uint32_t s_qm[kTieCap];
    __shared__ int s_qn;
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    // issued before the control block is read: the table by replica is right
    // unless this launch closes an exchange
    uint32_t thr_pre = 0u;
    if (a.thr_rep && threadIdx.x < kIsingTab) thr_pre = a.thr_rep[(int64_t)grep * kIsingTab + threadIdx.x];
    const SweepView v = view_of(a.ctl, color);
    int tidx = 0;
    if (v.exch) {
        const int64_t po = (v.round & 1) * a.total_reps;
        const int64_t* pacc = acc_of(a, v.t - 1);
        if (G.R <= kThreads) {
            // one round of loads (the group's permutation and bond sums, into the
            // tie queue's LDS, unused until the barrier below) instead of the
            // dependent chain replica -> temperature -> partner -> bond sums
            int64_t* s_bond = reinterpret_cast<int64_t*>(s_qh);
            int32_t* s_t2r = reinterpret_cast<int32_t*>(s_qm);
            int i = 0;
            if (threadIdx.x < kIsingTab) i = a.r2t[po + grep];
            if ((int)threadIdx.x < G.R) {
                s_bond[threadIdx.x] = acc_total(a, pacc, G.rep0 + (int)threadIdx.x, 0);
                s_t2r[threadIdx.x] = a.t2r[po + G.rep0 + threadIdx.x];
            }
            __syncthreads();
            if (threadIdx.x < kIsingTab) tidx = exch_tidx_gathered(a, G, i, v.round, s_t2r, s_bond);
        } else {
            tidx = exch_tidx(a, G, r, v.round, a.r2t + po, a.t2r + po, pacc);
        }
        if (b == G.blk0) exch_write_group(a, G, v.round, pacc);
    } else if (!a.thr_rep) {
        tidx = a.r2t[(int64_t)v.par * a.total_reps + grep];
    }
    if (b == 0) sweep_bookkeeping(a, v, color);
    if (threadIdx.x < kIsingTab)
        s_thr[threadIdx.x] = a.sched                  ? a.sched[(v.t - a.sched_t0) * kIsingTab + threadIdx.x]
                             : a.inline_thr           ? a.thr_inline[threadIdx.x]
                             : (a.thr_rep && !v.exch) ? thr_pre
                                                      : a.thr[(int64_t)(G.rep0 + tidx) * kIsingTab + threadIdx.x];
    if (threadIdx.x == 0) {
        s_qn = 0;
        s_red[0] = s_red[1] = 0;
        s_waves = 0;
    }
    const int64_t t = v.t;
    const bool need = (v.flags & (F_MEAS | F_EXCH)) != 0;
    __syncthreads();

    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);
    uint8_t* C = a.pool + G.off + (int64_t)(2 * r + color) * G.stride;
    const uint8_t* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * G.stride;
    const int lane = threadIdx.x & 63;

    int bsum = 0, msum = 0;
    if (G.fast == 3) {
        fast_tile<8, 1, 0u>(a, G, tile, C, O, s_thr, color, hs, stream, need, s_qh, s_qm, &s_qn, &msum, &bsum);
    } else if (G.fast == 2 && G.runs == 2) {
        fast_tile<4, 2, kThreads * 16u>(a, G, tile, C, O, s_thr, color, hs, stream, need, s_qh, s_qm, &s_qn, &msum,
                                        &bsum);
    } else if (G.fast == 2) {
        fast_tile<4, 1, 0u>(a, G, tile, C, O, s_thr, color, hs, stream, need, s_qh, s_qm, &s_qn, &msum, &bsum);
    } else if (G.fast == 1) {
        fast_tile<2, 2, 8u>(a, G, tile, C, O, s_thr, color, hs, stream, need, s_qh, s_qm, &s_qn, &msum, &bsum);
    } else {
        // Per-site addressing (Lh % 8 != 0, and 2D): same random-number assignment.
        const int64_t h0 = (int64_t)tile * G.tsz + (int64_t)threadIdx.x * kSitesPerThread;
        if (h0 < G.n_half) {
            const int L = G.L, Lh = G.Lh;
            const u32x4 w = philox((uint32_t)(h0 >> 4), hs, stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
#pragma unroll 4
            for (int j = 0; j < kSitesPerThread; ++j) {
                const int64_t h = h0 + j;
                if (h >= G.n_half) break;
                const int64_t row = Lh >= 2 ? (int64_t)fastdiv((uint32_t)h, G.mag_Lh, G.sh_Lh) : h;
                const int xh = (int)(h - row * Lh);
                const int z = (int)fastdiv((uint32_t)row, G.mag_L, G.sh_L), y = (int)(row - (int64_t)z * L);
                const int p = (y + z + color) & 1;
                const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
                const int zm = z == 0 ? L - 1 : z - 1, zp = z == L - 1 ? 0 : z + 1;
                const int xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
                uint32_t n = (uint32_t)O[row * Lh + xh] + O[row * Lh + xo] + O[((int64_t)z * L + ym) * Lh + xh] +
                             O[((int64_t)z * L + yp) * Lh + xh];
                if (G.dims == 3) n += (uint32_t)O[((int64_t)zm * L + y) * Lh + xh] + O[((int64_t)zp * L + y) * Lh + xh];
                const uint32_t bb = C[h] & 1u;
                const uint32_t k = bb * (uint32_t)(2 * G.dims + 1) + n;  // table row: b * (z + 1) + n_up
                const uint32_t thi = s_thr[k];
                const uint32_t r7 = (word_of(w, j >> 2) >> (8 * (j & 3))) & 0x7fu;
                bool acc = r7 < thi;
                if (r7 == thi) acc = tie_bits((uint32_t)h, hs, stream, a.k0, a.k1) < s_thr[14 + k];
                const uint32_t nb = bb ^ (uint32_t)acc;
                C[h] = (uint8_t)nb;
                if (need) {
                    const int sn = 2 * (int)nb - 1;
                    msum += sn;
                    if (color) bsum += sn * (2 * (int)n - 2 * G.dims);
                }
            }
        }
    }

    // Fused measurement: wave -> block -> one int64 atomic per block and
    // quantity.  No barrier: each wave adds its sums in LDS and the wave that
    // finishes last (workgroup-scope acquire-release counter) issues the
    // atomics, so waves without tied sites need not wait for the tie pass.
    if (need) {
        const int wb = wave_total_last(bsum), wm = wave_total_last(msum);
        if (lane == 63) {
            atomicAdd(&s_red[0], wb);
            atomicAdd(&s_red[1], wm);
            const int done = __hip_atomic_fetch_add(&s_waves, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (done == kThreads / 64 - 1) {
                const int64_t B = __hip_atomic_load(&s_red[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const int64_t M = __hip_atomic_load(&s_red[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                // accumulator copy by block (copies are summed where read): fewer
                // atomics per address than one copy taking every tile of a replica
                int64_t* acc = acc_of(a, t) + 2 * (int64_t)(b & (a.acc_copies - 1)) * a.total_reps;
                if (color) atomicAdd((unsigned long long*)&acc[2 * grep], (unsigned long long)B);
                atomicAdd((unsigned long long*)&acc[2 * grep + 1], (unsigned long long)M);
            }
        }
    }
}

}  // namespace ising
