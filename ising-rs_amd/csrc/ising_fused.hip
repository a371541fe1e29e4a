// One launch per int8 Ising sweep (both colours), gfx950.
//
// A block owns a box of one replica: every x row of a (y, z) square of B x B
// rows, plus a halo of two rows on each side, staged in LDS (both colour
// half-lattices).  It updates colour 0 on the box and the first halo ring,
// then colour 1 on the box alone, and writes the box back.  The halo's colour-0
// values are recomputed exactly as the box owning them computes them (same
// Philox counters, same inputs), so the sweep equals the two-launch sweep bit
// for bit; lattices with L <= 32 are staged whole (one box per replica, no
// halo, periodic in LDS).  With no grid-wide dependency between the colours
// the sweep needs one launch instead of two: the closing of the previous
// sweep (records, replica exchange) is done at the start as in a colour-0
// launch, and the block that finishes last writes the next sweep's control
// fields (a relaxed counter; every other block has read them by then).  Boxes
// read one lattice buffer and write the other (a neighbour's halo rows must
// be the old sweep's); the host alternates the two and copies back after an
// odd number of sweeps, so between calls the lattice is where it always is.
//
// Bytes per site and sweep: (B + 4)^2 / B^2 read + 1 written, against 2 + 1 for
// two launches.
#include "sweepctl.h"

namespace ising {

namespace {

constexpr int kFT = 256;  // threads per block

__device__ __forceinline__ uint32_t tie_draw(uint32_t h, uint32_t hs, uint32_t stream, uint32_t k0, uint32_t k1) {
    return philox(h, hs, stream, tag(P_TIE, MODEL_ISING), k0, k1).x >> 7;
}

__device__ __forceinline__ uint32_t word4(const u32x4& w, int d) {
    return d == 0 ? w.x : (d == 1 ? w.y : (d == 2 ? w.z : w.w));
}

// Box geometry of one block.
struct Box {
    int BR;      // LDS rows per axis (y and z)
    int off;     // LDS row = global row - (origin - off)
    int y0, z0;  // global origin of the box
    bool whole;  // the whole replica (periodic in LDS)
    int lo0, hi0, lo1, hi1;  // LDS row ranges updated in colour 0 / colour 1
};

__device__ __forceinline__ Box box_of(const DevGroup& G, int box) {
    Box X;
    X.whole = G.fbox == 0;
    if (X.whole) {
        X.BR = G.L;
        X.off = 0;
        X.y0 = X.z0 = 0;
        X.lo0 = X.lo1 = 0;
        X.hi0 = X.hi1 = G.L;
    } else {
        const int nb = G.L / G.fbox;
        X.BR = G.fbox + 4;
        X.off = 2;
        X.y0 = (box % nb) * G.fbox;
        X.z0 = (box / nb) * G.fbox;
        X.lo0 = 1;
        X.hi0 = X.BR - 1;
        X.lo1 = 2;
        X.hi1 = X.BR - 2;
    }
    return X;
}

__device__ __forceinline__ int wrap(int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); }

// One colour pass over LDS rows [lo, hi)^2 of the box: the 16-site chunks of
// half-index (one Philox call each), in place in LDS.
__device__ __forceinline__ void colour_pass(const IsingArgs& a, const DevGroup& G, const Box& X, uint8_t* sb,
                                            const uint32_t* s_thr, int color, int lo, int hi, uint32_t hs,
                                            uint32_t stream, bool need, int* bsum) {
    const int L = G.L, Lh = G.Lh, BR = X.BR;
    uint8_t* Cs = sb + (size_t)color * BR * BR * Lh;
    const uint8_t* Os = sb + (size_t)(color ^ 1) * BR * BR * Lh;
    const int nr = hi - lo;
    const int cpr = Lh >= 16 ? Lh / 16 : 0;  // chunks per row (rows of 16 or more half-sites)
    const int nchunks = cpr ? nr * nr * cpr : (nr * nr * Lh) / 16;
    for (int k = threadIdx.x; k < nchunks; k += kFT) {
        int ly0, lz0, xh0;
        if (cpr) {
            const int row = k / cpr;
            xh0 = (k - row * cpr) * 16;
            ly0 = lo + row % nr;
            lz0 = lo + row / nr;
        } else {  // whole lattice with rows shorter than 16: LDS index = half index
            const int h = k * 16;
            const int row = h / Lh;
            xh0 = h - row * Lh;
            ly0 = row % L;
            lz0 = row / L;
        }
        const int yg0 = X.whole ? ly0 : wrap(X.y0 - X.off + ly0, L);
        const int zg0 = X.whole ? lz0 : wrap(X.z0 - X.off + lz0, L);
        const uint32_t h0 = (uint32_t)(((zg0 * L) + yg0) * Lh + xh0);
        const u32x4 w = philox(h0 >> 4, hs, stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
        for (int j = 0; j < 16; ++j) {
            int ly = ly0, lz = lz0, xh = xh0 + j;
            if (!cpr && xh >= Lh) {  // the chunk runs on into the next rows
                const int rr = xh / Lh;
                xh -= rr * Lh;
                ly = ly0 + rr;
                lz = lz0;
                while (ly >= L) {
                    ly -= L;
                    ++lz;
                }
            }
            const int yg = X.whole ? ly : wrap(X.y0 - X.off + ly, L);
            const int zg = X.whole ? lz : wrap(X.z0 - X.off + lz, L);
            const int p = (yg + zg + color) & 1;
            const int xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
            const int lym = X.whole ? wrap(ly - 1, L) : ly - 1, lyp = X.whole ? wrap(ly + 1, L) : ly + 1;
            const int lzm = X.whole ? wrap(lz - 1, L) : lz - 1, lzp = X.whole ? wrap(lz + 1, L) : lz + 1;
            const int rb = (lz * BR + ly) * Lh;
            const uint32_t n = (uint32_t)Os[rb + xh] + Os[rb + xo] + Os[(lz * BR + lym) * Lh + xh] +
                               Os[(lz * BR + lyp) * Lh + xh] + Os[(lzm * BR + ly) * Lh + xh] +
                               Os[(lzp * BR + ly) * Lh + xh];
            const uint32_t b = Cs[rb + xh];
            const uint32_t kk = b * 7u + n;
            const uint32_t thi = s_thr[kk];
            const uint32_t r7 = (word4(w, j >> 2) >> (8 * (j & 3))) & 0x7fu;
            bool acc = r7 < thi;
            if (r7 == thi) {
                const uint32_t h = (uint32_t)(((zg * L) + yg) * Lh + xh);
                acc = tie_draw(h, hs, stream, a.k0, a.k1) < s_thr[14 + kk];
            }
            const uint32_t nb = b ^ (uint32_t)acc;
            Cs[rb + xh] = (uint8_t)nb;
            if (need && color) *bsum += (2 * (int)nb - 1) * (2 * (int)n - 6);
        }
    }
}

// The same pass, four sites per 32-bit word (rows of 16 or more half-sites):
// each thread takes 16-site chunks of one row, counts up neighbours per byte
// from five row words and the x-shifted word (edge byte across the chunk),
// looks the thresholds up by byte permutes of the packed table and compares
// all four 7-bit draws at once; a draw equal to its threshold's high part (a
// byte of exactly 0x80 after the subtraction) is decided on the spot with the
// site's 25 further bits.
__device__ __forceinline__ void colour_pass_words(const IsingArgs& a, const DevGroup& G, const Box& X, uint8_t* sb,
                                                  const uint32_t* s_thr, int color, int lo, int hi, uint32_t hs,
                                                  uint32_t stream, bool need, int* bsum) {
    const int L = G.L, Lh = G.Lh, BR = X.BR;
    uint8_t* Cs = sb + (size_t)color * BR * BR * Lh;
    const uint8_t* Os = sb + (size_t)(color ^ 1) * BR * BR * Lh;
    const int nr = hi - lo;
    const int cpr = Lh / 16;
    const int nchunks = nr * nr * cpr;
    const uint32_t T0l = s_thr[28], T0h = s_thr[29], T1l = s_thr[30], T1h = s_thr[31];
    for (int k = threadIdx.x; k < nchunks; k += kFT) {
        const int row = k / cpr;
        const int xh0 = (k - row * cpr) * 16;
        const int ly = lo + row % nr, lz = lo + row / nr;
        const int yg = X.whole ? ly : wrap(X.y0 - X.off + ly, L);
        const int zg = X.whole ? lz : wrap(X.z0 - X.off + lz, L);
        const int p = (yg + zg + color) & 1;
        const int lym = X.whole ? wrap(ly - 1, L) : ly - 1, lyp = X.whole ? wrap(ly + 1, L) : ly + 1;
        const int lzm = X.whole ? wrap(lz - 1, L) : lz - 1, lzp = X.whole ? wrap(lz + 1, L) : lz + 1;
        const int rb = (lz * BR + ly) * Lh;
        typedef uint32_t v4u __attribute__((ext_vector_type(4)));
        const v4u vc = *(const v4u*)(Cs + rb + xh0);
        const v4u vx = *(const v4u*)(Os + rb + xh0);
        const v4u v1 = *(const v4u*)(Os + (lz * BR + lym) * Lh + xh0);
        const v4u v2 = *(const v4u*)(Os + (lz * BR + lyp) * Lh + xh0);
        const v4u v3 = *(const v4u*)(Os + (lzm * BR + ly) * Lh + xh0);
        const v4u v4 = *(const v4u*)(Os + (lzp * BR + ly) * Lh + xh0);
        const uint32_t xe = p ? (xh0 + 16 == Lh ? 0u : (uint32_t)xh0 + 16u) : (xh0 == 0 ? (uint32_t)Lh - 1u : (uint32_t)xh0 - 1u);
        const uint32_t e = Os[rb + xe];
        const uint32_t pm = 0u - (uint32_t)p;
        const uint32_t h0 = (uint32_t)(((zg * L) + yg) * Lh + xh0);
        const u32x4 w = philox(h0 >> 4, hs, stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
        const uint32_t rw[4] = {w.x, w.y, w.z, w.w};
        v4u nc;
        uint32_t bn = 0, sn = 0;
        int sb = 0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const uint32_t up = __builtin_amdgcn_alignbyte(d < 3 ? vx[d < 3 ? d + 1 : 0] : e, vx[d], 1u);
            const uint32_t dn = __builtin_amdgcn_alignbyte(vx[d], d > 0 ? vx[d > 0 ? d - 1 : 0] : e << 24, 3u);
            const uint32_t xs = (up & pm) | (dn & ~pm);
            const uint32_t ns = vx[d] + xs + v1[d] + v2[d] + v3[d] + v4[d];  // per byte 0..6
            const uint32_t c = vc[d];
            const uint32_t m = __builtin_amdgcn_perm(0u, 0x0000ff00u, c);   // 0xff in up-spin bytes
            const uint32_t thr = (m & __builtin_amdgcn_perm(T1h, T1l, ns)) | (~m & __builtin_amdgcn_perm(T0h, T0l, ns));
            const uint32_t zz = (rw[d] | 0x80808080u) - thr;
            uint32_t n2 = c ^ (((zz >> 7) & 0x01010101u) ^ 0x01010101u);
            const uint32_t uu = (zz & 0x7f7f7f7fu) + 0x7f7f7f7fu;
            uint32_t tie = zz & ~uu & 0x80808080u;
            while (tie) {
                const uint32_t bit = (uint32_t)__builtin_ctz(tie);
                tie &= tie - 1u;
                const uint32_t j = bit >> 3;
                const uint32_t bb = (c >> (8 * j)) & 1u, nn = (ns >> (8 * j)) & 7u;
                const uint32_t kk = bb * 7u + nn;
                const uint32_t hh = h0 + 4u * (uint32_t)d + j;
                // the provisional byte keeps the spin; the 25 low bits decide
                if (tie_draw(hh, hs, stream, a.k0, a.k1) < s_thr[14 + kk]) n2 ^= 1u << (8 * j);
            }
            nc[d] = n2;
            if (need && color) {
                sb += __builtin_popcount(n2);
                bn += __builtin_amdgcn_perm(0u, 0x0000ff00u, n2) & ns;
                sn += ns;
            }
        }
        *(v4u*)(Cs + rb + xh0) = nc;
        if (need && color)
            *bsum += 4 * (int)__builtin_amdgcn_sad_u8(bn, 0u, 0u) - 12 * sb - 2 * (int)__builtin_amdgcn_sad_u8(sn, 0u, 0u) +
                     96;
    }
}

__device__ __forceinline__ uint32_t udiv(uint32_t n, uint32_t d, uint32_t m, uint32_t sh) {
    return d == 1u ? n : fastdiv(n, m, sh);
}

// Box <-> lattice copies in VB-byte units (VB divides Lh).  load: the box with
// its halo (both colours) into LDS; else the box rows back to the lattice,
// returning the sum of their spins (+-1).
template <int VB>
__device__ __forceinline__ int copy_box(uint8_t* P0, const DevGroup& G, const Box& X, uint8_t* sb, bool load) {
    typedef uint32_t vec __attribute__((ext_vector_type(VB / 4)));
    const int L = G.L, Lh = G.Lh, BR = X.BR;
    const int nr = load ? BR : X.hi1 - X.lo1, base = load ? 0 : X.lo1;
    const uint32_t upr = (uint32_t)(Lh / VB), nr2 = (uint32_t)(nr * nr);
    uint32_t mu = 0, su = 0, m2 = 0, s2 = 0, mr = 0, sr = 0;
    if (upr > 1u) fastdiv_magic(upr, &mu, &su);
    fastdiv_magic(nr2, &m2, &s2);
    fastdiv_magic((uint32_t)nr, &mr, &sr);
    const uint32_t units = 2u * nr2 * upr;
    int msum = 0;
    for (uint32_t i = threadIdx.x; i < units; i += kFT) {
        const uint32_t rowa = udiv(i, upr, mu, su), u = i - rowa * upr;
        const uint32_t c = udiv(rowa, nr2, m2, s2), rr = rowa - c * nr2;
        const uint32_t lzr = udiv(rr, (uint32_t)nr, mr, sr);
        const int lz = base + (int)lzr, ly = base + (int)(rr - lzr * (uint32_t)nr);
        const int yg = wrap(X.y0 - X.off + ly, L), zg = wrap(X.z0 - X.off + lz, L);
        vec* gp = (vec*)(P0 + c * G.stride + ((int64_t)zg * L + yg) * Lh) + u;
        vec* lp = (vec*)(sb + ((c * BR + lz) * BR + ly) * Lh) + u;
        if (load) {
            *lp = *gp;
        } else {
            const vec v = *lp;
            *gp = v;
#pragma unroll
            for (int k = 0; k < VB / 4; ++k) msum += 2 * __builtin_popcount(v[k]) - 4;
        }
    }
    return msum;
}

}  // namespace

__global__ __launch_bounds__(kFT) void ising_fused_kernel(IsingArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t s_box[];
    __shared__ uint32_t s_thr[kIsingTab];
    __shared__ int64_t s_red[2][kFT / 64];
    __shared__ int s_last;
    const int per = (int)(gridDim.x >> 3);
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (b >= a.fnblocks) return;
    int g = 0;
    if (a.gk_on) {
        while (g + 1 < a.ngroups && b >= a.gk[g + 1].fblk0) ++g;
    } else {
        while (g + 1 < a.ngroups && b >= a.groups[g + 1].fblk0) ++g;
    }
    const DevGroup G = a.gk_on ? a.gk[g] : a.groups[g];
    const int local = b - G.fblk0;
    const int r = local / G.fboxes, box = local - r * G.fboxes;
    const int grep = G.rep0 + r;
    uint32_t thr_pre = 0u;
    if (a.thr_rep && threadIdx.x < kIsingTab) thr_pre = a.thr_rep[(int64_t)grep * kIsingTab + threadIdx.x];
    const SweepView v = view_of(a.ctl, 0);
    int tidx = 0;
    if (v.exch) {
        const int64_t po = (v.round & 1) * a.total_reps;
        tidx = exch_tidx(a, G, r, v.round, a.r2t + po, a.t2r + po, acc_of(a, v.t - 1));
        if (local == 0) exch_write_group(a, G, v.round, acc_of(a, v.t - 1));
    } else if (!a.thr_rep) {
        tidx = a.r2t[(int64_t)v.par * a.total_reps + grep];
    }
    if (b == 0) sweep_bookkeeping(a, v, 0);
    if (threadIdx.x < kIsingTab)
        s_thr[threadIdx.x] = a.sched                  ? a.sched[(v.t - a.sched_t0) * kIsingTab + threadIdx.x]
                             : a.inline_thr           ? a.thr_inline[threadIdx.x]
                             : (a.thr_rep && !v.exch) ? thr_pre
                                                      : a.thr[(int64_t)(G.rep0 + tidx) * kIsingTab + threadIdx.x];
    const int64_t t = v.t;
    const bool need = (v.flags & (F_MEAS | F_EXCH)) != 0;

    // stage the box (both colours) in LDS: 16-byte vectors where rows allow, else words
    const Box X = box_of(G, box);
    const int L = G.L, Lh = G.Lh;
    // the lattice is read from a.pool and written to a.pool_out: neighbouring
    // boxes read this box's rows as their halo during the same launch
    uint8_t* P0 = a.pool + G.off + 2 * (int64_t)r * G.stride;
    uint8_t* Pout = a.pool_out + G.off + 2 * (int64_t)r * G.stride;
    if (Lh % 16 == 0) copy_box<16>(P0, G, X, s_box, true);
    else copy_box<4>(P0, G, X, s_box, true);
    __syncthreads();

    const uint32_t stream = stream_id((uint32_t)L, (uint32_t)r);
    int bsum = 0;
    if (Lh >= 16) {
        colour_pass_words(a, G, X, s_box, s_thr, 0, X.lo0, X.hi0, (uint32_t)(2 * t), stream, need, &bsum);
        __syncthreads();
        colour_pass_words(a, G, X, s_box, s_thr, 1, X.lo1, X.hi1, (uint32_t)(2 * t + 1), stream, need, &bsum);
    } else {
        colour_pass(a, G, X, s_box, s_thr, 0, X.lo0, X.hi0, (uint32_t)(2 * t), stream, need, &bsum);
        __syncthreads();
        colour_pass(a, G, X, s_box, s_thr, 1, X.lo1, X.hi1, (uint32_t)(2 * t + 1), stream, need, &bsum);
    }
    __syncthreads();

    // write the box back; magnetisation of the box rows on the way
    const int msum = Lh % 16 == 0 ? copy_box<16>(Pout, G, X, s_box, false) : copy_box<4>(Pout, G, X, s_box, false);

    if (need) {
        const int wb = wave_sum(bsum), wm = wave_sum(msum);
        if ((threadIdx.x & 63) == 0) {
            s_red[0][threadIdx.x >> 6] = wb;
            s_red[1][threadIdx.x >> 6] = wm;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t B = 0, M = 0;
            for (int i = 0; i < kFT / 64; ++i) {
                B += s_red[0][i];
                M += s_red[1][i];
            }
            int64_t* acc = acc_of(a, t);
            atomicAdd((unsigned long long*)&acc[2 * grep], (unsigned long long)B);
            atomicAdd((unsigned long long*)&acc[2 * grep + 1], (unsigned long long)M);
        }
    }

    // the block that finishes last advances the control block to the next sweep
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&a.ctl->fdone, 1) == a.fnblocks - 1;
    __syncthreads();
    if (s_last) {
        SweepView v1;
        v1.t = t;
        v1.flags = v.flags;
        v1.closing = false;
        v1.pflags = 0;
        v1.round = v.round + (v.exch ? 1 : 0);
        v1.exch = false;
        v1.par = (int)(v1.round & 1);
        sweep_bookkeeping(a, v1, 1);
        if (threadIdx.x == 0) a.ctl->fdone = 0;
    }
}

hipError_t launch_ising_fused(const IsingArgs& a, int lds_bytes, hipStream_t st) {
    const int grid = (a.fnblocks + 7) / 8 * 8;
    hipLaunchKernelGGL(ising_fused_kernel, dim3(grid), dim3(kFT), lds_bytes, st, a);
    return hipGetLastError();
}

}  // namespace ising
