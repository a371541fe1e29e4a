// One launch per int8 Ising sweep (both colours), gfx950.
//
// Sizes with L % 32 == 0 (L <= 128) are streamed in z.  A block owns Z
// consecutive z planes of P replicas (P = 16384 / L^2, capped by R, so that
// every thread takes one 16-site chunk of a row per colour pass and one
// 16-byte unit of a colour plane per load) and walks them through LDS: colour
// 0 in a ring of four plane slots, colour 1 in a ring of five, loading each
// plane two steps ahead into registers.  Colour 1 lags colour 0 by two planes,
// so a step has two phases: plane writes into LDS plus the decisions of the
// ties queued in the previous phase, then the two colour passes and the write
// back of a finished plane (strm_sweep).  Colour 0 of the planes z0 - 1 and
// z0 + Z, which the neighbouring blocks own, is recomputed here from the same
// inputs with the same Philox counters, so the sweep equals the two-launch
// sweep bit for bit.  A draw equal to its threshold's high part (probability
// 2^-7 per site) is queued in LDS and decided in the next phase by one lane
// each, so the pass itself has no divergent second Philox call.  Sizes with
// L < 32 are staged whole, one replica per block, four sites per thread.
//
// With no grid-wide dependency between the colours the sweep needs one launch
// instead of two: the previous sweep is closed (records, replica exchange) at
// the start as in a colour-0 launch, and the block that finishes last writes
// the next sweep's control fields (a relaxed counter; every other block has
// read them by then).  Blocks read one lattice buffer and write the other (a
// neighbour's boundary planes must be the old sweep's); the host alternates
// the two and copies back after an odd number of sweeps, so between calls the
// lattice is where it always is.
//
// Bytes per site and sweep: (Z + 4) / Z read + 1 written, against 2 + 1 for two
// launches.
#include "sweepctl.h"

namespace ising {

namespace {

constexpr int kFT = 512;      // threads per block
constexpr int kLanes = 16;    // replicas per block at most (L = 32)
constexpr int kTieCap = 512;  // queued ties per pass (further ties are decided on the spot)

typedef uint32_t v4u __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t tie_draw(uint32_t h, uint32_t hs, uint32_t stream, uint32_t k0, uint32_t k1) {
    return philox(h, hs, stream, tag(P_TIE, MODEL_ISING), k0, k1).x >> 7;
}

__device__ __forceinline__ uint32_t word4(const u32x4& w, int d) {
    return d == 0 ? w.x : (d == 1 ? w.y : (d == 2 ? w.z : w.w));
}

__device__ __forceinline__ int wrap(int v, int L) { return v < 0 ? v + L : (v >= L ? v - L : v); }

struct Shared {
    uint32_t thr[kLanes * kIsingTab];  // acceptance tables of the block's replicas
    int tidx[kLanes];
    int acc[kLanes][2];                // bond sum, magnetisation per replica lane
    uint32_t tie_w[kTieCap];           // LDS byte offset | table index << 16 | lane << 20
    uint32_t tie_h[kTieCap];           // half index of the site
    int nt[2];                         // queued ties, by pass parity
    int last;
};

// ---------------------------------------------------------------- streaming

// What one thread of a streaming block keeps through all steps.
struct Strm {
    int L, Lh, PB, P, z0, rg;
    int hs_bytes;    // one colour of one plane of the block's replicas: P * L * L / 2
    // the thread's chunk
    bool cv;         // a chunk of a valid replica
    int lane, y, x0, rym, ryp;
    uint32_t stream;
    // the thread's 16-byte unit of a colour plane
    bool uv;
    int64_t gbase;   // byte offset of the block's first replica in the lattice buffer
    uint32_t ug;     // byte offset of the unit's colour-0 plane at z = 0 from there
    int64_t stride;  // colour 1 lies one stride further
};

// LDS: colour 0 in a ring of four plane slots, colour 1 in a ring of five.
__device__ __forceinline__ uint8_t* slot0(uint8_t* ring, const Strm& S, int j) { return ring + (j & 3) * S.hs_bytes; }
__device__ __forceinline__ uint8_t* slot1(uint8_t* ring, const Strm& S, int j) {
    return ring + (4 + j % 5) * S.hs_bytes;
}
__device__ __forceinline__ uint8_t* slotc(uint8_t* ring, const Strm& S, int j, int c) {
    return c ? slot1(ring, S, j) : slot0(ring, S, j);
}

__device__ __forceinline__ v4u unit_load(const uint8_t* P0, const Strm& S, int j, int c) {
    const int64_t zo = (int64_t)wrap(S.z0 - 2 + j, S.L) * S.PB;
    v4u r = {0u, 0u, 0u, 0u};
    if (S.uv) r = *(const v4u*)(P0 + S.gbase + (c ? S.stride : 0) + zo + S.ug);
    return r;
}

__device__ __forceinline__ void unit_put(uint8_t* ring, const Strm& S, int j, int c, const v4u& r) {
    if (S.uv) *(v4u*)(slotc(ring, S, j, c) + 16 * threadIdx.x) = r;
}

// Writes both colours of plane j back; returns the sum of their spins (+-1).
__device__ __forceinline__ int unit_store(uint8_t* P1, uint8_t* ring, const Strm& S, int j) {
    if (!S.uv) return 0;
    const int64_t zo = (int64_t)wrap(S.z0 - 2 + j, S.L) * S.PB;
    const v4u c0 = *(const v4u*)(slot0(ring, S, j) + 16 * threadIdx.x);
    const v4u c1 = *(const v4u*)(slot1(ring, S, j) + 16 * threadIdx.x);
    *(v4u*)(P1 + S.gbase + zo + S.ug) = c0;
    *(v4u*)(P1 + S.gbase + S.stride + zo + S.ug) = c1;
    int m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) m += __builtin_popcount(c0[k]) + __builtin_popcount(c1[k]);
    return 2 * m - 32;
}

// Colour pass over this thread's chunk of plane j: four sites per 32-bit word,
// up-neighbour counts per byte from five row words and the x-shifted word
// (edge byte across the chunk), thresholds by byte permutes of the packed
// table, all four 7-bit draws compared at once.  A byte of exactly 0x80 after
// the subtraction is a tie: the provisional byte keeps the spin and the site
// is queued (LDS offset, table index, lane, colour) for the next phase.
// count_b: add the chunk's bonds to bsum.
__device__ __forceinline__ void strm_pass(const IsingArgs& a, uint8_t* ring, const Strm& S, Shared& sh, int j, int c,
                                          uint32_t hs, int tq, bool count_b, int& bsum) {
    if (!S.cv) return;
    const int L = S.L, Lh = S.Lh;
    const int zq = wrap(S.z0 - 2 + j, L);
    const int lo = S.lane * S.PB;
    uint8_t* Cs = slotc(ring, S, j, c) + lo;
    const uint8_t* Os = slotc(ring, S, j, c ^ 1) + lo;
    const uint8_t* Om = slotc(ring, S, j - 1, c ^ 1) + lo;
    const uint8_t* Op = slotc(ring, S, j + 1, c ^ 1) + lo;
    const int rb = S.y * Lh + S.x0;
    const int p = (S.y + zq + c) & 1;
    const v4u vc = *(const v4u*)(Cs + rb);
    const v4u vx = *(const v4u*)(Os + rb);
    const v4u v1 = *(const v4u*)(Os + S.rym + S.x0);
    const v4u v2 = *(const v4u*)(Os + S.ryp + S.x0);
    const v4u v3 = *(const v4u*)(Om + rb);
    const v4u v4 = *(const v4u*)(Op + rb);
    const int xe = p ? (S.x0 + 16 == Lh ? 0 : S.x0 + 16) : (S.x0 == 0 ? Lh - 1 : S.x0 - 1);
    const uint32_t e = Os[S.y * Lh + xe];
    const uint32_t pm = 0u - (uint32_t)p;
    const uint32_t h0 = (uint32_t)((zq * L + S.y) * Lh + S.x0);
    const u32x4 w = philox(h0 >> 4, hs, S.stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
    const uint32_t* tb = sh.thr + S.lane * kIsingTab;
    const uint32_t T0l = tb[28], T0h = tb[29], T1l = tb[30], T1h = tb[31];
    v4u nc;
    uint32_t bn = 0, sn = 0;
    int ups = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const uint32_t up = __builtin_amdgcn_alignbyte(d < 3 ? vx[d < 3 ? d + 1 : 0] : e, vx[d], 1u);
        const uint32_t dn = __builtin_amdgcn_alignbyte(vx[d], d > 0 ? vx[d > 0 ? d - 1 : 0] : e << 24, 3u);
        const uint32_t xs = (up & pm) | (dn & ~pm);
        const uint32_t ns = vx[d] + xs + v1[d] + v2[d] + v3[d] + v4[d];  // per byte 0..6
        const uint32_t cw = vc[d];
        const uint32_t m = __builtin_amdgcn_perm(0u, 0x0000ff00u, cw);  // 0xff in up-spin bytes
        const uint32_t thr = (m & __builtin_amdgcn_perm(T1h, T1l, ns)) | (~m & __builtin_amdgcn_perm(T0h, T0l, ns));
        const uint32_t zz = (word4(w, d) | 0x80808080u) - thr;
        uint32_t n2 = cw ^ (((zz >> 7) & 0x01010101u) ^ 0x01010101u);
        const uint32_t uu = (zz & 0x7f7f7f7fu) + 0x7f7f7f7fu;
        uint32_t tie = zz & ~uu & 0x80808080u;
        while (tie) {
            const uint32_t bit = (uint32_t)__builtin_ctz(tie);
            tie &= tie - 1u;
            const uint32_t jj = bit >> 3;
            const uint32_t bb = (cw >> (8 * jj)) & 1u, nn = (ns >> (8 * jj)) & 7u;
            const uint32_t kk = bb * 7u + nn;
            const uint32_t hh = h0 + 4u * (uint32_t)d + jj;
            const int q = atomicAdd(&sh.nt[tq], 1);
            if (q < kTieCap) {
                sh.tie_w[q] = (uint32_t)(Cs - ring + rb + 4 * d + (int)jj) | (kk << 17) |
                              ((uint32_t)S.lane << 21) | ((uint32_t)c << 25);
                sh.tie_h[q] = hh;
            } else if (tie_draw(hh, hs, S.stream, a.k0, a.k1) < tb[14 + kk]) {
                n2 ^= 1u << (8 * jj);  // queue full: decided on the spot
            }
        }
        nc[d] = n2;
        if (count_b) {
            ups += __builtin_popcount(n2);
            bn += __builtin_amdgcn_perm(0u, 0x0000ff00u, n2) & ns;
            sn += ns;
        }
    }
    *(v4u*)(Cs + rb) = nc;
    if (count_b)
        bsum += 4 * (int)__builtin_amdgcn_sad_u8(bn, 0u, 0u) - 12 * ups - 2 * (int)__builtin_amdgcn_sad_u8(sn, 0u, 0u) + 96;
}

// Decides the ties queued in the last phase (one lane each) and patches their
// bytes; the bond sums move by the colour-1 flips.
__device__ __forceinline__ void strm_ties(const IsingArgs& a, uint8_t* ring, const Strm& S, Shared& sh, int nt,
                                          uint32_t hs0, bool need) {
    for (int k = threadIdx.x; k < nt; k += kFT) {
        const uint32_t w = sh.tie_w[k];
        const uint32_t kk = (w >> 17) & 15u, lane = (w >> 21) & 15u, c = (w >> 25) & 1u;
        const uint32_t stream = stream_id((uint32_t)S.L, (uint32_t)(S.rg * S.P + (int)lane));
        if (tie_draw(sh.tie_h[k], hs0 + c, stream, a.k0, a.k1) < sh.thr[lane * kIsingTab + 14 + kk]) {
            ring[w & 0x1ffffu] ^= 1u;
            if (need && c) {
                const int b = kk >= 7u ? 1 : 0, n = (int)kk - 7 * b;
                atomicAdd(&sh.acc[lane][0], (2 - 4 * b) * (2 * n - 6));
            }
        }
    }
}

// The steps of a streaming block.  With j the plane index relative to z0 - 2,
// step s has two phases separated by barriers:
//   A: colour 0 of plane s + 1 and colour 1 of plane s + 2 go from registers
//      into their slots (loads of the planes two steps ahead start), and the
//      ties queued in the previous phase B are decided;
//   B: colour 0 of plane s + 1 is updated (j = 1 .. Z + 2: the planes z0 - 1
//      and z0 + Z are recomputed), colour 1 of plane s - 1 (j = 2 .. Z + 1),
//      and plane s - 2 (final in both colours) is written back.
// Colour 1 of plane j reads colour 0 of j - 1 .. j + 1, updated at steps
// j - 2 .. j with their ties decided by step j + 1; colour 0 of plane j reads
// colour 1 of j - 1 .. j + 1 before any of them is updated.
__device__ __forceinline__ void strm_sweep(const IsingArgs& a, uint8_t* s_ring, const Strm& S, Shared& sh, int Z,
                                           uint32_t hs0, uint32_t hs1, bool need, int& msum, int& bsum) {
    const uint8_t* P0 = a.pool;
    uint8_t* P1 = a.pool_out;
    // colour 1 of planes 0 and 1 straight into their slots; registers hold
    // (colour 0 of j + 1, colour 1 of j + 2) for the next two steps
    unit_put(s_ring, S, 0, 1, unit_load(P0, S, 0, 1));
    unit_put(s_ring, S, 1, 1, unit_load(P0, S, 1, 1));
    v4u a0 = unit_load(P0, S, 1, 0), a1 = unit_load(P0, S, 2, 1);
    v4u b0 = unit_load(P0, S, 2, 0), b1 = unit_load(P0, S, 3, 1);
    auto step = [&](int s, v4u& r0, v4u& r1) {
        // phase A
        if (s + 1 <= Z + 2) unit_put(s_ring, S, s + 1, 0, r0);
        if (s + 2 <= Z + 3) unit_put(s_ring, S, s + 2, 1, r1);
        r0 = unit_load(P0, S, min(s + 3, Z + 2), 0);
        r1 = unit_load(P0, S, min(s + 4, Z + 3), 1);
        const int nt = min(sh.nt[(s + 1) & 1], kTieCap);
        strm_ties(a, s_ring, S, sh, nt, hs0, need);
        __syncthreads();
        // phase B
        if (threadIdx.x == 0) sh.nt[(s + 1) & 1] = 0;
        if (s <= Z + 1) strm_pass(a, s_ring, S, sh, s + 1, 0, hs0, s & 1, false, bsum);
        if (s >= 3 && s <= Z + 2) strm_pass(a, s_ring, S, sh, s - 1, 1, hs1, s & 1, need, bsum);
        if (s >= 4) {
            const int m = unit_store(P1, s_ring, S, s - 2);
            if (need) msum += m;
        }
        __syncthreads();
    };
    for (int s = 0; s < Z + 4; s += 2) {
        step(s, a0, a1);
        step(s + 1, b0, b1);
    }
}

// ------------------------------------------------------------ whole replica

// One colour pass over a whole replica staged in LDS (L < 32: rows shorter than
// a 16-site chunk), four sites per thread (the chunk's Philox output is
// computed by each of its four threads), ties decided on the spot.
__device__ __forceinline__ void whole_pass(const IsingArgs& a, int L, int Lh, uint8_t* sb, const uint32_t* s_thr,
                                           int color, uint32_t hs, uint32_t stream, bool need, int* bsum) {
    uint8_t* Cs = sb + color * L * L * Lh;
    const uint8_t* Os = sb + (color ^ 1) * L * L * Lh;
    const int nquads = (L * L * Lh) / 4;
    for (int k = threadIdx.x; k < nquads; k += kFT) {
        const uint32_t h0 = (uint32_t)k * 4u;
        const u32x4 w = philox(h0 >> 4, hs, stream, tag(P_ISING, MODEL_ISING), a.k0, a.k1);
        const uint32_t rw = word4(w, k & 3);
        for (int j = 0; j < 4; ++j) {
            const int h = (int)h0 + j;
            const int row = h / Lh, xh = h - row * Lh;
            const int y = row % L, z = row / L;
            const int p = (y + z + color) & 1;
            const int xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
            const int ym = wrap(y - 1, L), yp = wrap(y + 1, L), zm = wrap(z - 1, L), zp = wrap(z + 1, L);
            const int rb = (z * L + y) * Lh;
            const uint32_t n = (uint32_t)Os[rb + xh] + Os[rb + xo] + Os[(z * L + ym) * Lh + xh] +
                               Os[(z * L + yp) * Lh + xh] + Os[(zm * L + y) * Lh + xh] + Os[(zp * L + y) * Lh + xh];
            const uint32_t b = Cs[rb + xh];
            const uint32_t kk = b * 7u + n;
            const uint32_t thi = s_thr[kk];
            const uint32_t r7 = (rw >> (8 * j)) & 0x7fu;
            bool acc = r7 < thi;
            if (r7 == thi) acc = tie_draw((uint32_t)h, hs, stream, a.k0, a.k1) < s_thr[14 + kk];
            const uint32_t nb = b ^ (uint32_t)acc;
            Cs[rb + xh] = (uint8_t)nb;
            if (need && color) *bsum += (2 * (int)nb - 1) * (2 * (int)n - 6);
        }
    }
}

}  // namespace

__global__ __launch_bounds__(kFT) void ising_fused_kernel(IsingArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t s_ring[];
    __shared__ Shared sh;
    const int per = (int)(gridDim.x >> 3);
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (b >= a.fnblocks) return;
    // sizes hold consecutive block ranges, the largest first
    int g = 0;
    if (a.gk_on) {
        while (g + 1 < a.ngroups && (unsigned)(b - a.gk[g].fblk0) >= (unsigned)a.gk[g].fblocks) ++g;
    } else {
        while (g + 1 < a.ngroups && (unsigned)(b - a.groups[g].fblk0) >= (unsigned)a.groups[g].fblocks) ++g;
    }
    const DevGroup G = a.gk_on ? a.gk[g] : a.groups[g];
    const int local = b - G.fblk0;
    const int L = G.L, Lh = G.Lh;
    const bool strm = G.fz > 0;
    const int P = strm ? G.frep : 1;
    int rg, zc;  // replica group (replicas rg * P + lane) and z chunk
    if (strm) {
        const int nzc = L / G.fz;
        rg = local / nzc;
        zc = local - rg * nzc;
    } else {
        rg = local;
        zc = 0;
    }
    const int nlanes = min(P, G.R - rg * P);

    // tables of the block's replicas (prefetched when no exchange intervenes)
    const int tl = (int)threadIdx.x >> 5, te = (int)threadIdx.x & 31;
    uint32_t thr_pre = 0u;
    if (a.thr_rep && tl < nlanes) thr_pre = a.thr_rep[(int64_t)(G.rep0 + rg * P + tl) * kIsingTab + te];
    const SweepView v = view_of(a.ctl, 0);
    if (v.exch) {
        if ((int)threadIdx.x < nlanes) {
            const int64_t po = (v.round & 1) * a.total_reps;
            sh.tidx[threadIdx.x] =
                exch_tidx(a, G, rg * P + (int)threadIdx.x, v.round, a.r2t + po, a.t2r + po, acc_of(a, v.t - 1));
        }
        if (local == 0) exch_write_group(a, G, v.round, acc_of(a, v.t - 1));
    } else if (!a.thr_rep && !a.sched && !a.inline_thr && (int)threadIdx.x < nlanes) {
        sh.tidx[threadIdx.x] = a.r2t[(int64_t)v.par * a.total_reps + G.rep0 + rg * P + (int)threadIdx.x];
    }
    if (b == 0) sweep_bookkeeping(a, v, 0);
    if (threadIdx.x < kLanes) {
        sh.acc[threadIdx.x][0] = 0;
        sh.acc[threadIdx.x][1] = 0;
    }
    if (threadIdx.x < 2) sh.nt[threadIdx.x] = 0;
    __syncthreads();
    if (tl < nlanes)
        sh.thr[threadIdx.x] = a.sched                  ? a.sched[(v.t - a.sched_t0) * kIsingTab + te]
                              : a.inline_thr           ? a.thr_inline[te]
                              : (a.thr_rep && !v.exch) ? thr_pre
                                                       : a.thr[(int64_t)(G.rep0 + sh.tidx[tl]) * kIsingTab + te];
    const int64_t t = v.t;
    const bool need = (v.flags & (F_MEAS | F_EXCH)) != 0;
    const uint32_t hs0 = (uint32_t)(2 * t), hs1 = (uint32_t)(2 * t + 1);
    int msum = 0, bsum = 0;
    int my_lane = 0;

    if (strm) {
        Strm S;
        S.L = L;
        S.Lh = Lh;
        S.PB = L * Lh;
        S.P = P;
        S.z0 = zc * G.fz;
        S.rg = rg;
        S.hs_bytes = P * S.PB;
        {
            const int cpr = Lh >> 4, cpl = L * cpr;
            const int i = (int)threadIdx.x;
            S.lane = i / cpl;
            const int rem = i - S.lane * cpl;
            S.y = rem / cpr;
            S.x0 = (rem - S.y * cpr) * 16;
            S.cv = S.lane < nlanes;
            if (!S.cv) S.lane = 0;
            S.rym = wrap(S.y - 1, L) * Lh;
            S.ryp = wrap(S.y + 1, L) * Lh;
            S.stream = stream_id((uint32_t)L, (uint32_t)(rg * P + S.lane));
            my_lane = S.lane;
        }
        {
            const int upc = S.PB >> 4;  // 16-byte units per colour plane of one replica
            const int u = (int)threadIdx.x;
            const int ln = u / upc, o = u - ln * upc;
            S.uv = ln < nlanes;
            S.gbase = G.off + 2 * (int64_t)rg * P * G.stride;
            S.ug = (uint32_t)(2 * ln * G.stride + 16 * o);  // < 2 P stride <= 16384 L
            S.stride = G.stride;
        }
        strm_sweep(a, s_ring, S, sh, G.fz, hs0, hs1, need, msum, bsum);
    } else {
        // whole replica (one per block)
        const int r = rg;
        const int nb = L * L * Lh;  // bytes per colour
        const uint8_t* src = a.pool + G.off + 2 * (int64_t)r * G.stride;
        uint8_t* dst = a.pool_out + G.off + 2 * (int64_t)r * G.stride;
        for (int i = threadIdx.x; i < nb / 2; i += kFT) {  // 4-byte words of both colours
            const int c = (4 * i) / nb, o = 4 * i - c * nb;
            *(uint32_t*)(s_ring + 4 * i) = *(const uint32_t*)(src + c * G.stride + o);
        }
        __syncthreads();
        const uint32_t stream = stream_id((uint32_t)L, (uint32_t)r);
        whole_pass(a, L, Lh, s_ring, sh.thr, 0, hs0, stream, need, &bsum);
        __syncthreads();
        whole_pass(a, L, Lh, s_ring, sh.thr, 1, hs1, stream, need, &bsum);
        __syncthreads();
        for (int i = threadIdx.x; i < nb / 2; i += kFT) {
            const int c = (4 * i) / nb, o = 4 * i - c * nb;
            const uint32_t w = *(const uint32_t*)(s_ring + 4 * i);
            *(uint32_t*)(dst + c * G.stride + o) = w;
            msum += 2 * __builtin_popcount(w) - 4;
        }
    }

    if (need) {
        // per replica lane; waves that lie in one lane reduce first
        const int cpl = strm ? L * (Lh >> 4) : kFT;
        if (P == 1 || cpl % 64 == 0) {
            const int wb = wave_sum(bsum), wm = wave_sum(msum);
            if ((threadIdx.x & 63) == 0 && (wb | wm)) {
                atomicAdd(&sh.acc[my_lane][0], wb);
                atomicAdd(&sh.acc[my_lane][1], wm);
            }
        } else if (bsum | msum) {
            atomicAdd(&sh.acc[my_lane][0], bsum);
            atomicAdd(&sh.acc[my_lane][1], msum);
        }
        __syncthreads();
        if ((int)threadIdx.x < nlanes) {
            const int gr = G.rep0 + rg * P + (int)threadIdx.x;
            // accumulator copy by block, as in the two-launch kernel (copies are summed where read)
            int64_t* acc = acc_of(a, t) + 2 * (int64_t)(b & (a.acc_copies - 1)) * a.total_reps;
            atomicAdd((unsigned long long*)&acc[2 * gr], (unsigned long long)(int64_t)sh.acc[threadIdx.x][0]);
            atomicAdd((unsigned long long*)&acc[2 * gr + 1], (unsigned long long)(int64_t)sh.acc[threadIdx.x][1]);
        }
    }

    // the block that finishes last advances the control block to the next sweep
    __syncthreads();
    if (threadIdx.x == 0) sh.last = atomicAdd(&a.ctl->fdone, 1) == a.fnblocks - 1;
    __syncthreads();
    if (sh.last) {
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
