// Shared device-side building blocks of libising_hip.so (gfx950 only).
//
// Counter-based random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Every draw is a pure
// function of (seed, counter), so results do not depend on tiling, launch
// shape or scheduling.  Counter words:
//   c0 = site / site-group / pair index
//   c1 = half-sweep id 2t + colour (t = global sweep index), or exchange round
//   c2 = stream id (L << 16) | replica
//   c3 = (purpose << 8) | model
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ising {

enum : uint32_t { MODEL_ISING = 0, MODEL_XY = 1, MODEL_HEIS = 2 };
enum : uint32_t {
    P_INIT = 1, P_ISING = 2, P_TIE = 3, P_EXCH = 4, P_CONT = 5, P_MSC = 6,
    P_BOND = 7, P_BTIE = 8, P_CDIR = 9, P_WSEED = 10,  // cluster updates
    P_HEAT = 11                                        // heat-bath moves (continuous models)
};

__host__ __device__ constexpr uint32_t tag(uint32_t purpose, uint32_t model) { return (purpose << 8) | model; }
__host__ __device__ constexpr uint32_t stream_id(uint32_t L, uint32_t replica) { return (L << 16) | replica; }

struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                        uint32_t k1) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        // three-input XORs: one v_bitop3_b32 each (truth table 0x96) instead of two XORs
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {c0, c1, c2, c3};
}

// 53-bit uniform in [0,1) from two words.
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
    return (double)m * (1.0 / 9007199254740992.0);
}
// 24-bit uniform in [0,1), exact in FP32.
__device__ __forceinline__ float u24f(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }

// FP64 exp from explicit fma, rint and ldexp only (Cody-Waite reduction by ln2,
// degree-13 Taylor polynomial in Horner form).  Built this way so the host and
// the device produce identical bits for the replica-exchange decision.
__device__ __forceinline__ double det_exp(double x) {
    if (x != x) return x;
    if (x < -700.0) return 0.0;
    if (x > 700.0) return __builtin_huge_val();
    const double inv_ln2 = 1.4426950408889634074;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double kd = __builtin_rint(x * inv_ln2);
    double r = __builtin_fma(-kd, ln2_hi, x);
    r = __builtin_fma(-kd, ln2_lo, r);
    double p = 1.0 / 6227020800.0;
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return __builtin_ldexp(p, (int)kd);
}

// Sum of a 32-bit integer over the 64 lanes of a wave, valid in lane 63 only
// (every lane must be active): pairs, quads, half rows and rows of 16 lanes by
// DPP lane permutations, then the four rows by DPP row broadcasts.  Six
// VALU adds, no LDS traffic.
__device__ __forceinline__ int wave_total_last(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false);  // row_half_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false);  // row_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Wave-wide sums (64 lanes) by xor shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Wave sums of a continuous pass's measurement partials (E, Mx, My, Mz) that
// can be non-zero: E only in colour-1 passes, Mz only for Heisenberg (N = 3).
// The others are +0.0 in every lane, which is also their sum.
template <int N>
__device__ __forceinline__ void wave_sum_parts(double v[4], int color) {
    if (color) v[0] = wave_sum(v[0]);
#pragma unroll
    for (int k = 1; k <= N; ++k) v[k] = wave_sum(v[k]);
}

// ---------------------------------------------------------------------------
// Ensemble descriptors shared by the host API and the kernels.
// ---------------------------------------------------------------------------

// Division by a run-time invariant d >= 2 of any 32-bit n: the round-up
// method of Granlund & Montgomery (PLDI'94).  l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, q = (t + ((n - t) >> 1)) >> (l - 1), t = mulhi(m, n).
__host__ __device__ inline void fastdiv_magic(uint32_t d, uint32_t* m, uint32_t* l) {
    uint32_t k = 0;
    while ((1ull << k) < d) ++k;
    *l = k;
    *m = (uint32_t)(((1ull << 32) * ((1ull << k) - d)) / d + 1);
}
__device__ __forceinline__ uint32_t fastdiv(uint32_t n, uint32_t m, uint32_t l) {
    const uint32_t t = __umulhi(m, n);
    return (t + ((n - t) >> 1)) >> (l - 1);
}

// One lattice size of an ensemble: R replicas of an L^3 lattice, each stored
// as two colour half-lattices (red/black) of L^3/2 sites, row-major in
// (z, y, xh) with x = 2*xh + ((y + z + colour) & 1).
struct DevGroup {
    int32_t L, Lh, R, fast;
    int32_t rep0;      // global index of replica 0 (tables are indexed by rep0 + i)
    int32_t tiles;     // work tiles per replica half-sweep
    int32_t blk0;      // first block of this group in a half-sweep launch
    int32_t dims;      // 3: L^3 cube; 2: L^2 square lattice (Ising only; z dropped)
    int64_t n_half;    // L^3 / 2
    int64_t stride;    // elements between consecutive half-lattices (padded)
    int64_t off;       // element offset of replica 0 colour 0 in the pool
    uint32_t mag_Lh, sh_Lh;  // n / Lh == fastdiv(n, mag_Lh, sh_Lh) for n < 2^32 (Lh >= 2)
    uint32_t mag_L, sh_L;    // n / L  likewise
    uint32_t mag_T, sh_T;    // n / tiles likewise (tiles >= 2)
    int32_t spt;             // Ising: thread steps per block tile
    int32_t tsz;             // Ising: half-sites per block tile (256 threads x step width x spt)
    int32_t runs;            // Ising: 16-site runs per thread step (1 or 2, a block row apart)
    int32_t pad2;
    // one-launch sweeps (int8 Ising): z planes per block (0: the whole replica
    // in one block), blocks of this size, its first block, replicas per block
    int32_t fz, fblocks, fblk0, frep;
};

// The six neighbour half-indices (in the other colour) of half-site h of
// colour `color` of group G: same row, x-shifted, y -/+ 1, z -/+ 1.  32-bit
// throughout (L^3 / 2 < 2^32 for every accepted L), divisions by the group's
// magic numbers (Lh = 1 only for L = 2).
__device__ __forceinline__ void site_nbrs(const DevGroup& G, uint32_t h, int color, uint32_t nb[6]) {
    const uint32_t L = (uint32_t)G.L, Lh = (uint32_t)G.Lh;
    const uint32_t row = Lh >= 2u ? fastdiv(h, G.mag_Lh, G.sh_Lh) : h;
    const uint32_t xh = h - row * Lh;
    const uint32_t z = fastdiv(row, G.mag_L, G.sh_L), y = row - z * L;
    const uint32_t p = (y + z + (uint32_t)color) & 1u;
    const uint32_t ym = y == 0u ? L - 1u : y - 1u, yp = y + 1u == L ? 0u : y + 1u;
    const uint32_t zm = z == 0u ? L - 1u : z - 1u, zp = z + 1u == L ? 0u : z + 1u;
    const uint32_t xo = p ? (xh + 1u == Lh ? 0u : xh + 1u) : (xh == 0u ? Lh - 1u : xh - 1u);
    // 24-bit products (one full-rate instruction each): every factor is below
    // 2^24 for L <= 2046 (rows < L^2 < 2^22)
    const uint32_t rb = __umul24(row, Lh);
    nb[0] = rb + xh;
    nb[1] = rb + xo;
    nb[2] = __umul24(__umul24(z, L) + ym, Lh) + xh;
    nb[3] = __umul24(__umul24(z, L) + yp, Lh) + xh;
    nb[4] = __umul24(__umul24(zm, L) + y, Lh) + xh;
    nb[5] = __umul24(__umul24(zp, L) + y, Lh) + xh;
}

// Component k of half-site h of a colour array C (components `stride` apart).
// O32: every half-lattice of the ensemble holds fewer than 2^30 sites, so the
// byte offset of h fits in 32 bits and a load is the component's scalar base
// plus a 32-bit vector offset (no 64-bit address arithmetic per load).
template <bool O32, class T>
__device__ __forceinline__ T* comp_at(T* C, int64_t stride, int k, uint32_t h) {
    static_assert(sizeof(T) == 4, "4-byte elements");
    if (O32) {
        typedef typename std::conditional<std::is_const<T>::value, const char, char>::type B;
        return reinterpret_cast<T*>(reinterpret_cast<B*>(C + k * stride) + (h << 2));
    }
    return C + k * stride + h;
}

// Replica (or replica word) and tile of a block's local index.
__device__ __forceinline__ void split_tiles(const DevGroup& G, int local, int* unit, int* tile) {
    const uint32_t u = G.tiles >= 2 ? fastdiv((uint32_t)local, G.mag_T, G.sh_T) : (uint32_t)local;
    *unit = (int)u;
    *tile = local - (int)u * G.tiles;
}


// Sweep control block, resident on the device.  Kernels read the sweep index
// and cadence from here, so a captured launch sequence replays unchanged.
struct DevCtl {
    int64_t t;             // global index of the sweep in progress
    int64_t sample_start;  // first sample sweep; INT64_MAX while warming up
    int32_t measure_every;
    int32_t exchange_every;
    int32_t n_rec;         // measurement records written
    int32_t rec_cap;
    int32_t flags;         // F_* decisions for sweep t (set by the host and the closing kernels)
    int32_t overflow;      // set when a record did not fit
    int64_t exch_tries;
    int64_t exch_accepts;
    int64_t mh_tries;      // learned-proposal Metropolis-Hastings moves attempted
    int64_t mh_accepts;
    int64_t round;         // exchange round of sweep t
    int32_t renorm_every;
    int32_t adapt_every;
    // phase counters of sweep t, advanced without division by the closing kernel
    int64_t s;             // t - sample_start (negative while warming up)
    int32_t ph_me, ph_ee;  // (s + 1) % measure_every, (s + 1) % exchange_every
    int32_t ph_rn, ph_ad;  // (t + 1) % renorm_every, (t + 1) % adapt_every
    // Ising sweeps close in the next colour-0 launch (see sweepctl.h)
    int64_t t1;            // sweep index seen by colour-1 launches
    int64_t round1;        // exchanges committed, as seen by colour-1 launches
    int64_t closed;        // last sweep closed by a closing kernel (-1: none)
    int32_t fl1;           // flags seen by colour-1 launches
    int32_t flp;           // flags of the previous sweep
    int32_t fdone;         // one-launch sweeps: blocks finished in the running launch
    int32_t cdone;         // continuous closing launches: blocks done with their sums
};

enum : int32_t { F_MEAS = 1, F_EXCH = 2, F_RENORM = 4, F_ADAPT = 8, F_WARM = 16 };

// Per-sweep decisions for sweep index t, computed once per sweep (by the
// closing kernel for the next sweep, by the host when a phase starts), so the
// sweep kernels never divide 64-bit integers.
__host__ __device__ inline void flags_from_phases(DevCtl* c) {
    int32_t f = 0;
    if (c->s >= 0) {
        if (c->ph_me == 0) f |= F_MEAS;
        if (c->ph_ee == 0) f |= F_EXCH;
    } else {
        f |= F_WARM;
        if (c->adapt_every > 0 && c->ph_ad == 0) f |= F_ADAPT;
    }
    if (c->renorm_every > 0 && c->ph_rn == 0) f |= F_RENORM;
    c->flags = f;
}

// Exact phases for sweep t (host side, where dividing is free).
inline void set_flags(DevCtl* c, int64_t t) {
    const bool sampling = c->sample_start <= t;
    c->s = sampling ? t - c->sample_start : -1;
    c->ph_me = sampling ? (int32_t)((c->s + 1) % c->measure_every) : 1;
    c->ph_ee = sampling ? (int32_t)((c->s + 1) % c->exchange_every) : 1;
    c->round = sampling ? c->s / c->exchange_every : 0;
    c->ph_rn = c->renorm_every > 0 ? (int32_t)((t + 1) % c->renorm_every) : 1;
    c->ph_ad = c->adapt_every > 0 ? (int32_t)((t + 1) % c->adapt_every) : 1;
    flags_from_phases(c);
}

__device__ __forceinline__ int32_t phase_step(int32_t ph, int32_t period) {
    return ph + 1 == period ? 0 : ph + 1;
}

// Advance the phases from sweep t to t + 1 (device side, no division).  Once
// sampling has started, s counts up; during warm-up it stays negative (the
// host re-derives every phase exactly when sampling starts).
// Phase counters and flags of the next sweep; the exchange round is left to
// the caller (Ising launches commit it when they close a sweep).
__device__ __forceinline__ void advance_phases(DevCtl* c) {
    if (c->s >= 0) {
        c->s += 1;
        c->ph_me = phase_step(c->ph_me, c->measure_every);
        c->ph_ee = phase_step(c->ph_ee, c->exchange_every);
    }
    if (c->renorm_every > 0) c->ph_rn = phase_step(c->ph_rn, c->renorm_every);
    if (c->adapt_every > 0) c->ph_ad = phase_step(c->ph_ad, c->adapt_every);
    flags_from_phases(c);
}

__device__ __forceinline__ void advance_flags(DevCtl* c) {
    if (c->s >= 0) {
        if (c->ph_ee == 0) c->round += 1;
        c->s += 1;
        c->ph_me = phase_step(c->ph_me, c->measure_every);
        c->ph_ee = phase_step(c->ph_ee, c->exchange_every);
    }
    if (c->renorm_every > 0) c->ph_rn = phase_step(c->ph_rn, c->renorm_every);
    if (c->adapt_every > 0) c->ph_ad = phase_step(c->ph_ad, c->adapt_every);
    flags_from_phases(c);
}

struct IsingRec {
    int64_t bond;  // forward-bond sum B (E = -J B)
    int64_t mag;   // sum of spins
    int32_t tidx;  // temperature index of the replica when measured
    int32_t pad;
};

struct ContRec {
    double e, mx, my, mz;
    int32_t tidx;
    int32_t pad;
};


}  // namespace ising
