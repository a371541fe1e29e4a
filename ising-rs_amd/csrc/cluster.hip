// Cluster updates for the Ising (int8, square and cubic), XY and Heisenberg
// ensembles, gfx950: Swendsen-Wang (every Fortuin-Kasteleyn cluster flipped
// with probability 1/2) and single-cluster Wolff (the cluster of one random
// seed site flipped).  Both label one bond configuration per update:
//
//   cluster_bond_kernel   one block per (replica, tile): spins of the tile and
//                         of its +x/+y/+z face planes into LDS, the forward
//                         bonds of every site drawn and stored (one byte per
//                         site: bits 0-2 bond +x/+y/+z, bit 3 the site's
//                         cluster coin), union-find over the bonds inside the
//                         tile in LDS, global label = index of the tile-local root
//   cluster_merge_kernel  union-find over the bonds that leave a tile through
//                         its + faces, on the global labels with atomicMin
//                         (a link always points to a smaller index, so every
//                         cluster ends with its smallest site index as root)
//   cluster_flip_kernel   root of every site; Swendsen-Wang flips it when the
//                         root's coin is set, Wolff when the root is the
//                         seed's; flipped sites counted per replica
//
// The labelling is the block-local + global union-find of Playne & Hawick,
// "A new algorithm for parallel connected-component labelling on GPUs",
// IEEE TPDS 29 (2018) 1217.  The partition is unique and its roots are the
// cluster minima, so the result is a pure function of (seed, update index,
// state): tiling and scheduling do not change it.
//
// Continuous spins use Wolff's embedding (Wolff, PRL 62 (1989) 361): a random
// unit vector r per replica and update, projections a_i = s_i . r, bond (i,j)
// active with probability 1 - exp(-2 beta J a_i a_j) when J a_i a_j > 0, and
// a flip reflects s -> s - 2 (s . r) r.
//
// The reference plans this as a label-propagation Wolff step
// (docs/plans/2026-03-09-gpu-compute-optimizations-design.md:82-116,
// gpu-compute-optimizations-impl.md:857-1010); its CPU cluster algorithms are
// src/wolff.rs:27-77 and src/xy/wolff.rs:16-73.
//
// Random numbers (Philox4x32-10, c1 = update index u, c2 = stream(L, replica)):
//   Ising bond   c0 = site / 2, tag P_BOND: words (w0, w1) belong to the even
//                site of the pair, (w2, w3) to the odd one; with (wa, wb) the
//                site's words the 16-bit draws of bonds +x, +y, +z are
//                wa & 0xffff, wa >> 16, wb & 0xffff and bit 31 of wb is the coin.
//                Bond active iff aligned (J > 0) / anti-aligned (J < 0) and
//                draw < hi, or draw == hi and the 16-bit tie draw < lo, where
//                hi * 2^16 + lo = floor(2^32 (1 - exp(-2 beta |J|))).
//   Ising tie    c0 = site, tag (P_BTIE << 8) | d: word 0 >> 16.
//   cont. bond   c0 = site index, tag P_BOND; word d -> 24-bit uniform, bit 31 of word 3 the coin.
//   cont. r      c0 = 0, tag P_CDIR: XY phi = 2 pi u01(w0, w1); Heisenberg
//                z = 2 u01(w0, w1) - 1, phi = 2 pi u01(w2, w3).
//   Wolff seed   c0 = 0, tag P_WSEED: site = floor(N * ((w0 << 32) | w1) / 2^64).
#include "kernels.h"

namespace ising {

namespace {

constexpr int kCB = 256;  // threads per block
constexpr int kPer = kClusterTile / kCB;  // sites per thread
static_assert(kClusterFace <= kCB, "one face site per thread");

__device__ __forceinline__ int find_cgroup(const ClusterGroup* cg, int ngroups, int b) {
    int g = 0;
    while (g + 1 < ngroups && b >= cg[g + 1].blk0) ++g;
    return g;
}

// A block's tile: replica, origin, extents and the divisions by them.
struct Tile {
    int r, x0, y0, z0, ex, ey, ez, n;
    uint32_t mx, sx, my, sy;  // n / ex, n / ey by fastdiv
};

__device__ __forceinline__ Tile tile_of(const ClusterGroup& C, int local) {
    Tile T;
    T.r = local / C.tiles;
    int t = local - T.r * C.tiles;
    const int itx = t % C.ntx;
    t /= C.ntx;
    const int ity = t % C.nty, itz = t / C.nty;
    T.x0 = itx * C.tx;
    T.y0 = ity * C.ty;
    T.z0 = itz * C.tz;
    T.ex = min(C.tx, C.L - T.x0);
    T.ey = min(C.ty, C.L - T.y0);
    T.ez = C.dims == 3 ? min(C.tz, C.L - T.z0) : 1;
    T.n = T.ex * T.ey * T.ez;
    const int lx = itx == C.ntx - 1, ly = ity == C.nty - 1;
    T.mx = C.mag_x[lx];
    T.sx = C.sh_x[lx];
    T.my = C.mag_y[ly];
    T.sy = C.sh_y[ly];
    return T;
}

// Tile-local index li = (zl * ey + yl) * ex + xl -> (xl, yl, zl).
__device__ __forceinline__ void decode(const Tile& T, uint32_t li, int* xl, int* yl, int* zl) {
    const uint32_t row = fastdiv(li, T.mx, T.sx);
    *xl = (int)(li - row * (uint32_t)T.ex);
    const uint32_t z = fastdiv(row, T.my, T.sy);
    *yl = (int)(row - z * (uint32_t)T.ey);
    *zl = (int)z;
}

// Site index within a replica (L^dims < 2^31).
__device__ __forceinline__ uint32_t site_of(const ClusterGroup& C, int x, int y, int z) {
    return ((uint32_t)z * (uint32_t)C.L + (uint32_t)y) * (uint32_t)C.L + (uint32_t)x;
}

// Element of site (x, y, z) in the colour half-lattice base[(x + y + z) & 1].
template <typename P>
__device__ __forceinline__ P* at(P* const base[2], const DevGroup& G, int x, int y, int z) {
    const uint32_t h = ((uint32_t)z * (uint32_t)G.L + (uint32_t)y) * (uint32_t)G.Lh + ((uint32_t)x >> 1);
    return base[(x + y + z) & 1] + h;
}

// Random reflection direction of replica r for update u (FP32).
template <int NC>
__device__ __forceinline__ void direction(const ClusterArgs& a, uint32_t stream, float rv[3]) {
    const u32x4 w = philox(0u, a.u, stream, tag(P_CDIR, a.model), a.k0, a.k1);
    if (NC == 2) {
        const float phi = (float)(6.283185307179586 * u01(w.x, w.y));
        rv[0] = cosf(phi);
        rv[1] = sinf(phi);
        rv[2] = 0.f;
    } else {
        const float zc = (float)(2.0 * u01(w.x, w.y) - 1.0);
        const float phi = (float)(6.283185307179586 * u01(w.z, w.w));
        const float s = sqrtf(fmaxf(0.f, 1.f - zc * zc));
        rv[0] = s * cosf(phi);
        rv[1] = s * sinf(phi);
        rv[2] = zc;
    }
}

template <int NC>
__device__ __forceinline__ float project(const float* p, int64_t stride, const float rv[3]) {
    float v = p[0] * rv[0] + p[stride] * rv[1];
    if (NC == 3) v += p[2 * stride] * rv[2];
    return v;
}

// Temperature slot of replica r (current permutation).
__device__ __forceinline__ int tslot(const ClusterArgs& a, const ClusterGroup& C, int r) {
    const int par = a.ctl ? (int)(a.ctl->round & 1) : 0;
    return C.rep0 + a.r2t[(int64_t)par * a.total_reps + C.rep0 + r];
}

// --- union-find in LDS -------------------------------------------------------

// Path halving: a parent is only ever replaced by one of its ancestors, which
// keeps every set intact under concurrent unions.
__device__ __forceinline__ uint32_t lds_find(uint32_t* lab, uint32_t x) {
    uint32_t p = __hip_atomic_load(&lab[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        const uint32_t gp = __hip_atomic_load(&lab[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (gp != p) __hip_atomic_store(&lab[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void lds_union(uint32_t* lab, uint32_t a, uint32_t b) {
    for (;;) {
        a = lds_find(lab, a);
        b = lds_find(lab, b);
        if (a == b) return;
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicMin(&lab[b], a);
        if (old == b) return;
        b = old;
    }
}

// --- union-find on the global labels -----------------------------------------
// Parents only decrease and always stay inside their set, so a stale (cached)
// parent is still an ancestor: finds may read through the caches, and a union
// that acts on a stale root is corrected by the atomicMin's return value
// (the link fails, and the union retries with the parent that won).  Loads
// and the path-halving stores are relaxed atomics so the compiler keeps every
// access; only the link itself is a device-wide atomic.

__device__ __forceinline__ int32_t g_load(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int32_t g_find(int32_t* lab, int32_t x) {
    int32_t p = g_load(&lab[x]);
    while (p != x) {
        const int32_t gp = g_load(&lab[p]);
        if (gp != p) __hip_atomic_store(&lab[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void g_union(int32_t* lab, int32_t a, int32_t b) {
    for (;;) {
        a = g_find(lab, a);
        b = g_find(lab, b);
        if (a == b) return;
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        const int32_t old = atomicMin(&lab[b], a);
        if (old == b) return;
        b = old;
    }
}

// Wave-aggregated append of `v` (lanes with act set) to an LDS list.
__device__ __forceinline__ void append(uint16_t* list, int* n, bool act, uint16_t v) {
    const uint64_t m = __ballot(act);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == __ffsll((unsigned long long)m) - 1) base = atomicAdd(n, __popcll(m));
    base = __shfl(base, __ffsll((unsigned long long)m) - 1, 64);
    if (act) list[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

template <int NC>
struct SpinT {
    using type = float;
};
template <>
struct SpinT<0> {
    using type = int8_t;
};

}  // namespace

// ---------------------------------------------------------------------------
// Bonds and tile-local labels.
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(kCB) void cluster_bond_kernel(ClusterArgs a) {
    using S = typename SpinT<NC>::type;
    using P = typename std::conditional<NC == 0, uint8_t, float>::type;
    __shared__ S s_v[kClusterTile];
    __shared__ S s_h[3][kClusterFace];
    __shared__ uint8_t s_bond[kClusterTile];
    __shared__ uint32_t s_lab[kClusterTile];
    __shared__ uint16_t s_list[kClusterTile];
    __shared__ int s_nlist;

    const int g = find_cgroup(a.cg, a.ngroups, blockIdx.x);
    const ClusterGroup C = a.cg[g];
    const DevGroup G = a.groups[g];
    const Tile T = tile_of(C, (int)blockIdx.x - C.blk0);
    const int L = C.L, dims = C.dims;
    const uint32_t stream = stream_id((uint32_t)L, (uint32_t)T.r);
    const int tid = threadIdx.x;

    float rv[3] = {0.f, 0.f, 0.f};
    if constexpr (NC > 0) direction<NC>(a, stream, rv);
    P* base[2];
    if constexpr (NC == 0) {
        base[0] = a.ipool + G.off + 2 * (int64_t)T.r * G.stride;
        base[1] = base[0] + G.stride;
    } else {
        base[0] = a.fpool + G.off + 2 * (int64_t)T.r * NC * G.stride;
        base[1] = base[0] + NC * G.stride;
    }
    auto value = [&](int x, int y, int z) -> S {
        if constexpr (NC == 0) return (S)*at(base, G, x, y, z);
        else return project<NC>(at(base, G, x, y, z), G.stride, rv);
    };

    // every load of the tile and its three + faces first (a face has at most
    // kCB sites), then the stores to LDS
    {
        S v[kPer], hv[3];
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int li = tid + k * kCB;
            if (li < T.n) {
                int xl, yl, zl;
                decode(T, (uint32_t)li, &xl, &yl, &zl);
                v[k] = value(T.x0 + xl, T.y0 + yl, T.z0 + zl);
            }
        }
        const int xn = T.x0 + T.ex == L ? 0 : T.x0 + T.ex;
        const int yn = T.y0 + T.ey == L ? 0 : T.y0 + T.ey;
        const int zn = T.z0 + T.ez == L ? 0 : T.z0 + T.ez;
        const int f0 = T.ey * T.ez, f1 = T.ex * T.ez, f2 = dims == 3 ? T.ex * T.ey : 0;
        if (tid < f0) {
            const int zl = (int)fastdiv((uint32_t)tid, T.my, T.sy), yl = tid - zl * T.ey;
            hv[0] = value(xn, T.y0 + yl, T.z0 + zl);
        }
        if (tid < f1) {
            const int zl = (int)fastdiv((uint32_t)tid, T.mx, T.sx), xl = tid - zl * T.ex;
            hv[1] = value(T.x0 + xl, yn, T.z0 + zl);
        }
        if (tid < f2) {
            const int yl = (int)fastdiv((uint32_t)tid, T.mx, T.sx), xl = tid - yl * T.ex;
            hv[2] = value(T.x0 + xl, T.y0 + yl, zn);
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int li = tid + k * kCB;
            if (li < T.n) {
                s_v[li] = v[k];
                s_lab[li] = (uint32_t)li;
            }
        }
        if (tid < f0) s_h[0][tid] = hv[0];
        if (tid < f1) s_h[1][tid] = hv[1];
        if (tid < f2) s_h[2][tid] = hv[2];
    }
    __syncthreads();

    // neighbour value of tile site (xl, yl, zl) in direction d
    auto nbr = [&](int li, int xl, int yl, int zl, int d) -> S {
        if (d == 0) return xl + 1 < T.ex ? s_v[li + 1] : s_h[0][zl * T.ey + yl];
        if (d == 1) return yl + 1 < T.ey ? s_v[li + T.ex] : s_h[1][zl * T.ex + xl];
        return zl + 1 < T.ez ? s_v[li + T.ex * T.ey] : s_h[2][yl * T.ex + xl];
    };

    uint8_t* bond = a.bond + C.lab0 + (int64_t)T.r * C.nsites;
    if constexpr (NC == 0) {
        const uint32_t thr = a.inline_p ? a.bthr_inline : a.bthr[tslot(a, C, T.r)];
        const uint32_t hi = thr >> 16, lo = thr & 0xffffu;
        const bool ferro = a.J > 0.0;
        // one Philox call per pair of sites (x even, x + 1): ex and x0 are even
        for (int k = tid; 2 * k < T.n; k += kCB) {
            const int li0 = 2 * k;
            int xl, yl, zl;
            decode(T, (uint32_t)li0, &xl, &yl, &zl);
            const uint32_t site0 = site_of(C, T.x0 + xl, T.y0 + yl, T.z0 + zl);
            const u32x4 w4 = philox(site0 >> 1, a.u, stream, tag(P_BOND, MODEL_ISING), a.k0, a.k1);
            for (int s = 0; s < 2; ++s) {
                const int li = li0 + s;
                const uint32_t wa = s ? w4.z : w4.x, wb = s ? w4.w : w4.y;
                const S si = s_v[li];
                uint32_t bits = (wb >> 31) << 3;
                for (int d = 0; d < dims; ++d) {
                    const S sj = nbr(li, xl + s, yl, zl, d);
                    if ((si == sj) != ferro) continue;
                    const uint32_t r16 = d == 0 ? wa & 0xffffu : (d == 1 ? wa >> 16 : wb & 0xffffu);
                    bool act = r16 < hi;
                    if (r16 == hi && lo != 0u)  // 1 in 65536: 16 more bits decide
                        act = (philox(site0 + s, a.u, stream, tag(P_BTIE, d), a.k0, a.k1).x >> 16) < lo;
                    bits |= (uint32_t)act << d;
                }
                s_bond[li] = (uint8_t)bits;
            }
        }
    } else {
        const double beta = a.inline_p ? a.beta_inline : a.betas[tslot(a, C, T.r)];
        const float kappa = (float)(2.0 * beta * a.J);
        for (int li = tid; li < T.n; li += kCB) {
            int xl, yl, zl;
            decode(T, (uint32_t)li, &xl, &yl, &zl);
            const uint32_t site = site_of(C, T.x0 + xl, T.y0 + yl, T.z0 + zl);
            const u32x4 w4 = philox(site, a.u, stream, tag(P_BOND, a.model), a.k0, a.k1);
            const uint32_t wk[3] = {w4.x, w4.y, w4.z};
            const float ai = s_v[li];
            uint32_t bits = (w4.w >> 31) << 3;
            for (int d = 0; d < 3; ++d) {
                const float x = kappa * ai * nbr(li, xl, yl, zl, d);
                if (x > 0.f && u24f(wk[d]) < -expm1f(-x)) bits |= 1u << d;
            }
            s_bond[li] = (uint8_t)bits;
        }
    }
    __syncthreads();

    // union-find over the bonds inside the tile, one direction per round: the
    // active bonds of the round are first compacted into a list (so every
    // lane of a wave has a union to do), then united
    uint32_t inner[kPer];  // bit d: the +d neighbour lies in the tile
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int li = tid + k * kCB;
        inner[k] = 0u;
        if (li < T.n) {
            int xl, yl, zl;
            decode(T, (uint32_t)li, &xl, &yl, &zl);
            inner[k] = (uint32_t)(xl + 1 < T.ex) | ((uint32_t)(yl + 1 < T.ey) << 1) | ((uint32_t)(zl + 1 < T.ez) << 2);
        }
    }
    const int step[3] = {1, T.ex, T.ex * T.ey};
    for (int d = 0; d < dims; ++d) {
        if (tid == 0) s_nlist = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int li = tid + k * kCB;
            const bool act = li < T.n && ((s_bond[li] & inner[k]) >> d) & 1u;
            append(s_list, &s_nlist, act, (uint16_t)li);
        }
        __syncthreads();
        const int n = s_nlist;
        for (int e = tid; e < n; e += kCB) {
            const uint32_t li = s_list[e];
            lds_union(s_lab, li, li + (uint32_t)step[d]);
        }
        __syncthreads();
    }

    // every site's label: the site index of its tile-local root (the smallest
    // index of its tile-local cluster); the bond byte gains bit 4 at those roots
    int32_t* lab = a.label + C.lab0 + (int64_t)T.r * C.nsites;
    for (int li = tid; li < T.n; li += kCB) {
        int xl, yl, zl, rx, ry, rz;
        decode(T, (uint32_t)li, &xl, &yl, &zl);
        const uint32_t rt = lds_find(s_lab, (uint32_t)li);
        decode(T, rt, &rx, &ry, &rz);
        const uint32_t site = site_of(C, T.x0 + xl, T.y0 + yl, T.z0 + zl);
        lab[site] = (int32_t)site_of(C, T.x0 + rx, T.y0 + ry, T.z0 + rz);
        bond[site] = (uint8_t)(s_bond[li] | (rt == (uint32_t)li ? 16u : 0u));
    }
}

// ---------------------------------------------------------------------------
// Bonds that leave a tile through its + faces: union of the two tile-local
// roots (finds start at the roots, so the labels of the other sites keep
// pointing at their tile-local root).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kCB) void cluster_merge_kernel(ClusterArgs a) {
    const int g = find_cgroup(a.cg, a.ngroups, blockIdx.x);
    const ClusterGroup C = a.cg[g];
    const Tile T = tile_of(C, (int)blockIdx.x - C.blk0);
    const int L = C.L, tid = threadIdx.x;
    const uint8_t* bond = a.bond + C.lab0 + (int64_t)T.r * C.nsites;
    int32_t* lab = a.label + C.lab0 + (int64_t)T.r * C.nsites;
    // this thread's site on each + face (a face has at most kCB sites) and its neighbour across
    uint32_t fi[3], fj[3];
    bool on[3] = {false, false, false};
    if (tid < T.ey * T.ez) {
        const int zl = (int)fastdiv((uint32_t)tid, T.my, T.sy), yl = tid - zl * T.ey;
        const int x = T.x0 + T.ex - 1;
        fi[0] = site_of(C, x, T.y0 + yl, T.z0 + zl);
        fj[0] = site_of(C, x + 1 == L ? 0 : x + 1, T.y0 + yl, T.z0 + zl);
        on[0] = true;
    }
    if (tid < T.ex * T.ez) {
        const int zl = (int)fastdiv((uint32_t)tid, T.mx, T.sx), xl = tid - zl * T.ex;
        const int y = T.y0 + T.ey - 1;
        fi[1] = site_of(C, T.x0 + xl, y, T.z0 + zl);
        fj[1] = site_of(C, T.x0 + xl, y + 1 == L ? 0 : y + 1, T.z0 + zl);
        on[1] = true;
    }
    if (C.dims == 3 && tid < T.ex * T.ey) {
        const int yl = (int)fastdiv((uint32_t)tid, T.mx, T.sx), xl = tid - yl * T.ex;
        const int z = T.z0 + T.ez - 1;
        fi[2] = site_of(C, T.x0 + xl, T.y0 + yl, z);
        fj[2] = site_of(C, T.x0 + xl, T.y0 + yl, z + 1 == L ? 0 : z + 1);
        on[2] = true;
    }
    uint32_t bb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) bb[d] = on[d] ? bond[fi[d]] : 0u;
    int32_t ra[3], rb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        on[d] = on[d] && ((bb[d] >> d) & 1u);
        if (on[d]) {
            ra[d] = g_load(&lab[fi[d]]);
            rb[d] = g_load(&lab[fj[d]]);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        if (on[d]) g_union(lab, ra[d], rb[d]);
}

// ---------------------------------------------------------------------------
// Roots and flips.  Phase 1 resolves the cluster root of every tile-local root
// and decides its flip (memoised in LDS); phase 2 applies it to every site
// through its tile-local root.  Each thread's sites are processed together so
// their loads overlap.
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(kCB) void cluster_flip_kernel(ClusterArgs a) {
    using P = typename std::conditional<NC == 0, uint8_t, float>::type;
    __shared__ uint8_t s_flip[kClusterTile];
    __shared__ uint16_t s_list[kClusterTile];
    __shared__ int s_nlist;
    __shared__ int32_t s_seed;
    __shared__ int64_t s_cnt[kCB / 64];
    const int g = find_cgroup(a.cg, a.ngroups, blockIdx.x);
    const ClusterGroup C = a.cg[g];
    const DevGroup G = a.groups[g];
    const Tile T = tile_of(C, (int)blockIdx.x - C.blk0);
    const uint32_t stream = stream_id((uint32_t)C.L, (uint32_t)T.r);
    const uint8_t* bond = a.bond + C.lab0 + (int64_t)T.r * C.nsites;
    int32_t* lab = a.label + C.lab0 + (int64_t)T.r * C.nsites;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_nlist = 0;
        if (a.kind == 1) {
            const u32x4 w = philox(0u, a.u, stream, tag(P_WSEED, NC == 0 ? MODEL_ISING : a.model), a.k0, a.k1);
            const uint64_t v = ((uint64_t)w.x << 32) | w.y;
            const int32_t seed = (int32_t)__umul64hi(v, (uint64_t)C.nsites);
            s_seed = g_find(lab, g_load(&lab[seed]));
        }
    }
    __syncthreads();
    uint32_t site[kPer], bb[kPer];
    int xs[kPer], ys[kPer], zs[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int li = tid + k * kCB;
        bb[k] = 0u;
        if (li < T.n) {
            decode(T, (uint32_t)li, &xs[k], &ys[k], &zs[k]);
            site[k] = site_of(C, T.x0 + xs[k], T.y0 + ys[k], T.z0 + zs[k]);
            bb[k] = bond[site[k]] | 32u;  // bit 5: in the tile
        }
    }
    // phase 1: the tile-local roots, compacted into a list, walk their parent
    // chains in step (kPer per thread at most: a tile has n <= kPer * kCB sites)
#pragma unroll
    for (int k = 0; k < kPer; ++k) append(s_list, &s_nlist, (bb[k] & 16u) != 0u, (uint16_t)(tid + k * kCB));
    __syncthreads();
    const int nroots = s_nlist;
    int32_t cur[kPer], own[kPer];
    bool walk[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int e = tid + k * kCB;
        walk[k] = e < nroots;
        if (walk[k]) {
            int xl, yl, zl;
            decode(T, s_list[e], &xl, &yl, &zl);
            own[k] = cur[k] = (int32_t)site_of(C, T.x0 + xl, T.y0 + yl, T.z0 + zl);
        }
    }
    bool mine[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) mine[k] = walk[k];
    for (;;) {
        int32_t par[kPer];
        bool any = false;
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (walk[k]) par[k] = g_load(&lab[cur[k]]);
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if (walk[k]) {
                if (par[k] == cur[k]) walk[k] = false;
                else cur[k] = par[k];
                any |= walk[k];
            }
        if (!any) break;
    }
    // point each tile-local root straight at its cluster root (later walks of
    // other tiles through it get shorter), and decide its flip
    uint32_t coin[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (mine[k]) {
            if (cur[k] != own[k])
                __hip_atomic_store(&lab[own[k]], cur[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (a.kind == 0) coin[k] = bond[cur[k]];
        }
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if (mine[k])
            s_flip[s_list[tid + k * kCB]] = (uint8_t)(a.kind == 1 ? cur[k] == s_seed : (coin[k] >> 3) & 1u);
    __syncthreads();

    // phase 2: every site takes the decision of its tile-local root
    float rv[3] = {0.f, 0.f, 0.f};
    if constexpr (NC > 0) direction<NC>(a, stream, rv);
    P* base[2];
    if constexpr (NC == 0) {
        base[0] = a.ipool + G.off + 2 * (int64_t)T.r * G.stride;
        base[1] = base[0] + G.stride;
    } else {
        base[0] = a.fpool + G.off + 2 * (int64_t)T.r * NC * G.stride;
        base[1] = base[0] + NC * G.stride;
    }
    const uint32_t tbase = site_of(C, T.x0, T.y0, T.z0);
    const uint32_t L = (uint32_t)C.L;
    int32_t lr[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k)
        if ((bb[k] & 48u) == 32u) lr[k] = lab[site[k]];  // not a tile-local root: its label names that root
    int cnt = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (!(bb[k] & 32u)) continue;
        int li = tid + k * kCB;
        if (!(bb[k] & 16u)) {
            const uint32_t d = (uint32_t)lr[k] - tbase;
            const uint32_t q = fastdiv(d, C.mag_L, C.sh_L);
            const uint32_t dz = fastdiv(q, C.mag_L, C.sh_L);
            li = (int)((dz * (uint32_t)T.ey + (q - dz * L)) * (uint32_t)T.ex + (d - q * L));
        }
        if (!s_flip[li]) continue;
        ++cnt;
        const int x = T.x0 + xs[k], y = T.y0 + ys[k], z = T.z0 + zs[k];
        if constexpr (NC == 0) {
            uint8_t* p = at(base, G, x, y, z);
            *p ^= 1u;
        } else {
            float* p = at(base, G, x, y, z);
            const float s2 = 2.f * project<NC>(p, G.stride, rv);
#pragma unroll
            for (int c = 0; c < NC; ++c) p[c * G.stride] -= s2 * rv[c];
        }
    }
    if (!a.flipped) return;
    const int w = wave_sum(cnt);
    if ((tid & 63) == 0) s_cnt[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) {
        int64_t t = 0;
        for (int i = 0; i < kCB / 64; ++i) t += s_cnt[i];
        if (t) atomicAdd((unsigned long long*)&a.flipped[C.rep0 + T.r], (unsigned long long)t);
    }
}

// Per-replica sums of continuous measurement partials (blocks of a replica are
// consecutive), in a fixed order: one wave per replica.
__global__ __launch_bounds__(kCB) void cont_fold_kernel(const DevGroup* groups, int ngroups, int total_reps,
                                                        const double* part, double* out) {
    const int lane = threadIdx.x & 63;
    const int gr = blockIdx.x * (kCB / 64) + (threadIdx.x >> 6);
    if (gr >= total_reps) return;
    int g = 0;
    while (g + 1 < ngroups && gr >= groups[g + 1].rep0) ++g;
    const int64_t tiles = groups[g].tiles;
    const int64_t b0 = groups[g].blk0 + (int64_t)(gr - groups[g].rep0) * tiles;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t b = lane; b < tiles; b += 64)
        for (int k = 0; k < 4; ++k) v[k] += part[4 * (b0 + b) + k];
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 4; ++k) v[k] += __shfl_down(v[k], off);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) out[4 * (int64_t)gr + k] = v[k];
}

hipError_t launch_cluster_update(const ClusterArgs& a, int nblocks, hipStream_t st) {
    if (a.ncomp == 0) hipLaunchKernelGGL(cluster_bond_kernel<0>, dim3(nblocks), dim3(kCB), 0, st, a);
    else if (a.ncomp == 2) hipLaunchKernelGGL(cluster_bond_kernel<2>, dim3(nblocks), dim3(kCB), 0, st, a);
    else hipLaunchKernelGGL(cluster_bond_kernel<3>, dim3(nblocks), dim3(kCB), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cluster_merge_kernel, dim3(nblocks), dim3(kCB), 0, st, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.ncomp == 0) hipLaunchKernelGGL(cluster_flip_kernel<0>, dim3(nblocks), dim3(kCB), 0, st, a);
    else if (a.ncomp == 2) hipLaunchKernelGGL(cluster_flip_kernel<2>, dim3(nblocks), dim3(kCB), 0, st, a);
    else hipLaunchKernelGGL(cluster_flip_kernel<3>, dim3(nblocks), dim3(kCB), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_cont_fold(const DevGroup* groups, int ngroups, int total_reps, const double* part, double* out,
                            hipStream_t st) {
    const int nb = (total_reps + kCB / 64 - 1) / (kCB / 64);
    hipLaunchKernelGGL(cont_fold_kernel, dim3(nb), dim3(kCB), 0, st, groups, ngroups, total_reps, part, out);
    return hipGetLastError();
}

}  // namespace ising
