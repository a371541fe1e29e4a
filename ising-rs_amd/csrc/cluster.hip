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
//   Ising bond   c0 = row * ceil(L/4) + x/4, tag P_BOND; word x & 3 belongs to
//                site x: byte d is the 8-bit draw of bond d, bit 31 the coin.
//                Bond active iff aligned (J > 0) / anti-aligned (J < 0) and
//                draw < hi, or draw == hi and the 24-bit tie draw < lo, where
//                hi * 2^24 + lo = floor(2^32 (1 - exp(-2 beta |J|))).
//   Ising tie    c0 = 4 * (Philox group) + (x & 3), tag P_BTIE; word d >> 8.
//   cont. bond   c0 = site index, tag P_BOND; word d -> 24-bit uniform, bit 31 of word 3 the coin.
//   cont. r      c0 = 0, tag P_CDIR: XY phi = 2 pi u01(w0, w1); Heisenberg
//                z = 2 u01(w0, w1) - 1, phi = 2 pi u01(w2, w3).
//   Wolff seed   c0 = 0, tag P_WSEED: site = floor(N * ((w0 << 32) | w1) / 2^64).
#include "kernels.h"

namespace ising {

namespace {

constexpr int kCB = 256;  // threads per block

__device__ __forceinline__ int find_cgroup(const ClusterGroup* cg, int ngroups, int b) {
    int g = 0;
    while (g + 1 < ngroups && b >= cg[g + 1].blk0) ++g;
    return g;
}

// A block's tile: replica, origin and extents.
struct Tile {
    int r, x0, y0, z0, ex, ey, ez, n;
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
    return T;
}

__device__ __forceinline__ int64_t site_of(const ClusterGroup& C, int x, int y, int z) {
    return ((int64_t)z * C.L + y) * C.L + x;
}

// Half-lattice element of site (x, y, z): colour and index within the colour.
__device__ __forceinline__ int64_t half_index(const DevGroup& G, int x, int y, int z, int* colour) {
    *colour = (x + y + z) & 1;
    return ((int64_t)z * G.L + y) * G.Lh + (x >> 1);
}

__device__ __forceinline__ uint8_t* ising_byte(const ClusterArgs& a, const DevGroup& G, int r, int x, int y, int z) {
    int c;
    const int64_t h = half_index(G, x, y, z, &c);
    return a.ipool + G.off + (2 * (int64_t)r + c) * G.stride + h;
}

template <int NC>
__device__ __forceinline__ float* cont_ptr(const ClusterArgs& a, const DevGroup& G, int r, int x, int y, int z) {
    int c;
    const int64_t h = half_index(G, x, y, z, &c);
    return a.fpool + G.off + (2 * (int64_t)r + c) * NC * G.stride + h;
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

__device__ __forceinline__ uint32_t lds_find(uint32_t* lab, uint32_t x) {
    uint32_t p = __hip_atomic_load(&lab[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&lab[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
// Loads bypass the non-coherent vector L1 (agent-scope atomics), so links made
// by other compute units are seen.  Path halving stores only ever replace a
// parent by one of its ancestors.

__device__ __forceinline__ int32_t g_load(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int32_t g_find(int32_t* lab, int32_t x) {
    int32_t p = g_load(&lab[x]);
    while (p != x) {
        const int32_t gp = g_load(&lab[p]);
        if (gp != p) __hip_atomic_store(&lab[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    __shared__ S s_v[kClusterTile];
    __shared__ S s_h[3][kClusterFace];
    __shared__ uint8_t s_bond[kClusterTile];
    __shared__ uint32_t s_lab[kClusterTile];

    const int g = find_cgroup(a.cg, a.ngroups, blockIdx.x);
    const ClusterGroup C = a.cg[g];
    const DevGroup G = a.groups[g];
    const Tile T = tile_of(C, (int)blockIdx.x - C.blk0);
    const int L = C.L, dims = C.dims;
    const uint32_t stream = stream_id((uint32_t)L, (uint32_t)T.r);
    const int tid = threadIdx.x;

    float rv[3] = {0.f, 0.f, 0.f};
    if constexpr (NC > 0) direction<NC>(a, stream, rv);

    auto value = [&](int x, int y, int z) -> S {
        if constexpr (NC == 0) return (S)*ising_byte(a, G, T.r, x, y, z);
        else return project<NC>(cont_ptr<NC>(a, G, T.r, x, y, z), G.stride, rv);
    };

    for (int li = tid; li < T.n; li += kCB) {
        const int row = li / T.ex, xl = li - row * T.ex;
        const int zl = row / T.ey, yl = row - zl * T.ey;
        s_v[li] = value(T.x0 + xl, T.y0 + yl, T.z0 + zl);
        s_lab[li] = (uint32_t)li;
    }
    {
        const int xn = (T.x0 + T.ex) % L, yn = (T.y0 + T.ey) % L, zn = (T.z0 + T.ez) % L;
        for (int k = tid; k < T.ey * T.ez; k += kCB) {
            const int zl = k / T.ey, yl = k - zl * T.ey;
            s_h[0][k] = value(xn, T.y0 + yl, T.z0 + zl);
        }
        for (int k = tid; k < T.ex * T.ez; k += kCB) {
            const int zl = k / T.ex, xl = k - zl * T.ex;
            s_h[1][k] = value(T.x0 + xl, yn, T.z0 + zl);
        }
        if (dims == 3)
            for (int k = tid; k < T.ex * T.ey; k += kCB) {
                const int yl = k / T.ex, xl = k - yl * T.ex;
                s_h[2][k] = value(T.x0 + xl, T.y0 + yl, zn);
            }
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
        const uint32_t hi = thr >> 24, lo = thr & 0xffffffu;
        const bool ferro = a.J > 0.0;
        const int gq = (T.ex + 3) >> 2;  // Philox groups per tile row (x0 is a multiple of 4)
        for (int k = tid; k < T.ey * T.ez * gq; k += kCB) {
            const int row = k / gq, q = k - row * gq;
            const int zl = row / T.ey, yl = row - zl * T.ey;
            const int y = T.y0 + yl, z = T.z0 + zl;
            const uint32_t grp = (uint32_t)(((int64_t)z * L + y) * C.q4 + (T.x0 >> 2) + q);
            const u32x4 w4 = philox(grp, a.u, stream, tag(P_BOND, MODEL_ISING), a.k0, a.k1);
            const uint32_t wk[4] = {w4.x, w4.y, w4.z, w4.w};
            for (int s = 0; s < 4; ++s) {
                const int xl = 4 * q + s;
                if (xl >= T.ex) break;
                const int li = (zl * T.ey + yl) * T.ex + xl;
                const S si = s_v[li];
                uint32_t bits = (wk[s] >> 31) << 3;
                for (int d = 0; d < dims; ++d) {
                    const S sj = nbr(li, xl, yl, zl, d);
                    if ((si == sj) != ferro) continue;
                    const uint32_t r8 = (wk[s] >> (8 * d)) & 0xffu;
                    bool act = r8 < hi;
                    if (r8 == hi && lo != 0u) {
                        const u32x4 t4 = philox(4u * grp + (uint32_t)s, a.u, stream, tag(P_BTIE, MODEL_ISING), a.k0,
                                                a.k1);
                        const uint32_t tw = d == 0 ? t4.x : (d == 1 ? t4.y : t4.z);
                        act = (tw >> 8) < lo;
                    }
                    bits |= (uint32_t)act << d;
                }
                s_bond[li] = (uint8_t)bits;
                bond[site_of(C, T.x0 + xl, y, z)] = (uint8_t)bits;
            }
        }
    } else {
        const double beta = a.inline_p ? a.beta_inline : a.betas[tslot(a, C, T.r)];
        const float kappa = (float)(2.0 * beta * a.J);
        for (int li = tid; li < T.n; li += kCB) {
            const int row = li / T.ex, xl = li - row * T.ex;
            const int zl = row / T.ey, yl = row - zl * T.ey;
            const int64_t site = site_of(C, T.x0 + xl, T.y0 + yl, T.z0 + zl);
            const u32x4 w4 = philox((uint32_t)site, a.u, stream, tag(P_BOND, a.model), a.k0, a.k1);
            const uint32_t wk[3] = {w4.x, w4.y, w4.z};
            const float ai = s_v[li];
            uint32_t bits = (w4.w >> 31) << 3;
            for (int d = 0; d < 3; ++d) {
                const float x = kappa * ai * nbr(li, xl, yl, zl, d);
                if (x > 0.f && u24f(wk[d]) < -expm1f(-x)) bits |= 1u << d;
            }
            s_bond[li] = (uint8_t)bits;
            bond[site] = (uint8_t)bits;
        }
    }
    __syncthreads();

    // union-find over the bonds inside the tile
    const int step[3] = {1, T.ex, T.ex * T.ey};
    for (int li = tid; li < T.n; li += kCB) {
        const uint32_t bits = s_bond[li];
        if (!(bits & 7u)) continue;
        const int row = li / T.ex, xl = li - row * T.ex;
        const int zl = row / T.ey, yl = row - zl * T.ey;
        const int cc[3] = {xl + 1 < T.ex, yl + 1 < T.ey, zl + 1 < T.ez};
        for (int d = 0; d < dims; ++d)
            if (((bits >> d) & 1u) && cc[d]) lds_union(s_lab, (uint32_t)li, (uint32_t)(li + step[d]));
    }
    __syncthreads();

    int32_t* lab = a.label + C.lab0 + (int64_t)T.r * C.nsites;
    for (int li = tid; li < T.n; li += kCB) {
        const int row = li / T.ex, xl = li - row * T.ex;
        const int zl = row / T.ey, yl = row - zl * T.ey;
        const int rt = (int)lds_find(s_lab, (uint32_t)li);
        const int rrow = rt / T.ex, rxl = rt - rrow * T.ex;
        const int rzl = rrow / T.ey, ryl = rrow - rzl * T.ey;
        lab[site_of(C, T.x0 + xl, T.y0 + yl, T.z0 + zl)] =
            (int32_t)site_of(C, T.x0 + rxl, T.y0 + ryl, T.z0 + rzl);
    }
}

// ---------------------------------------------------------------------------
// Bonds that leave a tile through its + faces.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kCB) void cluster_merge_kernel(ClusterArgs a) {
    const int g = find_cgroup(a.cg, a.ngroups, blockIdx.x);
    const ClusterGroup C = a.cg[g];
    const Tile T = tile_of(C, (int)blockIdx.x - C.blk0);
    const int L = C.L;
    const uint8_t* bond = a.bond + C.lab0 + (int64_t)T.r * C.nsites;
    int32_t* lab = a.label + C.lab0 + (int64_t)T.r * C.nsites;
    const int tid = threadIdx.x;
    {
        const int x = T.x0 + T.ex - 1, xn = (x + 1) % L;
        for (int k = tid; k < T.ey * T.ez; k += kCB) {
            const int zl = k / T.ey, yl = k - zl * T.ey;
            const int64_t i = site_of(C, x, T.y0 + yl, T.z0 + zl);
            if (bond[i] & 1u) g_union(lab, (int32_t)i, (int32_t)site_of(C, xn, T.y0 + yl, T.z0 + zl));
        }
    }
    {
        const int y = T.y0 + T.ey - 1, yn = (y + 1) % L;
        for (int k = tid; k < T.ex * T.ez; k += kCB) {
            const int zl = k / T.ex, xl = k - zl * T.ex;
            const int64_t i = site_of(C, T.x0 + xl, y, T.z0 + zl);
            if (bond[i] & 2u) g_union(lab, (int32_t)i, (int32_t)site_of(C, T.x0 + xl, yn, T.z0 + zl));
        }
    }
    if (C.dims == 3) {
        const int z = T.z0 + T.ez - 1, zn = (z + 1) % L;
        for (int k = tid; k < T.ex * T.ey; k += kCB) {
            const int yl = k / T.ex, xl = k - yl * T.ex;
            const int64_t i = site_of(C, T.x0 + xl, T.y0 + yl, z);
            if (bond[i] & 4u) g_union(lab, (int32_t)i, (int32_t)site_of(C, T.x0 + xl, T.y0 + yl, zn));
        }
    }
}

// ---------------------------------------------------------------------------
// Roots and flips.
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(kCB) void cluster_flip_kernel(ClusterArgs a) {
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
    if (a.kind == 1) {
        if (tid == 0) {
            const u32x4 w = philox(0u, a.u, stream, tag(P_WSEED, NC == 0 ? MODEL_ISING : a.model), a.k0, a.k1);
            const uint64_t v = ((uint64_t)w.x << 32) | w.y;
            const int32_t seed = (int32_t)__umul64hi(v, (uint64_t)C.nsites);
            s_seed = g_find(lab, seed);
        }
        __syncthreads();
    }
    float rv[3] = {0.f, 0.f, 0.f};
    if constexpr (NC > 0) direction<NC>(a, stream, rv);
    int cnt = 0;
    for (int li = tid; li < T.n; li += kCB) {
        const int row = li / T.ex, xl = li - row * T.ex;
        const int zl = row / T.ey, yl = row - zl * T.ey;
        const int x = T.x0 + xl, y = T.y0 + yl, z = T.z0 + zl;
        const int32_t rt = g_find(lab, (int32_t)site_of(C, x, y, z));
        const bool flip = a.kind == 1 ? rt == s_seed : ((bond[rt] >> 3) & 1u) != 0u;
        if (!flip) continue;
        ++cnt;
        if constexpr (NC == 0) {
            uint8_t* p = ising_byte(a, G, T.r, x, y, z);
            *p ^= 1u;
        } else {
            float* p = cont_ptr<NC>(a, G, T.r, x, y, z);
            const float s2 = 2.f * project<NC>(p, G.stride, rv);
#pragma unroll
            for (int k = 0; k < NC; ++k) p[k * G.stride] -= s2 * rv[k];
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
