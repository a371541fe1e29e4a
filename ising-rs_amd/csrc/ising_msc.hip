// Multispin-coded 3D Ising (h = 0, J > 0), gfx950.
//
// The replicas of one lattice size are packed across the bits of a 32-bit
// word per site: bit r % 32 of word r / 32 holds replica r.  A colour
// half-sweep of one word updates 32 replicas of one site at once:
//
//   A_i   = ~(S ^ N_i)              aligned-neighbour bits, i = 1..6
//   a     = sum_i A_i               bit-sliced adder tree -> 3 bit planes
//   dE    = 4 a - 12                (J = 1): dE > 0 iff a >= 4, i.e. plane a2 set
//   flip  = ~a2 | (u < T(a, beta_r))
//
// The 32-bit uniforms u_r are compared with the per-replica thresholds bit
// plane by bit plane from the most significant end, drawing four planes per
// Philox call (counter (h, 2t + c, (L << 16) | word, (6 << 8) | q)) only while
// some replica of the word is undecided.  This is exactly the comparison of a
// full 32-bit uniform, so the chain is the same Metropolis chain as the int8
// path, with far fewer random bits per update.  Storage: 4 bytes per site for
// 32 replicas, i.e. 1 bit per spin (a half-sweep moves 0.375 B per update).
#include "sweepctl.h"

namespace ising {

constexpr int kMT = 256;  // threads per block, one site word per thread

__device__ __forceinline__ int msc_block() {
    const int per = (int)(gridDim.x >> 3);
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

__device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t* s, uint32_t* k) {
    const uint32_t t = a ^ b;
    *s = t ^ c;
    *k = (a & b) | (t & c);
}

// 3-bit plane count (x2 x1 x0) of six one-bit inputs per lane.
__device__ __forceinline__ void count6(const uint32_t v[6], uint32_t* x0, uint32_t* x1, uint32_t* x2) {
    uint32_t s1, k1, s2, k2;
    full_add(v[0], v[1], v[2], &s1, &k1);
    full_add(v[3], v[4], v[5], &s2, &k2);
    *x0 = s1 ^ s2;
    const uint32_t k0 = s1 & s2;
    full_add(k1, k2, k0, x1, x2);
}

// Column sums of the wave's 64 x 32 bit matrix (one word per lane): lane l < 32
// returns the number of lanes whose word has bit l set.  Each 32-lane half is
// transposed as a 32 x 32 bit matrix in five butterfly stages (partner lane
// l ^ j by ds_swizzle, a rotation by j, a masked merge that swaps the j x j
// off-diagonal blocks), after which lane l holds column l of its half; one
// popcount per half and a cross-half add give the counts.  All 64 lanes must
// be active.
template <int J>
__device__ __forceinline__ uint32_t transpose_stage(uint32_t w, uint32_t lane) {
    constexpr uint32_t m = J == 16 ? 0x0000ffffu : J == 8 ? 0x00ff00ffu : J == 4 ? 0x0f0f0f0fu
                                                       : J == 2 ? 0x33333333u : 0x55555555u;
    const uint32_t p = (uint32_t)__builtin_amdgcn_ds_swizzle((int)w, 0x1f | (J << 10));  // lane ^ J
    const bool upper = (lane & (uint32_t)J) != 0u;
    // lower lanes take the partner's low block shifted up, upper lanes its high block shifted down
    const uint32_t rot = __builtin_amdgcn_alignbit(p, p, upper ? (uint32_t)J : 32u - (uint32_t)J);
    const uint32_t keep = upper ? ~m : m;
    return (w & keep) | (rot & ~keep);
}

__device__ __forceinline__ int column_count(uint32_t w) {
    const uint32_t lane = threadIdx.x & 63u;
    w = transpose_stage<16>(w, lane);
    w = transpose_stage<8>(w, lane);
    w = transpose_stage<4>(w, lane);
    w = transpose_stage<2>(w, lane);
    w = transpose_stage<1>(w, lane);
    const int c = __builtin_popcount(w);
    return c + __shfl_xor(c, 32);
}

__device__ __forceinline__ int find_group_m(const DevGroup* groups, int ngroups, int b) {
    int g = 0;
    while (g + 1 < ngroups && b >= groups[g + 1].blk0) ++g;
    return g;
}

__device__ __forceinline__ void msc_nbrs(const DevGroup& G, uint32_t h, int color, uint32_t nb[6]) {
    const uint32_t L = (uint32_t)G.L, Lh = (uint32_t)G.Lh;
    const uint32_t row = Lh >= 2 ? fastdiv(h, G.mag_Lh, G.sh_Lh) : h;
    const uint32_t xh = h - row * Lh;
    const uint32_t z = fastdiv(row, G.mag_L, G.sh_L), y = row - z * L;
    const uint32_t p = (y + z + (uint32_t)color) & 1u;
    const uint32_t ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
    const uint32_t zm = z == 0 ? L - 1 : z - 1, zp = z == L - 1 ? 0 : z + 1;
    const uint32_t xo = p ? (xh + 1 == Lh ? 0 : xh + 1) : (xh == 0 ? Lh - 1 : xh - 1);
    nb[0] = row * Lh + xh;
    nb[1] = row * Lh + xo;
    nb[2] = (z * L + ym) * Lh + xh;
    nb[3] = (z * L + yp) * Lh + xh;
    nb[4] = (zm * L + y) * Lh + xh;
    nb[5] = (zp * L + y) * Lh + xh;
}

// One colour half-sweep of every replica word of every size.
// Grid: (group, word, tile of 256 sites), padded to a multiple of 8 blocks.
__global__ __launch_bounds__(kMT) void msc_half_kernel(IsingArgs a, int color) {
    __shared__ uint32_t s_tp[3][32];   // threshold bit planes (MSB first) per class a = 4, 5, 6
    __shared__ int s_acc[32][2];
    const int b = msc_block();
    if (b >= a.nblocks) return;
    const int g = find_group_m(a.groups, a.ngroups, b);
    const DevGroup G = a.groups[g];
    int wi, tile;
    split_tiles(G, b - G.blk0, &wi, &tile);
    const int nrep = min(32, G.R - 32 * wi);
    const uint32_t valid = nrep >= 32 ? 0xffffffffu : ((1u << nrep) - 1u);
    const int tid = threadIdx.x;
    const SweepView v = view_of(a.ctl, color);
    if (v.exch) {
        // this launch closes an exchange sweep: the word's threshold planes
        // for the new temperatures, one ballot per plane (lane = replica bit)
        const int64_t po = (v.round & 1) * a.total_reps;
        const int64_t* pacc = acc_of(a, v.t - 1);
        if (b == G.blk0) exch_write_group(a, G, v.round, pacc);
        const int lane = tid & 63, wave = tid >> 6;
        uint32_t T[3] = {0u, 0u, 0u};
        if (lane < nrep) {
            const int ti = exch_tidx(a, G, 32 * wi + lane, v.round, a.r2t + po, a.t2r + po, pacc);
            for (int k = 0; k < 3; ++k) T[k] = a.mthr[3 * (G.rep0 + ti) + k];
        }
        for (int k = 0; k < 3; ++k)
            for (int j = 8 * wave; j < 8 * wave + 8; ++j) {
                const uint64_t m = __ballot((T[k] >> (31 - j)) & 1u);
                if (lane == 0) s_tp[k][j] = (uint32_t)m;
            }
    } else if (tid < 96) {
        s_tp[tid >> 5][tid & 31] = a.mplanes[(int64_t)(G.rep0 + 32 * wi) * 96 + tid];
    }
    if (b == 0) sweep_bookkeeping(a, v, color);
    if (tid < 64) s_acc[tid >> 1][tid & 1] = 0;
    const int64_t t = v.t;
    const bool need = (v.flags & (F_MEAS | F_EXCH)) != 0;
    __syncthreads();
    if (v.exch && tile == 0 && tid < 96) a.mplanes[(int64_t)(G.rep0 + 32 * wi) * 96 + tid] = s_tp[tid >> 5][tid & 31];

    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = ((uint32_t)G.L << 16) | (uint32_t)wi;
    uint32_t* pool = (uint32_t*)a.pool;
    uint32_t* C = pool + G.off + (int64_t)(2 * wi + color) * G.stride;
    const uint32_t* O = pool + G.off + (int64_t)(2 * wi + (color ^ 1)) * G.stride;
    const uint32_t h = (uint32_t)tile * kMT + (uint32_t)tid;
    const bool live = (int64_t)h < G.n_half;
    uint32_t ns = 0, x0 = 0, x1 = 0, x2 = 0;
    if (live) {
        uint32_t nb[6];
        msc_nbrs(G, h, color, nb);
        const uint32_t S = C[h];
        uint32_t A[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) A[i] = ~(S ^ O[nb[i]]);
        uint32_t a0, a1, a2;
        count6(A, &a0, &a1, &a2);
        uint32_t und = a2 & valid;
        const uint32_t K4 = und & ~a1 & ~a0, K5 = und & a0, K6 = und & a1;
        uint32_t acc = ~a2 & valid;
#pragma unroll 1
        for (int q = 0; q < 8 && und; ++q) {
            const u32x4 w = philox(h, hs, stream, (P_MSC << 8) | (uint32_t)q, a.k0, a.k1);
            const uint32_t U[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                const int j = 4 * q + p;
                const uint32_t P = (K4 & s_tp[0][j]) | (K5 & s_tp[1][j]) | (K6 & s_tp[2][j]);
                acc |= und & ~U[p] & P;
                und &= ~(U[p] ^ P);
            }
        }
        ns = S ^ acc;
        C[h] = ns;
        if (need && color) {
            uint32_t D[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) D[i] = ns ^ O[nb[i]];  // anti-aligned bits of the new spin
            count6(D, &x0, &x1, &x2);
        }
    }
    if (need) {
        // per replica r: M_r += 2 * #up - #sites, B_r += sum over sites of (6 - 2 * anti_r)
        // (ns and x0..x2 are 0 on lanes without a site); lane r of the wave gets replica r's counts
        const uint64_t lv = __ballot(live);
        const int nl = __popcll(lv);
        const int up = column_count(ns);
        int anti = 0;
        if (color) anti = column_count(x0) + 2 * column_count(x1) + 4 * column_count(x2);
        const int r = tid & 63;
        if (r < nrep) {
            atomicAdd(&s_acc[r][1], 2 * up - nl);
            if (color) atomicAdd(&s_acc[r][0], 6 * nl - 2 * anti);
        }
        __syncthreads();
        if (tid < 2 * nrep) {
            const int r = tid >> 1, q = tid & 1;
            const int val = s_acc[r][q];
            if (val != 0 && (q == 1 || color)) {
                const int copy = b & (kMscAccCopies - 1);
                atomicAdd((unsigned long long*)&acc_of(a, t)[2 * ((int64_t)copy * a.total_reps + G.rep0 + 32 * wi + r) + q],
                          (unsigned long long)(int64_t)val);
            }
        }
    }
}

// Hot start: replica r's spin at site s is Philox(s, 0, (L << 16) | r, init).x >> 31,
// the same start as the int8 path.
__global__ __launch_bounds__(kMT) void msc_init_kernel(const DevGroup* groups, int ngroups, uint32_t* pool,
                                                     int nblocks, uint32_t k0, uint32_t k1) {
    const int b = blockIdx.x;
    if (b >= nblocks) return;
    const int g = find_group_m(groups, ngroups, b);
    const DevGroup G = groups[g];
    int wi, tile;
    split_tiles(G, b - G.blk0, &wi, &tile);
    const int nrep = min(32, G.R - 32 * wi);
    const int64_t h = (int64_t)tile * kMT + threadIdx.x;
    if (h >= G.n_half) return;
    const int64_t row = h / G.Lh;
    const int xh = (int)(h - row * G.Lh);
    const int z = (int)(row / G.L), y = (int)(row - (int64_t)z * G.L);
    for (int c = 0; c < 2; ++c) {
        const int x = 2 * xh + ((y + z + c) & 1);
        const uint32_t site = (uint32_t)(((int64_t)z * G.L + y) * G.L + x);
        uint32_t v = 0;
        for (int bit = 0; bit < nrep; ++bit) {
            const uint32_t st = stream_id((uint32_t)G.L, (uint32_t)(32 * wi + bit));
            v |= (philox(site, 0u, st, tag(P_INIT, MODEL_ISING), k0, k1).x >> 31) << bit;
        }
        pool[G.off + (int64_t)(2 * wi + c) * G.stride + h] = v;
    }
}

// Standalone E/M of every replica: acc copy 0 += (B, M) per replica.
__global__ __launch_bounds__(kMT) void msc_measure_kernel(const DevGroup* groups, int ngroups, const uint32_t* pool,
                                                        int total_reps, int nblocks, int64_t* acc) {
    __shared__ int s_acc[32][2];
    const int b = blockIdx.x;
    if (b >= nblocks) return;
    const int g = find_group_m(groups, ngroups, b);
    const DevGroup G = groups[g];
    int wi, tile;
    split_tiles(G, b - G.blk0, &wi, &tile);
    const int nrep = min(32, G.R - 32 * wi);
    const int tid = threadIdx.x;
    if (tid < 64) s_acc[tid >> 1][tid & 1] = 0;
    __syncthreads();
    const uint32_t h = (uint32_t)tile * kMT + (uint32_t)tid;
    const bool live = (int64_t)h < G.n_half;
    uint32_t s1 = 0, s0 = 0, x0 = 0, x1 = 0, x2 = 0;
    if (live) {
        uint32_t nb[6];
        msc_nbrs(G, h, 1, nb);
        const uint32_t* C1 = pool + G.off + (int64_t)(2 * wi + 1) * G.stride;
        const uint32_t* C0 = pool + G.off + (int64_t)(2 * wi) * G.stride;
        s1 = C1[h];
        s0 = C0[h];
        uint32_t D[6];
        for (int i = 0; i < 6; ++i) D[i] = s1 ^ C0[nb[i]];
        count6(D, &x0, &x1, &x2);
    }
    const int nl = __popcll(__ballot(live));
    for (int r = 0; r < nrep; ++r) {
        const int up = __popcll(__ballot(live && ((s1 >> r) & 1u))) + __popcll(__ballot(live && ((s0 >> r) & 1u)));
        const int anti = __popcll(__ballot((x0 >> r) & 1u)) + 2 * __popcll(__ballot((x1 >> r) & 1u)) +
                         4 * __popcll(__ballot((x2 >> r) & 1u));
        if ((tid & 63) == 0) {
            atomicAdd(&s_acc[r][1], 2 * up - 2 * nl);
            atomicAdd(&s_acc[r][0], 6 * nl - 2 * anti);
        }
    }
    __syncthreads();
    if (tid < 2 * nrep) {
        const int r = tid >> 1, q = tid & 1;
        atomicAdd((unsigned long long*)&acc[2 * (G.rep0 + 32 * wi + r) + q], (unsigned long long)(int64_t)s_acc[r][q]);
    }
    (void)total_reps;
}

hipError_t launch_msc_half(const IsingArgs& a, int nblocks, int color, hipStream_t st) {
    const int grid = (nblocks + 7) / 8 * 8;
    hipLaunchKernelGGL(msc_half_kernel, dim3(grid), dim3(kMT), 0, st, a, color);
    return hipGetLastError();
}

hipError_t launch_msc_init(const DevGroup* groups, int ngroups, uint32_t* pool, int nblocks, uint32_t k0, uint32_t k1,
                           hipStream_t st) {
    hipLaunchKernelGGL(msc_init_kernel, dim3(nblocks), dim3(kMT), 0, st, groups, ngroups, pool, nblocks, k0, k1);
    return hipGetLastError();
}

hipError_t launch_msc_measure(const DevGroup* groups, int ngroups, const uint32_t* pool, int total_reps, int nblocks,
                              int64_t* acc, hipStream_t st) {
    hipLaunchKernelGGL(msc_measure_kernel, dim3(nblocks), dim3(kMT), 0, st, groups, ngroups, pool, total_reps,
                       nblocks, acc);
    return hipGetLastError();
}

}  // namespace ising
