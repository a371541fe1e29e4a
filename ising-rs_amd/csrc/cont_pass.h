// Passes of the continuous models (XY, Heisenberg) over one block tile, and
// the helpers they share: the bodies of the per-ensemble launches in
// cont.hip and of the work-list launch of several ensembles.  Storage,
// proposals and measurement are described in cont.hip.
#pragma once
#include "kernels.h"

namespace ising {

constexpr int kCT = 256;        // sites per block tile
constexpr int kCloseT = 256;  // threads of the closing block (one replica or pair each at a time)

// A value the whole block reads at one address and no block of the running
// launch writes (size descriptors, the ensemble's tables, the sweep counter),
// read through the constant address space: a scalar load also where the
// compiler cannot rule out an earlier store of the kernel to it -- inside a
// persistent grid's item loop, or in the second of two members' bodies laid
// out one after the other -- and issued vector loads (then readfirstlane),
// with the Metropolis items' parameters at several times the latency.
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// A size descriptor by scalar loads, word by word.
__device__ __forceinline__ DevGroup ld_group(const DevGroup* p) {
    static_assert(sizeof(DevGroup) % 4 == 0, "whole words");
    uint32_t w[sizeof(DevGroup) / 4];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(DevGroup) / 4); ++i) w[i] = ld_uniform(reinterpret_cast<const uint32_t*>(p) + i);
    DevGroup G;
    __builtin_memcpy(&G, w, sizeof(G));
    return G;
}

// The size group of block b: a count over every group, with no loop exit that
// depends on the data, so the index stays uniform (scalar loads of its
// descriptor) also inside a persistent grid's item loop.
__device__ __forceinline__ int find_group_c(const DevGroup* groups, int ngroups, int b) {
    int g = 0;
    for (int i = 1; i < ngroups; ++i) g += b >= ld_uniform(&groups[i].blk0) ? 1 : 0;
    return g;
}

// Blocks are dealt to the 8 XCDs round-robin (block i runs on XCD i % 8).  The
// sweep launches pad their grid to a multiple of 8 and renumber the blocks so
// each XCD walks one contiguous range of tiles: the other colour's rows and
// planes a tile reads then stay in that XCD's L2 (blocks past nblocks return).
__device__ __forceinline__ int xcd_block() {
    const int per = (int)(gridDim.x >> 3);
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Hardware transcendentals of the Metropolis passes: v_sin_f32 / v_cos_f32 take
// revolutions, v_exp_f32 is 2^x.  Measured on gfx950 against FP64: |error| of
// sin, cos <= 1.3e-7 for every 24-bit u in [0, 1) (2.7e-7 for rotations up to 3
// rad); relative error of exp2, rsq, sqrt, rcp <= 1e-7.
__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }
__device__ __forceinline__ float exp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float rsqrt_fast(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
constexpr float kInv2Pi = 0.15915494309189533577f;

__device__ __forceinline__ void neighbours(const DevGroup& G, int64_t h, int color, uint32_t nb[6]) {
    site_nbrs(G, (uint32_t)h, color, nb);
}

template <int N, bool O32>
__device__ __forceinline__ void load_nbrs(const float* O, int64_t stride, const uint32_t nb[6], float nv[3][6]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int q = 0; q < 6; ++q) nv[k][q] = *comp_at<O32>(O, stride, k, nb[q]);
    }
}

template <int N>
__device__ __forceinline__ void field_f32(const float nv[3][6], float hf[3]) {
    hf[0] = hf[1] = hf[2] = 0.f;
#pragma unroll
    for (int k = 0; k < N; ++k) hf[k] = ((nv[k][0] + nv[k][1]) + (nv[k][2] + nv[k][3])) + (nv[k][4] + nv[k][5]);
}

template <int N, bool O32>
__device__ __forceinline__ void load_field(const float* O, int64_t stride, const uint32_t nb[6], float hf[3]) {
    float nv[3][6];
    load_nbrs<N, O32>(O, stride, nb, nv);
    field_f32<N>(nv, hf);
}

// Block sum of up to 4 doubles; result valid in thread 0.  A Metropolis pass
// passes its model's N and colour: only the partials that can be non-zero are
// summed across the wave (wave_sum_parts).
template <int NT = kCT, int N = 0>
__device__ __forceinline__ void block_sum4(double v[4], double (*s_red)[kCT / 64], int color = 1) {
    if (N > 0) {
        wave_sum_parts<N>(v, color);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 4; ++k) s_red[k][wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            double s = 0.0;
            for (int i = 0; i < NT / 64; ++i) s += s_red[k][i];
            v[k] = s;
        }
}

// The same over the waves of one of the block's sub-blocks of NT threads
// (sub-block q: threads q NT ... (q + 1) NT - 1); result valid in the
// sub-block's thread 0.  Every thread of the block calls it (one barrier).
template <int NT, int N>
__device__ __forceinline__ void sub_block_sum4(double v[4], double (*s_red)[kCT / 64], int color, int q, int tid) {
    wave_sum_parts<N>(v, color);
    const int lane = tid & 63, wid = q * (NT / 64) + (tid >> 6);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) s_red[k][wid] = v[k];
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < 4; ++k) {
            double s = 0.0;
            for (int i = 0; i < NT / 64; ++i) s += s_red[k][q * (NT / 64) + i];
            v[k] = s;
        }
}

// Per-replica FP64 sums (E, Mx, My, Mz) of the block partials of a measured
// or exchange sweep, one wave per replica in a fixed order (lane l takes
// tiles l, l + 64, ...; then a fixed shuffle tree), into dacc; the blocks of
// the closing launch share the replicas, so the sums proceed on many CUs.
__device__ __forceinline__ void reduce_replica(const ContArgs& a, const int lb) {
    const int lane = (int)threadIdx.x & 63;
    const int gr = lb * (kCT / 64) + ((int)threadIdx.x >> 6);
    if (gr >= a.total_reps) return;  // whole waves
    int g = 0;
    while (g + 1 < a.ngroups && gr >= a.groups[g + 1].rep0) ++g;
    const int64_t tiles = a.groups[g].tiles;
    const int64_t b0 = a.groups[g].blk0 + (int64_t)(gr - a.groups[g].rep0) * tiles;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    // unrolled so that a lane's loads are in flight together (the sums keep their order)
#pragma unroll 8
    for (int64_t b = lane; b < tiles; b += 64) {
        const double* p0 = a.part0 + 4 * (b0 + b);
        const double* p1 = a.part1 + 4 * (b0 + b);
        v[0] += p1[0];
        v[1] += p0[1] + p1[1];
        v[2] += p0[2] + p1[2];
        v[3] += p0[3] + p1[3];
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 4; ++q) v[q] += __shfl_down(v[q], off);
    if (lane == 0)
        for (int q = 0; q < 4; ++q) a.dacc[4 * gr + q] = v[q];
}

// Closing step of a continuous sweep (one block of cont_close_kernel, after the sums):
// records from the per-replica sums, exchange, delta adaptation during
// warm-up, sweep counter.
static __device__ void close_cont(const ContArgs& a, int64_t t, int32_t fl) {
    const bool meas = (fl & F_MEAS) != 0, exch = (fl & F_EXCH) != 0;
    DevCtl* ctl = a.ctl;
    const int tid = threadIdx.x;
    const int nrec = ctl->n_rec;
    const bool room = nrec < ctl->rec_cap;
    if (meas && room) {
        // the per-replica sums were written by every block of this launch (reduce_replica)
        for (int gr = tid; gr < a.total_reps; gr += kCloseT) {
            ContRec rc;
            rc.e = a.dacc[4 * gr];
            rc.mx = a.dacc[4 * gr + 1];
            rc.my = a.dacc[4 * gr + 2];
            rc.mz = a.dacc[4 * gr + 3];
            rc.tidx = a.r2t[gr];
            rc.pad = 0;
            a.rec[(int64_t)nrec * a.total_reps + gr] = rc;
        }
    }
    __syncthreads();  // records read the permutation before the exchange rewrites it
    if (exch) {
        const int64_t round = ctl->round;
        for (int g = 0; g < a.ngroups; ++g) {
            const DevGroup G = a.groups[g];
            const int npairs = (G.R - (int)(round & 1)) / 2;
            for (int q = tid; q < npairs; q += kCloseT) {
                const int i = (int)(round & 1) + 2 * q;
                const int ri = a.t2r[G.rep0 + i], rj = a.t2r[G.rep0 + i + 1];
                const double Ei = a.dacc[4 * (G.rep0 + ri)];
                const double Ej = a.dacc[4 * (G.rep0 + rj)];
                const double delta = (a.betas[G.rep0 + i] - a.betas[G.rep0 + i + 1]) * (Ej - Ei);
                const u32x4 w = philox((uint32_t)i, (uint32_t)round, stream_id((uint32_t)G.L, 0u),
                                       tag(P_EXCH, a.model), a.k0, a.k1);
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
    // Warm-up only: scale each temperature's proposal size towards the target
    // acceptance.  Frozen during sampling, so the sampled chain is stationary.
    if (a.adapt_every > 0 && (fl & F_ADAPT)) {
        for (int gr = tid; gr < a.total_reps; gr += kCloseT) {
            int g = 0;
            while (g + 1 < a.ngroups && gr >= a.groups[g + 1].rep0) ++g;
            const DevGroup G = a.groups[g];
            const double sites = (double)G.n_half * 2.0 * (double)a.adapt_every;
            int64_t* cnt = a.acc_cnt + (int64_t)gr * kCntSlots;
            int64_t n = 0;
            for (int k = 0; k < kCntSlots; ++k) n += __hip_atomic_load(&cnt[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const double rate = (double)n / sites;
            const int ti = G.rep0 + a.r2t[gr];
            double f = rate / (double)a.adapt_target;
            f = f < 0.5 ? 0.5 : (f > 2.0 ? 2.0 : f);
            double d = (double)a.deltas[ti] * f;
            d = d < 1e-3 ? 1e-3 : (d > 3.14159265358979 ? 3.14159265358979 : d);
            a.deltas[ti] = (float)d;
            a.omcd[ti] = (float)(1.0 - cos((double)(float)d));
            for (int k = 0; k < kCntSlots; ++k) cnt[k] = 0;
        }
    }
    if (tid == 0) {
        if (meas) {
            if (room) ctl->n_rec = nrec + 1;
            else ctl->overflow = 1;
        }
        ctl->t = t + 1;
        advance_flags(ctl);
    }
}

// Metropolis pass over colour `color`: work block b (in the XCD-aware order,
// b < the pass's work blocks), one site per thread.
template <int N, bool O32>
__device__ __forceinline__ void cont_metro_block(const ContArgs& a, const int b, const int color, const int closes) {
    __shared__ double s_red[4][kCT / 64];
    const int g = find_group_c(a.groups, a.ngroups, b);
    const DevGroup G = ld_group(&a.groups[g]);
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int ti = G.rep0 + ld_uniform(&a.r2t[grep]);
    const float beta = a.inline_params ? a.beta_inline : ld_uniform(&a.betas_f[ti]);
    const float dlt = a.inline_params ? a.delta_inline : ld_uniform(&a.deltas[ti]);
    const float omc = N == 3 ? (a.inline_params ? a.omcd_inline : ld_uniform(&a.omcd[ti])) : 0.f;
    const int64_t t = ld_uniform(&a.ctl->t);
    const int32_t fl = ld_uniform(&a.ctl->flags);
    const bool need = (fl & (F_MEAS | F_EXCH)) != 0 && closes;
    const bool renorm = (fl & F_RENORM) != 0;
    const bool count = a.adapt_every > 0 && (fl & F_WARM) != 0;
    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);

    float* C = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
    const float* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    int acc_n = 0;
    if (h < G.n_half) {
        uint32_t nb[6];
        neighbours(G, h, color, nb);
        float nv[3][6];
        load_nbrs<N, O32>(O, G.stride, nb, nv);
        float hf[3];
        field_f32<N>(nv, hf);
        float S[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < N; ++k) S[k] = *comp_at<O32>(C, G.stride, k, (uint32_t)h);
        const u32x4 w = philox((uint32_t)h, hs, stream, tag(P_CONT, a.model), a.k0, a.k1);
        float nw[3] = {0.f, 0.f, 0.f};
        if (N == 2) {
            const float ar = dlt * kInv2Pi * (2.0f * u24f(w.x) - 1.0f);  // revolutions
            const float sa = sin_rev(ar), ca = cos_rev(ar);
            nw[0] = S[0] * ca - S[1] * sa;
            nw[1] = S[1] * ca + S[0] * sa;
        } else {
            // 1 - cos(theta) = (1 - cos d)(1 - u), exact in 1 - u: sin(theta) keeps
            // its relative precision near the pole, where 1 - ct * ct would cancel
            const float am = omc * (1.0f - u24f(w.x));
            const float ct = 1.0f - am;
            const float st = sqrt_fast(am * (2.0f - am));
            const float sp = sin_rev(u24f(w.y)), cp = cos_rev(u24f(w.y));  // phi = 2 pi u
            float tv0 = 0.f, tv1 = 0.f;
            if (fabsf(S[0]) < 0.9f) tv0 = 1.f; else tv1 = 1.f;
            const float ts = tv0 * S[0] + tv1 * S[1];
            float u0 = tv0 - ts * S[0], u1 = tv1 - ts * S[1], u2 = -ts * S[2];
            const float ui = rsqrt_fast(u0 * u0 + u1 * u1 + u2 * u2);
            u0 *= ui; u1 *= ui; u2 *= ui;
            const float v0 = S[1] * u2 - S[2] * u1, v1 = S[2] * u0 - S[0] * u2, v2 = S[0] * u1 - S[1] * u0;
            const float a0 = st * cp, a1 = st * sp;
            nw[0] = a0 * u0 + a1 * v0 + ct * S[0];
            nw[1] = a0 * u1 + a1 * v1 + ct * S[1];
            nw[2] = a0 * u2 + a1 * v2 + ct * S[2];
            const float ni = rsqrt_fast(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
            nw[0] *= ni; nw[1] *= ni; nw[2] *= ni;
        }
        float dE = 0.f;
#pragma unroll
        for (int k = 0; k < N; ++k) dE += (nw[k] - S[k]) * hf[k];
        dE *= -a.Jf;
        const bool accept = dE <= 0.0f || u24f(w.z) < exp_fast(-beta * dE);
        if (accept) {
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] = nw[k];
            acc_n = 1;
        }
        if (renorm) {
            float q = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) q += S[k] * S[k];
            const float qi = 1.0f / sqrtf(q);
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] *= qi;
        }
        if (accept || renorm) {
#pragma unroll
            for (int k = 0; k < N; ++k) *comp_at<O32>(C, G.stride, k, (uint32_t)h) = S[k];
        }
        if (need) {
            // FP64 from the stored FP32 components (the field is re-summed in FP64)
            double e = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                part[1 + k] = (double)S[k];
                if (color) {
                    double f = 0.0;
#pragma unroll
                    for (int q = 0; q < 6; ++q) f += (double)nv[k][q];
                    e += (double)S[k] * f;
                }
            }
            part[0] = color ? -a.J * e : 0.0;
        }
    }
    if (count) {
        const int wc = wave_sum(acc_n);
        if ((threadIdx.x & 63) == 0 && wc)
            atomicAdd((unsigned long long*)&a.acc_cnt[(int64_t)grep * kCntSlots + ((b * (kCT / 64) + (threadIdx.x >> 6)) & (kCntSlots - 1))],
                      (unsigned long long)wc);
    }
    if (need) {
        block_sum4<kCT, N>(part, s_red, color);
        if (threadIdx.x == 0) {
            double* dst = (color ? a.part1 : a.part0) + 4 * (int64_t)b;
            for (int k = 0; k < 4; ++k) dst[k] = part[k];
        }
    }
}

// The same pass for XY with two sites per thread (h and h + kCT / 2 of the
// block's 256-site tile), so each thread has both sites' loads in flight.  A
// kernel of its own: sharing the loop in the template above changed the
// Heisenberg instance's code and slowed it.
//
// Tile b for the kCT / 2 threads of sub-block q of the calling block (tid:
// thread index in the sub-block).  A block of kCT threads runs two tiles, one
// per sub-block; `valid` false for a sub-block without a tile, which still
// takes part in the block's barrier.
template <int N, bool O32>
__device__ __forceinline__ void cont_metro_pair_block(const ContArgs& a, const int b, const int color, const int closes,
                                                      const bool valid, const int q, const int tid) {
    __shared__ double s_red[4][kCT / 64];
    const int bb = valid ? b : 0;  // an absent tile reads tile 0's parameters and updates nothing
    const int g = find_group_c(a.groups, a.ngroups, bb);
    const DevGroup G = ld_group(&a.groups[g]);
    int r, tile;
    split_tiles(G, bb - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int ti = G.rep0 + ld_uniform(&a.r2t[grep]);
    const float beta = a.inline_params ? a.beta_inline : ld_uniform(&a.betas_f[ti]);
    const float dlt = a.inline_params ? a.delta_inline : ld_uniform(&a.deltas[ti]);
    const float omc = N == 3 ? (a.inline_params ? a.omcd_inline : ld_uniform(&a.omcd[ti])) : 0.f;
    const int64_t t = ld_uniform(&a.ctl->t);
    const int32_t fl = ld_uniform(&a.ctl->flags);
    const bool need = (fl & (F_MEAS | F_EXCH)) != 0 && closes;  // the same for the whole block
    const bool renorm = (fl & F_RENORM) != 0;
    const bool count = a.adapt_every > 0 && (fl & F_WARM) != 0;
    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);

    float* C = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
    const float* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    int acc_n = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int64_t h = (int64_t)tile * kCT + j * (kCT / 2) + tid;
        if (valid && h < G.n_half) {
            uint32_t nb[6];
            neighbours(G, h, color, nb);
            float nv[3][6];
            load_nbrs<N, O32>(O, G.stride, nb, nv);
            float hf[3];
            field_f32<N>(nv, hf);
            float S[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] = *comp_at<O32>(C, G.stride, k, (uint32_t)h);
            const u32x4 w = philox((uint32_t)h, hs, stream, tag(P_CONT, a.model), a.k0, a.k1);
            float nw[3] = {0.f, 0.f, 0.f};
            if (N == 2) {
                const float ar = dlt * kInv2Pi * (2.0f * u24f(w.x) - 1.0f);  // revolutions
                const float sa = sin_rev(ar), ca = cos_rev(ar);
                nw[0] = S[0] * ca - S[1] * sa;
                nw[1] = S[1] * ca + S[0] * sa;
            } else {
                // 1 - cos(theta) = (1 - cos d)(1 - u), exact in 1 - u: sin(theta) keeps
                // its relative precision near the pole, where 1 - ct * ct would cancel
                const float am = omc * (1.0f - u24f(w.x));
                const float ct = 1.0f - am;
                const float st = sqrt_fast(am * (2.0f - am));
                const float sp = sin_rev(u24f(w.y)), cp = cos_rev(u24f(w.y));  // phi = 2 pi u
                float tv0 = 0.f, tv1 = 0.f;
                if (fabsf(S[0]) < 0.9f) tv0 = 1.f; else tv1 = 1.f;
                const float ts = tv0 * S[0] + tv1 * S[1];
                float u0 = tv0 - ts * S[0], u1 = tv1 - ts * S[1], u2 = -ts * S[2];
                const float ui = rsqrt_fast(u0 * u0 + u1 * u1 + u2 * u2);
                u0 *= ui; u1 *= ui; u2 *= ui;
                const float v0 = S[1] * u2 - S[2] * u1, v1 = S[2] * u0 - S[0] * u2, v2 = S[0] * u1 - S[1] * u0;
                const float a0 = st * cp, a1 = st * sp;
                nw[0] = a0 * u0 + a1 * v0 + ct * S[0];
                nw[1] = a0 * u1 + a1 * v1 + ct * S[1];
                nw[2] = a0 * u2 + a1 * v2 + ct * S[2];
                const float ni = rsqrt_fast(nw[0] * nw[0] + nw[1] * nw[1] + nw[2] * nw[2]);
                nw[0] *= ni; nw[1] *= ni; nw[2] *= ni;
            }
            float dE = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) dE += (nw[k] - S[k]) * hf[k];
            dE *= -a.Jf;
            const bool accept = dE <= 0.0f || u24f(w.z) < exp_fast(-beta * dE);
            if (accept) {
#pragma unroll
                for (int k = 0; k < N; ++k) S[k] = nw[k];
                acc_n += 1;
            }
            if (renorm) {
                float q = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) q += S[k] * S[k];
                const float qi = 1.0f / sqrtf(q);
#pragma unroll
                for (int k = 0; k < N; ++k) S[k] *= qi;
            }
            if (accept || renorm) {
#pragma unroll
                for (int k = 0; k < N; ++k) *comp_at<O32>(C, G.stride, k, (uint32_t)h) = S[k];
            }
            if (need) {
                // FP64 from the stored FP32 components (the field is re-summed in FP64)
                double e = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    part[1 + k] += (double)S[k];
                    if (color) {
                        double f = 0.0;
#pragma unroll
                        for (int q = 0; q < 6; ++q) f += (double)nv[k][q];
                        e += (double)S[k] * f;
                    }
                }
                part[0] += color ? -a.J * e : 0.0;
            }
        }
    }
    if (count && valid) {
        const int wc = wave_sum(acc_n);
        if ((tid & 63) == 0 && wc)
            atomicAdd((unsigned long long*)&a.acc_cnt[(int64_t)grep * kCntSlots + ((b * (kCT / 128) + (tid >> 6)) & (kCntSlots - 1))],
                      (unsigned long long)wc);
    }
    if (need) {
        sub_block_sum4<kCT / 2, N>(part, s_red, color, q, tid);
        if (tid == 0 && valid) {
            double* dst = (color ? a.part1 : a.part0) + 4 * (int64_t)b;
            for (int k = 0; k < 4; ++k) dst[k] = part[k];
        }
    }
}

// Closes a continuous sweep (after the colour-1 Metropolis pass).  A sweep
// that measures or exchanges first sums the block partials per replica over
// the whole grid.  Every block then counts itself done (ctl->cdone), on every
// sweep, and the block that counts last closes the sweep: by then every block
// of the launch has read t and the flags (each block uses them before it
// counts), so the close may rewrite them.  A close by a fixed block instead
// would race with blocks dispatched after it, which would read the next
// sweep's flags.
// Work block lb of the nb blocks of the closing launch.
__device__ __forceinline__ void cont_close_block(const ContArgs& a, const int lb, const int nb) {
    static_assert(kCloseT == kCT, "the closing step runs on one block of this launch");
    __shared__ int s_last;
    const int64_t t = a.ctl->t;
    const int32_t fl = a.ctl->flags;
    const bool sums = (fl & (F_MEAS | F_EXCH)) != 0;
    if (sums) reduce_replica(a, lb);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (sums) __threadfence();  // this block's sums are visible before it counts as done
        s_last = atomicAdd(&a.ctl->cdone, 1) == nb - 1;
    }
    __syncthreads();
    if (!s_last) return;
    if (sums) __threadfence();  // and the last block reads every block's sums
    close_cont(a, t, fl);
    if (threadIdx.x == 0) a.ctl->cdone = 0;
}


// Overrelaxation pass over colour `color` (no random numbers, no renormalisation).
// SPT sites per thread (h and h + kCT / SPT, ... of the block's 256-site tile), so
// a thread has SPT sites' loads in flight; the grid is the same tiles for every S
// and no result depends on it (each site's update reads only the other colour).
// XY defaults to two (60.1 against 57.8 Metropolis updates/ns in an A/B run on
// one device), Heisenberg, with 18 loads per site, to one; the autotuner times
// both.
template <int N, int SPT, bool O32>
__device__ __forceinline__ void cont_or_block(const ContArgs& a, const int b, const int color,
                                              const int tid = (int)threadIdx.x) {
    const int g = find_group_c(a.groups, a.ngroups, b);
    const DevGroup G = ld_group(&a.groups[g]);
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    float* C = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
    const float* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int64_t h = (int64_t)tile * kCT + j * (kCT / SPT) + tid;
        if (h >= G.n_half) continue;
        uint32_t nb[6];
        neighbours(G, h, color, nb);
        float hf[3];
        load_field<N, O32>(O, G.stride, nb, hf);
        float h2 = 0.f, sh = 0.f, S[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < N; ++k) {
            hf[k] *= a.Jf;
            S[k] = *comp_at<O32>(C, G.stride, k, (uint32_t)h);
            h2 += hf[k] * hf[k];
            sh += S[k] * hf[k];
        }
        if (h2 < 1e-30f) continue;
        const float sc = 2.0f * sh * __builtin_amdgcn_rcpf(h2);
#pragma unroll
        for (int k = 0; k < N; ++k) *comp_at<O32>(C, G.stride, k, (uint32_t)h) = sc * hf[k] - S[k];
    }
}

}  // namespace ising
