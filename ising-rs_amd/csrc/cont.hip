// XY (n=2) and Heisenberg (n=3) checkerboard Metropolis + overrelaxation, gfx950.
//
// Replaces continuous_metropolis_kernel / continuous_overrelax_kernel /
// init_continuous_rng_kernel (src/cuda/continuous_spin_kernel.cu:9-152) and
// reduce_mag_continuous / reduce_energy_continuous (src/cuda/reduce_kernel.cu:70-151).
//
// Storage: FP32 unit vectors, structure of arrays per colour half-lattice:
// component k of colour c of replica r lives at off + ((2r + c) * n + k) * stride.
// A pass over colour c reads n floats of c and of 1-c and writes n floats of c:
// 3 * 4n bytes per update (XY 24 B, Heisenberg 36 B).
//
// Proposal (symmetric, unlike the reference GPU's cube perturbation):
//   Heisenberg: cap around the current spin, cos(theta) uniform in [cos d, 1],
//               phi uniform in [0, 2 pi)   (heisenberg/metropolis.rs:54-92)
//   XY:         rotation by alpha uniform in [-d, d)
// One Philox call per site and pass: words 0,1 -> proposal, word 2 -> acceptance.
// Overrelaxation: S' = 2 (S.h) h / |h|^2 - S, skipped when |h|^2 < 1e-30.
//
// Measurement is fused into the Metropolis passes and summed in FP64: block
// partials are written to a buffer and the closing launch after the colour-1
// pass adds them per replica in a fixed order (one wave per replica, spread over
// the launch's blocks), so every run with the same seed gives the same bits; the
// block of that launch that finishes last writes the records, decides the
// exchanges and advances the sweep counter.
//
// Threads: one site per thread, except XY's Metropolis (cont_metro_pair_kernel)
// and overrelaxation passes, which take two sites of the block's 256-site tile
// per thread.
#include "kernels.h"

namespace ising {

constexpr int kCT = 256;        // sites per block tile
constexpr int kCloseT = 256;  // threads of the closing block (one replica or pair each at a time)

__device__ __forceinline__ int find_group_c(const DevGroup* groups, int ngroups, int b) {
    int g = 0;
    while (g + 1 < ngroups && b >= groups[g + 1].blk0) ++g;
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

template <int N>
__device__ __forceinline__ void load_nbrs(const float* O, int64_t stride, const uint32_t nb[6], float nv[3][6]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const float* Ok = O + k * stride;
#pragma unroll
        for (int q = 0; q < 6; ++q) nv[k][q] = Ok[nb[q]];
    }
}

template <int N>
__device__ __forceinline__ void field_f32(const float nv[3][6], float hf[3]) {
    hf[0] = hf[1] = hf[2] = 0.f;
#pragma unroll
    for (int k = 0; k < N; ++k) hf[k] = ((nv[k][0] + nv[k][1]) + (nv[k][2] + nv[k][3])) + (nv[k][4] + nv[k][5]);
}

template <int N>
__device__ __forceinline__ void load_field(const float* O, int64_t stride, const uint32_t nb[6], float hf[3]) {
    float nv[3][6];
    load_nbrs<N>(O, stride, nb, nv);
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

// Per-replica FP64 sums (E, Mx, My, Mz) of the block partials of a measured
// or exchange sweep, one wave per replica in a fixed order (lane l takes
// tiles l, l + 64, ...; then a fixed shuffle tree), into dacc; the blocks of
// the closing launch share the replicas, so the sums proceed on many CUs.
__device__ __forceinline__ void reduce_replica(const ContArgs& a) {
    const int lane = (int)threadIdx.x & 63;
    const int gr = (int)blockIdx.x * (kCT / 64) + ((int)threadIdx.x >> 6);
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
__device__ void close_cont(const ContArgs& a, int64_t t, int32_t fl) {
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

// Metropolis pass over colour `color` of every replica.
template <int N>
__global__ __launch_bounds__(kCT) void cont_metro_kernel(ContArgs a, int nblocks, int color, int closes) {
    __shared__ double s_red[4][kCT / 64];
    const int b = xcd_block();
    if (b >= nblocks) return;
    const int g = find_group_c(a.groups, a.ngroups, b);
    const DevGroup G = a.groups[g];
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int ti = G.rep0 + a.r2t[grep];
    const float beta = a.inline_params ? a.beta_inline : a.betas_f[ti];
    const float dlt = a.inline_params ? a.delta_inline : a.deltas[ti];
    const float omc = N == 3 ? (a.inline_params ? a.omcd_inline : a.omcd[ti]) : 0.f;
    const int64_t t = a.ctl->t;
    const int32_t fl = a.ctl->flags;
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
        load_nbrs<N>(O, G.stride, nb, nv);
        float hf[3];
        field_f32<N>(nv, hf);
        float S[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < N; ++k) S[k] = C[k * G.stride + h];
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
            for (int k = 0; k < N; ++k) C[k * G.stride + h] = S[k];
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
template <int N>
__global__ __launch_bounds__(kCT / 2) void cont_metro_pair_kernel(ContArgs a, int nblocks, int color, int closes) {
    __shared__ double s_red[4][kCT / 64];
    const int b = xcd_block();
    if (b >= nblocks) return;
    const int g = find_group_c(a.groups, a.ngroups, b);
    const DevGroup G = a.groups[g];
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int ti = G.rep0 + a.r2t[grep];
    const float beta = a.inline_params ? a.beta_inline : a.betas_f[ti];
    const float dlt = a.inline_params ? a.delta_inline : a.deltas[ti];
    const float omc = N == 3 ? (a.inline_params ? a.omcd_inline : a.omcd[ti]) : 0.f;
    const int64_t t = a.ctl->t;
    const int32_t fl = a.ctl->flags;
    const bool need = (fl & (F_MEAS | F_EXCH)) != 0 && closes;
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
        const int64_t h = (int64_t)tile * kCT + j * (kCT / 2) + threadIdx.x;
        if (h < G.n_half) {
            uint32_t nb[6];
            neighbours(G, h, color, nb);
            float nv[3][6];
            load_nbrs<N>(O, G.stride, nb, nv);
            float hf[3];
            field_f32<N>(nv, hf);
            float S[3] = {0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] = C[k * G.stride + h];
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
                for (int k = 0; k < N; ++k) C[k * G.stride + h] = S[k];
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
    if (count) {
        const int wc = wave_sum(acc_n);
        if ((threadIdx.x & 63) == 0 && wc)
            atomicAdd((unsigned long long*)&a.acc_cnt[(int64_t)grep * kCntSlots + ((b * (kCT / 128) + (threadIdx.x >> 6)) & (kCntSlots - 1))],
                      (unsigned long long)wc);
    }
    if (need) {
        block_sum4<kCT / 2, N>(part, s_red, color);
        if (threadIdx.x == 0) {
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
__global__ __launch_bounds__(kCT) void cont_close_kernel(ContArgs a) {
    static_assert(kCloseT == kCT, "the closing step runs on one block of this launch");
    __shared__ int s_last;
    const int64_t t = a.ctl->t;
    const int32_t fl = a.ctl->flags;
    const bool sums = (fl & (F_MEAS | F_EXCH)) != 0;
    if (sums) reduce_replica(a);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (sums) __threadfence();  // this block's sums are visible before it counts as done
        s_last = atomicAdd(&a.ctl->cdone, 1) == (int)gridDim.x - 1;
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
template <int N, int SPT>
__global__ __launch_bounds__(kCT / SPT) void cont_or_kernel(ContArgs a, int nblocks, int color) {
    const int b = xcd_block();
    if (b >= nblocks) return;
    const int g = find_group_c(a.groups, a.ngroups, b);
    const DevGroup G = a.groups[g];
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    float* C = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
    const float* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int64_t h = (int64_t)tile * kCT + j * (kCT / SPT) + threadIdx.x;
        if (h >= G.n_half) continue;
        uint32_t nb[6];
        neighbours(G, h, color, nb);
        float hf[3];
        load_field<N>(O, G.stride, nb, hf);
        float h2 = 0.f, sh = 0.f, S[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < N; ++k) {
            hf[k] *= a.Jf;
            S[k] = C[k * G.stride + h];
            h2 += hf[k] * hf[k];
            sh += S[k] * hf[k];
        }
        if (h2 < 1e-30f) continue;
        const float sc = 2.0f * sh * __builtin_amdgcn_rcpf(h2);
#pragma unroll
        for (int k = 0; k < N; ++k) C[k * G.stride + h] = sc * hf[k] - S[k];
    }
}

// Random start: Heisenberg z = 2u-1, phi = 2 pi u' (uniform on S^2); XY theta = 2 pi u.
template <int N>
__global__ __launch_bounds__(kCT) void cont_init_kernel(const DevGroup* groups, int ngroups, float* pool, int model,
                                                      uint32_t k0, uint32_t k1) {
    const int g = find_group_c(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, (int)blockIdx.x - G.blk0, &r, &tile);
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    if (h >= G.n_half) return;
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);
    for (int c = 0; c < 2; ++c) {
        float* C = pool + G.off + (int64_t)(2 * r + c) * N * G.stride;
        const int64_t row = h / G.Lh;
        const int xh = (int)(h - row * G.Lh);
        const int z = (int)(row / G.L), y = (int)(row - (int64_t)z * G.L);
        const int x = 2 * xh + ((y + z + c) & 1);
        const uint32_t site = (uint32_t)(((int64_t)z * G.L + y) * G.L + x);
        const u32x4 w = philox(site, 0u, stream, tag(P_INIT, (uint32_t)model), k0, k1);
        if (N == 2) {
            float sn, cs;
            sincosf(6.28318530717958647692f * u24f(w.x), &sn, &cs);
            C[h] = cs;
            C[G.stride + h] = sn;
        } else {
            const float zz = 2.0f * u24f(w.x) - 1.0f;
            float sn, cs;
            sincosf(6.28318530717958647692f * u24f(w.y), &sn, &cs);
            const float rr = sqrtf(fmaxf(0.0f, 1.0f - zz * zz));
            C[h] = rr * cs;
            C[G.stride + h] = rr * sn;
            C[2 * G.stride + h] = zz;
        }
    }
}

// Standalone FP64 E/M of every replica: part[4*blk + k] partials (E, Mx, My, Mz).
template <int N>
__global__ __launch_bounds__(kCT) void cont_measure_kernel(const DevGroup* groups, int ngroups, const float* pool,
                                                         double J, double* part) {
    __shared__ double s_red[4][kCT / 64];
    const int g = find_group_c(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, (int)blockIdx.x - G.blk0, &r, &tile);
    const float* C1 = pool + G.off + (int64_t)(2 * r + 1) * N * G.stride;
    const float* C0 = pool + G.off + (int64_t)(2 * r) * N * G.stride;
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (h < G.n_half) {
        uint32_t nb[6];
        neighbours(G, h, 1, nb);
        double e = 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double f = 0.0;
            for (int q = 0; q < 6; ++q) f += (double)C0[k * G.stride + nb[q]];
            const double s1 = (double)C1[k * G.stride + h];
            e += s1 * f;
            v[1 + k] = s1 + (double)C0[k * G.stride + h];
        }
        v[0] = -J * e;
    }
    block_sum4(v, s_red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) part[4 * (int64_t)blockIdx.x + k] = v[k];
}

// XY spin stiffness sums of every replica: over the forward bonds (i, i + d),
// d = x, y, z, C_d = sum cos(theta_i - theta_j) = S_i . S_j and
// J_d = sum sin(theta_i - theta_j) = S_i^y S_j^x - S_i^x S_j^y, as
// part[6 * blk + {C_x, C_y, C_z, J_x, J_y, J_z}] block partials (FP64).
// The helicity modulus along d is (<C_d> - beta <J_d^2>) / N.
__global__ __launch_bounds__(kCT) void cont_helicity_kernel(const DevGroup* groups, int ngroups, const float* pool,
                                                          double* part) {
    __shared__ double s_red[4][kCT / 64];
    const int g = find_group_c(groups, ngroups, blockIdx.x);
    const DevGroup G = groups[g];
    int r, tile;
    split_tiles(G, (int)blockIdx.x - G.blk0, &r, &tile);
    const int64_t h = (int64_t)tile * kCT + threadIdx.x;
    double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (h < G.n_half) {
        for (int c = 0; c < 2; ++c) {
            const float* Ci = pool + G.off + (int64_t)(2 * r + c) * 2 * G.stride;
            const float* Cj = pool + G.off + (int64_t)(2 * r + (c ^ 1)) * 2 * G.stride;
            uint32_t nb[6];
            neighbours(G, h, c, nb);
            const int64_t row = h / G.Lh;
            const int z = (int)(row / G.L), y = (int)(row - (int64_t)z * G.L);
            const int p = (y + z + c) & 1;
            const int64_t fwd[3] = {p ? nb[1] : nb[0], nb[3], nb[5]};  // +x, +y, +z
            const double xi = (double)Ci[h], yi = (double)Ci[G.stride + h];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double xj = (double)Cj[fwd[d]], yj = (double)Cj[G.stride + fwd[d]];
                v[d] += xi * xj + yi * yj;
                v[3 + d] += yi * xj - xi * yj;
            }
        }
    }
    block_sum4(v, s_red);
    __syncthreads();
    block_sum4(v + 4, s_red);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * (int64_t)blockIdx.x + k] = v[k];
}

hipError_t launch_cont_helicity(const DevGroup* groups, int ngroups, const float* pool, double* part, int nblocks,
                                hipStream_t st) {
    hipLaunchKernelGGL(cont_helicity_kernel, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, part);
    return hipGetLastError();
}

static inline int xcd_grid(int nblocks) { return (nblocks + 7) & ~7; }

hipError_t launch_cont_metro(const ContArgs& a, int nblocks, int color, int closes, hipStream_t st) {
    const dim3 grid(xcd_grid(nblocks));
    if (a.ncomp == 2)
        hipLaunchKernelGGL(cont_metro_pair_kernel<2>, grid, dim3(kCT / 2), 0, st, a, nblocks, color, closes);
    else hipLaunchKernelGGL(cont_metro_kernel<3>, grid, dim3(kCT), 0, st, a, nblocks, color, closes);
    return hipGetLastError();
}

hipError_t launch_cont_close(const ContArgs& a, hipStream_t st) {
    const int nb = (a.total_reps + kCT / 64 - 1) / (kCT / 64);
    hipLaunchKernelGGL(cont_close_kernel, dim3(nb), dim3(kCT), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_cont_or(const ContArgs& a, int nblocks, int color, hipStream_t st) {
    const dim3 grid(xcd_grid(nblocks));
    const int spt = a.or_spt ? a.or_spt : (a.ncomp == 2 ? 2 : 1);
    if (a.ncomp == 2 && spt == 2) hipLaunchKernelGGL((cont_or_kernel<2, 2>), grid, dim3(kCT / 2), 0, st, a, nblocks, color);
    else if (a.ncomp == 2) hipLaunchKernelGGL((cont_or_kernel<2, 1>), grid, dim3(kCT), 0, st, a, nblocks, color);
    else if (spt == 2) hipLaunchKernelGGL((cont_or_kernel<3, 2>), grid, dim3(kCT / 2), 0, st, a, nblocks, color);
    else hipLaunchKernelGGL((cont_or_kernel<3, 1>), grid, dim3(kCT), 0, st, a, nblocks, color);
    return hipGetLastError();
}

hipError_t launch_cont_init(const DevGroup* groups, int ngroups, float* pool, int ncomp, int model, int nblocks,
                            uint32_t k0, uint32_t k1, hipStream_t st) {
    if (ncomp == 2)
        hipLaunchKernelGGL(cont_init_kernel<2>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, model, k0, k1);
    else
        hipLaunchKernelGGL(cont_init_kernel<3>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, model, k0, k1);
    return hipGetLastError();
}

hipError_t launch_cont_measure(const DevGroup* groups, int ngroups, const float* pool, int ncomp, double J,
                               double* part, int nblocks, hipStream_t st) {
    if (ncomp == 2)
        hipLaunchKernelGGL(cont_measure_kernel<2>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, J, part);
    else
        hipLaunchKernelGGL(cont_measure_kernel<3>, dim3(nblocks), dim3(kCT), 0, st, groups, ngroups, pool, J, part);
    return hipGetLastError();
}

}  // namespace ising
