// Learned proposal for XY / Heisenberg Metropolis-Hastings, gfx950.
//
// A small MLP maps rotation-invariant features of a site's six neighbours to a
// two-component mixture around the local-field direction h^ = h/|h|:
//   x = [beta, beta|h|, |h|/6, <q^2>, <q^3>, <q^4>],  q_j = S_j . h^
//   6 -> H (ReLU) -> H (ReLU) -> H (ReLU) -> 3:   kappa_k = softplus(o_k) + 1e-3,
//   w_1 = sigmoid(o_2), w_2 = 1 - w_1
// Components, in c = s . h^:  XY von Mises  f = e^{kappa c} / (2 pi I0(kappa)),
//                             Heisenberg Fisher f = kappa e^{kappa c} / (4 pi sinh kappa).
// Weights are stored in BF16 and every product is accumulated in FP32.
//
// The proposal depends only on the neighbours, which a checkerboard pass does
// not change, so the move s -> s' is accepted with the exact Metropolis-Hastings
// ratio  r = exp(beta J |h| (c' - c)) q(c) / q(c').  Proposals are drawn exactly:
// Heisenberg by inverting the Fisher CDF; XY by the Best-Fisher rejection
// sampler (Appl. Stat. 28:152, 1979) with at most kVmAttempts attempts, after
// which the site keeps its spin -- that failure probability depends only on
// the neighbours, so it cancels between forward and reverse moves.
//
// Training: maximum likelihood of the single-site conditional on sites of the
// current PT configurations (loss = -log q(s | neighbours)); the minimiser is
// the exact conditional (heat bath).  Per-block gradient partials are reduced
// in a fixed order by the Adam kernel, so training is deterministic.
#include <cstddef>
#include <type_traits>

#include "kernels.h"

namespace ising {

constexpr int kLearnThreads = 64;  // training: one wave per block, one sample per lane
// v_cos_f32 / v_sin_f32 take revolutions (|error| <= 1.3e-7 on [0, 1) against FP64 on gfx950).
__device__ __forceinline__ float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }
__device__ __forceinline__ float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
// Natural log, square root and reciprocal square root straight from v_log_f32
// (log2), v_sqrt_f32 and v_rsq_f32, for normal inputs: no denormal rescaling
// and no correction steps.  Relative error ~1e-7; log_hw(0) = -inf.
__device__ __forceinline__ float log_hw(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }
constexpr int kVmAttempts = 7;  // von Mises attempts per draw: 3 together, then 2 pairs
constexpr int kNIn = 6;

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}

// exp(-x) I0(x) and exp(-x) I1(x), x >= 0: Abramowitz & Stegun 9.8.1-9.8.4.
__device__ __forceinline__ float i0e_f(float x) {
    if (x <= 3.75f) {
        const float t = (x * (1.0f / 3.75f)) * (x * (1.0f / 3.75f));
        const float p = 1.0f + t * (3.5156229f + t * (3.0899424f + t * (1.2067492f + t * (0.2659732f + t * (0.0360768f + t * 0.0045813f)))));
        return p * __expf(-x);
    }
    const float y = 3.75f * __builtin_amdgcn_rcpf(x);
    const float p = 0.39894228f + y * (0.01328592f + y * (0.00225319f + y * (-0.00157565f + y * (0.00916281f + y * (-0.02057706f + y * (0.02635537f + y * (-0.01647633f + y * 0.00392377f)))))));
    return p * rsq_hw(x);
}
__device__ __forceinline__ float i1e_f(float x) {
    if (x <= 3.75f) {
        const float t = (x * (1.0f / 3.75f)) * (x * (1.0f / 3.75f));
        const float p = 0.5f + t * (0.87890594f + t * (0.51498869f + t * (0.15084934f + t * (0.02658733f + t * (0.00301532f + t * 0.00032411f)))));
        return x * p * __expf(-x);
    }
    const float y = 3.75f * __builtin_amdgcn_rcpf(x);
    const float p = 0.39894228f + y * (-0.03988024f + y * (-0.00362018f + y * (0.00163801f + y * (-0.01031555f + y * (0.02282967f + y * (-0.02895312f + y * (0.01787654f + y * -0.00420059f)))))));
    return p * rsq_hw(x);
}

// Hidden activation: ReLU (tanh would take an exponential, a reciprocal and
// three more operations, for 48 activations per site at H = 16).  One
// v_max_i32 on the bits: a float with the sign bit set is a negative int32, so
// max(bits, 0) keeps z > 0 and gives +0 otherwise (fmaxf would add a NaN
// canonicalisation).  Its derivative from the activation itself: 1 where a > 0,
// else 0.
__device__ __forceinline__ float act_f(float z) { return __int_as_float(max(__float_as_int(z), 0)); }
__device__ __forceinline__ float act_grad(float a) { return a > 0.0f ? 1.0f : 0.0f; }
__device__ __forceinline__ float softplus_f(float x) { return x > 15.0f ? x : log_hw(1.0f + __expf(x)); }
__device__ __forceinline__ float sigmoid_f(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

// log f_k(c) per component (see header).
template <int N>
__device__ __forceinline__ float log_comp(float kap, float c) {
    if (N == 2) return kap * (c - 1.0f) - 1.8378770664093453f - logf(i0e_f(kap));
    return logf(kap) + kap * (c - 1.0f) - 1.8378770664093453f - log1pf(-expf(-2.0f * kap));
}
// E[c] under one component: I1/I0 (XY), coth k - 1/k (Heisenberg); derivative of log f is c - mean_c.
template <int N>
__device__ __forceinline__ float mean_c(float kap) {
    if (N == 2) return i1e_f(kap) / i0e_f(kap);
    if (kap < 1e-2f) return kap * (1.0f / 3.0f - kap * kap / 45.0f);
    const float e = __expf(-2.0f * kap);
    return (1.0f + e) / (1.0f - e) - 1.0f / kap;
}

struct Mix {
    float k0, k1, w0;
};

__device__ __forceinline__ Mix mixture(const float o[3]) {
    Mix m;
    m.k0 = softplus_f(o[0]) + 1e-3f;
    m.k1 = softplus_f(o[1]) + 1e-3f;
    m.w0 = sigmoid_f(o[2]);
    return m;
}

// log q(c) = logsumexp(A_0 + kappa_0 c, A_1 + kappa_1 c): the constants of a
// site's mixture, computed once and shared by the forward and reverse moves.
struct MixLog {
    float k0, k1, a0, a1;
};

// log f_k(0), the sweep's form of log_comp(kap, 0) from the hardware log:
//   XY          -kappa - log(2 pi) - log(e^-kappa I0(kappa))
//   Heisenberg  log(kappa / (1 - e^{-2 kappa})) - kappa - log(2 pi); below
//               kappa = 1/16 the series -kappa^2/6 + kappa^4/180 - log(4 pi)
//               (error < 2e-11), which avoids the cancellation in 1 - e^{-2 kappa}.
template <int N>
__device__ __forceinline__ float log_comp0(float kap) {
    if (N == 2) return -kap - 1.8378770664093453f - log_hw(i0e_f(kap));
    if (kap < 0.0625f) {
        const float k2 = kap * kap;
        return k2 * (k2 * (1.0f / 180.0f) - (1.0f / 6.0f)) - 2.5310242469692907f;
    }
    return log_hw(kap) - log_hw(1.0f - __expf(-2.0f * kap)) - kap - 1.8378770664093453f;
}

template <int N>
__device__ __forceinline__ MixLog mix_log(const Mix& m) {
    MixLog q;
    q.k0 = m.k0;
    q.k1 = m.k1;
    q.a0 = log_hw(m.w0) + log_comp0<N>(m.k0);
    q.a1 = log_hw(1.0f - m.w0) + log_comp0<N>(m.k1);
    return q;
}

__device__ __forceinline__ float log_q_fast(const MixLog& q, float c) {
    const float l0 = q.a0 + q.k0 * c, l1 = q.a1 + q.k1 * c;
    const float mx = fmaxf(l0, l1);
    return mx + log_hw(__expf(l0 - mx) + __expf(l1 - mx));
}

template <int N>
__device__ __forceinline__ float log_q(const Mix& m, float c) {
    const float l0 = logf(m.w0) + log_comp<N>(m.k0, c);
    const float l1 = logf(1.0f - m.w0) + log_comp<N>(m.k1, c);
    const float mx = fmaxf(l0, l1);
    return mx + logf(expf(l0 - mx) + expf(l1 - mx));
}

// Parameter offsets of the flat vector W1,b1,W2,b2,W3,b3,W4,b4 (row-major).
template <int H>
struct Off {
    static constexpr int W1 = 0, b1 = W1 + H * kNIn, W2 = b1 + H, b2 = W2 + H * H, W3 = b2 + H, b3 = W3 + H * H,
                         W4 = b3 + H, b4 = W4 + 3 * H, P = b4 + 3;
};

// Forward pass for one sample; activations written to act[3][H] when non-null.
template <int H>
__device__ __forceinline__ void mlp_forward(const uint16_t* __restrict__ w, const float x[kNIn], float o[3],
                                            float* act, int act_stride) {
    using F = Off<H>;
    float a[H], b[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float s = bf2f(w[F::b1 + j]);
#pragma unroll
        for (int i = 0; i < kNIn; ++i) s += bf2f(w[F::W1 + j * kNIn + i]) * x[i];
        a[j] = act_f(s);
        if (act) act[(0 * H + j) * act_stride] = a[j];
    }
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float s = bf2f(w[F::b2 + j]);
#pragma unroll
        for (int i = 0; i < H; ++i) s += bf2f(w[F::W2 + j * H + i]) * a[i];
        b[j] = act_f(s);
        if (act) act[(1 * H + j) * act_stride] = b[j];
    }
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float s = bf2f(w[F::b3 + j]);
#pragma unroll
        for (int i = 0; i < H; ++i) s += bf2f(w[F::W3 + j * H + i]) * b[i];
        a[j] = act_f(s);
        if (act) act[(2 * H + j) * act_stride] = a[j];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float s = bf2f(w[F::b4 + k]);
#pragma unroll
        for (int i = 0; i < H; ++i) s += bf2f(w[F::W4 + k * H + i]) * a[i];
        o[k] = s;
    }
}

// Features of a site from its neighbour vectors (N components each).
template <int N>
__device__ __forceinline__ float site_features(float beta, const float nv[3][6], float x[kNIn], float hhat[3]) {
    float h[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < N; ++k) h[k] = ((nv[k][0] + nv[k][1]) + (nv[k][2] + nv[k][3])) + (nv[k][4] + nv[k][5]);
    float hn2 = 0.f;
#pragma unroll
    for (int k = 0; k < N; ++k) hn2 += h[k] * h[k];
    const float hn = sqrt_hw(hn2);
    const float inv = hn > 1e-20f ? __builtin_amdgcn_rcpf(hn) : 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) hhat[k] = k < N ? h[k] * inv : 0.0f;
    float m2 = 0.f, m3 = 0.f, m4 = 0.f;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        float d = 0.f;
#pragma unroll
        for (int k = 0; k < N; ++k) d += nv[k][q] * hhat[k];
        const float d2 = d * d;
        m2 += d2;
        m3 += d2 * d;
        m4 += d2 * d2;
    }
    x[0] = beta;
    x[1] = beta * hn;
    x[2] = hn * (1.0f / 6.0f);
    x[3] = m2 * (1.0f / 6.0f);
    x[4] = m3 * (1.0f / 6.0f);
    x[5] = m4 * (1.0f / 6.0f);
    return hn;
}

// ---------------------------------------------------------------------------
// Inference on the matrix cores.  A wave evaluates the MLP for its 64 sites as
// four tiles of 16 sites with v_mfma_f32_16x16x16_bf16, in the transposed form
// D[neuron][site] = W[neuron][k] X[k][site]: each lane's accumulator (rows
// 4 (lane / 16) + r, column lane % 16) is exactly the B operand the next layer
// needs (rows k = 4 (lane / 16) + r of column lane % 16), so activations stay
// in registers between layers.  Layer 1 takes the 6 features plus a constant 1
// (its bias column) padded to K = 16; later biases seed the accumulators.
// Weights are BF16 as stored; inputs and activations are rounded to BF16 for
// the products, accumulation is FP32.
// ---------------------------------------------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// A lane's operand fragments.  Layer 1 (K = 6 features + a constant 1, padded to
// 16) uses v_mfma_f32_16x16x16_bf16: 4 BF16 per lane.  For H >= 32 the layers
// that contract over the hidden width use v_mfma_f32_16x16x32_bf16 (gfx950),
// KC = H / 32 products per 16 outputs, 8 BF16 per lane; for H <= 16 the
// 16x16x16 form.
template <int H>
struct NetFrag {
    static constexpr int T = (H + 15) / 16;  // 16-wide tiles of the hidden width
    static constexpr bool K32 = T >= 2;
    static constexpr int KC = K32 ? T / 2 : 1;  // K chunks of the hidden-width contractions
    using WK = std::conditional_t<K32, v8s, v4s>;
    v4s w1[T];
    WK w2[T][KC], w3[T][KC], w4[KC];
    v4f b2[T], b3[T], b4;
};

// Position p = 8 q + i of a 32-wide K chunk j (q = lane / 16, i = 0..7) stands
// for this hidden neuron: i < 4 -> row 4 q + i of output tile 2 j, i >= 4 ->
// row 4 q + i - 4 of tile 2 j + 1.  Those are exactly the values a lane holds
// in its accumulators of the two tiles, so a layer's activations become the
// next product's operand without moving between lanes; the weights are laid
// out with the same order of K, which leaves the sum unchanged.
__host__ __device__ constexpr int k32_neuron(int j, int q, int i) {
    return 32 * j + (i < 4 ? 4 * q + i : 16 + 4 * q + (i - 4));
}

// This lane's operand fragments of every layer (zero padding past H and 3).
template <int H>
__device__ __forceinline__ void load_net(const uint16_t* __restrict__ w, NetFrag<H>& F) {
    using O = Off<H>;
    constexpr int T = NetFrag<H>::T;
    const int l = threadIdx.x & 63, col = l & 15, kq = 4 * (l >> 4);
#pragma unroll
    for (int mt = 0; mt < T; ++mt) {
        const int j = 16 * mt + col;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = kq + q;
            uint16_t v = 0;
            if (j < H) v = k < kNIn ? w[O::W1 + j * kNIn + k] : (k == kNIn ? w[O::b1 + j] : (uint16_t)0);
            F.w1[mt][q] = (short)v;
        }
        if constexpr (NetFrag<H>::K32) {
#pragma unroll
            for (int c = 0; c < NetFrag<H>::KC; ++c)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int i = k32_neuron(c, l >> 4, q);
                    const bool ok = j < H && i < H;
                    F.w2[mt][c][q] = (short)(ok ? w[O::W2 + j * H + i] : (uint16_t)0);
                    F.w3[mt][c][q] = (short)(ok ? w[O::W3 + j * H + i] : (uint16_t)0);
                }
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = kq + q;
                const bool ok = j < H && i < H;
                F.w2[mt][0][q] = (short)(ok ? w[O::W2 + j * H + i] : (uint16_t)0);
                F.w3[mt][0][q] = (short)(ok ? w[O::W3 + j * H + i] : (uint16_t)0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int n = 16 * mt + kq + r;
            F.b2[mt][r] = n < H ? bf2f(w[O::b2 + n]) : 0.0f;
            F.b3[mt][r] = n < H ? bf2f(w[O::b3 + n]) : 0.0f;
        }
    }
    if constexpr (NetFrag<H>::K32) {
#pragma unroll
        for (int c = 0; c < NetFrag<H>::KC; ++c)
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = k32_neuron(c, l >> 4, q);
                F.w4[c][q] = (short)((col < 3 && i < H) ? w[O::W4 + col * H + i] : (uint16_t)0);
            }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = kq + q;
            F.w4[0][q] = (short)((col < 3 && i < H) ? w[O::W4 + col * H + i] : (uint16_t)0);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) F.b4[r] = kq + r < 3 ? bf2f(w[O::b4 + kq + r]) : 0.0f;
}

// The fragments in device memory, as 8-byte units: unit u of lane l at byte
// 8 (64 u + l) (every fragment is 8 bytes, a bias 16 = two units).  A
// network's fragments are built once per parameter update (learn_frag_kernel);
// a block copies the whole buffer into LDS, where a wave reads any unit of its
// 64 lanes as one conflict-free 512-byte ds_read_b64.
template <int H>
struct NetUnits {
    static_assert(sizeof(NetFrag<H>) % 16 == 0, "fragments are whole 16-byte words");
    static constexpr int U = (int)(sizeof(NetFrag<H>) / 8);
    static constexpr int BYTES = 64 * (int)sizeof(NetFrag<H>);  // the whole buffer (64 lanes)
    uint2 u[U];
};

template <int H>
__global__ __launch_bounds__(64) void learn_frag_kernel(const uint16_t* __restrict__ net, uint2* __restrict__ frag) {
    NetFrag<H> F;
    load_net<H>(net, F);
    const NetUnits<H> W = __builtin_bit_cast(NetUnits<H>, F);
#pragma unroll
    for (int u = 0; u < NetUnits<H>::U; ++u) frag[64 * u + threadIdx.x] = W.u[u];
}

// The block's LDS copy of the fragment buffer (every thread takes part; a
// barrier must follow before any wave reads it).
template <int H>
__device__ __forceinline__ void stage_frag(const uint4* __restrict__ frag, uint4* s) {
    for (int i = threadIdx.x; i < NetUnits<H>::BYTES / 16; i += blockDim.x) s[i] = frag[i];
}

// This lane's 8-byte unit at byte offset b of its NetFrag (b a multiple of 8).
__device__ __forceinline__ v4s frag8(const char* s, int lane, int b) {
    return *(const v4s*)(s + 8 * (64 * (b >> 3) + lane));
}
__device__ __forceinline__ v8s fragw(const char* s, int lane, int b, v8s) {
    return __builtin_shufflevector(frag8(s, lane, b), frag8(s, lane, b + 8), 0, 1, 2, 3, 4, 5, 6, 7);
}
__device__ __forceinline__ v4s fragw(const char* s, int lane, int b, v4s) { return frag8(s, lane, b); }
__device__ __forceinline__ v4f frag16(const char* s, int lane, int b) {
    const v4s lo = frag8(s, lane, b), hi = frag8(s, lane, b + 8);
    const v2u a = __builtin_bit_cast(v2u, lo), c = __builtin_bit_cast(v2u, hi);
    return (v4f){__uint_as_float(a[0]), __uint_as_float(a[1]), __uint_as_float(c[0]), __uint_as_float(c[1])};
}

__device__ __forceinline__ short bf16_bits(float f) { return __builtin_bit_cast(short, (__bf16)f); }  // RNE

// BF16 activations of four accumulators: two v_cvt_pk_bf16_f32 (round to
// nearest even), then ReLU as one v_pk_max_i16 per pair on the BF16 bits (a
// negative BF16 is a negative int16).  Rounding keeps the sign, so this equals
// rounding max(z, 0).
typedef short v2s __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_pk_bf16(float a, float b) {
    const v2s p = __builtin_bit_cast(v2s, __builtin_convertvector((v2f){a, b}, v2bf));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(p, (v2s){0, 0}));
}
__device__ __forceinline__ v4s act_bf16(v4f d) {
    const v2u p = {relu_pk_bf16(d[0], d[1]), relu_pk_bf16(d[2], d[3])};
    return __builtin_bit_cast(v4s, p);
}

// Outputs o[3] for the site of every lane of the block's waves, from the
// fragments staged in LDS (sf).  All threads of the block call it together.
// feat: 8 halves per thread; out: 3 floats per thread.  A wave evaluates its
// 64 sites as four 16-site tiles layer by layer: each weight fragment is read
// from LDS once and applied to all four tiles, and only the activations of two
// layers (4 tiles x H/16 fragments each) stay in registers.
template <int H>
__device__ __forceinline__ void mlp_block(const char* sf, const float x[kNIn], float o[3], uint16_t* feat,
                                          float* out) {
    using NF = NetFrag<H>;
    constexpr int T = NF::T;
    const int t = threadIdx.x, l = t & 63, w0 = t - l;
#pragma unroll
    for (int i = 0; i < kNIn; ++i) feat[8 * t + i] = (uint16_t)bf16_bits(x[i]);
    feat[8 * t + kNIn] = 0x3f80;  // 1.0
    feat[8 * t + kNIn + 1] = 0;
    __syncthreads();
    // input fragments: features 4 (lane / 16) .. + 3 of site 16 st + lane % 16
    v4s xb[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) {
        xb[st] = (v4s){0, 0, 0, 0};
        if ((l >> 4) < 2) {
            const uint16_t* p = feat + 8 * (w0 + 16 * st + (l & 15)) + 4 * (l >> 4);
#pragma unroll
            for (int q = 0; q < 4; ++q) xb[st][q] = (short)p[q];
        }
    }
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    v4s a[4][T], c[4][T];
#pragma unroll
    for (int mt = 0; mt < T; ++mt) {
        const v4s A = frag8(sf, l, (int)offsetof(NF, w1) + 8 * mt);
#pragma unroll
        for (int st = 0; st < 4; ++st) a[st][mt] = act_bf16(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, xb[st], zero, 0, 0, 0));
    }
    // products over the hidden width: K = 16 (T = 1) or KC chunks of 32
    using WK = typename NF::WK;
    constexpr int KC = NF::KC;
    auto contract = [](const WK (&A)[KC], const v4s (&in)[T], v4f d) -> v4f {
        if constexpr (NF::K32) {
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const v8s b = __builtin_shufflevector(in[2 * c], in[2 * c + 1], 0, 1, 2, 3, 4, 5, 6, 7);
                d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, A[c]), __builtin_bit_cast(v8bf, b),
                                                            d, 0, 0, 0);
            }
            return d;
        } else {
            return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A[0], in[0], d, 0, 0, 0);
        }
    };
    // hidden layers 2 and 3 (a -> c -> a)
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const int wo = layer == 0 ? (int)offsetof(NF, w2) : (int)offsetof(NF, w3);
        const int bo = layer == 0 ? (int)offsetof(NF, b2) : (int)offsetof(NF, b3);
#pragma unroll
        for (int mt = 0; mt < T; ++mt) {
            WK A[KC];
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) A[kc] = fragw(sf, l, wo + (int)sizeof(WK) * (mt * KC + kc), WK{});
            const v4f bias = frag16(sf, l, bo + 16 * mt);
#pragma unroll
            for (int st = 0; st < 4; ++st) {
                const v4f d = contract(A, layer == 0 ? a[st] : c[st], bias);
                if (layer == 0) c[st][mt] = act_bf16(d);
                else a[st][mt] = act_bf16(d);
            }
        }
    }
    {
        WK A[KC];
#pragma unroll
        for (int kc = 0; kc < KC; ++kc) A[kc] = fragw(sf, l, (int)offsetof(NF, w4) + (int)sizeof(WK) * kc, WK{});
        const v4f bias = frag16(sf, l, (int)offsetof(NF, b4));
#pragma unroll
        for (int st = 0; st < 4; ++st) {
            const v4f d = contract(A, a[st], bias);
            // lanes 0-15 hold o_0..o_2 of site 16 st + lane in rows 0-2
            if (l < 16)
#pragma unroll
                for (int r = 0; r < 3; ++r) out[3 * (w0 + 16 * st + l) + r] = d[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = out[3 * t + r];
}

// Draws s' around the direction hh from one mixture component of concentration
// kap, c' = s'.hh into *cn_out.  Heisenberg: c' = 1 + log(u + (1 - u) e^{-2 kappa}) / kappa
// with u = w.y (the Fisher CDF inverted), azimuth 2 pi w.z.  XY: von Mises by
// Best-Fisher rejection, the first attempt from words y, z of w, later ones
// from Philox calls whose counter word 3 is tg with the call index in bits 16
// and up; returns false when all kVmAttempts attempts failed (the site keeps
// its spin, with a probability that depends on the neighbours only).
template <int N>
__device__ __forceinline__ bool draw_around(const float hh[3], float kap, const u32x4& w, uint32_t h, uint32_t hs,
                                            uint32_t stream, uint32_t tg, uint32_t k0, uint32_t k1, float nw[3],
                                            float* cn_out) {
    float cn = 0.f;
    bool have = false;
    if (N == 3) {
        // Fisher: c = 1 + log(u + (1 - u) e^{-2 kappa}) / kappa, azimuth uniform
        const float u = u24f(w.y);
        cn = 1.0f + log_hw(u + (1.0f - u) * __expf(-2.0f * kap)) * __builtin_amdgcn_rcpf(kap);
        cn = fminf(1.0f, fmaxf(-1.0f, cn));
        const float sp = sin_rev(u24f(w.z)), cp = cos_rev(u24f(w.z));  // azimuth 2 pi u
        // orthonormal e1, e2 perpendicular to hh
        float t0 = 0.f, t1 = 0.f;
        if (fabsf(hh[0]) < 0.9f) t0 = 1.f; else t1 = 1.f;
        const float th = t0 * hh[0] + t1 * hh[1];
        float e0 = t0 - th * hh[0], e1 = t1 - th * hh[1], e2 = -th * hh[2];
        const float ei = rsq_hw(e0 * e0 + e1 * e1 + e2 * e2);
        e0 *= ei; e1 *= ei; e2 *= ei;
        const float f0 = hh[1] * e2 - hh[2] * e1, f1 = hh[2] * e0 - hh[0] * e2, f2 = hh[0] * e1 - hh[1] * e0;
        const float sn = sqrt_hw(fmaxf(0.0f, 1.0f - cn * cn));
        nw[0] = cn * hh[0] + sn * (cp * e0 + sp * f0);
        nw[1] = cn * hh[1] + sn * (cp * e1 + sp * f1);
        nw[2] = cn * hh[2] + sn * (cp * e2 + sp * f2);
        have = true;
    } else {
        // von Mises by Best-Fisher rejection.  An attempt takes two words (24-bit
        // uniforms for the angle and the test; the low bit of the first word
        // gives the attempt's sign): attempt 0 words y, z of the caller's call
        // (w), attempts 1 and 2 words x, y and z, w of one more call.  These
        // three are evaluated together, with no chain of dependent draws, and
        // the first accepted one wins, as if they had been made in turn.  Later
        // attempts run pair by pair, one call per pair.
        const float sr = 0.5f * __builtin_amdgcn_rcpf(kap);
        const float rr = sr + sqrt_hw(1.0f + sr * sr);
        auto try_vm = [&](uint32_t ua, uint32_t uv, float* W) -> bool {
            const float zc = cos_rev(0.5f * u24f(ua));  // cos(pi u)
            *W = (1.0f + rr * zc) * __builtin_amdgcn_rcpf(rr + zc);
            const float Y = kap * (rr - *W);
            const float V = fmaxf(u24f(uv), 1e-30f);
            return Y * (2.0f - Y) - V >= 0.0f || log_hw(Y) - log_hw(V) + 1.0f - Y >= 0.0f;
        };
        const u32x4 v1 = philox(h, hs, stream, tg | (1u << 16), k0, k1);
        const uint32_t ua[3] = {w.y, v1.x, v1.z}, uv[3] = {w.z, v1.y, v1.w};
        float Wsel = 0.f;
        uint32_t zsel = 0u;
#pragma unroll
        for (int k = 2; k >= 0; --k) {
            float Wk;
            if (try_vm(ua[k], uv[k], &Wk)) {
                Wsel = Wk;
                zsel = ua[k];
                have = true;
            }
        }
#pragma unroll 1
        for (int at = 2; 2 * at + 1 <= kVmAttempts && !have; ++at) {
            const u32x4 v = philox(h, hs, stream, tg | ((uint32_t)at << 16), k0, k1);
            float Wk;
            if (try_vm(v.x, v.y, &Wk)) {
                Wsel = Wk;
                zsel = v.x;
                have = true;
            } else if (try_vm(v.z, v.w, &Wk)) {
                Wsel = Wk;
                zsel = v.z;
                have = true;
            }
        }
        if (have) {
            cn = fminf(1.0f, fmaxf(-1.0f, Wsel));
            const float sn = sqrt_hw(fmaxf(0.0f, 1.0f - cn * cn)) * ((zsel & 1u) ? -1.0f : 1.0f);
            nw[0] = cn * hh[0] - sn * hh[1];
            nw[1] = sn * hh[0] + cn * hh[1];
        }
    }
    *cn_out = cn;
    return have;
}

// ---------------------------------------------------------------------------
// Learned Metropolis-Hastings pass over one colour (replaces the cap/rotation
// proposal when the ensemble's learned mode is on).  Same measurement fusion
// as the plain pass.
// ---------------------------------------------------------------------------
template <int N, int H, bool O32>
__device__ __forceinline__ void learned_tile(const ContArgs& a, int b, int color, int closes, int64_t t_override,
                                             double (*s_red)[4], uint16_t* s_feat, float* s_out, const char* s_frag) {
    int g = 0;
    while (g + 1 < a.ngroups && b >= a.groups[g + 1].blk0) ++g;
    const DevGroup G = a.groups[g];
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int ti = G.rep0 + a.r2t[grep];
    const float beta = a.betas_f[ti];
    const int64_t t = t_override >= 0 ? t_override : a.ctl->t;
    const int32_t fl = t_override >= 0 ? 0 : a.ctl->flags;
    const bool need = (fl & (F_MEAS | F_EXCH)) != 0 && closes;
    const bool renorm = (fl & F_RENORM) != 0;
    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);

    float* C = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
    const float* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
    const int64_t h = (int64_t)tile * 256 + threadIdx.x;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    int acc_n = 0, tried = 0;
    const bool live = h < G.n_half;
    uint32_t nb[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    float S[3] = {0.f, 0.f, 0.f};
    float x[kNIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, hh[3] = {0.f, 0.f, 0.f};
    float hn = 0.f;
    if (live) {
        // neighbours (same addressing as the plain pass); their values are not
        // kept across the network (registers), a measured sweep reads them again
        site_nbrs(G, (uint32_t)h, color, nb);
        float nv[3][6] = {};
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int q = 0; q < 6; ++q) nv[k][q] = *comp_at<O32>(O, G.stride, k, nb[q]);
#pragma unroll
        for (int k = 0; k < N; ++k) S[k] = *comp_at<O32>(C, G.stride, k, (uint32_t)h);
        hn = site_features<N>(beta, nv, x, hh);
    }
    // the proposal's parameters for every site of the block, on the matrix cores
    float o[3];
    mlp_block<H>(s_frag, x, o, s_feat, s_out);
    if (live) {
        bool accept = false;
        float nw[3] = {S[0], S[1], S[2]};
        if (hn > 1e-6f) {
            const Mix m = mixture(o);
            const u32x4 w = philox((uint32_t)h, hs, stream, (uint32_t)((7u << 8) | a.model), a.k0, a.k1);
            const float kap = u24f(w.x) < m.w0 ? m.k0 : m.k1;
            float cn = 0.f;
            const bool have = draw_around<N>(hh, kap, w, (uint32_t)h, hs, stream, (7u << 8) | a.model, a.k0, a.k1,
                                             nw, &cn);
            if (have) {
                float co = 0.f;
#pragma unroll
                for (int k = 0; k < N; ++k) co += S[k] * hh[k];
                const MixLog ql = mix_log<N>(m);
                const float lr = beta * a.Jf * hn * (cn - co) + log_q_fast(ql, co) - log_q_fast(ql, cn);
                accept = lr >= 0.0f || log_hw(fmaxf(u24f(w.w), 1e-30f)) < lr;
                tried = 1;
            }
        }
        if (accept) {
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] = nw[k];
            acc_n = 1;
        }
        if (renorm) {
            float q2 = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) q2 += S[k] * S[k];
            const float qi = rsqrtf(q2);
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] *= qi;
        }
        if (accept || renorm) {
#pragma unroll
            for (int k = 0; k < N; ++k) *comp_at<O32>(C, G.stride, k, (uint32_t)h) = S[k];
        }
        if (need) {
            double e = 0.0;
#pragma unroll
            for (int k = 0; k < N; ++k) {
                part[1 + k] = (double)S[k];
                if (color) {
                    double f = 0.0;
#pragma unroll
                    for (int q = 0; q < 6; ++q) f += (double)*comp_at<O32>(O, G.stride, k, nb[q]);
                    e += (double)S[k] * f;
                }
            }
            part[0] = color ? -a.J * e : 0.0;
        }
    }
    // acceptance statistics (every pass)
    {
        // 0/1 per lane: a ballot and a scalar bit count each
        const int wa = __popcll(__ballot(acc_n != 0)), wt = __popcll(__ballot(tried != 0));
        if ((threadIdx.x & 63) == 0 && wt) {
            int64_t* slot = a.mh + 8 * ((b * 4 + (threadIdx.x >> 6)) & (kMhSlots - 1));
            atomicAdd((unsigned long long*)&slot[0], (unsigned long long)wt);
            atomicAdd((unsigned long long*)&slot[1], (unsigned long long)wa);
        }
    }
    if (need) {
        wave_sum_parts<N>(part, color);
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < 4; ++k) s_red[k][threadIdx.x >> 6] = part[k];
        __syncthreads();
        if (threadIdx.x == 0) {
            double* dst = (color ? a.part1 : a.part0) + 4 * (int64_t)b;
            for (int k = 0; k < 4; ++k) dst[k] = s_red[k][0] + s_red[k][1] + s_red[k][2] + s_red[k][3];
        }
    }
}

// The pass's kernel: one tile per block, in the XCD-contiguous order, with the
// network's fragments copied into LDS per block.  (Four tiles per block, sharing
// one copy, measured slower at H = 64: the loop raised the kernel from 116 to
// 144 VGPRs, 3 waves per SIMD instead of 4; XY pass 163 against 148 us.)
template <int N, int H, bool O32>
__global__ __launch_bounds__(256) void cont_learned_kernel(ContArgs a, int nblocks, int color, int closes,
                                                           int64_t t_override) {
    __shared__ double s_red[4][4];
    __shared__ uint16_t s_feat[8 * 256];
    __shared__ float s_out[3 * 256];
    __shared__ uint4 s_frag[NetUnits<H>::BYTES / 16];
    const int per = (int)(gridDim.x >> 3);
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);  // XCD-contiguous tiles
    if (b >= nblocks) return;  // whole padding blocks, before any barrier
    stage_frag<H>(a.lfrag, s_frag);  // read after the first barrier in mlp_block
    learned_tile<N, H, O32>(a, b, color, closes, t_override, s_red, s_feat, s_out, (const char*)s_frag);
}

// ---------------------------------------------------------------------------
// Heat-bath pass over one colour: every site draws its new spin from the exact
// single-site conditional p(s | neighbours) ~ exp(beta J s.h) -- Fisher(h^, kappa)
// on S^2, von Mises(arg h, kappa) on S^1, kappa = beta |J| |h| -- with the
// samplers of the learned proposal (draw_around), and keeps it (the
// Metropolis-Hastings ratio of the exact conditional is 1).  The learned
// proposal depends on the neighbours only, like this one, so by Peskun's
// ordering no setting of its weights mixes faster per pass: this pass is the
// ceiling of that family.  Same measurement fusion and block order as the
// plain pass.  Draws: Philox(h, 2t + colour, stream, (11 << 8) | model).
// ---------------------------------------------------------------------------
template <int N, bool O32>
__global__ __launch_bounds__(256) void cont_heatbath_kernel(ContArgs a, int nblocks, int color, int closes,
                                                            int64_t t_override) {
    __shared__ double s_red[4][4];
    const int per = (int)(gridDim.x >> 3);
    const int b = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);  // XCD-contiguous tiles
    if (b >= nblocks) return;
    int g = 0;
    while (g + 1 < a.ngroups && b >= a.groups[g + 1].blk0) ++g;
    const DevGroup G = a.groups[g];
    int r, tile;
    split_tiles(G, b - G.blk0, &r, &tile);
    const int grep = G.rep0 + r;
    const int ti = G.rep0 + a.r2t[grep];
    const float beta = a.inline_params ? a.beta_inline : a.betas_f[ti];
    const int64_t t = t_override >= 0 ? t_override : a.ctl->t;
    const int32_t fl = t_override >= 0 ? 0 : a.ctl->flags;
    const bool need = (fl & (F_MEAS | F_EXCH)) != 0 && closes;
    const bool renorm = (fl & F_RENORM) != 0;
    const uint32_t hs = (uint32_t)(2 * t + color);
    const uint32_t stream = stream_id((uint32_t)G.L, (uint32_t)r);
    const uint32_t tg = tag(P_HEAT, a.model);

    float* C = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
    const float* O = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
    const int64_t h = (int64_t)tile * 256 + threadIdx.x;
    double part[4] = {0.0, 0.0, 0.0, 0.0};
    if (h < G.n_half) {
        uint32_t nb[6];
        site_nbrs(G, (uint32_t)h, color, nb);
        float nv[3][6] = {};
#pragma unroll
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int q = 0; q < 6; ++q) nv[k][q] = *comp_at<O32>(O, G.stride, k, nb[q]);
        float S[3] = {0.f, 0.f, 0.f}, hf[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < N; ++k) {
            S[k] = *comp_at<O32>(C, G.stride, k, (uint32_t)h);
            hf[k] = ((nv[k][0] + nv[k][1]) + (nv[k][2] + nv[k][3])) + (nv[k][4] + nv[k][5]);
        }
        float hn2 = 0.f;
#pragma unroll
        for (int k = 0; k < N; ++k) hn2 += hf[k] * hf[k];
        bool moved = false;
        if (hn2 > 1e-12f) {
            // the conditional leans along sign(J) h
            const float hi = __builtin_amdgcn_rsqf(hn2) * (a.Jf < 0.0f ? -1.0f : 1.0f);
            const float hh[3] = {hf[0] * hi, hf[1] * hi, hf[2] * hi};
            const float kap = beta * fabsf(a.Jf) * hn2 * fabsf(hi);
            const u32x4 w = philox((uint32_t)h, hs, stream, tg, a.k0, a.k1);
            float nw[3] = {S[0], S[1], S[2]}, cn = 0.f;
            if (kap > 1e-6f && draw_around<N>(hh, kap, w, (uint32_t)h, hs, stream, tg, a.k0, a.k1, nw, &cn)) {
#pragma unroll
                for (int k = 0; k < N; ++k) S[k] = nw[k];
                moved = true;
            }
        }
        if (renorm) {
            float q2 = 0.f;
#pragma unroll
            for (int k = 0; k < N; ++k) q2 += S[k] * S[k];
            const float qi = rsqrtf(q2);
#pragma unroll
            for (int k = 0; k < N; ++k) S[k] *= qi;
        }
        if (moved || renorm) {
#pragma unroll
            for (int k = 0; k < N; ++k) *comp_at<O32>(C, G.stride, k, (uint32_t)h) = S[k];
        }
        if (need) {
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
    if (need) {
        wave_sum_parts<N>(part, color);
        if ((threadIdx.x & 63) == 0)
            for (int k = 0; k < 4; ++k) s_red[k][threadIdx.x >> 6] = part[k];
        __syncthreads();
        if (threadIdx.x == 0) {
            double* dst = (color ? a.part1 : a.part0) + 4 * (int64_t)b;
            for (int k = 0; k < 4; ++k) dst[k] = s_red[k][0] + s_red[k][1] + s_red[k][2] + s_red[k][3];
        }
    }
}

template <bool O32>
static void launch_heatbath(const ContArgs& a, int nblocks, int color, int closes, int64_t t_override,
                            hipStream_t st) {
    const dim3 grid((nblocks + 7) & ~7);
    if (a.ncomp == 2)
        hipLaunchKernelGGL((cont_heatbath_kernel<2, O32>), grid, dim3(256), 0, st, a, nblocks, color, closes,
                           t_override);
    else
        hipLaunchKernelGGL((cont_heatbath_kernel<3, O32>), grid, dim3(256), 0, st, a, nblocks, color, closes,
                           t_override);
}

hipError_t launch_cont_heatbath(const ContArgs& a, int nblocks, int color, int closes, int64_t t_override,
                                hipStream_t st) {
    if (a.off32) launch_heatbath<true>(a, nblocks, color, closes, t_override, st);
    else launch_heatbath<false>(a, nblocks, color, closes, t_override, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Training: one sample per lane, 64 samples per block.  Samples are sites of
// the current configurations (mode 0) or explicit feature rows (mode 1, for
// gradient checks).  Block partial gradients go to gpart[block][P].
// ---------------------------------------------------------------------------
template <int N, int H>
__global__ __launch_bounds__(kLearnThreads) void learn_grad_kernel(ContArgs a, int64_t step, int mode,
                                                                    const float* __restrict__ xin,
                                                                    const float* __restrict__ cin, int64_t n_in,
                                                                    float* __restrict__ gpart,
                                                                    double* __restrict__ lpart) {
    using F = Off<H>;
    __shared__ float s_x[kNIn][kLearnThreads];
    __shared__ float s_a[3][H][kLearnThreads];   // activations of layers 1..3
    __shared__ float s_d[H][kLearnThreads];      // back-propagated pre-activation gradient
    __shared__ float s_do[3][kLearnThreads];
    __shared__ uint16_t s_w[F::P];
    const int t = threadIdx.x;
    for (int i = t; i < F::P; i += kLearnThreads) s_w[i] = a.net[i];
    __syncthreads();
    float x[kNIn], c = 0.f;
    bool live = true;
    if (mode == 1) {
        const int64_t idx = (int64_t)blockIdx.x * kLearnThreads + t;
        live = idx < n_in;
        for (int i = 0; i < kNIn; ++i) x[i] = live ? xin[idx * kNIn + i] : 0.0f;
        c = live ? cin[idx] : 0.0f;
    } else {
        // a uniformly random site of a uniformly random replica of the ensemble
        const u32x4 w = philox((uint32_t)(blockIdx.x * kLearnThreads + t), (uint32_t)step, 0xFFFF0000u,
                               (uint32_t)((8u << 8) | a.model), a.k0, a.k1);
        const int gr = (int)(((uint64_t)w.x * (uint32_t)a.total_reps) >> 32);
        int g = 0;
        while (g + 1 < a.ngroups && gr >= a.groups[g + 1].rep0) ++g;
        const DevGroup G = a.groups[g];
        const int r = gr - G.rep0;
        const int color = (int)(w.z & 1u);
        const int64_t h = (int64_t)(((uint64_t)w.y * (uint64_t)G.n_half) >> 32);
        uint32_t nb[6];
        site_nbrs(G, (uint32_t)h, color, nb);
        const float* Cp = a.pool + G.off + (int64_t)(2 * r + color) * N * G.stride;
        const float* Op = a.pool + G.off + (int64_t)(2 * r + (color ^ 1)) * N * G.stride;
        float nv[3][6];
        for (int k = 0; k < N; ++k)
            for (int q = 0; q < 6; ++q) nv[k][q] = Op[k * G.stride + nb[q]];
        float hh[3];
        const float beta = a.betas_f[G.rep0 + a.r2t[gr]];
        const float hn = site_features<N>(beta, nv, x, hh);
        live = hn > 1e-6f;
        for (int k = 0; k < N; ++k) c += Cp[k * G.stride + h] * hh[k];
    }
    for (int i = 0; i < kNIn; ++i) s_x[i][t] = x[i];
    float o[3];
    mlp_forward<H>(s_w, x, o, &s_a[0][0][t], kLearnThreads);
    // loss and d loss / d o for this sample
    const Mix m = mixture(o);
    const float l0 = __logf(m.w0) + log_comp<N>(m.k0, c);
    const float l1 = __logf(1.0f - m.w0) + log_comp<N>(m.k1, c);
    const float mx = fmaxf(l0, l1);
    const float e0 = __expf(l0 - mx), e1 = __expf(l1 - mx);
    const float lq = mx + __logf(e0 + e1);
    const float r0 = e0 / (e0 + e1), r1 = e1 / (e0 + e1);  // responsibilities
    const float f0 = __expf(log_comp<N>(m.k0, c) - lq), f1 = __expf(log_comp<N>(m.k1, c) - lq);
    const float sc = live ? 1.0f : 0.0f;
    s_do[0][t] = sc * -r0 * (c - mean_c<N>(m.k0)) * sigmoid_f(o[0]);
    s_do[1][t] = sc * -r1 * (c - mean_c<N>(m.k1)) * sigmoid_f(o[1]);
    s_do[2][t] = sc * -(f0 - f1) * m.w0 * (1.0f - m.w0);
    const double lsum = wave_sum(live ? (double)(-lq) : 0.0);
    if (t == 0) lpart[blockIdx.x] = lsum;
    __syncthreads();

    float* gb = gpart + (int64_t)blockIdx.x * F::P;
    // output layer: dW4[k][i] = sum_t do[k] a3[i]; db4
    for (int e = t; e < 3 * H; e += kLearnThreads) {
        const int k = e / H, i = e - k * H;
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_do[k][u] * s_a[2][i][u];
        gb[F::W4 + e] = sacc;
    }
    if (t < 3) {
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_do[t][u];
        gb[F::b4 + t] = sacc;
    }
    // d z3 = (W4^T do) act'(a3)
    for (int i = 0; i < H; ++i) {
        float v = 0.f;
        for (int k = 0; k < 3; ++k) v += bf2f(s_w[F::W4 + k * H + i]) * s_do[k][t];
        const float ai = s_a[2][i][t];
        s_d[i][t] = v * act_grad(ai);
    }
    __syncthreads();
    // layer 3 weights from d z3 and a2, then d z2 into s_a[2] (a3 no longer needed)
    for (int e = t; e < H * H; e += kLearnThreads) {
        const int j = e / H, i = e - j * H;
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_d[j][u] * s_a[1][i][u];
        gb[F::W3 + e] = sacc;
    }
    for (int j = t; j < H; j += kLearnThreads) {
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_d[j][u];
        gb[F::b3 + j] = sacc;
    }
    __syncthreads();
    for (int i = 0; i < H; ++i) {
        float v = 0.f;
        for (int j = 0; j < H; ++j) v += bf2f(s_w[F::W3 + j * H + i]) * s_d[j][t];
        const float ai = s_a[1][i][t];
        s_a[2][i][t] = v * act_grad(ai);
    }
    __syncthreads();
    // layer 2 weights from d z2 (in s_a[2]) and a1; then d z1 into s_d
    for (int e = t; e < H * H; e += kLearnThreads) {
        const int j = e / H, i = e - j * H;
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_a[2][j][u] * s_a[0][i][u];
        gb[F::W2 + e] = sacc;
    }
    for (int j = t; j < H; j += kLearnThreads) {
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_a[2][j][u];
        gb[F::b2 + j] = sacc;
    }
    __syncthreads();
    for (int i = 0; i < H; ++i) {
        float v = 0.f;
        for (int j = 0; j < H; ++j) v += bf2f(s_w[F::W2 + j * H + i]) * s_a[2][j][t];
        const float ai = s_a[0][i][t];
        s_d[i][t] = v * act_grad(ai);
    }
    __syncthreads();
    for (int e = t; e < H * kNIn; e += kLearnThreads) {
        const int j = e / kNIn, i = e - j * kNIn;
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_d[j][u] * s_x[i][u];
        gb[F::W1 + e] = sacc;
    }
    for (int j = t; j < H; j += kLearnThreads) {
        float sacc = 0.f;
        for (int u = 0; u < kLearnThreads; ++u) sacc += s_d[j][u];
        gb[F::b1 + j] = sacc;
    }
}

// Sum block partials in block order (deterministic), scale by 1/n, then Adam
// (or, with lr == 0, just write the mean gradient to `gout`).
__global__ void learn_adam_kernel(float* master, uint16_t* net, float* m1, float* m2, const float* gpart, int nblk,
                                  int P, float inv_n, float lr, float b1, float b2, float eps, int64_t step,
                                  float* gout) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += gpart[(int64_t)b * P + p];
    g *= inv_n;
    if (gout) gout[p] = g;
    if (lr == 0.0f) return;
    const float mm = b1 * m1[p] + (1.0f - b1) * g;
    const float vv = b2 * m2[p] + (1.0f - b2) * g * g;
    m1[p] = mm;
    m2[p] = vv;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    const float th = master[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    master[p] = th;
    net[p] = f2bf(th);
}

__global__ void learn_pack_kernel(const float* master, uint16_t* net, int P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) net[p] = f2bf(master[p]);
}

template <int H>
int learn_param_count() {
    return Off<H>::P;
}

int learn_params(int H) {
    switch (H) {
        case 8: return Off<8>::P;
        case 16: return Off<16>::P;
        case 32: return Off<32>::P;
        case 64: return Off<64>::P;
    }
    return -1;
}

// Network outputs o[3] for explicit feature rows, evaluated exactly as the
// fused sweep does (checks of the matrix-core path).
template <int H>
__global__ __launch_bounds__(256) void learn_infer_kernel(const uint4* __restrict__ frag, const float* __restrict__ xin,
                                                          int64_t n, float* __restrict__ out) {
    __shared__ uint16_t s_feat[8 * 256];
    __shared__ float s_out[3 * 256];
    __shared__ uint4 s_frag[NetUnits<H>::BYTES / 16];
    stage_frag<H>(frag, s_frag);  // read after the barrier in mlp_block
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float x[kNIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (i < n)
#pragma unroll
        for (int k = 0; k < kNIn; ++k) x[k] = xin[i * kNIn + k];
    float o[3];
    mlp_block<H>((const char*)s_frag, x, o, s_feat, s_out);
    if (i < n)
#pragma unroll
        for (int r = 0; r < 3; ++r) out[i * 3 + r] = o[r];
}

hipError_t launch_learn_infer(const uint4* frag, int H, const float* x, int64_t n, float* out, hipStream_t st) {
    const int nb = (int)((n + 255) / 256);
    switch (H) {
        case 8: hipLaunchKernelGGL(learn_infer_kernel<8>, dim3(nb), dim3(256), 0, st, frag, x, n, out); break;
        case 16: hipLaunchKernelGGL(learn_infer_kernel<16>, dim3(nb), dim3(256), 0, st, frag, x, n, out); break;
        case 32: hipLaunchKernelGGL(learn_infer_kernel<32>, dim3(nb), dim3(256), 0, st, frag, x, n, out); break;
        case 64: hipLaunchKernelGGL(learn_infer_kernel<64>, dim3(nb), dim3(256), 0, st, frag, x, n, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

size_t learn_frag_bytes(int H) {
    switch (H) {
        case 8: return NetUnits<8>::BYTES;
        case 16: return NetUnits<16>::BYTES;
        case 32: return NetUnits<32>::BYTES;
        case 64: return NetUnits<64>::BYTES;
    }
    return 0;
}

hipError_t launch_learn_frag(const uint16_t* net, int H, uint4* frag, hipStream_t st) {
    uint2* f = (uint2*)frag;
    switch (H) {
        case 8: hipLaunchKernelGGL(learn_frag_kernel<8>, dim3(1), dim3(64), 0, st, net, f); break;
        case 16: hipLaunchKernelGGL(learn_frag_kernel<16>, dim3(1), dim3(64), 0, st, net, f); break;
        case 32: hipLaunchKernelGGL(learn_frag_kernel<32>, dim3(1), dim3(64), 0, st, net, f); break;
        case 64: hipLaunchKernelGGL(learn_frag_kernel<64>, dim3(1), dim3(64), 0, st, net, f); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int N, int H>
static hipError_t launch_learned_nh(const ContArgs& a, int nblocks, int color, int closes, int64_t t_override,
                                    hipStream_t st) {
    const dim3 grid((nblocks + 7) & ~7);  // a multiple of the 8 XCDs
    if (a.off32)
        hipLaunchKernelGGL((cont_learned_kernel<N, H, true>), grid, dim3(256), 0, st, a, nblocks, color, closes,
                           t_override);
    else
        hipLaunchKernelGGL((cont_learned_kernel<N, H, false>), grid, dim3(256), 0, st, a, nblocks, color, closes,
                           t_override);
    return hipGetLastError();
}

hipError_t launch_cont_learned(const ContArgs& a, int nblocks, int color, int closes, int64_t t_override,
                               hipStream_t st) {
#define LEARN_CASE(HH)                                                                              \
    case HH:                                                                                        \
        return a.ncomp == 2 ? launch_learned_nh<2, HH>(a, nblocks, color, closes, t_override, st)   \
                            : launch_learned_nh<3, HH>(a, nblocks, color, closes, t_override, st);
    switch (a.learn_H) {
        LEARN_CASE(8)
        LEARN_CASE(16)
        LEARN_CASE(32)
        LEARN_CASE(64)
    }
#undef LEARN_CASE
    return hipErrorInvalidValue;
}

template <int N, int H>
static hipError_t launch_grad_nh(const ContArgs& a, int nblk, int64_t step, int mode, const float* xin,
                                 const float* cin, int64_t n_in, float* gpart, double* lpart, hipStream_t st) {
    hipLaunchKernelGGL((learn_grad_kernel<N, H>), dim3(nblk), dim3(kLearnThreads), 0, st, a, step, mode, xin, cin,
                       n_in, gpart, lpart);
    return hipGetLastError();
}

hipError_t launch_learn_grad(const ContArgs& a, int nblk, int64_t step, int mode, const float* xin, const float* cin,
                             int64_t n_in, float* gpart, double* lpart, hipStream_t st) {
#define GRAD_CASE(HH)                                                                                            \
    case HH:                                                                                                     \
        return a.ncomp == 2 ? launch_grad_nh<2, HH>(a, nblk, step, mode, xin, cin, n_in, gpart, lpart, st)       \
                            : launch_grad_nh<3, HH>(a, nblk, step, mode, xin, cin, n_in, gpart, lpart, st);
    switch (a.learn_H) {
        GRAD_CASE(8)
        GRAD_CASE(16)
        GRAD_CASE(32)
        GRAD_CASE(64)
    }
#undef GRAD_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_learn_adam(float* master, uint16_t* net, float* m1, float* m2, const float* gpart, int nblk, int P,
                             float inv_n, float lr, int64_t step, float* gout, hipStream_t st) {
    hipLaunchKernelGGL(learn_adam_kernel, dim3((P + 255) / 256), dim3(256), 0, st, master, net, m1, m2, gpart, nblk, P,
                       inv_n, lr, 0.9f, 0.999f, 1e-8f, step, gout);
    return hipGetLastError();
}

hipError_t launch_learn_pack(const float* master, uint16_t* net, int P, hipStream_t st) {
    hipLaunchKernelGGL(learn_pack_kernel, dim3((P + 255) / 256), dim3(256), 0, st, master, net, P);
    return hipGetLastError();
}

}  // namespace ising
