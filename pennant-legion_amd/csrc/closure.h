// Learned QCS closure: the network, its packed weights and the wave-level
// MFMA inference shared by the fused predictor kernel and the test kernels.
//
// The network maps 10 per-corner features to 2 coefficients:
//   x (10 features, slot 10 = 1 for the first layer's bias, slots 11-15 = 0)
//   h1 = relu(W1 x)            W1: 64 x 16 (column 10 is the bias)
//   h2 = relu(W2 h1 + b2)      W2: 64 x 64
//   y  = W3 h2 + b3            W3: 2 x 64
// BF16 weights and activations, FP32 accumulation (v_mfma_f32_32x32x16_bf16).
//
// Features of a compressing corner (cdiv < 0) at the predictor half step,
// with zss the zone sound speed, u the corner point velocities (k the
// corner, km/kp the previous/next corner of the zone), evol the corner
// characteristic length, du the corner velocity jump (setCornerDiv,
// src/QCS.cc:164-241), G = q2 (qgamma + 1) / 4 and slog(x) = sign(x)
// log(1 + |x|) (which keeps the unbounded ratios within a few units):
//   f0,f1 = slog((u_k - u_km) / zss)    f2,f3 = slog((u_kp - u_k) / zss)
//   f4 = slog(cdiv * evol / zss)        f5 = evol / min(elen_km, elen_k)
//   f6 = zss / (zss + |duav|)           f7 = zrp / zr
//   f8 = slog(du / zss)                 f9 = G du / (q1 zss + G du)
// Outputs: the linear and quadratic viscosity coefficients, scaled by q1 and
// G, so that the Kurapatenko coefficient of setQCnForce (src/QCS.cc:285-296)
// becomes  zkur = q1 y0 zss + G y1 du.  The exact values the network is
// trained on split zkur = t2 + sqrt(t1^2 + t2^2) (t1 = q1 zss, t2 = G du)
// with s = sqrt(t1^2 + t2^2) as
//   y0 = t1 / s,  y1 = 1 + t2 / s        (y0 = y1 = 1 when s = 0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

#define PN_MLP_NF 10          // features
#define PN_MLP_NH 64          // hidden units per layer
#define PN_MLP_NO 2           // outputs
#define PN_MLP_K1 16          // padded input width (features, bias slot, zeros)
// flat FP32 parameter vector: W1[64][16], W2[64][64], b2[64], W3[2][64], b3[2]
#define PN_MLP_OFF_W1 0
#define PN_MLP_OFF_W2 (PN_MLP_OFF_W1 + PN_MLP_NH * PN_MLP_K1)
#define PN_MLP_OFF_B2 (PN_MLP_OFF_W2 + PN_MLP_NH * PN_MLP_NH)
#define PN_MLP_OFF_W3 (PN_MLP_OFF_B2 + PN_MLP_NH)
#define PN_MLP_OFF_B3 (PN_MLP_OFF_W3 + PN_MLP_NO * PN_MLP_NH)
#define PN_MLP_NPARAM (PN_MLP_OFF_B3 + PN_MLP_NO)
#define PN_MLP_TUPLE 12       // a training tuple: 10 features, 2 targets

// Weights in MFMA A-operand order, one 16-byte fragment per lane:
//   w1[mb][lane]     rows 32mb + lane%32 of W1, slots 8(lane/32)..+7
//   w2[mb][c][lane]  rows 32mb + lane%32 of W2, inputs 16c + 8(t/4) + 4h + t%4
//   w3[c][lane]      rows lane%32 < 2 of W3 (other rows zero), same inputs
// The hidden-unit order of a k-chunk follows the accumulator layout of the
// previous layer (unit 8(r/4) + 4h + r%4 in register r of lane half h), so
// an accumulator feeds the next MFMA as its B operand without a transpose.
struct ClosurePack {
    bf16x8 w1[2][64];
    bf16x8 w2[2][4][64];
    bf16x8 w3[4][64];
    float b2c[2][2][16];      // b2 in accumulator order [mb][lane half][register]
    float b3[2];
    float lo[PN_MLP_NF];      // feature envelope of the training data
    float hi[PN_MLP_NF];
    float pad[10];
};

// sign(x) log(1 + |x|), from the hardware base-2 logarithm of 1 + |x| (the
// rounding of 1 + |x| bounds the absolute error by about 1e-7, far below the
// BF16 resolution of the network inputs)
__device__ __forceinline__ float slog(float x) {
    return copysignf(__builtin_amdgcn_logf(1.f + fabsf(x)) * 0.693147180559945309f, x);
}

// a and b rounded to BF16 (to nearest even), a in the low half: one
// conversion of the pair
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
    const f32x2 v = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// relu(a), relu(b) as a BF16 pair: the pair is rounded first, then every
// negative half (sign bit set, so negative as a 16-bit integer, -0 included)
// becomes +0 by an integer max.  Rounding keeps the sign, so this equals
// rounding relu(a), relu(b).
__device__ __forceinline__ unsigned relu_bf16(float a, float b) {
    const i16x2 h = __builtin_bit_cast(i16x2, pack_bf16(a, b));
    const i16x2 z = {0, 0};
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(h, z));
}

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// relu of 8 accumulator registers, rounded to BF16 as a B operand
__device__ __forceinline__ bf16x8 relu_chunk(const f32x16& a, int base) {
    u32x4 v;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        v[t] = relu_bf16(a[base + 2 * t], a[base + 2 * t + 1]);
    return as_bf16x8(v);
}

// Where a lane reads its weight fragments: straight from the packed weights
// (global memory or an LDS copy), or, for the hidden layer, from registers
// loaded once (MlpRegs), for a wave that evaluates the network several times
// in a row.  Both give the same operands, so the same results.
struct MlpPackSrc {
    const ClosurePack* __restrict__ w;
    int lane;
    __device__ __forceinline__ bf16x8 w1(int mb) const { return w->w1[mb][lane]; }
    __device__ __forceinline__ bf16x8 w2(int mb, int c) const { return w->w2[mb][c][lane]; }
    __device__ __forceinline__ bf16x8 w3(int c) const { return w->w3[c][lane]; }
    __device__ __forceinline__ float b2(int mb, int r) const { return w->b2c[mb][lane >> 5][r]; }
    __device__ __forceinline__ float b3(int o) const { return w->b3[o]; }
};
struct MlpRegs {
    const ClosurePack* __restrict__ w;
    int lane;
    bf16x8 w2_[2][4];     // the hidden layer's fragments (8 of the 14 MFMAs)
    __device__ __forceinline__ bf16x8 w1(int mb) const { return w->w1[mb][lane]; }
    __device__ __forceinline__ bf16x8 w2(int mb, int c) const { return w2_[mb][c]; }
    __device__ __forceinline__ bf16x8 w3(int c) const { return w->w3[c][lane]; }
    __device__ __forceinline__ float b2(int mb, int r) const { return w->b2c[mb][lane >> 5][r]; }
    __device__ __forceinline__ float b3(int o) const { return w->b3[o]; }
};
__device__ __forceinline__ void mlp_load_regs(const ClosurePack* __restrict__ w, int lane, MlpRegs& r) {
    r.w = w;
    r.lane = lane;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int c = 0; c < 4; ++c) r.w2_[mb][c] = w->w2[mb][c][lane];
}

// One group of 32 corners (the B operand columns): returns y0, y1 in lanes
// 0-31 (column lane), unspecified in lanes 32-63.  Wave-uniform control flow.
template <class W>
__device__ __forceinline__ void mlp_group(const W& w, bf16x8 x, float& y0, float& y1) {
    f32x16 a1[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        f32x16 z = {};
        a1[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w1(mb), x, z, 0, 0, 0);
    }
    bf16x8 h1[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) h1[c] = relu_chunk(a1[c >> 1], 8 * (c & 1));
    f32x16 a2[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) a2[mb][r] = w.b2(mb, r);
#pragma unroll
        for (int c = 0; c < 4; ++c)
            a2[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w2(mb, c), h1[c], a2[mb], 0, 0, 0);
    }
    bf16x8 h2[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) h2[c] = relu_chunk(a2[c >> 1], 8 * (c & 1));
    f32x16 a3 = {};
#pragma unroll
    for (int c = 0; c < 4; ++c)
        a3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.w3(c), h2[c], a3, 0, 0, 0);
    y0 = a3[0] + w.b3(0);
    y1 = a3[1] + w.b3(1);
}

// Both groups of a wave in one instruction stream, so that the MFMAs of one
// group issue while the other's wait on their inputs (the same operations
// as two mlp_group calls, so the same results).
template <class W>
__device__ __forceinline__ void mlp_group2(const W& w, bf16x8 xa, bf16x8 xb, float& a0, float& a1,
                                           float& b0, float& b1) {
    f32x16 pa[2], pb[2];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        f32x16 z = {};
        const bf16x8 w1 = w.w1(mb);
        pa[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, xa, z, 0, 0, 0);
        pb[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w1, xb, z, 0, 0, 0);
    }
    bf16x8 ha[4], hb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        ha[c] = relu_chunk(pa[c >> 1], 8 * (c & 1));
        hb[c] = relu_chunk(pb[c >> 1], 8 * (c & 1));
    }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            pa[mb][r] = w.b2(mb, r);
            pb[mb][r] = pa[mb][r];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bf16x8 w2 = w.w2(mb, c);
            pa[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2, ha[c], pa[mb], 0, 0, 0);
            pb[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w2, hb[c], pb[mb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        ha[c] = relu_chunk(pa[c >> 1], 8 * (c & 1));
        hb[c] = relu_chunk(pb[c >> 1], 8 * (c & 1));
    }
    f32x16 qa = {}, qb = {};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const bf16x8 w3 = w.w3(c);
        qa = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w3, ha[c], qa, 0, 0, 0);
        qb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w3, hb[c], qb, 0, 0, 0);
    }
    a0 = qa[0] + w.b3(0);
    a1 = qa[1] + w.b3(1);
    b0 = qb[0] + w.b3(0);
    b1 = qb[1] + w.b3(1);
}

// Evaluates the network on one corner per lane (64 corners).  lo/hi are the
// lane's packed input slots 0-7 and 8-15.  needA / needB (wave-uniform) say
// whether any lane of the low / high half needs a result; a half that needs
// none skips its MFMA group.  Each lane receives its own corner's outputs.
// With PAIR, a wave that needs both groups runs them as one stream
// (mlp_group2).
template <class W, bool PAIR = false>
__device__ __forceinline__ void mlp_wave_w(const W& w, u32x4 lo, u32x4 hi, bool needA, bool needB,
                                           float& y0, float& y1) {
    // group A: columns are the corners of lanes 0-31 (slots 0-7 from the low
    // lanes, slots 8-15 from the high lanes); group B: the corners of lanes
    // 32-63.  One swap of the upper half of `lo` with the lower half of `hi`
    // builds both operands.
    u32x4 opA, opB;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const auto r = __builtin_amdgcn_permlane32_swap(lo[i], hi[i], false, false);
        opA[i] = r[0];
        opB[i] = r[1];
    }
    float a0 = 0.f, a1 = 0.f, b0 = 0.f, b1 = 0.f;
    if (PAIR && needA && needB) {
        mlp_group2(w, as_bf16x8(opA), as_bf16x8(opB), a0, a1, b0, b1);
    } else {
        if (needA) mlp_group(w, as_bf16x8(opA), a0, a1);
        if (needB) mlp_group(w, as_bf16x8(opB), b0, b1);
    }
    // lanes 0-31 keep group A's results; lanes 32-63 take group B's results
    // from lanes 0-31
    const auto s0 = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a0),
                                                     __builtin_bit_cast(unsigned, b0), false, false);
    const auto s1 = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a1),
                                                     __builtin_bit_cast(unsigned, b1), false, false);
    y0 = __builtin_bit_cast(float, (unsigned)s0[0]);
    y1 = __builtin_bit_cast(float, (unsigned)s1[0]);
}
__device__ __forceinline__ void mlp_wave(const ClosurePack* __restrict__ w, u32x4 lo, u32x4 hi,
                                         bool needA, bool needB, float& y0, float& y1) {
    mlp_wave_w(MlpPackSrc{w, (int)(threadIdx.x & 63)}, lo, hi, needA, needB, y0, y1);
}

// Packs 10 features into the two input halves (slot 10 = 1 for the bias).
__device__ __forceinline__ void pack_features(const float (&f)[PN_MLP_NF], u32x4& lo, u32x4& hi) {
    lo[0] = pack_bf16(f[0], f[1]);
    lo[1] = pack_bf16(f[2], f[3]);
    lo[2] = pack_bf16(f[4], f[5]);
    lo[3] = pack_bf16(f[6], f[7]);
    hi[0] = pack_bf16(f[8], f[9]);
    hi[1] = pack_bf16(1.f, 0.f);
    hi[2] = 0u;
    hi[3] = 0u;
}

}  // namespace pn
