// MLP surrogate of the Herschel-Bulkley return map: 12 -> 64 -> 64 -> 64 -> 7,
// SiLU, bf16 weights and activations on MFMA with f32 accumulation.
//
// The layers run transposed, Z^T = W^T H^T + b, with samples along the MFMA's
// N dimension.  For v_mfma_f32_16x16x16_bf16 (wave64) the operand layouts are
//   A[i][k] (16 x 16): lane i + 16 (k / 4), element k % 4
//   B[k][j] (16 x 16): lane j + 16 (k / 4), element k % 4
//   D[i][j] (16 x 16): lane j + 16 (i / 4), element i % 4
// so the D tile of one layer (rows = units, columns = samples) is, element for
// element, the B operand of the next layer: activations never leave registers
// between layers.  Weights are pre-packed as A fragments (64 lanes x 4 bf16).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 7;
// f32 master parameters: W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][7] b4[7]
// (W[k][n]: input k to output n).
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + 12 * 64, kMlpW2 = kMlpB1 + 64, kMlpB2 = kMlpW2 + 64 * 64,
              kMlpW3 = kMlpB2 + 64, kMlpB3 = kMlpW3 + 64 * 64, kMlpW4 = kMlpB3 + 64, kMlpB4 = kMlpW4 + 64 * 7,
              kMlpParams = kMlpB4 + 7;
// Packed forward fragments (W^T as A operands), 256 bf16 each:
//   layer 1: 4 output tiles x 1 K step, layers 2 and 3: 4 x 4, layer 4: 1 x 4.
constexpr int kFragElems = 256;
constexpr int kFwdFrag1 = 0, kFwdFrag2 = 4, kFwdFrag3 = 20, kFwdFrag4 = 36, kFwdFrags = 40;
// Packed backward fragments (W as A operands) for layers 4, 3, 2: dH^T = W dZ^T.
//   layer 4: 4 row tiles x 1 K step, layers 3 and 2: 4 x 4.
constexpr int kBwdFrag4 = 0, kBwdFrag3 = 4, kBwdFrag2 = 20, kBwdFrags = 36;
// Packed biases (f32): b1[64] b2[64] b3[64] b4[16].
constexpr int kBias1 = 0, kBias2 = 64, kBias3 = 128, kBias4 = 192, kBiasPacked = 208;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// f32 -> bf16 bits, round to nearest even (inputs here are finite).
__host__ __device__ inline uint16_t bf16_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_value(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// sigmoid(z) = 1 / (1 + e^-z) with the hardware reciprocal (1 ulp): the result
// is rounded to bf16 (8 bits) before use, so a correctly rounded division buys
// nothing.
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid(z); }
// d silu / dz = s (1 + z (1 - s)), s = sigmoid(z)
__device__ inline float silu_grad(float z) {
    float s = sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// Four f32 to bf16, round to nearest even (v_cvt_pk_bf16_f32, the same rounding as bf16_bits).
__device__ inline bf16x4 pack4(f32x4 v) {
    const bf16x2v lo = __builtin_convertvector(f32x2{v[0], v[1]}, bf16x2v);
    const bf16x2v hi = __builtin_convertvector(f32x2{v[2], v[3]}, bf16x2v);
    bf16x4 r;
    __builtin_memcpy(&r, &lo, 4);
    __builtin_memcpy(reinterpret_cast<char*>(&r) + 4, &hi, 4);
    return r;
}

// Position of the n-th (from 0) set bit of m; n < popcount(m).
__host__ __device__ inline int nth_set_bit(unsigned long long m, int n) {
    unsigned w = (unsigned)m;
    int pos = 0, c = __builtin_popcount(w);
    if (n >= c) n -= c, w = (unsigned)(m >> 32), pos = 32;
    c = __builtin_popcount(w & 0xffffu);
    if (n >= c) n -= c, w >>= 16, pos += 16;
    c = __builtin_popcount(w & 0xffu);
    if (n >= c) n -= c, w >>= 8, pos += 8;
    c = __builtin_popcount(w & 0xfu);
    if (n >= c) n -= c, w >>= 4, pos += 4;
    c = __builtin_popcount(w & 0x3u);
    if (n >= c) n -= c, w >>= 2, pos += 2;
    return pos + (n >= (int)(w & 1u) ? 1 : 0);
}

// Orders this wavefront's LDS writes before other lanes of the same wavefront read them.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Fragment f of a packed buffer (LDS or global), this lane's 4 elements.
__device__ inline bf16x4 frag(const uint16_t* frags, int f, int lane) {
    return *reinterpret_cast<const bf16x4*>(frags + f * kFragElems + lane * 4);
}

__device__ inline f32x4 bias4(const float* bias, int base, int lane) {
    int n = base + 4 * (lane >> 4);
    return f32x4{bias[n], bias[n + 1], bias[n + 2], bias[n + 3]};
}

// Activations of one 16-sample tile through the network, in D layout.
struct MlpTile {
    f32x4 z[3][4];  // pre-activations of the hidden layers (unit tile, 4 per lane)
    bf16x4 h[3][4]; // SiLU outputs as the next layer's B operand
    f32x4 y;        // outputs (units 0..15 of which 0..6 are used)
};

// Forward pass of one 16-sample tile.  x: this lane's 4 input features (B
// operand of layer 1: features 4 (lane / 16) .. + 3 of sample lane % 16).
// keep_z: keep pre-activations for the backward pass.
template <bool keep_z>
__device__ inline void mlp_tile_forward(const uint16_t* fwd, const float* bias, bf16x4 x, int lane, MlpTile& t) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = bias4(bias, kBias1 + nt * 16, lane);
        acc = mfma(frag(fwd, kFwdFrag1 + nt, lane), x, acc);
        if (keep_z) t.z[0][nt] = acc;
        t.h[0][nt] = pack4(f32x4{silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3])});
    }
    // The barriers keep the weight loads of later tiles from being hoisted
    // ahead: 40 fragments in flight would take 80 registers.
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int l = 1; l < 3; ++l) {
        const int fbase = l == 1 ? kFwdFrag2 : kFwdFrag3;
        const int bbase = l == 1 ? kBias2 : kBias3;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = bias4(bias, bbase + nt * 16, lane);
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma(frag(fwd, fbase + nt * 4 + kt, lane), t.h[l - 1][kt], acc);
            if (keep_z) t.z[l][nt] = acc;
            t.h[l][nt] = pack4(f32x4{silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3])});
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    f32x4 acc = bias4(bias, kBias4, lane);
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) acc = mfma(frag(fwd, kFwdFrag4 + kt, lane), t.h[2][kt], acc);
    t.y = acc;
}

// Weight W_l[k][n] of the f32 master parameters (zero outside the layer).
__host__ __device__ inline float mlp_weight(const float* p, int l, int k, int n) {
    const int base[4] = {kMlpW1, kMlpW2, kMlpW3, kMlpW4}, nin[4] = {12, 64, 64, 64}, nout[4] = {64, 64, 64, 7};
    return (k < nin[l] && n < nout[l]) ? p[base[l] + k * nout[l] + n] : 0.0f;
}

// Element idx of the packed forward fragments: fragment f = (layer, output tile
// nt, K step kt), A operand W^T: lane, e hold W[kt*16 + 4 (lane/16) + e][nt*16 + lane%16].
__host__ __device__ inline uint16_t mlp_fwd_elem(const float* p, int idx) {
    int f = idx / kFragElems, w = idx % kFragElems, lane = w / 4, e = w % 4;
    int l, nt, kt;
    if (f < kFwdFrag2) {
        l = 0, nt = f, kt = 0;
    } else if (f < kFwdFrag4) {
        l = f < kFwdFrag3 ? 1 : 2;
        int g = f - (l == 1 ? kFwdFrag2 : kFwdFrag3);
        nt = g / 4, kt = g % 4;
    } else {
        l = 3, nt = 0, kt = f - kFwdFrag4;
    }
    return bf16_bits(mlp_weight(p, l, kt * 16 + 4 * (lane / 16) + e, nt * 16 + lane % 16));
}

// Element idx of the packed backward fragments: fragment (layer, row tile rt,
// K step ct), A operand W: lane, e hold W[rt*16 + lane%16][ct*16 + 4 (lane/16) + e].
__host__ __device__ inline uint16_t mlp_bwd_elem(const float* p, int idx) {
    int f = idx / kFragElems, w = idx % kFragElems, lane = w / 4, e = w % 4;
    int l, rt, ct;
    if (f < kBwdFrag3) {
        l = 3, rt = f, ct = 0;
    } else {
        l = f < kBwdFrag2 ? 2 : 1;
        int g = f - (l == 2 ? kBwdFrag3 : kBwdFrag2);
        rt = g / 4, ct = g % 4;
    }
    return bf16_bits(mlp_weight(p, l, rt * 16 + lane % 16, ct * 16 + 4 * (lane / 16) + e));
}

__host__ __device__ inline float mlp_bias_elem(const float* p, int i) {
    if (i < kBias2) return p[kMlpB1 + i];
    if (i < kBias3) return p[kMlpB2 + i - kBias2];
    if (i < kBias4) return p[kMlpB3 + i - kBias3];
    return i - kBias4 < 7 ? p[kMlpB4 + i - kBias4] : 0.0f;
}

}  // namespace mpm
