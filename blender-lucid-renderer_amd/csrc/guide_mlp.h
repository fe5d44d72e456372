// guide_mlp.h -- the guiding network on 16 samples per wavefront, with
// v_mfma_f32_16x16x32_bf16 (gfx950).  Lane l works on sample l % 16 and, in
// every operand, on the K / M slice of group g = l / 16 (layouts in
// guiding.h).
#pragma once

#include <hip/hip_runtime.h>

#include "guiding.h"

namespace lucid {
namespace guide {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct GuideDev {
    const uint4* fwd;     // packed forward A operands, 8 bf16 per lane per tile
    const uint4* bwd;     // packed backward A operands
    const float* params;  // FP32 master parameters (biases are read from here)
    float bmin[3], binv[3];  // scene box: pn = (p - bmin) * binv
    float alpha;             // guide-sampling probability of the guided bounce
    int learn_sel;           // 1: alpha per surface point from output kSel
};

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) {
    bf16x8 r;
    __builtin_memcpy(&r, &u, 16);
    return r;
}
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// B operand of the next layer from the accumulators of tiles 2kb, 2kb+1
__device__ __forceinline__ bf16x8 pack_b(f32x4 lo, f32x4 hi) {
    bf16x8 r;
    for (int i = 0; i < 4; ++i) {
        r[i] = (__bf16)lo[i];
        r[4 + i] = (__bf16)hi[i];
    }
    return r;
}

// ReLU, then the B operand of the next layer, on the packed BF16 values: a
// BF16 has the sign in its top bit, so max(x, 0) as signed 16-bit integers
// is the ReLU (v_pk_max_i16, two values per instruction; -0 becomes +0).
// Rounding commutes with it, so this equals pack_b of the ReLU'd floats.
__device__ __forceinline__ uint32_t relu_bf16x2(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef short s16x2 __attribute__((ext_vector_type(2)));
    const f32x2 f = {a, b};
    const bf16x2 p = __builtin_convertvector(f, bf16x2);  // v_cvt_pk_bf16_f32
    s16x2 v;
    __builtin_memcpy(&v, &p, 4);
    v = __builtin_elementwise_max(v, (s16x2)0);
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return u;
}
__device__ __forceinline__ bf16x8 pack_b_relu(f32x4 lo, f32x4 hi) {
    const uint4 u = {relu_bf16x2(lo[0], lo[1]), relu_bf16x2(lo[2], lo[3]), relu_bf16x2(hi[0], hi[1]),
                     relu_bf16x2(hi[2], hi[3])};
    return as_bf16x8(u);
}

// Forward pass.  `x` holds the lane's input operand for K blocks 0 and 1
// (features perm(kb, g, i) of its sample).  Outputs: the three output tiles
// (features 16m + 4g + j).  When `store` is given, the post-ReLU hidden
// activations of layers 0..2 are returned in store[l][m] (D layout).
__device__ __forceinline__ void mlp_forward(const GuideDev& G, bf16x8 x0, bf16x8 x1, f32x4 out[3],
                                            f32x4 (*store)[4] = nullptr) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    bf16x8 b0 = x0, b1 = x1;
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        const int mt = layer_mtiles(l);
        const float* bias = G.params + b_offset(l);
        f32x4 acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (m >= mt) break;
            for (int j = 0; j < 4; ++j) acc[m][j] = bias[16 * m + 4 * g + j];
            const uint4* a = G.fwd + (fwd_offset(l) + (m * 2) * 512) / 8 + lane;
            acc[m] = mfma(as_bf16x8(a[0]), b0, acc[m]);
            acc[m] = mfma(as_bf16x8(a[64]), b1, acc[m]);
        }
        if (l < kLayers - 1) {
#pragma unroll
            for (int m = 0; m < 4; ++m)
                for (int j = 0; j < 4; ++j) acc[m][j] = fmaxf(acc[m][j], 0.0f);
            if (store)
                for (int m = 0; m < 4; ++m) store[l][m] = acc[m];
            b0 = pack_b(acc[0], acc[1]);
            b1 = pack_b(acc[2], acc[3]);
        } else {
            out[0] = acc[0];
            out[1] = acc[1];
            out[2] = acc[2];
        }
    }
}

// Forward pass of NT tiles (16 * NT samples) per wavefront: each weight
// operand is loaded once and applied to every tile.  No activations kept.
template <int NT>
__device__ __forceinline__ void mlp_forward_tiles(const GuideDev& G, bf16x8 (&b0)[NT], bf16x8 (&b1)[NT],
                                                  f32x4 (&out)[NT][3]);

// The same with the packed weights at `fwd` and the biases of layer l at
// bias_of(l) (both may live in LDS; see stage_weights).
constexpr int kBiasStride = 64;
template <int NT, class BiasOf>
__device__ __forceinline__ void mlp_forward_tiles_at(const uint4* fwd, BiasOf bias_of, bf16x8 (&b0)[NT],
                                                     bf16x8 (&b1)[NT], f32x4 (&out)[NT][3]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        const int mt = layer_mtiles(l);
        const float* bias = bias_of(l);
        f32x4 acc[NT][4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            if (m >= mt) break;
            const uint4* a = fwd + (fwd_offset(l) + (m * 2) * 512) / 8 + lane;
            const bf16x8 a0 = as_bf16x8(a[0]), a1 = as_bf16x8(a[64]);
            f32x4 bv;
            for (int j = 0; j < 4; ++j) bv[j] = bias[16 * m + 4 * g + j];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                acc[t][m] = mfma(a0, b0[t], bv);
                acc[t][m] = mfma(a1, b1[t], acc[t][m]);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            if (l < kLayers - 1) {
                b0[t] = pack_b_relu(acc[t][0], acc[t][1]);
                b1[t] = pack_b_relu(acc[t][2], acc[t][3]);
            } else {
                out[t][0] = acc[t][0];
                out[t][1] = acc[t][1];
                out[t][2] = acc[t][2];
            }
        }
    }
}

// Copies the packed forward operands and the biases (kBiasStride apart per
// layer) into the workgroup's LDS arrays; the caller synchronises after.
__device__ __forceinline__ void stage_weights(const GuideDev& G, uint4* fwd, float* bias) {
    for (int i = threadIdx.x; i < kPackedFwd / 8; i += blockDim.x) fwd[i] = G.fwd[i];
    for (int i = threadIdx.x; i < kLayers * kBiasStride; i += blockDim.x) {
        const int l = i / kBiasStride, o = i % kBiasStride;
        bias[i] = o < layer_out(l) ? G.params[b_offset(l) + o] : 0.0f;
    }
}

// Forward pass with the weights read from global memory (tests).
template <int NT>
__device__ __forceinline__ void mlp_forward_tiles(const GuideDev& G, bf16x8 (&b0)[NT], bf16x8 (&b1)[NT],
                                                  f32x4 (&out)[NT][3]) {
    mlp_forward_tiles_at<NT>(G.fwd, [&](int l) { return G.params + b_offset(l); }, b0, b1, out);
}

}  // namespace guide
}  // namespace lucid
