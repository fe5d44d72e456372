// guiding.h -- learned path guiding: network layout, encoding, vMF mixture.
//
// A small MLP maps a surface point to a mixture of 8 von Mises-Fisher lobes
// over directions:
//   input  64 features: frequency-encoded position (scene box -> [0,1]^3,
//          8 octaves), shading normal and outgoing direction
//   hidden 3 x 64, ReLU
//   output 40 = 8 lobes x (mean direction xyz, log concentration, weight logit),
//          plus output 40, the logit of the guide-selection probability
//          (learned when selection learning is on)
// Weights are BF16, arithmetic FP32 (MFMA accumulate).  The guided bounce
// mixes the network's density with BSDF sampling by one-sample MIS:
// pdf = alpha * p_guide + (1 - alpha) * p_bsdf, alpha fixed or learned per
// surface point.
//
// Device layout of the MFMA operands (v_mfma_f32_16x16x32_bf16, lane l,
// group g = l / 16): the network runs transposed, out^T = W^T in^T, with
// samples as the 16 columns.  The accumulator of output tile m holds
// features 16m + 4g + j (j < 4) of sample l % 16, and the next layer takes
// the two accumulators of tiles 2kb, 2kb+1 as its B operand for K block kb,
// i.e. the K index of block kb, group g, slot i is
//   perm(kb, g, i) = 32kb + (i < 4 ? 4g + i : 16 + 4g + i - 4).
// Weights are packed with the same permutation so no data moves between
// layers:
//   forward  A[l][m][kb][lane][8] = W_l[in = perm(kb, g, i)][out = 16m + l % 16]
//   backward B[l][m][kb][lane][8] = W_l[in = 16m + l % 16][out = perm(kb, g, i)]
#pragma once

#include <stdint.h>

namespace lucid {
namespace guide {

constexpr int kIn = 64, kHid = 64, kOut = 40, kOutPad = 48, kLobes = 8, kLayers = 4;
// output 40: selection logit; the render-side inference stores outputs
// 0..43 per path (FP16, 88 bytes: whole 8-byte runs of 4)
constexpr int kSel = 40, kOutTrained = 41, kGoutStride = 44;
// the learned selection probability is kept inside [kSelMin, kSelMax], so
// each technique always keeps some of the samples
constexpr float kSelMin = 0.02f, kSelMax = 0.98f;
constexpr int kOctaves = 8, kFeatPos = 6 * kOctaves;  // frequency features of the position

// per layer: input width, output width (padded to whole 16-column tiles)
constexpr int layer_in(int l) { return l == 0 ? kIn : kHid; }
constexpr int layer_out(int l) { return l == kLayers - 1 ? kOutPad : kHid; }
constexpr int layer_mtiles(int l) { return layer_out(l) / 16; }
// element offsets of each layer's weights [in][out] and biases in the
// master parameter vector
constexpr int w_offset(int l) {
    return l == 0 ? 0 : w_offset(l - 1) + layer_in(l - 1) * layer_out(l - 1) + layer_out(l - 1);
}
constexpr int b_offset(int l) { return w_offset(l) + layer_in(l) * layer_out(l); }
constexpr int kParams = w_offset(kLayers);
// packed BF16 operands per layer: 64 lanes x 8 values per (m, kb) tile
constexpr int fwd_tiles(int l) { return layer_mtiles(l) * 2; }
constexpr int bwd_tiles(int l) { return (layer_in(l) / 16) * 2; }
constexpr int fwd_offset(int l) { return l == 0 ? 0 : fwd_offset(l - 1) + fwd_tiles(l - 1) * 512; }
constexpr int bwd_offset(int l) { return l == 0 ? 0 : bwd_offset(l - 1) + bwd_tiles(l - 1) * 512; }
constexpr int kPackedFwd = fwd_offset(kLayers), kPackedBwd = bwd_offset(kLayers);

// K index of block kb, lane group g, slot i (see the header comment)
#if defined(__HIPCC__)
__host__ __device__
#endif
inline int perm(int kb, int g, int i) { return 32 * kb + (i < 4 ? 4 * g + i : 16 + 4 * g + i - 4); }

// Training sample: where, how it was reached, which way it went, the
// weight lum((L_end - L_vertex) / T_in) = lum(W_bsdf * L_incident), and the
// BSDF sampler's density of wi there (pbsdf: selection learning).
struct alignas(16) Sample {
    float p[3], n[3], wo[3], wi[3], weight, pbsdf, pad[2];
};
static_assert(sizeof(Sample) == 64, "64-byte training samples");

}  // namespace guide
}  // namespace lucid
