// guide_device.h -- device side of path guiding: input encoding and the
// von Mises-Fisher mixture the network outputs (see guiding.h for the
// network).  Used by the inference / training kernels and by the guided
// shading bounce.
#pragma once

#include <hip/hip_fp16.h>

#include "device_math.h"
#include "guiding.h"

namespace lucid {
namespace guide {

// Input feature idx (0..63) of a surface point: pn = position in the scene
// box mapped to [0,1]^3, n = shading normal, wo = outgoing direction.
//   0..47   sin / cos (2^k pi pn_c), k = 0..7: feature 3 q + c, q = 2k (sin) or 2k + 1 (cos)
//   48..50  2 pn - 1
//   51..53  n,  54..56 wo
//   57..59  sin(pi n),  60..62 sin(pi wo),  63 n . wo
__device__ __forceinline__ float feature(int idx, const float pn[3], v3f n, v3f wo) {
    if (idx < kFeatPos) {
        const int c = idx % 3, q = idx / 3, k = q >> 1;
        const float a = pn[c] * (float)(1 << k);  // in half turns
        return (q & 1) ? cospif(a) : sinpif(a);
    }
    const float nn[3] = {n.x, n.y, n.z}, ww[3] = {wo.x, wo.y, wo.z};
    if (idx < 51) return 2.0f * pn[idx - 48] - 1.0f;
    if (idx < 54) return nn[idx - 51];
    if (idx < 57) return ww[idx - 54];
    if (idx < 60) return sinpif(nn[idx - 57]);
    if (idx < 63) return sinpif(ww[idx - 60]);
    return dot(n, wo);
}

// sin(pi x) and cos(pi x) by the hardware's v_sin_f32 / v_cos_f32 (their
// argument is in revolutions; absolute error about 1e-6 for |x| <= 2, far
// below the BF16 rounding the network inputs take).
__device__ __forceinline__ float hw_sinpi(float x) { return __builtin_amdgcn_sinf(0.5f * x); }
__device__ __forceinline__ float hw_cospi(float x) { return __builtin_amdgcn_cosf(0.5f * x); }

// All 64 features at once: sin / cos (2^k pi pn_c) from one hardware sine
// and cosine per coordinate and the double-angle recurrence (the values
// feature() defines, to about 1e-5 absolute).
__device__ __forceinline__ void features(const float pn[3], v3f n, v3f wo, float f[kIn]) {
    for (int c = 0; c < 3; ++c) {
        float sn = hw_sinpi(pn[c]), cs = hw_cospi(pn[c]);
        for (int k = 0; k < kOctaves; ++k) {
            f[3 * (2 * k) + c] = sn;
            f[3 * (2 * k + 1) + c] = cs;
            const float s2 = 2.0f * sn * cs, c2 = cs * cs - sn * sn;
            sn = s2;
            cs = c2;
        }
    }
    f[48] = 2.0f * pn[0] - 1.0f; f[49] = 2.0f * pn[1] - 1.0f; f[50] = 2.0f * pn[2] - 1.0f;
    f[51] = n.x; f[52] = n.y; f[53] = n.z;
    f[54] = wo.x; f[55] = wo.y; f[56] = wo.z;
    f[57] = hw_sinpi(n.x); f[58] = hw_sinpi(n.y); f[59] = hw_sinpi(n.z);
    f[60] = hw_sinpi(wo.x); f[61] = hw_sinpi(wo.y); f[62] = hw_sinpi(wo.z);
    f[63] = dot(n, wo);
}

// features() written straight into a row of BF16 (no 64-float array in
// registers): the same values.
__device__ __forceinline__ void features_bf16(const float pn[3], v3f n, v3f wo, __bf16* row) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float sn = hw_sinpi(pn[c]), cs = hw_cospi(pn[c]);
#pragma unroll
        for (int k = 0; k < kOctaves; ++k) {
            row[3 * (2 * k) + c] = (__bf16)sn;
            row[3 * (2 * k + 1) + c] = (__bf16)cs;
            const float s2 = 2.0f * sn * cs, c2 = cs * cs - sn * sn;
            sn = s2;
            cs = c2;
        }
    }
    row[48] = (__bf16)(2.0f * pn[0] - 1.0f); row[49] = (__bf16)(2.0f * pn[1] - 1.0f);
    row[50] = (__bf16)(2.0f * pn[2] - 1.0f);
    row[51] = (__bf16)n.x; row[52] = (__bf16)n.y; row[53] = (__bf16)n.z;
    row[54] = (__bf16)wo.x; row[55] = (__bf16)wo.y; row[56] = (__bf16)wo.z;
    row[57] = (__bf16)hw_sinpi(n.x); row[58] = (__bf16)hw_sinpi(n.y); row[59] = (__bf16)hw_sinpi(n.z);
    row[60] = (__bf16)hw_sinpi(wo.x); row[61] = (__bf16)hw_sinpi(wo.y); row[62] = (__bf16)hw_sinpi(wo.z);
    row[63] = (__bf16)dot(n, wo);
}

// The mixture decoded from the 40 raw outputs: lobe k has mean direction
// normalize(o[5k..5k+2]), concentration exp(clamp(o[5k+3], -2, 7)) and
// weight softmax(o[5k+4]).
struct Mixture {
    v3f mu[kLobes];
    float kappa[kLobes], logc[kLobes], w[kLobes];
};
constexpr float kLogKappaMin = -2.0f, kLogKappaMax = 7.0f;

// log of the vMF normalisation kappa / (2 pi (1 - exp(-2 kappa)))
__device__ __forceinline__ float vmf_logc(float kappa) {
    return logf(kappa) - 1.8378770664093453f - logf(-expm1f(-2.0f * kappa));
}

__device__ __forceinline__ float out_f32(float x) { return x; }
__device__ __forceinline__ float out_f32(__half x) { return __half2float(x); }

// Raw outputs as FP32 (training) or FP16 (what the render-side inference
// stores per path: half the bytes; sampling and pdf read the same values).
// Returns false for a degenerate mixture (a non-finite weight, mean,
// concentration or normalisation), which the caller must not sample.
template <class T>
__device__ __forceinline__ bool decode(const T* o, Mixture& m) {
    float mx = -1e30f;
    for (int k = 0; k < kLobes; ++k) mx = fmaxf(mx, out_f32(o[5 * k + 4]));
    float sum = 0.0f;
    for (int k = 0; k < kLobes; ++k) {
        const v3f r = mk(out_f32(o[5 * k]), out_f32(o[5 * k + 1]), out_f32(o[5 * k + 2]));
        const float len = sqrtf(dot(r, r)) + 1e-6f;
        m.mu[k] = r * (1.0f / len);
        const float lk = fminf(fmaxf(out_f32(o[5 * k + 3]), kLogKappaMin), kLogKappaMax);
        m.kappa[k] = expf(lk);
        m.logc[k] = vmf_logc(m.kappa[k]);
        m.w[k] = expf(out_f32(o[5 * k + 4]) - mx);
        sum += m.w[k];
    }
    const float inv = 1.0f / sum;
    bool ok = isfinite(inv) && sum > 0.0f;
    for (int k = 0; k < kLobes; ++k) {
        m.w[k] *= inv;
        ok = ok && isfinite(m.w[k]) && isfinite(m.logc[k]) && isfinite(m.mu[k].x) && isfinite(m.mu[k].y) &&
             isfinite(m.mu[k].z);
    }
    return ok;
}

// Guide-selection probability at a surface point: the fixed fraction, or
// the network's sigmoid(o[kSel]) kept inside [kSelMin, kSelMax].
template <class T>
__device__ __forceinline__ float selection(const T* o, float alpha, int learn_sel) {
    if (!learn_sel) return alpha;
    const float a = 1.0f / (1.0f + expf(-out_f32(o[kSel])));
    return isfinite(a) ? fminf(fmaxf(a, kSelMin), kSelMax) : alpha;
}

__device__ __forceinline__ float mixture_pdf(const Mixture& m, v3f w) {
    float p = 0.0f;
    for (int k = 0; k < kLobes; ++k) p += m.w[k] * expf(m.logc[k] + m.kappa[k] * (dot(m.mu[k], w) - 1.0f));
    return p;
}

// Samples a lobe with u0, then the lobe with (u1, u2) (the inversion
// w = 1 + log(u1 + (1 - u1) exp(-2 kappa)) / kappa of the vMF cosine).
__device__ __forceinline__ v3f mixture_sample(const Mixture& m, float u0, float u1, float u2) {
    // the first lobe whose running weight sum exceeds u0 (the last one if
    // none does), selected without a dynamic register index
    float c = 0.0f, ka = m.kappa[kLobes - 1];
    v3f mu = m.mu[kLobes - 1];
    bool found = false;
#pragma unroll
    for (int i = 0; i < kLobes - 1; ++i) {
        c += m.w[i];
        const bool take = !found && u0 < c;
        ka = take ? m.kappa[i] : ka;
        mu = take ? m.mu[i] : mu;
        found = found || take;
    }
    const float wz = fminf(fmaxf(1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * ka)) / ka, -1.0f), 1.0f);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - wz * wz));
    const float phi = 2.0f * PI_F * u2;
    // orthonormal basis around mu (Duff et al. 2017, "Building an orthonormal basis, revisited")
    const float s = copysignf(1.0f, mu.z);
    const float a = -1.0f / (s + mu.z), b = mu.x * mu.y * a;
    const v3f t = mk(1.0f + s * mu.x * mu.x * a, s * b, -s * mu.x);
    const v3f bb = mk(b, s + mu.y * mu.y * a, -mu.y);
    return t * (r * cosf(phi)) + bb * (r * sinf(phi)) + mu * wz;
}

// ---- render side: the FP16 outputs of one path, decoded where they are
// used.  A guided vertex keeps the 44 raw outputs packed in 22 registers
// (instead of 48 decoded floats), takes the lobe weights in one pass,
// decodes only the chosen lobe to sample it, and evaluates the density at
// both of its directions (the NEE direction and the sampled one) in one
// more pass.  The transcendentals are the hardware's (v_exp_f32,
// v_log_f32, v_sqrt_f32, v_rcp_f32; about 1 ulp): the guide is this
// build's own, and sampling and density read the same decoded values, so
// the density stays the one the sampler draws from up to float rounding.
struct GuideOut {
    uint32_t h[kGoutStride / 2];  // outputs 2i (low half) and 2i + 1 (high half)
};
constexpr float kLog2E = 1.4426950408889634f, kLn2 = 0.6931471805599453f;
constexpr float kLog2TwoPi = 2.6514961294723187f;  // log2(2 pi)

__device__ __forceinline__ GuideOut load_out(const __half* go) {
    GuideOut g;
    const uint2* p = reinterpret_cast<const uint2*>(go);  // rows are 88 bytes: 8-byte aligned
#pragma unroll
    for (int i = 0; i < kGoutStride / 4; ++i) {
        const uint2 v = p[i];
        g.h[2 * i] = v.x;
        g.h[2 * i + 1] = v.y;
    }
    return g;
}
__device__ __forceinline__ float out_at(const GuideOut& g, int i) {
    const uint32_t w = g.h[i >> 1];
    const unsigned short b = (unsigned short)((i & 1) ? (w >> 16) : (w & 0xffffu));
    return __half2float(__ushort_as_half(b));
}
// false if any of the trained outputs is not finite (exponent all ones):
// such a vertex is left to plain sampling, as decode() leaves it
__device__ __forceinline__ bool out_finite(const GuideOut& g) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < kOutTrained; ++i) {
        const uint32_t b = (i & 1) ? (g.h[i >> 1] >> 16) : g.h[i >> 1];
        ok = ok && (b & 0x7c00u) != 0x7c00u;
    }
    return ok;
}
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// softmax of the weight logits: w_k = exp2((o_k - mx) log2 e) * inv_sum
struct LobeWeights {
    float mx, inv_sum, log2_sum;
};
__device__ __forceinline__ LobeWeights lobe_weights(const GuideOut& g) {
    LobeWeights lw;
    float mx = -1e30f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) mx = fmaxf(mx, out_at(g, 5 * k + 4));
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) sum += fast_exp2((out_at(g, 5 * k + 4) - mx) * kLog2E);
    lw.mx = mx;
    lw.inv_sum = fast_rcp(sum);
    lw.log2_sum = fast_log2(sum);
    return lw;
}
// mean direction r / (|r| + 1e-6) and concentration exp(clamp(o, -2, 7)) of
// a lobe from its raw outputs (decode()'s mapping)
__device__ __forceinline__ v3f lobe_mu(float rx, float ry, float rz) {
    const float il = fast_rcp(fast_sqrt(rx * rx + ry * ry + rz * rz) + 1e-6f);
    return mk(rx * il, ry * il, rz * il);
}
__device__ __forceinline__ float lobe_log_kappa(float o) { return fminf(fmaxf(o, kLogKappaMin), kLogKappaMax); }

// The mixture density at directions a and b (one pass over the lobes):
// p = sum_k exp2(log2(w_k C(kappa_k)) + kappa_k log2 e (mu_k . x - 1)),
// C(kappa) = kappa / (2 pi (1 - exp(-2 kappa))).
__device__ __forceinline__ void mixture_pdf2(const GuideOut& g, const LobeWeights& lw, v3f a, v3f b, float& pa,
                                             float& pb) {
    pa = 0.0f;
    pb = 0.0f;
#pragma unroll
    for (int k = 0; k < kLobes; ++k) {
        const v3f mu = lobe_mu(out_at(g, 5 * k), out_at(g, 5 * k + 1), out_at(g, 5 * k + 2));
        const float lk = lobe_log_kappa(out_at(g, 5 * k + 3));
        const float ka = fast_exp2(lk * kLog2E), k2 = ka * kLog2E;
        const float logc2 = lk * kLog2E - kLog2TwoPi - fast_log2(1.0f - fast_exp2(-2.0f * k2));
        const float lw2 = logc2 + (out_at(g, 5 * k + 4) - lw.mx) * kLog2E - lw.log2_sum;
        pa += fast_exp2(lw2 + k2 * (dot(mu, a) - 1.0f));
        pb += fast_exp2(lw2 + k2 * (dot(mu, b) - 1.0f));
    }
}

// Draws a direction: the lobe with u0 (the first whose running weight sum
// exceeds u0, the last if none does), then the lobe's cosine with u1 by
// inversion, w = 1 + log(u1 + (1 - u1) exp(-2 kappa)) / kappa, and its
// azimuth 2 pi u2.
__device__ __forceinline__ v3f mixture_sample_out(const GuideOut& g, const LobeWeights& lw, float u0, float u1,
                                                  float u2) {
    float c = 0.0f;
    float rx = out_at(g, 5 * (kLobes - 1)), ry = out_at(g, 5 * (kLobes - 1) + 1);
    float rz = out_at(g, 5 * (kLobes - 1) + 2), olk = out_at(g, 5 * (kLobes - 1) + 3);
    bool found = false;
#pragma unroll
    for (int i = 0; i < kLobes - 1; ++i) {
        c += fast_exp2((out_at(g, 5 * i + 4) - lw.mx) * kLog2E) * lw.inv_sum;
        const bool take = !found && u0 < c;
        rx = take ? out_at(g, 5 * i) : rx;
        ry = take ? out_at(g, 5 * i + 1) : ry;
        rz = take ? out_at(g, 5 * i + 2) : rz;
        olk = take ? out_at(g, 5 * i + 3) : olk;
        found = found || take;
    }
    const v3f mu = lobe_mu(rx, ry, rz);
    const float k2 = fast_exp2(lobe_log_kappa(olk) * kLog2E) * kLog2E;  // kappa log2 e
    // 1 + log(x) / kappa = 1 + log2(x) / (kappa log2 e)
    const float wz = fminf(fmaxf(1.0f + fast_log2(u1 + (1.0f - u1) * fast_exp2(-2.0f * k2)) * fast_rcp(k2), -1.0f),
                           1.0f);
    const float r = fast_sqrt(fmaxf(0.0f, 1.0f - wz * wz));
    // v_cos_f32 / v_sin_f32 take their argument in revolutions
    const float cp = __builtin_amdgcn_cosf(u2), sp = __builtin_amdgcn_sinf(u2);
    const float s = copysignf(1.0f, mu.z);
    const float a = -1.0f / (s + mu.z), bb0 = mu.x * mu.y * a;
    const v3f t = mk(1.0f + s * mu.x * mu.x * a, s * bb0, -s * mu.x);
    const v3f bb = mk(bb0, s + mu.y * mu.y * a, -mu.y);
    return t * (r * cp) + bb * (r * sp) + mu * wz;
}
// the guide-selection probability from the packed outputs (selection())
__device__ __forceinline__ float selection_out(const GuideOut& g, float alpha, int learn_sel) {
    if (!learn_sel) return alpha;
    const float a = 1.0f / (1.0f + expf(-out_at(g, kSel)));
    return isfinite(a) ? fminf(fmaxf(a, kSelMin), kSelMax) : alpha;
}

}  // namespace guide
}  // namespace lucid
