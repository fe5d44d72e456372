// guiding.hip -- online training of the path-guiding network (gfx950).
//
// One training step on a batch of samples drawn from the sample ring:
//   k_select      batch indices (uniform over the valid ring entries)
//   k_wsum        sum of the batch's sample weights (loss normalisation)
//   k_fwd         MFMA forward pass, stores BF16 activations and FP32 outputs
//   k_loss        d loss / d outputs of  L = -sum_s w_s log p(wi_s) / sum_s w_s,
//                 plus, with selection learning, the selection logit's part of
//                 L_sel = -sum_s c_s log(a p(wi_s) + (1 - a) pbsdf_s)
//   k_bwd         MFMA backward pass through the hidden layers (BF16 deltas)
//   k_wgrad       weight and bias gradients dW_l = X_l^T dY_l, db_l = 1^T dY_l,
//                 split-K MFMA, FP32 atomics
//   k_adam        Adam on the FP32 master parameters, then k_pack repacks
//                 the BF16 MFMA operands
// Samples come from the tracer (weight lum(W_bsdf * L_incident) of a guided
// bounce): minimising the loss is minimising KL(target || guide) with the
// target proportional to f cos L_i (neural importance sampling).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "guide_device.h"
#include "guide_host.h"
#include "guide_mlp.h"

namespace lucid {
namespace guide {

namespace {

__device__ __forceinline__ uint32_t hash32(uint32_t a) {
    a ^= a >> 16; a *= 0x7feb352dU; a ^= a >> 15; a *= 0x846ca68bU; a ^= a >> 16;
    return a;
}
__device__ __forceinline__ float hash_unit(uint32_t a) { return (float)(hash32(a) >> 8) * 0x1p-24f; }

// He-uniform weights, zero biases, zero output padding (columns 40..47);
// the selection column (40) starts as the bias sel_bias alone, i.e. the
// configured guide fraction at every surface point.
__global__ void k_init(float* params, float* m, float* v, uint32_t seed, float sel_bias) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kParams) return;
    m[i] = 0.0f;
    v[i] = 0.0f;
    float val = 0.0f;
    for (int l = 0; l < kLayers; ++l) {
        const int w0 = w_offset(l), b0 = b_offset(l);
        if (i >= w0 && i < b0) {
            const int out = (i - w0) % layer_out(l);
            const float lim = sqrtf(6.0f / (float)layer_in(l)) * (l == kLayers - 1 ? 0.1f : 1.0f);
            val = (l == kLayers - 1 && out >= kOut) ? 0.0f : (2.0f * hash_unit(seed * 0x9e3779b9U + (uint32_t)i) - 1.0f) * lim;
        }
    }
    if (i == b_offset(kLayers - 1) + kSel) val = sel_bias;
    params[i] = val;
}

// Packed BF16 operands: thread per packed element.
__global__ void k_pack(const float* params, __bf16* fwd, __bf16* bwd) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < kPackedFwd) {
        int l = 0;
        while (l < kLayers - 1 && t >= fwd_offset(l + 1)) ++l;
        const int e = t - fwd_offset(l);
        const int tile = e / 512, r = e % 512, lane = r / 8, i = r % 8;
        const int m = tile / 2, kb = tile % 2, g = lane / 16;
        const int in = perm(kb, g, i), out = 16 * m + lane % 16;
        fwd[t] = (__bf16)params[w_offset(l) + in * layer_out(l) + out];
    }
    if (t < kPackedBwd) {
        int l = 0;
        while (l < kLayers - 1 && t >= bwd_offset(l + 1)) ++l;
        const int e = t - bwd_offset(l);
        const int tile = e / 512, r = e % 512, lane = r / 8, i = r % 8;
        const int m = tile / 2, kb = tile % 2, g = lane / 16;
        const int in = 16 * m + lane % 16, out = perm(kb, g, i);
        bwd[t] = out < layer_out(l) ? (__bf16)params[w_offset(l) + in * layer_out(l) + out] : (__bf16)0.0f;
    }
}

__global__ void k_select(int32_t* sel, int batch, uint32_t step, const unsigned long long* total, int capacity) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= batch) return;
    const unsigned long long n = *total;
    const uint32_t avail = (uint32_t)(n < (unsigned long long)capacity ? n : (unsigned long long)capacity);
    sel[b] = avail ? (int32_t)(hash32(step * 0x85ebca6bU ^ hash32((uint32_t)b + 0x1234567U)) % avail) : 0;
}

__global__ void k_wsum(const Sample* samples, const int32_t* sel, int batch, float* wsum) {
    __shared__ float part[256];
    const int b = blockIdx.x * 256 + threadIdx.x;
    part[threadIdx.x] = b < batch ? samples[sel[b]].weight : 0.0f;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(wsum, part[0]);
}

// Encoded input operand of the lane's sample.
__device__ __forceinline__ void encode(const GuideDev& G, const Sample& s, bf16x8& x0, bf16x8& x1) {
    const int g = (threadIdx.x & 63) >> 4;
    float pn[3];
    for (int c = 0; c < 3; ++c) pn[c] = (s.p[c] - G.bmin[c]) * G.binv[c];
    const v3f n = mk(s.n[0], s.n[1], s.n[2]), wo = mk(s.wo[0], s.wo[1], s.wo[2]);
    float f[kIn];
    features(pn, n, wo, f);
    for (int i = 0; i < 8; ++i) {
        x0[i] = (__bf16)f[perm(0, g, i)];
        x1[i] = (__bf16)f[perm(1, g, i)];
    }
}

// Forward of 16-sample tiles; one tile per wavefront per iteration.
// act[b][l][64] (BF16): l = 0 encoding, 1..3 hidden layers; out[b][48] FP32.
__global__ __launch_bounds__(64) void k_fwd(GuideDev G, const Sample* samples, const int32_t* sel, int batch,
                                            __bf16* act, float* out) {
    const int lane = threadIdx.x, g = lane >> 4, s = lane & 15;
    for (int tile = blockIdx.x; tile * 16 < batch; tile += gridDim.x) {
        const int b = tile * 16 + s;
        const bool ok = b < batch;
        Sample smp;
        if (ok) smp = samples[sel[b]];
        else __builtin_memset(&smp, 0, sizeof(smp));
        bf16x8 x0, x1;
        encode(G, smp, x0, x1);
        f32x4 o[3], h[3][4];
        mlp_forward(G, x0, x1, o, h);
        if (!ok) continue;
        __bf16* a = act + (size_t)b * 4 * 64;
        for (int i = 0; i < 8; ++i) {
            a[perm(0, g, i)] = x0[i];
            a[perm(1, g, i)] = x1[i];
        }
        for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 4; ++m)
                for (int j = 0; j < 4; ++j) a[(l + 1) * 64 + 16 * m + 4 * g + j] = (__bf16)h[l][m][j];
        for (int m = 0; m < 3; ++m)
            for (int j = 0; j < 4; ++j) out[(size_t)b * kOutPad + 16 * m + 4 * g + j] = o[m][j];
    }
}

// d loss / d raw outputs for L = -sum_s c_s log p(wi_s), c_s = w_s / sum w.
// With responsibilities r_k = w_k f_k / p:
//   d/d mean_k   = -c r_k kappa_k (wi - (mu_k . wi) mu_k) / |raw mean_k|
//   d/d logk_k   = -c r_k (1 - 2 kappa_k / (exp(2 kappa_k) - 1) + kappa_k (mu_k . wi - 1))  (0 outside the clamp)
//   d/d logit_k  = -c (r_k - w_k)
// Selection learning: the one-sample mixture q = a p + (1 - a) pbsdf with
// a = sigmoid(o_sel) is fitted to the same target through a alone (the
// guide's own loss stays the one above):
//   d/d o_sel    = -c a (1 - a) (p - pbsdf) / q
// and, like the loss value, only the guide's part enters `loss`.
__global__ void k_loss(const Sample* samples, const int32_t* sel, int batch, const float* out, const float* wsum,
                       float* dout, float* loss, int learn_sel) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= batch) return;
    const Sample s = samples[sel[b]];
    const float* o = out + (size_t)b * kOutPad;
    float* d = dout + (size_t)b * kOutPad;
    const float c = *wsum > 0.0f ? s.weight / *wsum : 0.0f;
    Mixture mx;
    (void)decode(o, mx);  // training takes the raw outputs as they are
    const v3f wi = mk(s.wi[0], s.wi[1], s.wi[2]);
    float lf[kLobes], p = 0.0f;
    for (int k = 0; k < kLobes; ++k) {
        lf[k] = mx.logc[k] + mx.kappa[k] * (dot(mx.mu[k], wi) - 1.0f);
        p += mx.w[k] * expf(lf[k]);
    }
    p = fmaxf(p, 1e-30f);
    for (int k = 0; k < kLobes; ++k) {
        const float r = mx.w[k] * expf(lf[k]) / p;
        const v3f raw = mk(o[5 * k], o[5 * k + 1], o[5 * k + 2]);
        const float len = sqrtf(dot(raw, raw)) + 1e-6f;
        const float ka = mx.kappa[k], mw = dot(mx.mu[k], wi);
        const v3f gm = (wi - mx.mu[k] * mw) * (-c * r * ka / len);
        d[5 * k] = gm.x; d[5 * k + 1] = gm.y; d[5 * k + 2] = gm.z;
        const float lk = o[5 * k + 3];
        const float t2 = 2.0f * ka, q = t2 / expm1f(t2);
        d[5 * k + 3] = (lk > kLogKappaMin && lk < kLogKappaMax) ? -c * r * (1.0f - q + ka * (mw - 1.0f)) : 0.0f;
        d[5 * k + 4] = -c * (r - mx.w[k]);
    }
    for (int f = kOut; f < kOutPad; ++f) d[f] = 0.0f;
    if (learn_sel && s.pbsdf > 0.0f) {
        const float a = 1.0f / (1.0f + expf(-o[kSel]));
        const float q = a * p + (1.0f - a) * s.pbsdf;
        if (q > 0.0f && isfinite(q)) d[kSel] = -c * a * (1.0f - a) * (p - s.pbsdf) / q;
    }
    atomicAdd(loss, -c * logf(p));
}

// Backward through the hidden layers: dY[l] = (W_{l+1} dY[l+1]) * relu'(h_{l+1}),
// l = 2, 1, 0, starting from dY[3] = dout.  dy[b][l][64] (BF16).
__global__ __launch_bounds__(64) void k_bwd(GuideDev G, int batch, const __bf16* act, const float* dout, __bf16* dy) {
    const int lane = threadIdx.x, g = lane >> 4, s = lane & 15;
    for (int tile = blockIdx.x; tile * 16 < batch; tile += gridDim.x) {
        const int b = tile * 16 + s;
        const bool ok = b < batch;
        // B operand: dY[3] of the lane's sample, K = output features (48, padded to 64)
        bf16x8 y0, y1;
        for (int i = 0; i < 8; ++i) {
            const int f0 = perm(0, g, i), f1 = perm(1, g, i);
            y0[i] = (__bf16)(ok ? dout[(size_t)b * kOutPad + f0] : 0.0f);
            y1[i] = (__bf16)(ok && f1 < kOutPad ? dout[(size_t)b * kOutPad + f1] : 0.0f);
        }
        if (ok) {
            __bf16* d3 = dy + ((size_t)b * 4 + 3) * 64;
            for (int i = 0; i < 8; ++i) {
                d3[perm(0, g, i)] = y0[i];
                d3[perm(1, g, i)] = y1[i];
            }
        }
        for (int l = kLayers - 1; l >= 1; --l) {
            // dX = W_l dY[l]: M = inputs of layer l (4 tiles), K = its outputs
            f32x4 acc[4];
            for (int m = 0; m < 4; ++m) {
                acc[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                const uint4* a = G.bwd + (bwd_offset(l) + (m * 2) * 512) / 8 + lane;
                acc[m] = mfma(as_bf16x8(a[0]), y0, acc[m]);
                acc[m] = mfma(as_bf16x8(a[64]), y1, acc[m]);
            }
            // relu' of layer l's input = hidden activation act[b][l] > 0
            for (int m = 0; m < 4; ++m)
                for (int j = 0; j < 4; ++j) {
                    const int f = 16 * m + 4 * g + j;
                    const float h = ok ? (float)act[((size_t)b * 4 + l) * 64 + f] : 0.0f;
                    acc[m][j] = h > 0.0f ? acc[m][j] : 0.0f;
                }
            y0 = pack_b(acc[0], acc[1]);
            y1 = pack_b(acc[2], acc[3]);
            if (ok) {
                __bf16* dl = dy + ((size_t)b * 4 + (l - 1)) * 64;
                for (int m = 0; m < 4; ++m)
                    for (int j = 0; j < 4; ++j) dl[16 * m + 4 * g + j] = (__bf16)acc[m][j];
            }
        }
    }
}

// dW_l[in][out] = sum_b X_l[b][in] dY_l[b][out]; X_l = act[b][l].  One
// wavefront per (layer, m, n) output tile and K range; FP32 atomics.  Tile
// row m = in / 16 (4 per layer) plus m = 4, the bias row: its A operand is a
// constant 1 in row 0, so row 0 of the product is db_l[out] = sum_b dY_l[b][out].
__global__ __launch_bounds__(64) void k_wgrad(int batch, int chunks_per_wave, const __bf16* act, const __bf16* dy,
                                              float* grad) {
    const int lane = threadIdx.x, g = lane >> 4, s = lane & 15;
    int t = blockIdx.x, l = 0;
    for (; l < kLayers; ++l) {
        const int nt = (layer_in(l) / 16 + 1) * layer_mtiles(l);
        if (t < nt) break;
        t -= nt;
    }
    if (l >= kLayers) return;
    const int ntile = layer_mtiles(l);
    const int m = t / ntile, n = t % ntile;
    const bool bias = m == layer_in(l) / 16;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int nchunks = (batch + 31) / 32;
    const int c0 = blockIdx.y * chunks_per_wave, c1 = min(nchunks, c0 + chunks_per_wave);
    for (int c = c0; c < c1; ++c) {
        bf16x8 a, b;
        for (int i = 0; i < 8; ++i) {
            const int bb = 32 * c + 8 * g + i;
            const bool ok = bb < batch;
            if (bias) a[i] = (ok && s == 0) ? (__bf16)1.0f : (__bf16)0.0f;
            else a[i] = ok ? act[((size_t)bb * 4 + l) * 64 + 16 * m + s] : (__bf16)0.0f;
            b[i] = ok ? dy[((size_t)bb * 4 + l) * 64 + 16 * n + s] : (__bf16)0.0f;
        }
        acc = mfma(a, b, acc);
    }
    // acc[j] = product row 4g + j, column 16n + s
    const int out = 16 * n + s;
    if (out >= (l == kLayers - 1 ? kOutTrained : layer_out(l))) return;
    if (bias) {
        if (g == 0) atomicAdd(&grad[b_offset(l) + out], acc[0]);
        return;
    }
    for (int j = 0; j < 4; ++j) atomicAdd(&grad[w_offset(l) + (16 * m + 4 * g + j) * layer_out(l) + out], acc[j]);
}

__global__ void k_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kParams) return;
    const float gr = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gr;
    const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_zero(float* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}
__global__ void k_zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0) *p = 0ull;
}

// The rendering-side forward (mlp_forward_tiles, 4 tiles per wavefront) on
// explicit samples, for tests.
__global__ __launch_bounds__(64) void k_infer(GuideDev G, const Sample* samples, int n, float* out) {
    const int lane = threadIdx.x, g = lane >> 4, s = lane & 15;
    for (int base = blockIdx.x * 64; base < n; base += gridDim.x * 64) {
        bf16x8 x0[4], x1[4];
        for (int t = 0; t < 4; ++t) {
            const int b = base + 16 * t + s;
            Sample smp;
            if (b < n) smp = samples[b];
            else __builtin_memset(&smp, 0, sizeof(smp));
            encode(G, smp, x0[t], x1[t]);
        }
        f32x4 o[4][3];
        mlp_forward_tiles<4>(G, x0, x1, o);
        for (int t = 0; t < 4; ++t) {
            const int b = base + 16 * t + s;
            if (b >= n) continue;
            for (int m = 0; m < 3; ++m)
                for (int j = 0; j < 4; ++j) {
                    const int f = 16 * m + 4 * g + j;
                    if (f < kOut) out[(size_t)b * kOut + f] = o[t][m][j];
                }
        }
    }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ host --
struct State {
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    __bf16 *fwd = nullptr, *bwd = nullptr;
    Sample* ring = nullptr;
    int capacity = 0;
    unsigned long long* total = nullptr;  // samples ever appended (ring head = total % capacity)
    int32_t* sel = nullptr;
    __bf16 *act = nullptr, *dy = nullptr;
    float *out = nullptr, *dout = nullptr, *scal = nullptr;  // scal: [0] weight sum, [1] loss
    int batch_cap = 0;
    int adam_t = 0;
    std::vector<void*> allocs;
};

static bool dalloc(State* s, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return false;
    s->allocs.push_back(*p);
    return true;
}

State* create(int capacity, int batch_cap) {
    State* s = new State();
    s->capacity = capacity;
    s->batch_cap = batch_cap;
    bool ok = dalloc(s, (void**)&s->params, kParams * 4) && dalloc(s, (void**)&s->m, kParams * 4) &&
              dalloc(s, (void**)&s->v, kParams * 4) && dalloc(s, (void**)&s->grad, kParams * 4) &&
              dalloc(s, (void**)&s->fwd, kPackedFwd * 2) && dalloc(s, (void**)&s->bwd, kPackedBwd * 2) &&
              dalloc(s, (void**)&s->ring, (size_t)capacity * sizeof(Sample)) &&
              dalloc(s, (void**)&s->total, 8) && dalloc(s, (void**)&s->sel, (size_t)batch_cap * 4) &&
              dalloc(s, (void**)&s->act, (size_t)batch_cap * 256 * 2) &&
              dalloc(s, (void**)&s->dy, (size_t)batch_cap * 256 * 2) &&
              dalloc(s, (void**)&s->out, (size_t)batch_cap * kOutPad * 4) &&
              dalloc(s, (void**)&s->dout, (size_t)batch_cap * kOutPad * 4) && dalloc(s, (void**)&s->scal, 16);
    if (!ok) {
        destroy(s);
        return nullptr;
    }
    return s;
}

void destroy(State* s) {
    if (!s) return;
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
}

void reset(State* s, uint32_t seed, float alpha0, hipStream_t st) {
    const float a = fminf(fmaxf(alpha0, kSelMin), kSelMax);
    k_init<<<nblk(kParams, 256), 256, 0, st>>>(s->params, s->m, s->v, seed, logf(a / (1.0f - a)));
    k_pack<<<nblk(kPackedFwd > kPackedBwd ? kPackedFwd : kPackedBwd, 256), 256, 0, st>>>(s->params, s->fwd, s->bwd);
    k_zero_u64<<<1, 64, 0, st>>>(s->total);
    s->adam_t = 0;
}

GuideDev dev(const State* s, const float bmin[3], const float bmax[3], float alpha, bool learn_sel) {
    GuideDev g;
    g.fwd = reinterpret_cast<const uint4*>(s->fwd);
    g.bwd = reinterpret_cast<const uint4*>(s->bwd);
    g.params = s->params;
    for (int c = 0; c < 3; ++c) {
        g.bmin[c] = bmin[c];
        const float e = bmax[c] - bmin[c];
        g.binv[c] = e > 0.0f ? 1.0f / e : 1.0f;
    }
    g.alpha = alpha;
    g.learn_sel = learn_sel ? 1 : 0;
    return g;
}

Sample* ring(State* s) { return s->ring; }
unsigned long long* ring_total(State* s) { return s->total; }
int ring_capacity(const State* s) { return s->capacity; }

// Forward + loss + backward + gradients of one batch given by `sel` (already
// on the device); gradients land in s->grad.
static void grad_step(State* s, const GuideDev& g, int batch, hipStream_t st) {
    k_zero<<<nblk(kParams, 256), 256, 0, st>>>(s->grad, kParams);
    k_zero<<<1, 256, 0, st>>>(s->scal, 4);
    k_wsum<<<nblk(batch, 256), 256, 0, st>>>(s->ring, s->sel, batch, s->scal);
    const unsigned tiles = nblk(batch, 16);
    k_fwd<<<tiles < 4096 ? tiles : 4096, 64, 0, st>>>(g, s->ring, s->sel, batch, s->act, s->out);
    k_loss<<<nblk(batch, 256), 256, 0, st>>>(s->ring, s->sel, batch, s->out, s->scal, s->dout, s->scal + 1,
                                             g.learn_sel);
    k_bwd<<<tiles < 4096 ? tiles : 4096, 64, 0, st>>>(g, batch, s->act, s->dout, s->dy);
    int wtiles = 0;
    for (int l = 0; l < kLayers; ++l) wtiles += (layer_in(l) / 16 + 1) * layer_mtiles(l);
    // 8 chunks of 32 samples per wavefront (128 K ranges of a 32,768-sample
    // batch): 2.4 ms of training per guided room pass, against 3.0 ms with
    // 32 chunks (fewer wavefronts, each a longer chain of dependent loads)
    const int nchunks = (batch + 31) / 32;
    const int cpw = 8, splits = (nchunks + cpw - 1) / cpw;
    k_wgrad<<<dim3(wtiles, splits), 64, 0, st>>>(batch, cpw, s->act, s->dy, s->grad);
}

void train(State* s, const GuideDev& g, int steps, int batch, float lr, hipStream_t st) {
    if (batch > s->batch_cap) batch = s->batch_cap;
    for (int i = 0; i < steps; ++i) {
        ++s->adam_t;
        k_select<<<nblk(batch, 256), 256, 0, st>>>(s->sel, batch, (uint32_t)s->adam_t, s->total, s->capacity);
        grad_step(s, g, batch, st);
        const float b1 = 0.9f, b2 = 0.999f;
        k_adam<<<nblk(kParams, 256), 256, 0, st>>>(s->params, s->grad, s->m, s->v, lr, b1, b2, 1e-8f,
                                                   1.0f - powf(b1, (float)s->adam_t), 1.0f - powf(b2, (float)s->adam_t));
        k_pack<<<nblk(kPackedFwd > kPackedBwd ? kPackedFwd : kPackedBwd, 256), 256, 0, st>>>(s->params, s->fwd,
                                                                                           s->bwd);
    }
}

int debug_step(State* s, const float* params, const Sample* samples, int n, const float bmin[3], const float bmax[3],
               float* out, float* grad, float* loss, hipStream_t st) {
    if (n <= 0 || n > s->batch_cap || n > s->capacity) return 1;
    if (hipMemcpyAsync(s->params, params, kParams * 4, hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    k_pack<<<nblk(kPackedFwd > kPackedBwd ? kPackedFwd : kPackedBwd, 256), 256, 0, st>>>(s->params, s->fwd, s->bwd);
    if (hipMemcpyAsync(s->ring, samples, (size_t)n * sizeof(Sample), hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    std::vector<int32_t> id(n);
    for (int i = 0; i < n; ++i) id[i] = i;
    if (hipMemcpyAsync(s->sel, id.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    GuideDev g = dev(s, bmin, bmax, 0.5f, true);
    grad_step(s, g, n, st);
    std::vector<float> o((size_t)n * kOutPad);
    if (hipMemcpyAsync(o.data(), s->out, o.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    if (hipMemcpyAsync(grad, s->grad, kParams * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    float sc[2];
    if (hipMemcpyAsync(sc, s->scal, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    if (hipStreamSynchronize(st) != hipSuccess) return 2;
    for (int i = 0; i < n; ++i)
        for (int f = 0; f < kOut; ++f) out[(size_t)i * kOut + f] = o[(size_t)i * kOutPad + f];
    *loss = sc[1];
    // the device ring now holds the debug samples: forget them
    k_zero_u64<<<1, 64, 0, st>>>(s->total);
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

int debug_infer(State* s, const float* params, const Sample* samples, int n, const float bmin[3], const float bmax[3],
                float* out, hipStream_t st) {
    if (n <= 0 || n > s->capacity || n > s->batch_cap) return 1;
    if (hipMemcpyAsync(s->params, params, kParams * 4, hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    k_pack<<<nblk(kPackedFwd > kPackedBwd ? kPackedFwd : kPackedBwd, 256), 256, 0, st>>>(s->params, s->fwd, s->bwd);
    if (hipMemcpyAsync(s->ring, samples, (size_t)n * sizeof(Sample), hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    GuideDev g = dev(s, bmin, bmax, 0.5f, false);
    k_infer<<<nblk(n, 64), 64, 0, st>>>(g, s->ring, n, s->out);
    if (hipMemcpyAsync(out, s->out, (size_t)n * kOut * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    k_zero_u64<<<1, 64, 0, st>>>(s->total);
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

int get_params(State* s, float* params, hipStream_t st) {
    if (hipMemcpyAsync(params, s->params, kParams * 4, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}
int set_params(State* s, const float* params, hipStream_t st) {
    if (hipMemcpyAsync(s->params, params, kParams * 4, hipMemcpyHostToDevice, st) != hipSuccess) return 2;
    k_pack<<<nblk(kPackedFwd > kPackedBwd ? kPackedFwd : kPackedBwd, 256), 256, 0, st>>>(s->params, s->fwd, s->bwd);
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

float last_loss(State* s, hipStream_t st) {
    float sc[2] = {0.0f, 0.0f};
    if (hipMemcpyAsync(sc, s->scal, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return 0.0f;
    if (hipStreamSynchronize(st) != hipSuccess) return 0.0f;
    return sc[1];
}

}  // namespace guide
}  // namespace lucid
