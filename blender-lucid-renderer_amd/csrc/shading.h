// shading.h -- Principled BSDF, light sampling and volume helpers (device).
//
// Restates, for gfx950 device code, the reference functions cited below
// (paths relative to LucidRenderer/cpp_renderer/include/).  Expression order
// follows the reference so that IEEE-exact steps give identical bits.
#pragma once

#include <hip/hip_fp16.h>

#include "device_math.h"
#include "scene_pack.h"

namespace lucid {

constexpr float GGX_EPS = 1e-6f;
constexpr float MIN_ROUGH = 0.01f;
constexpr float MIN_PDF = 1e-6f;
constexpr float RAY_T_MIN = 1e-4f;
constexpr float RAY_T_MAX = 1e30f;

struct Mat {  // bsdf/bsdf.hpp:24-37 (roughness already clamped)
    v3f albedo;
    float metallic, roughness, transmission, ior;
};

// fresnel.hpp:40-48, :74-97
__device__ __forceinline__ v3f fresnel_schlick_color(float c, v3f f0) {
    float f = powf(1.0f - c, 5.0f);
    return mk(f0.x + (1.0f - f0.x) * f, f0.y + (1.0f - f0.y) * f, f0.z + (1.0f - f0.z) * f);
}
__device__ __forceinline__ float fresnel_dielectric(float ci, float eta) {
    ci = fabsf(ci);
    ci = (ci < 0.0f) ? 0.0f : ((1.0f < ci) ? 1.0f : ci);
    float si2 = 1.0f - ci * ci;
    float st2 = eta * eta * si2;
    if (st2 >= 1.0f) return 1.0f;
    float ct = sqrtf(smax(0.0f, 1.0f - st2));
    float eci = eta * ci, ect = eta * ct;
    float rs = (eci - ct) / (eci + ct + 1e-6f);
    float rp = (ci - ect) / (ci + ect + 1e-6f);
    return 0.5f * (rs * rs + rp * rp);
}
// ggx.hpp:28-206
__device__ __forceinline__ float safe_sqrt(float x) { return sqrtf(smax(0.0f, x)); }
__device__ __forceinline__ void onb(v3f n, v3f& t, v3f& b) {
    if (n.z < -0.9999f) {
        t = mk(0.0f, -1.0f, 0.0f);
        b = mk(-1.0f, 0.0f, 0.0f);
        return;
    }
    float a = 1.0f / (1.0f + n.z);
    float bb = -n.x * n.y * a;
    t = make_dir_or(mk(1.0f - n.x * n.x * a, bb, -n.x), mk(0.0f, 0.0f, 1.0f));
    b = make_dir_or(mk(bb, 1.0f - n.y * n.y * a, -n.y), mk(0.0f, 0.0f, 1.0f));
}
__device__ __forceinline__ v3f local_to_world(v3f l, v3f n, v3f t, v3f b) {
    return make_dir_or(mk(t.x * l.x + b.x * l.y + n.x * l.z, t.y * l.x + b.y * l.y + n.y * l.z,
                          t.z * l.x + b.z * l.y + n.z * l.z),
                       mk(0.0f, 0.0f, 1.0f));
}
__device__ __forceinline__ float ggx_d(float nh, float r) {
    float al = r * r, a2 = al * al;
    float nh2 = nh * nh;
    float den = nh2 * (a2 - 1.0f) + 1.0f;
    return a2 * INV_PI_F / (den * den + GGX_EPS);
}
__device__ __forceinline__ float ggx_g1(float nv, float r) {
    float al = r * r, a2 = al * al;
    return 2.0f * nv / (nv + safe_sqrt(a2 + (1.0f - a2) * nv * nv));
}
__device__ __forceinline__ float ggx_g2(float nl, float nv, float r) {
    if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
    float al = r * r, a2 = al * al;
    float ll = safe_sqrt(a2 + (1.0f - a2) * nl * nl);
    float lv = safe_sqrt(a2 + (1.0f - a2) * nv * nv);
    float den = nv * ll + nl * lv;
    if (den < GGX_EPS) return 0.0f;
    return 2.0f * nl * nv / den;
}
__device__ __forceinline__ v3f sample_vndf(v3f wo, float r, float u1, float u2, v3f n, v3f t, v3f b) {
    v3f wl = make_dir_or(mk(dot(wo, t), dot(wo, b), dot(wo, n)), mk(0.0f, 0.0f, 1.0f));
    float al = r * r;
    v3f ws = make_dir_or(mk(wl.x * al, wl.y * al, wl.z), mk(0.0f, 0.0f, 1.0f));
    v3f t1 = mk(1.0f, 0.0f, 0.0f);
    if (ws.z < 0.9999f) {
        v3f c;
        if (make_dir(cross(mk(0.0f, 0.0f, 1.0f), ws), c)) t1 = c;
    }
    v3f t2;
    if (!make_dir(cross(ws, t1), t2)) t2 = mk(0.0f, 1.0f, 0.0f);
    float rr = sqrtf(u1);
    float phi = 2.0f * PI_F * u2;
    float p1 = rr * cosf(phi);
    float p2 = rr * sinf(phi);
    float s = 0.5f * (1.0f + ws.z);
    p2 = (1.0f - s) * sqrtf(1.0f - p1 * p1) + s * p2;
    v3f h = t1 * p1 + t2 * p2 + ws * sqrtf(smax(0.0f, 1.0f - p1 * p1 - p2 * p2));
    h = mk(h.x * al, h.y * al, smax(0.0f, h.z));
    return local_to_world(make_dir_or(h, mk(0.0f, 0.0f, 1.0f)), n, t, b);
}
__device__ __forceinline__ float pdf_vndf(v3f wo, v3f h, float r, v3f n) {
    float nh = dot(n, h), vh = dot(wo, h), nv = dot(n, wo);
    if (nh <= 0.0f || vh <= 0.0f || nv <= 0.0f) return 0.0f;
    return ggx_d(nh, r) * ggx_g1(nv, r) / (4.0f * nv);
}
__device__ __forceinline__ v3f sample_cosine(float u1, float u2, v3f n) {
    float r = sqrtf(u1);
    float th = 2.0f * PI_F * u2;
    float x = r * cosf(th), y = r * sinf(th);
    float z = sqrtf(smax(0.0f, 1.0f - u1));
    v3f t, b;
    onb(n, t, b);
    return local_to_world(mk(x, y, z), n, t, b);
}
__device__ __forceinline__ float pdf_cosine(float nl) { return smax(0.0f, nl) * INV_PI_F; }

// evalSpecular + evalBSDF, bsdf/bsdf.hpp:126-190
__device__ __forceinline__ v3f eval_bsdf(const Mat& m, v3f wo, v3f wi, v3f n) {
    float nl = dot(n, wi), nv = dot(n, wo);
    if (nl <= 0.0f || nv <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    v3f h = make_dir_or(wo + wi, mk(0.0f, 0.0f, 1.0f));
    float nh = smax(dot(n, h), 0.0f);
    float vh = smax(dot(wo, h), 0.0f);
    float r = smax(m.roughness, MIN_ROUGH);
    float D = ggx_d(nh, r);
    float al = r * r, a2 = al * al;
    float ll = safe_sqrt(a2 + (1.0f - a2) * nl * nl);
    float lv = safe_sqrt(a2 + (1.0f - a2) * nv * nv);
    float g2d = 0.5f / (nv * ll + nl * lv + GGX_EPS);
    v3f f0 = m.albedo * m.metallic + mk(0.04f, 0.04f, 0.04f) * (1.0f - m.metallic);
    v3f F = fresnel_schlick_color(vh, f0);
    float spec = D * g2d;
    v3f sp = mk(spec * F.x, spec * F.y, spec * F.z);
    v3f Fd = fresnel_schlick_color(vh, mk(0.04f, 0.04f, 0.04f));
    const float INV_PI_D = 0.31830988618f;  // render_units.hpp:1240-1247
    v3f diff = mk(m.albedo.x * INV_PI_D, m.albedo.y * INV_PI_D, m.albedo.z * INV_PI_D);
    v3f kd = mk((1.0f - Fd.x) * (1.0f - m.metallic), (1.0f - Fd.y) * (1.0f - m.metallic),
                (1.0f - Fd.z) * (1.0f - m.metallic));
    return cmul(diff, kd) + sp;
}
__device__ __forceinline__ float spec_prob(const Mat& m) {  // bsdf.hpp:279-285
    if (m.metallic > 0.99f) return 1.0f;
    float sp = 0.5f * (1.0f + m.metallic) * (1.0f - m.roughness * 0.5f);
    return (sp < 0.1f) ? 0.1f : ((0.9f < sp) ? 0.9f : sp);
}
// pdfBSDF, bsdf.hpp:352-376
__device__ __forceinline__ float pdf_bsdf(const Mat& m, v3f wo, v3f wi, v3f n) {
    float nl = dot(n, wi);
    if (nl <= 0.0f) return 0.0f;
    float sp = spec_prob(m);
    float r = smax(m.roughness, MIN_ROUGH);
    v3f h = make_dir_or(wo + wi, mk(0.0f, 0.0f, 1.0f));
    return (sp * pdf_vndf(wo, h, r, n) + (1.0f - sp) * pdf_cosine(nl)) * (1.0f - m.transmission);
}
// The density with which sample_bsdf's opaque lobes actually draw wi.  Its
// VNDF branch (Heitz 2018) samples the exact GGX distribution of visible
// normals, while ggx_d above carries the reference's 1e-6 guard in the
// denominator (ggx.hpp:87-99), which flattens the peak of D below roughness
// ~0.2 (at 0.1 by two orders of magnitude).  So pdf_bsdf is the sampler's
// density only where the guard is negligible.  The plain tracers weigh a
// BSDF sample by eval / pdf_bsdf exactly as the reference does; the guided
// bounce needs the true density to estimate the same thing (shade_one).
__device__ __forceinline__ float pdf_bsdf_sampled(const Mat& m, v3f wo, v3f wi, v3f n) {
    float nl = dot(n, wi);
    if (nl <= 0.0f) return 0.0f;
    float sp = spec_prob(m);
    float r = smax(m.roughness, MIN_ROUGH);
    v3f h = make_dir_or(wo + wi, mk(0.0f, 0.0f, 1.0f));
    float nh = dot(n, h), vh = dot(wo, h), nv = dot(n, wo);
    float pv = 0.0f;
    if (nh > 0.0f && vh > 0.0f && nv > 0.0f) {
        float al = r * r, a2 = al * al;
        float den = nh * nh * (a2 - 1.0f) + 1.0f;
        pv = a2 * INV_PI_F / (den * den) * ggx_g1(nv, r) / (4.0f * nv);
    }
    return (sp * pv + (1.0f - sp) * pdf_cosine(nl)) * (1.0f - m.transmission);
}
// bsdf_sample_weight, render_units.hpp:1413-1420
__device__ __forceinline__ v3f bsdf_weight(v3f f, float c, float pdf) {
    if (pdf < MIN_PDF) return mk(0.0f, 0.0f, 0.0f);
    return mk((f.x / pdf) * c, (f.y / pdf) * c, (f.z / pdf) * c);
}
// mis_power_heuristic, render_units.hpp:1427-1432
__device__ __forceinline__ float power_heuristic(float pf, float pg) {
    float f2 = pf * pf, g2 = pg * pg;
    return f2 / (f2 + g2 + 1e-10f);
}

struct BSample {
    v3f wi;
    v3f fw;       // f (evaluated) or the precomputed weight
    float pdf;
    int kind;     // 0 evaluated, 1 precomputed weight
    int type;     // 0 diffuse 1 specular 2 transmission
};
// sampleBSDF, bsdf/bsdf.hpp:198-345 (u1..u3 drawn left to right by the caller;
// the glass branch draws one more number, bsdf.hpp:221)
__device__ __forceinline__ BSample sample_bsdf(const Mat& m, v3f wo, v3f n, float u1, float u2, float u3, Pcg& rng) {
    BSample s;
    s.wi = mk(0.0f, 0.0f, 1.0f);
    s.fw = mk(0.0f, 0.0f, 0.0f);
    s.pdf = 0.0f;
    s.kind = 0;
    s.type = 0;
    float r = smax(m.roughness, MIN_ROUGH);
    if (m.transmission > 0.0f && u3 < m.transmission) {
        bool front = dot(wo, n) > 0.0f;
        v3f fnrm = front ? n : -n;
        float eta = front ? (1.0f / m.ior) : m.ior;
        v3f t, b;
        onb(fnrm, t, b);
        v3f h = sample_vndf(wo, r, u1, u2, fnrm, t, b);
        float F = fresnel_dielectric(fabsf(dot(wo, h)), eta);
        v3f inc = -wo;
        bool refl = randf(rng) < F;
        v3f rf;
        if (!refl && refract(inc, h, eta, rf)) {
            s.wi = rf;
            float nl = fabsf(dot(s.wi, fnrm)), nv = fabsf(dot(wo, fnrm));
            float w = ggx_g2(nl, nv, r) / (ggx_g1(nv, r) + GGX_EPS);
            s.kind = 1; s.fw = mk(w, w, w); s.type = 2;
            return s;
        }
        // Fresnel reflection, or total internal reflection
        s.wi = reflect(inc, h);
        float nl = dot(fnrm, s.wi), nv = dot(fnrm, wo);
        if (nl <= 0.0f || nv <= 0.0f) return s;  // EvaluatedBSDF{0, 0}
        float w = ggx_g2(nl, nv, r) / (ggx_g1(nv, r) + GGX_EPS);
        s.kind = 1; s.fw = mk(w, w, w); s.type = 1;
        return s;
    }
    float sp = spec_prob(m);
    v3f t, b;
    onb(n, t, b);
    if (u1 < sp) {
        v3f h = sample_vndf(wo, r, u1 / sp, u2, n, t, b);
        s.wi = reflect(-wo, h);
        float nl = dot(n, s.wi);
        s.type = 1;
        if (nl <= 0.0f) return s;
        v3f f = eval_bsdf(m, wo, s.wi, n);
        float pdf = (sp * pdf_vndf(wo, h, r, n) + (1.0f - sp) * pdf_cosine(nl)) * (1.0f - m.transmission);
        if (r < 0.05f && m.metallic > 0.5f) {
            s.kind = 1; s.fw = bsdf_weight(f, nl, pdf);
        } else {
            s.fw = f; s.pdf = pdf;
        }
    } else {
        float u1a = (u1 - sp) / (1.0f - sp);
        s.wi = sample_cosine(u1a, u2, n);
        float nl = dot(n, s.wi);
        if (nl <= 0.0f) return s;
        v3f f = eval_bsdf(m, wo, s.wi, n);
        v3f h = make_dir_or(wo + s.wi, mk(0.0f, 0.0f, 1.0f));
        float pdf = (sp * pdf_vndf(wo, h, r, n) + (1.0f - sp) * pdf_cosine(nl)) * (1.0f - m.transmission);
        s.fw = f; s.pdf = pdf;
    }
    return s;
}

// ---- lights: light/light.hpp:146-358, render_units.hpp:1393-1448 ----
__device__ __forceinline__ float pdf_w_from_area(float dist, float area, float c) {
    if (area < 1e-6f || c < 1e-6f) return 0.0f;
    float d2 = dist * dist;
    return (d2 / area) / c;
}
__device__ __forceinline__ float inv_sq(float d) {
    if (d < 1e-6f) return 0.0f;
    float d2 = d * d;
    return 1.0f / d2;
}
struct LSample {
    v3f pos, dir, emission;
    float pdf, dist;
};
__device__ __forceinline__ LSample sample_light(const GLight& l, v3f p, float u1, float u2) {
    LSample s;
    s.emission = mk(l.emission[0], l.emission[1], l.emission[2]);
    s.pdf = 0.0f;
    v3f nrm = mk(l.normal[0], l.normal[1], l.normal[2]);
    switch (l.type) {
        case 1: {  // POINT
            v3f c = mk(l.pos[0], l.pos[1], l.pos[2]);
            v3f sn = mk(0.0f, 1.0f, 0.0f);
            if (l.radius > 1e-6f) {
                float z = 1.0f - 2.0f * u1;
                float r = sqrtf(smax(0.0f, 1.0f - z * z));
                float phi = 2.0f * PI_F * u2;
                v3f od = mk(r * cosf(phi), r * sinf(phi), z);
                s.pos = c + od * l.radius;
                sn = od;
            } else {
                s.pos = c;
            }
            normalize_len(s.pos - p, s.dir, s.dist);
            if (l.radius < 1e-6f) {
                float f = inv_sq(s.dist);
                s.emission = mk(s.emission.x * f, s.emission.y * f, s.emission.z * f);
                s.pdf = 1.0f;
            } else {
                float cl = dot(sn, -s.dir);
                s.pdf = (cl < 1e-6f) ? 0.0f : pdf_w_from_area(s.dist, l.area, cl);
            }
            break;
        }
        case 2: {  // SUN
            s.dir = -nrm;
            s.pos = p + s.dir * 1000000.0f;
            s.dist = 1000000.0f;
            s.pdf = 1.0f;
            break;
        }
        case 3: {  // SPOT
            s.pos = mk(l.pos[0], l.pos[1], l.pos[2]);
            normalize_len(s.pos - p, s.dir, s.dist);
            float f = inv_sq(s.dist);
            s.emission = mk(s.emission.x * f, s.emission.y * f, s.emission.z * f);
            float ca = dot(nrm, -s.dir);
            float cc = l.spot_cos;
            if (ca < cc) {
                s.emission = mk(0.0f, 0.0f, 0.0f);
            } else if (l.spot_blend > 0.0f) {
                float t = (ca - cc) / (1.0f - cc);
                float fo = smin(1.0f, t / l.spot_blend);
                s.emission = s.emission * fo;
            }
            s.pdf = 1.0f;
            break;
        }
        case 4: {  // AREA
            float lx, ly;
            if (l.shape == 2 || l.shape == 3) {
                float rs = sqrtf(u1);
                float th = 2.0f * PI_F * u2;
                lx = l.sx * 0.5f * rs * cosf(th);
                ly = l.sy * 0.5f * rs * sinf(th);
            } else {
                lx = l.sx * (u1 - 0.5f);
                ly = l.sy * (u2 - 0.5f);
            }
            s.pos = mk(l.pos[0], l.pos[1], l.pos[2]) + mk(l.right[0], l.right[1], l.right[2]) * lx +
                    mk(l.up[0], l.up[1], l.up[2]) * ly;
            normalize_len(s.pos - p, s.dir, s.dist);
            float cl = dot(nrm, -s.dir);
            s.pdf = (cl < 1e-6f) ? 0.0f : pdf_w_from_area(s.dist, l.area, cl);
            break;
        }
        default: {  // emissive triangle: sampleTrianglePoint (light.hpp:146-165)
            v3f v0 = mk(l.v0[0], l.v0[1], l.v0[2]);
            float su1 = sqrtf(u1);
            float b0 = 1.0f - su1;
            float b1 = u2 * su1;
            float b2 = 1.0f - b0 - b1;
            v3f d1, d2;
            float l1, l2;
            normalize_len(mk(l.v1[0], l.v1[1], l.v1[2]) - v0, d1, l1);
            normalize_len(mk(l.v2[0], l.v2[1], l.v2[2]) - v0, d2, l2);
            s.pos = v0 + d1 * (l1 * b1) + d2 * (l2 * b2);
            normalize_len(s.pos - p, s.dir, s.dist);
            float cl = fabsf(dot(nrm, -s.dir));
            s.pdf = (cl < 1e-6f) ? 0.0f : pdf_w_from_area(s.dist, l.area, cl);
            break;
        }
    }
    return s;
}
// pdfLightSampleTyped, light.hpp:344-358
__device__ __forceinline__ float pdf_light(float area, v3f p, v3f lp, v3f ln) {
    v3f d;
    float dist;
    normalize_len(lp - p, d, dist);
    float cl = fabsf(dot(ln, -d));
    if (cl < 1e-6f || area < 1e-6f) return 0.0f;
    return pdf_w_from_area(dist, area, cl);
}
__device__ __forceinline__ float area_ratio(float a, float total) { return total < 1e-6f ? 0.0f : a / total; }

// ---- volumes: volume/phase.hpp:32-77, volume/transmittance.hpp:35-352 ----
__device__ __forceinline__ float hg_eval(float g, float c) {
    float gg = g * g;
    float den = 1.0f + gg - 2.0f * g * c;
    float d = smax(den, 1e-8f);
    const float INV4PI = 1.0f / (4.0f * PI_F);
    return INV4PI * (1.0f - gg) / (d * sqrtf(d));
}
__device__ __forceinline__ v3f hg_sample(float g, v3f wo, float u1, float u2, float& pdf) {
    float ct;
    if (fabsf(g) < 1e-3f) {
        ct = 1.0f - 2.0f * u1;
    } else {
        float gg = g * g;
        float sq = (1.0f - gg) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f + gg - sq * sq) / (2.0f * g);
        ct = (ct < -1.0f) ? -1.0f : ((1.0f < ct) ? 1.0f : ct);
    }
    float st = sqrtf(smax(0.0f, 1.0f - ct * ct));
    float phi = 2.0f * PI_F * u2;
    v3f t, b;
    onb(wo, t, b);
    float lx = st * cosf(phi), ly = st * sinf(phi), lz = ct;
    pdf = hg_eval(g, ct);
    return mk(t.x * lx + b.x * ly + wo.x * lz, t.y * lx + b.y * ly + wo.y * lz, t.z * lx + b.z * ly + wo.z * lz);
}
__device__ __forceinline__ bool ray_aabb_segment(v3f o, v3f d, float tmax, const float* bmin, const float* bmax,
                                                 float& tin, float& tout) {
    float ro[3] = {o.x, o.y, o.z}, rd[3] = {d.x, d.y, d.z};
    float t0 = 0.0f, t1 = tmax;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (fabsf(rd[a]) < 1e-12f) {
            if (ro[a] < bmin[a] || ro[a] > bmax[a]) return false;
            continue;
        }
        float inv = 1.0f / rd[a];
        float ta = (bmin[a] - ro[a]) * inv;
        float tb = (bmax[a] - ro[a]) * inv;
        if (ta > tb) { float x = ta; ta = tb; tb = x; }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (t0 > t1) return false;
    }
    tin = smax(0.0f, t0);
    tout = t1;
    return tout > tin;
}
// luminance of an RGB radiance (units/render_units.hpp:1192-1194)
__device__ __forceinline__ float luminance(v3f c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// ---- node-tree programs (GNodeOp; the host compiler folds everything
// but Checker textures, node_evaluator.cpp:352-401) ----

// Low bit of floor(x): the reference's floor(u * scale) % 2 with its
// negative fix-up.  |floor(x)| >= 2^24 is even, and NaN reads as even as the
// reference's x86 float-to-int conversion (INT_MIN) makes it.
__device__ __forceinline__ int checker_parity(float x) {
    const float f = floorf(x);
    return (fabsf(f) < 16777216.0f) ? ((int)f & 1) : 0;
}

// Evaluation slots are registers: every access is a per-component select
// on the slot index, never a dynamically indexed (scratch) array.
__device__ __forceinline__ v3f pick(bool c, v3f a, v3f b) {
    return mk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
struct NodeStack {
    v3f s[kNodeStack];
    __device__ __forceinline__ v3f get(int i) const {
        v3f r = s[0];
#pragma unroll
        for (int k = 1; k < kNodeStack; ++k) r = pick(i == k, s[k], r);
        return r;
    }
    __device__ __forceinline__ void put(int i, v3f v) {
#pragma unroll
        for (int k = 0; k < kNodeStack; ++k) s[k] = pick(i == k, v, s[k]);
    }
};

// The RGB a program leaves at NOP_END, at texture coordinate (u, v).
__device__ __forceinline__ v3f run_node_program(const GNodeOp* ops, float u, float v) {
    NodeStack st;
#pragma unroll
    for (int k = 0; k < kNodeStack; ++k) st.s[k] = mk(0.0f, 0.0f, 0.0f);
    int sp = 0;
    for (;; ++ops) {
        const GNodeOp o = *ops;
        const int code = o.op & 0xff;
        if (code == NOP_END) break;
        if (code == NOP_CONST) {
            st.put(sp, mk(o.a, o.b, o.c));
            ++sp;
        } else if (code == NOP_SCALE) {
            st.put(sp - 1, st.get(sp - 1) * o.a);
        } else {
            const bool sw = (o.op & kNopSwapped) != 0;
            const v3f top = st.get(sp - 1), below = st.get(sp - 2);
            const v3f x = pick(sw, top, below);  // color1 / A
            const v3f y = pick(sw, below, top);  // color2 / B
            v3f r;
            if (code == NOP_CHECKER) {
                r = pick(checker_parity(u * o.a) == checker_parity(v * o.a), x, y);
            } else {
                const float f1 = 1.0f - o.a;
                r = mk(x.x * f1 + y.x * o.a, x.y * f1 + y.y * o.a, x.z * f1 + y.z * o.a);
            }
            --sp;
            st.put(sp - 1, r);
        }
    }
    return st.s[0];
}

__device__ __forceinline__ float grid_f32(float x) { return x; }
__device__ __forceinline__ float grid_f32(__half x) { return __half2float(x); }

// Density grids are stored as FP32 (the reference's values) or FP16 (half
// the bytes; computed in FP32 after widening, which is exact).
template <class G>
__device__ __forceinline__ float grid_trilinear(const GVolume& v, const G* grid, float wx, float wy, float wz) {
    float u = (wx - v.wmin[0]) / smax(1e-12f, v.wmax[0] - v.wmin[0]);
    float vv = (wy - v.wmin[1]) / smax(1e-12f, v.wmax[1] - v.wmin[1]);
    float w = (wz - v.wmin[2]) / smax(1e-12f, v.wmax[2] - v.wmin[2]);
    if (u < 0.0f || u > 1.0f || vv < 0.0f || vv > 1.0f || w < 0.0f || w > 1.0f) return 0.0f;
    int nx = v.dims[0], ny = v.dims[1], nz = v.dims[2];
    float fx = u * (float)nx - 0.5f, fy = vv * (float)ny - 0.5f, fz = w * (float)nz - 0.5f;
    if (fx < 0.0f || fx > (float)(nx - 1) || fy < 0.0f || fy > (float)(ny - 1) || fz < 0.0f || fz > (float)(nz - 1))
        return 0.0f;
    int ix = min(max((int)fx, 0), nx - 2);
    int iy = min(max((int)fy, 0), ny - 2);
    int iz = min(max((int)fz, 0), nz - 2);
    float tx = fx - (float)ix, ty = fy - (float)iy, tz = fz - (float)iz;
    const G* g = grid + v.grid_offset;
    size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    size_t b0 = (size_t)ix + sy * iy + sz * iz;
    float c000 = grid_f32(g[b0]), c100 = grid_f32(g[b0 + sx]);
    float c010 = grid_f32(g[b0 + sy]), c110 = grid_f32(g[b0 + sx + sy]);
    float c001 = grid_f32(g[b0 + sz]), c101 = grid_f32(g[b0 + sx + sz]);
    float c011 = grid_f32(g[b0 + sy + sz]), c111 = grid_f32(g[b0 + sx + sy + sz]);
    float c00 = c000 * (1.0f - tx) + c100 * tx;
    float c10 = c010 * (1.0f - tx) + c110 * tx;
    float c01 = c001 * (1.0f - tx) + c101 * tx;
    float c11 = c011 * (1.0f - tx) + c111 * tx;
    float c0 = c00 * (1.0f - ty) + c10 * ty;
    float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}

}  // namespace lucid
