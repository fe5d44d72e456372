// diag_film.h -- DiagnosticFilm per-pixel path statistics, host and device.
//
// Byte layout of PixelPathData<16, 4> and its parts, as the reference
// defines them (LucidRenderer/cpp_renderer/include/diagnostics/):
//   PathVertex      8 B  {i32 object_id, i16 material_id, u8 bsdf_type, u8 flags}
//                        (path_types.hpp:178-243)
//   PathStatistics 40 B  {u32 count, RGB3f mean, m2, variance} (path_stats.hpp:51-98)
//   PathGroup     568 B  {u64 hash, stats, u8 depth, coarse, light, strategy,
//                        pad[4], vertices[16], positions[16], normals[16]}
//                        (path_stats.hpp:133-268)
//   PixelPathData 9576 B {groups[16], coarse[8], outliers[4] (24 B each),
//                        u64 num_groups, num_outliers, total_samples,
//                        overflow_count, total_stats} (pixel_path_data.hpp:38-227)
// so a film read back from the device is the reference's film byte for byte.
//
// record_path() is PixelPathData::record_path (pixel_path_data.hpp:53-86):
// float Welford totals and coarse-type stats, 3-sigma outlier ring after 10
// samples, FNV-1a group key over (depth, vertices), first free slot or
// eviction of the group with the least variance x log2(count + 1).  One
// thread owns a film pixel and applies its records in order, so the result
// is deterministic.  The same code runs on the host for the host-side
// DiagnosticFilm class and in the device film kernel.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LUCID_DIAG_HD __host__ __device__ __forceinline__
#else
#define LUCID_DIAG_HD inline
#endif

namespace lucid {
namespace diag {

constexpr int kMaxDepth = 16;    // MAX_PATH_DEPTH / MAX_GROUP_DEPTH
constexpr int kMaxGroups = 16;   // DiagnosticFilm::PixelData = PixelPathData<16, 4>
constexpr int kMaxOutliers = 4;

enum BsdfType : uint8_t { kDiffuse = 0, kGlossy = 1, kMirror = 2, kGlass = 3, kEmission = 4, kEnvironment = 5 };
enum LightSource : uint8_t { kLightUnknown = 0, kLightPoint = 1, kLightArea = 2, kLightDirectional = 3,
                             kLightSpot = 4, kLightEnvironment = 5, kLightEmissive = 6 };
enum Strategy : uint8_t { kBSDF = 0, kNEE = 1, kMisBSDF = 2, kMisNEE = 3 };
constexpr uint8_t kFlagDelta = 0x01, kFlagLightSampled = 0x02;

struct F3 {
    float x, y, z;
};
struct Vertex {
    int32_t object_id;
    int16_t material_id;
    uint8_t bsdf_type;
    uint8_t flags;
};
struct Stats {
    uint32_t count;
    F3 mean, m2, variance;
};
struct Group {
    uint64_t path_hash;
    Stats stats;
    uint8_t depth, coarse_type, light_type, strategy;
    uint8_t pad[4];
    Vertex vertices[kMaxDepth];
    F3 positions[kMaxDepth];
    F3 normals[kMaxDepth];
};
struct Outlier {
    uint64_t path_hash;
    F3 contribution;
    uint32_t pad;
};
struct PixelData {
    Group groups[kMaxGroups];
    Stats coarse[8];
    Outlier outliers[kMaxOutliers];
    uint64_t num_groups, num_outliers, total_samples, overflow_count;
    Stats total;
};
static_assert(sizeof(Vertex) == 8, "PathVertex is 8 bytes");
static_assert(sizeof(Stats) == 40, "PathStatistics is 40 bytes");
static_assert(sizeof(Group) == 568, "PathGroup is 568 bytes");
static_assert(sizeof(Outlier) == 24, "OutlierSample is 24 bytes");
static_assert(sizeof(PixelData) == 9576, "PixelPathData<16,4> is 9576 bytes");

// FP64 Welford per film pixel and sampling strategy: n, mean[3], M2[3].
struct StrategyWelford {
    double n, mean[3], m2[3];
};
static_assert(sizeof(StrategyWelford) == 56, "56 bytes per strategy");

// record_path and its helpers take the completed path through an accessor V
// with depth(), vertex(k), position(k), normal(k), contribution(),
// light_type() and strategy().
// PathStatistics::add_sample (path_stats.hpp:63-98)
LUCID_DIAG_HD void stats_add(Stats& s, float r, float g, float b) {
    ++s.count;
    if (s.count == 1) {
        s.mean = F3{r, g, b};
        s.m2 = F3{0.0f, 0.0f, 0.0f};
        s.variance = F3{0.0f, 0.0f, 0.0f};
        return;
    }
    const float dr = r - s.mean.x, dg = g - s.mean.y, db = b - s.mean.z;
    const float n = (float)s.count;
    s.mean.x += dr / n;
    s.mean.y += dg / n;
    s.mean.z += db / n;
    const float er = r - s.mean.x, eg = g - s.mean.y, eb = b - s.mean.z;
    s.m2.x += dr * er;
    s.m2.y += dg * eg;
    s.m2.z += db * eb;
    const float nm1 = (float)(s.count - 1);
    s.variance.x = s.m2.x / nm1;
    s.variance.y = s.m2.y / nm1;
    s.variance.z = s.m2.z / nm1;
}
LUCID_DIAG_HD float total_variance(const Stats& s) { return s.variance.x + s.variance.y + s.variance.z; }
LUCID_DIAG_HD float mean_luminance(const Stats& s) {
    return 0.2126f * s.mean.x + 0.7152f * s.mean.y + 0.0722f * s.mean.z;
}
LUCID_DIAG_HD float variance_luminance(const Stats& s) {
    return 0.2126f * 0.2126f * s.variance.x + 0.7152f * 0.7152f * s.variance.y + 0.0722f * 0.0722f * s.variance.z;
}

// PathTrace::hash (path_types.hpp:321-343): FNV-1a, one 64-bit XOR per
// field, signed fields widened with their sign.
template <class V>
LUCID_DIAG_HD uint64_t path_hash(const V& t) {
    uint64_t h = 14695981039346656037ULL;
    const int d = t.depth();
    h ^= (uint64_t)d;
    h *= 1099511628211ULL;
    for (int i = 0; i < d; ++i) {
        const Vertex v = t.vertex(i);
        h ^= (uint64_t)(int64_t)v.object_id;
        h *= 1099511628211ULL;
        h ^= (uint64_t)(int64_t)v.material_id;
        h *= 1099511628211ULL;
        h ^= (uint64_t)v.bsdf_type;
        h *= 1099511628211ULL;
        h ^= (uint64_t)v.flags;
        h *= 1099511628211ULL;
    }
    return h;
}
// PathTrace::coarse_type (path_types.hpp:354-372)
template <class V>
LUCID_DIAG_HD uint8_t coarse_type(const V& t) {
    uint8_t c = 0;
    const int d = t.depth();
    for (int i = 0; i < d; ++i) {
        const uint8_t f = t.vertex(i).flags;
        if (f & kFlagLightSampled) c |= 0x01;
        if (f & kFlagDelta) c |= 0x02;
    }
    if (d > 2) c |= 0x04;
    return c;
}

// PathGroup::init_from (path_stats.hpp:160-176): only the first `depth`
// vertices are written.
template <class V>
LUCID_DIAG_HD void group_init(Group& g, const V& t, uint64_t h, uint8_t coarse) {
    g.path_hash = h;
    const int d = t.depth() < kMaxDepth ? t.depth() : kMaxDepth;
    g.depth = (uint8_t)d;
    g.coarse_type = coarse;
    g.light_type = t.light_type();
    g.strategy = t.strategy();
    for (int i = 0; i < d; ++i) {
        g.vertices[i] = t.vertex(i);
        g.positions[i] = t.position(i);
        g.normals[i] = t.normal(i);
    }
    g.stats.count = 0;
    g.stats.mean = F3{0.0f, 0.0f, 0.0f};
    g.stats.m2 = F3{0.0f, 0.0f, 0.0f};
    g.stats.variance = F3{0.0f, 0.0f, 0.0f};
    const F3 c = t.contribution();
    stats_add(g.stats, c.x, c.y, c.z);
}

// PixelPathData() default state: zero statistics, default PathVertex
// (object -1, material -1, Diffuse, no flags) in every group slot.
LUCID_DIAG_HD void pixel_init(PixelData& p) {
    uint8_t* b = reinterpret_cast<uint8_t*>(&p);
    for (unsigned i = 0; i < sizeof(PixelData); ++i) b[i] = 0;
    for (int gi = 0; gi < kMaxGroups; ++gi)
        for (int k = 0; k < kMaxDepth; ++k) {
            p.groups[gi].vertices[k].object_id = -1;
            p.groups[gi].vertices[k].material_id = -1;
        }
}

LUCID_DIAG_HD float diag_log2(float x) { return log2f(x); }
LUCID_DIAG_HD float diag_sqrt(float x) { return sqrtf(x); }

// PixelPathData::record_path (pixel_path_data.hpp:53-86) with is_outlier
// (:172-189), add_outlier (:194-198) and evict_and_add (:203-224).
template <class V>
LUCID_DIAG_HD void record_path(PixelData& p, const V& t) {
    ++p.total_samples;
    const F3 c = t.contribution();
    stats_add(p.total, c.x, c.y, c.z);
    const uint8_t coarse = coarse_type(t);
    stats_add(p.coarse[coarse], c.x, c.y, c.z);
    const uint64_t h = path_hash(t);
    if (p.total_samples >= 10) {
        const float sr = diag_sqrt(p.total.variance.x + 1e-10f);
        const float sg = diag_sqrt(p.total.variance.y + 1e-10f);
        const float sb = diag_sqrt(p.total.variance.z + 1e-10f);
        const bool out = fabsf(c.x - p.total.mean.x) > 3.0f * sr || fabsf(c.y - p.total.mean.y) > 3.0f * sg ||
                         fabsf(c.z - p.total.mean.z) > 3.0f * sb;
        if (out) {
            Outlier& o = p.outliers[p.num_outliers % kMaxOutliers];
            o.path_hash = h;
            o.contribution = c;
            ++p.num_outliers;
        }
    }
    for (uint64_t i = 0; i < p.num_groups; ++i) {
        if (p.groups[i].path_hash == h) {
            stats_add(p.groups[i].stats, c.x, c.y, c.z);
            return;
        }
    }
    if (p.num_groups < (uint64_t)kMaxGroups) {
        group_init(p.groups[p.num_groups], t, h, coarse);
        ++p.num_groups;
        return;
    }
    uint64_t mi = 0;
    float mv = total_variance(p.groups[0].stats);
    for (uint64_t i = 1; i < p.num_groups; ++i) {
        const float v = total_variance(p.groups[i].stats);
        const float adj = v * diag_log2((float)(p.groups[i].stats.count + 1));
        const float madj = mv * diag_log2((float)(p.groups[mi].stats.count + 1));
        if (adj < madj) {
            mi = i;
            mv = v;
        }
    }
    ++p.overflow_count;
    group_init(p.groups[mi], t, h, coarse);
}

// FP64 Welford update of one strategy sum.
LUCID_DIAG_HD void welford_add(StrategyWelford& w, double r, double g, double b) {
    w.n += 1.0;
    const double x[3] = {r, g, b};
    for (int k = 0; k < 3; ++k) {
        const double d = x[k] - w.mean[k];
        w.mean[k] += d / w.n;
        w.m2[k] += d * (x[k] - w.mean[k]);
    }
}

// A path held in host memory (the host-side PathTrace class).
struct HostTrace {
    int n = 0;
    Vertex v[kMaxDepth];
    F3 pos[kMaxDepth], nrm[kMaxDepth];
    F3 c{0.0f, 0.0f, 0.0f};
    uint8_t light = kLightUnknown, strat = kBSDF;
    LUCID_DIAG_HD int depth() const { return n; }
    LUCID_DIAG_HD Vertex vertex(int i) const { return v[i]; }
    LUCID_DIAG_HD F3 position(int i) const { return pos[i]; }
    LUCID_DIAG_HD F3 normal(int i) const { return nrm[i]; }
    LUCID_DIAG_HD F3 contribution() const { return c; }
    LUCID_DIAG_HD uint8_t light_type() const { return light; }
    LUCID_DIAG_HD uint8_t strategy() const { return strat; }
};

}  // namespace diag
}  // namespace lucid
