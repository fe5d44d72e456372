// diag_host.h -- host side of the DiagnosticFilm API (lucidrenderer module).
//
// The reference's diagnostics classes as the Python module exposes them
// (src/pybind_module.cpp:316-541): PathRecordingConfig
// (diagnostics/path_stats_config.hpp), PathTrace / PathDiagnosticRecorder
// (path_types.hpp, diagnostic_integrator.hpp:87-389), DiagnosticFilm
// (diagnostic_film.hpp), DiagnosticPathTracer (diagnostic_integrator.hpp:
// 391-465) and DiagnosticExporter with its export structures
// (diagnostic_export.hpp).  Per-pixel data is PixelPathData<16, 4> in the
// reference's byte layout (diag_film.h), recorded with the same code the
// device film kernel runs.  The queries work over any film source: a host
// DiagnosticFilm, or the device film of a Renderer streamed through the C
// ABI in chunks (FilmSource).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "diag_film.h"

namespace lucid {
namespace diag {

// PathRecordingConfig (path_stats_config.hpp:27-182)
struct RecordingConfig {
    uint8_t subsample_factor = 2;
    uint8_t max_groups_per_pixel = 16;
    uint8_t max_outliers_per_pixel = 4;
    uint8_t max_depth = 16;
    float variance_threshold = 0.001f;
    float outlier_threshold = 3.0f;

    static RecordingConfig minimal() {
        RecordingConfig c;
        c.subsample_factor = 4; c.max_groups_per_pixel = 8; c.max_outliers_per_pixel = 2; c.max_depth = 8;
        c.variance_threshold = 0.01f; c.outlier_threshold = 5.0f;
        return c;
    }
    static RecordingConfig standard() { return RecordingConfig(); }
    static RecordingConfig detailed() {
        RecordingConfig c;
        c.subsample_factor = 1; c.max_groups_per_pixel = 32; c.max_outliers_per_pixel = 8; c.max_depth = 16;
        c.variance_threshold = 0.0001f; c.outlier_threshold = 2.0f;
        return c;
    }
    size_t sampled_pixel_count(size_t w, size_t h) const { return (w / subsample_factor) * (h / subsample_factor); }
    size_t per_pixel_data_size() const {
        return 8 * sizeof(Stats) + max_groups_per_pixel * sizeof(Group) + max_outliers_per_pixel * (8 + 12) + 32;
    }
    size_t estimate_memory_bytes(size_t w, size_t h) const {
        return sampled_pixel_count(w, h) * per_pixel_data_size() + 1024 * 1024;
    }
};

// Names (path_types.hpp:38-157, path_stats.hpp:26-41)
inline char bsdf_short(uint8_t t) {
    static const char c[] = {'D', 'G', 'M', 'T', 'L', 'E'};
    return t < 6 ? c[t] : '?';
}
inline char bsdf_heckbert(uint8_t t) {
    static const char c[] = {'D', 'S', 'S', 'S', 'E', 'E'};
    return t < 6 ? c[t] : '?';
}
inline const char* light_heckbert(uint8_t t) {
    switch (t) {
        case kLightPoint: return "LSD";
        case kLightArea: return "LDD";
        case kLightDirectional: return "LSD";
        case kLightSpot: return "LSD";
        case kLightEnvironment: return "LDE";
        default: return "L??";  // Unknown, and Emissive, which the reference does not list
    }
}
inline const char* coarse_name(uint8_t t) {
    static const char* n[] = {"Direct-NoDelta-BSDF",   "Direct-NoDelta-NEE",   "Direct-Delta-BSDF",
                              "Direct-Delta-NEE",      "Indirect-NoDelta-BSDF", "Indirect-NoDelta-NEE",
                              "Indirect-Delta-BSDF",   "Indirect-Delta-NEE"};
    return n[t & 7];
}
inline const char* strategy_name(uint8_t s) {
    static const char* n[] = {"BSDF", "NEE", "MIS_BSDF", "MIS_NEE"};
    return s < 4 ? n[s] : "Unknown";
}
// PathGroup::signature / signature_heckbert (path_stats.hpp:196-242)
inline std::string signature(const Group& g) {
    std::string s;
    for (int i = 0; i < g.depth; ++i) s += bsdf_short(g.vertices[i].bsdf_type);
    return s;
}
inline std::string signature_heckbert(const Group& g) {
    std::string s = light_heckbert(g.light_type);
    int end = g.depth;
    while (end > 0 && (g.vertices[end - 1].bsdf_type == kEmission || g.vertices[end - 1].bsdf_type == kEnvironment))
        --end;
    if (end > 0) {
        s += ' ';
        for (int i = 0; i < end; ++i) {
            s += bsdf_heckbert(g.vertices[i].bsdf_type);
            if (i + 1 < end) s += ' ';
        }
    }
    s += " DSE";
    return s;
}

struct GlobalStats {  // GlobalDiagnosticStats (diagnostic_film.hpp:27-34)
    size_t total_samples = 0, active_pixels = 0, total_groups = 0, total_overflow = 0;
    float total_variance = 0.0f;  // never filled by the reference either
};
struct GroupInfo {  // ExportedGroupInfo (diagnostic_export.hpp:26-50)
    size_t pixel_x = 0, pixel_y = 0;
    std::string signature, signature_heckbert;
    std::vector<int32_t> object_ids;
    std::string object_path;
    uint32_t sample_count = 0;
    float mean_luminance = 0.0f, variance_luminance = 0.0f;
    std::vector<float> mean_rgb{0.0f, 0.0f, 0.0f}, variance_rgb{0.0f, 0.0f, 0.0f};
    uint8_t depth = 0, coarse_type = 0;
    std::string coarse_type_name;
    uint8_t strategy = kBSDF;
    std::string strategy_name;
    std::vector<std::vector<float>> positions, normals;
};
struct CoarseInfo {  // ExportedCoarseStats (:55-63)
    uint8_t type_id = 0;
    std::string name;
    uint32_t sample_count = 0;
    float mean_luminance = 0.0f, variance_luminance = 0.0f;
    std::vector<float> mean_rgb{0.0f, 0.0f, 0.0f}, variance_rgb{0.0f, 0.0f, 0.0f};
};
struct Suggestion {  // DiagnosticSuggestion (:68-73)
    std::string category, severity, message, action;
};

// Read access to a film: dimensions, config, and pixels in chunks.
struct FilmSource {
    size_t width = 0, height = 0, sampled_w = 0, sampled_h = 0, total_recorded = 0;
    RecordingConfig config;
    // visit(first, count, fn): calls fn(index, pixel) for sampled pixels
    // [first, first + count) in row-major order
    std::function<void(size_t, size_t, const std::function<void(size_t, const PixelData&)>&)> visit;
    // pixel_data(x, y) of DiagnosticFilm (:118-127): index or -1
    long long index_of(size_t x, size_t y) const {
        const size_t sx = x / config.subsample_factor, sy = y / config.subsample_factor;
        if (sx >= sampled_w || sy >= sampled_h) return -1;
        return (long long)(sy * sampled_w + sx);
    }
    void for_each(const std::function<void(size_t, const PixelData&)>& fn) const {
        const size_t n = sampled_w * sampled_h, chunk = 4096;
        for (size_t f = 0; f < n; f += chunk) visit(f, std::min(chunk, n - f), fn);
    }
};

// DiagnosticFilm::global_stats (:135-155)
inline GlobalStats global_stats(const FilmSource& f) {
    GlobalStats s;
    f.for_each([&](size_t, const PixelData& p) {
        if (p.total_samples > 0) {
            ++s.active_pixels;
            s.total_samples += p.total_samples;
            s.total_groups += p.num_groups;
            s.total_overflow += p.overflow_count;
        }
    });
    return s;
}

struct GroupRef {
    size_t px, py;
    Group g;
};
// top_groups_by_variance / _by_mean (:160-224): groups with count >= 2
// (variance) or >= 1 (mean) over the film in row-major order, top n.
inline std::vector<GroupRef> top_groups(const FilmSource& f, size_t n, bool by_mean) {
    std::vector<GroupRef> all;
    const unsigned min_count = by_mean ? 1 : 2;
    auto key = [by_mean](const Group& g) { return by_mean ? mean_luminance(g.stats) : total_variance(g.stats); };
    f.for_each([&](size_t i, const PixelData& p) {
        const size_t sx = i % f.sampled_w, sy = i / f.sampled_w;
        for (uint64_t k = 0; k < p.num_groups && k < (uint64_t)kMaxGroups; ++k) {
            if (p.groups[k].stats.count >= min_count)
                all.push_back({sx * f.config.subsample_factor, sy * f.config.subsample_factor, p.groups[k]});
        }
        // keep memory bounded on large films: trim to the best n now and then
        if (all.size() > 4 * n + 65536) {
            std::nth_element(all.begin(), all.begin() + (long)n, all.end(),
                             [&](const GroupRef& a, const GroupRef& b) { return key(a.g) > key(b.g); });
            all.resize(n);
        }
    });
    const size_t lim = std::min(n, all.size());
    std::partial_sort(all.begin(), all.begin() + (long)lim, all.end(),
                      [&](const GroupRef& a, const GroupRef& b) { return key(a.g) > key(b.g); });
    all.resize(lim);
    return all;
}
// DiagnosticExporter::convert_refs (diagnostic_export.hpp:254-284)
inline std::vector<GroupInfo> convert_refs(const std::vector<GroupRef>& refs) {
    std::vector<GroupInfo> out;
    for (const GroupRef& r : refs) {
        GroupInfo gi;
        gi.pixel_x = r.px; gi.pixel_y = r.py;
        gi.signature = signature(r.g);
        gi.sample_count = r.g.stats.count;
        gi.mean_luminance = mean_luminance(r.g.stats);
        gi.variance_luminance = variance_luminance(r.g.stats);
        gi.mean_rgb = {r.g.stats.mean.x, r.g.stats.mean.y, r.g.stats.mean.z};
        gi.variance_rgb = {r.g.stats.variance.x, r.g.stats.variance.y, r.g.stats.variance.z};
        gi.depth = r.g.depth;
        gi.coarse_type = r.g.coarse_type;
        gi.coarse_type_name = coarse_name(r.g.coarse_type);
        out.push_back(gi);
    }
    return out;
}
// DiagnosticExporter::get_coarse_stats (:109-155), including its fixed
// factor-2 pixel lookup.
inline std::vector<CoarseInfo> coarse_stats(const FilmSource& f) {
    Stats agg[8];
    std::memset(agg, 0, sizeof(agg));
    std::vector<long long> want;
    for (size_t sy = 0; sy < f.sampled_h; ++sy)
        for (size_t sx = 0; sx < f.sampled_w; ++sx) want.push_back(f.index_of(sx * 2, sy * 2));
    std::vector<PixelData> cache;
    // pixels are visited in want-order; fetch each needed index once
    std::vector<long long> order;
    for (long long w : want) if (w >= 0) order.push_back(w);
    for (long long idx : order) {
        f.visit((size_t)idx, 1, [&](size_t, const PixelData& p) {
            for (int t = 0; t < 8; ++t) {
                const Stats& cs = p.coarse[t];
                if (cs.count == 0) continue;
                agg[t].count += cs.count;
                const float w = (float)cs.count / (float)agg[t].count;
                agg[t].mean.x = agg[t].mean.x * (1.0f - w) + cs.mean.x * w;
                agg[t].mean.y = agg[t].mean.y * (1.0f - w) + cs.mean.y * w;
                agg[t].mean.z = agg[t].mean.z * (1.0f - w) + cs.mean.z * w;
                agg[t].variance.x = std::max(agg[t].variance.x, cs.variance.x);
                agg[t].variance.y = std::max(agg[t].variance.y, cs.variance.y);
                agg[t].variance.z = std::max(agg[t].variance.z, cs.variance.z);
            }
        });
    }
    std::vector<CoarseInfo> out;
    for (int t = 0; t < 8; ++t) {
        CoarseInfo c;
        c.type_id = (uint8_t)t;
        c.name = coarse_name((uint8_t)t);
        c.sample_count = agg[t].count;
        c.mean_luminance = mean_luminance(agg[t]);
        c.variance_luminance = variance_luminance(agg[t]);
        c.mean_rgb = {agg[t].mean.x, agg[t].mean.y, agg[t].mean.z};
        c.variance_rgb = {agg[t].variance.x, agg[t].variance.y, agg[t].variance.z};
        out.push_back(c);
    }
    return out;
}
// DiagnosticFilm::metadata_json (:230-244)
inline std::string metadata_json(const FilmSource& f) {
    std::ostringstream o;
    o << "{\n";
    o << "  \"width\": " << f.width << ",\n";
    o << "  \"height\": " << f.height << ",\n";
    o << "  \"sampled_width\": " << f.sampled_w << ",\n";
    o << "  \"sampled_height\": " << f.sampled_h << ",\n";
    o << "  \"subsample_factor\": " << (int)f.config.subsample_factor << ",\n";
    o << "  \"max_groups\": " << (int)f.config.max_groups_per_pixel << ",\n";
    o << "  \"max_outliers\": " << (int)f.config.max_outliers_per_pixel << ",\n";
    o << "  \"total_samples\": " << f.total_recorded << "\n";
    o << "}\n";
    return o.str();
}
// DiagnosticFilm::export_binary "PVAR" v1 (:247-299)
inline std::string export_binary(const FilmSource& f) {
    std::string b;
    auto put = [&b](const void* p, size_t n) { b.append(static_cast<const char*>(p), n); };
    b += "PVAR";
    const uint32_t version = 1, w = (uint32_t)f.width, h = (uint32_t)f.height, sw = (uint32_t)f.sampled_w,
                   sh = (uint32_t)f.sampled_h;
    put(&version, 4); put(&w, 4); put(&h, 4); put(&sw, 4); put(&sh, 4);
    const uint8_t cfg[4] = {f.config.subsample_factor, f.config.max_groups_per_pixel, f.config.max_outliers_per_pixel,
                            0};
    put(cfg, 4);
    f.for_each([&](size_t, const PixelData& p) {
        const uint16_t gc = (uint16_t)p.num_groups;
        const uint32_t ts = (uint32_t)p.total_samples;
        put(&gc, 2); put(&ts, 4);
        put(&p.total.mean, 12);
        const float var = total_variance(p.total);
        put(&var, 4);
    });
    return b;
}
// DiagnosticExporter::generate_suggestions (:188-249)
inline std::vector<Suggestion> suggestions(const FilmSource& f) {
    std::vector<Suggestion> out;
    auto top = top_groups(f, 10, false);
    if (top.empty()) return out;
    size_t glossy = 0, diffuse = 0, indirect = 0;
    for (const GroupRef& r : top) {
        bool has_glossy = false, has_indirect = r.g.depth > 2;
        for (int i = 0; i < r.g.depth; ++i)
            if (r.g.vertices[i].bsdf_type == kGlossy) has_glossy = true;
        if (has_glossy) ++glossy;
        if (has_indirect) ++indirect;
        if (!has_glossy && !has_indirect) ++diffuse;
    }
    if (glossy > 3)
        out.push_back({"Settings", "warning", "High variance detected in glossy reflections",
                       "Consider increasing samples or using adaptive sampling for glossy materials"});
    if (indirect > 3)
        out.push_back({"Algorithm", "info", "Significant variance from indirect illumination",
                       "MIS algorithm recommended; consider path guiding for complex scenes"});
    if (diffuse > 3)
        out.push_back({"Scene", "info", "High variance in diffuse lighting",
                       "Check for small or distant light sources; consider area lights"});
    GlobalStats s = global_stats(f);
    if (s.total_samples > 0 && (double)s.total_overflow > (double)s.active_pixels * 0.1)
        out.push_back({"Settings", "warning", "Many path types exceeded tracking capacity",
                       "Consider using 'Detailed' preset for more accurate analysis"});
    return out;
}

// Host film (DiagnosticFilm): pixels allocated on first record.
class HostFilm {
public:
    HostFilm(size_t w, size_t h, const RecordingConfig& c) : w_(w), h_(h), cfg_(c) {
        const int sub = c.subsample_factor ? c.subsample_factor : 1;
        sw_ = w / sub;
        sh_ = h / sub;
    }
    size_t width() const { return w_; }
    size_t height() const { return h_; }
    size_t sampled_width() const { return sw_; }
    size_t sampled_height() const { return sh_; }
    size_t total_paths_recorded() const { return total_; }
    void record_path(size_t x, size_t y, const HostTrace& t) {
        const size_t sx = x / cfg_.subsample_factor, sy = y / cfg_.subsample_factor;
        if (sx >= sw_ || sy >= sh_) return;
        ensure();
        record_path_into(px_[sy * sw_ + sx], t);
        ++total_;
    }
    void clear() {
        px_.clear();
        total_ = 0;
    }
    FilmSource source() const {
        FilmSource f;
        f.width = w_; f.height = h_; f.sampled_w = sw_; f.sampled_h = sh_; f.total_recorded = total_;
        f.config = cfg_;
        f.visit = [this](size_t first, size_t count, const std::function<void(size_t, const PixelData&)>& fn) {
            PixelData def;
            pixel_init(def);
            for (size_t i = first; i < first + count; ++i) fn(i, px_.empty() ? def : px_[i]);
        };
        return f;
    }

private:
    size_t w_, h_, sw_, sh_, total_ = 0;
    RecordingConfig cfg_;
    std::vector<PixelData> px_;
    void ensure() {
        if (!px_.empty()) return;
        px_.resize(sw_ * sh_);
        for (PixelData& p : px_) pixel_init(p);
    }
    static void record_path_into(PixelData& p, const HostTrace& t) { ::lucid::diag::record_path(p, t); }
};

// PathDiagnosticRecorder (diagnostic_integrator.hpp:87-389), the parts the
// module binds.
class HostRecorder {
public:
    void begin_path() {
        cur_ = HostTrace();
    }
    void record_vertex(int32_t obj, int16_t mat, uint8_t type, bool delta, bool light_sampled) {
        if (cur_.n >= kMaxDepth) return;
        Vertex v{obj, mat, type, (uint8_t)((delta ? kFlagDelta : 0) | (light_sampled ? kFlagLightSampled : 0))};
        cur_.v[cur_.n] = v;
        cur_.pos[cur_.n] = F3{0.0f, 0.0f, 0.0f};
        cur_.nrm[cur_.n] = F3{0.0f, 0.0f, 0.0f};
        ++cur_.n;
    }
    void record_environment_hit() {
        if (cur_.n >= kMaxDepth) return;
        cur_.v[cur_.n++] = Vertex{-1, -1, kEnvironment, 0};
        cur_.light = kLightEnvironment;
    }
    void record_light_hit(int32_t light_id, uint8_t type) {
        if (cur_.n >= kMaxDepth) return;
        cur_.v[cur_.n++] = Vertex{-(light_id + 2), -1, kEmission, 0};
        cur_.light = type;
    }
    void record_emissive_hit(int32_t obj, int16_t mat) {
        if (cur_.n >= kMaxDepth) return;
        cur_.v[cur_.n++] = Vertex{obj, mat, kEmission, 0};
        cur_.light = kLightEmissive;
    }
    size_t current_depth() const { return (size_t)cur_.n; }
    HostTrace end_path(float r, float g, float b) {
        cur_.c = F3{r, g, b};
        return cur_;
    }

private:
    HostTrace cur_;
};

}  // namespace diag
}  // namespace lucid
