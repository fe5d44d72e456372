// lucid_oracle.cpp -- CPU restatement of Lucid's path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
// smoke check and the CPU-baseline leg of the benchmark load it.  The product
// (liblucid_hip.so / lucidrenderer) never links, loads or calls it.
//
// It restates, operation by operation and in plain float arithmetic, the
// reference CPU tracer of daikiad/blender-lucid-renderer.  Every function
// cites the reference file:line it follows; paths are relative to
// LucidRenderer/cpp_renderer/.  The reference types its maths with mp-units,
// whose wrappers are size-identical float wrappers
// (include/units/render_units.hpp:1513-1524), so every operation reduces to
// the float expressions below.  Evaluation order of every expression follows
// the reference source left to right, and random numbers are drawn left to
// right (Apple Clang, the reference's compiler, evaluates call arguments in
// that order).
//
// Build (parity build): g++ -O2 -fno-fast-math -ffp-contract=off -fopenmp.
// The baseline build adds -O3 -ffast-math, like the reference module flags.
//
// Parity status of each part is listed in the header of oracle.py.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- vectors --
// render_units.hpp:56-164 (Vec3f, dot, cross, normalize)
struct V3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
static inline V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 mul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 divs(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V3 cmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// std::min / std::max semantics (first argument returned on ties / NaN)
static inline float fmin_std(float a, float b) { return (b < a) ? b : a; }
static inline float fmax_std(float a, float b) { return (a < b) ? b : a; }

static const float kDirEps2 = 1e-6f * 1e-6f;  // render_units.hpp:461,485-486

// make_direction / make_normal: render_units.hpp:498-507, 714-723
static inline bool make_dir(V3 v, V3& out) {
    float l2 = len2(v);
    if (l2 < kDirEps2) return false;
    out = divs(v, std::sqrt(l2));
    return true;
}
static inline V3 make_dir_or(V3 v, V3 fb) {
    V3 d;
    return make_dir(v, d) ? d : fb;
}
// render::normalize (render_units.hpp:157-163)
static inline V3 normalize_plain(V3 v) {
    float l = std::sqrt(len2(v));
    if (l > 0.0f) return divs(v, l);
    return v;
}
// normalize_with_length (render_units.hpp:653-658)
static inline void normalize_with_length(V3 v, V3& dir, float& len) {
    len = std::sqrt(len2(v));
    dir = divs(v, len);
}
// reflect / refract (render_units.hpp:836-858)
static inline V3 reflect(V3 i, V3 n) {
    float c = dot(i, n);
    V3 r = sub(i, mul(n, 2.0f * c));
    return make_dir_or(r, v3(0, 0, 1));
}
static inline bool refract(V3 i, V3 n, float eta, V3& out) {
    float cos_i = -dot(i, n);
    float sin2_t = eta * eta * (1.0f - cos_i * cos_i);
    if (sin2_t > 1.0f) return false;
    float cos_t = std::sqrt(1.0f - sin2_t);
    V3 r = add(mul(i, eta), mul(n, eta * cos_i - cos_t));
    return make_dir(r, out);
}

// -------------------------------------------------------------------- RNG --
// math/random.hpp:19-120
struct Rng {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;
};
static inline uint32_t murmur3_fmix(uint32_t h) {  // random.hpp:32-39
    h ^= h >> 16; h *= 0x85ebca6bU; h ^= h >> 13; h *= 0xc2b2ae35U; h ^= h >> 16;
    return h;
}
static inline uint64_t splitmix64(uint64_t x) {  // random.hpp:42-47
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
static inline uint32_t jenkins32(uint32_t a) {  // random.hpp:50-58
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    return a;
}
static const uint64_t PCG_MULT = 6364136223846793005ULL;
static inline void seed_xyz(Rng& s, uint32_t x, uint32_t y, uint32_t si) {  // random.hpp:64-91
    uint32_t h1 = jenkins32(x + 1u);
    uint32_t h2 = jenkins32(y + 1u + h1);
    uint32_t h3 = jenkins32(si + 1u + h2);
    uint64_t seed1 = ((uint64_t)murmur3_fmix(h1 ^ h3) << 32) | murmur3_fmix(h2);
    uint64_t seed2 = ((uint64_t)murmur3_fmix(h2 ^ h1) << 32) | murmur3_fmix(h3);
    uint64_t stream = splitmix64(seed1);
    uint64_t initseq = splitmix64(seed2);
    s.inc = (stream << 1u) | 1u;
    s.state = 0;
    s.state = s.state * PCG_MULT + s.inc;
    s.state += initseq;
    s.state = s.state * PCG_MULT + s.inc;
    for (int i = 0; i < 8; ++i) s.state = s.state * PCG_MULT + s.inc;
}
static inline uint32_t pcg32(Rng& s) {  // random.hpp:105-112
    uint64_t old = s.state;
    s.state = old * PCG_MULT + s.inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
static inline float randf(Rng& s) {  // random.hpp:118-120
    return std::ldexp((float)(pcg32(s) >> 9), -23);
}

// --------------------------------------------------------------- geometry --
static const float RAY_T_MIN = 1e-4f;   // intersection.hpp:24
static const float RAY_T_MAX = 1e30f;   // intersection.hpp:25
static const float AREA_EPS = 1e-6f * 1e-6f;  // render_units.hpp:474-476
static const float GEOM_EPS = 1e-6f;    // render_units.hpp:479

struct Ray {
    V3 o, d;
};
static inline V3 ray_at(const Ray& r, float t) { return add(r.o, mul(r.d, t)); }

// intersection.hpp:412-419
static inline V3 inv_dir(V3 d) {
    return v3(std::fabs(d.x) > 1e-8f ? 1.0f / d.x : 0.0f,
              std::fabs(d.y) > 1e-8f ? 1.0f / d.y : 0.0f,
              std::fabs(d.z) > 1e-8f ? 1.0f / d.z : 0.0f);
}
// intersection.hpp:257-302
static inline bool axis_slab(float dc, float ic, float dmin, float dmax, float& tmin, float& tmax) {
    if (std::fabs(dc) < 1e-8f) {
        bool outP = (dmin > 0.0f && dmax > 0.0f);
        bool outN = (dmin < 0.0f && dmax < 0.0f);
        return !(outP || outN);
    }
    float t1 = dmin * ic;
    float t2 = dmax * ic;
    float amin = fmin_std(t1, t2);
    float amax = fmax_std(t1, t2);
    tmin = fmax_std(tmin, amin);
    tmax = fmin_std(tmax, amax);
    return tmax >= tmin;
}
static inline bool intersect_aabb(const Ray& r, V3 id, V3 bmin, V3 bmax) {
    V3 tomin = sub(bmin, r.o), tomax = sub(bmax, r.o);
    float tmin = RAY_T_MIN * 0.0f, tmax = RAY_T_MAX;
    if (!axis_slab(r.d.x, id.x, tomin.x, tomax.x, tmin, tmax)) return false;
    if (!axis_slab(r.d.y, id.y, tomin.y, tomax.y, tmin, tmax)) return false;
    if (!axis_slab(r.d.z, id.z, tomin.z, tomax.z, tmin, tmax)) return false;
    return tmax >= fmax_std(tmin, 0.0f);
}
// Moller-Trumbore, intersection.hpp:335-396
static inline bool intersect_tri(const Ray& r, V3 a, V3 b, V3 c, float tmin, float tmax,
                                 float& t_out, float& u_out, float& v_out) {
    V3 e1 = sub(b, a), e2 = sub(c, a);
    V3 pvec = cross(r.d, e2);
    float det = dot(e1, pvec);
    if (det > -AREA_EPS && det < AREA_EPS) return false;
    V3 tvec = sub(r.o, a);
    float u_num = dot(tvec, pvec);
    float u = u_num / det;
    if (u < 0.0f || u > 1.0f) return false;
    V3 qvec = cross(tvec, e1);
    float v_num = dot(r.d, qvec);
    float v = v_num / det;
    if (v < 0.0f || u + v > 1.0f) return false;
    float t_num = dot(e2, qvec);
    float t = t_num / det;
    if (t >= tmin && t <= tmax) {
        t_out = t; u_out = u; v_out = v;
        return true;
    }
    return false;
}
// intersection.hpp:58-104
static inline bool intersect_sphere(const Ray& r, V3 c, float radius, float& t_hit, V3& n_hit) {
    V3 oc = sub(r.o, c);
    const float a = 1.0f;
    float b = 2.0f * dot(oc, r.d);
    float cc = len2(oc) - radius * radius;
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    float sq = std::sqrt(fmax_std(0.0f, disc));
    float t = (-b - sq) / (2.0f * a);
    if (t < RAY_T_MIN) {
        t = (-b + sq) / (2.0f * a);
        if (t < RAY_T_MIN) return false;
    }
    t_hit = t;
    V3 hp = add(r.o, mul(r.d, t));
    n_hit = normalize_plain(sub(hp, c));
    return true;
}
// intersection.hpp:128-174 (rectangle) and :192-236 (ellipse)
static inline bool intersect_planar(const Ray& r, V3 c, V3 n, V3 right, V3 up, float sx, float sy,
                                    bool ellipse, float& t_hit, V3& p_hit) {
    float denom = dot(r.d, n);
    if (std::fabs(denom) < GEOM_EPS) return false;
    V3 cp = sub(c, r.o);
    float num = dot(cp, n);
    float t = num / denom;
    if (t < RAY_T_MIN) return false;
    V3 hp = add(r.o, mul(r.d, t));
    V3 lh = sub(hp, c);
    float x = dot(lh, right);
    float y = dot(lh, up);
    if (ellipse) {
        float nx = x / sx, ny = y / sy;
        if (nx * nx + ny * ny > 1.0f) return false;
    } else {
        float hx = sx * 0.5f, hy = sy * 0.5f;
        if (x * x > hx * hx || y * y > hy * hy) return false;
    }
    t_hit = t;
    p_hit = hp;
    return true;
}

// ----------------------------------------------------------------- scene --
// core/scene.hpp:32-44, 48-58, 157-165; core/material.hpp:139-217
struct Tri {
    int i0 = 0, i1 = 0, i2 = 0;
    V3 fn = v3(0, 1, 0), n0 = v3(0, 1, 0), n1 = v3(0, 1, 0), n2 = v3(0, 1, 0);
    V3 centroid;
    bool smooth = false;
    float uv0[2] = {0, 0}, uv1[2] = {0, 0}, uv2[2] = {0, 0};  // core/scene.hpp:36
    bool has_uv = false;
};
struct Node {
    V3 bmin, bmax;
    int left = -1, right = -1, tri_start = 0, tri_count = 0;
};
struct Volume {
    V3 color = v3(1, 1, 1), absorption = v3(0, 0, 0);
    float density = 0.0f, anisotropy = 0.0f;
    std::vector<float> grid;
    int dims[3] = {0, 0, 0};
    float wmin[3] = {0, 0, 0}, wmax[3] = {0, 0, 0};
    float grid_max = 0.0f;
    bool present() const { return density > 0.0f || has_grid(); }
    bool has_grid() const { return !grid.empty() && dims[0] > 0 && dims[1] > 0 && dims[2] > 0; }
};
// Shader-node graph, core/material.hpp:29-127 (SocketValue, NodeSocket,
// MaterialNode, NodeTree), filled as parseNodeTree does
// (src/node_evaluator.cpp:82-149).
enum SockType { SV_FLOAT, SV_VEC3, SV_VEC4, SV_NONE };
struct SockVal {
    SockType type = SV_NONE;
    float f = 0.0f;
    V3 v3_ = v3(0, 0, 0);   // VEC3
    V3 v4 = v3(0, 0, 0);    // VEC4 colour part
};
struct NodeSocket {
    std::string name;
    SockVal dv;
    bool is_linked = false;
    std::string linked_node, linked_socket;
};
struct MatNode {
    std::string name, type;
    std::vector<NodeSocket> inputs, outputs;
    const NodeSocket* find_input(const std::string& n) const {
        for (const NodeSocket& x : inputs)
            if (x.name == n) return &x;
        return nullptr;
    }
};
struct NodeTree {
    std::vector<MatNode> nodes;
    bool valid = false;
    const MatNode* find_node(const std::string& n) const {
        for (const MatNode& x : nodes)
            if (x.name == n) return &x;
        return nullptr;
    }
    const MatNode* find_output() const {
        for (const MatNode& x : nodes)
            if (x.type == "ShaderNodeOutputMaterial") return &x;
        return nullptr;
    }
};
struct Material {
    V3 albedo = v3(0.8f, 0.8f, 0.8f);
    float metallic = 0.0f, roughness = 0.5f;
    V3 emission = v3(0, 0, 0);
    float transmission = 0.0f, ior = 1.45f;
    Volume vol;
    bool volume_boundary_only = false;
    NodeTree tree;
    bool use_nodes = false;   // "use_nodes" && tree.valid
};
struct Mesh {
    std::vector<V3> verts;
    std::vector<Tri> tris;
    Material mat;
    std::vector<Node> nodes;
    std::vector<int> tri_idx;
    V3 bmin = v3(1e30f, 1e30f, 1e30f), bmax = v3(-1e30f, -1e30f, -1e30f);
};
enum LightType { L_EMISSIVE = 0, L_POINT = 1, L_SUN = 2, L_SPOT = 3, L_AREA = 4 };
enum AreaShape { S_SQUARE = 0, S_RECTANGLE = 1, S_DISK = 2, S_ELLIPSE = 3 };
struct Light {  // light/light.hpp:42-87
    int type = L_EMISSIVE;
    V3 pos, normal = v3(0, 0, 1), emission;
    float area = 0.0f;
    int mesh_index = -1, tri_index = -1;
    V3 v0, v1, v2;
    float energy = 1.0f, radius = 0.0f, spot_angle = 0.0f, spot_blend = 0.0f;
    V3 right = v3(1, 0, 0), up = v3(0, 1, 0);
    float sx = 1.0f, sy = 1.0f;
    int shape = S_SQUARE;
};
struct Camera {  // core/ray.hpp:43-72; pybind_renderer.hpp:131-160
    V3 pos, forward = v3(0, 0, 1), right = v3(1, 0, 0), up = v3(0, 1, 0);
    float fov = 60.0f * (3.14159265358979323846f / 180.0f);
    float exposure_ev = 0.0f;
    bool set = false;
};
struct SceneLights {  // light/scene_lights.hpp:44-167
    std::vector<Light> lights;
    std::vector<float> cdf;
    float total_area = 0.0f;
    int native_start = 0;
};
struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Light> native;
    V3 env_color = v3(0.05f, 0.05f, 0.05f);
    float env_strength = 1.0f;
    Camera cam;
    SceneLights sl;
    // per mesh: triangle -> index in sl.lights (or -1); replaces the O(L)
    // findLightIndex scan of path_tracer.hpp:414-422 with an equal result.
    std::vector<std::vector<int>> tri_light;
};

// BVH::build / buildRecursive, core/scene.hpp:65-152
static int bvh_build_rec(Mesh& m, int start, int end) {
    int idx = (int)m.nodes.size();
    m.nodes.push_back(Node());
    Node node;
    const float inf = 1e30f;
    node.bmin = v3(inf, inf, inf);
    node.bmax = v3(-inf, -inf, -inf);
    for (int i = start; i < end; ++i) {
        const Tri& t = m.tris[m.tri_idx[i]];
        for (int k = 0; k < 3; ++k) {
            V3 p = m.verts[k == 0 ? t.i0 : (k == 1 ? t.i1 : t.i2)];
            node.bmin = v3(fmin_std(node.bmin.x, p.x), fmin_std(node.bmin.y, p.y), fmin_std(node.bmin.z, p.z));
            node.bmax = v3(fmax_std(node.bmax.x, p.x), fmax_std(node.bmax.y, p.y), fmax_std(node.bmax.z, p.z));
        }
    }
    int count = end - start;
    if (count <= 4) {
        node.tri_start = start;
        node.tri_count = count;
        node.left = node.right = -1;
        m.nodes[idx] = node;
        return idx;
    }
    V3 ext = sub(node.bmax, node.bmin);
    int axis = 0;
    if (ext.y > ext.x) axis = 1;
    if (ext.z > (axis == 0 ? ext.x : ext.y)) axis = 2;
    int mid = (start + end) / 2;
    std::nth_element(m.tri_idx.begin() + start, m.tri_idx.begin() + mid, m.tri_idx.begin() + end,
                     [&](int a, int b) {
                         const V3& ca = m.tris[a].centroid;
                         const V3& cb = m.tris[b].centroid;
                         float fa = axis == 0 ? ca.x : (axis == 1 ? ca.y : ca.z);
                         float fb = axis == 0 ? cb.x : (axis == 1 ? cb.y : cb.z);
                         return fa < fb;
                     });
    node.tri_start = 0;
    node.tri_count = 0;
    m.nodes[idx] = node;
    int l = bvh_build_rec(m, start, mid);
    int r = bvh_build_rec(m, mid, end);
    m.nodes[idx].left = l;
    m.nodes[idx].right = r;
    return idx;
}
static void bvh_build(Mesh& m) {
    if (m.tris.empty()) return;
    m.tri_idx.resize(m.tris.size());
    for (size_t i = 0; i < m.tris.size(); ++i) {
        m.tri_idx[i] = (int)i;
        Tri& t = m.tris[i];
        V3 a = m.verts[t.i0], b = m.verts[t.i1], c = m.verts[t.i2];
        V3 ab = sub(b, a), ac = sub(c, a);
        t.centroid = add(a, divs(add(ab, ac), 3.0f));
    }
    m.nodes.reserve(2 * m.tris.size());
    bvh_build_rec(m, 0, (int)m.tris.size());
}
// finalizeMeshBounds, core/scene.hpp:191-227
static void finalize_mesh(Mesh& m) {
    for (const V3& v : m.verts) {
        m.bmin = v3(fmin_std(m.bmin.x, v.x), fmin_std(m.bmin.y, v.y), fmin_std(m.bmin.z, v.z));
        m.bmax = v3(fmax_std(m.bmax.x, v.x), fmax_std(m.bmax.y, v.y), fmax_std(m.bmax.z, v.z));
    }
    const float eps = 0.001f;
    V3 ext = sub(m.bmax, m.bmin);
    if (ext.x < eps) { m.bmin.x = m.bmin.x - 1.0f * eps; m.bmax.x = m.bmax.x + 1.0f * eps; }
    if (ext.y < eps) { m.bmin.y = m.bmin.y - 1.0f * eps; m.bmax.y = m.bmax.y + 1.0f * eps; }
    if (ext.z < eps) { m.bmin.z = m.bmin.z - 1.0f * eps; m.bmax.z = m.bmax.z + 1.0f * eps; }
    for (Tri& t : m.tris) {
        V3 e1 = sub(m.verts[t.i1], m.verts[t.i0]);
        V3 e2 = sub(m.verts[t.i2], m.verts[t.i0]);
        V3 n;
        if (make_dir(cross(e1, e2), n)) t.fn = n;
    }
    bvh_build(m);
}

// Hit record: core/material.hpp:447-485
struct Hit {
    bool hit = false;
    float t = 1e30f;
    V3 point, normal = v3(0, 1, 0);
    int mesh = -1, tri = -1;   // tri = original triangle index
    int bvh_pos = -1;          // position in the mesh's triIndices (tie key)
    float u = 0.0f, v = 0.0f;
    float uv[2] = {0.0f, 0.0f};  // texture coordinate; kept from an earlier
    bool has_uv = false;         // candidate when the closest triangle has none
};

// traverseBVH, geometry/bvh.hpp:34-142
static void traverse_bvh(const Mesh& m, const Ray& r, V3 id, Hit& res, int mesh_idx, float tmin, float tmax) {
    float cur = fmin_std(tmax, res.t);
    if (m.nodes.empty()) return;  // meshes here always have triangles
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        int ni = stack[--sp];
        const Node& nd = m.nodes[ni];
        if (!intersect_aabb(r, id, nd.bmin, nd.bmax)) continue;
        if (nd.tri_count > 0) {
            for (int i = 0; i < nd.tri_count; ++i) {
                int ti = m.tri_idx[nd.tri_start + i];
                const Tri& tr = m.tris[ti];
                float t, u, v;
                if (intersect_tri(r, m.verts[tr.i0], m.verts[tr.i1], m.verts[tr.i2], tmin, cur, t, u, v)) {
                    cur = t;
                    res.t = t;
                    res.point = ray_at(r, t);
                    float w = 1.0f - u - v;  // bary_w
                    if (tr.smooth) {
                        V3 ni3 = add(add(mul(tr.n0, w), mul(tr.n1, u)), mul(tr.n2, v));
                        res.normal = make_dir_or(ni3, tr.fn);
                    } else {
                        res.normal = tr.fn;
                    }
                    res.hit = true;
                    res.mesh = mesh_idx;
                    res.tri = ti;
                    res.bvh_pos = nd.tri_start + i;
                    res.u = u;
                    res.v = v;
                    // bvh.hpp:121-128: uv0 * w + uv1 * u + uv2 * v, left unchanged
                    // for a triangle without UVs
                    if (tr.has_uv) {
                        res.uv[0] = tr.uv0[0] * w + tr.uv1[0] * u + tr.uv2[0] * v;
                        res.uv[1] = tr.uv0[1] * w + tr.uv1[1] * u + tr.uv2[1] * v;
                        res.has_uv = true;
                    } else {
                        res.has_uv = false;
                    }
                }
            }
        } else {
            if (sp < 62) {
                if (nd.right >= 0) stack[sp++] = nd.right;
                if (nd.left >= 0) stack[sp++] = nd.left;
            }
        }
    }
}
// intersectScene, geometry/bvh.hpp:155-177
static Hit intersect_scene(const Scene& s, const Ray& r, float tmin = RAY_T_MIN, float tmax = RAY_T_MAX) {
    Hit res;
    res.t = tmax;
    res.point = v3(0, 0, 0);
    res.normal = v3(0, 1, 0);
    res.hit = false;
    V3 id = inv_dir(r.d);
    for (size_t mi = 0; mi < s.meshes.size(); ++mi) {
        const Mesh& m = s.meshes[mi];
        if (!intersect_aabb(r, id, m.bmin, m.bmax)) continue;
        traverse_bvh(m, r, id, res, (int)mi, tmin, res.t);
    }
    return res;
}

struct LightHit {  // light/light.hpp:126-139
    bool hit = false;
    float t = 1e30f;
    V3 point, normal = v3(0, 0, 1), emission;
    int light_index = -1;
};
// intersectNativeLights, light/scene_lights.hpp:176-246
static LightHit intersect_native_lights(const Scene& s, const Ray& r, int depth) {
    LightHit res;
    if (depth == 0) return res;
    for (size_t i = 0; i < s.native.size(); ++i) {
        const Light& l = s.native[i];
        float t = 0.0f;
        V3 hp, hn;
        if (l.type == L_POINT) {
            if (l.radius > 1e-6f) {
                if (intersect_sphere(r, l.pos, l.radius, t, hn)) {
                    if (t < res.t) {
                        res.hit = true; res.t = t; res.point = ray_at(r, t); res.normal = hn;
                        res.emission = l.emission; res.light_index = (int)i;
                    }
                }
            }
        } else if (l.type == L_AREA) {
            bool h;
            if (l.shape == S_DISK || l.shape == S_ELLIPSE) {
                h = intersect_planar(r, l.pos, l.normal, l.right, l.up, l.sx * 0.5f, l.sy * 0.5f, true, t, hp);
            } else {
                h = intersect_planar(r, l.pos, l.normal, l.right, l.up, l.sx, l.sy, false, t, hp);
            }
            if (h && t < res.t) {
                float facing = dot(l.normal, r.d);
                if (facing < 0) {
                    res.hit = true; res.t = t; res.point = hp; res.normal = l.normal;
                    res.emission = l.emission; res.light_index = (int)i;
                }
            }
        }
    }
    return res;
}

// path_tracer.hpp:82-95
static void intersect_skip_vb(const Scene& s, const Ray& r, float tmin, int depth, Hit& hit, LightHit& lh) {
    hit = Hit();
    hit.hit = false;
    float smin = tmin;
    for (int peek = 0; peek < 8; ++peek) {
        Hit h = intersect_scene(s, r, smin);
        if (!h.hit) break;
        if (!s.meshes[h.mesh].mat.volume_boundary_only) { hit = h; break; }
        smin = h.t + 1e-3f;
    }
    lh = intersect_native_lights(s, r, depth);
}
// path_tracer.hpp:103-119
static bool is_shadowed_skip_vb(const Scene& s, const Ray& ray, float tmax) {
    Ray probe = ray;
    float toff = 0.0f;
    for (int peek = 0; peek < 8; ++peek) {
        float remaining = tmax - toff;
        if (remaining <= 0.0f) return false;
        Hit h = intersect_scene(s, probe, RAY_T_MIN, remaining);
        if (!h.hit) return false;
        if (!s.meshes[h.mesh].mat.volume_boundary_only) return true;
        float skip = h.t + 1e-3f;
        toff += skip;
        probe.o = ray_at(probe, skip);
    }
    return false;
}

// ------------------------------------------------------------------ BSDF --
static const float PI_F = 3.14159265358979323846f;       // std::numbers::pi_v<float>
static const float INV_PI_F = 0.318309886183790671538f;  // std::numbers::inv_pi_v<float>
static const float GGX_EPS = 1e-6f;
static const float MIN_ROUGH = 0.01f;
static const float MIN_PDF = 1e-6f;

struct MatParams {  // bsdf/bsdf.hpp:24-37
    V3 albedo;
    float metallic, roughness, transmission, ior;
};
// getMaterialParams, path_tracer.hpp:45-65 (defined after the node-tree
// evaluator below)
static inline MatParams mat_params(const Material& m, float u = 0.0f, float v = 0.0f);
static inline float lum_f(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
static inline bool is_emissive(V3 e) { return lum_f(e) > 1e-6f; }  // render_units.hpp:1197-1202

// ---- node-tree evaluation, src/node_evaluator.cpp:168-740 ----
static const V3 NODE_GRAY = {0.8f, 0.8f, 0.8f};
// A socket's colour: VEC4 -> v4, VEC3 -> v3; false for other types.
static bool sock_color(const NodeSocket* s, V3& c) {
    if (s->dv.type == SV_VEC4) { c = s->dv.v4; return true; }
    if (s->dv.type == SV_VEC3) { c = s->dv.v3_; return true; }
    return false;
}
// evaluateNode, node_evaluator.cpp:168-415.  The reference recurses without
// a bound; the depth guard only keeps a cyclic test tree from overflowing
// this process's stack.
static V3 eval_node(const NodeTree& t, const std::string& name, const std::string& sock, float u, float v,
                    int depth = 0) {
    const MatNode* n = t.find_node(name);
    if (!n || depth > 4096) return NODE_GRAY;
    V3 c;
    if (n->type == "ShaderNodeBsdfPrincipled") {
        if (sock == "BSDF") {
            const NodeSocket* bc = n->find_input("Base Color");
            if (!bc) return NODE_GRAY;
            if (bc->is_linked) return eval_node(t, bc->linked_node, bc->linked_socket, u, v, depth + 1);
            return sock_color(bc, c) ? c : NODE_GRAY;
        } else if (sock == "Base Color") {
            const NodeSocket* bc = n->find_input("Base Color");
            if (bc && !bc->is_linked && sock_color(bc, c)) return c;
            return NODE_GRAY;
        }
    } else if (n->type == "ShaderNodeEmission") {
        if (sock == "Emission") {
            const NodeSocket* cs = n->find_input("Color");
            const NodeSocket* ss = n->find_input("Strength");
            V3 color = v3(1.0f, 1.0f, 1.0f);
            float strength = 1.0f;
            if (cs) {
                if (cs->is_linked) color = eval_node(t, cs->linked_node, cs->linked_socket, u, v, depth + 1);
                else if (sock_color(cs, c)) color = c;
            }
            if (ss && !ss->is_linked && ss->dv.type == SV_FLOAT) strength = ss->dv.f;
            return mul(color, strength);
        }
    } else if (n->type == "ShaderNodeBsdfDiffuse") {
        if (sock == "BSDF") {
            const NodeSocket* cs = n->find_input("Color");
            if (!cs) return NODE_GRAY;
            if (cs->is_linked) return eval_node(t, cs->linked_node, cs->linked_socket, u, v, depth + 1);
            return sock_color(cs, c) ? c : NODE_GRAY;
        }
    } else if (n->type == "ShaderNodeRGB") {
        for (const NodeSocket& o : n->outputs)
            if (o.name == "Color" && o.dv.type != SV_NONE && sock_color(&o, c)) return c;
        return v3(1.0f, 1.0f, 1.0f);
    } else if (n->type == "ShaderNodeMix" || n->type == "ShaderNodeMixRGB") {
        const NodeSocket* fs = n->find_input("Fac");
        const NodeSocket* as = n->find_input("A");
        const NodeSocket* bs = n->find_input("B");
        float fac = 0.5f;
        V3 a = v3(0, 0, 0), b = v3(1.0f, 1.0f, 1.0f);
        if (fs && !fs->is_linked && fs->dv.type == SV_FLOAT) fac = fs->dv.f;
        if (as) {
            if (as->is_linked) a = eval_node(t, as->linked_node, as->linked_socket, u, v, depth + 1);
            else if (sock_color(as, c)) a = c;
        }
        if (bs) {
            if (bs->is_linked) b = eval_node(t, bs->linked_node, bs->linked_socket, u, v, depth + 1);
            else if (sock_color(bs, c)) b = c;
        }
        return add(mul(a, 1.0f - fac), mul(b, fac));
    } else if (n->type == "ShaderNodeTexChecker") {
        const NodeSocket* c1s = n->find_input("Color1");
        const NodeSocket* c2s = n->find_input("Color2");
        const NodeSocket* ss = n->find_input("Scale");
        V3 c1 = v3(0.8f, 0.8f, 0.8f), c2 = v3(0.2f, 0.2f, 0.2f);
        float scale = 5.0f;
        if (c1s) {
            if (c1s->is_linked) c1 = eval_node(t, c1s->linked_node, c1s->linked_socket, u, v, depth + 1);
            else if (sock_color(c1s, c)) c1 = c;
        }
        if (c2s) {
            if (c2s->is_linked) c2 = eval_node(t, c2s->linked_node, c2s->linked_socket, u, v, depth + 1);
            else if (sock_color(c2s, c)) c2 = c;
        }
        if (ss && !ss->is_linked && ss->dv.type == SV_FLOAT) scale = ss->dv.f;
        float fu = u * scale, fv = v * scale;
        // static_cast<int>(std::floor(x)) as x86 computes it: out-of-range
        // and NaN convert to INT_MIN
        auto to_int = [](float x) -> long long {
            float f = std::floor(x);
            return (f >= -2147483648.0f && f < 2147483648.0f) ? (long long)f : -2147483648LL;
        };
        long long cu = to_int(fu) % 2, cv = to_int(fv) % 2;
        if (cu < 0) cu += 2;
        if (cv < 0) cv += 2;
        return ((cu + cv) % 2 == 0) ? c1 : c2;
    } else if (n->type == "ShaderNodeOutputMaterial") {
        const NodeSocket* ss = n->find_input("Surface");
        if (ss && ss->is_linked) return eval_node(t, ss->linked_node, ss->linked_socket, u, v, depth + 1);
    }
    return NODE_GRAY;
}
// The Principled node the Material Output's linked Surface comes from.
static const MatNode* surface_principled(const NodeTree& t, const NodeSocket** surf = nullptr) {
    const MatNode* out = t.find_output();
    if (!out) return nullptr;
    const NodeSocket* s = out->find_input("Surface");
    if (!s || !s->is_linked) return nullptr;
    if (surf) *surf = s;
    const MatNode* sh = t.find_node(s->linked_node);
    return (sh && sh->type == "ShaderNodeBsdfPrincipled") ? sh : nullptr;
}
// getAlbedoFromNodeTree, :436-478
static V3 tree_albedo(const NodeTree& t, float u, float v) {
    if (!t.valid) return NODE_GRAY;
    const MatNode* out = t.find_output();
    if (!out) return NODE_GRAY;
    const NodeSocket* s = out->find_input("Surface");
    if (!s || !s->is_linked) return NODE_GRAY;
    return eval_node(t, s->linked_node, s->linked_socket, u, v);
}
// getEmissionFromNodeTree, :501-565
static V3 tree_emission(const NodeTree& t, float u, float v) {
    if (!t.valid) return v3(0, 0, 0);
    const MatNode* out = t.find_output();
    if (!out) return v3(0, 0, 0);
    const NodeSocket* s = out->find_input("Surface");
    if (!s || !s->is_linked) return v3(0, 0, 0);
    const MatNode* sh = t.find_node(s->linked_node);
    if (sh && sh->type == "ShaderNodeEmission") return eval_node(t, s->linked_node, "Emission", u, v);
    if (sh && sh->type == "ShaderNodeBsdfPrincipled") {
        const NodeSocket* es = sh->find_input("Emission Color");
        if (!es) es = sh->find_input("Emission");
        if (es) {
            const NodeSocket* st = sh->find_input("Emission Strength");
            float strength = (st && st->dv.type == SV_FLOAT) ? st->dv.f : 0.0f;
            if (es->is_linked) return mul(eval_node(t, es->linked_node, es->linked_socket, u, v), strength);
            if (es->dv.type == SV_VEC4) return mul(es->dv.v4, strength);
            if (es->dv.type == SV_VEC3) return mul(es->dv.v3_, strength);
        }
    }
    return v3(0, 0, 0);
}
// getEmission, path_tracer.hpp:67-72
static inline V3 mat_emission(const Material& m, float u = 0.0f, float v = 0.0f);
// get{Transmission,IOR,Metallic,Roughness}FromNodeTree, :579-740
static float tree_scalar(const NodeTree& t, const char* name, const char* alt, float dflt) {
    if (!t.valid) return dflt;
    const MatNode* sh = surface_principled(t);
    if (!sh) return dflt;
    const NodeSocket* s = sh->find_input(name);
    if (!s && alt) s = sh->find_input(alt);
    if (s && !s->is_linked && s->dv.type == SV_FLOAT) return s->dv.f;
    return dflt;
}
static inline MatParams mat_params(const Material& m, float u, float v) {
    MatParams p;
    p.albedo = m.albedo; p.metallic = m.metallic; p.roughness = m.roughness;
    p.transmission = m.transmission; p.ior = m.ior;
    if (m.use_nodes && m.tree.valid) {
        p.albedo = tree_albedo(m.tree, u, v);
        p.metallic = tree_scalar(m.tree, "Metallic", nullptr, 0.0f);
        p.roughness = tree_scalar(m.tree, "Roughness", nullptr, 0.5f);
        p.transmission = tree_scalar(m.tree, "Transmission", "Transmission Weight", 0.0f);
        p.ior = tree_scalar(m.tree, "IOR", nullptr, 1.45f);
    }
    p.roughness = fmax_std(p.roughness, MIN_ROUGH);
    return p;
}
static inline V3 mat_emission(const Material& m, float u, float v) {
    if (m.use_nodes && m.tree.valid) return tree_emission(m.tree, u, v);
    return m.emission;
}

// fresnel.hpp:40-48
static inline V3 fresnel_schlick_color(float c, V3 f0) {
    float f = std::pow(1.0f - c, 5.0f);
    return v3(f0.x + (1.0f - f0.x) * f, f0.y + (1.0f - f0.y) * f, f0.z + (1.0f - f0.z) * f);
}
static inline float fresnel_schlick(float c, float f0) { return f0 + (1.0f - f0) * std::pow(1.0f - c, 5.0f); }
// fresnel.hpp:74-97
static inline float fresnel_dielectric(float ci, float eta) {
    ci = std::min(std::max(std::fabs(ci), 0.0f), 1.0f);
    float si2 = 1.0f - ci * ci;
    float st2 = eta * eta * si2;
    if (st2 >= 1.0f) return 1.0f;
    float ct = std::sqrt(fmax_std(0.0f, 1.0f - st2));
    float eci = eta * ci, ect = eta * ct;
    float rs = (eci - ct) / (eci + ct + 1e-6f);
    float rp = (ci - ect) / (ci + ect + 1e-6f);
    return 0.5f * (rs * rs + rp * rp);
}
// ggx.hpp:28-122
static inline float safe_sqrt(float x) { return std::sqrt(fmax_std(0.0f, x)); }
static inline void onb(V3 n, V3& t, V3& b) {  // ggx.hpp:33-43
    if (n.z < -0.9999f) {
        t = v3(0.0f, -1.0f, 0.0f);
        b = v3(-1.0f, 0.0f, 0.0f);
        return;
    }
    float a = 1.0f / (1.0f + n.z);
    float bb = -n.x * n.y * a;
    t = make_dir_or(v3(1.0f - n.x * n.x * a, bb, -n.x), v3(0, 0, 1));
    b = make_dir_or(v3(bb, 1.0f - n.y * n.y * a, -n.y), v3(0, 0, 1));
}
static inline V3 local_to_world(V3 l, V3 n, V3 t, V3 b) {  // ggx.hpp:46-53
    return make_dir_or(v3(t.x * l.x + b.x * l.y + n.x * l.z,
                          t.y * l.x + b.y * l.y + n.y * l.z,
                          t.z * l.x + b.z * l.y + n.z * l.z), v3(0, 0, 1));
}
static inline V3 world_to_local(V3 w, V3 n, V3 t, V3 b) {  // ggx.hpp:56-63
    return make_dir_or(v3(dot(w, t), dot(w, b), dot(w, n)), v3(0, 0, 1));
}
static inline float ggx_d(float nh, float r) {  // ggx.hpp:87-93
    float alpha = r * r;
    float a2 = alpha * alpha;
    float nh2 = nh * nh;
    float den = nh2 * (a2 - 1.0f) + 1.0f;
    return a2 * INV_PI_F / (den * den + GGX_EPS);
}
static inline float ggx_g1(float nv, float r) {  // ggx.hpp:100-104
    float alpha = r * r;
    float a2 = alpha * alpha;
    return 2.0f * nv / (nv + safe_sqrt(a2 + (1.0f - a2) * nv * nv));
}
static inline float ggx_g2(float nl, float nv, float r) {  // ggx.hpp:109-122
    if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
    float alpha = r * r;
    float a2 = alpha * alpha;
    float ll = safe_sqrt(a2 + (1.0f - a2) * nl * nl);
    float lv = safe_sqrt(a2 + (1.0f - a2) * nv * nv);
    float den = nv * ll + nl * lv;
    if (den < GGX_EPS) return 0.0f;
    return 2.0f * nl * nv / den;
}
// ggx.hpp:130-168
static inline V3 sample_vndf(V3 wo, float r, float u1, float u2, V3 n, V3 t, V3 b) {
    V3 wl = world_to_local(wo, n, t, b);
    float alpha = r * r;
    V3 ws = make_dir_or(v3(wl.x * alpha, wl.y * alpha, wl.z), v3(0, 0, 1));
    V3 t1;
    if (ws.z < 0.9999f) {
        V3 c;
        t1 = make_dir(cross(v3(0, 0, 1), ws), c) ? c : v3(1, 0, 0);
    } else {
        t1 = v3(1, 0, 0);
    }
    V3 t2;
    if (!make_dir(cross(ws, t1), t2)) t2 = v3(0, 1, 0);
    float rr = std::sqrt(u1);
    float phi = 2.0f * PI_F * u2;
    float p1 = rr * std::cos(phi);
    float p2 = rr * std::sin(phi);
    float s = 0.5f * (1.0f + ws.z);
    p2 = (1.0f - s) * std::sqrt(1.0f - p1 * p1) + s * p2;
    V3 h = add(add(mul(t1, p1), mul(t2, p2)), mul(ws, std::sqrt(fmax_std(0.0f, 1.0f - p1 * p1 - p2 * p2))));
    h = v3(h.x * alpha, h.y * alpha, fmax_std(0.0f, h.z));
    V3 hl = make_dir_or(h, v3(0, 0, 1));
    return local_to_world(hl, n, t, b);
}
static inline float pdf_vndf(V3 wo, V3 h, float r, V3 n) {  // ggx.hpp:173-183
    float nh = dot(n, h), vh = dot(wo, h), nv = dot(n, wo);
    if (nh <= 0.0f || vh <= 0.0f || nv <= 0.0f) return 0.0f;
    float D = ggx_d(nh, r);
    float G1 = ggx_g1(nv, r);
    return D * G1 / (4.0f * nv);
}
static inline V3 sample_cosine(float u1, float u2, V3 n) {  // ggx.hpp:191-202
    float r = std::sqrt(u1);
    float th = 2.0f * PI_F * u2;
    float x = r * std::cos(th), y = r * std::sin(th);
    float z = std::sqrt(fmax_std(0.0f, 1.0f - u1));
    V3 t, b;
    onb(n, t, b);
    return local_to_world(v3(x, y, z), n, t, b);
}
static inline float pdf_cosine(float nl) { return fmax_std(0.0f, nl) * INV_PI_F; }

// evalSpecular / evalBSDF, bsdf.hpp:126-190
static inline V3 eval_specular(const MatParams& m, V3 wo, V3 wi, V3 n, float nl, float nv) {
    if (nl <= 0.0f || nv <= 0.0f) return v3(0, 0, 0);
    V3 h = make_dir_or(add(wo, wi), v3(0, 0, 1));
    float nh = fmax_std(dot(n, h), 0.0f);
    float vh = fmax_std(dot(wo, h), 0.0f);
    float r = fmax_std(m.roughness, MIN_ROUGH);
    float D = ggx_d(nh, r);
    float alpha = r * r, a2 = alpha * alpha;
    float ll = safe_sqrt(a2 + (1.0f - a2) * nl * nl);
    float lv = safe_sqrt(a2 + (1.0f - a2) * nv * nv);
    float g2d = 0.5f / (nv * ll + nl * lv + GGX_EPS);
    V3 f0 = add(mul(m.albedo, m.metallic), mul(v3(0.04f, 0.04f, 0.04f), 1.0f - m.metallic));
    V3 F = fresnel_schlick_color(vh, f0);
    float spec = D * g2d;
    return cmul(v3(spec, spec, spec), F);
}
static inline V3 eval_bsdf(const MatParams& m, V3 wo, V3 wi, V3 n) {
    float nl = dot(n, wi), nv = dot(n, wo);
    if (nl <= 0.0f || nv <= 0.0f) return v3(0, 0, 0);
    V3 spec = eval_specular(m, wo, wi, n, nl, nv);
    V3 h = make_dir_or(add(wo, wi), v3(0, 0, 1));
    float vh = fmax_std(dot(wo, h), 0.0f);
    V3 F = fresnel_schlick_color(vh, v3(0.04f, 0.04f, 0.04f));
    // evalDiffuse: albedo * INV_PI (render_units.hpp:1240-1247 uses 0.31830988618f)
    const float INV_PI_D = 0.31830988618f;
    V3 diff = v3(m.albedo.x * INV_PI_D, m.albedo.y * INV_PI_D, m.albedo.z * INV_PI_D);
    V3 kd = v3((1.0f - F.x) * (1.0f - m.metallic), (1.0f - F.y) * (1.0f - m.metallic),
               (1.0f - F.z) * (1.0f - m.metallic));
    diff = cmul(diff, kd);
    return add(diff, spec);
}
static inline float spec_prob_of(const MatParams& m) {  // bsdf.hpp:279-285
    if (m.metallic > 0.99f) return 1.0f;
    float sp = 0.5f * (1.0f + m.metallic) * (1.0f - m.roughness * 0.5f);
    return std::min(std::max(sp, 0.1f), 0.9f);
}
// pdfBSDF, bsdf.hpp:352-376
static inline float pdf_bsdf(const MatParams& m, V3 wo, V3 wi, V3 n) {
    float nl = dot(n, wi);
    if (nl <= 0.0f) return 0.0f;
    float sp = spec_prob_of(m);
    float r = fmax_std(m.roughness, MIN_ROUGH);
    V3 h = make_dir_or(add(wo, wi), v3(0, 0, 1));
    float ps = pdf_vndf(wo, h, r, n);
    float pd = pdf_cosine(nl);
    return (sp * ps + (1.0f - sp) * pd) * (1.0f - m.transmission);
}

enum SampleKind { SK_EVAL = 0, SK_WEIGHT = 1 };
enum BType { BT_DIFFUSE = 0, BT_SPECULAR = 1, BT_TRANSMISSION = 2 };
struct BSample {  // bsdf.hpp:65-89
    V3 wi = v3(0, 0, 1);
    int kind = SK_EVAL;
    V3 f;          // EvaluatedBSDF::f
    float pdf = 0; // EvaluatedBSDF::pdf
    V3 weight;     // PrecomputedWeight::weight
    int type = BT_DIFFUSE;
    bool valid() const { return kind == SK_WEIGHT ? true : pdf > MIN_PDF; }
};
// bsdf_sample_weight, render_units.hpp:1413-1420
static inline V3 bsdf_weight(V3 f, float c, float pdf) {
    if (pdf < MIN_PDF) return v3(0, 0, 0);
    V3 base = v3(f.x / pdf, f.y / pdf, f.z / pdf);
    return v3(base.x * c, base.y * c, base.z * c);
}
// sampleBSDF, bsdf.hpp:198-345.  u1,u2,u3 are drawn by the caller left to
// right; the glass branch draws one more inside (bsdf.hpp:221).
static BSample sample_bsdf(const MatParams& m, V3 wo, V3 n, float u1, float u2, float u3, Rng& rng) {
    BSample s;
    s.f = v3(0, 0, 0);
    float r = fmax_std(m.roughness, MIN_ROUGH);
    if (m.transmission > 0.0f && u3 < m.transmission) {
        float won = dot(wo, n);
        bool front = won > 0.0f;
        V3 fnrm = front ? n : neg(n);
        float eta = front ? (1.0f / m.ior) : m.ior;
        V3 t, b;
        onb(fnrm, t, b);
        V3 h = sample_vndf(wo, r, u1, u2, fnrm, t, b);
        float ci = std::fabs(dot(wo, h));
        float F = fresnel_dielectric(ci, eta);
        V3 inc = neg(wo);
        if (randf(rng) < F) {
            s.wi = reflect(inc, h);
            float nl = dot(fnrm, s.wi), nv = dot(fnrm, wo);
            if (nl <= 0.0f || nv <= 0.0f) { s.kind = SK_EVAL; s.f = v3(0, 0, 0); s.pdf = 0.0f; return s; }
            float G2 = ggx_g2(nl, nv, r), G1 = ggx_g1(nv, r);
            float w = G2 / (G1 + GGX_EPS);
            s.kind = SK_WEIGHT; s.weight = v3(w, w, w); s.type = BT_SPECULAR;
        } else {
            V3 rf;
            if (!refract(inc, h, eta, rf)) {
                s.wi = reflect(inc, h);
                float nl = dot(fnrm, s.wi), nv = dot(fnrm, wo);
                if (nl <= 0.0f || nv <= 0.0f) { s.kind = SK_EVAL; s.f = v3(0, 0, 0); s.pdf = 0.0f; return s; }
                float G2 = ggx_g2(nl, nv, r), G1 = ggx_g1(nv, r);
                float w = G2 / (G1 + GGX_EPS);
                s.kind = SK_WEIGHT; s.weight = v3(w, w, w); s.type = BT_SPECULAR;
            } else {
                s.wi = rf;
                float nl = std::fabs(dot(s.wi, fnrm));
                float nv = std::fabs(dot(wo, fnrm));
                float G2 = ggx_g2(nl, nv, r), G1 = ggx_g1(nv, r);
                float w = G2 / (G1 + GGX_EPS);
                s.kind = SK_WEIGHT; s.weight = v3(w, w, w); s.type = BT_TRANSMISSION;
            }
        }
        return s;
    }
    float sp = spec_prob_of(m);
    V3 t, b;
    onb(n, t, b);
    (void)fmax_std(dot(n, wo), GGX_EPS);  // NdotV (bsdf.hpp:289) is unused afterwards
    if (u1 < sp) {
        V3 h = sample_vndf(wo, r, u1 / sp, u2, n, t, b);
        s.wi = reflect(neg(wo), h);
        float nl = dot(n, s.wi);
        if (nl <= 0.0f) { s.kind = SK_EVAL; s.f = v3(0, 0, 0); s.pdf = 0.0f; return s; }
        V3 f = eval_bsdf(m, wo, s.wi, n);
        float ps = pdf_vndf(wo, h, r, n);
        float pd = pdf_cosine(nl);
        float pdf = (sp * ps + (1.0f - sp) * pd) * (1.0f - m.transmission);
        if (r < 0.05f && m.metallic > 0.5f) {
            s.kind = SK_WEIGHT; s.weight = bsdf_weight(f, nl, pdf);
        } else {
            s.kind = SK_EVAL; s.f = f; s.pdf = pdf;
        }
        s.type = BT_SPECULAR;
    } else {
        float u1a = (u1 - sp) / (1.0f - sp);
        s.wi = sample_cosine(u1a, u2, n);
        float nl = dot(n, s.wi);
        if (nl <= 0.0f) { s.kind = SK_EVAL; s.f = v3(0, 0, 0); s.pdf = 0.0f; return s; }
        V3 f = eval_bsdf(m, wo, s.wi, n);
        V3 h = make_dir_or(add(wo, s.wi), v3(0, 0, 1));
        float ps = pdf_vndf(wo, h, r, n);
        float pd = pdf_cosine(nl);
        float pdf = (sp * ps + (1.0f - sp) * pd) * (1.0f - m.transmission);
        s.kind = SK_EVAL; s.f = f; s.pdf = pdf;
        s.type = BT_DIFFUSE;
    }
    return s;
}
// compute_throughput_update, bsdf.hpp:92-104
static inline bool throughput_update(const BSample& s, V3 n, V3& w) {
    if (s.kind == SK_WEIGHT) { w = s.weight; return true; }
    float c = std::fabs(dot(n, s.wi));
    if (c <= 1e-6f || s.pdf < MIN_PDF) return false;
    w = bsdf_weight(s.f, c, s.pdf);
    return true;
}
// mis_pdf_for_next_bounce, bsdf.hpp:108-113
static inline float mis_pdf_next(const BSample& s) { return s.kind == SK_EVAL ? s.pdf : 0.0f; }
// mis_power_heuristic, render_units.hpp:1427-1432
static inline float power_heuristic(float pf, float pg) {
    float f2 = pf * pf, g2 = pg * pg;
    return f2 / (f2 + g2 + 1e-10f);
}
// compute_pdf_w_from_area, render_units.hpp:1440-1448
static inline float pdf_w_from_area(float dist, float area, float c) {
    if (area < 1e-6f || c < 1e-6f) return 0.0f;
    float d2 = dist * dist;
    return (d2 / area) / c;
}
// inverse_square_factor, render_units.hpp:1393-1397
static inline float inv_sq(float d) {
    if (d < 1e-6f) return 0.0f;
    float d2 = d * d;
    return 1.0f / d2;
}

// ---------------------------------------------------------------- lights --
struct LSample {  // light.hpp:100-119
    V3 pos, normal = v3(0, 0, 1), emission;
    float pdf = 0.0f, dist = 0.0f;
    V3 dir = v3(0, 0, 1);
};
// sampleTrianglePoint, light.hpp:146-165
static inline V3 sample_tri_point(V3 v0, V3 v1, V3 v2, float u1, float u2) {
    float su1 = std::sqrt(u1);
    float b0 = 1.0f - su1;
    float b1 = u2 * su1;
    float b2 = 1.0f - b0 - b1;
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 d1, d2;
    float l1, l2;
    normalize_with_length(e1, d1, l1);
    normalize_with_length(e2, d2, l2);
    return add(add(v0, mul(d1, l1 * b1)), mul(d2, l2 * b2));
}
// triangleArea, light.hpp:172-180
static inline float tri_area(V3 v0, V3 v1, V3 v2) {
    V3 c = cross(sub(v1, v0), sub(v2, v0));
    return std::sqrt(len2(c)) * 0.5f;
}
// sampleLightTyped, light.hpp:187-335
static LSample sample_light(const Light& l, V3 p, float u1, float u2) {
    LSample s;
    s.emission = l.emission;
    switch (l.type) {
        case L_POINT: {
            if (l.radius > 1e-6f) {
                float z = 1.0f - 2.0f * u1;
                float r = std::sqrt(fmax_std(0.0f, 1.0f - z * z));
                float phi = 2.0f * PI_F * u2;
                V3 od = v3(r * std::cos(phi), r * std::sin(phi), z);
                s.pos = add(l.pos, mul(od, l.radius));
                s.normal = od;
            } else {
                s.pos = l.pos;
                s.normal = v3(0, 1, 0);
            }
            normalize_with_length(sub(s.pos, p), s.dir, s.dist);
            if (l.radius < 1e-6f) {
                float f = inv_sq(s.dist);
                s.emission = v3(l.emission.x * f, l.emission.y * f, l.emission.z * f);
                s.pdf = 1.0f;
            } else {
                float cl = dot(s.normal, neg(s.dir));
                s.pdf = (cl < 1e-6f) ? 0.0f : pdf_w_from_area(s.dist, l.area, cl);
            }
            break;
        }
        case L_SUN: {
            s.dir = neg(l.normal);
            const float sd = 1000000.0f;
            s.pos = add(p, mul(s.dir, sd));
            s.normal = l.normal;
            s.dist = sd;
            s.pdf = 1.0f;
            break;
        }
        case L_SPOT: {
            s.pos = l.pos;
            s.normal = l.normal;
            normalize_with_length(sub(s.pos, p), s.dir, s.dist);
            float f = inv_sq(s.dist);
            s.emission = v3(l.emission.x * f, l.emission.y * f, l.emission.z * f);
            float ca = dot(l.normal, neg(s.dir));
            float cc = std::cos(l.spot_angle * 0.5f);
            if (ca < cc) {
                s.emission = v3(0, 0, 0);
            } else {
                float blend = l.spot_blend;
                if (blend > 0.0f) {
                    float t = (ca - cc) / (1.0f - cc);
                    float fo = fmin_std(1.0f, t / blend);
                    s.emission = mul(s.emission, fo);
                }
            }
            s.pdf = 1.0f;
            break;
        }
        case L_AREA: {
            float lx, ly;
            if (l.shape == S_DISK || l.shape == S_ELLIPSE) {
                float rs = std::sqrt(u1);
                float th = 2.0f * PI_F * u2;
                lx = l.sx * 0.5f * rs * std::cos(th);
                ly = l.sy * 0.5f * rs * std::sin(th);
            } else {
                lx = l.sx * (u1 - 0.5f);
                ly = l.sy * (u2 - 0.5f);
            }
            s.pos = add(add(l.pos, mul(l.right, lx)), mul(l.up, ly));
            s.normal = l.normal;
            normalize_with_length(sub(s.pos, p), s.dir, s.dist);
            float cl = dot(s.normal, neg(s.dir));
            s.pdf = (cl < 1e-6f) ? 0.0f : pdf_w_from_area(s.dist, l.area, cl);
            break;
        }
        default: {
            s.pos = sample_tri_point(l.v0, l.v1, l.v2, u1, u2);
            s.normal = l.normal;
            normalize_with_length(sub(s.pos, p), s.dir, s.dist);
            float cl = std::fabs(dot(s.normal, neg(s.dir)));
            s.pdf = (cl < 1e-6f) ? 0.0f : pdf_w_from_area(s.dist, l.area, cl);
            break;
        }
    }
    return s;
}
// pdfLightSampleTyped, light.hpp:344-358
static inline float pdf_light(const Light& l, V3 p, V3 lp, V3 ln) {
    V3 d;
    float dist;
    normalize_with_length(sub(lp, p), d, dist);
    float cl = std::fabs(dot(ln, neg(d)));
    if (cl < 1e-6f || l.area < 1e-6f) return 0.0f;
    return pdf_w_from_area(dist, l.area, cl);
}
// area_ratio, render_units.hpp:1401-1404
static inline float area_ratio(float num, float den) { return den < 1e-6f ? 0.0f : num / den; }

// SceneLights::buildFromScene, scene_lights.hpp:56-122
static void build_scene_lights(Scene& s) {
    SceneLights& sl = s.sl;
    sl.lights.clear();
    sl.total_area = 0.0f;
    s.tri_light.assign(s.meshes.size(), std::vector<int>());
    for (size_t mi = 0; mi < s.meshes.size(); ++mi) {
        const Mesh& m = s.meshes[mi];
        s.tri_light[mi].assign(m.tris.size(), -1);
        // node-tree emission at uv (0, 0), scene_lights.hpp:64-67
        const V3 em = mat_emission(m.mat, 0.0f, 0.0f);
        if (!is_emissive(em)) continue;
        for (size_t ti = 0; ti < m.tris.size(); ++ti) {
            const Tri& t = m.tris[ti];
            Light l;
            l.v0 = m.verts[t.i0]; l.v1 = m.verts[t.i1]; l.v2 = m.verts[t.i2];
            l.normal = t.fn;
            l.emission = em;
            l.area = tri_area(l.v0, l.v1, l.v2);
            l.mesh_index = (int)mi;
            l.tri_index = (int)ti;
            l.pos = divs(add(add(l.v0, l.v1), l.v2), 3.0f);
            if (l.area > 1e-6f) {
                s.tri_light[mi][ti] = (int)sl.lights.size();
                sl.lights.push_back(l);
                sl.total_area += l.area;
            }
        }
    }
    sl.native_start = (int)sl.lights.size();
    for (const Light& nl : s.native) {
        Light l = nl;
        if (l.area < 1e-6f) l.area = 1.0f;
        sl.lights.push_back(l);
        sl.total_area += l.area;
    }
    sl.cdf.resize(sl.lights.size());
    float cum = 0.0f;
    for (size_t i = 0; i < sl.lights.size(); ++i) {
        cum += sl.lights[i].area;
        sl.cdf[i] = area_ratio(cum, sl.total_area);
    }
}
// selectLight, scene_lights.hpp:130-144
static inline int select_light(const SceneLights& sl, float u, float& prob) {
    if (sl.lights.empty()) { prob = 0.0f; return -1; }
    int idx = (int)(std::lower_bound(sl.cdf.begin(), sl.cdf.end(), u) - sl.cdf.begin());
    idx = std::min(idx, (int)sl.lights.size() - 1);
    prob = area_ratio(sl.lights[idx].area, sl.total_area);
    return idx;
}
static inline float light_select_pdf(const SceneLights& sl, int i) {  // scene_lights.hpp:150-154
    if (i < 0 || i >= (int)sl.lights.size()) return 0.0f;
    return area_ratio(sl.lights[i].area, sl.total_area);
}
static inline int find_native_light_index(const SceneLights& sl, int ni) {  // scene_lights.hpp:161-164
    int idx = sl.native_start + ni;
    return (idx >= 0 && idx < (int)sl.lights.size()) ? idx : -1;
}

// ----------------------------------------------------------------- volume --
// volume/phase.hpp:32-77
static inline float hg_eval(float g, float c) {
    float gg = g * g;
    float den = 1.0f + gg - 2.0f * g * c;
    float d = fmax_std(den, 1e-8f);
    const float INV4PI = 1.0f / (4.0f * PI_F);
    return INV4PI * (1.0f - gg) / (d * std::sqrt(d));
}
static inline V3 hg_sample(float g, V3 wo, float u1, float u2, float& pdf) {
    float ct;
    if (std::fabs(g) < 1e-3f) {
        ct = 1.0f - 2.0f * u1;
    } else {
        float gg = g * g;
        float sq = (1.0f - gg) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f + gg - sq * sq) / (2.0f * g);
        ct = std::min(std::max(ct, -1.0f), 1.0f);
    }
    float st = std::sqrt(fmax_std(0.0f, 1.0f - ct * ct));
    float phi = 2.0f * PI_F * u2;
    V3 t, b;
    onb(wo, t, b);
    V3 loc = v3(st * std::cos(phi), st * std::sin(phi), ct);
    V3 wi = v3(t.x * loc.x + b.x * loc.y + wo.x * loc.z,
               t.y * loc.x + b.y * loc.y + wo.y * loc.z,
               t.z * loc.x + b.z * loc.y + wo.z * loc.z);
    pdf = hg_eval(g, ct);
    return wi;
}
// volume/transmittance.hpp:35-61
static bool ray_aabb_segment(const Ray& r, float tmax, const float* bmin, const float* bmax, float& tin, float& tout) {
    float ro[3] = {r.o.x, r.o.y, r.o.z}, rd[3] = {r.d.x, r.d.y, r.d.z};
    float t0 = 0.0f, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        if (std::fabs(rd[a]) < 1e-12f) {
            if (ro[a] < bmin[a] || ro[a] > bmax[a]) return false;
            continue;
        }
        float inv = 1.0f / rd[a];
        float ta = (bmin[a] - ro[a]) * inv;
        float tb = (bmax[a] - ro[a]) * inv;
        if (ta > tb) std::swap(ta, tb);
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (t0 > t1) return false;
    }
    tin = fmax_std(0.0f, t0);
    tout = t1;
    return tout > tin;
}
// volume/transmittance.hpp:298-352
static float grid_trilinear(const Volume& v, float wx, float wy, float wz) {
    float u = (wx - v.wmin[0]) / fmax_std(1e-12f, v.wmax[0] - v.wmin[0]);
    float vv = (wy - v.wmin[1]) / fmax_std(1e-12f, v.wmax[1] - v.wmin[1]);
    float w = (wz - v.wmin[2]) / fmax_std(1e-12f, v.wmax[2] - v.wmin[2]);
    if (u < 0.0f || u > 1.0f || vv < 0.0f || vv > 1.0f || w < 0.0f || w > 1.0f) return 0.0f;
    int nx = v.dims[0], ny = v.dims[1], nz = v.dims[2];
    float fx = u * (float)nx - 0.5f, fy = vv * (float)ny - 0.5f, fz = w * (float)nz - 0.5f;
    if (fx < 0.0f || fx > (float)(nx - 1) || fy < 0.0f || fy > (float)(ny - 1) || fz < 0.0f || fz > (float)(nz - 1))
        return 0.0f;
    int ix = std::min(std::max((int)fx, 0), nx - 2);
    int iy = std::min(std::max((int)fy, 0), ny - 2);
    int iz = std::min(std::max((int)fz, 0), nz - 2);
    float tx = fx - (float)ix, ty = fy - (float)iy, tz = fz - (float)iz;
    auto at = [&](int i, int j, int k) {
        return v.grid[(size_t)i + (size_t)nx * (size_t)j + (size_t)nx * (size_t)ny * (size_t)k];
    };
    float c000 = at(ix, iy, iz), c100 = at(ix + 1, iy, iz), c010 = at(ix, iy + 1, iz), c110 = at(ix + 1, iy + 1, iz);
    float c001 = at(ix, iy, iz + 1), c101 = at(ix + 1, iy, iz + 1), c011 = at(ix, iy + 1, iz + 1),
          c111 = at(ix + 1, iy + 1, iz + 1);
    float c00 = c000 * (1.0f - tx) + c100 * tx;
    float c10 = c010 * (1.0f - tx) + c110 * tx;
    float c01 = c001 * (1.0f - tx) + c101 * tx;
    float c11 = c011 * (1.0f - tx) + c111 * tx;
    float c0 = c00 * (1.0f - ty) + c10 * ty;
    float c1 = c01 * (1.0f - ty) + c11 * ty;
    return c0 * (1.0f - tz) + c1 * tz;
}
static inline void grid_march_params(const Volume& v, float seg, float& step, int& n) {
    float ext = std::max({v.wmax[0] - v.wmin[0], v.wmax[1] - v.wmin[1], v.wmax[2] - v.wmin[2]});
    int md = std::max({v.dims[0], v.dims[1], v.dims[2]});
    float vox = ext / std::max(1, md);
    step = fmax_std(1e-4f, vox * 0.5f);
    n = std::min(2048, (int)(seg / step) + 1);
}
// transmittance_along_segment, transmittance.hpp:77-128
static V3 transmittance_seg(const Scene& s, const Ray& r, float tmax) {
    float tr = 0, tg = 0, tb = 0;
    for (const Mesh& m : s.meshes) {
        const Volume& v = m.mat.vol;
        if (!v.present() || !v.has_grid()) continue;
        float tin = 0, tout = 0;
        if (!ray_aabb_segment(r, tmax, v.wmin, v.wmax, tin, tout)) continue;
        float seg = tout - tin;
        if (seg <= 0.0f) continue;
        float step; int n;
        grid_march_params(v, seg, step, n);
        for (int i = 0; i < n; ++i) {
            float t = tin + ((float)i + 0.5f) * step;
            if (t > tout) break;
            float d = grid_trilinear(v, r.o.x + r.d.x * t, r.o.y + r.d.y * t, r.o.z + r.d.z * t);
            if (d <= 0.0f) continue;
            float sa = v.density * d * step;
            tr += sa * (1.0f - v.color.x);
            tg += sa * (1.0f - v.color.y);
            tb += sa * (1.0f - v.color.z);
        }
    }
    return v3(std::exp(-tr), std::exp(-tg), std::exp(-tb));
}
// extinction_transmittance_along_segment, transmittance.hpp:145-209
static V3 ext_transmittance_seg(const Scene& s, const Ray& r, float tmax) {
    float tr = 0, tg = 0, tb = 0;
    for (const Mesh& m : s.meshes) {
        const Volume& v = m.mat.vol;
        if (!v.present() || !v.has_grid()) continue;
        float tin = 0, tout = 0;
        if (!ray_aabb_segment(r, tmax, v.wmin, v.wmax, tin, tout)) continue;
        float seg = tout - tin;
        if (seg <= 0.0f) continue;
        float step; int n;
        grid_march_params(v, seg, step, n);
        float er = 1.0f + v.absorption.x, eg = 1.0f + v.absorption.y, eb = 1.0f + v.absorption.z;
        for (int i = 0; i < n; ++i) {
            float t = tin + ((float)i + 0.5f) * step;
            if (t > tout) break;
            float d = grid_trilinear(v, r.o.x + r.d.x * t, r.o.y + r.d.y * t, r.o.z + r.d.z * t);
            if (d <= 0.0f) continue;
            float base = v.density * d * step;
            tr += base * er; tg += base * eg; tb += base * eb;
        }
    }
    return v3(std::exp(-tr), std::exp(-tg), std::exp(-tb));
}
// sample_volume_distance, transmittance.hpp:220-289
struct VolEvent { bool happened = false; float t = 0.0f; const Volume* vol = nullptr; };
static VolEvent sample_volume_distance(const Scene& s, const Ray& r, float tmax, Rng& rng) {
    VolEvent out;
    float best = tmax;
    for (const Mesh& m : s.meshes) {
        const Volume& v = m.mat.vol;
        if (!v.present() || !v.has_grid()) continue;
        if (v.density <= 0.0f || v.grid_max <= 0.0f) continue;
        float tin = 0, tout = 0;
        if (!ray_aabb_segment(r, tmax, v.wmin, v.wmax, tin, tout)) continue;
        if (tin >= best) continue;
        float seg_end = fmin_std(tout, best);
        if (seg_end <= tin) continue;
        float emax = std::max({1.0f + v.absorption.x, 1.0f + v.absorption.y, 1.0f + v.absorption.z});
        float smax = v.density * v.grid_max * emax;
        if (smax <= 0.0f) continue;
        float t = tin;
        bool sc = false;
        for (int it = 0; it < 4096; ++it) {
            float us = fmax_std(1e-12f, 1.0f - randf(rng));
            t += -std::log(us) / smax;
            if (t >= seg_end) break;
            float d = grid_trilinear(v, r.o.x + r.d.x * t, r.o.y + r.d.y * t, r.o.z + r.d.z * t);
            float st = v.density * d * emax;
            if (randf(rng) * smax < st) { sc = true; break; }
        }
        if (sc && t < best) { best = t; out.happened = true; out.t = t; out.vol = &v; }
    }
    return out;
}

// ------------------------------------------------------------ integrators --
// Per-sample output.  `strat` receives the strategy-separated contributions
// (0 = emission/environment/light hit via the BSDF strategy, 1 = NEE); their
// sum equals the returned radiance term by term.
struct SampleOut {
    V3 L;
    V3 strat[2];
    int rays_closest = 0, rays_shadow = 0;
};
static inline void acc(V3& L, V3 c) { L = add(L, c); }

static inline V3 env_color(const Scene& s) { return mul(s.env_color, s.env_strength); }

// traceSimple, path_tracer.hpp:128-198
static V3 trace_simple(const Scene& s, Ray ray, int max_depth, Rng& rng, SampleOut& so) {
    V3 result, thr = v3(1, 1, 1);
    float tmin = 0.0f;
    for (int depth = 0; depth < max_depth; ++depth) {
        Hit hit; LightHit lh;
        intersect_skip_vb(s, ray, tmin, depth, hit, lh);
        so.rays_closest++;
        if (lh.hit && (!hit.hit || lh.t < hit.t)) { V3 c = cmul(lh.emission, thr); acc(result, c); acc(so.strat[0], c); break; }
        if (!hit.hit) { V3 c = cmul(env_color(s), thr); acc(result, c); acc(so.strat[0], c); break; }
        MatParams mat = mat_params(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 em = mat_emission(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        { V3 c = cmul(em, thr); acc(result, c); acc(so.strat[0], c); }
        V3 n = hit.normal, wo = neg(ray.d);
        bool front = dot(wo, n) > 0;
        V3 sn = n;
        if (mat.transmission < 0.5f && !front) sn = neg(n);
        V3 smp_n = (mat.transmission > 0.5f) ? n : sn;
        float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
        BSample bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
        if (!bs.valid()) break;
        V3 w;
        if (!throughput_update(bs, smp_n, w)) break;
        thr = cmul(thr, w);
        if (depth >= 3) {
            float mx = std::max({thr.x, thr.y, thr.z});
            float rr = fmin_std(mx, 0.95f);
            if (randf(rng) > rr) break;
            thr = mul(thr, 1.0f / rr);
        }
        if (!std::isfinite(thr.x) || !std::isfinite(thr.y) || !std::isfinite(thr.z)) break;
        ray.o = hit.point; ray.d = bs.wi;
        tmin = RAY_T_MIN;
    }
    return result;
}

// traceNEE, path_tracer.hpp:292-397
static V3 trace_nee(const Scene& s, Ray ray, int max_depth, Rng& rng, SampleOut& so) {
    const SceneLights& sl = s.sl;
    V3 result, thr = v3(1, 1, 1);
    float tmin = 0.0f;
    for (int depth = 0; depth < max_depth; ++depth) {
        Hit hit; LightHit lh;
        intersect_skip_vb(s, ray, tmin, depth, hit, lh);
        so.rays_closest++;
        if (lh.hit && (!hit.hit || lh.t < hit.t)) { V3 c = cmul(lh.emission, thr); acc(result, c); acc(so.strat[0], c); break; }
        if (!hit.hit) { V3 c = cmul(env_color(s), thr); acc(result, c); acc(so.strat[0], c); break; }
        MatParams mat = mat_params(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 em = mat_emission(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 n = hit.normal, wo = neg(ray.d);
        bool front = dot(wo, n) > 0;
        V3 sn = n;
        if (mat.transmission < 0.5f && !front) sn = neg(n);
        bool has_em = is_emissive(em);
        if (has_em && depth == 0) { V3 c = cmul(em, thr); acc(result, c); acc(so.strat[0], c); }
        bool transm = mat.transmission > 0.5f;
        if (!sl.lights.empty() && !has_em && !transm) {
            float sp;
            int li = select_light(sl, randf(rng), sp);
            if (li >= 0 && sp > 1e-6f) {
                const Light& l = sl.lights[li];
                float a1 = randf(rng), a2 = randf(rng);
                LSample ls = sample_light(l, hit.point, a1, a2);
                if (ls.pdf > MIN_PDF) {
                    float nl = dot(sn, ls.dir);
                    if (nl > 1e-6f) {
                        Ray sr; sr.o = hit.point; sr.d = ls.dir;
                        so.rays_shadow++;
                        if (!is_shadowed_skip_vb(s, sr, ls.dist)) {
                            V3 f = eval_bsdf(mat, wo, ls.dir, sn);
                            float pl = ls.pdf * sp;
                            V3 bw = bsdf_weight(f, nl, pl);
                            V3 contrib = cmul(ls.emission, bw);
                            V3 c = cmul(contrib, thr);
                            acc(result, c); acc(so.strat[1], c);
                        }
                    }
                }
            }
        }
        V3 smp_n = (mat.transmission > 0.5f) ? n : sn;
        float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
        BSample bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
        if (!bs.valid()) break;
        V3 w;
        if (!throughput_update(bs, smp_n, w)) break;
        thr = cmul(thr, w);
        if (depth >= 3) {
            float mx = std::max({thr.x, thr.y, thr.z});
            float rr = fmin_std(mx, 0.95f);
            if (randf(rng) > rr) break;
            thr = mul(thr, 1.0f / rr);
        }
        if (!std::isfinite(thr.x) || !std::isfinite(thr.y) || !std::isfinite(thr.z)) break;
        ray.o = hit.point; ray.d = bs.wi;
        tmin = RAY_T_MIN;
    }
    return result;
}

// traceMIS, path_tracer.hpp:406-570, and traceVolumeMIS, :586-835
// (volume = true).  The MIS emission/NEE arithmetic is identical in both;
// the volume variant adds the delta-tracking step and multiplies surface NEE
// by the shadow-ray extinction transmittance.
static V3 trace_mis(const Scene& s, Ray ray, int max_depth, Rng& rng, SampleOut& so, bool volume) {
    const SceneLights& sl = s.sl;
    V3 result, thr = v3(1, 1, 1);
    float last_pdf = 0.0f;
    float tmin = 0.0f;
    for (int depth = 0; depth < max_depth; ++depth) {
        Hit hit; LightHit lh;
        intersect_skip_vb(s, ray, tmin, depth, hit, lh);
        so.rays_closest++;
        if (volume) {
            float ts = 1e6f;
            if (hit.hit) ts = fmin_std(ts, hit.t);
            if (lh.hit) ts = fmin_std(ts, lh.t);
            VolEvent vs = sample_volume_distance(s, ray, ts, rng);
            if (vs.happened && vs.vol != nullptr) {
                const Volume& v = *vs.vol;
                float ar = v.color.x / (1.0f + v.absorption.x);
                float ag = v.color.y / (1.0f + v.absorption.y);
                float ab = v.color.z / (1.0f + v.absorption.z);
                thr = cmul(thr, v3(ar, ag, ab));
                V3 x = ray_at(ray, vs.t);
                V3 wo = neg(ray.d);
                float g = v.anisotropy;
                if (!sl.lights.empty()) {
                    float sp;
                    int li = select_light(sl, randf(rng), sp);
                    if (li >= 0 && sp > 1e-6f) {
                        const Light& l = sl.lights[li];
                        float a1 = randf(rng), a2 = randf(rng);
                        LSample ls = sample_light(l, x, a1, a2);
                        if (ls.pdf > MIN_PDF) {
                            Ray sr; sr.o = x; sr.d = ls.dir;
                            so.rays_shadow++;
                            if (!is_shadowed_skip_vb(s, sr, ls.dist)) {
                                V3 T = ext_transmittance_seg(s, sr, ls.dist);
                                float ct = dot(wo, ls.dir);
                                float pe = hg_eval(g, ct);
                                float pp = hg_eval(g, ct);
                                float pl = ls.pdf * sp;
                                float mw = power_heuristic(pl, pp);
                                V3 pw = bsdf_weight(v3(pe, pe, pe), 1.0f, pl);
                                V3 c = cmul(mul(pw, mw), cmul(ls.emission, T));
                                V3 cc = cmul(c, thr);
                                acc(result, cc); acc(so.strat[1], cc);
                            }
                        }
                    }
                }
                float a1 = randf(rng), a2 = randf(rng);
                float ppdf;
                V3 wi = hg_sample(g, wo, a1, a2, ppdf);
                last_pdf = ppdf;
                ray.o = x; ray.d = wi;
                tmin = RAY_T_MIN;
                if (depth >= 3) {
                    float mx = std::max({thr.x, thr.y, thr.z});
                    float rr = fmin_std(mx, 0.95f);
                    if (randf(rng) > rr) break;
                    thr = mul(thr, 1.0f / rr);
                }
                if (!std::isfinite(thr.x) || !std::isfinite(thr.y) || !std::isfinite(thr.z)) break;
                continue;
            }
        }
        if (lh.hit && (!hit.hit || lh.t < hit.t)) {
            float lp = 0.0f;
            if (!sl.lights.empty() && lh.light_index >= 0) {
                int sli = find_native_light_index(sl, lh.light_index);
                if (sli >= 0) {
                    float spb = light_select_pdf(sl, sli);
                    if (spb > 0.0f) lp = pdf_light(sl.lights[sli], ray.o, lh.point, lh.normal) * spb;
                }
            }
            float bp = last_pdf;
            float mw = (bp > MIN_PDF && lp > MIN_PDF) ? power_heuristic(bp, lp) : 1.0f;
            V3 c = mul(cmul(lh.emission, thr), mw);
            acc(result, c); acc(so.strat[0], c);
            break;
        }
        if (!hit.hit) { V3 c = cmul(env_color(s), thr); acc(result, c); acc(so.strat[0], c); break; }
        MatParams mat = mat_params(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 em = mat_emission(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 n = hit.normal, wo = neg(ray.d);
        bool front = dot(wo, n) > 0;
        V3 sn = n;
        if (mat.transmission < 0.5f && !front) sn = neg(n);
        bool has_em = is_emissive(em);
        if (has_em) {
            float mw = 1.0f;
            if (!sl.lights.empty() && last_pdf > MIN_PDF) {
                int li = s.tri_light[hit.mesh][hit.tri];
                if (li >= 0) {
                    float spb = light_select_pdf(sl, li);
                    if (spb > 0.0f) {
                        float lp = pdf_light(sl.lights[li], ray.o, hit.point, hit.normal) * spb;
                        if (lp > MIN_PDF) mw = power_heuristic(last_pdf, lp);
                    }
                }
            }
            V3 c = mul(cmul(em, thr), mw);
            acc(result, c); acc(so.strat[0], c);
        }
        bool transm = mat.transmission > 0.5f;
        if (!sl.lights.empty() && !has_em && !transm) {
            float sp;
            int li = select_light(sl, randf(rng), sp);
            if (li >= 0 && sp > 1e-6f) {
                const Light& l = sl.lights[li];
                float a1 = randf(rng), a2 = randf(rng);
                LSample ls = sample_light(l, hit.point, a1, a2);
                if (ls.pdf > MIN_PDF) {
                    float nl = dot(sn, ls.dir);
                    if (nl > 1e-6f) {
                        Ray sr; sr.o = hit.point; sr.d = ls.dir;
                        so.rays_shadow++;
                        if (!is_shadowed_skip_vb(s, sr, ls.dist)) {
                            V3 f = eval_bsdf(mat, wo, ls.dir, sn);
                            float pl = ls.pdf * sp;
                            float pb = pdf_bsdf(mat, wo, ls.dir, sn);
                            float mw = power_heuristic(pl, pb);
                            V3 bw = bsdf_weight(f, nl, pl);
                            V3 c;
                            if (volume) {
                                V3 T = ext_transmittance_seg(s, sr, ls.dist);
                                c = cmul(mul(bw, mw), cmul(ls.emission, T));
                            } else {
                                c = cmul(mul(bw, mw), ls.emission);
                            }
                            V3 cc = cmul(c, thr);
                            acc(result, cc); acc(so.strat[1], cc);
                        }
                    }
                }
            }
        }
        V3 smp_n = (mat.transmission > 0.5f) ? n : sn;
        float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
        BSample bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
        if (!bs.valid()) break;
        V3 w;
        if (!throughput_update(bs, smp_n, w)) break;
        thr = cmul(thr, w);
        last_pdf = mis_pdf_next(bs);
        if (depth >= 3) {
            float mx = std::max({thr.x, thr.y, thr.z});
            float rr = fmin_std(mx, 0.95f);
            if (randf(rng) > rr) break;
            thr = mul(thr, 1.0f / rr);
        }
        if (!std::isfinite(thr.x) || !std::isfinite(thr.y) || !std::isfinite(thr.z)) break;
        ray.o = hit.point; ray.d = bs.wi;
        tmin = RAY_T_MIN;
    }
    return result;
}

// traceVolumeSimple, path_tracer.hpp:209-283
static V3 trace_volume_simple(const Scene& s, Ray ray, int max_depth, Rng& rng, SampleOut& so) {
    V3 result, thr = v3(1, 1, 1);
    float tmin = 0.0f;
    for (int depth = 0; depth < max_depth; ++depth) {
        Hit hit; LightHit lh;
        intersect_skip_vb(s, ray, tmin, depth, hit, lh);
        so.rays_closest++;
        float ts = 1e6f;
        if (hit.hit) ts = fmin_std(ts, hit.t);
        if (lh.hit) ts = fmin_std(ts, lh.t);
        thr = cmul(thr, transmittance_seg(s, ray, ts));
        if (lh.hit && (!hit.hit || lh.t < hit.t)) { V3 c = cmul(lh.emission, thr); acc(result, c); acc(so.strat[0], c); break; }
        if (!hit.hit) { V3 c = cmul(env_color(s), thr); acc(result, c); acc(so.strat[0], c); break; }
        MatParams mat = mat_params(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 em = mat_emission(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        { V3 c = cmul(em, thr); acc(result, c); acc(so.strat[0], c); }
        V3 n = hit.normal, wo = neg(ray.d);
        bool front = dot(wo, n) > 0;
        V3 sn = n;
        if (mat.transmission < 0.5f && !front) sn = neg(n);
        V3 smp_n = (mat.transmission > 0.5f) ? n : sn;
        float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
        BSample bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
        if (!bs.valid()) break;
        V3 w;
        if (!throughput_update(bs, smp_n, w)) break;
        thr = cmul(thr, w);
        if (depth >= 3) {
            float mx = std::max({thr.x, thr.y, thr.z});
            float rr = fmin_std(mx, 0.95f);
            if (randf(rng) > rr) break;
            thr = mul(thr, 1.0f / rr);
        }
        if (!std::isfinite(thr.x) || !std::isfinite(thr.y) || !std::isfinite(thr.z)) break;
        ray.o = hit.point; ray.d = bs.wi;
        tmin = RAY_T_MIN;
    }
    return result;
}

enum Algo { A_SIMPLE = 0, A_NEE = 1, A_MIS = 2, A_VOLUME_SIMPLE = 3, A_VOLUME_MIS = 4 };
// algorithm string dispatch, pybind_renderer.hpp:301-327 (unknown -> NEE)
static int algo_of(const char* a) {
    std::string s = a ? a : "";
    if (s == "simple") return A_SIMPLE;
    if (s == "mis") return A_MIS;
    if (s == "volume_simple") return A_VOLUME_SIMPLE;
    if (s == "volume_mis") return A_VOLUME_MIS;
    return A_NEE;
}
static V3 trace(const Scene& s, int algo, const Ray& r, int md, Rng& rng, SampleOut& so) {
    switch (algo) {
        case A_SIMPLE: return trace_simple(s, r, md, rng, so);
        case A_MIS: return trace_mis(s, r, md, rng, so, false);
        case A_VOLUME_SIMPLE: return trace_volume_simple(s, r, md, rng, so);
        case A_VOLUME_MIS: return trace_mis(s, r, md, rng, so, true);
        default: return trace_nee(s, r, md, rng, so);
    }
}

// Camera ray of render_tile, pybind_renderer.hpp:253-293.
struct RayGen {
    V3 pos, fwd, right, up;
    float aspect, scale;
    int fw, fh;
};
static RayGen make_raygen(const Scene& s, int fw, int fh) {
    RayGen g;
    g.pos = s.cam.pos; g.fwd = s.cam.forward; g.right = s.cam.right; g.up = s.cam.up;
    g.aspect = (float)fw / (float)fh;
    g.scale = std::tan(s.cam.fov * 0.5f);
    g.fw = fw; g.fh = fh;
    return g;
}
static inline V3 cam_dir(const RayGen& g, float nx, float ny) {
    V3 d = v3(g.fwd.x + g.right.x * (nx * g.scale) + g.up.x * (ny * g.scale),
              g.fwd.y + g.right.y * (nx * g.scale) + g.up.y * (ny * g.scale),
              g.fwd.z + g.right.z * (nx * g.scale) + g.up.z * (ny * g.scale));
    return make_dir_or(d, g.fwd);
}
// Seeds the RNG for (gx, gy, sample) and returns the (possibly jittered)
// camera ray, drawing the two jitter numbers when samples_in_call > 1.
static Ray camera_ray(const RayGen& g, int gx, int gy, uint32_t sidx, int samples_in_call, Rng& rng) {
    float ndc_x = (2.0f * (gx + 0.5f) / g.fw - 1.0f) * g.aspect;
    float ndc_y = 1.0f - 2.0f * (gy + 0.5f) / g.fh;
    seed_xyz(rng, (uint32_t)gx, (uint32_t)gy, sidx);
    Ray r;
    r.o = g.pos;
    if (samples_in_call > 1) {
        float jx = (randf(rng) - 0.5f) / g.fw;
        float jy = (randf(rng) - 0.5f) / g.fh;
        r.d = cam_dir(g, ndc_x + jx, ndc_y + jy);
    } else {
        r.d = cam_dir(g, ndc_x, ndc_y);
    }
    return r;
}

// ------------------------------------------------------------ diagnostics --
// Path records of the diagnostic tracers (traceSimpleWithDiagnostics,
// traceNEEWithDiagnostics, traceMISWithDiagnostics; path_tracer.hpp:880-1514)
// and the per-pixel statistics of PixelPathData
// (diagnostics/pixel_path_data.hpp), restated over std::vector and written
// out at the end in the reference's byte layout for comparison.
enum DBsdf : uint8_t { DB_DIFFUSE = 0, DB_GLOSSY = 1, DB_MIRROR = 2, DB_GLASS = 3, DB_EMISSION = 4, DB_ENV = 5 };
enum DLight : uint8_t { DL_UNKNOWN = 0, DL_POINT = 1, DL_AREA = 2, DL_DIRECTIONAL = 3, DL_SPOT = 4, DL_ENV = 5,
                        DL_EMISSIVE = 6 };
enum DStrat : uint8_t { DS_BSDF = 0, DS_NEE = 1, DS_MIS_BSDF = 2, DS_MIS_NEE = 3 };
static const size_t D_MAX_DEPTH = 16;  // path_types.hpp:22

struct DVert {  // PathVertex, path_types.hpp:178-243, plus its geometry
    int32_t obj = -1;
    int16_t mat = -1;
    uint8_t type = DB_DIFFUSE, flags = 0;
    V3 pos, nrm;
};
struct DRec {  // one completed path (PathTrace)
    std::vector<DVert> v;
    V3 contrib;
    uint8_t light_type = DL_UNKNOWN, strategy = DS_BSDF;
};
// PathDiagnosticRecorder, diagnostic_integrator.hpp:87-389
struct DRecorder {
    std::vector<DVert> cur;
    std::vector<DRec> done;
    void vertex(int32_t obj, uint8_t type, bool delta, V3 pos, V3 nrm) {  // :112-128
        if (cur.size() >= D_MAX_DEPTH) return;
        DVert x;
        x.obj = obj; x.mat = 0; x.type = type; x.flags = delta ? 1 : 0; x.pos = pos; x.nrm = nrm;
        cur.push_back(x);
    }
    void complete(uint8_t strat, int32_t light_id, uint8_t ltype, V3 c, V3 lpos) {  // :338-368
        DRec r;
        r.v = cur;
        if (r.v.size() < D_MAX_DEPTH) {
            DVert x;
            x.obj = ltype == DL_ENV ? -1 : -(light_id + 2);
            x.mat = -1; x.type = DB_EMISSION; x.flags = 0; x.pos = lpos; x.nrm = v3(0, 0, 0);
            r.v.push_back(x);
        }
        r.contrib = c; r.light_type = ltype; r.strategy = strat;
        done.push_back(r);
    }
};
// classify_bsdf / is_delta_bsdf, diagnostic_integrator.hpp:24-60
static uint8_t d_classify(const MatParams& m) {
    if (m.transmission > 0.5f) return DB_GLASS;
    if (m.roughness < 0.01f && m.metallic > 0.5f) return DB_MIRROR;
    if (m.metallic > 0.5f || m.roughness < 0.5f) return DB_GLOSSY;
    return DB_DIFFUSE;
}
static bool d_is_delta(const MatParams& m) {
    if (m.transmission > 0.5f && m.roughness < 0.01f) return true;
    if (m.metallic > 0.5f && m.roughness < 0.01f) return true;
    return false;
}
// to_light_source_type, path_types.hpp:160-169
static uint8_t d_light_type(int t) {
    switch (t) {
        case L_EMISSIVE: return DL_AREA;
        case L_POINT: return DL_POINT;
        case L_SUN: return DL_DIRECTIONAL;
        case L_SPOT: return DL_SPOT;
        case L_AREA: return DL_AREA;
        default: return DL_UNKNOWN;
    }
}
static V3 d_far(const Ray& r) { return v3(r.o.x + r.d.x * 1000.0f, r.o.y + r.d.y * 1000.0f, r.o.z + r.d.z * 1000.0f); }

// The three diagnostic tracers.  kind: 0 simple, 1 NEE, 2 MIS.  They use
// intersectScene directly (no volume-boundary skipping) for camera,
// extension and shadow rays.
static V3 trace_diag(const Scene& s, int kind, Ray ray, int max_depth, Rng& rng, DRecorder& rec) {
    const SceneLights& sl = s.sl;
    V3 result, thr = v3(1, 1, 1);
    float last_pdf = 0.0f, tmin = 0.0f;
    for (int depth = 0; depth < max_depth; ++depth) {
        Hit hit = intersect_scene(s, ray, tmin);
        LightHit lh = intersect_native_lights(s, ray, depth);
        if (lh.hit && (!hit.hit || lh.t < hit.t)) {
            if (kind == 1 && depth != 0) break;  // :1060-1091
            uint8_t lt = DL_UNKNOWN;
            if (lh.light_index >= 0 && lh.light_index < (int)s.native.size()) lt = d_light_type(s.native[lh.light_index].type);
            V3 lpos = v3(ray.o.x + ray.d.x * lh.t, ray.o.y + ray.d.y * lh.t, ray.o.z + ray.d.z * lh.t);
            V3 c = cmul(lh.emission, thr);
            if (kind == 2) {  // :1282-1322
                float lp = 0.0f;
                if (!sl.lights.empty() && lh.light_index >= 0) {
                    int sli = find_native_light_index(sl, lh.light_index);
                    if (sli >= 0) {
                        float spb = light_select_pdf(sl, sli);
                        if (spb > 0.0f) lp = pdf_light(sl.lights[sli], ray.o, lh.point, lh.normal) * spb;
                    }
                }
                float mw = (last_pdf > MIN_PDF && lp > MIN_PDF) ? power_heuristic(last_pdf, lp) : 1.0f;
                c = mul(c, mw);
            }
            rec.complete(kind == 2 ? DS_MIS_BSDF : DS_BSDF, lh.light_index, lt, c, lpos);
            acc(result, c);
            break;
        }
        if (!hit.hit) {
            if (kind == 1 && depth != 0) break;  // :1094-1116
            V3 c = cmul(env_color(s), thr);
            rec.complete(kind == 2 ? DS_MIS_BSDF : DS_BSDF, -1, DL_ENV, c, d_far(ray));
            acc(result, c);
            break;
        }
        MatParams mat = mat_params(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        V3 em = mat_emission(s.meshes[hit.mesh].mat, hit.uv[0], hit.uv[1]);
        uint8_t bt = d_classify(mat);
        bool dl = d_is_delta(mat);
        bool has_em = is_emissive(em);
        V3 n = hit.normal, wo = neg(ray.d);
        bool front = dot(wo, n) > 0;
        V3 sn = n;
        if (mat.transmission < 0.5f && !front) sn = neg(n);
        if (kind == 0) {  // :949-985
            if (has_em) {
                V3 c = cmul(em, thr);
                rec.complete(DS_BSDF, hit.mesh, DL_EMISSIVE, c, hit.point);
                acc(result, c);
                break;
            }
            rec.vertex(hit.mesh, bt, dl, hit.point, hit.normal);
        } else {
            if (has_em) {
                if (kind == 1 && depth == 0) {  // :1131-1146
                    V3 c = cmul(em, thr);
                    rec.complete(DS_BSDF, hit.mesh, DL_EMISSIVE, c, hit.point);
                    acc(result, c);
                } else if (kind == 2) {  // :1386-1406
                    float mw = 1.0f;
                    if (!sl.lights.empty() && last_pdf > MIN_PDF) {
                        int li = s.tri_light[hit.mesh][hit.tri];
                        if (li >= 0) {
                            float spb = light_select_pdf(sl, li);
                            if (spb > 0.0f) {
                                float lp = pdf_light(sl.lights[li], ray.o, hit.point, hit.normal) * spb;
                                if (lp > MIN_PDF) mw = power_heuristic(last_pdf, lp);
                            }
                        }
                    }
                    V3 c = mul(cmul(em, thr), mw);
                    rec.complete(DS_MIS_BSDF, hit.mesh, DL_EMISSIVE, c, hit.point);
                    acc(result, c);
                }
            } else {
                rec.vertex(hit.mesh, bt, dl, hit.point, hit.normal);
                bool transm = mat.transmission > 0.5f;
                if (!sl.lights.empty() && !transm) {
                    float sp;
                    int li = select_light(sl, randf(rng), sp);
                    if (li >= 0 && sp > 1e-6f) {
                        const Light& l = sl.lights[li];
                        float a1 = randf(rng), a2 = randf(rng);
                        LSample ls = sample_light(l, hit.point, a1, a2);
                        if (ls.pdf > MIN_PDF) {
                            float nl = dot(sn, ls.dir);
                            if (nl > 1e-6f) {
                                Ray sr; sr.o = hit.point; sr.d = ls.dir;
                                if (!intersect_scene(s, sr, RAY_T_MIN, ls.dist).hit) {
                                    V3 f = eval_bsdf(mat, wo, ls.dir, sn);
                                    float pl = ls.pdf * sp;
                                    V3 nee;
                                    if (kind == 2) {
                                        float pb = pdf_bsdf(mat, wo, ls.dir, sn);
                                        float mw = power_heuristic(pl, pb);
                                        nee = cmul(mul(bsdf_weight(f, nl, pl), mw), ls.emission);
                                    } else {
                                        nee = cmul(bsdf_weight(f, nl, pl), ls.emission);
                                    }
                                    V3 c = cmul(thr, nee);
                                    rec.complete(kind == 2 ? DS_MIS_NEE : DS_NEE, li - sl.native_start,
                                                 d_light_type(l.type), c, ls.pos);
                                    acc(result, c);
                                }
                            }
                        }
                    }
                }
            }
        }
        V3 smp_n = (mat.transmission > 0.5f) ? n : sn;
        float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
        BSample bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
        if (!bs.valid()) break;
        V3 w;
        if (!throughput_update(bs, smp_n, w)) break;
        thr = cmul(thr, w);
        if (kind == 2) last_pdf = mis_pdf_next(bs);
        if (depth >= 3) {
            float mx = std::max({thr.x, thr.y, thr.z});
            float rr = fmin_std(mx, 0.95f);
            if (randf(rng) > rr) break;
            thr = mul(thr, 1.0f / rr);
        }
        if (!std::isfinite(thr.x) || !std::isfinite(thr.y) || !std::isfinite(thr.z)) break;
        ray.o = hit.point; ray.d = bs.wi;
        tmin = RAY_T_MIN;
    }
    return result;
}

// PathTrace::hash (FNV-1a over depth and every vertex field, each widened
// to uint64) and coarse_type, path_types.hpp:321-372
static uint64_t d_hash(const std::vector<DVert>& v) {
    uint64_t h = 14695981039346656037ULL;
    h ^= (uint64_t)v.size();
    h *= 1099511628211ULL;
    for (const DVert& x : v) {
        h ^= (uint64_t)(int64_t)x.obj; h *= 1099511628211ULL;
        h ^= (uint64_t)(int64_t)x.mat; h *= 1099511628211ULL;
        h ^= (uint64_t)x.type; h *= 1099511628211ULL;
        h ^= (uint64_t)x.flags; h *= 1099511628211ULL;
    }
    return h;
}
static uint8_t d_coarse(const std::vector<DVert>& v) {
    uint8_t t = 0;
    for (const DVert& x : v) {
        if (x.flags & 2) t |= 1;
        if (x.flags & 1) t |= 2;
    }
    if (v.size() > 2) t |= 4;
    return t;
}
// PathStatistics::add_sample, path_stats.hpp:63-98 (float Welford, n-1)
struct DStats {
    uint32_t n = 0;
    float mean[3] = {0, 0, 0}, m2[3] = {0, 0, 0}, var[3] = {0, 0, 0};
    void add(V3 c) {
        const float x[3] = {c.x, c.y, c.z};
        ++n;
        if (n == 1) {
            for (int k = 0; k < 3; ++k) { mean[k] = x[k]; m2[k] = 0.0f; var[k] = 0.0f; }
            return;
        }
        float d1[3], d2[3];
        for (int k = 0; k < 3; ++k) d1[k] = x[k] - mean[k];
        for (int k = 0; k < 3; ++k) mean[k] += d1[k] / (float)n;
        for (int k = 0; k < 3; ++k) d2[k] = x[k] - mean[k];
        for (int k = 0; k < 3; ++k) m2[k] += d1[k] * d2[k];
        float nm1 = (float)(n - 1);
        for (int k = 0; k < 3; ++k) var[k] = m2[k] / nm1;
    }
    float total_var() const { return var[0] + var[1] + var[2]; }
};
// PathGroup, path_stats.hpp:133-268: fixed arrays, init_from copies only
// the first `depth` vertices (later ones keep what was there).
struct DGroup {
    uint64_t hash = 0;
    DStats st;
    uint8_t depth = 0, coarse = 0, ltype = DL_UNKNOWN, strat = DS_BSDF;
    DVert v[D_MAX_DEPTH];
    void init_from(const DRec& r) {
        hash = d_hash(r.v);
        depth = (uint8_t)std::min(r.v.size(), D_MAX_DEPTH);
        coarse = d_coarse(r.v);
        ltype = r.light_type;
        strat = r.strategy;
        for (size_t i = 0; i < depth; ++i) v[i] = r.v[i];
        st = DStats();
        st.add(r.contrib);
    }
};
// PixelPathData<G, O>::record_path, pixel_path_data.hpp:53-224
struct DPixel {
    size_t cap_g = 16, cap_o = 4;
    std::vector<DGroup> g;
    size_t n_g = 0, n_o = 0, total = 0, overflow = 0;
    DStats coarse[8], tot;
    std::vector<uint64_t> o_hash;
    std::vector<V3> o_c;
    DPixel(size_t cg = 16, size_t co = 4) : cap_g(cg), cap_o(co), g(cg), o_hash(co, 0), o_c(co) {}
    void record(const DRec& r) {
        ++total;
        tot.add(r.contrib);
        coarse[d_coarse(r.v)].add(r.contrib);
        if (total >= 10) {  // is_outlier, :172-189 (3 sigma on any channel)
            const float x[3] = {r.contrib.x, r.contrib.y, r.contrib.z};
            bool out = false;
            for (int k = 0; k < 3; ++k) {
                float sd = std::sqrt(tot.var[k] + 1e-10f);
                if (std::fabs(x[k] - tot.mean[k]) > 3.0f * sd) out = true;
            }
            if (out) {
                size_t i = n_o % cap_o;
                o_hash[i] = d_hash(r.v);
                o_c[i] = r.contrib;
                ++n_o;
            }
        }
        uint64_t h = d_hash(r.v);
        for (size_t i = 0; i < n_g; ++i) {
            if (g[i].hash == h) { g[i].st.add(r.contrib); return; }
        }
        if (n_g < cap_g) {
            g[n_g].init_from(r);
            ++n_g;
            return;
        }
        size_t mi = 0;  // evict_and_add, :203-224
        float mv = g[0].st.total_var();
        for (size_t i = 1; i < n_g; ++i) {
            float vv = g[i].st.total_var();
            float adj = vv * std::log2((float)(g[i].st.n + 1));
            float madj = mv * std::log2((float)(g[mi].st.n + 1));
            if (adj < madj) { mi = i; mv = vv; }
        }
        ++overflow;
        g[mi].init_from(r);
    }
};
// The film pixel in PixelPathData<16, 4>'s byte layout (9576 bytes).
static const size_t D_PIXEL_BYTES = 9576;
static void d_put_stats(unsigned char* p, const DStats& s) {
    std::memcpy(p, &s.n, 4);
    std::memcpy(p + 4, s.mean, 12);
    std::memcpy(p + 16, s.m2, 12);
    std::memcpy(p + 28, s.var, 12);
}
static void d_serialise(const DPixel& px, unsigned char* out) {
    std::memset(out, 0, D_PIXEL_BYTES);
    for (size_t i = 0; i < 16; ++i) {
        unsigned char* q = out + i * 568;
        const DGroup& G = px.g[i];
        std::memcpy(q, &G.hash, 8);
        d_put_stats(q + 8, G.st);
        q[48] = G.depth; q[49] = G.coarse; q[50] = G.ltype; q[51] = G.strat;
        for (size_t k = 0; k < D_MAX_DEPTH; ++k) {
            std::memcpy(q + 56 + k * 8, &G.v[k].obj, 4);
            std::memcpy(q + 60 + k * 8, &G.v[k].mat, 2);
            q[62 + k * 8] = G.v[k].type;
            q[63 + k * 8] = G.v[k].flags;
            const float pp[3] = {G.v[k].pos.x, G.v[k].pos.y, G.v[k].pos.z};
            const float nn[3] = {G.v[k].nrm.x, G.v[k].nrm.y, G.v[k].nrm.z};
            std::memcpy(q + 184 + k * 12, pp, 12);
            std::memcpy(q + 376 + k * 12, nn, 12);
        }
    }
    for (int t = 0; t < 8; ++t) d_put_stats(out + 9088 + t * 40, px.coarse[t]);
    for (size_t i = 0; i < 4; ++i) {
        std::memcpy(out + 9408 + i * 24, &px.o_hash[i], 8);
        const float c[3] = {px.o_c[i].x, px.o_c[i].y, px.o_c[i].z};
        std::memcpy(out + 9416 + i * 24, c, 12);
    }
    const uint64_t cnt[4] = {px.n_g, px.n_o, px.total, px.overflow};
    std::memcpy(out + 9504, cnt, 32);
    d_put_stats(out + 9536, px.tot);
}

}  // namespace orc

// =========================================================== C interface ==
using namespace orc;

extern "C" {

struct orc_mesh_desc {
    int32_t num_vertices;
    const float* vertices;       // 3 per vertex
    int32_t num_triangles;
    const int32_t* indices;      // 3 per triangle
    const float* tri_normals;    // 9 per triangle or NULL
    int32_t num_tri_normals;     // triangles [0, num_tri_normals) carry normals
    int32_t smooth;
    float base_color[3];
    float metallic, roughness;
    float emission[3];
    float transmission, ior;
    int32_t has_volume;
    float vol_color[3], vol_absorption[3];
    float vol_density, vol_anisotropy;
    int32_t grid_dims[3];
    float grid_min[3], grid_max[3];
    const float* grid;           // dims[0]*dims[1]*dims[2] floats, x fastest, or NULL
    int32_t surface_linked;      // 0 => volume_boundary_only when a volume is present
    int32_t use_nodes;           // "use_nodes"
    const struct orc_node* nodes;  // "node_tree" nodes, links resolved per input
    int32_t num_nodes;
    const float* tri_uvs;        // 6 per triangle or NULL
    int32_t num_tri_uvs;
};
// A node-tree socket: kind 0 none, 1 float, 2 vec3, 3 vec4 (parseSocketValue,
// node_evaluator.cpp:47-69); linked_node NULL when not linked.
struct orc_node_socket {
    const char* name;
    int32_t kind;
    float value[4];
    const char* linked_node;
    const char* linked_socket;
};
struct orc_node {
    const char* name;
    const char* type;
    const orc_node_socket* inputs;
    int32_t num_inputs;
    const orc_node_socket* outputs;
    int32_t num_outputs;
};

struct orc_light_desc {
    int32_t type;                // 1 POINT, 2 SUN, 3 SPOT, 4 AREA
    float position[3];
    int32_t has_direction;
    float direction[3];
    float color[3];
    float energy;
    float radius;
    float spot_size;
    float spot_blend;
    float size, size_y;
    int32_t shape;               // 0 SQUARE 1 RECTANGLE 2 DISK 3 ELLIPSE
    int32_t has_right, has_up;
    float right[3], up[3];
};

void* orc_scene_new(void) { return new Scene(); }
void orc_scene_free(void* p) { delete static_cast<Scene*>(p); }

// Mesh part of loadSceneFromJsonString, pybind_renderer.hpp:1266-1507.
int orc_scene_add_mesh(void* p, const orc_mesh_desc* d) {
    Scene& s = *static_cast<Scene*>(p);
    Mesh m;
    m.verts.resize(d->num_vertices);
    for (int i = 0; i < d->num_vertices; ++i)
        m.verts[i] = v3(d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2]);
    m.tris.resize(d->num_triangles);
    const V3 up = v3(0.0f, 0.0f, 1.0f);
    for (int i = 0; i < d->num_triangles; ++i) {
        Tri& t = m.tris[i];
        t.i0 = d->indices[3 * i]; t.i1 = d->indices[3 * i + 1]; t.i2 = d->indices[3 * i + 2];
        if (t.i0 < 0 || t.i1 < 0 || t.i2 < 0 || t.i0 >= d->num_vertices || t.i1 >= d->num_vertices ||
            t.i2 >= d->num_vertices)
            return -1;
        if (d->tri_normals && i < d->num_tri_normals) {
            const float* n = d->tri_normals + 9 * i;
            t.n0 = make_dir_or(v3(n[0], n[1], n[2]), up);
            t.n1 = make_dir_or(v3(n[3], n[4], n[5]), up);
            t.n2 = make_dir_or(v3(n[6], n[7], n[8]), up);
            t.smooth = d->smooth != 0;
        } else {
            t.smooth = false;
        }
        if (d->tri_uvs && i < d->num_tri_uvs) {  // pybind_renderer.hpp:1330-1337
            const float* q = d->tri_uvs + 6 * (size_t)i;
            t.uv0[0] = q[0]; t.uv0[1] = q[1]; t.uv1[0] = q[2]; t.uv1[1] = q[3]; t.uv2[0] = q[4]; t.uv2[1] = q[5];
            t.has_uv = true;
        }
    }
    Material& mt = m.mat;
    // parseNodeTree (node_evaluator.cpp:82-149), useNodes (pybind_renderer.hpp:1381-1384)
    auto str = [](const char* c) { return std::string(c ? c : ""); };
    auto sockval = [](const orc_node_socket& q) {
        SockVal sv;
        if (q.kind == 1) { sv.type = SV_FLOAT; sv.f = q.value[0]; }
        else if (q.kind == 2) { sv.type = SV_VEC3; sv.v3_ = v3(q.value[0], q.value[1], q.value[2]); }
        else if (q.kind == 3) { sv.type = SV_VEC4; sv.v4 = v3(q.value[0], q.value[1], q.value[2]); }
        return sv;
    };
    for (int i = 0; d->nodes && i < d->num_nodes; ++i) {
        const orc_node& nj = d->nodes[i];
        MatNode nd;
        nd.name = str(nj.name);
        nd.type = str(nj.type);
        for (int k = 0; k < nj.num_inputs; ++k) {
            NodeSocket so;
            so.name = str(nj.inputs[k].name);
            so.dv = sockval(nj.inputs[k]);
            if (nj.inputs[k].linked_node) {
                so.is_linked = true;
                so.linked_node = str(nj.inputs[k].linked_node);
                so.linked_socket = str(nj.inputs[k].linked_socket);
            }
            nd.inputs.push_back(so);
        }
        for (int k = 0; k < nj.num_outputs; ++k) {
            NodeSocket so;
            so.name = str(nj.outputs[k].name);
            so.dv = sockval(nj.outputs[k]);
            nd.outputs.push_back(so);
        }
        mt.tree.nodes.push_back(nd);
    }
    mt.tree.valid = !mt.tree.nodes.empty();
    mt.use_nodes = d->use_nodes != 0 && mt.tree.valid;
    mt.albedo = v3(d->base_color[0], d->base_color[1], d->base_color[2]);
    mt.metallic = d->metallic; mt.roughness = d->roughness;
    mt.emission = v3(d->emission[0], d->emission[1], d->emission[2]);
    mt.transmission = d->transmission; mt.ior = d->ior;
    if (d->has_volume) {
        Volume& v = mt.vol;
        v.color = v3(d->vol_color[0], d->vol_color[1], d->vol_color[2]);
        v.absorption = v3(d->vol_absorption[0], d->vol_absorption[1], d->vol_absorption[2]);
        v.density = d->vol_density;
        v.anisotropy = d->vol_anisotropy;
        if (d->grid && d->grid_dims[0] > 0 && d->grid_dims[1] > 0 && d->grid_dims[2] > 0) {
            for (int k = 0; k < 3; ++k) {
                v.dims[k] = d->grid_dims[k]; v.wmin[k] = d->grid_min[k]; v.wmax[k] = d->grid_max[k];
            }
            size_t n = (size_t)v.dims[0] * v.dims[1] * v.dims[2];
            v.grid.assign(d->grid, d->grid + n);
            float mx = 0.0f;
            for (float x : v.grid) if (x > mx) mx = x;
            v.grid_max = mx;
        }
        if (v.present() && !d->surface_linked) mt.volume_boundary_only = true;
    }
    finalize_mesh(m);
    s.meshes.push_back(std::move(m));
    return 0;
}

// parseNativeLights, pybind_renderer.hpp:1544-1746 (float/double mix kept:
// M_PI promotes the listed expressions to double before the float store).
int orc_scene_add_light(void* p, const orc_light_desc* d) {
    Scene& s = *static_cast<Scene*>(p);
    Light l;
    l.pos = v3(d->position[0], d->position[1], d->position[2]);
    if (d->has_direction)
        l.normal = make_dir_or(v3(d->direction[0], d->direction[1], d->direction[2]), v3(0.0f, 0.0f, -1.0f));
    float cr = d->color[0], cg = d->color[1], cb = d->color[2];
    float energy = d->energy;
    l.energy = energy;
    const double MPI = 3.14159265358979323846;
    const float EPS = 1e-6f;
    if (d->type == L_POINT) {
        l.type = L_POINT;
        float radius = d->radius;
        l.radius = radius;
        float area = (float)(4.0f * MPI * radius * radius);
        if (area < EPS) area = 1.0f;
        l.area = area;
        if (radius > EPS) {
            l.emission = v3((float)(cr * (energy / (MPI * area))), (float)(cg * (energy / (MPI * area))),
                            (float)(cb * (energy / (MPI * area))));
        } else {
            float factor = (float)(energy / (4.0f * MPI));
            l.emission = v3(cr * factor, cg * factor, cb * factor);
        }
    } else if (d->type == L_SUN) {
        l.type = L_SUN;
        l.area = 1.0f;
        l.emission = v3(cr * energy, cg * energy, cb * energy);
    } else if (d->type == L_SPOT) {
        l.type = L_SPOT;
        l.radius = d->radius;
        float sa = d->spot_size;
        l.spot_angle = sa;
        l.spot_blend = d->spot_blend;
        l.area = 1.0f;
        float half = sa * 0.5f;
        float solid = (float)(2.0f * MPI * (1.0f - std::cos(half)));
        float factor = energy / solid;
        l.emission = v3(cr * factor, cg * factor, cb * factor);
    } else if (d->type == L_AREA) {
        l.type = L_AREA;
        float sx = d->size, sy = d->size_y;
        l.sx = sx; l.sy = sy;
        float area;
        if (d->shape == S_DISK) {
            l.shape = S_DISK;
            float r = sx * 0.5f;
            area = (float)(MPI * r * r);
        } else if (d->shape == S_ELLIPSE) {
            l.shape = S_ELLIPSE;
            float rx = sx * 0.5f, ry = sy * 0.5f;
            area = (float)(MPI * rx * ry);
        } else if (d->shape == S_RECTANGLE) {
            l.shape = S_RECTANGLE;
            area = sx * sy;
        } else {
            l.shape = S_SQUARE;
            area = sx * sy;
        }
        l.area = area;
        if (d->has_right) l.right = make_dir_or(v3(d->right[0], d->right[1], d->right[2]), v3(1, 0, 0));
        if (d->has_up) l.up = make_dir_or(v3(d->up[0], d->up[1], d->up[2]), v3(0, 1, 0));
        float factor = (float)(energy / (MPI * area));
        l.emission = v3(cr * factor, cg * factor, cb * factor);
    } else {
        return -1;
    }
    l.mesh_index = -1; l.tri_index = -1;
    s.native.push_back(l);
    return 0;
}

void orc_scene_set_env(void* p, float r, float g, float b, float strength) {
    Scene& s = *static_cast<Scene*>(p);
    s.env_color = v3(r, g, b);
    s.env_strength = strength;
}

void orc_scene_finalize(void* p) { build_scene_lights(*static_cast<Scene*>(p)); }

// set_camera, pybind_renderer.hpp:131-160
void orc_set_camera(void* p, float px, float py, float pz, float dx, float dy, float dz, float ux, float uy,
                    float uz, float fov_deg) {
    Scene& s = *static_cast<Scene*>(p);
    Camera& c = s.cam;
    c.pos = v3(px, py, pz);
    V3 dir = make_dir_or(v3(dx, dy, dz), v3(0.0f, 0.0f, -1.0f));
    V3 up = make_dir_or(v3(ux, uy, uz), v3(0.0f, 1.0f, 0.0f));
    c.fov = fov_deg * (PI_F / 180.0f);
    V3 right = make_dir_or(cross(dir, up), v3(1.0f, 0.0f, 0.0f));
    V3 upc = make_dir_or(cross(right, dir), v3(0.0f, 1.0f, 0.0f));
    c.forward = dir; c.right = right; c.up = upc;
    c.set = true;
}
void orc_set_exposure(void* p, float ev) { static_cast<Scene*>(p)->cam.exposure_ev = ev; }

int orc_num_lights(void* p) { return (int)static_cast<Scene*>(p)->sl.lights.size(); }
int orc_native_light_start(void* p) { return static_cast<Scene*>(p)->sl.native_start; }
float orc_light_cdf(void* p, int i) { return static_cast<Scene*>(p)->sl.cdf[i]; }
float orc_light_total_area(void* p) { return static_cast<Scene*>(p)->sl.total_area; }
void orc_light_emission(void* p, int i, float* out) {
    const Light& l = static_cast<Scene*>(p)->sl.lights[i];
    out[0] = l.emission.x; out[1] = l.emission.y; out[2] = l.emission.z; out[3] = l.area;
}

// render_tile, pybind_renderer.hpp:185-361: OpenMP over 8x8 micro-tiles,
// per-sample seeding, jitter, SUM x 2^EV, clamp >= 0, alpha 1, Y flip.
// accumulate_fp64 != 0 sums the per-sample radiance in double (the
// high-spp reference mode); 0 sums in float as the reference does.
// stats (optional, 2*3 doubles per pixel, tile order, not flipped): sum and
// sum of squares of per-sample radiance for variance estimates.
// strat (optional, 2 strategies * 6 doubles per pixel): per-strategy sum and
// sum of squares.  counts (optional, 2 int64): closest and shadow queries.
int orc_render_tile(void* p, const char* algorithm, int tx, int ty, int tw, int th, int fw, int fh, int samples,
                    int sample_offset, int max_depth, int accumulate_fp64, int nthreads, float* out, double* stats,
                    double* strat, long long* counts) {
    Scene& s = *static_cast<Scene*>(p);
    std::vector<float> px((size_t)tw * th * 4, 0.0f);
    if (!s.cam.set) {
        std::memcpy(out, px.data(), px.size() * sizeof(float));
        return 1;
    }
    int algo = algo_of(algorithm);
    RayGen g = make_raygen(s, fw, fh);
    const float sens = std::pow(2.0f, s.cam.exposure_ev);
    long long n_closest = 0, n_shadow = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const int micro = 8;
    for (int my = 0; my < th; my += micro) {
        for (int mx = 0; mx < tw; mx += micro) {
            int mh = std::min(micro, th - my), mw = std::min(micro, tw - mx);
#pragma omp parallel for collapse(2) schedule(dynamic, 1) reduction(+ : n_closest, n_shadow)
            for (int j = 0; j < mh; ++j) {
                for (int i = 0; i < mw; ++i) {
                    int pxi = mx + i, pyi = my + j;
                    int gx = tx + pxi, gy = ty + pyi;
                    size_t pix = (size_t)pyi * tw + pxi;
                    V3 radf;
                    double rd[3] = {0, 0, 0}, rq[3] = {0, 0, 0}, sd[2][6] = {{0}};
                    for (int si = 0; si < samples; ++si) {
                        Rng rng;
                        Ray r = camera_ray(g, gx, gy, (uint32_t)(sample_offset + si), samples, rng);
                        SampleOut so;
                        V3 L = trace(s, algo, r, max_depth, rng, so);
                        n_closest += so.rays_closest;
                        n_shadow += so.rays_shadow;
                        radf = add(radf, L);
                        rd[0] += L.x; rd[1] += L.y; rd[2] += L.z;
                        rq[0] += (double)L.x * L.x; rq[1] += (double)L.y * L.y; rq[2] += (double)L.z * L.z;
                        for (int k = 0; k < 2; ++k) {
                            V3 c = so.strat[k];
                            sd[k][0] += c.x; sd[k][1] += c.y; sd[k][2] += c.z;
                            sd[k][3] += (double)c.x * c.x; sd[k][4] += (double)c.y * c.y; sd[k][5] += (double)c.z * c.z;
                        }
                    }
                    V3 col;
                    if (accumulate_fp64) col = v3((float)(rd[0] * sens), (float)(rd[1] * sens), (float)(rd[2] * sens));
                    else col = v3(radf.x * sens, radf.y * sens, radf.z * sens);
                    col = v3(fmax_std(col.x, 0.0f), fmax_std(col.y, 0.0f), fmax_std(col.z, 0.0f));
                    px[pix * 4 + 0] = col.x; px[pix * 4 + 1] = col.y; px[pix * 4 + 2] = col.z; px[pix * 4 + 3] = 1.0f;
                    if (stats) {
                        for (int k = 0; k < 3; ++k) { stats[pix * 6 + k] = rd[k]; stats[pix * 6 + 3 + k] = rq[k]; }
                    }
                    if (strat) {
                        for (int k = 0; k < 2; ++k)
                            for (int q = 0; q < 6; ++q) strat[pix * 12 + k * 6 + q] = sd[k][q];
                    }
                }
            }
        }
    }
    for (int y = 0; y < th; ++y)
        std::memcpy(out + (size_t)(th - 1 - y) * tw * 4, px.data() + (size_t)y * tw * 4, (size_t)tw * 4 * sizeof(float));
    if (counts) { counts[0] = n_closest; counts[1] = n_shadow; }
    return 0;
}

// Per-sample radiance of one pixel, exactly as render_tile computes it
// (samples_in_call decides the jitter).  out: 3 floats per sample.
int orc_trace_pixel(void* p, const char* algorithm, int gx, int gy, int fw, int fh, int samples_in_call,
                    int sample_offset, int nsamples, int max_depth, float* out) {
    Scene& s = *static_cast<Scene*>(p);
    if (!s.cam.set) return 1;
    int algo = algo_of(algorithm);
    RayGen g = make_raygen(s, fw, fh);
    for (int si = 0; si < nsamples; ++si) {
        Rng rng;
        Ray r = camera_ray(g, gx, gy, (uint32_t)(sample_offset + si), samples_in_call, rng);
        SampleOut so;
        V3 L = trace(s, algo, r, max_depth, rng, so);
        out[3 * si] = L.x; out[3 * si + 1] = L.y; out[3 * si + 2] = L.z;
    }
    return 0;
}

// Primary (camera) hits of a tile: for each pixel (tile order) and sample,
// the mesh index, original triangle index, BVH position, raw u, v and t
// bits of intersectScene on the (jittered) camera ray.
int orc_primary_hits(void* p, int tx, int ty, int tw, int th, int fw, int fh, int samples, int sample_offset,
                     int32_t* mesh, int32_t* tri, int32_t* pos, float* u, float* v, float* t) {
    Scene& s = *static_cast<Scene*>(p);
    if (!s.cam.set) return 1;
    RayGen g = make_raygen(s, fw, fh);
#pragma omp parallel for schedule(dynamic, 16)
    for (int pyi = 0; pyi < th; ++pyi) {
        for (int pxi = 0; pxi < tw; ++pxi) {
            for (int si = 0; si < samples; ++si) {
                Rng rng;
                Ray r = camera_ray(g, tx + pxi, ty + pyi, (uint32_t)(sample_offset + si), samples, rng);
                Hit h = intersect_scene(s, r, 0.0f);
                size_t k = ((size_t)pyi * tw + pxi) * samples + si;
                mesh[k] = h.hit ? h.mesh : -1;
                tri[k] = h.hit ? h.tri : -1;
                pos[k] = h.hit ? h.bvh_pos : -1;
                u[k] = h.u; v[k] = h.v; t[k] = h.hit ? h.t : 0.0f;
            }
        }
    }
    return 0;
}

// render_debug (pybind_renderer.hpp:371-439): one unjittered camera ray per
// pixel, intersectScene from t = 0, coloured by traceNormal (renderer.hpp:
// 58-65, then n * 0.5 + 0.5), traceAlbedo (:67-78), traceEmission (:80-93)
// or traceVolume (:109-190); any other mode magenta.  mode: 0 normal,
// 1 albedo, 2 emission, 3 volume.  out: tw*th RGBA, rows bottom first.
static V3 debug_volume(const Scene& s, const Ray& ray) {
    Hit hit = intersect_scene(s, ray, 0.0f);
    if (!hit.hit) return v3(0, 0, 0);
    const Volume& vol = s.meshes[hit.mesh].mat.vol;
    if (!vol.present()) return v3(0, 0, 0);
    const float kEntryOffset = 1e-4f;
    V3 entry = ray_at(ray, hit.t);
    Ray in;
    in.o = add(entry, mul(ray.d, kEntryOffset));
    in.d = ray.d;
    Hit ex = intersect_scene(s, in, 0.0f);
    if (vol.has_grid()) {
        const float t_inside = ex.hit ? ex.t : 0.0f;
        if (t_inside <= 0.0f) return v3(0, 0, 0);
        const float extent = std::max({vol.wmax[0] - vol.wmin[0], vol.wmax[1] - vol.wmin[1], vol.wmax[2] - vol.wmin[2]});
        const int max_dim = std::max({vol.dims[0], vol.dims[1], vol.dims[2]});
        const float voxel = extent / (float)std::max(1, max_dim);
        const float step = fmax_std(1e-4f, voxel * 0.5f);
        const int n_steps = std::min(2048, (int)(t_inside / step) + 1);
        float integrated = 0.0f, max_sampled = 0.0f;
        for (int i = 0; i < n_steps; ++i) {
            const float t = ((float)i + 0.5f) * step;
            if (t > t_inside) break;
            const float d = grid_trilinear(vol, in.o.x + in.d.x * t, in.o.y + in.d.y * t, in.o.z + in.d.z * t);
            integrated += d * step;
            if (d > max_sampled) max_sampled = d;
        }
        const float opacity = 1.0f - std::exp(-vol.density * integrated * 10.0f);
        const float vis = std::max(max_sampled, opacity);
        return v3(vol.color.x * vis, vol.color.y * vis, vol.color.z * vis);
    }
    const float t = ex.hit ? 1.0f - std::exp(-vol.density * ex.t) : 1.0f - std::exp(-vol.density);
    return v3(vol.color.x * t, vol.color.y * t, vol.color.z * t);
}
int orc_render_debug(void* p, int mode, int tx, int ty, int tw, int th, int fw, int fh, float* out) {
    const Scene& s = *static_cast<Scene*>(p);
    for (size_t i = 0; i < (size_t)tw * th * 4; ++i) out[i] = 0.0f;
    if (!s.cam.set) return 1;
    RayGen g = make_raygen(s, fw, fh);
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int py = 0; py < th; ++py) {
        for (int px = 0; px < tw; ++px) {
            Rng rng;
            Ray r = camera_ray(g, tx + px, ty + py, 0u, 1, rng);
            V3 c = v3(0, 0, 0);
            if (mode == 0) {
                Hit h = intersect_scene(s, r, 0.0f);
                V3 n = h.hit ? h.normal : v3(0, 0, 0);
                c = v3(n.x * 0.5f + 0.5f, n.y * 0.5f + 0.5f, n.z * 0.5f + 0.5f);
            } else if (mode == 1 || mode == 2) {
                Hit h = intersect_scene(s, r, 0.0f);
                if (h.hit) {
                    const Material& m = s.meshes[h.mesh].mat;
                    c = mode == 1 ? mat_params(m, h.uv[0], h.uv[1]).albedo : mat_emission(m, h.uv[0], h.uv[1]);
                }
            } else if (mode == 3) {
                c = debug_volume(s, r);
            } else {
                c = v3(1, 0, 1);
            }
            float* o = out + ((size_t)(th - 1 - py) * tw + px) * 4;
            o[0] = c.x; o[1] = c.y; o[2] = c.z; o[3] = 1.0f;
        }
    }
    return 0;
}

// getMaterialParams + getEmission (path_tracer.hpp:45-72) of mesh `mi` at n
// texture coordinates: albedo RGB, emission RGB, metallic, roughness,
// transmission, ior per coordinate.
int orc_material_eval(void* p, int mi, int n, const float* uv, float* out) {
    const Scene& s = *static_cast<Scene*>(p);
    if (mi < 0 || mi >= (int)s.meshes.size()) return -1;
    const Material& m = s.meshes[mi].mat;
    for (int i = 0; i < n; ++i) {
        MatParams mp = mat_params(m, uv[2 * i], uv[2 * i + 1]);
        V3 em = mat_emission(m, uv[2 * i], uv[2 * i + 1]);
        float* o = out + 10 * (size_t)i;
        o[0] = mp.albedo.x; o[1] = mp.albedo.y; o[2] = mp.albedo.z;
        o[3] = em.x; o[4] = em.y; o[5] = em.z;
        o[6] = mp.metallic; o[7] = mp.roughness; o[8] = mp.transmission; o[9] = mp.ior;
    }
    return 0;
}

// intersectScene on explicit rays (closest hit); occluded-only when any_hit.
int orc_intersect(void* p, int n, const float* o, const float* d, const float* tmin, const float* tmax,
                  int32_t* mesh, int32_t* tri, int32_t* pos, float* u, float* v, float* t) {
    Scene& s = *static_cast<Scene*>(p);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.o = v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        r.d = v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        Hit h = intersect_scene(s, r, tmin[i], tmax[i]);
        mesh[i] = h.hit ? h.mesh : -1;
        tri[i] = h.hit ? h.tri : -1;
        pos[i] = h.hit ? h.bvh_pos : -1;
        u[i] = h.u; v[i] = h.v; t[i] = h.hit ? h.t : 0.0f;
    }
    return 0;
}

// Texture coordinate of the closest hit (Hit::uv, bvh.hpp:121-128): 2
// floats per ray, (0, 0) on a miss.
int orc_intersect_uv(void* p, int n, const float* o, const float* d, float* uv) {
    Scene& s = *static_cast<Scene*>(p);
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.o = v3(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        r.d = v3(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        Hit h = intersect_scene(s, r, RAY_T_MIN, RAY_T_MAX);
        uv[2 * i] = h.hit ? h.uv[0] : 0.0f;
        uv[2 * i + 1] = h.hit ? h.uv[1] : 0.0f;
    }
    return 0;
}

// BVH of one mesh as the reference builds it: node boxes, children, leaf
// ranges and the reordered triangle indices.
int orc_mesh_bvh(void* p, int mi, int* n_nodes, float* boxes, int32_t* links, int32_t* tri_idx) {
    Scene& s = *static_cast<Scene*>(p);
    const Mesh& m = s.meshes[mi];
    *n_nodes = (int)m.nodes.size();
    if (boxes) {
        for (size_t i = 0; i < m.nodes.size(); ++i) {
            const Node& nd = m.nodes[i];
            float* b = boxes + 6 * i;
            b[0] = nd.bmin.x; b[1] = nd.bmin.y; b[2] = nd.bmin.z; b[3] = nd.bmax.x; b[4] = nd.bmax.y; b[5] = nd.bmax.z;
            int32_t* l = links + 4 * i;
            l[0] = nd.left; l[1] = nd.right; l[2] = nd.tri_start; l[3] = nd.tri_count;
        }
        for (size_t i = 0; i < m.tri_idx.size(); ++i) tri_idx[i] = m.tri_idx[i];
    }
    return 0;
}

// ---- unit entry points for known-answer tests ----
void orc_seed(uint32_t x, uint32_t y, uint32_t si, uint64_t* state, uint64_t* inc) {
    Rng r;
    seed_xyz(r, x, y, si);
    *state = r.state; *inc = r.inc;
}
void orc_pcg32_stream(uint64_t state, uint64_t inc, int n, uint32_t* out) {
    Rng r; r.state = state; r.inc = inc;
    for (int i = 0; i < n; ++i) out[i] = pcg32(r);
}
void orc_randf_stream(uint32_t x, uint32_t y, uint32_t si, int n, float* out) {
    Rng r;
    seed_xyz(r, x, y, si);
    for (int i = 0; i < n; ++i) out[i] = randf(r);
}
uint32_t orc_jenkins(uint32_t a) { return jenkins32(a); }
uint32_t orc_fmix32(uint32_t a) { return murmur3_fmix(a); }
uint64_t orc_splitmix64(uint64_t a) { return splitmix64(a); }

int orc_intersect_triangle(const float* o, const float* d, const float* a, const float* b, const float* c,
                           float tmin, float tmax, float* tuv) {
    Ray r; r.o = v3(o[0], o[1], o[2]); r.d = v3(d[0], d[1], d[2]);
    float t, u, v;
    bool h = intersect_tri(r, v3(a[0], a[1], a[2]), v3(b[0], b[1], b[2]), v3(c[0], c[1], c[2]), tmin, tmax, t, u, v);
    if (h) { tuv[0] = t; tuv[1] = u; tuv[2] = v; }
    return h ? 1 : 0;
}
int orc_intersect_aabb(const float* o, const float* d, const float* bmin, const float* bmax) {
    Ray r; r.o = v3(o[0], o[1], o[2]); r.d = v3(d[0], d[1], d[2]);
    return intersect_aabb(r, inv_dir(r.d), v3(bmin[0], bmin[1], bmin[2]), v3(bmax[0], bmax[1], bmax[2])) ? 1 : 0;
}
int orc_intersect_sphere(const float* o, const float* d, const float* c, float radius, float* t_out) {
    Ray r; r.o = v3(o[0], o[1], o[2]); r.d = v3(d[0], d[1], d[2]);
    V3 n;
    return intersect_sphere(r, v3(c[0], c[1], c[2]), radius, *t_out, n) ? 1 : 0;
}
float orc_fresnel_dielectric(float c, float eta) { return fresnel_dielectric(c, eta); }
float orc_fresnel_schlick(float c, float f0) { return fresnel_schlick(c, f0); }
float orc_ggx_d(float nh, float r) { return ggx_d(nh, r); }
float orc_ggx_g1(float nv, float r) { return ggx_g1(nv, r); }
float orc_ggx_g2(float nl, float nv, float r) { return ggx_g2(nl, nv, r); }
float orc_power_heuristic(float a, float b) { return power_heuristic(a, b); }
void orc_bsdf_weight(const float* f, float c, float pdf, float* out) {
    V3 w = bsdf_weight(v3(f[0], f[1], f[2]), c, pdf);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
}
float orc_pdf_w_from_area(float dist, float area, float c) { return pdf_w_from_area(dist, area, c); }
float orc_hg_eval(float g, float c) { return hg_eval(g, c); }
// mat: albedo3, metallic, roughness(raw), transmission, ior
void orc_eval_bsdf(const float* mat, const float* wo, const float* wi, const float* n, float* f, float* pdf) {
    MatParams m;
    m.albedo = v3(mat[0], mat[1], mat[2]); m.metallic = mat[3]; m.roughness = fmax_std(mat[4], MIN_ROUGH);
    m.transmission = mat[5]; m.ior = mat[6];
    V3 W = v3(wo[0], wo[1], wo[2]), I = v3(wi[0], wi[1], wi[2]), N = v3(n[0], n[1], n[2]);
    V3 e = eval_bsdf(m, W, I, N);
    f[0] = e.x; f[1] = e.y; f[2] = e.z;
    *pdf = pdf_bsdf(m, W, I, N);
}
// out: wi3, kind, f3|weight3, pdf, type ; the glass branch draws its extra
// number from the stream seeded by (x, y, s).
void orc_sample_bsdf(const float* mat, const float* wo, const float* n, float u1, float u2, float u3, uint32_t sx,
                     uint32_t sy, uint32_t ss, float* out) {
    MatParams m;
    m.albedo = v3(mat[0], mat[1], mat[2]); m.metallic = mat[3]; m.roughness = fmax_std(mat[4], MIN_ROUGH);
    m.transmission = mat[5]; m.ior = mat[6];
    Rng r;
    seed_xyz(r, sx, sy, ss);
    BSample b = sample_bsdf(m, v3(wo[0], wo[1], wo[2]), v3(n[0], n[1], n[2]), u1, u2, u3, r);
    out[0] = b.wi.x; out[1] = b.wi.y; out[2] = b.wi.z; out[3] = (float)b.kind;
    V3 q = b.kind == SK_WEIGHT ? b.weight : b.f;
    out[4] = q.x; out[5] = q.y; out[6] = q.z; out[7] = b.pdf; out[8] = (float)b.type;
}
// Samples light i of the scene's light list at point p; out: pos3, dir3, dist, pdf, emission3
void orc_sample_light(void* sp, int i, const float* pt, float u1, float u2, float* out) {
    Scene& s = *static_cast<Scene*>(sp);
    LSample ls = sample_light(s.sl.lights[i], v3(pt[0], pt[1], pt[2]), u1, u2);
    out[0] = ls.pos.x; out[1] = ls.pos.y; out[2] = ls.pos.z;
    out[3] = ls.dir.x; out[4] = ls.dir.y; out[5] = ls.dir.z;
    out[6] = ls.dist; out[7] = ls.pdf;
    out[8] = ls.emission.x; out[9] = ls.emission.y; out[10] = ls.emission.z;
}
int orc_select_light(void* sp, float u, float* prob) { return select_light(static_cast<Scene*>(sp)->sl, u, *prob); }
float orc_pdf_light_area(float area, const float* p, const float* lp, const float* ln) {
    Light l; l.area = area;
    return pdf_light(l, v3(p[0], p[1], p[2]), v3(lp[0], lp[1], lp[2]), v3(ln[0], ln[1], ln[2]));
}

// ----------------------------------------------------------- diagnostics --
// render_tile with diagnostics enabled (pybind_renderer.hpp:206-208,
// 296-317): the diagnostic tracer of the algorithm ("simple", "mis", any
// other name -> NEE), radiance as render_tile, and a DiagnosticFilm sized to
// the tile with subsample factor `sub`.  Film pixel (fx, fy) receives the
// records of the tile pixels (fx*sub+dx, fy*sub+dy), applied sample by
// sample, and within a sample in row-major (dy, dx) order, each path's
// records in the order it produced them.  Alongside, an FP64 Welford
// (n, mean[3], M2[3]) per film pixel and sampling strategy over the
// per-path strategy sums.
// film: sw*sh pixels of 9576 bytes; strat: sw*sh*4*7 doubles.
int orc_render_tile_diag(void* p, const char* algorithm, int tx, int ty, int tw, int th, int fw, int fh, int samples,
                         int sample_offset, int max_depth, int sub, int nthreads, float* out, unsigned char* film,
                         double* strat, long long* total_recorded) {
    Scene& s = *static_cast<Scene*>(p);
    if (sub < 1) return 2;
    std::string a = algorithm ? algorithm : "";
    const int kind = a == "simple" ? 0 : (a == "mis" ? 2 : 1);
    std::vector<float> px((size_t)tw * th * 4, 0.0f);
    const int sw = tw / sub, sh = th / sub;
    if (!s.cam.set) {
        std::memcpy(out, px.data(), px.size() * sizeof(float));
        return 1;
    }
    RayGen g = make_raygen(s, fw, fh);
    const float sens = std::pow(2.0f, s.cam.exposure_ev);
    std::vector<std::vector<DRec>> recs((size_t)tw * th * samples);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int pi = 0; pi < tw * th; ++pi) {
        const int pxi = pi % tw, pyi = pi / tw;
        V3 radf;
        for (int si = 0; si < samples; ++si) {
            Rng rng;
            Ray r = camera_ray(g, tx + pxi, ty + pyi, (uint32_t)(sample_offset + si), samples, rng);
            DRecorder rec;
            V3 L = trace_diag(s, kind, r, max_depth, rng, rec);
            radf = add(radf, L);
            recs[(size_t)pi * samples + si] = std::move(rec.done);
        }
        V3 col = v3(radf.x * sens, radf.y * sens, radf.z * sens);
        col = v3(fmax_std(col.x, 0.0f), fmax_std(col.y, 0.0f), fmax_std(col.z, 0.0f));
        px[(size_t)pi * 4 + 0] = col.x; px[(size_t)pi * 4 + 1] = col.y; px[(size_t)pi * 4 + 2] = col.z;
        px[(size_t)pi * 4 + 3] = 1.0f;
    }
    long long total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int fi = 0; fi < sw * sh; ++fi) {
        const int fx = fi % sw, fy = fi / sw;
        DPixel d;
        double w[4][7] = {{0}};
        for (int si = 0; si < samples; ++si) {
            for (int dy = 0; dy < sub; ++dy) {
                for (int dx = 0; dx < sub; ++dx) {
                    const size_t pi = (size_t)(fy * sub + dy) * tw + (fx * sub + dx);
                    double sum[4][3] = {{0}};
                    for (const DRec& r : recs[pi * samples + si]) {
                        d.record(r);
                        ++total;
                        sum[r.strategy][0] += r.contrib.x; sum[r.strategy][1] += r.contrib.y;
                        sum[r.strategy][2] += r.contrib.z;
                    }
                    for (int k = 0; k < 4; ++k) {
                        double* W = w[k];
                        W[0] += 1.0;
                        for (int c = 0; c < 3; ++c) {
                            double dl = sum[k][c] - W[1 + c];
                            W[1 + c] += dl / W[0];
                            W[4 + c] += dl * (sum[k][c] - W[1 + c]);
                        }
                    }
                }
            }
        }
        if (film) d_serialise(d, film + (size_t)fi * D_PIXEL_BYTES);
        if (strat) std::memcpy(strat + (size_t)fi * 28, w, sizeof(w));
    }
    if (total_recorded) *total_recorded = total;
    for (int y = 0; y < th; ++y)
        std::memcpy(out + (size_t)(th - 1 - y) * tw * 4, px.data() + (size_t)y * tw * 4, (size_t)tw * 4 * sizeof(float));
    return 0;
}

// PixelPathData<max_groups, max_outliers> driven by explicit records, for
// the known-answer tests of the reference's diagnostics suite.  Record i has
// depth[i] vertices (obj/mat/type/flags at i*16 + k), contribution
// contrib[3i..], light type and strategy.  Outputs: hash[i], coarse[i] per
// record; summary = {groups, outliers, samples, overflow}; per group g:
// ghash, gcount, gmean[3], gvar[3], gdepth, gtype0 (bsdf type of vertex 0);
// totals: tmean[3], tvar[3]; coarse stats: ccount[8], cmean[24]; for
// <16, 4>, the pixel in the reference's byte layout when bytes is given.
// std::nth_element over triangle ids compared by their keys, as the
// reference's BVH build calls it (core/scene.hpp:138-142); keys and ids are
// permuted together.  depth_limit >= 0 runs the library's introselect with
// that partition budget (GNU libstdc++; tests of the heap-select fallback).
void orc_nth_element(float* keys, int32_t* idx, int n, int nth, int depth_limit) {
    std::vector<int32_t> ord(n);
    std::vector<float> k0(keys, keys + n);
    std::vector<int32_t> i0(idx, idx + n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    auto comp = [&](int32_t a, int32_t b) { return k0[a] < k0[b]; };
    if (depth_limit < 0)
        std::nth_element(ord.begin(), ord.begin() + nth, ord.end(), comp);
    else if (n > 0 && nth < n)
        std::__introselect(ord.begin(), ord.begin() + nth, ord.end(), depth_limit,
                           __gnu_cxx::__ops::__iter_comp_iter(comp));
    for (int i = 0; i < n; ++i) {
        keys[i] = k0[ord[i]];
        idx[i] = i0[ord[i]];
    }
}

void orc_diag_pixel_run(int max_groups, int max_outliers, int n, const int32_t* depth, const int32_t* obj,
                        const int16_t* mat, const uint8_t* type, const uint8_t* flags, const float* contrib,
                        const uint8_t* ltype, const uint8_t* strat, uint64_t* hash, uint8_t* coarse, long long* summary,
                        uint64_t* ghash, uint32_t* gcount, float* gmean, float* gvar, int32_t* gdepth, int32_t* gtype0,
                        float* tmean, float* tvar, uint32_t* ccount, float* cmean, unsigned char* bytes) {
    DPixel d((size_t)max_groups, (size_t)max_outliers);
    for (int i = 0; i < n; ++i) {
        DRec r;
        for (int k = 0; k < depth[i] && k < (int)D_MAX_DEPTH; ++k) {
            DVert x;
            x.obj = obj[i * 16 + k]; x.mat = mat[i * 16 + k]; x.type = type[i * 16 + k]; x.flags = flags[i * 16 + k];
            r.v.push_back(x);
        }
        r.contrib = v3(contrib[3 * i], contrib[3 * i + 1], contrib[3 * i + 2]);
        r.light_type = ltype[i]; r.strategy = strat[i];
        hash[i] = d_hash(r.v);
        coarse[i] = d_coarse(r.v);
        d.record(r);
    }
    summary[0] = (long long)d.n_g; summary[1] = (long long)d.n_o; summary[2] = (long long)d.total;
    summary[3] = (long long)d.overflow;
    for (int gi = 0; gi < max_groups; ++gi) {
        const DGroup& G = d.g[gi];
        ghash[gi] = G.hash; gcount[gi] = G.st.n; gdepth[gi] = G.depth; gtype0[gi] = G.v[0].type;
        for (int c = 0; c < 3; ++c) { gmean[gi * 3 + c] = G.st.mean[c]; gvar[gi * 3 + c] = G.st.var[c]; }
    }
    for (int c = 0; c < 3; ++c) { tmean[c] = d.tot.mean[c]; tvar[c] = d.tot.var[c]; }
    for (int t = 0; t < 8; ++t) {
        ccount[t] = d.coarse[t].n;
        for (int c = 0; c < 3; ++c) cmean[t * 3 + c] = d.coarse[t].mean[c];
    }
    if (bytes && max_groups == 16 && max_outliers == 4) d_serialise(d, bytes);
}

// Per-pixel FP64 sums of the radiance of `samples` samples (render_tile's
// camera rays with samples_in_call = `samples_in_call`, seeds
// sample_offset + s) at the listed pixels: sum[3n], sumsq[3n].  For the
// high-spp reference of the relMSE measurement.
int orc_render_pixels(void* p, const char* algorithm, int fw, int fh, int n, const int32_t* xs, const int32_t* ys,
                      int samples, int samples_in_call, long long sample_offset, int max_depth, int nthreads,
                      double* sum, double* sumsq) {
    Scene& s = *static_cast<Scene*>(p);
    if (!s.cam.set) return 1;
    const int algo = algo_of(algorithm);
    RayGen g = make_raygen(s, fw, fh);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < n; ++i) {
        double a[3] = {0, 0, 0}, q[3] = {0, 0, 0};
        for (int si = 0; si < samples; ++si) {
            Rng rng;
            Ray r = camera_ray(g, xs[i], ys[i], (uint32_t)(sample_offset + si), samples_in_call, rng);
            SampleOut so;
            V3 L = trace(s, algo, r, max_depth, rng, so);
            const double v[3] = {L.x, L.y, L.z};
            for (int c = 0; c < 3; ++c) { a[c] += v[c]; q[c] += v[c] * v[c]; }
        }
        for (int c = 0; c < 3; ++c) { sum[3 * i + c] += a[c]; sumsq[3 * i + c] += q[c]; }
    }
    return 0;
}

}  // extern "C"
