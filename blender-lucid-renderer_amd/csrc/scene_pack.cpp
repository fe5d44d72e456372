// scene_pack.cpp -- builds the device scene layout on the host.
//
// Compiled with g++ -O2 -ffp-contract=off: every float computed here
// (bounds, centroids, normals, light radiance, camera basis) must equal the
// value the reference computes in IEEE float, and the per-mesh BVH must have
// the reference's node order (median split, std::nth_element under the same
// comparator), because the BVH-order triangle position is the exact-tie key
// of the closest-hit query.
#include "scene_pack.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <cstring>
#include <functional>

namespace lucid {
namespace {

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 fsub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 fadd(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 fdiv(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
inline F3 fcross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float flen2(F3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline float lo(float a, float b) { return (b < a) ? b : a; }   // std::min
inline float hi(float a, float b) { return (a < b) ? b : a; }   // std::max
const float kDirEps2 = 1e-6f * 1e-6f;

// make_direction_or_default (render_units.hpp:498-515)
inline F3 unit_or(F3 v, F3 fb) {
    float l2 = flen2(v);
    if (l2 < kDirEps2) return fb;
    return fdiv(v, std::sqrt(l2));
}
inline bool unit(F3 v, F3& out) {
    float l2 = flen2(v);
    if (l2 < kDirEps2) return false;
    out = fdiv(v, std::sqrt(l2));
    return true;
}
inline void put3(float* d, F3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
inline F3 get3(const float* s) { return f3(s[0], s[1], s[2]); }

struct RefNode {  // one node of the reference per-mesh BVH
    F3 bmin, bmax;
    int left = -1, right = -1, start = 0, count = 0;
};

struct MeshBvh {
    std::vector<RefNode> nodes;
    std::vector<int> order;   // triIndices
};

// Per-mesh median-split BVH with the reference's semantics
// (core/scene.hpp:65-152): node bounds over the triangles' vertices, leaf at
// <= 4 triangles, longest axis with strict comparisons (x, then y if ey > ex,
// then z if ez exceeds the chosen extent), split index (start+end)/2 chosen
// with std::nth_element on the centroid coordinate, left subtree built first.
void build_mesh_bvh(const float* verts, const int32_t* idx, int ntri, MeshBvh& out) {
    out.nodes.clear();
    out.order.resize(ntri);
    if (ntri == 0) return;
    std::vector<F3> cen(ntri);
    for (int i = 0; i < ntri; ++i) {
        out.order[i] = i;
        F3 a = get3(verts + 3 * idx[3 * i]), b = get3(verts + 3 * idx[3 * i + 1]), c = get3(verts + 3 * idx[3 * i + 2]);
        cen[i] = fadd(a, fdiv(fadd(fsub(b, a), fsub(c, a)), 3.0f));
    }
    out.nodes.reserve(2 * (size_t)ntri);
    std::function<int(int, int)> rec = [&](int start, int end) -> int {
        int id = (int)out.nodes.size();
        out.nodes.push_back(RefNode());
        RefNode nd;
        nd.bmin = f3(1e30f, 1e30f, 1e30f);
        nd.bmax = f3(-1e30f, -1e30f, -1e30f);
        for (int i = start; i < end; ++i) {
            int t = out.order[i];
            for (int k = 0; k < 3; ++k) {
                F3 p = get3(verts + 3 * idx[3 * t + k]);
                nd.bmin = f3(lo(nd.bmin.x, p.x), lo(nd.bmin.y, p.y), lo(nd.bmin.z, p.z));
                nd.bmax = f3(hi(nd.bmax.x, p.x), hi(nd.bmax.y, p.y), hi(nd.bmax.z, p.z));
            }
        }
        int count = end - start;
        if (count <= 4) {
            nd.start = start;
            nd.count = count;
            out.nodes[id] = nd;
            return id;
        }
        F3 ext = fsub(nd.bmax, nd.bmin);
        int axis = 0;
        if (ext.y > ext.x) axis = 1;
        if (ext.z > (axis == 0 ? ext.x : ext.y)) axis = 2;
        int mid = (start + end) / 2;
        std::nth_element(out.order.begin() + start, out.order.begin() + mid, out.order.begin() + end,
                         [&](int a, int b) {
                             const F3& ca = cen[a];
                             const F3& cb = cen[b];
                             float fa = axis == 0 ? ca.x : (axis == 1 ? ca.y : ca.z);
                             float fb = axis == 0 ? cb.x : (axis == 1 ? cb.y : cb.z);
                             return fa < fb;
                         });
        out.nodes[id] = nd;
        int l = rec(start, mid);
        int r = rec(mid, end);
        out.nodes[id].left = l;
        out.nodes[id].right = r;
        return id;
    };
    rec(0, ntri);
}

// Converts reference node `r` of a mesh into the child encoding, appending
// GNodes for internal nodes.  Returns the encoding and tracks the depth.
int32_t emit_blas(const MeshBvh& bvh, int r, int32_t tri_off, std::vector<GNode>& nodes, int depth,
                  int& max_depth) {
    const RefNode& n = bvh.nodes[r];
    if (n.count > 0) return encode_leaf(tri_off + n.start, n.count);
    max_depth = std::max(max_depth, depth + 1);
    int32_t id = (int32_t)nodes.size();
    nodes.push_back(GNode());
    const RefNode& L = bvh.nodes[n.left];
    const RefNode& R = bvh.nodes[n.right];
    GNode g;
    std::memset(&g, 0, sizeof(g));
    put3(g.b + 0, L.bmin); put3(g.b + 3, L.bmax); put3(g.b + 6, R.bmin); put3(g.b + 9, R.bmax);
    nodes[id] = g;
    int32_t cl = emit_blas(bvh, n.left, tri_off, nodes, depth + 1, max_depth);
    int32_t cr = emit_blas(bvh, n.right, tri_off, nodes, depth + 1, max_depth);
    nodes[id].child[0] = cl;
    nodes[id].child[1] = cr;
    return id;
}

struct MeshRoot {
    F3 bmin, bmax, centroid;
    int32_t enc;  // encoding of the mesh's root (node index or leaf)
    int depth;    // depth of the mesh subtree below its root
};

// Top-level tree over mesh roots: median split on the longest axis of the
// centroid bounds.  Its boxes are unions of mesh root boxes; the slab test is
// monotone in the box, so these extra boxes never reject a ray that the
// reference's per-mesh test accepts.
int32_t emit_tlas(std::vector<MeshRoot>& roots, int b, int e, std::vector<GNode>& nodes, int depth,
                  int& max_depth, F3& bmin, F3& bmax) {
    if (e - b == 1) {
        bmin = roots[b].bmin;
        bmax = roots[b].bmax;
        max_depth = std::max(max_depth, depth + roots[b].depth);
        return roots[b].enc;
    }
    F3 cmin = f3(1e30f, 1e30f, 1e30f), cmax = f3(-1e30f, -1e30f, -1e30f);
    for (int i = b; i < e; ++i) {
        F3 c = roots[i].centroid;
        cmin = f3(lo(cmin.x, c.x), lo(cmin.y, c.y), lo(cmin.z, c.z));
        cmax = f3(hi(cmax.x, c.x), hi(cmax.y, c.y), hi(cmax.z, c.z));
    }
    F3 ext = fsub(cmax, cmin);
    int axis = (ext.y > ext.x) ? 1 : 0;
    if (ext.z > (axis == 0 ? ext.x : ext.y)) axis = 2;
    int mid = (b + e) / 2;
    std::nth_element(roots.begin() + b, roots.begin() + mid, roots.begin() + e,
                     [&](const MeshRoot& x, const MeshRoot& y) {
                         float fx = axis == 0 ? x.centroid.x : (axis == 1 ? x.centroid.y : x.centroid.z);
                         float fy = axis == 0 ? y.centroid.x : (axis == 1 ? y.centroid.y : y.centroid.z);
                         return fx < fy;
                     });
    max_depth = std::max(max_depth, depth + 1);
    int32_t id = (int32_t)nodes.size();
    nodes.push_back(GNode());
    F3 lmin, lmax, rmin, rmax;
    int32_t cl = emit_tlas(roots, b, mid, nodes, depth + 1, max_depth, lmin, lmax);
    int32_t cr = emit_tlas(roots, mid, e, nodes, depth + 1, max_depth, rmin, rmax);
    GNode& g = nodes[id];
    std::memset(&g, 0, sizeof(g));
    put3(g.b + 0, lmin); put3(g.b + 3, lmax); put3(g.b + 6, rmin); put3(g.b + 9, rmax);
    g.child[0] = cl;
    g.child[1] = cr;
    bmin = f3(lo(lmin.x, rmin.x), lo(lmin.y, rmin.y), lo(lmin.z, rmin.z));
    bmax = f3(hi(lmax.x, rmax.x), hi(lmax.y, rmax.y), hi(lmax.z, rmax.z));
    return id;
}

inline float luminance(const float* c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

// parseNativeLights conversions (pybind_renderer.hpp:1544-1746).  The
// reference mixes float and M_PI (double); the promoted expressions are kept.
bool convert_light(const lucid_light_desc& d, GLight& l, std::string& err) {
    std::memset(&l, 0, sizeof(l));
    const double MPI = 3.14159265358979323846;
    const float EPS = 1e-6f;  // pbr.hpp:34
    put3(l.pos, get3(d.position));
    F3 n = f3(0.0f, 0.0f, 1.0f);  // Light() default normal (light.hpp:72)
    if (d.has_direction) n = unit_or(get3(d.direction), f3(0.0f, 0.0f, -1.0f));
    put3(l.normal, n);
    put3(l.right, f3(1, 0, 0));
    put3(l.up, f3(0, 1, 0));
    l.sx = 1.0f;
    l.sy = 1.0f;
    float cr = d.color[0], cg = d.color[1], cb = d.color[2];
    float energy = d.energy;
    switch (d.type) {
        case LUCID_LIGHT_POINT: {
            l.type = 1;
            float radius = d.radius;
            l.radius = radius;
            float area = (float)(4.0f * MPI * radius * radius);
            if (area < EPS) area = 1.0f;
            l.area = area;
            if (radius > EPS) {
                l.emission[0] = (float)(cr * (energy / (MPI * area)));
                l.emission[1] = (float)(cg * (energy / (MPI * area)));
                l.emission[2] = (float)(cb * (energy / (MPI * area)));
            } else {
                float factor = (float)(energy / (4.0f * MPI));
                l.emission[0] = cr * factor; l.emission[1] = cg * factor; l.emission[2] = cb * factor;
            }
            break;
        }
        case LUCID_LIGHT_SUN:
            l.type = 2;
            l.area = 1.0f;
            l.emission[0] = cr * energy; l.emission[1] = cg * energy; l.emission[2] = cb * energy;
            break;
        case LUCID_LIGHT_SPOT: {
            l.type = 3;
            l.radius = d.radius;
            float sa = d.spot_size;
            l.spot_blend = d.spot_blend;
            l.area = 1.0f;
            // render::cos(spotAngle * 0.5f) of sampleLightTyped (light.hpp:262)
            l.spot_cos = std::cos(sa * 0.5f);
            float half = sa * 0.5f;
            float solid = (float)(2.0f * MPI * (1.0f - std::cos(half)));
            float factor = energy / solid;
            l.emission[0] = cr * factor; l.emission[1] = cg * factor; l.emission[2] = cb * factor;
            break;
        }
        case LUCID_LIGHT_AREA: {
            l.type = 4;
            float sx = d.size, sy = d.size_y;
            l.sx = sx;
            l.sy = sy;
            float area;
            if (d.shape == LUCID_SHAPE_DISK) {
                l.shape = LUCID_SHAPE_DISK;
                float r = sx * 0.5f;
                area = (float)(MPI * r * r);
            } else if (d.shape == LUCID_SHAPE_ELLIPSE) {
                l.shape = LUCID_SHAPE_ELLIPSE;
                float rx = sx * 0.5f, ry = sy * 0.5f;
                area = (float)(MPI * rx * ry);
            } else {
                l.shape = d.shape == LUCID_SHAPE_RECTANGLE ? LUCID_SHAPE_RECTANGLE : LUCID_SHAPE_SQUARE;
                area = sx * sy;
            }
            l.area = area;
            if (d.has_right) put3(l.right, unit_or(get3(d.right), f3(1, 0, 0)));
            if (d.has_up) put3(l.up, unit_or(get3(d.up), f3(0, 1, 0)));
            float factor = (float)(energy / (MPI * area));
            l.emission[0] = cr * factor; l.emission[1] = cg * factor; l.emission[2] = cb * factor;
            break;
        }
        default:
            err = "unknown native light type " + std::to_string(d.type);
            return false;
    }
    return true;
}

}  // namespace

bool pack_scene(const lucid_mesh_desc* meshes, int32_t num_meshes, const lucid_light_desc* lights,
                int32_t num_lights, const float env_color[3], float env_strength, PackedScene& ps,
                std::string& err, bool geometry) {
    ps = PackedScene();
    if (num_meshes < 0 || num_lights < 0 || (num_meshes > 0 && !meshes) || (num_lights > 0 && !lights)) {
        err = "invalid mesh or light array";
        return false;
    }
    // ---- validate and size ----
    int64_t total = 0;
    for (int m = 0; m < num_meshes; ++m) {
        const lucid_mesh_desc& d = meshes[m];
        if (d.num_vertices < 0 || d.num_triangles < 0 || (d.num_vertices > 0 && !d.vertices) ||
            (d.num_triangles > 0 && !d.indices)) {
            err = "mesh " + std::to_string(m) + ": invalid vertex or index array";
            return false;
        }
        for (int64_t i = 0; i < 3LL * d.num_triangles; ++i) {
            if (d.indices[i] < 0 || d.indices[i] >= d.num_vertices) {
                err = "mesh " + std::to_string(m) + ": triangle index out of range";
                return false;
            }
        }
        total += d.num_triangles;
    }
    if (total > (1LL << 27)) {
        err = "scene exceeds 134M triangles";
        return false;
    }
    ps.num_triangles = (int32_t)total;
    ps.num_meshes = num_meshes;
    if (geometry) {
        ps.tri_v0.resize(4 * total);
        ps.tri_e1.resize(4 * total);
        ps.tri_e2.resize(4 * total);
        ps.tri_fn.resize(4 * total);
    }
    for (int m = 0; m < num_meshes; ++m)
        if (meshes[m].smooth && meshes[m].tri_normals && meshes[m].num_tri_normals > 0) ps.any_smooth = true;
    if (ps.any_smooth && geometry) ps.tri_vn.assign(12 * total, 0.0f);

    // ---- per-mesh BVHs (independent, built in parallel) ----
    std::vector<MeshBvh> bvhs(num_meshes);
    if (geometry) {
        std::atomic<int> next{0};
        auto worker = [&]() {
            for (int m = next.fetch_add(1); m < num_meshes; m = next.fetch_add(1))
                build_mesh_bvh(meshes[m].vertices, meshes[m].indices, meshes[m].num_triangles, bvhs[m]);
        };
        unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        unsigned nt = std::min<unsigned>(hw, (unsigned)std::max(1, num_meshes));
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < nt; ++i) pool.emplace_back(worker);
        worker();
        for (auto& t : pool) t.join();
    }

    std::vector<int32_t> tri_off(num_meshes + 1, 0);
    for (int m = 0; m < num_meshes; ++m) tri_off[m + 1] = tri_off[m] + meshes[m].num_triangles;

    // ---- materials and volumes ----
    ps.mats.resize(num_meshes);
    for (int m = 0; m < num_meshes; ++m) {
        const lucid_mesh_desc& d = meshes[m];
        GMat& g = ps.mats[m];
        std::memset(&g, 0, sizeof(g));
        // getMaterialParams / getEmission (path_tracer.hpp:45-72): the node
        // tree, when used, replaces every constant field
        CompiledMaterial cm;
        if (!compile_material(d, cm, err)) {
            err = "mesh " + std::to_string(m) + ": " + err;
            return false;
        }
        put3(g.albedo, get3(cm.albedo));
        g.metallic = cm.metallic;
        g.roughness = hi(cm.roughness, 0.01f);
        g.transmission = cm.transmission;
        g.ior = cm.ior;
        g.volume = -1;
        g.prog_albedo = g.prog_emission = -1;
        if (!cm.albedo_prog.empty()) {
            g.prog_albedo = (int32_t)ps.node_ops.size();
            ps.node_ops.insert(ps.node_ops.end(), cm.albedo_prog.begin(), cm.albedo_prog.end());
            g.flags |= MF_NODES;
        }
        // the light list takes the emission at uv (0, 0) (scene_lights.hpp:64-67)
        float em00[3] = {cm.emission[0], cm.emission[1], cm.emission[2]};
        if (!cm.emission_prog.empty()) {
            g.prog_emission = (int32_t)ps.node_ops.size();
            ps.node_ops.insert(ps.node_ops.end(), cm.emission_prog.begin(), cm.emission_prog.end());
            run_node_program(&ps.node_ops[g.prog_emission], 0.0f, 0.0f, em00);
            g.flags |= MF_NODES | MF_EM_UV;
        }
        put3(g.emission, get3(em00));
        if (luminance(em00) > 1e-6f) g.flags |= MF_EMISSIVE;
        if (g.flags & MF_NODES) ps.any_nodes = true;
        g.shade_class = (g.flags & MF_NODES) ? kClassNodes
                        : (g.flags & MF_EMISSIVE) ? kClassEmitter
                        : g.transmission > 0.5f ? kClassGlass
                        : g.transmission > 0.0f ? kClassMixed
                        : g.metallic > 0.5f ? kClassMetal
                                              : kClassDielectric;
        if (d.has_volume) {
            GVolume v;
            std::memset(&v, 0, sizeof(v));
            put3(v.color, get3(d.vol_color));
            put3(v.absorption, get3(d.vol_absorption));
            v.density = d.vol_density;
            v.anisotropy = d.vol_anisotropy;
            bool grid = d.grid && d.grid_dims[0] > 0 && d.grid_dims[1] > 0 && d.grid_dims[2] > 0;
            if (grid && (d.grid_dims[0] < 2 || d.grid_dims[1] < 2 || d.grid_dims[2] < 2)) {
                // The trilinear lookup (transmittance.hpp:298-345) clamps the
                // cell index to [0, n-2], an empty range for a 1-voxel axis
                // that the reference leaves undefined; refuse such grids.
                err = "mesh " + std::to_string(m) + ": density grid needs at least 2 voxels on every axis";
                return false;
            }
            if (grid) {
                size_t n = (size_t)d.grid_dims[0] * d.grid_dims[1] * d.grid_dims[2];
                v.has_grid = 1;
                v.grid_offset = (int32_t)ps.grid.size();
                for (int k = 0; k < 3; ++k) {
                    v.dims[k] = d.grid_dims[k];
                    v.wmin[k] = d.grid_min[k];
                    v.wmax[k] = d.grid_max[k];
                }
                float mx = 0.0f;
                for (size_t i = 0; i < n; ++i) mx = d.grid[i] > mx ? d.grid[i] : mx;
                v.grid_max = mx;
                ps.grid.insert(ps.grid.end(), d.grid, d.grid + n);
            }
            bool present = v.density > 0.0f || grid;
            if (present) {
                g.flags |= MF_VOLUME;
                g.volume = (int32_t)ps.volumes.size();
                ps.volumes.push_back(v);
                ps.any_volume = true;
                if (!d.surface_linked) g.flags |= MF_VOLUME_BOUNDARY;
            }
        }
    }

    // ---- lights: emissive triangles (mesh, triangle order), then native ----
    std::vector<std::vector<int32_t>> tri_light(num_meshes);
    for (int m = 0; m < num_meshes; ++m) {
        const lucid_mesh_desc& d = meshes[m];
        tri_light[m].assign(d.num_triangles, -1);
        if (!(ps.mats[m].flags & MF_EMISSIVE)) continue;
        for (int t = 0; t < d.num_triangles; ++t) {
            F3 a = get3(d.vertices + 3 * d.indices[3 * t]);
            F3 b = get3(d.vertices + 3 * d.indices[3 * t + 1]);
            F3 c = get3(d.vertices + 3 * d.indices[3 * t + 2]);
            float area = std::sqrt(flen2(fcross(fsub(b, a), fsub(c, a)))) * 0.5f;  // light.hpp:172-180
            if (!(area > 1e-6f)) continue;
            GLight l;
            std::memset(&l, 0, sizeof(l));
            l.type = 0;
            put3(l.v0, a); put3(l.v1, b); put3(l.v2, c);
            F3 fn = f3(0, 1, 0);
            unit(fcross(fsub(b, a), fsub(c, a)), fn);
            put3(l.normal, fn);
            put3(l.emission, get3(ps.mats[m].emission));
            l.area = area;
            put3(l.pos, fdiv(fadd(fadd(a, b), c), 3.0f));
            tri_light[m][t] = (int32_t)ps.lights.size();
            ps.lights.push_back(l);
            ps.total_area += area;
        }
    }
    ps.native_start = (int32_t)ps.lights.size();
    for (int i = 0; i < num_lights; ++i) {
        GLight l;
        if (!convert_light(lights[i], l, err)) return false;
        if (l.area < 1e-6f) l.area = 1.0f;
        ps.native_light_ids.push_back((int32_t)ps.lights.size());
        ps.lights.push_back(l);
        ps.total_area += l.area;
    }
    ps.cdf.resize(ps.lights.size());
    {
        float cum = 0.0f;
        for (size_t i = 0; i < ps.lights.size(); ++i) {
            cum += ps.lights[i].area;
            ps.cdf[i] = ps.total_area < 1e-6f ? 0.0f : cum / ps.total_area;
        }
    }

    if (!geometry) {  // the device builds the BVHs and triangle records
        ps.tri_light.reserve((size_t)total);
        for (int m = 0; m < num_meshes; ++m) ps.tri_light.insert(ps.tri_light.end(), tri_light[m].begin(), tri_light[m].end());
        ps.env[0] = env_color[0] * env_strength;
        ps.env[1] = env_color[1] * env_strength;
        ps.env[2] = env_color[2] * env_strength;
        return true;
    }

    // ---- triangles in BVH order ----
    if (ps.any_nodes) ps.tri_uv.assign(6 * total, 0.0f);
    for (int m = 0; m < num_meshes; ++m) {
        const lucid_mesh_desc& d = meshes[m];
        const MeshBvh& bvh = bvhs[m];
        bool smooth_mesh = d.smooth && d.tri_normals;
        const bool uvs = (ps.mats[m].flags & MF_NODES) && d.tri_uvs;
        for (int p = 0; p < d.num_triangles; ++p) {
            int t = bvh.order[p];
            size_t g = (size_t)tri_off[m] + p;
            F3 a = get3(d.vertices + 3 * d.indices[3 * t]);
            F3 b = get3(d.vertices + 3 * d.indices[3 * t + 1]);
            F3 c = get3(d.vertices + 3 * d.indices[3 * t + 2]);
            F3 e1 = fsub(b, a), e2 = fsub(c, a);
            float* v0 = &ps.tri_v0[4 * g];
            float* pe1 = &ps.tri_e1[4 * g];
            float* pe2 = &ps.tri_e2[4 * g];
            put3(v0, a); put3(pe1, e1); put3(pe2, e2);
            int32_t mi = m, ti = t, li = tri_light[m][t];
            std::memcpy(v0 + 3, &mi, 4);
            std::memcpy(pe1 + 3, &ti, 4);
            std::memcpy(pe2 + 3, &li, 4);
            F3 fn = f3(0, 1, 0);  // Normal() default when the cross product degenerates
            unit(fcross(e1, e2), fn);
            float* pfn = &ps.tri_fn[4 * g];
            put3(pfn, fn);
            int32_t sm = (smooth_mesh && t < d.num_tri_normals) ? 1 : 0;
            std::memcpy(pfn + 3, &sm, 4);
            if (sm) {
                const float* tn = d.tri_normals + 9 * (size_t)t;
                const F3 up = f3(0.0f, 0.0f, 1.0f);
                for (int k = 0; k < 3; ++k) put3(&ps.tri_vn[12 * g + 4 * k], unit_or(get3(tn + 3 * k), up));
            }
            // a triangle without "triangle_uvs" reads uv (0, 0)
            if (uvs && t < d.num_tri_uvs) std::memcpy(&ps.tri_uv[6 * g], d.tri_uvs + 6 * (size_t)t, 6 * sizeof(float));
        }
    }

    // ---- node array: TLAS on top, mesh BLASes below ----
    ps.nodes.clear();
    ps.nodes.push_back(GNode());  // root placeholder at index 0
    std::vector<TopRoot> tops;
    for (int m = 0; m < num_meshes; ++m) {
        if (meshes[m].num_triangles == 0) continue;
        const MeshBvh& bvh = bvhs[m];
        TopRoot t;
        put3(t.bmin, bvh.nodes[0].bmin);
        put3(t.bmax, bvh.nodes[0].bmax);
        int d = 0;
        t.enc = emit_blas(bvh, 0, tri_off[m], ps.nodes, 0, d);
        t.depth = d;
        tops.push_back(t);
    }
    ps.root = 0;
    ps.max_depth = build_top_level(tops, ps.nodes);

    // ---- environment: to_rgb3f(color) * strength (bvh.hpp:187-189) ----
    ps.env[0] = env_color[0] * env_strength;
    ps.env[1] = env_color[1] * env_strength;
    ps.env[2] = env_color[2] * env_strength;
    return true;
}

int build_top_level(const std::vector<TopRoot>& tops, std::vector<GNode>& nodes) {
    std::vector<MeshRoot> roots;
    for (const TopRoot& t : tops) {
        MeshRoot r;
        r.bmin = get3(t.bmin);
        r.bmax = get3(t.bmax);
        r.centroid = f3(0.5f * (r.bmin.x + r.bmax.x), 0.5f * (r.bmin.y + r.bmax.y), 0.5f * (r.bmin.z + r.bmax.z));
        r.enc = t.enc;
        r.depth = t.depth;
        roots.push_back(r);
    }
    GNode root;
    std::memset(&root, 0, sizeof(root));
    root.child[0] = root.child[1] = kEmptyChild;
    int depth = 1;
    if (roots.size() == 1) {
        put3(root.b + 0, roots[0].bmin);
        put3(root.b + 3, roots[0].bmax);
        root.child[0] = roots[0].enc;
        depth = 1 + roots[0].depth;
    } else if (roots.size() > 1) {
        int mid = (int)roots.size() / 2;
        // split the top level once here so that the root holds two children
        F3 cmin = f3(1e30f, 1e30f, 1e30f), cmax = f3(-1e30f, -1e30f, -1e30f);
        for (auto& r : roots) {
            cmin = f3(lo(cmin.x, r.centroid.x), lo(cmin.y, r.centroid.y), lo(cmin.z, r.centroid.z));
            cmax = f3(hi(cmax.x, r.centroid.x), hi(cmax.y, r.centroid.y), hi(cmax.z, r.centroid.z));
        }
        F3 ext = fsub(cmax, cmin);
        int axis = (ext.y > ext.x) ? 1 : 0;
        if (ext.z > (axis == 0 ? ext.x : ext.y)) axis = 2;
        std::nth_element(roots.begin(), roots.begin() + mid, roots.end(), [&](const MeshRoot& x, const MeshRoot& y) {
            float fx = axis == 0 ? x.centroid.x : (axis == 1 ? x.centroid.y : x.centroid.z);
            float fy = axis == 0 ? y.centroid.x : (axis == 1 ? y.centroid.y : y.centroid.z);
            return fx < fy;
        });
        F3 lmin, lmax, rmin, rmax;
        int dl = 0;
        int32_t cl = emit_tlas(roots, 0, mid, nodes, 1, dl, lmin, lmax);
        int32_t cr = emit_tlas(roots, mid, (int)roots.size(), nodes, 1, dl, rmin, rmax);
        put3(root.b + 0, lmin); put3(root.b + 3, lmax); put3(root.b + 6, rmin); put3(root.b + 9, rmax);
        root.child[0] = cl;
        root.child[1] = cr;
        depth = std::max(1, dl);
    }
    nodes[0] = root;
    return depth + 1;
}

CameraBasis camera_basis(const lucid_camera& cam, int32_t full_w, int32_t full_h) {
    CameraBasis b;
    const float PI_F = 3.14159265358979323846f;
    F3 dir = unit_or(get3(cam.dir), f3(0.0f, 0.0f, -1.0f));
    F3 up = unit_or(get3(cam.up), f3(0.0f, 1.0f, 0.0f));
    float fov = cam.fov_deg * (PI_F / 180.0f);
    F3 right = unit_or(fcross(dir, up), f3(1.0f, 0.0f, 0.0f));
    F3 upc = unit_or(fcross(right, dir), f3(0.0f, 1.0f, 0.0f));
    put3(b.pos, get3(cam.pos));
    put3(b.fwd, dir);
    put3(b.right, right);
    put3(b.up, upc);
    b.aspect = (float)full_w / (float)full_h;
    b.scale = std::tan(fov * 0.5f);
    b.sensitivity = std::pow(2.0f, cam.exposure_ev);
    return b;
}

}  // namespace lucid
