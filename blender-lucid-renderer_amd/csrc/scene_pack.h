// scene_pack.h -- host-side packing of a Lucid scene into the device layout.
//
// The device layout (all arrays 16-byte aligned, structure of arrays):
//   nodes   GNode[]   binary BVH, child boxes stored in the parent (64 B).
//                     Top levels: a BVH over mesh root boxes; below: the
//                     per-mesh BVHs exactly as the reference builds them
//                     (median split on the longest axis, leaves <= 4 triangles;
//                     include/core/scene.hpp:65-152), so a triangle's global
//                     position in BVH order is the reference's exact-tie key.
//   tri_v0/e1/e2      float4 per triangle in BVH order: vertex 0, edge 1,
//                     edge 2 (the Moller-Trumbore operands), with the mesh
//                     index, original triangle index and light index in .w.
//   tri_fn            float4 face normal (w = smooth flag)
//   tri_vn            3 x float4 vertex normals per triangle (smooth meshes)
//   mats              GMat per mesh; lights GLight[] (emissive triangles
//                     first, then native lights, scene_lights.hpp:56-122) with
//                     the area CDF.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lucid_hip.h"

namespace lucid {

constexpr int32_t kEmptyChild = 0x7fffffff;

// Leaf child encoding: ~((first_triangle << 3) | count), count <= 4.
inline int32_t encode_leaf(int32_t start, int32_t count) { return ~((start << 3) | count); }

struct alignas(16) GNode {
    float b[12];        // child 0 min xyz, max xyz, child 1 min xyz, max xyz
    int32_t child[2];   // >= 0 node index, < 0 leaf (encode_leaf), kEmptyChild
    int32_t pad[2];
};
static_assert(sizeof(GNode) == 64, "GNode is 64 bytes");

// MF_EMISSIVE: the material's emission at uv (0, 0) is emissive, so its
// triangles are in the light list (scene_lights.hpp:59-70).  MF_NODES: the
// albedo or the emission is a node-tree program of the hit's uv (a Checker
// texture below the Surface link); MF_EM_UV: the emission is such a program,
// so is_emissive is decided per hit.
enum MatFlags : int32_t {
    MF_EMISSIVE = 1, MF_VOLUME_BOUNDARY = 2, MF_VOLUME = 4, MF_NODES = 8, MF_EM_UV = 16
};

// Shading classes: hits of one class take the same branches of a shading
// bounce (class 0 is a miss).  The key of the regrouped shading order
// (lucid_set_shade_order).
enum ShadeClass : int32_t {
    kClassMiss = 0, kClassEmitter = 1, kClassDielectric = 2, kClassMetal = 3, kClassMixed = 4, kClassGlass = 5,
    kClassNodes = 6, kShadeClasses = 8
};

struct alignas(16) GMat {
    float albedo[3];
    float metallic;
    float roughness;      // clamped to >= 0.01 (getMaterialParams, path_tracer.hpp:61-62)
    float transmission;
    float ior;
    int32_t flags;
    float emission[3];
    int32_t volume;       // index into volumes, -1 if none
    int32_t prog_albedo;  // first GNodeOp of the albedo program, -1: albedo above
    int32_t prog_emission;// first GNodeOp of the emission program, -1: emission above
    int32_t shade_class;  // ShadeClass of the material
    int32_t pad;
};
static_assert(sizeof(GMat) == 64, "GMat is 64 bytes");

// Node-tree programs: postfix code over RGB values, one 16-byte op each.
// A program leaves one RGB on its stack at NOP_END.  The host constant-folds
// every subtree that does not depend on uv, so programs remain only for
// trees with a Checker texture (node_evaluator.cpp:352-401).
enum NodeOpcode : int32_t {
    NOP_END = 0,
    NOP_CONST = 1,    // push (a, b, c)
    NOP_CHECKER = 2,  // pop color2, color1; push the checker colour at uv * a
    NOP_MIX = 3,      // pop B, A; push A * (1 - a) + B * a
    NOP_SCALE = 4,    // pop C; push C * a
};
constexpr int32_t kNopSwapped = 0x100;  // binary op: the operands were pushed B first
constexpr int kNodeStack = 4;           // evaluation slots a program may use
struct alignas(16) GNodeOp {
    int32_t op;
    float a, b, c;
};
static_assert(sizeof(GNodeOp) == 16, "GNodeOp is 16 bytes");

struct alignas(16) GLight {
    int32_t type;         // 0 emissive triangle, 1 point, 2 sun, 3 spot, 4 area
    int32_t shape;
    float area;
    float radius;
    float pos[3];
    float spot_cos;       // cos(spotAngle * 0.5), computed on the host
    float normal[3];
    float spot_blend;
    float emission[3];
    float sx;
    float v0[3];
    float sy;
    float v1[3];
    int32_t pad0;
    float v2[3];
    int32_t pad1;
    float right[3];
    int32_t pad2;
    float up[3];
    int32_t pad3;
};
static_assert(sizeof(GLight) == 144, "GLight is 144 bytes");

struct alignas(16) GVolume {
    float color[3];
    float density;
    float absorption[3];
    float anisotropy;
    float wmin[3];
    float grid_max;
    float wmax[3];
    int32_t has_grid;
    int32_t dims[3];
    int32_t grid_offset;  // element offset into the scene's grid buffer
};
static_assert(sizeof(GVolume) == 80, "GVolume is 80 bytes");

struct PackedScene {
    std::vector<GNode> nodes;
    std::vector<float> tri_v0, tri_e1, tri_e2, tri_fn, tri_vn;  // 4 floats per entry
    std::vector<GMat> mats;
    std::vector<GNodeOp> node_ops;  // node-tree programs of all materials
    std::vector<float> tri_uv;      // 6 floats per triangle in BVH order (only with MF_NODES materials)
    std::vector<GVolume> volumes;
    std::vector<float> grid;     // all heterogeneous grids, concatenated
    std::vector<GLight> lights;
    std::vector<float> cdf;
    std::vector<int32_t> native_light_ids;  // light index of native light i (i in JSON order)
    int32_t native_start = 0;
    float total_area = 0.0f;
    float env[3] = {0.05f, 0.05f, 0.05f};    // color * strength
    int32_t num_triangles = 0;
    int32_t num_meshes = 0;
    int32_t root = 0;
    int32_t max_depth = 0;                   // deepest node path (for the traversal stack)
    bool any_smooth = false;
    bool any_volume = false;
    bool any_nodes = false;                  // a material with MF_NODES (tri_uv filled)
    // without geometry: the light index of every triangle (mesh, triangle
    // order; -1: not a light), for the device build
    std::vector<int32_t> tri_light;
};

// Packs the scene; returns false and sets err on invalid input.  With
// geometry = false the BVHs, the node array and the triangle records are
// left to the device build (scene_build.hip); materials, lights, volumes and
// the environment are packed as always.
bool pack_scene(const lucid_mesh_desc* meshes, int32_t num_meshes, const lucid_light_desc* lights,
                int32_t num_lights, const float env_color[3], float env_strength, PackedScene& out,
                std::string& err, bool geometry = true);

// A mesh's BVH root as the top-level tree sees it: its box, its child
// encoding (node index or leaf) and the depth of its subtree.
struct TopRoot {
    float bmin[3], bmax[3];
    int32_t enc;
    int depth;
};
// The top-level tree over the mesh roots (median split on the longest axis
// of the root-box centroids, std::nth_element): its nodes are appended to
// `nodes`, whose index 0 (reserved by the caller) becomes the root.  Returns
// the deepest node path, for the traversal stack.
int build_top_level(const std::vector<TopRoot>& roots, std::vector<GNode>& nodes);

// A material as getMaterialParams / getEmission see it at one uv
// (include/integrator/path_tracer.hpp:45-72), compiled from a mesh
// descriptor: the constant parameters plus programs for albedo and emission.
struct CompiledMaterial {
    float albedo[3], emission[3];      // values when the program is empty
    float metallic, roughness, transmission, ior;  // roughness not yet clamped
    std::vector<GNodeOp> albedo_prog, emission_prog;  // empty: constant
};
bool compile_material(const lucid_mesh_desc& d, CompiledMaterial& out, std::string& err);
// Runs a program (ending in NOP_END) at uv, as the device does.
void run_node_program(const GNodeOp* ops, float u, float v, float rgb[3]);

// Camera basis and ray-generation constants of render_tile, computed on the
// host in IEEE float exactly as set_camera / render_tile do
// (pybind_renderer.hpp:131-160, 211-215).
struct CameraBasis {
    float pos[3], fwd[3], right[3], up[3];
    float aspect, scale, sensitivity;
};
CameraBasis camera_basis(const lucid_camera& cam, int32_t full_w, int32_t full_h);

}  // namespace lucid
