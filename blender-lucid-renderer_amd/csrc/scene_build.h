// scene_build.h -- device build of the per-mesh BVHs, the node array and the
// triangle records (scene_build.hip), equal byte for byte to the host
// packer's (scene_pack.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "scene_pack.h"

namespace lucid {
namespace build {

// One BVH node of one mesh, as the host lays the tree out from the triangle
// counts: its range of the mesh's BVH order, its preorder slot among all
// meshes' nodes (ref; its left child is ref + 1, its right child rref), and
// for internal nodes its traversal-node slot and the encodings of its two
// children (node slot, or leaf: first BVH-order triangle and count).
struct Item {
    int32_t mesh, start, end, ref;
    int32_t gnode;    // -1: leaf
    int32_t rref;
    int32_t cenc[2];
};
static_assert(sizeof(Item) == 32, "32-byte build items");

// nodes of at least this many triangles get a workgroup of their own
constexpr int kBigRange = 4096;

struct DeviceGeometry {
    GNode* nodes = nullptr;   // node array: 0 the root, then the meshes' nodes, then the top level
    float4* tri = nullptr;    // 3 per triangle: v0 | mesh, e1 | triangle, e2 | light
    float4* fn = nullptr;     // face normal | smooth flag
    float4* vn = nullptr;     // 3 vertex normals per triangle (scenes with smooth meshes), or null
    float2* tuv = nullptr;    // 3 UVs per triangle (scenes with node materials), or null
    int64_t n_nodes = 0;
    int32_t root = 0;
    int32_t max_depth = 0;
    GNode root_node;          // host copy of node 0 (the scene box)
    std::vector<void*> allocs;  // device buffers, owned by the caller
};

// Builds the geometry of `meshes` on the device on stream `st`; `meta` is
// the host packer's output without geometry (pack_scene(..., false)).
bool build_geometry(const lucid_mesh_desc* meshes, int32_t num_meshes, const PackedScene& meta, hipStream_t st,
                    DeviceGeometry& out, std::string& err);

}  // namespace build
}  // namespace lucid
