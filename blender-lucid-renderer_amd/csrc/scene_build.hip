// scene_build.hip -- the per-mesh BVHs and triangle records built on the GPU
// (gfx950), equal byte for byte to the host packer's (scene_pack.cpp).
//
// The reference builds one median-split BVH per mesh (core/scene.hpp:65-152):
// node bounds over the triangles' vertices, a leaf at <= 4 triangles, the
// longest axis, and std::nth_element on the triangle centroids splitting
// the range at (start + end) / 2, left subtree first.  The BVH-order
// position of a triangle is the exact-tie key of the closest-hit query, so
// the device build must leave every range in the order nth_element leaves
// it, ties included.  Three facts make that a data-parallel job:
//   * the tree's shape depends on the triangle counts alone: a range of c
//     triangles splits into floor(c / 2) and c - floor(c / 2), so every node's
//     range, preorder index and node-array slot is known before any geometry
//     is read (the host lists them, level by level);
//   * a node's bounds and split depend only on the set and order of its own
//     range, which its parent's split fixed: the nodes of one level are
//     independent;
//   * nth_element is a deterministic function of the range's keys in their
//     current order.  select_nth below restates the introselect of the C++
//     library the host packer is compiled against (GNU libstdc++ 11:
//     median-of-three pivot moved to the front, unguarded Hoare partition,
//     heap select past 2 floor(log2 n) partitions, insertion sort of the
//     last <= 3 elements), so it moves elements exactly as std::nth_element.
// Kernels, one launch per tree level over all meshes at once:
//   k_centroids   centroid a + ((b - a) + (c - a)) / 3 of every triangle
//   k_level       one lane per node (k_level_big: one workgroup per node of
//                 >= kBigRange triangles for the bounds and the key gather):
//                 bounds, split axis, keys, select_nth
//   k_emit_nodes  the 64-byte traversal nodes of the internal nodes
//   k_tri_records triangle records in BVH order (v0, e1, e2 with mesh,
//                 triangle and light ids; face normal; vertex normals; UVs)
// All float arithmetic is IEEE (-ffp-contract=off, correctly rounded / and
// sqrt), in the host packer's operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "scene_build.h"

namespace lucid {
namespace build {
namespace {

// ---- libstdc++ 11 introselect on parallel key / index arrays ------------
// comp(a, b) = key[a] < key[b]; positions are absolute indices.
struct Seq {
    float* k;
    int32_t* v;
    __host__ __device__ __forceinline__ bool lt(int a, int b) const { return k[a] < k[b]; }
    __host__ __device__ __forceinline__ void swap(int a, int b) const {
        const float tk = k[a]; k[a] = k[b]; k[b] = tk;
        const int32_t tv = v[a]; v[a] = v[b]; v[b] = tv;
    }
    __host__ __device__ __forceinline__ void put(int a, float kk, int32_t vv) const { k[a] = kk; v[a] = vv; }
    __host__ __device__ __forceinline__ void copy(int dst, int src) const { k[dst] = k[src]; v[dst] = v[src]; }
};

__host__ __device__ void move_median_to_first(const Seq& s, int result, int a, int b, int c) {
    if (s.lt(a, b)) {
        if (s.lt(b, c)) s.swap(result, b);
        else if (s.lt(a, c)) s.swap(result, c);
        else s.swap(result, a);
    } else if (s.lt(a, c)) {
        s.swap(result, a);
    } else if (s.lt(b, c)) {
        s.swap(result, c);
    } else {
        s.swap(result, b);
    }
}

__host__ __device__ int unguarded_partition(const Seq& s, int first, int last, int pivot) {
    while (true) {
        while (s.lt(first, pivot)) ++first;
        --last;
        while (s.lt(pivot, last)) --last;
        if (!(first < last)) return first;
        s.swap(first, last);
        ++first;
    }
}

__host__ __device__ int unguarded_partition_pivot(const Seq& s, int first, int last) {
    const int mid = first + (last - first) / 2;
    move_median_to_first(s, first, first + 1, mid, last - 1);
    return unguarded_partition(s, first + 1, last, first);
}

// heap of [first, first + len) under comp, with the hole / value moves of
// std::__push_heap and std::__adjust_heap
__host__ __device__ void push_heap(const Seq& s, int first, int hole, int top, float vk, int32_t vv) {
    int parent = (hole - 1) / 2;
    while (hole > top && s.k[first + parent] < vk) {
        s.copy(first + hole, first + parent);
        hole = parent;
        parent = (hole - 1) / 2;
    }
    s.put(first + hole, vk, vv);
}

__host__ __device__ void adjust_heap(const Seq& s, int first, int hole, int len, float vk, int32_t vv) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (s.lt(first + second, first + second - 1)) second--;
        s.copy(first + hole, first + second);
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        s.copy(first + hole, first + second - 1);
        hole = second - 1;
    }
    push_heap(s, first, hole, top, vk, vv);
}

__host__ __device__ void heap_select(const Seq& s, int first, int middle, int last) {
    const int len = middle - first;
    if (len >= 2) {  // make_heap
        int parent = (len - 2) / 2;
        while (true) {
            adjust_heap(s, first, parent, len, s.k[first + parent], s.v[first + parent]);
            if (parent == 0) break;
            parent--;
        }
    }
    for (int i = middle; i < last; ++i)
        if (s.lt(i, first)) {  // pop_heap(first, middle, i)
            const float vk = s.k[i];
            const int32_t vv = s.v[i];
            s.copy(i, first);
            adjust_heap(s, first, 0, len, vk, vv);
        }
}

__host__ __device__ void insertion_sort(const Seq& s, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        const float vk = s.k[i];
        const int32_t vv = s.v[i];
        if (vk < s.k[first]) {
            for (int j = i; j > first; --j) s.copy(j, j - 1);
            s.put(first, vk, vv);
        } else {
            int hole = i, next = i - 1;
            while (vk < s.k[next]) {
                s.copy(hole, next);
                hole = next;
                --next;
            }
            s.put(hole, vk, vv);
        }
    }
}

__host__ __device__ void select_nth(const Seq& s, int first, int nth, int last, int depth_limit = -1) {
    if (first == last || nth == last) return;
    if (depth_limit < 0) depth_limit = 2 * (31 - __builtin_clz((unsigned)(last - first)));  // 2 floor(log2 n)
    while (last - first > 3) {
        if (depth_limit == 0) {
            heap_select(s, first, nth + 1, last);
            s.swap(first, nth);
            return;
        }
        --depth_limit;
        const int cut = unguarded_partition_pivot(s, first, last);
        if (cut <= nth) first = cut;
        else last = cut;
    }
    insertion_sort(s, first, last);
}

// ---- kernels ------------------------------------------------------------
struct MeshIn {
    int32_t vtx;      // first vertex (float3) of the mesh in the vertex array
    int32_t tri;      // first triangle (global, input order = BVH-order offset)
    int32_t ntri;
    int32_t nrm;      // first triangle of its normals in the normal array, or -1
    int32_t n_nrm;    // triangles carrying normals
    int32_t uv;       // first triangle of its UVs in the UV array, or -1
    int32_t n_uv;
    int32_t pad;
};

__device__ __forceinline__ float3 ld3(const float* p, size_t i) { return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float lo(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ float hi(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ float comp(float3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// mesh of global triangle g: the last m with tri[m] <= g
__device__ __forceinline__ int mesh_of(const MeshIn* M, int nm, int g) {
    int lo_ = 0, hi_ = nm - 1;
    while (lo_ < hi_) {
        const int mid = (lo_ + hi_ + 1) >> 1;
        if (M[mid].tri <= g) lo_ = mid;
        else hi_ = mid - 1;
    }
    return lo_;
}

__global__ void k_centroids(const MeshIn* M, int nm, const float* verts, const int32_t* idx, int n, float* cen,
                            int32_t* order) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= n) return;
    const MeshIn& m = M[mesh_of(M, nm, g)];
    const int t = g - m.tri;
    const float3 a = ld3(verts, (size_t)m.vtx + idx[3 * (size_t)g]);
    const float3 b = ld3(verts, (size_t)m.vtx + idx[3 * (size_t)g + 1]);
    const float3 c = ld3(verts, (size_t)m.vtx + idx[3 * (size_t)g + 2]);
    const float3 ab = sub3(b, a), ac = sub3(c, a);
    const float3 s = make_float3(ab.x + ac.x, ab.y + ac.y, ab.z + ac.z);
    cen[3 * (size_t)g] = a.x + s.x / 3.0f;
    cen[3 * (size_t)g + 1] = a.y + s.y / 3.0f;
    cen[3 * (size_t)g + 2] = a.z + s.z / 3.0f;
    order[g] = t;
}

// Bounds of the range's triangles (order holds mesh-local triangle ids).
struct Box {
    float mn[3], mx[3];
};
__device__ __forceinline__ void box_init(Box& b) {
    b.mn[0] = b.mn[1] = b.mn[2] = 1e30f;
    b.mx[0] = b.mx[1] = b.mx[2] = -1e30f;
}
__device__ __forceinline__ void box_add(Box& b, float3 p) {
    b.mn[0] = lo(b.mn[0], p.x); b.mn[1] = lo(b.mn[1], p.y); b.mn[2] = lo(b.mn[2], p.z);
    b.mx[0] = hi(b.mx[0], p.x); b.mx[1] = hi(b.mx[1], p.y); b.mx[2] = hi(b.mx[2], p.z);
}
__device__ __forceinline__ void box_tri(Box& b, const MeshIn& m, const float* verts, const int32_t* idx, int t) {
    const size_t g = (size_t)m.tri + t;
    for (int k = 0; k < 3; ++k) box_add(b, ld3(verts, (size_t)m.vtx + idx[3 * g + k]));
}
__device__ __forceinline__ int split_axis(const Box& b) {
    const float ex = b.mx[0] - b.mn[0], ey = b.mx[1] - b.mn[1], ez = b.mx[2] - b.mn[2];
    int axis = 0;
    if (ey > ex) axis = 1;
    if (ez > (axis == 0 ? ex : ey)) axis = 2;
    return axis;
}

// Min / max are exact, so the bounds do not depend on the order in which
// the vertices are folded in (no NaN coordinates: the packer refuses none,
// but a NaN vertex would make the reference's own bounds order-dependent).
__global__ void k_level(const Item* items, int n_items, const MeshIn* M, const float* verts, const int32_t* idx,
                        const float* cen, int32_t* order, float* keys, float* boxes) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n_items) return;
    const Item it = items[i];
    const MeshIn m = M[it.mesh];
    int32_t* ord = order + m.tri;
    Box b;
    box_init(b);
    for (int p = it.start; p < it.end; ++p) box_tri(b, m, verts, idx, ord[p]);
    float* bx = boxes + 6 * (size_t)it.ref;
    for (int k = 0; k < 3; ++k) { bx[k] = b.mn[k]; bx[3 + k] = b.mx[k]; }
    if (it.end - it.start <= 4) return;
    const int axis = split_axis(b);
    float* key = keys + m.tri;
    for (int p = it.start; p < it.end; ++p) key[p] = cen[3 * ((size_t)m.tri + ord[p]) + axis];
    select_nth(Seq{key, ord}, it.start, (it.start + it.end) / 2, it.end);
}

// Nodes of >= kBigRange triangles: the workgroup folds the bounds and
// gathers the keys, lane 0 selects.
__global__ __launch_bounds__(256) void k_level_big(const Item* items, const int32_t* big, const MeshIn* M,
                                                   const float* verts, const int32_t* idx, const float* cen,
                                                   int32_t* order, float* keys, float* boxes) {
    __shared__ float red[6][256];
    __shared__ int s_axis;
    const Item it = items[big[blockIdx.x]];
    const MeshIn m = M[it.mesh];
    int32_t* ord = order + m.tri;
    Box b;
    box_init(b);
    for (int p = it.start + (int)threadIdx.x; p < it.end; p += 256) box_tri(b, m, verts, idx, ord[p]);
    for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = b.mn[k]; red[3 + k][threadIdx.x] = b.mx[k]; }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = lo(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = hi(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Box t;
        for (int k = 0; k < 3; ++k) { t.mn[k] = red[k][0]; t.mx[k] = red[3 + k][0]; }
        float* bx = boxes + 6 * (size_t)it.ref;
        for (int k = 0; k < 6; ++k) bx[k] = red[k][0];
        s_axis = split_axis(t);
    }
    __syncthreads();
    const int axis = s_axis;
    float* key = keys + m.tri;
    for (int p = it.start + (int)threadIdx.x; p < it.end; p += 256)
        key[p] = cen[3 * ((size_t)m.tri + ord[p]) + axis];
    __syncthreads();
    if (threadIdx.x == 0) select_nth(Seq{key, ord}, it.start, (it.start + it.end) / 2, it.end);
}

__global__ void k_emit_nodes(const Item* items, int n_items, const float* boxes, GNode* nodes) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n_items) return;
    const Item it = items[i];
    if (it.gnode < 0) return;
    GNode g;
    const float* l = boxes + 6 * (size_t)(it.ref + 1);  // the left child follows in preorder
    const float* r = boxes + 6 * (size_t)it.rref;
    for (int k = 0; k < 6; ++k) { g.b[k] = l[k]; g.b[6 + k] = r[k]; }
    g.child[0] = it.cenc[0];
    g.child[1] = it.cenc[1];
    g.pad[0] = g.pad[1] = 0;
    nodes[it.gnode] = g;
}

__device__ __forceinline__ float3 unit_or(float3 v, float3 fb, bool* ok) {
    const float l2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (l2 < 1e-6f * 1e-6f) {
        if (ok) *ok = false;
        return fb;
    }
    if (ok) *ok = true;
    const float l = sqrtf(l2);
    return make_float3(v.x / l, v.y / l, v.z / l);
}

__global__ void k_tri_records(const MeshIn* M, int nm, const float* verts, const int32_t* idx, const float* nrm,
                              const float* uv, const int32_t* tri_light, const int32_t* order, int n, float4* tri,
                              float4* fn, float4* vn, float2* tuv) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= n) return;
    const int mi = mesh_of(M, nm, g);
    const MeshIn& m = M[mi];
    const int t = order[g];
    const size_t gi = (size_t)m.tri + t;  // input position of the triangle
    const float3 a = ld3(verts, (size_t)m.vtx + idx[3 * gi]);
    const float3 b = ld3(verts, (size_t)m.vtx + idx[3 * gi + 1]);
    const float3 c = ld3(verts, (size_t)m.vtx + idx[3 * gi + 2]);
    const float3 e1 = sub3(b, a), e2 = sub3(c, a);
    const int32_t li = tri_light ? tri_light[gi] : -1;
    tri[3 * (size_t)g] = make_float4(a.x, a.y, a.z, __int_as_float(mi));
    tri[3 * (size_t)g + 1] = make_float4(e1.x, e1.y, e1.z, __int_as_float(t));
    tri[3 * (size_t)g + 2] = make_float4(e2.x, e2.y, e2.z, __int_as_float(li));
    const float3 cr = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
    const float3 f = unit_or(cr, make_float3(0.0f, 1.0f, 0.0f), nullptr);
    const int sm = (m.nrm >= 0 && t < m.n_nrm) ? 1 : 0;
    fn[g] = make_float4(f.x, f.y, f.z, __int_as_float(sm));
    if (vn) {
        for (int k = 0; k < 3; ++k) {
            float3 v = make_float3(0.0f, 0.0f, 0.0f);
            if (sm) v = unit_or(ld3(nrm, 3 * ((size_t)m.nrm + t) + k), make_float3(0.0f, 0.0f, 1.0f), nullptr);
            vn[3 * (size_t)g + k] = make_float4(v.x, v.y, v.z, 0.0f);
        }
    }
    if (tuv) {
        const bool has = m.uv >= 0 && t < m.n_uv;
        for (int k = 0; k < 3; ++k)
            tuv[3 * (size_t)g + k] = has ? make_float2(uv[6 * ((size_t)m.uv + t) + 2 * k], uv[6 * ((size_t)m.uv + t) + 2 * k + 1])
                                         : make_float2(0.0f, 0.0f);
    }
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- tree shape from the counts alone ----------------------------------
// nodes(c) = 1 (leaf, c <= 4) or 1 + nodes(floor(c/2)) + nodes(c - floor(c/2));
// internal(c) likewise without the leaves; depth(c) = internal levels.
struct Shape {
    std::unordered_map<int, int64_t> nodes_, internal_;
    int64_t nodes(int c) {
        if (c <= 4) return 1;
        auto f = nodes_.find(c);
        if (f != nodes_.end()) return f->second;
        const int64_t v = 1 + nodes(c / 2) + nodes(c - c / 2);
        nodes_[c] = v;
        return v;
    }
    int64_t internal(int c) {
        if (c <= 4) return 0;
        auto f = internal_.find(c);
        if (f != internal_.end()) return f->second;
        const int64_t v = 1 + internal(c / 2) + internal(c - c / 2);
        internal_[c] = v;
        return v;
    }
    static int depth(int c) { return c <= 4 ? 0 : 1 + depth(c - c / 2); }
};

struct Gen {
    Shape& sh;
    std::vector<std::vector<Item>>& levels;
    int32_t mesh, ref_base, node_base, out_off;
    // range [start, end) at `level`, preorder index `id`, internal index `iid`
    void run(int start, int end, int64_t id, int64_t iid, int level) {
        struct Frame { int start, end; int64_t id, iid; int level; };
        std::vector<Frame> st{{start, end, id, iid, level}};
        while (!st.empty()) {
            const Frame f = st.back();
            st.pop_back();
            if ((int)levels.size() <= f.level) levels.resize(f.level + 1);
            Item it;
            std::memset(&it, 0, sizeof(it));
            it.mesh = mesh;
            it.start = f.start;
            it.end = f.end;
            it.ref = (int32_t)(ref_base + f.id);
            it.gnode = -1;
            const int c = f.end - f.start;
            if (c > 4) {
                const int mid = (f.start + f.end) / 2, lc = mid - f.start, rc = f.end - mid;
                const int64_t lid = f.id + 1, rid = f.id + 1 + sh.nodes(lc);
                const int64_t liid = f.iid + 1, riid = f.iid + 1 + sh.internal(lc);
                it.gnode = (int32_t)(node_base + f.iid);
                it.rref = (int32_t)(ref_base + rid);
                it.cenc[0] = lc <= 4 ? encode_leaf(out_off + f.start, lc) : (int32_t)(node_base + liid);
                it.cenc[1] = rc <= 4 ? encode_leaf(out_off + mid, rc) : (int32_t)(node_base + riid);
                st.push_back({mid, f.end, rid, riid, f.level + 1});
                st.push_back({f.start, mid, lid, liid, f.level + 1});
            }
            levels[f.level].push_back(it);
        }
    }
};

struct Allocs {
    std::vector<void*> v;
    ~Allocs() {
        for (void* p : v) (void)hipFree(p);
    }
    template <class T>
    bool get(T** p, size_t count) {
        *p = nullptr;
        if (count == 0) return true;
        if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)) != hipSuccess) return false;
        v.push_back(*p);
        return true;
    }
    template <class T>
    bool put(T** p, const T* src, size_t count, hipStream_t st) {
        if (!get(p, count)) return false;
        return count == 0 || hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, st) == hipSuccess;
    }
};

}  // namespace

bool build_geometry(const lucid_mesh_desc* meshes, int32_t num_meshes, const PackedScene& meta, hipStream_t st,
                    DeviceGeometry& out, std::string& err) {
    out = DeviceGeometry();
    const int64_t n = meta.num_triangles;
    // ---- inputs: vertices, indices, normals, UVs, light ids ----
    std::vector<MeshIn> M(num_meshes > 0 ? num_meshes : 1);
    int64_t nv = 0, nn = 0, nuv = 0, tri = 0;
    for (int m = 0; m < num_meshes; ++m) {
        const lucid_mesh_desc& d = meshes[m];
        MeshIn& mi = M[m];
        std::memset(&mi, 0, sizeof(mi));
        mi.vtx = (int32_t)nv;
        mi.tri = (int32_t)tri;
        mi.ntri = d.num_triangles;
        const bool smooth = d.smooth && d.tri_normals && d.num_tri_normals > 0;
        mi.nrm = smooth ? (int32_t)nn : -1;
        mi.n_nrm = smooth ? std::min(d.num_tri_normals, d.num_triangles) : 0;
        const bool uvs = (meta.mats[m].flags & MF_NODES) && d.tri_uvs && d.num_tri_uvs > 0;
        mi.uv = uvs ? (int32_t)nuv : -1;
        mi.n_uv = uvs ? std::min(d.num_tri_uvs, d.num_triangles) : 0;
        nv += d.num_vertices;
        nn += mi.n_nrm;
        nuv += mi.n_uv;
        tri += d.num_triangles;
    }
    if (nv > (1LL << 31) / 3 || tri != n) {
        err = "device build: scene too large";
        return false;
    }
    std::vector<float> hv((size_t)nv * 3), hn((size_t)nn * 9), huv((size_t)nuv * 6);
    std::vector<int32_t> hi((size_t)n * 3);
    for (int m = 0; m < num_meshes; ++m) {
        const lucid_mesh_desc& d = meshes[m];
        const MeshIn& mi = M[m];
        if (d.num_vertices) std::memcpy(&hv[3 * (size_t)mi.vtx], d.vertices, 12 * (size_t)d.num_vertices);
        if (d.num_triangles) std::memcpy(&hi[3 * (size_t)mi.tri], d.indices, 12 * (size_t)d.num_triangles);
        if (mi.nrm >= 0) std::memcpy(&hn[9 * (size_t)mi.nrm], d.tri_normals, 36 * (size_t)mi.n_nrm);
        if (mi.uv >= 0) std::memcpy(&huv[6 * (size_t)mi.uv], d.tri_uvs, 24 * (size_t)mi.n_uv);
    }
    bool any_light = false;
    for (int32_t l : meta.tri_light) any_light = any_light || l >= 0;

    // ---- tree shape: every node's range and slots, level by level ----
    Shape shape;
    std::vector<std::vector<Item>> levels;
    std::vector<TopRoot> tops;
    int64_t refs = 0, gnodes = 1;  // node 0: the top-level root
    for (int m = 0; m < num_meshes; ++m) {
        const int c = meshes[m].num_triangles;
        if (c == 0) continue;
        Gen gen{shape, levels, m, (int32_t)refs, (int32_t)gnodes, M[m].tri};
        gen.run(0, c, 0, 0, 0);
        TopRoot t;
        std::memset(&t, 0, sizeof(t));
        t.enc = c <= 4 ? encode_leaf(M[m].tri, c) : (int32_t)gnodes;
        t.depth = Shape::depth(c);
        tops.push_back(t);
        refs += shape.nodes(c);
        gnodes += shape.internal(c);
        if (refs > (1LL << 30) || gnodes > (1LL << 30)) {
            err = "device build: BVH too large";
            return false;
        }
    }

    // ---- device buffers ----
    Allocs tmp;
    MeshIn* dM;
    float *dv, *dn, *duv, *cen, *keys, *boxes;
    int32_t *di, *dl = nullptr, *order;
    bool ok = tmp.put(&dM, M.data(), M.size(), st) && tmp.put(&dv, hv.data(), hv.size(), st) &&
              tmp.put(&di, hi.data(), hi.size(), st) && tmp.put(&dn, hn.data(), hn.size(), st) &&
              tmp.put(&duv, huv.data(), huv.size(), st) && tmp.get(&cen, (size_t)n * 3) &&
              tmp.get(&keys, (size_t)n) && tmp.get(&order, (size_t)n) && tmp.get(&boxes, (size_t)refs * 6);
    if (ok && any_light) ok = tmp.put(&dl, meta.tri_light.data(), meta.tri_light.size(), st);
    std::vector<Item> all;
    std::vector<size_t> level_off;
    for (auto& lv : levels) {
        level_off.push_back(all.size());
        all.insert(all.end(), lv.begin(), lv.end());
    }
    level_off.push_back(all.size());
    Item* ditems = nullptr;
    if (ok) ok = tmp.put(&ditems, all.data(), all.size(), st);
    // outputs (owned by the scene); the node array also holds the top-level
    // tree: its root in slot 0 and its other r - 2 internal nodes (r mesh
    // roots) after the meshes' nodes
    const int64_t n_top = tops.size() >= 2 ? (int64_t)tops.size() - 2 : 0;
    Allocs keep;
    ok = ok && keep.get(&out.nodes, (size_t)(gnodes + n_top)) && keep.get(&out.tri, (size_t)n * 3) &&
         keep.get(&out.fn, (size_t)n);
    if (ok && meta.any_smooth) ok = keep.get(&out.vn, (size_t)n * 3);
    if (ok && meta.any_nodes) ok = keep.get(&out.tuv, (size_t)n * 3);
    if (!ok) {
        err = "device build: out of device memory";
        return false;
    }
    // ---- levels ----
    if (n > 0) k_centroids<<<nblk(n, 256), 256, 0, st>>>(dM, num_meshes, dv, di, (int)n, cen, order);
    std::vector<int32_t> big_idx;
    int32_t* dbig = nullptr;
    size_t big_cap = 0;
    for (size_t L = 0; L + 1 < level_off.size(); ++L) {
        const size_t a = level_off[L], b = level_off[L + 1];
        // large nodes go to one workgroup each, the rest one lane each
        big_idx.clear();
        std::vector<Item> small;
        for (size_t i = a; i < b; ++i) {
            if (all[i].end - all[i].start >= kBigRange) big_idx.push_back((int32_t)i);
        }
        if (!big_idx.empty()) {
            if (big_idx.size() > big_cap) {
                if (!tmp.get(&dbig, big_idx.size())) {
                    err = "device build: out of device memory";
                    return false;
                }
                big_cap = big_idx.size();
            }
            if (hipMemcpyAsync(dbig, big_idx.data(), big_idx.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
                err = "device build: copy failed";
                return false;
            }
            k_level_big<<<(unsigned)big_idx.size(), 256, 0, st>>>(ditems, dbig, dM, dv, di, cen, order, keys, boxes);
            // the host vector dies with this iteration: wait for the copy
            if (hipStreamSynchronize(st) != hipSuccess) {
                err = "device build: level failed";
                return false;
            }
        }
        // one lane per node; nodes handled by k_level_big are skipped by
        // running the small kernel only over the whole level when no node
        // is big, else over the individual small nodes
        if (big_idx.empty()) {
            k_level<<<nblk((int64_t)(b - a), 64), 64, 0, st>>>(ditems + a, (int)(b - a), dM, dv, di, cen, order, keys,
                                                               boxes);
        } else {
            // levels holding a big node are near the root: few nodes; copy
            // the small ones into a compact list
            std::vector<Item> rest;
            for (size_t i = a; i < b; ++i)
                if (all[i].end - all[i].start < kBigRange) rest.push_back(all[i]);
            if (!rest.empty()) {
                Item* dr = nullptr;
                if (!tmp.put(&dr, rest.data(), rest.size(), st)) {
                    err = "device build: out of device memory";
                    return false;
                }
                k_level<<<nblk((int64_t)rest.size(), 64), 64, 0, st>>>(dr, (int)rest.size(), dM, dv, di, cen, order,
                                                                      keys, boxes);
                if (hipStreamSynchronize(st) != hipSuccess) {
                    err = "device build: level failed";
                    return false;
                }
            }
        }
    }
    if (!all.empty()) k_emit_nodes<<<nblk((int64_t)all.size(), 256), 256, 0, st>>>(ditems, (int)all.size(), boxes, out.nodes);
    if (n > 0)
        k_tri_records<<<nblk(n, 256), 256, 0, st>>>(dM, num_meshes, dv, di, dn, duv, dl, order, (int)n, out.tri,
                                                     out.fn, out.vn, out.tuv);
    if (hipGetLastError() != hipSuccess) {
        err = "device build: launch failed";
        return false;
    }
    // ---- top level on the host from the mesh root boxes ----
    std::vector<float> root_box(6 * tops.size());
    {
        size_t k = 0;
        int64_t r = 0;
        for (int m = 0; m < num_meshes; ++m) {
            const int c = meshes[m].num_triangles;
            if (c == 0) continue;
            if (hipMemcpyAsync(&root_box[6 * k], boxes + 6 * r, 24, hipMemcpyDeviceToHost, st) != hipSuccess) {
                err = "device build: copy failed";
                return false;
            }
            r += shape.nodes(c);
            ++k;
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "device build: kernels failed";
        return false;
    }
    for (size_t k = 0; k < tops.size(); ++k)
        for (int c = 0; c < 3; ++c) {
            tops[k].bmin[c] = root_box[6 * k + c];
            tops[k].bmax[c] = root_box[6 * k + 3 + c];
        }
    std::vector<GNode> top_nodes((size_t)gnodes);  // slots 1 .. gnodes - 1 are the device's
    out.max_depth = build_top_level(tops, top_nodes);
    out.root = 0;
    if ((int64_t)top_nodes.size() != gnodes + n_top) {
        err = "device build: top-level tree of unexpected size";
        return false;
    }
    bool cp = hipMemcpyAsync(out.nodes, &top_nodes[0], sizeof(GNode), hipMemcpyHostToDevice, st) == hipSuccess;
    if (n_top)
        cp = cp && hipMemcpyAsync(out.nodes + gnodes, &top_nodes[(size_t)gnodes], (size_t)n_top * sizeof(GNode),
                                  hipMemcpyHostToDevice, st) == hipSuccess;
    cp = cp && hipStreamSynchronize(st) == hipSuccess;
    if (!cp) {
        err = "device build: copy failed";
        return false;
    }
    out.root_node = top_nodes[0];
    out.n_nodes = gnodes + n_top;
    out.allocs = keep.v;  // owned by the scene from here on
    keep.v.clear();
    return true;
}

}  // namespace build
}  // namespace lucid

// select_nth run on the host, for the tests that compare it with
// std::nth_element (no device needed).
// depth_limit < 0: the library's own limit, 2 floor(log2 n).
extern "C" int lucid_debug_select_nth(float* keys, int32_t* idx, int32_t n, int32_t nth, int32_t depth_limit) {
    if (!keys || !idx || n < 0 || nth < 0 || nth > n) return 1;
    lucid::build::select_nth(lucid::build::Seq{keys, idx}, 0, nth, n, depth_limit);
    return 0;
}
