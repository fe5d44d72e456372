// pathtrace.hip -- gfx950 wavefront path tracer behind the lucid_hip.h C ABI.
//
// One render = a sequence of wavefront batches ("waves").  A wave holds
// P pixels x B consecutive sample indices as independent paths with their own
// PCG state, so the random stream of every path is the one the reference's
// render_tile seeds for (x, y, sample_offset + s) (pybind_renderer.hpp:274-293).
// Per bounce the host enqueues, on one stream and without synchronising:
//
//   k_intersect  closest hit of every live path (LDS traversal stack)
//   k_shade      emission / native light / environment with MIS weights,
//                NEE light sampling (shadow ray queued), BSDF sampling,
//                Russian roulette; live paths are compacted into the next
//                queue with a wavefront ballot + one atomic per wavefront
//   k_shadow     occlusion of the queued shadow rays; unoccluded NEE
//                contributions are added to the path radiance before the next
//                bounce's emission, i.e. in the reference's summation order
//
// and after the last bounce k_film adds the B per-sample radiances of each
// pixel to the tile sum in sample order (the reference's `radiance = radiance
// + sample_radiance`).  k_output applies 2^EV, clamps at 0, writes alpha 1
// and flips rows (pybind_renderer.hpp:331-360).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "lucid_hip.h"
#include "scene_build.h"
#include "scene_pack.h"
#include "shading.h"
#include "diag_film.h"
#include "guide_device.h"
#include "guide_host.h"
#include "guide_mlp.h"

namespace lucid {

constexpr int kBlock = 128;      // 2 wavefronts per workgroup
constexpr int kStackMax = 64;    // LDS traversal stack entries per lane, upper bound
constexpr int kMinRaysPerLane = 8;  // persistent traversal: grid shrinks for small waves
constexpr int64_t kWaveBytesPerPath = 10 * 16 + 7 * 4;  // WaveDev SoA bytes per path slot

// Queue counters.  Device-scope atomics on one 128-byte line serialize (about
// 11 ns each on MI355X, measured with 4,096 wavefronts; 64 lines take 0.25 ns
// per atomic in aggregate), and a pass issues millions of them: the queue
// appends of every shading wavefront and the ray fetches of the persistent
// traversal.  So each launch spreads them over kShards counters with a line
// each.  The counter block (cnt):
//   [0], [1]  live-queue counts      [2] shadow-queue count
//   [3]       stack overflows        [7] cancel latch      bytes 64..127: ray counters
//   cnt[kFetchBase + (kind * kShards + s) * kLineInts]: fetch counter of range s
//             of a persistent launch's queue (kind 0 closest hit, 1 shadow, 2 volume march)
//   cnt[kAppendBase + (w * kShards + s) * kLineInts]: entries in segment s of
//             the next live queue (w 0) or of the shadow queue (w 1)
constexpr int kShardBits = 6;
constexpr int kShards = 1 << kShardBits;  // one per lane of a wavefront
constexpr int kLineInts = 32;             // 128 bytes
constexpr int kFetchBase = 32;
constexpr int kAppendBase = kFetchBase + 3 * kShards * kLineInts;
constexpr int kCounterInts = kAppendBase + 2 * kShards * kLineInts;
constexpr int64_t kCounterBytes = (int64_t)kCounterInts * 4;
// entries per append segment of a wave of n paths: a shading launch appends
// the entries of input chunk c (kBlock consecutive queue positions) to
// segment c % kShards, at most one per position and queue
inline int64_t segment_capacity(int64_t n) {
    const int64_t chunks = (n + kBlock - 1) / kBlock;
    return (chunks + kShards - 1) / kShards * kBlock;
}

// ------------------------------------------------------------- device scene --
struct SceneDev {
    const GNode* nodes;
    const float4* tri;      // per triangle: v0, e1, e2 (the Moller-Trumbore operands), 48 B
    const float4* fn;
    const float4* vn;       // 3 per triangle, or null
    const GMat* mats;
    const GNodeOp* prog;    // node-tree programs (MF_NODES materials), or null
    const float2* tuv;      // 3 UVs per triangle, BVH order, when a material has MF_NODES
    const GLight* lights;
    const float* cdf;
    const int32_t* native_ids;
    const GVolume* vols;
    const float* grid;      // FP32 density grids, or null when
    const __half* grid16;   // the scene stores them as FP16
    int32_t n_lights, n_native, n_vols, root;
    float total_area;
    float env[3];
    int32_t any_vbo;
    int32_t native_start;  // index of the first native light in `lights`
};

// Diagnostic path storage (only in diagnostic runs), indexed [k * dstride + p]:
//   dv0/dv1  vertex k: position + object id, normal + packed
//            (material & 0xffff) | bsdf_type << 16 | flags << 24
//   dr0/dr1  record slot k (= bounce): contribution + packed
//            valid | strategy << 8 | light type << 16 | prefix vertices << 24,
//            light vertex position + light object id
//   dcnt     vertices | slots << 8
struct WaveDev {
    float4* ray_o;   // origin, tmin
    float4* ray_d;   // direction
    float4* thr;     // throughput, last BSDF pdf
    float4* rad;     // path radiance
    ulonglong2* rng; // PCG state, increment
    int4* meta;      // pixel, sample, depth, flags
    float4* hit;     // t, primitive (bits), u, v
    float4* sh_o;    // shadow origin, distance
    float4* sh_d;    // shadow direction
    float4* sh_c;    // pending NEE contribution
    // traceVolumeMIS with gridded media: the NEE contribution's factors
    // until k_vol_march has the shadow ray's transmittance T:
    // sh_c = (sh_a * (sh_e * T)) * sh_t
    float4 *sh_a, *sh_e, *sh_t;
    int32_t* q[2];   // live-path queues
    int32_t* qs;     // shadow queue
    int32_t* sq;     // the live queue regrouped by ray bin (lucid_set_ray_order)
    uint32_t* skey;  // bin of queue entry i
    uint32_t* sbins; // kRayBins counters, then start offsets
    // append segments of the next live queue (aq) and of the shadow queue
    // (aqs): kShards segments of seg_cap entries, concatenated into q / qs
    // by k_queue_concat after each shading launch
    int32_t* aq;
    int32_t* aqs;
    int64_t seg_cap;
    int32_t* cnt;    // counter block: queue counts, fetch and append counters (kCounterInts)
    unsigned long long* rays;  // [0] closest, [1] shadow, [2..3] closest nodes/tris, [4..5] shadow nodes/tris
    int32_t count_work;        // accumulate node / triangle fetch counts
    float4 *dv0, *dv1, *dr0, *dr1;
    int32_t* dcnt;
    int64_t dstride;
    // guided runs: network outputs (40 per path) and, for 1 path in
    // `tstride`, training slots of kTrainSlots bounces x 5 float4:
    // (p, T_in.x) (n, T_in.y) (wo, T_in.z) (wi, valid) (L_after, valid)
    __half* gout;
    int64_t gspare;  // a spare row of gout that slots past the live queue write
    float4* tslot;
    int32_t tstride;
    int64_t tnrec;  // records (paths with slots); slot arrays are [k * 5 + field][record]
    // render_tile: the context's cancel word in host memory the device reads
    // (null elsewhere).  One thread of each bounce's closest-hit launch reads
    // it and latches a set word in cnt[7] (device memory, cleared at the
    // render's start, never cleared within it).  Shading ends every path of
    // a latched wave and the film skips it.  Every other kernel reads only
    // the latch, written by an earlier launch of the same stream, so all of
    // them agree: a path ends early only if the film drops its wave.
    const int32_t* stop;
};
// The latch is written and read with device-scope atomics: a plain store
// can sit in the writer's L2 and a plain load of this uniform address
// compiles to a scalar-cache load, and neither cache is coherent with the
// CUs of the other XCDs that run the later kernels.
__device__ __forceinline__ int32_t latch_read(const int32_t* w) {
    return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void poll_stop(const WaveDev& W) {
    if (W.stop && *reinterpret_cast<const volatile int32_t*>(W.stop) != 0)
        __hip_atomic_store(&W.cnt[7], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool stop_latched(const WaveDev& W) { return W.stop && latch_read(&W.cnt[7]) != 0; }
constexpr int kTrainSlots = 6;
// Paths that keep training slots: one per block of `tstride` consecutive
// paths, at a hashed offset so that every pixel column contributes.
__device__ __forceinline__ int train_record_path(const WaveDev& W, int r) {
    uint32_t h = (uint32_t)r * 0x9e3779b1u;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return r * W.tstride + (int)(h % (uint32_t)W.tstride);
}
__device__ __forceinline__ bool keeps_record(const WaveDev& W, int p) {
    return W.tslot && train_record_path(W, p / W.tstride) == p;
}
// Field f (0..4) of training slot k of record r: structure of arrays, so
// consecutive records are consecutive float4s.
__device__ __forceinline__ float4* tslot_at(const WaveDev& W, int r, int k, int f) {
    return W.tslot + (size_t)(k * 5 + f) * (size_t)W.tnrec + (size_t)r;
}
constexpr int32_t kRecValid = 1;  // bit 0 of a diagnostic record's packed meta

struct CamDev {
    float pos[3], fwd[3], right[3], up[3];
    float aspect, scale;
    int32_t fw, fh, tx, ty, tw, th;
    int32_t samples_in_call;  // render_tile's `samples` (jitter only when > 1)
    int32_t s0;               // absolute sample index of wave slot b = 0
    int32_t P, B, max_depth;
    const int32_t* s0p;       // when set, s0 is read here (a session pass replayed from a graph)
};

// ---------------------------------------------------------------- helpers --
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Appends `pred` lanes to a queue: one atomic per wavefront.  Every lane of
// the wavefront must call it (convergent).
__device__ __forceinline__ int wave_append(int32_t* counter, bool pred) {
    unsigned long long mask = __ballot(pred);
    int total = __popcll(mask);
    int base = 0;
    int leader = mask ? __ffsll((long long)mask) - 1 : 0;
    if (total > 0 && lane_id() == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    int rank = __popcll(mask & ((1ull << lane_id()) - 1ull));
    return base + rank;
}

// ---- sharded queue counters (see kShards) ----
// Persistent launches hand out the positions [0, n) of their input queue
// through kShards fetch counters.  The queue is cut into chunks of
// kFetchChunk positions dealt round-robin to kShards ranges: range s holds
// chunks s, s + kShards, s + 2 kShards, ... in that order.  All ranges
// advance together, so the rays in flight stay near the front of the queue
// (the order the queue was built in: neighbouring pixels, or paths queued
// together) as with one counter, while each counter takes 1/kShards of the
// atomics.  A wavefront fetches from one range until it is handed out, then
// moves on to the next range that has positions left.
constexpr int kFetchChunkBits = 8;
constexpr int kFetchChunk = 1 << kFetchChunkBits;
__device__ __forceinline__ int range_len(int n, int s) {
    const int full = n >> (kFetchChunkBits + kShardBits);
    const int rem = n - (full << (kFetchChunkBits + kShardBits));
    return (full << kFetchChunkBits) + min(max(rem - (s << kFetchChunkBits), 0), kFetchChunk);
}
// queue position of entry j of range s
__device__ __forceinline__ int range_pos(int s, int j) {
    return ((j >> kFetchChunkBits) << (kFetchChunkBits + kShardBits)) + (s << kFetchChunkBits) + (j & (kFetchChunk - 1));
}
__device__ __forceinline__ int32_t* fetch_counters(const WaveDev& W, int kind) {
    return W.cnt + kFetchBase + kind * kShards * kLineInts;
}
struct RangeFetch {
    int cur;                     // range the wavefront fetches from (uniform)
    unsigned long long drained;  // ranges it has seen handed out
};
__device__ __forceinline__ RangeFetch first_range() {
    RangeFetch f;
    f.cur = (int)((blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) & (kShards - 1));
    f.drained = 0ull;
    return f;
}
// Range f.cur is handed out: move to the first range after it (cyclically)
// that still has positions, reading all kShards fetch counters at once (one
// per lane).  Returns true when no range has any left.  Every lane must call
// it.  A range is left at most once, so a wavefront's fetches end.
__device__ __forceinline__ bool next_range(const int32_t* fetch, int n, RangeFetch& f) {
    f.drained |= 1ull << f.cur;
    const int s = lane_id();
    const int taken = __hip_atomic_load(fetch + s * kLineInts, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long avail = __ballot(taken < range_len(n, s)) & ~f.drained;
    if (avail == 0) return true;
    const unsigned long long after = avail & (~0ull << f.cur);
    f.cur = __ffsll((long long)(after ? after : avail)) - 1;
    return false;
}
// One wavefront fetch of `want` entries of range f.cur (for the lanes that
// ask); entries base + k < len are valid (queue position range_pos(s, base + k)).
// Every lane must call it.
__device__ __forceinline__ int range_fetch(int32_t* fetch, int n, RangeFetch& f, unsigned long long idle, int want,
                                           bool& exhausted, int& s, int& len) {
    const int leader = __ffsll((long long)idle) - 1;
    s = f.cur;
    len = range_len(n, s);
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(fetch + s * kLineInts, want);
    base = __shfl(base, leader);
    if (base + want >= len) exhausted = next_range(fetch, n, f);
    return base;
}
// Zeroes the fetch counters of the next persistent launches (called by one
// workgroup of the launch before them).
__device__ __forceinline__ void reset_fetch(const WaveDev& W) {
    for (int j = threadIdx.x; j < 3 * kShards; j += blockDim.x) W.cnt[kFetchBase + j * kLineInts] = 0;
}
// Zeroes the append counters of the next shading launch (one workgroup of
// the closest-hit launch before it).
__device__ __forceinline__ void reset_append(const WaveDev& W) {
    for (int j = threadIdx.x; j < 2 * kShards; j += blockDim.x) W.cnt[kAppendBase + j * kLineInts] = 0;
}

// Same for a 64-bit running total whose positions wrap around a ring of
// `capacity` entries: returns this lane's ring index.
__device__ __forceinline__ int wave_append_u64(unsigned long long* total, bool pred, int capacity) {
    unsigned long long mask = __ballot(pred);
    int n = __popcll(mask);
    unsigned long long base = 0;
    int leader = mask ? __ffsll((long long)mask) - 1 : 0;
    if (n > 0 && lane_id() == leader) base = atomicAdd(total, (unsigned long long)n);
    base = __shfl(base, leader);
    int rank = __popcll(mask & ((1ull << lane_id()) - 1ull));
    return (int)((base + (unsigned long long)rank) % (unsigned long long)capacity);
}

LUCID_HD v3f inv_dir(v3f d) {  // intersection.hpp:412-419
    return mk(fabsf(d.x) > 1e-8f ? 1.0f / d.x : 0.0f, fabsf(d.y) > 1e-8f ? 1.0f / d.y : 0.0f,
              fabsf(d.z) > 1e-8f ? 1.0f / d.z : 0.0f);
}
// intersectAABB slab test (intersection.hpp:257-302), with the entry
// distance returned for ordering and pruning.
LUCID_HD bool slab1(float dc, float ic, float dmin, float dmax, float& tmin, float& tmax) {
    if (fabsf(dc) < 1e-8f) return !((dmin > 0.0f && dmax > 0.0f) || (dmin < 0.0f && dmax < 0.0f));
    float t1 = dmin * ic, t2 = dmax * ic;
    float amin = smin(t1, t2), amax = smax(t1, t2);
    tmin = smax(tmin, amin);
    tmax = smin(tmax, amax);
    return tmax >= tmin;
}
__device__ __forceinline__ bool aabb(v3f o, v3f d, v3f id, const float* b, float& tentry) {
    float tmin = 0.0f, tmax = RAY_T_MAX;
    if (!slab1(d.x, id.x, b[0] - o.x, b[3] - o.x, tmin, tmax)) return false;
    if (!slab1(d.y, id.y, b[1] - o.y, b[4] - o.y, tmin, tmax)) return false;
    if (!slab1(d.z, id.z, b[2] - o.z, b[5] - o.z, tmin, tmax)) return false;
    tentry = tmin;
    return tmax >= smax(tmin, 0.0f);
}
// Moller-Trumbore with precomputed edges (intersection.hpp:335-396).
LUCID_HD bool tri_test(v3f o, v3f d, float4 A, float4 E1, float4 E2, float tmin, float tmax,
                                         float& t, float& u, float& v) {
    v3f e1 = mk(E1), e2 = mk(E2);
    v3f pvec = cross(d, e2);
    float det = dot(e1, pvec);
    const float AE = 1e-6f * 1e-6f;
    if (det > -AE && det < AE) return false;
    v3f tvec = o - mk(A);
    u = dot(tvec, pvec) / det;
    if (u < 0.0f || u > 1.0f) return false;
    v3f qvec = cross(tvec, e1);
    v = dot(d, qvec) / det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qvec) / det;
    return t >= tmin && t <= tmax;
}

// Pruning margin: a child box whose entry distance exceeds the current
// closest t by more than 1 % cannot hold a triangle whose Moller-Trumbore t
// ties or beats it; the reference never prunes, and the tie rule below
// reproduces its "last tested wins" order.
LUCID_HD float prune_bound(float best) { return best + smax(best * 0.01f, 1e-6f); }

struct TraceHit {
    float t, u, v;
    int32_t prim;
    int32_t mesh;  // of `prim`: the .w of its first record, which the test loads anyway
};
// the bits of a float as an int, on the host and the device
LUCID_HD int32_t float_bits(float x) {
    int32_t i;
    __builtin_memcpy(&i, &x, 4);
    return i;
}

// Slab test of one child box given as scalars (no local arrays, so nothing
// is spilled to scratch).
LUCID_HD bool box_hit(v3f o, v3f d, v3f id, float x0, float y0, float z0, float x1, float y1,
                                        float z1, float& tentry) {
    float tmin = 0.0f, tmax = RAY_T_MAX;
    if (!slab1(d.x, id.x, x0 - o.x, x1 - o.x, tmin, tmax)) return false;
    if (!slab1(d.y, id.y, y0 - o.y, y1 - o.y, tmin, tmax)) return false;
    if (!slab1(d.z, id.z, z0 - o.z, z1 - o.z, tmin, tmax)) return false;
    tentry = tmin;
    return tmax >= smax(tmin, 0.0f);
}

// BVH traversal over the whole scene (top-level tree over mesh roots, then
// the per-mesh trees).  ANY = false: closest hit, ties in t resolve to the
// larger BVH-order position = the triangle the reference tests last.
// ANY = true: returns at the first triangle with tmin <= t <= tmax (the
// reference's closest-hit shadow query reports a hit exactly when one
// exists).  stk: this lane's LDS stack (stride kBlock), depth `sd`.
// Host execution (tests only) checks every node and triangle index against
// the array sizes before it is used; device code does not carry the checks.
struct HostBounds {
    int32_t n_nodes, n_tris;
    bool violated;
};
#if defined(__HIP_DEVICE_COMPILE__)
#define LUCID_HOST_CHECK(cond) ((void)0)
#define LUCID_HOST_CHECK_VOID(cond) ((void)0)
#else
#define LUCID_HOST_CHECK(cond)                               \
    do {                                                      \
        if (hb && !(cond)) {                                  \
            hb->violated = true;                              \
            return false;                                     \
        }                                                     \
    } while (0)
#define LUCID_HOST_CHECK_VOID(cond)                          \
    do {                                                      \
        if (hb && !(cond)) {                                  \
            hb->violated = true;                              \
            node = kPop;                                      \
            sp = 0;                                           \
            return;                                           \
        }                                                     \
    } while (0)
#endif
LUCID_HD void note_overflow(int32_t* ovf) {
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(ovf, 1);
#else
    ++*ovf;
#endif
}

// Traversal with postponed leaves, the form the persistent kernels run: a
// ray's state is the internal node to visit next (or kPop: take the next
// stack entry), at most one leaf waiting to be tested, and the stack, whose
// entries are internal nodes or leaves.  A node visit tests both child
// boxes, goes on to the nearer hit child (or holds it, if it is a leaf) and
// pushes the farther one; so a wavefront can walk nodes until most of its
// lanes hold a leaf and then test those leaves together, instead of leaving
// most lanes idle while a few test triangles.  The result is the closest-hit
// rule's (least t, then the larger BVH-order position), which does not
// depend on the order of the tests: only a box whose entry distance exceeds
// the best t so far by more than the 1 % pruning margin is skipped, and a
// leaf taken from the stack is tested without that check (more work, the
// same hit).
constexpr int32_t kPop = INT32_MIN;  // no leaf encodes to it (count 0)

LUCID_HD bool ray_done(int32_t node, int32_t leaf, int sp) { return node == kPop && leaf == 0 && sp == 0; }

template <bool ANY>
LUCID_HD void node_visit(const SceneDev& S, v3f o, v3f d, v3f id, int32_t* stk, int sd, int32_t* ovf, uint32_t& nn,
                         const TraceHit& h, int32_t& node, int32_t& leaf, int& sp, HostBounds* hb) {
    (void)hb;
    if (node == kPop) {
        if (sp == 0) return;
        --sp;
        const int32_t e = stk[sp * kBlock];
        if (e < 0) leaf = e;
        else node = e;
        return;
    }
    LUCID_HOST_CHECK_VOID(node >= 0 && node < hb->n_nodes);
    const float4* n4 = reinterpret_cast<const float4*>(S.nodes + node);
    const float4 b0 = n4[0], b1 = n4[1], b2 = n4[2];
    const int2 cc = *reinterpret_cast<const int2*>(n4 + 3);
    ++nn;
    const float lim = prune_bound(h.t);
    float te0 = 0.0f, te1 = 0.0f;
    const bool h0 = cc.x != kEmptyChild && box_hit(o, d, id, b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, te0) && te0 <= lim;
    const bool h1 = cc.y != kEmptyChild && box_hit(o, d, id, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w, te1) && te1 <= lim;
    const bool sw = h1 && (!h0 || te1 < te0);
    const int32_t cn = sw ? cc.y : cc.x, cf = sw ? cc.x : cc.y;
    const bool hn = sw ? h1 : h0, hf = sw ? h0 : h1;
    if (hn && hf) {
        if (sp < sd) {
            stk[sp * kBlock] = cf;
            ++sp;
        } else {
            note_overflow(ovf);
        }
    }
    const int32_t next = hn ? cn : (hf ? cf : kPop);
    node = kPop;
    if (next != kPop) {
        if (next < 0) leaf = next;
        else node = next;
    }
}

// Tests up to `batch` of the leaf's triangles (in order); `leaf` becomes the
// rest of the leaf, 0 when none is left.  True when an any-hit query has its
// answer.
template <bool ANY>
LUCID_HD bool leaf_visit(const SceneDev& S, v3f o, v3f d, float tmin, uint32_t& nt, TraceHit& h, int32_t& leaf,
                         int batch, HostBounds* hb) {
    (void)hb;
    const int32_t enc = ~leaf;
    const int32_t start = enc >> 3, all = enc & 7;
    const int32_t count = all < batch ? all : batch;
    leaf = count < all ? ~(((start + count) << 3) | (all - count)) : 0;
    nt += count;
    for (int i = 0; i < count; ++i) {
        const int32_t prim = start + i;
        LUCID_HOST_CHECK(prim >= 0 && prim < hb->n_tris);
        float t, u, v;
        const float4* tp = S.tri + 3 * (size_t)prim;
        if (tri_test(o, d, tp[0], tp[1], tp[2], tmin, h.t, t, u, v)) {
            if (ANY) {
                h.t = t; h.u = u; h.v = v; h.prim = prim; h.mesh = float_bits(tp[0].w);
                return true;
            }
            if (t < h.t || prim > h.prim) {
                h.t = t; h.u = u; h.v = v; h.prim = prim; h.mesh = float_bits(tp[0].w);
            }
        }
    }
    return false;
}

template <bool ANY>
LUCID_HD bool traverse(const SceneDev& S, v3f o, v3f d, float tmin, float tmax, int32_t* stk, int sd, int32_t* ovf,
                       uint32_t& nn, uint32_t& nt, TraceHit& h, HostBounds* hb = nullptr) {
    h.t = tmax; h.u = 0.0f; h.v = 0.0f; h.prim = -1;
    const v3f id = inv_dir(d);
    int sp = 0;
    int32_t node = S.root, leaf = 0;
    while (!ray_done(node, leaf, sp)) {
        if (leaf) {
            if (leaf_visit<ANY>(S, o, d, tmin, nt, h, leaf, 4, hb)) break;
        } else {
            node_visit<ANY>(S, o, d, id, stk, sd, ovf, nn, h, node, leaf, sp, hb);
        }
#if !defined(__HIP_DEVICE_COMPILE__)
        if (hb && hb->violated) return false;
#endif
    }
    return h.prim >= 0;
}

__device__ __forceinline__ int32_t prim_mesh(const SceneDev& S, int32_t prim) {
    return __float_as_int(S.tri[3 * (size_t)prim].w);
}

// intersect_skipping_volume_boundaries (path_tracer.hpp:82-95): the closest
// real surface; boundary-only meshes are skipped by re-querying past them.
__device__ __forceinline__ TraceHit closest_real(const SceneDev& S, v3f o, v3f d, float tmin, int32_t* stk,
                                                 int sd, int32_t* ovf, uint32_t& nn, uint32_t& nt) {
    TraceHit h;
    traverse<false>(S, o, d, tmin, RAY_T_MAX, stk, sd, ovf, nn, nt, h);
    if (!S.any_vbo) return h;
    float smin_ = tmin;
    for (int peek = 0; peek < 8; ++peek) {
        if (h.prim < 0) return h;
        if (!(S.mats[h.mesh].flags & MF_VOLUME_BOUNDARY)) return h;
        smin_ = h.t + 1e-3f;
        traverse<false>(S, o, d, smin_, RAY_T_MAX, stk, sd, ovf, nn, nt, h);
    }
    h.prim = -1;
    return h;
}
// is_shadowed_skipping_volume_boundaries (path_tracer.hpp:103-119)
__device__ __forceinline__ bool shadowed(const SceneDev& S, v3f o, v3f d, float tmax, int32_t* stk, int sd,
                                         int32_t* ovf, uint32_t& nn, uint32_t& nt) {
    TraceHit h;
    if (!S.any_vbo) return traverse<true>(S, o, d, RAY_T_MIN, tmax, stk, sd, ovf, nn, nt, h);
    float toff = 0.0f;
    v3f po = o;
    for (int peek = 0; peek < 8; ++peek) {
        float remaining = tmax - toff;
        if (remaining <= 0.0f) return false;
        traverse<false>(S, po, d, RAY_T_MIN, remaining, stk, sd, ovf, nn, nt, h);
        if (h.prim < 0) return false;
        if (!(S.mats[h.mesh].flags & MF_VOLUME_BOUNDARY)) return true;
        float skip = h.t + 1e-3f;
        toff += skip;
        po = po + d * skip;
    }
    return false;
}


// ------------------------------------------------------------------ media --
__device__ __forceinline__ float grid_value(const SceneDev& S, const GVolume& v, float x, float y, float z) {
    return S.grid16 ? grid_trilinear(v, S.grid16, x, y, z) : grid_trilinear(v, S.grid, x, y, z);
}

// transmittance_along_segment (absorption, transmittance.hpp:77-128) and
// extinction_transmittance_along_segment (:145-209): midpoint ray march at
// half a voxel over every gridded medium, T = exp(-tau) per channel.
__device__ __noinline__ v3f vol_transmittance(const SceneDev& S, v3f o, v3f d, float tmax, bool extinction) {
    float tr = 0.0f, tg = 0.0f, tb = 0.0f;
    for (int vi = 0; vi < S.n_vols; ++vi) {
        const GVolume& v = S.vols[vi];
        if (!v.has_grid) continue;
        float tin = 0.0f, tout = 0.0f;
        if (!ray_aabb_segment(o, d, tmax, v.wmin, v.wmax, tin, tout)) continue;
        const float seg = tout - tin;
        if (seg <= 0.0f) continue;
        const float ext = max3(v.wmax[0] - v.wmin[0], v.wmax[1] - v.wmin[1], v.wmax[2] - v.wmin[2]);
        const int md = max(max(v.dims[0], v.dims[1]), v.dims[2]);
        const float vox = ext / (float)max(1, md);
        const float step = smax(1e-4f, vox * 0.5f);
        const int n = min(2048, (int)(seg / step) + 1);
        const float er = 1.0f + v.absorption[0], eg = 1.0f + v.absorption[1], eb = 1.0f + v.absorption[2];
        for (int i = 0; i < n; ++i) {
            const float t = tin + ((float)i + 0.5f) * step;
            if (t > tout) break;
            const float dd = grid_value(S, v, o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
            if (dd <= 0.0f) continue;
            const float base = v.density * dd * step;
            if (extinction) {
                tr += base * er; tg += base * eg; tb += base * eb;
            } else {
                tr += base * (1.0f - v.color[0]); tg += base * (1.0f - v.color[1]); tb += base * (1.0f - v.color[2]);
            }
        }
    }
    return mk(expf(-tr), expf(-tg), expf(-tb));
}

// The factors of a traceVolumeMIS NEE contribution whose transmittance is
// marched later by k_vol_march.
__device__ __forceinline__ void defer_transmittance(const WaveDev& W, int p, v3f a, v3f e, v3f thr) {
    W.sh_a[p] = make_float4(a.x, a.y, a.z, 0.0f);
    W.sh_e[p] = make_float4(e.x, e.y, e.z, 0.0f);
    W.sh_t[p] = make_float4(thr.x, thr.y, thr.z, 0.0f);
}

// sample_volume_distance (transmittance.hpp:220-289): Woodcock / delta
// tracking against the max-density majorant of each gridded medium, in mesh
// order, at most 4096 steps; draws from the path's PCG stream in the
// reference's order.  Returns the medium index of the nearest real collision.
__device__ __noinline__ int vol_sample_distance(const SceneDev& S, v3f o, v3f d, float tmax, Pcg& rng, float& t_out) {
    float best = tmax;
    int hit = -1;
    for (int vi = 0; vi < S.n_vols; ++vi) {
        const GVolume& v = S.vols[vi];
        if (!v.has_grid) continue;
        if (v.density <= 0.0f || v.grid_max <= 0.0f) continue;
        float tin = 0.0f, tout = 0.0f;
        if (!ray_aabb_segment(o, d, tmax, v.wmin, v.wmax, tin, tout)) continue;
        if (tin >= best) continue;
        const float seg_end = smin(tout, best);
        if (seg_end <= tin) continue;
        const float emax = max3(1.0f + v.absorption[0], 1.0f + v.absorption[1], 1.0f + v.absorption[2]);
        const float sig = v.density * v.grid_max * emax;
        if (sig <= 0.0f) continue;
        float t = tin;
        bool sc = false;
        for (int it = 0; it < 4096; ++it) {
            const float us = smax(1e-12f, 1.0f - randf(rng));
            t += -logf(us) / sig;
            if (t >= seg_end) break;
            const float dd = grid_value(S, v, o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
            const float st = v.density * dd * emax;
            if (randf(rng) * sig < st) {
                sc = true;
                break;
            }
        }
        if (sc && t < best) {
            best = t;
            hit = vi;
        }
    }
    t_out = best;
    return hit;
}

// ---------------------------------------------------------------- raygen --
// Camera ray of render_tile (pybind_renderer.hpp:253-293).
__device__ __forceinline__ v3f cam_dir(const CamDev& c, float nx, float ny) {
    v3f d = mk(c.fwd[0] + c.right[0] * (nx * c.scale) + c.up[0] * (ny * c.scale),
               c.fwd[1] + c.right[1] * (nx * c.scale) + c.up[1] * (ny * c.scale),
               c.fwd[2] + c.right[2] * (nx * c.scale) + c.up[2] * (ny * c.scale));
    return make_dir_or(d, mk(c.fwd[0], c.fwd[1], c.fwd[2]));
}

// One camera path: wave slot p, pixel `pix` of the tile, sample index s.
// meta.w and the direction's w carry `tag` and `dw` (0 for render_tile;
// the session and its BVH root in fused session passes).
__device__ __forceinline__ void raygen_path(const WaveDev& W, const CamDev& c, int p, int pix, int s, int tag,
                                            float dw) {
    int px = pix % c.tw, py = pix / c.tw;
    int gx = c.tx + px, gy = c.ty + py;
    float ndc_x = (2.0f * (gx + 0.5f) / c.fw - 1.0f) * c.aspect;
    float ndc_y = 1.0f - 2.0f * (gy + 0.5f) / c.fh;
    Pcg rng = pcg_seed((uint32_t)gx, (uint32_t)gy, (uint32_t)s);
    v3f d;
    if (c.samples_in_call > 1) {
        float jx = (randf(rng) - 0.5f) / c.fw;
        float jy = (randf(rng) - 0.5f) / c.fh;
        d = cam_dir(c, ndc_x + jx, ndc_y + jy);
    } else {
        d = cam_dir(c, ndc_x, ndc_y);
    }
    W.ray_o[p] = make_float4(c.pos[0], c.pos[1], c.pos[2], 0.0f);
    W.ray_d[p] = make_float4(d.x, d.y, d.z, dw);
    W.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    W.rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    W.rng[p] = make_ulonglong2(rng.state, rng.inc);
    W.meta[p] = make_int4(pix, s, 0, tag);
    W.q[0][p] = p;
    if (W.dcnt) W.dcnt[p] = 0;
    if (keeps_record(W, p)) {
        const int r = p / W.tstride;
        for (int k = 0; k < kTrainSlots; ++k) {
            *tslot_at(W, r, k, 3) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            *tslot_at(W, r, k, 4) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
}
__global__ __launch_bounds__(kBlock) void k_raygen(WaveDev W, CamDev c, int32_t n_paths) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    if (p == 0) { W.cnt[0] = n_paths; W.cnt[1] = 0; W.cnt[2] = 0; }
    if (blockIdx.x == 0) reset_fetch(W);
    if (p >= n_paths) return;
    int b = p / c.P, pix = p - b * c.P;
    raygen_path(W, c, p, pix, (c.s0p ? *c.s0p : c.s0) + b, 0, 0.0f);
}

// ------------------------------------------------------------- intersect --
__global__ __launch_bounds__(kBlock) void k_intersect(SceneDev S, WaveDev W, int qin, int sd) {
    extern __shared__ int32_t lds_stack[];
    const int n = W.cnt[qin];
    if (blockIdx.x == 0) reset_append(W);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        W.cnt[1 - qin] = 0;
        W.cnt[2] = 0;
        atomicAdd(&W.rays[0], (unsigned long long)n);
        poll_stop(W);
    }
    uint32_t nn = 0, nt = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int p = W.q[qin][i];
        const float4 o = W.ray_o[p], d = W.ray_d[p];
        TraceHit h = closest_real(S, mk(o), mk(d), o.w, lds_stack + threadIdx.x, sd, &W.cnt[3], nn, nt);
        W.hit[p] = make_float4(h.t, __int_as_float(h.prim), h.u, h.v);
    }
    if (W.count_work && (nn | nt)) {
        atomicAdd(&W.rays[2], (unsigned long long)nn);
        atomicAdd(&W.rays[3], (unsigned long long)nt);
    }
}

// Persistent traversal with ray refill.  A grid of a few workgroups per CU
// walks the queue; every lane traces one ray at a time, one node per loop
// iteration, and a wavefront fetches new rays for its idle lanes (one atomic
// per fetch) once `refill` or more of them are idle.  Lanes of a wavefront
// thus stay busy while their rays take different numbers of steps, instead
// of the wavefront waiting for its longest ray.  Fetch counters: cnt[4]
// (closest hit) and cnt[5] (shadow), zeroed by k_raygen and k_shade.
// MODE 0: closest hit with volume-boundary re-queries (closest_real);
// MODE 1: shadow any-hit; MODE 2: shadow past boundary-only meshes
// (shadowed with re-queries from a moving origin).  Per ray the steps and
// their order are those of closest_real / shadowed.
template <bool FUSED = false>
__device__ __forceinline__ void shadow_unoccluded(const WaveDev& W, int p) {
    const float4 L = W.rad[p], c = W.sh_c[p];
    W.rad[p] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, 0.0f);
    if (FUSED) return;  // the direction's w holds the BVH root there
    // diagnostic runs: the pending NEE record of this bounce is kept
    const int32_t tag = __float_as_int(W.sh_d[p].w);
    if (tag > 0) {
        float4* r = W.dr0 + (size_t)(tag - 1) * (size_t)W.dstride + (size_t)p;
        r->w = __int_as_float(__float_as_int(r->w) | kRecValid);
    }
}

// FUSED: a fused session pass over the sessions' merged scenes; every ray
// starts at its own scene's BVH root, stored in its direction's w.
template <int MODE, bool FUSED = false>
__global__ __launch_bounds__(kBlock) void k_trace_refill(SceneDev S, WaveDev W, int qin, int sd, int refill,
                                                         int walk_min, int leaf_tris) {
    constexpr bool SHADOW = MODE != 0, ANY = MODE == 1;
    extern __shared__ int32_t lds_stack[];
    int32_t* stk = lds_stack + threadIdx.x;
    const int n = SHADOW ? W.cnt[2] : W.cnt[qin];
    int32_t* fetch = fetch_counters(W, SHADOW ? 1 : 0);
    RangeFetch rf = first_range();
    if (!SHADOW && blockIdx.x == 0) reset_append(W);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (SHADOW) {
            atomicAdd(&W.rays[1], (unsigned long long)n);
        } else {
            W.cnt[1 - qin] = 0;
            W.cnt[2] = 0;
            atomicAdd(&W.rays[0], (unsigned long long)n);
            poll_stop(W);
        }
    }
    uint32_t nn = 0, nt = 0;
    int p = -1;  // path of this lane's ray, -1 idle
    bool exhausted = false;  // wavefront-uniform: the queue has been handed out
    v3f o = mk(0.0f, 0.0f, 0.0f), d = o, id = o;
    float tmin = 0.0f, tmax = 0.0f, toff = 0.0f;  // MODE 2: shadow distance, offset of the moving origin
    TraceHit h;
    h.t = 0.0f; h.u = 0.0f; h.v = 0.0f; h.prim = -1;
    int32_t node = kPop, leaf = 0, root = S.root;
    int sp = 0, requery = 0;
    for (;;) {
        const unsigned long long im = __ballot(p < 0);
        const int nidle = __popcll(im);
        if (!exhausted && (nidle >= refill || nidle == 64)) {
            int rs, len;
            const int base = range_fetch(fetch, n, rf, im, nidle, exhausted, rs, len);
            if (p < 0) {
                const int j = base + __popcll(im & ((1ull << lane_id()) - 1ull));
                if (j < len) {
                    const int i = range_pos(rs, j);
                    if (SHADOW) {
                        p = W.qs[i];
                        const float4 o4 = W.sh_o[p], d4 = W.sh_d[p];
                        o = mk(o4); d = mk(d4);
                        tmin = RAY_T_MIN;
                        tmax = o4.w;
                        toff = 0.0f;
                        h.t = o4.w;
                        if (FUSED) root = __float_as_int(d4.w);
                    } else {
                        p = W.q[qin][i];
                        const float4 o4 = W.ray_o[p], d4 = W.ray_d[p];
                        o = mk(o4); d = mk(d4);
                        tmin = o4.w;
                        h.t = RAY_T_MAX;
                        if (FUSED) root = __float_as_int(d4.w);
                    }
                    id = inv_dir(d);
                    h.u = 0.0f; h.v = 0.0f; h.prim = -1;
                    node = root;
                    leaf = 0;
                    sp = 0;
                    requery = 0;
                    if (MODE == 2 && !(tmax > 0.0f)) {  // nothing left to test: unoccluded
                        shadow_unoccluded<FUSED>(W, p);
                        p = -1;
                    }
                }
            }
        }
        if (__ballot(p >= 0) == 0) {
            if (exhausted) break;
            continue;
        }
        // node phase: lanes without a waiting leaf walk nodes, for as long as
        // more than walk_min lanes of the wavefront still do (and at least
        // until one lane holds a leaf, so every pass makes progress); then
        // the waiting leaves are tested together
        for (;;) {
            const bool walk = p >= 0 && leaf == 0 && !ray_done(node, leaf, sp);
            const unsigned long long wm = __ballot(walk);
            if (wm == 0) break;
            if (__popcll(wm) <= walk_min && __ballot(p >= 0 && leaf != 0) != 0) break;
            if (walk) {
                node_visit<ANY>(S, o, d, id, stk, sd, &W.cnt[3], nn, h, node, leaf, sp, nullptr);
                // a second visit per pass halves the ballots and branches
                // (1-2 % on the room and the icospheres)
                if (leaf == 0 && !ray_done(node, leaf, sp))
                    node_visit<ANY>(S, o, d, id, stk, sd, &W.cnt[3], nn, h, node, leaf, sp, nullptr);
            }
        }
        if (p < 0) continue;
        if (leaf != 0) {
            const bool found = leaf_visit<ANY>(S, o, d, tmin, nt, h, leaf, leaf_tris, nullptr);
            if (found) leaf = 0;
            if (!found && !ray_done(node, leaf, sp)) continue;
        } else if (!ray_done(node, leaf, sp)) {
            continue;
        }
        // the ray's query has ended (stack empty, or an any-hit found)
        if (MODE == 0) {
            // intersect_skipping_volume_boundaries: re-query past a
            // boundary-only mesh; the result of an 8th re-query is discarded
            // by the reference, so a 7th boundary hit ends with no hit
            if (S.any_vbo && h.prim >= 0 && (S.mats[h.mesh].flags & MF_VOLUME_BOUNDARY)) {
                if (requery < 7) {
                    ++requery;
                    tmin = h.t + 1e-3f;
                    h.t = RAY_T_MAX; h.u = 0.0f; h.v = 0.0f; h.prim = -1;
                    node = root;
                    leaf = 0;
                    sp = 0;
                    continue;
                }
                h.prim = -1;
            }
            W.hit[p] = make_float4(h.t, __int_as_float(h.prim), h.u, h.v);
        } else if (MODE == 1) {
            if (h.prim < 0) shadow_unoccluded<FUSED>(W, p);
        } else {
            // is_shadowed_skipping_volume_boundaries: a real surface within
            // the remaining distance occludes; a boundary-only one moves the
            // origin 1e-3 past its hit, at most 8 queries in all
            bool unoccluded = h.prim < 0;
            if (!unoccluded && (S.mats[h.mesh].flags & MF_VOLUME_BOUNDARY)) {
                const float skip = h.t + 1e-3f;
                toff += skip;
                o = o + d * skip;
                ++requery;
                const float remaining = tmax - toff;
                if (requery < 8 && remaining > 0.0f) {
                    h.t = remaining; h.u = 0.0f; h.v = 0.0f; h.prim = -1;
                    node = root;
                    leaf = 0;
                    sp = 0;
                    continue;
                }
                unoccluded = true;
            }
            if (unoccluded) shadow_unoccluded<FUSED>(W, p);
        }
        p = -1;
    }
    if (W.count_work && (nn | nt)) {
        atomicAdd(&W.rays[SHADOW ? 4 : 2], (unsigned long long)nn);
        atomicAdd(&W.rays[SHADOW ? 5 : 3], (unsigned long long)nt);
    }
}

// Extinction transmittance of the queued traceVolumeMIS shadow rays
// (vol_transmittance with extinction, the same march in the same order),
// persistent with refill like the traversal: a lane marches up to kMarchChunk
// steps per loop iteration, so lanes whose rays cross long stretches of
// medium do not hold back the rest of their wavefront.  Then
// sh_c = (sh_a * (sh_e * T)) * sh_t.  Fetch counter cnt[6].
constexpr int kMarchChunk = 8;
__global__ __launch_bounds__(kBlock) void k_vol_march(SceneDev S, WaveDev W, int refill) {
    const int n = W.cnt[2];
    int32_t* fetch = fetch_counters(W, 2);
    RangeFetch rf = first_range();
    int p = -1;
    bool exhausted = false;
    v3f o = mk(0.0f, 0.0f, 0.0f), d = o;
    float tmax = 0.0f, tin = 0.0f, tout = 0.0f, step = 0.0f, er = 0.0f, eg = 0.0f, eb = 0.0f;
    float tr = 0.0f, tg = 0.0f, tb = 0.0f;
    int vi = 0, i = 0, ns = 0;
    for (;;) {
        const unsigned long long im = __ballot(p < 0);
        const int nidle = __popcll(im);
        if (!exhausted && (nidle >= refill || nidle == 64)) {
            int rs, len;
            const int base = range_fetch(fetch, n, rf, im, nidle, exhausted, rs, len);
            if (p < 0) {
                const int j = base + __popcll(im & ((1ull << lane_id()) - 1ull));
                if (j < len) {
                    p = W.qs[range_pos(rs, j)];
                    const float4 o4 = W.sh_o[p], d4 = W.sh_d[p];
                    o = mk(o4); d = mk(d4);
                    tmax = o4.w;
                    tr = tg = tb = 0.0f;
                    vi = -1;
                    i = ns = 0;  // no medium entered yet
                }
            }
        }
        if (__ballot(p >= 0) == 0) {
            if (exhausted) break;
            continue;
        }
        if (p < 0) continue;
        if (i >= ns) {
            // next gridded medium the shadow segment crosses, or done
            ++vi;
            for (; vi < S.n_vols; ++vi) {
                const GVolume& v = S.vols[vi];
                if (!v.has_grid) continue;
                if (!ray_aabb_segment(o, d, tmax, v.wmin, v.wmax, tin, tout)) continue;
                const float seg = tout - tin;
                if (seg <= 0.0f) continue;
                const float ext = max3(v.wmax[0] - v.wmin[0], v.wmax[1] - v.wmin[1], v.wmax[2] - v.wmin[2]);
                const int md = max(max(v.dims[0], v.dims[1]), v.dims[2]);
                const float vox = ext / (float)max(1, md);
                step = smax(1e-4f, vox * 0.5f);
                ns = min(2048, (int)(seg / step) + 1);
                er = 1.0f + v.absorption[0]; eg = 1.0f + v.absorption[1]; eb = 1.0f + v.absorption[2];
                i = 0;
                break;
            }
            if (vi >= S.n_vols) {
                const v3f T = mk(expf(-tr), expf(-tg), expf(-tb));
                const float4 a = W.sh_a[p], e = W.sh_e[p], t = W.sh_t[p];
                const v3f cc = cmul(cmul(mk(a), cmul(mk(e), T)), mk(t));
                W.sh_c[p] = make_float4(cc.x, cc.y, cc.z, 0.0f);
                p = -1;
                continue;
            }
        }
        const GVolume& v = S.vols[vi];
        for (int k = 0; k < kMarchChunk && i < ns; ++k) {
            const float t = tin + ((float)i + 0.5f) * step;
            if (t > tout) {
                i = ns;
                break;
            }
            ++i;
            const float dd = grid_value(S, v, o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
            if (dd <= 0.0f) continue;
            const float base = v.density * dd * step;
            tr += base * er; tg += base * eg; tb += base * eb;
        }
    }
}

// ----------------------------------------------------------------- shade --
struct NativeHit {
    bool hit;
    float t;
    v3f point, normal, emission;
    int32_t index;  // native light index (JSON order)
};
// intersectNativeLights (scene_lights.hpp:176-246); only area and sphere
// lights are visible, and only for depth > 0.
__device__ __forceinline__ NativeHit native_lights(const SceneDev& S, v3f o, v3f d) {
    NativeHit r;
    r.hit = false; r.t = 1e30f; r.index = -1;
    for (int i = 0; i < S.n_native; ++i) {
        const GLight& l = S.lights[S.native_ids[i]];
        v3f c = mk(l.pos[0], l.pos[1], l.pos[2]);
        if (l.type == 1) {
            if (!(l.radius > 1e-6f)) continue;
            // intersectSphere (intersection.hpp:58-104)
            v3f oc = o - c;
            float bq = 2.0f * dot(oc, d);
            float cq = len2(oc) - l.radius * l.radius;
            float disc = bq * bq - 4.0f * 1.0f * cq;
            if (disc < 0.0f) continue;
            float sq = sqrtf(smax(0.0f, disc));
            float t = (-bq - sq) / (2.0f * 1.0f);
            if (t < RAY_T_MIN) {
                t = (-bq + sq) / (2.0f * 1.0f);
                if (t < RAY_T_MIN) continue;
            }
            if (t < r.t) {
                r.hit = true; r.t = t; r.point = o + d * t;
                r.normal = normalize_plain(r.point - c);
                r.emission = mk(l.emission[0], l.emission[1], l.emission[2]);
                r.index = i;
            }
        } else if (l.type == 4) {
            // intersectRectangle / intersectEllipse (intersection.hpp:128-236)
            v3f nrm = mk(l.normal[0], l.normal[1], l.normal[2]);
            float den = dot(d, nrm);
            if (fabsf(den) < 1e-6f) continue;
            float t = dot(c - o, nrm) / den;
            if (t < RAY_T_MIN) continue;
            v3f hp = o + d * t;
            v3f lh = hp - c;
            float x = dot(lh, mk(l.right[0], l.right[1], l.right[2]));
            float y = dot(lh, mk(l.up[0], l.up[1], l.up[2]));
            if (l.shape == 2 || l.shape == 3) {
                float nx = x / (l.sx * 0.5f), ny = y / (l.sy * 0.5f);
                if (nx * nx + ny * ny > 1.0f) continue;
            } else {
                float hx = l.sx * 0.5f, hy = l.sy * 0.5f;
                if (x * x > hx * hx || y * y > hy * hy) continue;
            }
            if (t < r.t && dot(nrm, d) < 0.0f) {
                r.hit = true; r.t = t; r.point = hp; r.normal = nrm;
                r.emission = mk(l.emission[0], l.emission[1], l.emission[2]);
                r.index = i;
            }
        }
    }
    return r;
}

// selectLight (scene_lights.hpp:130-144): std::lower_bound on the CDF.
__device__ __forceinline__ int select_light(const SceneDev& S, float u, float& prob) {
    int lo = 0, hi = S.n_lights;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (S.cdf[mid] < u) lo = mid + 1; else hi = mid;
    }
    int idx = min(lo, S.n_lights - 1);
    prob = area_ratio(S.lights[idx].area, S.total_area);
    return idx;
}

__device__ __forceinline__ bool finite3(v3f a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// Internal algorithm ids of the diagnostic tracers (traceSimpleWithDiagnostics,
// traceNEEWithDiagnostics, traceMISWithDiagnostics; path_tracer.hpp:880-1514).
constexpr int kAlgoDiagBase = 16;
constexpr int kAlgoDiagSimple = kAlgoDiagBase + LUCID_ALGO_SIMPLE;
constexpr int kAlgoDiagNee = kAlgoDiagBase + LUCID_ALGO_NEE;
constexpr int kAlgoDiagMis = kAlgoDiagBase + LUCID_ALGO_MIS;

// classify_bsdf / is_delta_bsdf (diagnostic_integrator.hpp:24-60) on the
// material parameters after getMaterialParams.
__device__ __forceinline__ uint8_t classify_bsdf(const Mat& m) {
    if (m.transmission > 0.5f) return diag::kGlass;
    if (m.roughness < 0.01f && m.metallic > 0.5f) return diag::kMirror;
    if (m.metallic > 0.5f || m.roughness < 0.5f) return diag::kGlossy;
    return diag::kDiffuse;
}
__device__ __forceinline__ bool is_delta_bsdf(const Mat& m) {
    return (m.transmission > 0.5f && m.roughness < 0.01f) || (m.metallic > 0.5f && m.roughness < 0.01f);
}
// to_light_source_type (path_types.hpp:160-169)
__device__ __forceinline__ uint8_t light_source_type(int32_t t) {
    switch (t) {
        case 0: return diag::kLightArea;
        case 1: return diag::kLightPoint;
        case 2: return diag::kLightDirectional;
        case 3: return diag::kLightSpot;
        case 4: return diag::kLightArea;
        default: return diag::kLightUnknown;
    }
}
// Record slot `slot` of path p: a completed path (record_completed_path_with_
// geometry, diagnostic_integrator.hpp:338-368) = the path's first `prefix`
// vertices plus a light vertex {light_obj, -1, Emission, 0} at lpos.
__device__ __forceinline__ void put_record(const WaveDev& W, int p, int slot, v3f c, uint8_t strategy,
                                           uint8_t light_type, int prefix, int32_t light_obj, v3f lpos, bool valid) {
    const int32_t meta = (valid ? kRecValid : 0) | ((int32_t)strategy << 8) | ((int32_t)light_type << 16) |
                         ((int32_t)prefix << 24);
    const size_t k = (size_t)slot * (size_t)W.dstride + (size_t)p;
    W.dr0[k] = make_float4(c.x, c.y, c.z, __int_as_float(meta));
    W.dr1[k] = make_float4(lpos.x, lpos.y, lpos.z, __int_as_float(light_obj));
}

// The NEE contribution of a guided vertex once the guide's density at the
// light sample's direction is known: c = bsdf_weight * emission, weighed by
// the power heuristic of the light pdf pl against the density q of the
// vertex's hemisphere technique, times the throughput.
__device__ __forceinline__ void finish_guided_nee(const WaveDev& W, int p, float q, float pl, v3f c, v3f thr) {
    const float mw = power_heuristic(pl, q);
    const v3f cc = cmul(c * mw, thr);
    W.sh_c[p] = make_float4(cc.x, cc.y, cc.z, 0.0f);
}

// One bounce of traceSimple / traceNEE / traceMIS (path_tracer.hpp:128-570),
// their volume variants (:209-283, 586-835), or the diagnostic tracers.
// NODES: the scene has node-tree programs (separate instantiations keep
// their registers out of the kernels of scenes without them).
// FUSED: a fused session pass; S is the path's session's view of the merged
// scenes (see batch_view), the path keeps its session in meta.w and its
// scene's BVH root in the w of its ray and shadow-ray directions.
template <int ALGO, bool NODES, bool FUSED = false>
__device__ __forceinline__ void shade_one(const SceneDev& S, const WaveDev& W, const guide::GuideDev& G, int qin,
                                          int32_t max_depth, int i, int n, bool stop) {
    constexpr bool DIAG = ALGO >= kAlgoDiagBase;
    constexpr bool GUIDED = ALGO == LUCID_ALGO_GUIDED_MIS;
    constexpr int A = DIAG ? ALGO - kAlgoDiagBase : (GUIDED ? LUCID_ALGO_MIS : ALGO);
    constexpr bool MIS = A == LUCID_ALGO_MIS || A == LUCID_ALGO_VOLUME_MIS;
    constexpr bool SIMPLE = A == LUCID_ALGO_SIMPLE || A == LUCID_ALGO_VOLUME_SIMPLE;
    bool live = i < n;
    bool cont = false, nee = false;
    int p = live ? W.q[qin][i] : 0;
    if (live) {
        float4 ro = W.ray_o[p], rdv = W.ray_d[p], tv = W.thr[p];
        int4 meta = W.meta[p];
        float4 hv = W.hit[p];
        v3f o = mk(ro), d = mk(rdv), thr = mk(tv);
        float last_pdf = tv.w;
        // guided runs: the previous bounce's plain BSDF pdf and 1 - rho (see
        // the guided BSDF sampling below); a path from raygen or a plain
        // bounce carries rho = 1
        const float last_pb = GUIDED ? rdv.w : 0.0f;
        const float rho = GUIDED ? 1.0f - __int_as_float(meta.w) : 1.0f;
        // MIS weight of an emitter hit of the previous bounce's BSDF strategy
        // against light pdf lp (power heuristic; 1 without a pdf); guided:
        // rho w(q) + (1 - rho) w(pb)
        const auto hit_mis = [&](float lp) {
            float w = (last_pdf > MIN_PDF && lp > MIN_PDF) ? power_heuristic(last_pdf, lp) : 1.0f;
            if (GUIDED && rho < 1.0f) {
                const float wp = (last_pb > MIN_PDF && lp > MIN_PDF) ? power_heuristic(last_pb, lp) : 1.0f;
                w = rho * w + (1.0f - rho) * wp;
            }
            return w;
        };
        int depth = meta.z;
        int32_t prim = __float_as_int(hv.y);
        ulonglong2 rs = W.rng[p];
        Pcg rng;
        rng.state = rs.x; rng.inc = rs.y;
        float4 Lv = W.rad[p];
        v3f L = mk(Lv);
        // guided runs: the radiance gathered up to here closes the previous
        // bounce's training slot (everything after it came through wi)
        int trec = -1;  // training record of this path, if it keeps one
        if (GUIDED && keeps_record(W, p)) {
            trec = p / W.tstride;
            if (depth >= 1 && depth - 1 < kTrainSlots)
                *tslot_at(W, trec, depth - 1, 4) = make_float4(L.x, L.y, L.z, 1.0f);
        }
        // diagnostic path: vertex count so far; this bounce's record slot
        // starts empty (it may hold a record of an earlier batch)
        int nvert = 0;
        if (DIAG) {
            nvert = W.dcnt[p] & 0xff;
            W.dr0[(size_t)depth * (size_t)W.dstride + (size_t)p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        bool has_hit = prim >= 0;
        NativeHit lh;
        lh.hit = false;
        if (depth > 0 && S.n_native > 0) lh = native_lights(S, o, d);
        bool done = false;
        constexpr bool VOLUME = A == LUCID_ALGO_VOLUME_MIS || A == LUCID_ALGO_VOLUME_SIMPLE;
        if (VOLUME) {
            float ts = 1e6f;
            if (has_hit) ts = smin(ts, hv.x);
            if (lh.hit) ts = smin(ts, lh.t);
            if (A == LUCID_ALGO_VOLUME_SIMPLE) {
                // traceVolumeSimple (path_tracer.hpp:226-230): absorption only
                if (S.n_vols > 0) thr = cmul(thr, vol_transmittance(S, o, d, ts, false));
            } else if (S.n_vols > 0) {
                float vt;
                const int vi = vol_sample_distance(S, o, d, ts, rng, vt);
                if (vi >= 0) {
                    // scatter event (path_tracer.hpp:617-699)
                    done = true;
                    const GVolume& vm = S.vols[vi];
                    thr = cmul(thr, mk(vm.color[0] / (1.0f + vm.absorption[0]), vm.color[1] / (1.0f + vm.absorption[1]),
                                       vm.color[2] / (1.0f + vm.absorption[2])));
                    const v3f x = o + d * vt;
                    const v3f wo = -d;
                    const float g = vm.anisotropy;
                    if (S.n_lights > 0) {
                        float sp;
                        const int li = select_light(S, randf(rng), sp);
                        if (sp > 1e-6f) {
                            const float a1 = randf(rng), a2 = randf(rng);
                            const LSample ls = sample_light(S.lights[li], x, a1, a2);
                            if (ls.pdf > MIN_PDF) {
                                const float ct = dot(wo, ls.dir);
                                const float pe = hg_eval(g, ct);
                                const float pp = hg_eval(g, ct);
                                const float pl = ls.pdf * sp;
                                const float mw = power_heuristic(pl, pp);
                                const v3f pw = bsdf_weight(mk(pe, pe, pe), 1.0f, pl);
                                // the transmittance march runs in k_vol_march
                                defer_transmittance(W, p, pw * mw, ls.emission, thr);
                                W.sh_o[p] = make_float4(x.x, x.y, x.z, ls.dist);
                                W.sh_d[p] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, FUSED ? __int_as_float(S.root) : 0.0f);
                                nee = true;
                            }
                        }
                    }
                    const float a1 = randf(rng), a2 = randf(rng);
                    float ppdf;
                    const v3f wi = hg_sample(g, wo, a1, a2, ppdf);
                    last_pdf = ppdf;
                    bool ok = true;
                    if (depth >= 3) {
                        const float mx = max3(thr.x, thr.y, thr.z);
                        const float rr = smin(mx, 0.95f);
                        if (randf(rng) > rr) ok = false;
                        else thr = thr * (1.0f / rr);
                    }
                    if (ok && finite3(thr) && depth + 1 < max_depth) {
                        cont = true;
                        W.ray_o[p] = make_float4(x.x, x.y, x.z, RAY_T_MIN);
                        W.ray_d[p] = make_float4(wi.x, wi.y, wi.z, FUSED ? __int_as_float(S.root) : 0.0f);
                        W.thr[p] = make_float4(thr.x, thr.y, thr.z, last_pdf);
                        W.rng[p] = make_ulonglong2(rng.state, rng.inc);
                        W.meta[p] = make_int4(meta.x, meta.y, depth + 1, FUSED ? meta.w : 0);
                    }
                }
            }
        }
        if (done) {
        } else if (lh.hit && (!has_hit || lh.t < hv.x)) {
            // the NEE diagnostic tracer counts light hits of camera rays only
            if (!(DIAG && A == LUCID_ALGO_NEE && depth != 0)) {
                v3f c = cmul(lh.emission, thr);
                if (MIS) {
                    float lp = 0.0f;
                    if (S.n_lights > 0 && lh.index >= 0) {
                        int sli = S.native_ids[lh.index];
                        float sp = area_ratio(S.lights[sli].area, S.total_area);
                        if (sp > 0.0f) lp = pdf_light(S.lights[sli].area, o, lh.point, lh.normal) * sp;
                    }
                    float mw = hit_mis(lp);
                    c = c * mw;
                }
                L = L + c;
                if (DIAG)
                    put_record(W, p, depth, c, MIS ? diag::kMisBSDF : diag::kBSDF,
                               light_source_type(S.lights[S.native_ids[lh.index]].type), nvert, -(lh.index + 2),
                               o + d * lh.t, true);
            }
            done = true;
        } else if (!has_hit) {
            if (!(DIAG && A == LUCID_ALGO_NEE && depth != 0)) {
                const v3f c = cmul(mk(S.env[0], S.env[1], S.env[2]), thr);
                L = L + c;
                if (DIAG)
                    put_record(W, p, depth, c, MIS ? diag::kMisBSDF : diag::kBSDF, diag::kLightEnvironment, nvert,
                               -1, o + d * 1000.0f, true);
            }
            done = true;
        }
        if (!done) {
            float t = hv.x, u = hv.z, v = hv.w;
            float4 v0 = S.tri[3 * (size_t)prim];
            int32_t mesh = __float_as_int(v0.w);
            GMat gm = S.mats[mesh];
            Mat mat;
            mat.albedo = mk(gm.albedo[0], gm.albedo[1], gm.albedo[2]);
            mat.metallic = gm.metallic; mat.roughness = gm.roughness;
            mat.transmission = gm.transmission; mat.ior = gm.ior;
            v3f em = mk(gm.emission[0], gm.emission[1], gm.emission[2]);
            bool has_em = (gm.flags & MF_EMISSIVE) != 0;
            if (NODES && (gm.flags & MF_NODES)) {
                // hit.uv = uv0 * w + uv1 * u + uv2 * v (geometry/bvh.hpp:67-69),
                // then the material's programs at that uv
                const float2* tu = S.tuv + 3 * (size_t)prim;
                const float2 t0 = tu[0], t1 = tu[1], t2 = tu[2];
                const float w = 1.0f - u - v;
                const float hu = t0.x * w + t1.x * u + t2.x * v;
                const float hv2 = t0.y * w + t1.y * u + t2.y * v;
                if (gm.prog_albedo >= 0) mat.albedo = run_node_program(S.prog + gm.prog_albedo, hu, hv2);
                if (gm.flags & MF_EM_UV) {
                    em = run_node_program(S.prog + gm.prog_emission, hu, hv2);
                    has_em = luminance(em) > 1e-6f;  // is_emissive, render_units.hpp:1200-1202
                }
            }
            v3f point = o + d * t;
            float4 fnv = S.fn[prim];
            v3f nrm = mk(fnv);
            if (__float_as_int(fnv.w)) {
                float w = 1.0f - u - v;
                const float4* vn = S.vn + 3 * (size_t)prim;
                v3f ni = mk(vn[0]) * w + mk(vn[1]) * u + mk(vn[2]) * v;
                nrm = make_dir_or(ni, nrm);
            }
            v3f wo = -d;
            bool front = dot(wo, nrm) > 0.0f;
            v3f sn = (mat.transmission < 0.5f && !front) ? -nrm : nrm;
            bool stop = false;  // traceSimpleWithDiagnostics ends at an emitter
            // guided bounce: opaque, non-delta surfaces (the BSDF sample is an
            // evaluated f / pdf pair there)
            const bool guidable = GUIDED && mat.transmission == 0.0f && !(mat.roughness < 0.05f && mat.metallic > 0.5f);
            bool guided = guidable && G.alpha > 0.0f;
            // the vertex's network outputs, packed (decoded where used)
            guide::GuideOut gout = {};
            guide::LobeWeights glw = {0.0f, 0.0f, 0.0f};
            float ga = G.alpha;  // guide-selection probability of this vertex
            // a degenerate learned density (non-finite outputs) leaves this
            // vertex to plain BSDF sampling and NEE
            if (GUIDED && guided) {
                gout = guide::load_out(W.gout + (size_t)p * guide::kGoutStride);
                guided = guide::out_finite(gout);
                glw = guide::lobe_weights(gout);
                ga = guide::selection_out(gout, G.alpha, G.learn_sel);
            }
            // guided vertex: the NEE sample's MIS weight needs the guide's
            // density at its direction, which the density pass after the
            // BSDF sampling evaluates together with the sampled direction's
            bool nee_wait = false;
            v3f nee_dir = mk(0.0f, 0.0f, 1.0f), nee_c = mk(0.0f, 0.0f, 0.0f);
            float nee_pl = 0.0f, nee_pt = 0.0f;
            if (SIMPLE) {
                if (DIAG) {
                    if (has_em) {
                        const v3f c = cmul(em, thr);
                        L = L + c;
                        put_record(W, p, depth, c, diag::kBSDF, diag::kLightEmissive, nvert, -(mesh + 2), point, true);
                        stop = true;
                    }
                } else {
                    L = L + cmul(em, thr);
                }
            } else if (A == LUCID_ALGO_NEE) {
                if (has_em && depth == 0) {
                    const v3f c = cmul(em, thr);
                    L = L + c;
                    if (DIAG)
                        put_record(W, p, depth, c, diag::kBSDF, diag::kLightEmissive, nvert, -(mesh + 2), point, true);
                }
            } else if (has_em) {  // traceMIS / traceVolumeMIS
                float mw = 1.0f;
                if (S.n_lights > 0 && (last_pdf > MIN_PDF || (GUIDED && rho < 1.0f))) {
                    int32_t li = __float_as_int(S.tri[3 * (size_t)prim + 2].w);
                    if (li >= 0) {
                        float sp = area_ratio(S.lights[li].area, S.total_area);
                        if (sp > 0.0f) {
                            float lp = pdf_light(S.lights[li].area, o, point, nrm) * sp;
                            if (lp > MIN_PDF) mw = hit_mis(lp);
                        }
                    }
                }
                const v3f c = cmul(em, thr) * mw;
                L = L + c;
                if (DIAG)
                    put_record(W, p, depth, c, diag::kMisBSDF, diag::kLightEmissive, nvert, -(mesh + 2), point, true);
            }
            // diagnostic vertex of a non-emissive surface (record_vertex_with_geometry)
            if (DIAG && !has_em && nvert < diag::kMaxDepth) {
                const int32_t packed = (0 & 0xffff) | ((int32_t)classify_bsdf(mat) << 16) |
                                       ((int32_t)(is_delta_bsdf(mat) ? diag::kFlagDelta : 0) << 24);
                const size_t k = (size_t)nvert * (size_t)W.dstride + (size_t)p;
                W.dv0[k] = make_float4(point.x, point.y, point.z, __int_as_float(mesh));
                W.dv1[k] = make_float4(nrm.x, nrm.y, nrm.z, __int_as_float(packed));
                ++nvert;
            }
            // next event estimation
            if (!SIMPLE && S.n_lights > 0 && !has_em && !(mat.transmission > 0.5f)) {
                float sp;
                int li = select_light(S, randf(rng), sp);
                if (sp > 1e-6f) {
                    float a1 = randf(rng), a2 = randf(rng);
                    LSample ls = sample_light(S.lights[li], point, a1, a2);
                    if (ls.pdf > MIN_PDF) {
                        float nl = dot(sn, ls.dir);
                        if (nl > 1e-6f) {
                            v3f f = eval_bsdf(mat, wo, ls.dir, sn);
                            float pl = ls.pdf * sp;
                            v3f cc;
                            if (A == LUCID_ALGO_MIS && GUIDED && guided) {
                                // guided bounce: weighted against the density of its
                                // hemisphere technique (see the BSDF sampling below);
                                // the contribution is written after the density pass
                                nee_wait = true;
                                nee_dir = ls.dir;
                                nee_pl = pl;
                                nee_pt = pdf_bsdf_sampled(mat, wo, ls.dir, sn);
                                nee_c = cmul(bsdf_weight(f, nl, pl), ls.emission);
                                cc = mk(0.0f, 0.0f, 0.0f);
                            } else if (A == LUCID_ALGO_MIS) {
                                float pb = pdf_bsdf(mat, wo, ls.dir, sn);
                                float mw = power_heuristic(pl, pb);
                                cc = cmul(bsdf_weight(f, nl, pl) * mw, ls.emission);
                            } else if (A == LUCID_ALGO_VOLUME_MIS) {
                                // surface NEE through media (path_tracer.hpp:783-797); with
                                // gridded media the transmittance march runs in k_vol_march
                                float pb = pdf_bsdf(mat, wo, ls.dir, sn);
                                float mw = power_heuristic(pl, pb);
                                if (S.n_vols > 0) {
                                    defer_transmittance(W, p, bsdf_weight(f, nl, pl) * mw, ls.emission, thr);
                                    cc = mk(0.0f, 0.0f, 0.0f);
                                } else {
                                    cc = cmul(bsdf_weight(f, nl, pl) * mw, cmul(ls.emission, mk(1.0f, 1.0f, 1.0f)));
                                }
                            } else {
                                cc = cmul(bsdf_weight(f, nl, pl), ls.emission);
                            }
                            cc = cmul(cc, thr);
                            float tag = 0.0f;
                            if (DIAG) {
                                // pending until the shadow ray finds the light unoccluded
                                put_record(W, p, depth, cc, MIS ? diag::kMisNEE : diag::kNEE,
                                           light_source_type(S.lights[li].type), nvert,
                                           -((li - S.native_start) + 2), ls.pos, false);
                                tag = __int_as_float(depth + 1);
                            }
                            W.sh_o[p] = make_float4(point.x, point.y, point.z, ls.dist);
                            W.sh_d[p] = make_float4(ls.dir.x, ls.dir.y, ls.dir.z, FUSED ? __int_as_float(S.root) : tag);
                            if (!(GUIDED && nee_wait)) W.sh_c[p] = make_float4(cc.x, cc.y, cc.z, 0.0f);
                            nee = true;
                        }
                    }
                }
            }
            // BSDF sampling for the next bounce (at the last bounce the reference
            // still samples, but nothing it draws reaches the result)
            if (!stop && depth + 1 < max_depth) {
            v3f smp_n = (mat.transmission > 0.5f) ? nrm : sn;
            const v3f thr_in = thr;
            BSample bs;
            // the density the throughput divides by, and the MIS pdf of the
            // next emitter hit (bs.pdf: the plain tracers' value for both)
            float pw = 0.0f, mis_pdf = 0.0f, next_1m_rho = 0.0f, pt_wi = -1.0f;
            if (GUIDED && guided) {
                // One-sample MIS of the guide and the BSDF sampler, unbiased
                // with respect to the plain tracer.  The plain bounce draws wi
                // with the sampler's true density pt but weighs it by
                // f cos / pb (pb = pdf_bsdf, the reference's pdf, below pt at
                // low roughness: pdf_bsdf_sampled): its BSDF strategy
                // integrates f_eff = f pt / pb, its NEE f.  Here wi is drawn
                // with q = a p_guide + (1 - a) pt and the throughput takes
                // f_eff cos / q.  Direct light then splits as
                //   NEE:  w'_l f                    w'_l = pl^2 / (pl^2 + q^2)
                //   hit:  w'_b f + w_b (f_eff - f)  w'_b = 1 - w'_l, w_b: plain's (pb)
                // which sums to plain's w_l f + w_b f_eff: the same image, with
                // the guide's own density in the weights.  At the hit, in units
                // of the throughput's f_eff: rho w'_b + (1 - rho) w_b, rho = pb / pt.
                if (randf(rng) < ga) {
                    const float g0 = randf(rng), g1 = randf(rng), g2 = randf(rng);
                    bs.wi = guide::mixture_sample_out(gout, glw, g0, g1, g2);
                    bs.kind = 0;
                    bs.type = 0;
                    bs.fw = eval_bsdf(mat, wo, bs.wi, smp_n);
                    bs.pdf = pdf_bsdf(mat, wo, bs.wi, smp_n);
                } else {
                    const float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
                    bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
                }
                const float pt = pdf_bsdf_sampled(mat, wo, bs.wi, smp_n);
                pt_wi = pt;
                float qg = 0.0f, qg_nee = 0.0f;
                guide::mixture_pdf2(gout, glw, bs.wi, nee_dir, qg, qg_nee);
                if (nee_wait) {
                    nee_wait = false;
                    finish_guided_nee(W, p, ga * qg_nee + (1.0f - ga) * nee_pt, nee_pl, nee_c, thr);
                }
                const float q = ga * qg + (1.0f - ga) * pt;
                pw = (pt > 0.0f && q > 0.0f) ? q * (bs.pdf / pt) : 0.0f;
                mis_pdf = q;
                next_1m_rho = pt > 0.0f ? 1.0f - bs.pdf / pt : 0.0f;
            } else {
                const float u1 = randf(rng), u2 = randf(rng), u3 = randf(rng);
                bs = sample_bsdf(mat, wo, smp_n, u1, u2, u3, rng);
                pw = bs.pdf;
                mis_pdf = bs.pdf;
            }
            bool ok = bs.kind == 1 || bs.pdf > MIN_PDF;
            v3f w = bs.fw;
            if (ok && bs.kind == 0) {
                float cs = fabsf(dot(smp_n, bs.wi));
                if (cs <= 1e-6f || bs.pdf < MIN_PDF || pw < MIN_PDF) ok = false;
                else w = bsdf_weight(bs.fw, cs, pw);
            }
            // training slot of this bounce: where, how much arrived, which way it left
            if (GUIDED && trec >= 0 && depth < kTrainSlots) {
                const bool rec = ok && guidable && (w.x > 0.0f || w.y > 0.0f || w.z > 0.0f);
                // the sampler's density of wi (> 0 wherever a sample is kept)
                float pts = 0.0f;
                if (rec) pts = pt_wi >= 0.0f ? pt_wi : pdf_bsdf_sampled(mat, wo, bs.wi, smp_n);
                if (rec) {
                    *tslot_at(W, trec, depth, 0) = make_float4(point.x, point.y, point.z, thr_in.x);
                    *tslot_at(W, trec, depth, 1) = make_float4(nrm.x, nrm.y, nrm.z, thr_in.y);
                    *tslot_at(W, trec, depth, 2) = make_float4(wo.x, wo.y, wo.z, thr_in.z);
                }
                *tslot_at(W, trec, depth, 3) = make_float4(bs.wi.x, bs.wi.y, bs.wi.z, rec ? fmaxf(pts, 1e-30f) : 0.0f);
                *tslot_at(W, trec, depth, 4) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
            if (ok) {
                thr = cmul(thr, w);
                if (MIS) last_pdf = bs.kind == 0 ? mis_pdf : 0.0f;
                if (depth >= 3) {
                    float mx = max3(thr.x, thr.y, thr.z);
                    float rr = smin(mx, 0.95f);
                    if (randf(rng) > rr) ok = false;
                    else thr = thr * (1.0f / rr);
                }
            }
            if (ok && finite3(thr)) {
                cont = true;
                W.ray_o[p] = make_float4(point.x, point.y, point.z, RAY_T_MIN);
                W.ray_d[p] = make_float4(bs.wi.x, bs.wi.y, bs.wi.z,
                                         FUSED ? __int_as_float(S.root) : (GUIDED && bs.kind == 0 ? bs.pdf : 0.0f));
                W.thr[p] = make_float4(thr.x, thr.y, thr.z, last_pdf);
                W.rng[p] = make_ulonglong2(rng.state, rng.inc);
                W.meta[p] = make_int4(meta.x, meta.y, depth + 1,
                                      FUSED ? meta.w : (GUIDED && bs.kind == 0 ? __float_as_int(next_1m_rho) : 0));
            }
            }
            if (GUIDED && nee_wait) {
                float qg_nee = 0.0f, unused = 0.0f;
                guide::mixture_pdf2(gout, glw, nee_dir, nee_dir, qg_nee, unused);
                finish_guided_nee(W, p, ga * qg_nee + (1.0f - ga) * nee_pt, nee_pl, nee_c, thr);
            }
        }
        W.rad[p] = make_float4(L.x, L.y, L.z, 0.0f);
        if (DIAG) W.dcnt[p] = nvert | ((depth + 1) << 8);
    }
    // a cancelled render ends its paths here (no next bounce, no shadow ray)
    cont = cont && !stop;
    nee = nee && !stop;
    // into the segments of this input chunk (the same for the whole
    // workgroup: i = chunk * kBlock + threadIdx.x); k_queue_concat joins them
    const int seg = (i / kBlock) & (kShards - 1);
    int slot = wave_append(W.cnt + kAppendBase + seg * kLineInts, cont);
    if (cont) W.aq[seg * W.seg_cap + slot] = p;
    int sslot = wave_append(W.cnt + kAppendBase + (kShards + seg) * kLineInts, nee);
    if (nee) W.aqs[seg * W.seg_cap + sslot] = p;
}

// Grid-stride over the live queue; the trip count is uniform per workgroup so
// every lane reaches the wavefront ballots of shade_one.
template <int ALGO, bool NODES>
__device__ __forceinline__ void shade_launch(const SceneDev& S, const WaveDev& W, const guide::GuideDev& G, int qin,
                                             int32_t max_depth) {
    const int n = W.cnt[qin];
    // ray fetch counters of the next persistent traversal launches
    if (blockIdx.x == 0) reset_fetch(W);
    // the cancel latch of this bounce (set by its closest-hit launch, not
    // written during this one): uniform over the launch, so the ballots in
    // shade_one stay convergent and every wavefront takes the same decision
    const bool stop = stop_latched(W);
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock)
        shade_one<ALGO, NODES>(S, W, G, qin, max_depth, base + (int)threadIdx.x, n, stop);
}
template <int ALGO, bool NODES>
__global__ __launch_bounds__(kBlock) void k_shade(SceneDev S, WaveDev W, guide::GuideDev G, int qin,
                                                  int32_t max_depth) {
    shade_launch<ALGO, NODES>(S, W, G, qin, max_depth);
}
// The guided shading kernel, held to kGuidedWaves wavefronts per SIMD: its
// register allocation would otherwise leave room for only two of them.
constexpr int kGuidedWaves = 3;
template <bool NODES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kGuidedWaves, kGuidedWaves))) void
k_shade_guided(SceneDev S, WaveDev W, guide::GuideDev G, int qin, int32_t max_depth) {
    shade_launch<LUCID_ALGO_GUIDED_MIS, NODES>(S, W, G, qin, max_depth);
}

// Joins the append segments of a shading launch into the dense next live
// queue q[qout] (blockIdx.y 0) and shadow queue qs (blockIdx.y 1), and
// writes their counts.  Workgroup (s, r) copies slice r of segment s to the
// segment's offset, the sum of the counts before it; gridDim.x is kShards
// times the number of slices.
__global__ __launch_bounds__(256) void k_queue_concat(WaveDev W, int qout) {
    __shared__ int s_off, s_cnt, s_total;
    const int w = blockIdx.y;
    const int slices = (int)gridDim.x / kShards;
    const int s = blockIdx.x / slices, r = blockIdx.x % slices;
    if (threadIdx.x < 64) {
        // lane l: count of segment l (written by the atomics of the shading launch)
        const int c = __hip_atomic_load(W.cnt + kAppendBase + (w * kShards + (int)threadIdx.x) * kLineInts,
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int x = c;  // inclusive prefix sum over the wavefront
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if ((int)threadIdx.x >= o) x += y;
        }
        if ((int)threadIdx.x == s) { s_off = x - c; s_cnt = c; }
        if (threadIdx.x == 63) s_total = x;
    }
    __syncthreads();
    const int32_t* src = (w ? W.aqs : W.aq) + (int64_t)s * W.seg_cap;
    int32_t* dst = (w ? W.qs : W.q[qout]) + s_off;
    const int lo = (int)((int64_t)s_cnt * r / slices), hi = (int)((int64_t)s_cnt * (r + 1) / slices);
    for (int j = lo + (int)threadIdx.x; j < hi; j += 256) dst[j] = src[j];
    if (blockIdx.x == 0 && threadIdx.x == 0) W.cnt[w ? 2 : qout] = s_total;
}
// n: the wave's paths (a bound on both queues); one slice per 32,768
// entries a segment may hold, at most 16 (few workgroups for small waves:
// on a lane they wait for the CUs other lanes' persistent launches hold)
inline void concat_queues(const WaveDev& W, int qout, int64_t n, hipStream_t st) {
    const int64_t per = (int64_t)kShards * 32768;
    const int slices = (int)std::min<int64_t>(16, std::max<int64_t>(1, (n + per - 1) / per));
    k_queue_concat<<<dim3(kShards * slices, 2), 256, 0, st>>>(W, qout);
}
static_assert(kShards == 64, "one shard per lane of a wavefront");

template <int ALGO>
void launch_shade(unsigned gs, hipStream_t st, const SceneDev& S, const WaveDev& W, const guide::GuideDev& G,
                  int qin, int32_t max_depth) {
    if constexpr (ALGO == LUCID_ALGO_GUIDED_MIS) {
        if (S.prog)
            k_shade_guided<true><<<gs, kBlock, 0, st>>>(S, W, G, qin, max_depth);
        else
            k_shade_guided<false><<<gs, kBlock, 0, st>>>(S, W, G, qin, max_depth);
    } else {
        if (S.prog)
            k_shade<ALGO, true><<<gs, kBlock, 0, st>>>(S, W, G, qin, max_depth);
        else
            k_shade<ALGO, false><<<gs, kBlock, 0, st>>>(S, W, G, qin, max_depth);
    }
}

// Shading workgroups per CU of the guided passes: as many as fit at
// kGuidedWaves wavefronts per SIMD (4 SIMDs, 2 wavefronts per workgroup).
constexpr int kGuidedShadeGroups = 2 * kGuidedWaves;

// Surface point and shading normal of a closest hit (as the shade kernel
// computes them).
__device__ __forceinline__ void hit_frame(const SceneDev& S, v3f o, v3f d, float4 hv, v3f& point, v3f& nrm) {
    const int32_t prim = __float_as_int(hv.y);
    point = o + d * hv.x;
    const float4 fnv = S.fn[prim];
    nrm = mk(fnv);
    if (__float_as_int(fnv.w)) {
        const float u = hv.z, v = hv.w, w = 1.0f - u - v;
        const float4* vn = S.vn + 3 * (size_t)prim;
        nrm = make_dir_or(mk(vn[0]) * w + mk(vn[1]) * u + mk(vn[2]) * v, nrm);
    }
}

// Guiding network on the live queue: 64 paths per wavefront, the weights
// staged once per workgroup in LDS (read there by every MFMA tile instead of
// through the cache hierarchy: 16.0 -> 12.7 ms of a 1080p 16-spp guided
// pass, bit-identical outputs).
// Each lane encodes one path's 64 input features into its LDS row (8 16-byte
// stores: rows are padded to 144 bytes, so they stay 16-byte aligned); the
// MFMA operands are then gathered from LDS, two 8-byte runs of 4 features
// each, as 4 tiles of 16 that share every weight load.
// Paths without a surface hit run on zero input and are not written.
// Surface point and shading normal of every traced path (into sh_o / sh_d),
// for the inference test hook.
__global__ __launch_bounds__(256) void k_debug_frame(SceneDev S, WaveDev W, int32_t n) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const float4 hv = W.hit[p];
    v3f point = mk(0.0f, 0.0f, 0.0f), nrm = point;
    if (__float_as_int(hv.y) >= 0) hit_frame(S, mk(W.ray_o[p]), mk(W.ray_d[p]), hv, point, nrm);
    W.sh_o[p] = make_float4(point.x, point.y, point.z, 0.0f);
    W.sh_d[p] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
}

// Orders this wavefront's LDS accesses across its lanes (a lane reads what
// another lane of the same wavefront wrote, or overwrites what another lane
// read): LDS instructions of one wavefront execute in order, so waiting for
// the LDS counter is enough; no workgroup barrier is involved.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt and expcnt untouched
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

constexpr int kInferTiles = 4;  // tiles of 16 samples per MLP pass
// Wavefronts per workgroup: they share one LDS copy of the weights (31 KB);
// with their rows a workgroup takes 68 KB, two per CU.
constexpr int kInferWaves = 4;
constexpr int kInferGroupsPerCU = 2;
// A lane's LDS row (36 dwords, 144 bytes) holds, between trips, the staged
// inputs of its next two trips, and during a trip its path's 64 BF16
// features (dwords 0..31):
//   0..10   this trip's path: index (-1: slot past the queue), hit, ray origin, direction
//   11..23  this trip's triangle normals: face normal + smooth flag, 3 vertex normals
//   24..34  the next trip's path, as 0..10
//   35      the queue entry of the trip after next
constexpr int kRowDw = 36;
struct StagedPath {
    int32_t pp;
    float4 hv;
    float o[3], d[3];
};
__global__ __launch_bounds__(kInferWaves * 64) void k_guide_infer(SceneDev S, WaveDev W, guide::GuideDev G, int qin) {
    __shared__ uint4 wfwd[guide::kPackedFwd / 8];
    __shared__ float wbias[guide::kLayers * guide::kBiasStride];
    __shared__ __attribute__((aligned(16))) uint32_t rows[kInferWaves][64][kRowDw];
    __shared__ int32_t paths[kInferWaves][64];
    const int n = W.cnt[qin];
    if (n <= 0) return;  // uniform: every path of the wave has ended
    guide::stage_weights(G, wfwd, wbias);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, s = lane & 15;
    uint32_t* row = rows[wave][lane];
    int32_t* path = paths[wave];
    const auto bias_of = [&](int l) { return wbias + guide::kBiasStride * l; };
    const int stride = gridDim.x * kInferWaves * 64, lane0 = wave * 64 + lane;
    // Software pipeline through LDS: a trip reads its inputs from the row,
    // issues the loads of later trips (the triangle normals of the next
    // trip, the hits and rays of the trip after it, a queue entry), runs the
    // MLP, and only then writes what arrived to the row.  Every load is
    // issued and waited for inside one trip, from a valid address (a slot
    // past the end reads the queue's last entry, a path without a hit reads
    // path 0's hit record; neither is used), so the compiler places exact
    // waits after the MLP; the wavefronts of a workgroup never wait for one
    // another.
    const auto load_q = [&](int b0) { return W.q[qin][min(b0 + lane0, n - 1)]; };
    const auto load_path = [&](int raw, int b0) {
        StagedPath r;
        r.pp = b0 + lane0 < n ? raw : -1;
        const int i = r.pp >= 0 ? r.pp : 0;
        const float* po = reinterpret_cast<const float*>(W.ray_o + i);
        const float* pd = reinterpret_cast<const float*>(W.ray_d + i);
        r.hv = W.hit[i];
        r.o[0] = po[0]; r.o[1] = po[1]; r.o[2] = po[2];
        r.d[0] = pd[0]; r.d[1] = pd[1]; r.d[2] = pd[2];
        return r;
    };
    struct Frame {
        float4 fn;
        float n0[3], n1[3], n2[3];
    };
    const auto load_frame = [&](const StagedPath& c) {
        const int32_t prim = __float_as_int(c.hv.y);
        const bool hit = c.pp >= 0 && prim >= 0;
        const bool smooth = hit && S.vn != nullptr;
        const float4* fs = hit ? S.fn + prim : W.hit;
        const float* vs = reinterpret_cast<const float*>(smooth ? S.vn + 3 * (size_t)prim : W.hit);
        const int k = smooth ? 4 : 0;  // floats between vertex normals
        Frame f;
        f.fn = fs[0];
        f.n0[0] = vs[0]; f.n0[1] = vs[1]; f.n0[2] = vs[2];
        f.n1[0] = vs[k]; f.n1[1] = vs[k + 1]; f.n1[2] = vs[k + 2];
        f.n2[0] = vs[2 * k]; f.n2[1] = vs[2 * k + 1]; f.n2[2] = vs[2 * k + 2];
        return f;
    };
    const auto put_path = [&](uint32_t* d, const StagedPath& c) {
        d[0] = (uint32_t)c.pp;
        d[1] = __float_as_uint(c.hv.x); d[2] = __float_as_uint(c.hv.y);
        d[3] = __float_as_uint(c.hv.z); d[4] = __float_as_uint(c.hv.w);
        for (int k = 0; k < 3; ++k) { d[5 + k] = __float_as_uint(c.o[k]); d[8 + k] = __float_as_uint(c.d[k]); }
    };
    const auto get_path = [&](const uint32_t* d) {
        StagedPath c;
        c.pp = (int32_t)d[0];
        c.hv = make_float4(__uint_as_float(d[1]), __uint_as_float(d[2]), __uint_as_float(d[3]), __uint_as_float(d[4]));
        for (int k = 0; k < 3; ++k) { c.o[k] = __uint_as_float(d[5 + k]); c.d[k] = __uint_as_float(d[8 + k]); }
        return c;
    };
    const auto put_frame = [&](uint32_t* d, const Frame& f) {
        d[0] = __float_as_uint(f.fn.x); d[1] = __float_as_uint(f.fn.y);
        d[2] = __float_as_uint(f.fn.z); d[3] = __float_as_uint(f.fn.w);
        for (int k = 0; k < 3; ++k) {
            d[4 + k] = __float_as_uint(f.n0[k]); d[7 + k] = __float_as_uint(f.n1[k]); d[10 + k] = __float_as_uint(f.n2[k]);
        }
    };
    // a row as 9 16-byte words
    const auto put_row = [&](const uint32_t (&v)[kRowDw]) {
        uint4* r4 = reinterpret_cast<uint4*>(row);
#pragma unroll
        for (int k = 0; k < kRowDw / 4; ++k) r4[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    };
    int b0 = blockIdx.x * kInferWaves * 64;
    {  // prologue: the first trip's inputs, the second trip's path
        uint32_t v[kRowDw];
        const StagedPath a0 = load_path(load_q(b0), b0);
        put_path(v, a0);
        put_frame(v + 11, load_frame(a0));
        put_path(v + 24, load_path(load_q(b0 + stride), b0 + stride));
        v[35] = (uint32_t)load_q(b0 + 2 * stride);
        put_row(v);
    }
    // a wavefront's trips end when its first slot is past the queue
    for (; b0 + wave * 64 < n; b0 += stride) {
        wave_lds_fence();
        uint32_t v[kRowDw];
        {
            const uint4* r4 = reinterpret_cast<const uint4*>(row);
#pragma unroll
            for (int k = 0; k < kRowDw / 4; ++k) {
                const uint4 t = r4[k];
                v[4 * k] = t.x; v[4 * k + 1] = t.y; v[4 * k + 2] = t.z; v[4 * k + 3] = t.w;
            }
        }
        const StagedPath cur = get_path(v);
        {
            int p = -1;
            float pn[3] = {0.0f, 0.0f, 0.0f};
            v3f nrm = mk(0.0f, 0.0f, 1.0f), wo = mk(0.0f, 0.0f, 1.0f);
            if (cur.pp >= 0 && __float_as_int(cur.hv.y) >= 0) {
                // hit_frame's point and shading normal, from the staged words
                const v3f o = mk(cur.o[0], cur.o[1], cur.o[2]), d = mk(cur.d[0], cur.d[1], cur.d[2]);
                const v3f point = o + d * cur.hv.x;
                nrm = mk(__uint_as_float(v[11]), __uint_as_float(v[12]), __uint_as_float(v[13]));
                if ((int32_t)v[14]) {
                    const float u = cur.hv.z, vv = cur.hv.w, w = 1.0f - u - vv;
                    const v3f n0 = mk(__uint_as_float(v[15]), __uint_as_float(v[16]), __uint_as_float(v[17]));
                    const v3f n1 = mk(__uint_as_float(v[18]), __uint_as_float(v[19]), __uint_as_float(v[20]));
                    const v3f n2 = mk(__uint_as_float(v[21]), __uint_as_float(v[22]), __uint_as_float(v[23]));
                    nrm = make_dir_or(n0 * w + n1 * u + n2 * vv, nrm);
                }
                wo = -d;
                pn[0] = (point.x - G.bmin[0]) * G.binv[0];
                pn[1] = (point.y - G.bmin[1]) * G.binv[1];
                pn[2] = (point.z - G.bmin[2]) * G.binv[2];
                p = cur.pp;
            }
            __bf16 frow[guide::kIn];
            guide::features_bf16(pn, nrm, wo, frow);
            uint4* dst = reinterpret_cast<uint4*>(row);
#pragma unroll
            for (int q = 0; q < guide::kIn / 8; ++q) {
                uint4 t;
                __builtin_memcpy(&t, frow + 8 * q, 16);
                dst[q] = t;
            }
            path[lane] = p;
        }
        // the loads of later trips, ahead of the MLP
        const StagedPath nxt = get_path(v + 24);
        const Frame fnew = load_frame(nxt);
        const StagedPath pnew = load_path((int32_t)v[35], b0 + 2 * stride);
        const int32_t qnew = load_q(b0 + 3 * stride);
        __builtin_amdgcn_sched_barrier(0);
        wave_lds_fence();  // every lane's features are in its row
        // the 4 tiles of 16 samples in passes of kInferTiles tiles (fewer
        // live accumulators per pass, more wavefronts per SIMD)
#pragma unroll
        for (int t0 = 0; t0 < 4; t0 += kInferTiles) {
            guide::bf16x8 x0[kInferTiles], x1[kInferTiles];
            int pt[kInferTiles];
#pragma unroll
            for (int t = 0; t < kInferTiles; ++t) {
                // perm(kb, g, i): features 32 kb + 4 g + i (i < 4) and
                // 32 kb + 16 + 4 g + i - 4, i.e. 8-byte runs 8 kb + g and 8 kb + 4 + g
                const uint2* r = reinterpret_cast<const uint2*>(rows[wave][16 * (t0 + t) + s]);
                const uint4 a = {r[g].x, r[g].y, r[4 + g].x, r[4 + g].y};
                const uint4 bq = {r[8 + g].x, r[8 + g].y, r[12 + g].x, r[12 + g].y};
                __builtin_memcpy(&x0[t], &a, 16);
                __builtin_memcpy(&x1[t], &bq, 16);
                pt[t] = path[16 * (t0 + t) + s];
            }
            guide::f32x4 o[kInferTiles][3];
            guide::mlp_forward_tiles_at<kInferTiles>(wfwd, bias_of, x0, x1, o);
#pragma unroll
            for (int t = 0; t < kInferTiles; ++t) {
                // a slot past the queue writes the spare row
                const size_t dst_path = pt[t] >= 0 ? (size_t)pt[t] : (size_t)W.gspare;
                __half* out = W.gout + dst_path * guide::kGoutStride;
                for (int m = 0; m < 3; ++m) {
                    const int f = 16 * m + 4 * g;  // 4 outputs per lane and tile, 8-byte aligned
                    if (f >= guide::kGoutStride) continue;
                    __half2 lo = __floats2half2_rn(o[t][m][0], o[t][m][1]);
                    __half2 hi = __floats2half2_rn(o[t][m][2], o[t][m][3]);
                    uint2 u2;
                    __builtin_memcpy(&u2.x, &lo, 4);
                    __builtin_memcpy(&u2.y, &hi, 4);
                    *reinterpret_cast<uint2*>(out + f) = u2;
                }
            }
        }
        // the next trip's inputs (every lane has read its operands)
        wave_lds_fence();
        for (int k = 0; k < 11; ++k) v[k] = v[24 + k];
        put_frame(v + 11, fnew);
        put_path(v + 24, pnew);
        v[35] = (uint32_t)qnew;
        put_row(v);
    }
}

// Training samples of the finished batch: for each recorded bounce of the
// recording paths, weight = lum((L_end - L_after) / T_in), the radiance that
// came back through wi in units of the throughput that reached the vertex.
__global__ __launch_bounds__(256) void k_guide_collect(WaveDev W, int32_t n_paths, guide::Sample* ring,
                                                       unsigned long long* total, int capacity) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int p = train_record_path(W, r);
    const bool live = p < n_paths;
    const float4 le = live ? W.rad[p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // the samples of this lane's slots: slot k is kept when bit k is set
    float wgt[kTrainSlots];
    unsigned keep = 0;
#pragma unroll
    for (int k = 0; k < kTrainSlots; ++k) {
        wgt[k] = 0.0f;
        if (!live) continue;
        const float4 a = *tslot_at(W, r, k, 0), b = *tslot_at(W, r, k, 1), c = *tslot_at(W, r, k, 2);
        const float4 d = *tslot_at(W, r, k, 3), e = *tslot_at(W, r, k, 4);
        if (d.w > 0.0f && e.w == 1.0f) {
            const float tr = a.w, tg = b.w, tb = c.w;
            const float cr = tr > 0.0f ? (le.x - e.x) / tr : 0.0f;
            const float cg = tg > 0.0f ? (le.y - e.y) / tg : 0.0f;
            const float cb = tb > 0.0f ? (le.z - e.z) / tb : 0.0f;
            wgt[k] = 0.2126f * cr + 0.7152f * cg + 0.0722f * cb;
        }
        if (wgt[k] > 0.0f && isfinite(wgt[k])) keep |= 1u << k;
    }
    // ring positions for all kept samples of the wavefront: one atomic
    // (lane-major: a lane's kept slots are consecutive)
    const int mine = __popc(keep);
    int x = mine;  // inclusive prefix sum over the wavefront
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane_id() >= o) x += y;
    }
    const int wave_total = __shfl(x, 63);
    unsigned long long base = 0;
    if (lane_id() == 0 && wave_total > 0) base = atomicAdd(total, (unsigned long long)wave_total);
    base = __shfl(base, 0);
    unsigned long long pos = base + (unsigned long long)(x - mine);
#pragma unroll
    for (int k = 0; k < kTrainSlots; ++k) {
        if (!(keep & (1u << k))) continue;
        const float4 a = *tslot_at(W, r, k, 0), b = *tslot_at(W, r, k, 1), c = *tslot_at(W, r, k, 2);
        const float4 d = *tslot_at(W, r, k, 3);
        guide::Sample sm;
        sm.p[0] = a.x; sm.p[1] = a.y; sm.p[2] = a.z;
        sm.n[0] = b.x; sm.n[1] = b.y; sm.n[2] = b.z;
        sm.wo[0] = c.x; sm.wo[1] = c.y; sm.wo[2] = c.z;
        sm.wi[0] = d.x; sm.wi[1] = d.y; sm.wi[2] = d.z;
        sm.weight = wgt[k];
        sm.pbsdf = d.w;
        sm.pad[0] = sm.pad[1] = 0.0f;
        ring[(int)(pos % (unsigned long long)capacity)] = sm;
        ++pos;
    }
}

// ---------------------------------------------------------------- shadow --
__global__ __launch_bounds__(kBlock) void k_shadow(SceneDev S, WaveDev W, int sd) {
    extern __shared__ int32_t lds_stack[];
    const int n = W.cnt[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&W.rays[1], (unsigned long long)n);
    uint32_t nn = 0, nt = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int p = W.qs[i];
        const float4 o = W.sh_o[p], d = W.sh_d[p];
        if (!shadowed(S, mk(o), mk(d), o.w, lds_stack + threadIdx.x, sd, &W.cnt[3], nn, nt)) {
            const float4 L = W.rad[p], c = W.sh_c[p];
            W.rad[p] = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, 0.0f);
            // diagnostic runs: the pending NEE record of this bounce is kept
            const int32_t tag = __float_as_int(d.w);
            if (tag > 0) {
                float4* r = W.dr0 + (size_t)(tag - 1) * (size_t)W.dstride + (size_t)p;
                r->w = __int_as_float(__float_as_int(r->w) | kRecValid);
            }
        }
    }
    if (W.count_work && (nn | nt)) {
        atomicAdd(&W.rays[4], (unsigned long long)nn);
        atomicAdd(&W.rays[5], (unsigned long long)nt);
    }
}

// ------------------------------------------------------------------ film --
__global__ __launch_bounds__(256) void k_film(WaveDev W, float4* film, int32_t P, int32_t B) {
    int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= P) return;
    // a latched wave may hold paths that shading ended early: the film keeps
    // the waves that finished before the latch only
    if (stop_latched(W)) return;
    float4 a = film[pix];
    for (int b = 0; b < B; ++b) {
        float4 r = W.rad[(size_t)b * P + pix];
        a.x = a.x + r.x; a.y = a.y + r.y; a.z = a.z + r.z;
    }
    film[pix] = a;
}
// ---- fused session passes: the sessions of a batched pass, their scenes
// merged into one set of arrays (their node and triangle ranges laid end to
// end; node links and leaf triangle ranges shifted to the merged ranges),
// share one launch per kernel and bounce.  A path carries its session in
// meta.w and its scene's BVH root in the w of its ray directions; the
// per-scene values (light list, CDF, materials, node programs, environment)
// are read through the session's view of the merged arrays, so each path
// computes exactly what a render of its session alone computes.
struct FusedSession {
    CamDev cam;            // the session's camera; cam.s0 = its first sample index of the pass
    float4* accum;         // its radiance sums
    int64_t pix_off;       // its first pixel in the pass's pixel range
    int32_t root, max_depth;
    int32_t mat_off, prog_off, light_off, native_off;
    int32_t n_lights, n_native, native_start, pad;
    float total_area, env[3];
};
struct FusedDev {
    const FusedSession* ses;
    int32_t n_ses;
    int64_t P;  // pixels of all sessions
};
// the session of pixel `pix` of the pass (sessions in pixel order)
__device__ __forceinline__ int fused_session_of(const FusedDev& F, int64_t pix) {
    int lo = 0, hi = F.n_ses - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (F.ses[mid].pix_off <= pix) lo = mid; else hi = mid - 1;
    }
    return lo;
}
// a session's view of the merged scene arrays: the per-triangle and node
// arrays are shared (global triangle and node indices), the per-scene lists
// start at the session's offsets, so the indices stored in the scene's own
// records (mesh, light, node program, native light) stay valid
__device__ __forceinline__ SceneDev fused_view(const SceneDev& A, const FusedSession& f) {
    SceneDev v = A;
    v.root = f.root;
    v.mats = A.mats + f.mat_off;
    v.prog = A.prog ? A.prog + f.prog_off : nullptr;
    v.lights = A.lights + f.light_off;
    v.cdf = A.cdf + f.light_off;
    v.native_ids = A.native_ids + f.native_off;
    v.n_lights = f.n_lights;
    v.n_native = f.n_native;
    v.native_start = f.native_start;
    v.total_area = f.total_area;
    v.env[0] = f.env[0]; v.env[1] = f.env[1]; v.env[2] = f.env[2];
    v.n_vols = 0;
    v.any_vbo = 0;
    return v;
}
// wave slot p = b * P + pixel: sample b of every pixel of every session
__global__ __launch_bounds__(kBlock) void k_raygen_fused(WaveDev W, FusedDev F, int32_t s_base, int64_t n_paths) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p == 0) { W.cnt[0] = (int32_t)n_paths; W.cnt[1] = 0; W.cnt[2] = 0; }
    if (blockIdx.x == 0) reset_fetch(W);
    if (p >= n_paths) return;
    const int64_t b = p / F.P, pix = p - b * F.P;
    const int sid = fused_session_of(F, pix);
    const FusedSession& f = F.ses[sid];
    raygen_path(W, f.cam, (int)p, (int)(pix - f.pix_off), f.cam.s0 + s_base + (int)b, sid, __int_as_float(f.root));
}
template <int ALGO, bool NODES>
__global__ __launch_bounds__(kBlock) void k_shade_fused(SceneDev A, WaveDev W, FusedDev F, int qin) {
    const int n = W.cnt[qin];
    if (blockIdx.x == 0) reset_fetch(W);
    guide::GuideDev G;
    __builtin_memset(&G, 0, sizeof(G));
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const int i = base + (int)threadIdx.x;
        int sid = 0;
        if (i < n) sid = W.meta[W.q[qin][i]].w;
        const FusedSession& f = F.ses[sid];
        shade_one<ALGO, NODES, true>(fused_view(A, f), W, G, qin, f.max_depth, i, n, false);
    }
}
template <int ALGO>
void launch_shade_fused(unsigned gs, hipStream_t st, const SceneDev& A, const WaveDev& W, const FusedDev& F, int qin) {
    if (A.prog)
        k_shade_fused<ALGO, true><<<gs, kBlock, 0, st>>>(A, W, F, qin);
    else
        k_shade_fused<ALGO, false><<<gs, kBlock, 0, st>>>(A, W, F, qin);
}
// k_film for the sessions of a fused pass: each pixel's B samples in order
// onto its session's sums
__global__ __launch_bounds__(256) void k_film_fused(WaveDev W, FusedDev F, int32_t B) {
    const int64_t pix = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (pix >= F.P) return;
    const FusedSession& f = F.ses[fused_session_of(F, pix)];
    float4* film = f.accum + (pix - f.pix_off);
    float4 a = *film;
    for (int b = 0; b < B; ++b) {
        const float4 r = W.rad[(size_t)b * F.P + pix];
        a.x = a.x + r.x; a.y = a.y + r.y; a.z = a.z + r.z;
    }
    *film = a;
}
// A scene's nodes copied into the merged array: child node indices and
// leaf triangle ranges shifted by the scene's offsets.
__global__ __launch_bounds__(256) void k_rebase_nodes(GNode* nodes, int64_t n, int32_t node_off, int32_t tri_off) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 2; ++c) {
        const int32_t ch = nodes[i].child[c];
        if (ch == kEmptyChild) continue;
        if (ch >= 0) {
            nodes[i].child[c] = ch + node_off;
        } else {
            const int32_t start = (~ch) >> 3, count = (~ch) & 7;
            nodes[i].child[c] = ~(((start + tri_off) << 3) | count);  // encode_leaf
        }
    }
}

// ---- ray binning: a ray queue regrouped before its traversal pass, so
// that a wavefront's rays start in the same region (and head the same way).
// Counting sort on a 12-bit key; the order within a bin is arbitrary.
// Bounce rays: direction octant, then the Morton code of an 8x8x8 grid over
// the scene box.  Shadow rays, which mostly head for the same few lights:
// the Morton code of a 16x16x16 grid.  Results do not depend on queue
// order: every path keeps its own state and RNG.
constexpr int kRayBinBits = 12;
constexpr int kRayBins = 1 << kRayBinBits;
constexpr int kBinBlock = 1024;  // threads per workgroup of the binning kernels
struct BinBox {
    float lo[3], sc[3];  // (origin - lo) * sc is in [0, 1) inside the scene box
};
enum BinKey { kBinBounce = 0, kBinShadow = 1 };
template <int KEY>
__device__ __forceinline__ uint32_t ray_bin(float4 o, float4 d, const BinBox& b) {
    constexpr uint32_t cells = KEY == kBinShadow ? 16u : 8u;
    constexpr int bits = KEY == kBinShadow ? 4 : 3;
    const float oc[3] = {o.x, o.y, o.z};
    uint32_t c[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float f = (oc[k] - b.lo[k]) * b.sc[k] * (float)cells;
        c[k] = f >= (float)(cells - 1u) ? cells - 1u : (f > 0.0f ? (uint32_t)f : 0u);  // NaN -> 0
    }
    uint32_t m = 0u;  // Morton code of the origin cell
#pragma unroll
    for (int bit = 0; bit < bits; ++bit)
        m |= (((c[0] >> bit) & 1u) << (3 * bit)) | (((c[1] >> bit) & 1u) << (3 * bit + 1)) |
             (((c[2] >> bit) & 1u) << (3 * bit + 2));
    if (KEY == kBinShadow) return m;
    const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    return (oct << 9) | m;
}
__global__ __launch_bounds__(256) void k_bin_clear(uint32_t* bins) {
    for (int i = threadIdx.x; i < kRayBins; i += 256) bins[i] = 0u;
}
// per-bin counts of the live queue; each entry's bin is kept in skey
// The queue to regroup: entries q[0, *count), rays org[p] + t dir[p].
struct BinSrc {
    const int32_t* q;
    const int32_t* count;
    const float4* org;
    const float4* dir;
};
template <int KEY>
__global__ __launch_bounds__(kBinBlock) void k_bin_count(WaveDev W, BinSrc src, BinBox b) {
    __shared__ uint32_t h[kRayBins];
    for (int i = threadIdx.x; i < kRayBins; i += kBinBlock) h[i] = 0u;
    __syncthreads();
    const int n = *src.count;
    for (int i = blockIdx.x * kBinBlock + (int)threadIdx.x; i < n; i += gridDim.x * kBinBlock) {
        const int p = src.q[i];
        const uint32_t k = ray_bin<KEY>(src.org[p], src.dir[p], b);
        W.skey[i] = k;
        atomicAdd(&h[k], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kRayBins; i += kBinBlock)
        if (h[i]) atomicAdd(&W.sbins[i], h[i]);
}
// exclusive prefix sum of the kRayBins counts (one workgroup, 4 bins a lane)
__global__ __launch_bounds__(1024) void k_bin_scan(uint32_t* bins) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    uint32_t v[4], sum = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = bins[4 * t + k];
        sum += v[k];
    }
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t x = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        bins[4 * t + k] = run;
        run += v[k];
    }
}
// Each workgroup moves one contiguous chunk of the queue: it reserves a
// range in every bin it holds entries of, then places them.
__global__ __launch_bounds__(kBinBlock) void k_bin_scatter(WaveDev W, BinSrc src) {
    __shared__ uint32_t h[kRayBins];
    __shared__ uint32_t base[kRayBins];
    for (int i = threadIdx.x; i < kRayBins; i += kBinBlock) h[i] = 0u;
    __syncthreads();
    const int n = *src.count;
    const int chunk = (n + (int)gridDim.x - 1) / (int)gridDim.x;
    const int i0 = (int)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int i = i0 + (int)threadIdx.x; i < i1; i += kBinBlock) atomicAdd(&h[W.skey[i]], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kRayBins; i += kBinBlock) {
        base[i] = h[i] ? atomicAdd(&W.sbins[i], h[i]) : 0u;
        h[i] = 0u;
    }
    __syncthreads();
    for (int i = i0 + (int)threadIdx.x; i < i1; i += kBinBlock) {
        const uint32_t k = W.skey[i];
        const uint32_t r = atomicAdd(&h[k], 1u);
        W.sq[base[k] + r] = src.q[i];
    }
}

// Regroups src into W.sq (stream-ordered; W.skey and W.sbins are scratch).
template <int KEY>
void bin_rays(const WaveDev& W, const BinSrc& src, const BinBox& box, int64_t n, int num_cu, hipStream_t st) {
    // one workgroup per CU at most: each workgroup adds its LDS histogram
    // into the 4096 global counters, so fewer, larger workgroups mean fewer
    // contended atomics on them
    const unsigned gb =
        (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kBinBlock - 1) / kBinBlock), (int64_t)num_cu);
    k_bin_clear<<<1, 256, 0, st>>>(W.sbins);
    k_bin_count<KEY><<<gb, kBinBlock, 0, st>>>(W, src, box);
    k_bin_scan<<<1, 1024, 0, st>>>(W.sbins);
    k_bin_scatter<<<gb, kBinBlock, 0, st>>>(W, src);
}

// ---- shading order: a bounce's live queue regrouped by the shading class
// of each path's hit (miss, emitter, dielectric, metal, partly transmissive,
// glass, node-tree material), so that the wavefronts of the shading kernel
// run paths that take the same branches.  A stable partition by wavefront
// ballots and prefix sums: each workgroup owns a contiguous chunk of the
// queue; per round of 256 entries every wavefront counts its lanes of each
// class with one ballot per class, the counts are prefix-summed over the
// workgroup's wavefronts, and one atomic per class reserves the round's
// range in that class's segment.  The order within a class follows the
// queue, so the shading kernel's per-path reads stay as coalesced as they
// were.  Results do not depend on the order.
constexpr int kClassBlock = 256;
__device__ __forceinline__ int shade_class_of(const SceneDev& S, float4 hv) {
    const int32_t prim = __float_as_int(hv.y);
    return prim < 0 ? (int)kClassMiss : S.mats[prim_mesh(S, prim)].shade_class;
}
__device__ __forceinline__ int class_chunk(int n) {
    return ((n + (int)gridDim.x - 1) / (int)gridDim.x + kClassBlock - 1) / kClassBlock * kClassBlock;
}
// per-class totals of the queue (added into totals[0, kShadeClasses)); each
// entry's class is kept in skey
__global__ __launch_bounds__(kClassBlock) void k_class_count(SceneDev S, WaveDev W, const int32_t* q, const int32_t* count,
                                                             uint32_t* totals) {
    __shared__ uint32_t h[kShadeClasses];
    if (threadIdx.x < kShadeClasses) h[threadIdx.x] = 0u;
    __syncthreads();
    const int n = *count, chunk = class_chunk(n);
    const int i0 = (int)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    for (int base = i0; base < i1; base += kClassBlock) {  // uniform over the workgroup
        const int i = base + (int)threadIdx.x;
        int c = -1;
        if (i < i1) {
            c = shade_class_of(S, W.hit[q[i]]);
            W.skey[i] = (uint32_t)c;
        }
#pragma unroll
        for (int k = 0; k < kShadeClasses; ++k) {
            const unsigned long long m = __ballot(c == k);
            if (m && lane_id() == 0) atomicAdd(&h[k], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x < kShadeClasses && h[threadIdx.x]) atomicAdd(&totals[threadIdx.x], h[threadIdx.x]);
}
// cursors[k] = start of class k's segment (exclusive prefix sum of totals)
__global__ __launch_bounds__(64) void k_class_scan(const uint32_t* totals, uint32_t* cursors) {
    if (threadIdx.x == 0) {
        uint32_t run = 0u;
        for (int k = 0; k < kShadeClasses; ++k) {
            cursors[k] = run;
            run += totals[k];
        }
    }
}
__global__ __launch_bounds__(kClassBlock) void k_class_scatter(WaveDev W, const int32_t* q, const int32_t* count,
                                                               uint32_t* cursors, int32_t* out) {
    constexpr int kWaves = kClassBlock / 64;
    __shared__ uint32_t wc[kWaves][kShadeClasses];
    __shared__ uint32_t rb[kShadeClasses];
    const int n = *count, chunk = class_chunk(n);
    const int i0 = (int)blockIdx.x * chunk, i1 = min(n, i0 + chunk);
    const int wave = (int)threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane_id()) - 1ull;
    for (int base = i0; base < i1; base += kClassBlock) {
        const int i = base + (int)threadIdx.x;
        const int c = i < i1 ? (int)W.skey[i] : -1;
        unsigned long long mine = 0ull;
#pragma unroll
        for (int k = 0; k < kShadeClasses; ++k) {
            const unsigned long long m = __ballot(c == k);
            if (lane_id() == 0) wc[wave][k] = (uint32_t)__popcll(m);
            if (c == k) mine = m;
        }
        __syncthreads();
        if (threadIdx.x < kShadeClasses) {
            uint32_t t = 0u;
            for (int w = 0; w < kWaves; ++w) t += wc[w][threadIdx.x];
            rb[threadIdx.x] = t ? atomicAdd(&cursors[threadIdx.x], t) : 0u;
        }
        __syncthreads();
        if (c >= 0) {
            uint32_t off = rb[c];
            for (int w = 0; w < wave; ++w) off += wc[w][c];
            out[off + (uint32_t)__popcll(mine & below)] = q[i];
        }
        __syncthreads();  // wc / rb are rewritten by the next round
    }
}
// Regroups queue q (count entries) by shading class into `out`; sbins and
// skey are scratch (free between a traversal pass and the next binning).
void class_order(const SceneDev& S, const WaveDev& W, const int32_t* q, const int32_t* count, int32_t* out, int64_t n,
                 int num_cu, hipStream_t st) {
    const unsigned g = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (n + kClassBlock - 1) / kClassBlock),
                                                   (int64_t)num_cu * 4);
    k_bin_clear<<<1, 256, 0, st>>>(W.sbins);
    k_class_count<<<g, kClassBlock, 0, st>>>(S, W, q, count, W.sbins);
    k_class_scan<<<1, 64, 0, st>>>(W.sbins, W.sbins + kShadeClasses);
    k_class_scatter<<<g, kClassBlock, 0, st>>>(W, q, count, W.sbins + kShadeClasses, out);
}

// One int to device memory, in stream order (vector store from lane 0).
__global__ __launch_bounds__(64) void k_set_i32(int32_t* p, int32_t v) {
    if (threadIdx.x == 0) p[threadIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_film_clear(float4* film, int32_t P) {
    int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix < P) film[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
// per-sample radiance out: dst[(pix*S + s)*3 + c]
__global__ __launch_bounds__(256) void k_samples_out(WaveDev W, float* dst, int32_t P, int32_t B, int32_t s_base,
                                                     int32_t S) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P * B) return;
    int b = p / P, pix = p - b * P;
    float4 r = W.rad[p];
    size_t k = ((size_t)pix * S + s_base + b) * 3;
    dst[k] = r.x; dst[k + 1] = r.y; dst[k + 2] = r.z;
}
// apply_camera_sensitivity + clamp + alpha + Y flip (pybind_renderer.hpp:331-360)
__global__ __launch_bounds__(256) void k_output(const float4* film, float* out, int32_t tw, int32_t th, float sens,
                                                int32_t flip, const int32_t* latch) {
    int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= tw * th) return;
    // a cancelled render returns its partial sums in row order, unflipped
    // (pybind_renderer.hpp:225-233 returns before the flip of :351-360)
    if (latch && latch_read(latch)) flip = 0;
    int px = pix % tw, py = pix / tw;
    float4 a = film[pix];
    float r = smax(a.x * sens, 0.0f), g = smax(a.y * sens, 0.0f), b = smax(a.z * sens, 0.0f);
    int row = flip ? (th - 1 - py) : py;
    size_t k = ((size_t)row * tw + px) * 4;
    *(float4*)(out + k) = make_float4(r, g, b, 1.0f);
}
// Session read-out: (sum * inv_samples) * 2^EV, clamp at 0, alpha 1, rows
// bottom first (the async accumulator snapshot of src/gpu/path_tracer.cpp:
// 1064-1075 followed by render_poll_async's sensitivity and clamp,
// pybind_renderer.hpp:766-783).  inv_samples = 1 reads the plain sum.
__global__ __launch_bounds__(256) void k_session_out(const float4* acc, float* out, int32_t w, int32_t h,
                                                     float inv_samples, float sens) {
    int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= w * h) return;
    int px = pix % w, py = pix / w;
    float4 a = acc[pix];
    float r = smax((a.x * inv_samples) * sens, 0.0f);
    float g = smax((a.y * inv_samples) * sens, 0.0f);
    float b = smax((a.z * inv_samples) * sens, 0.0f);
    size_t k = ((size_t)(h - 1 - py) * w + px) * 4;
    *(float4*)(out + k) = make_float4(r, g, b, 1.0f);
}
// primary-hit export: mesh, original triangle, u, v, t
__global__ __launch_bounds__(256) void k_hits_out(SceneDev S, WaveDev W, int32_t* mesh, int32_t* tri, float* u,
                                                  float* v, float* t, int32_t P, int32_t B, int32_t S_total,
                                                  int32_t s_base) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P * B) return;
    int b = p / P, pix = p - b * P;
    size_t k = (size_t)pix * S_total + s_base + b;
    float4 h = W.hit[p];
    int32_t prim = __float_as_int(h.y);
    if (prim >= 0) {
        mesh[k] = __float_as_int(S.tri[3 * (size_t)prim].w);
        tri[k] = __float_as_int(S.tri[3 * (size_t)prim + 1].w);
        u[k] = h.z; v[k] = h.w; t[k] = h.x;
    } else {
        mesh[k] = -1; tri[k] = -1; u[k] = 0.0f; v[k] = 0.0f; t[k] = 0.0f;
    }
}


// ---------------------------------------------------------- debug passes --
// render_debug (pybind_renderer.hpp:371-439): the closest hit of one
// unjittered camera ray per pixel, found by intersectScene from t = 0 (no
// volume-boundary skipping), shown as the shading normal mapped to
// n * 0.5 + 0.5 (traceNormal, renderer.hpp:58-65), the material's albedo
// (traceAlbedo :67-78), its emission (traceEmission :80-93) or a medium's
// thickness (traceVolume :109-190); any other mode is magenta.
constexpr int32_t kDebugNormal = 0, kDebugAlbedo = 1, kDebugEmission = 2, kDebugVolume = 3;
__device__ __forceinline__ bool medium_present(const GVolume& v) { return v.density > 0.0f || v.has_grid; }
// traceVolume's second query: a camera ray whose closest hit is a mesh with
// a medium is traced again from 1e-4 past the entry point (the first hit is
// kept in sh_o).  Queue 1 gets those paths.
__global__ __launch_bounds__(256) void k_debug_vol_prep(SceneDev S, WaveDev W, int32_t n) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    bool again = false;
    if (p < n) {
        const float4 hv = W.hit[p];
        W.sh_o[p] = hv;
        const int32_t prim = __float_as_int(hv.y);
        if (prim >= 0) {
            const int32_t vi = S.mats[prim_mesh(S, prim)].volume;
            if (vi >= 0 && medium_present(S.vols[vi])) {
                const v3f o = mk(W.ray_o[p]), d = mk(W.ray_d[p]);
                const v3f entry = o + d * hv.x;
                const v3f in = entry + d * 1e-4f;
                W.ray_o[p] = make_float4(in.x, in.y, in.z, 0.0f);
                again = true;
            }
        }
    }
    const int slot = wave_append(&W.cnt[1], again);
    if (again) W.q[1][slot] = p;
}
__global__ __launch_bounds__(256) void k_debug_out(SceneDev S, WaveDev W, float* out, int32_t tw, int32_t th,
                                                   int32_t mode) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= tw * th) return;
    const int px = p % tw, py = p / tw;
    const float4 hv = mode == kDebugVolume ? W.sh_o[p] : W.hit[p];
    const int32_t prim = __float_as_int(hv.y);
    v3f c = mk(0.0f, 0.0f, 0.0f);
    if (mode == kDebugNormal) {
        v3f nrm = c;
        if (prim >= 0) {
            v3f point;
            hit_frame(S, mk(W.ray_o[p]), mk(W.ray_d[p]), hv, point, nrm);
        }
        c = mk(nrm.x * 0.5f + 0.5f, nrm.y * 0.5f + 0.5f, nrm.z * 0.5f + 0.5f);
    } else if ((mode == kDebugAlbedo || mode == kDebugEmission) && prim >= 0) {
        const GMat& gm = S.mats[prim_mesh(S, prim)];
        const int32_t prog = mode == kDebugAlbedo ? gm.prog_albedo : gm.prog_emission;
        c = mode == kDebugAlbedo ? mk(gm.albedo[0], gm.albedo[1], gm.albedo[2])
                                 : mk(gm.emission[0], gm.emission[1], gm.emission[2]);
        if (S.prog && (gm.flags & MF_NODES) && prog >= 0) {
            // the hit's texture coordinate (geometry/bvh.hpp:67-69)
            const float2* tu = S.tuv + 3 * (size_t)prim;
            const float w = 1.0f - hv.z - hv.w;
            const float hu = tu[0].x * w + tu[1].x * hv.z + tu[2].x * hv.w;
            const float hv2 = tu[0].y * w + tu[1].y * hv.z + tu[2].y * hv.w;
            c = run_node_program(S.prog + prog, hu, hv2);
        }
    } else if (mode == kDebugVolume && prim >= 0) {
        const int32_t vi = S.mats[prim_mesh(S, prim)].volume;
        if (vi >= 0 && medium_present(S.vols[vi])) {
            const GVolume& v = S.vols[vi];
            const float4 ex = W.hit[p];  // the query from inside
            const bool exit_hit = __float_as_int(ex.y) >= 0;
            if (v.has_grid) {
                // ray march at half a voxel, max sampled density vs opacity
                const float t_inside = exit_hit ? ex.x : 0.0f;
                if (t_inside > 0.0f) {
                    const float extent = max3(v.wmax[0] - v.wmin[0], v.wmax[1] - v.wmin[1], v.wmax[2] - v.wmin[2]);
                    const int max_dim = max(max(v.dims[0], v.dims[1]), v.dims[2]);
                    const float voxel = extent / (float)max(1, max_dim);
                    const float step = smax(1e-4f, voxel * 0.5f);
                    const int n_steps = min(2048, (int)(t_inside / step) + 1);
                    const v3f ro = mk(W.ray_o[p]), rd = mk(W.ray_d[p]);
                    float integrated = 0.0f, max_sampled = 0.0f;
                    for (int i = 0; i < n_steps; ++i) {
                        const float t = ((float)i + 0.5f) * step;
                        if (t > t_inside) break;
                        const float d = grid_value(S, v, ro.x + rd.x * t, ro.y + rd.y * t, ro.z + rd.z * t);
                        integrated += d * step;
                        if (d > max_sampled) max_sampled = d;
                    }
                    const float opacity = 1.0f - expf(-v.density * integrated * 10.0f);
                    const float vis = smax(max_sampled, opacity);
                    c = mk(v.color[0] * vis, v.color[1] * vis, v.color[2] * vis);
                }
            } else {
                const float t = exit_hit ? 1.0f - expf(-v.density * ex.x) : 1.0f - expf(-v.density);
                c = mk(v.color[0] * t, v.color[1] * t, v.color[2] * t);
            }
        }
    } else if (mode < kDebugNormal || mode > kDebugVolume) {
        c = mk(1.0f, 0.0f, 1.0f);
    }
    const size_t k = ((size_t)(th - 1 - py) * tw + px) * 4;
    *(float4*)(out + k) = make_float4(c.x, c.y, c.z, 1.0f);
}

// ------------------------------------------------------------ diag film --
// A recorded completed path, read from the wave's diagnostic storage: the
// first `prefix` path vertices, then the light vertex when the path had room
// (record_completed_path_with_geometry, diagnostic_integrator.hpp:338-368).
struct DevTrace {
    const float4 *dv0, *dv1;
    int64_t stride;
    int p, prefix, n;
    float4 r0, r1;
    __device__ int depth() const { return n; }
    __device__ diag::Vertex vertex(int k) const {
        diag::Vertex v;
        if (k < prefix) {
            const size_t i = (size_t)k * (size_t)stride + (size_t)p;
            const int32_t packed = __float_as_int(dv1[i].w);
            v.object_id = __float_as_int(dv0[i].w);
            v.material_id = (int16_t)(packed & 0xffff);
            v.bsdf_type = (uint8_t)((packed >> 16) & 0xff);
            v.flags = (uint8_t)((packed >> 24) & 0xff);
        } else {
            v.object_id = __float_as_int(r1.w);
            v.material_id = -1;
            v.bsdf_type = diag::kEmission;
            v.flags = 0;
        }
        return v;
    }
    __device__ diag::F3 position(int k) const {
        if (k < prefix) {
            const float4 a = dv0[(size_t)k * (size_t)stride + (size_t)p];
            return diag::F3{a.x, a.y, a.z};
        }
        return diag::F3{r1.x, r1.y, r1.z};
    }
    __device__ diag::F3 normal(int k) const {
        if (k < prefix) {
            const float4 b = dv1[(size_t)k * (size_t)stride + (size_t)p];
            return diag::F3{b.x, b.y, b.z};
        }
        return diag::F3{0.0f, 0.0f, 0.0f};
    }
    __device__ diag::F3 contribution() const { return diag::F3{r0.x, r0.y, r0.z}; }
    __device__ uint8_t light_type() const { return (uint8_t)((__float_as_int(r0.w) >> 16) & 0xff); }
    __device__ uint8_t strategy() const { return (uint8_t)((__float_as_int(r0.w) >> 8) & 0xff); }
};

// DiagnosticFilm::record_path for one wavefront batch: one thread per film
// pixel applies, sample by sample and within a sample for the tile pixels
// (fx*sub+dx, fy*sub+dy) in row-major order, each path's valid records in
// bounce order; and adds each path's per-strategy sums to the FP64 Welford.
__global__ __launch_bounds__(64) void k_diag_film(WaveDev W, diag::PixelData* film, diag::StrategyWelford* strat,
                                                  unsigned long long* total, int32_t tw, int32_t sub, int32_t sw,
                                                  int32_t sh, int32_t P, int32_t B) {
    const int fi = blockIdx.x * 64 + threadIdx.x;
    if (fi >= sw * sh) return;
    const int fx = fi % sw, fy = fi / sw;
    diag::PixelData& px = film[fi];
    diag::StrategyWelford* w = strat + (size_t)fi * 4;
    unsigned long long count = 0;
    for (int b = 0; b < B; ++b) {
        for (int dy = 0; dy < sub; ++dy) {
            for (int dx = 0; dx < sub; ++dx) {
                const int pix = (fy * sub + dy) * tw + fx * sub + dx;
                const int p = b * P + pix;
                const int slots = W.dcnt[p] >> 8;
                double sum[4][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
                for (int sl = 0; sl < slots; ++sl) {
                    const size_t k = (size_t)sl * (size_t)W.dstride + (size_t)p;
                    const float4 r0 = W.dr0[k];
                    const int32_t meta = __float_as_int(r0.w);
                    if (!(meta & kRecValid)) continue;
                    DevTrace t;
                    t.dv0 = W.dv0; t.dv1 = W.dv1; t.stride = W.dstride; t.p = p;
                    t.prefix = (meta >> 24) & 0xff;
                    t.n = t.prefix < diag::kMaxDepth ? t.prefix + 1 : t.prefix;
                    t.r0 = r0; t.r1 = W.dr1[k];
                    diag::record_path(px, t);
                    ++count;
                    const int sg = (meta >> 8) & 3;
                    sum[sg][0] += (double)r0.x; sum[sg][1] += (double)r0.y; sum[sg][2] += (double)r0.z;
                }
                for (int g = 0; g < 4; ++g) diag::welford_add(w[g], sum[g][0], sum[g][1], sum[g][2]);
            }
        }
    }
    if (count) atomicAdd(total, count);
}

// Film reset: every pixel takes the default PixelPathData pattern (1197
// 8-byte words), every Welford accumulator zero.
__global__ __launch_bounds__(256) void k_diag_init(uint64_t* film, const uint64_t* pattern, int64_t words,
                                                   double* strat, int64_t n_strat) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    constexpr int64_t kWords = (int64_t)sizeof(diag::PixelData) / 8;
    if (i < words) film[i] = pattern[i % kWords];
    if (i < n_strat) strat[i] = 0.0;
}

}  // namespace lucid

// ============================================================ host side ==
using namespace lucid;

// Device memory of lucid_render_tile_device_noalloc: the counters, bins,
// film and wave buffers of that render are carved from this one file-scope
// array, so the call allocates nothing.  Every GPU the library runs on holds
// its own copy for the life of the process.  Host code never uses the name
// (on the host it is a stand-in, not a device address): lucid_create looks
// up each device's copy with hipGetSymbolAddress once, with that device
// current, and the renders use that address.
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

std::atomic<uint64_t> g_scene_serial{0};
struct lucid_scene {
    std::atomic<int> refs{1};
    int device = 0;
    uint64_t serial = ++g_scene_serial;  // identifies the scene in the fused-pass arena cache
    PackedScene host;
    std::vector<void*> allocs;
    SceneDev dev;
    size_t grid_bytes = 0;  // device bytes of the density grids
    float bmin[3] = {0.0f, 0.0f, 0.0f}, bmax[3] = {1.0f, 1.0f, 1.0f};  // scene box
    int64_t n_nodes = 0;                  // traversal nodes
    int32_t built_on_device = 0;
};

// A stream with its own wavefront buffers.  Renders on a context use lane 0,
// whose stream is the context's; a batched session pass spreads its sessions
// over all lanes, so that the short tail of one session's launches overlaps
// the work of another's.  Four lanes = HIP's hardware queues per process.
constexpr int kLanes = 4;
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t wave_ev[2] = {nullptr, nullptr};  // pacing: at most two waves in flight
    int64_t cap_paths = 0;
    std::vector<void*> allocs;
    void* ctr = nullptr;  // queue / ray counters, kept across reallocations
    WaveDev W{};
    std::vector<void*> vallocs;  // deferred-transmittance factors (traceVolumeMIS)
    int64_t vcap = 0;
};

// The merged scenes of a fused pass (cached: the next pass over the same
// scenes reuses them).
struct FusedArena {
    std::vector<uint64_t> key;  // serials of the scenes, in first-use order
    std::vector<void*> allocs;
    SceneDev dev{};
    std::vector<int32_t> node_off, tri_off, mat_off, prog_off, light_off, native_off;
    int sd = 2;  // traversal stack entries: the deepest scene's
};
struct lucid_ctx {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Lane lanes[kLanes];
    int32_t lanes_ready = 1;  // lanes with a stream (lane 0 always)
    std::string err;
    std::atomic<int> cancel{0};
    // the cancel word the kernels of render_tile read: pinned host memory
    // mapped into the device's address space
    int32_t* cancel_host = nullptr;
    const int32_t* cancel_dev = nullptr;
    void* scratch = nullptr;  // this device's `scratch` array (no-allocation renders), or null
    std::mutex mu;
    int64_t wave_paths = 0;
    // Launch shape: workgroups per CU of the persistent kernels, or 0 for
    // one lane per queued ray.  Persistent traversal with ray refill beats
    // one lane per ray 1.9-2.0x (1080p room, 1M-triangle spheres): a
    // wavefront no longer waits for its longest ray.  16 groups per CU
    // saturate the LDS stacks; a fetch once 48 of 64 lanes are idle
    // amortises the atomic.  Shading: 8 groups per CU = what its register
    // use lets run at once.
    int32_t traverse_groups = 16, shade_groups = 8;
    int32_t refill = 48;  // idle lanes that make a wavefront fetch new rays
    int32_t walk_min = 16;  // traversal: node walking stops once at most this many lanes still walk
    int32_t leaf_tris = 4;  // traversal: triangles of a waiting leaf tested per leaf phase
    // 1: bounce and shadow rays regrouped by ray bin before their traversal
    // passes (depth >= 1); 0: never; 2: for waves of at least 2^24 paths.
    // The 1080p occluded room (33M paths a wave) renders 8 % faster with it
    // (shadow pass -22 %, closest hit -5 % with the binning included), the
    // 1M-triangle spheres (67M) 1 % slower; batched sessions of small scenes
    // (waves of 0.1-4M) lose 9 % to the binning launches.
    int32_t ray_order = 2;
    // 1: each bounce's live queue is regrouped by the shading class of its
    // hits before the shading launch (class_order); 0: compaction order
    int32_t shade_order = 0;
    int32_t session_graphs = 1;  // replay plain one-wave session passes from captured graphs
    // 1: a batched pass whose sessions all qualify runs fused (one launch per
    // kernel and bounce over all of them, lucid_set_session_fusion); 0: the
    // sessions are spread over the stream lanes
    int32_t session_fusion = 1;
    FusedArena* arena = nullptr;  // the merged scenes of the last fused pass
    FusedSession* fses = nullptr;        // its per-session records
    int64_t fses_cap = 0;
    int32_t grid_bits = 32;  // density-grid storage of scenes created from now on
    int32_t scene_build = 1;  // 1: BVHs and triangle records built on the device, 0: on the host
    float4* film = nullptr;
    int64_t film_cap = 0;
    float* out_dev = nullptr;
    int64_t out_cap = 0;
    lucid_stats stats{};
    int32_t profiling = 0;
    // learned path guiding
    guide::State* gstate = nullptr;
    lucid_guiding_config gcfg{0.5f, 16, 32768, 1e-2f, 32, 1u, 0, 0};
    int64_t gsteps = 0;
    __half* gout = nullptr;  // FP16 network outputs, 40 per path
    int64_t gout_cap = 0;
    float4* tslot = nullptr;
    int64_t tslot_cap = 0;
    // DiagnosticFilm (rebuilt by every diagnostic render, sized to the tile)
    int32_t diag_on = 0, diag_sub = 1;
    int32_t diag_tw = 0, diag_th = 0;
    bool diag_valid = false;
    diag::PixelData* dfilm = nullptr;
    int64_t dfilm_cap = 0;
    double* dstrat = nullptr;
    int64_t dstrat_cap = 0;
    unsigned long long* dtotal = nullptr;
    uint64_t* dpattern = nullptr;
    std::vector<void*> diag_allocs;  // per-path diagnostic storage
    int64_t diag_cap_paths = 0;
    int32_t diag_slots = 0;
    // LUCID_PROFILE_KERNELS: one event pair per launch, kind 0/1/2
    std::vector<hipEvent_t> kev;
    std::vector<int> kev_kind;
    size_t kev_used = 0;
};

// A progressive accumulation: one frame of one scene and camera, its
// radiance sum in device memory, the number of samples in it, and, while an
// async worker runs, the worker thread and the latest averaged snapshot.
struct lucid_session {
    lucid_ctx* ctx = nullptr;
    lucid_scene* scene = nullptr;
    lucid_camera cam{};
    int32_t w = 0, h = 0, max_depth = 8, algorithm = LUCID_ALGO_NEE;
    float4* accum = nullptr;  // radiance sum per pixel, rows top first
    int64_t samples = 0;      // samples in `accum`
    // async accumulation (render_start_async & co.)
    std::thread worker;
    std::atomic<bool> running{false};
    std::atomic<int> stop{0};
    std::atomic<bool> reset_pending{false};
    std::mutex pmu;  // camera / depth of a pending reset
    lucid_camera pend_cam{};
    int32_t pend_depth = 8;
    std::mutex smu;  // the snapshot
    std::vector<float> snap;
    int64_t snap_samples = 0;
    std::atomic<uint32_t> revision{0};
    float* snap_dev = nullptr;
    float* snap_pinned = nullptr;
    std::atomic<bool> failed{false};
    std::string err;  // why the worker ended early: written before `failed`
    // a pass replayed from a captured graph (plain passes of one wave)
    hipGraphExec_t gexec = nullptr;
    std::vector<uint8_t> gsig;  // what the graph (or the last plain pass) was made for
    bool gseen = false;         // a plain pass ran with gsig: capture on the next
    int32_t* s0_dev = nullptr;  // first sample index of the replayed pass
};

namespace {

int fail(lucid_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
#define HIP_OK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, LUCID_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int upload(lucid_ctx* ctx, lucid_scene* s, const std::vector<T>& v, const T** out) {
    *out = nullptr;
    if (v.empty()) return LUCID_OK;
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess) return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (scene)");
    s->allocs.push_back(p);
    HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T*>(p);
    return LUCID_OK;
}

// v0, e1, e2 of every triangle as 12 consecutive floats.
std::vector<float> interleave_triangles(const PackedScene& ps) {
    std::vector<float> t((size_t)ps.num_triangles * 12);
    for (size_t i = 0; i < (size_t)ps.num_triangles; ++i) {
        std::memcpy(&t[12 * i], &ps.tri_v0[4 * i], 16);
        std::memcpy(&t[12 * i + 4], &ps.tri_e1[4 * i], 16);
        std::memcpy(&t[12 * i + 8], &ps.tri_e2[4 * i], 16);
    }
    return t;
}

void free_wave(Lane& L) {
    for (void* p : L.allocs) (void)hipFree(p);
    L.allocs.clear();
    L.cap_paths = 0;
    L.W = WaveDev{};
    if (L.ctr) {
        L.W.cnt = static_cast<int32_t*>(L.ctr);
        L.W.rays = reinterpret_cast<unsigned long long*>(static_cast<char*>(L.ctr) + 64);
    }
}

int ensure_counters(lucid_ctx* ctx, Lane& L) {
    if (L.ctr) return LUCID_OK;
    if (hipMalloc(&L.ctr, kCounterBytes) != hipSuccess) {
        L.ctr = nullptr;
        return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (counters)");
    }
    HIP_OK(hipMemset(L.ctr, 0, kCounterBytes));
    L.W.cnt = static_cast<int32_t*>(L.ctr);
    L.W.rays = reinterpret_cast<unsigned long long*>(static_cast<char*>(L.ctr) + 64);
    return LUCID_OK;
}

int ensure_wave(lucid_ctx* ctx, Lane& L, int64_t n) {
    int rc = ensure_counters(ctx, L);
    if (rc) return rc;
    if (n <= L.cap_paths) return LUCID_OK;
    // the previous buffers may still be in use by queued work of this lane
    HIP_OK(hipStreamSynchronize(L.stream));
    free_wave(L);
    auto alloc = [&](size_t bytes, void** out) -> bool {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return false;
        L.allocs.push_back(p);
        *out = p;
        return true;
    };
    WaveDev& W = L.W;
    bool ok = alloc(n * 16, (void**)&W.ray_o) && alloc(n * 16, (void**)&W.ray_d) && alloc(n * 16, (void**)&W.thr) &&
              alloc(n * 16, (void**)&W.rad) && alloc(n * 16, (void**)&W.rng) && alloc(n * 16, (void**)&W.meta) &&
              alloc(n * 16, (void**)&W.hit) && alloc(n * 16, (void**)&W.sh_o) && alloc(n * 16, (void**)&W.sh_d) &&
              alloc(n * 16, (void**)&W.sh_c) && alloc(n * 4, (void**)&W.q[0]) && alloc(n * 4, (void**)&W.q[1]) &&
              alloc(n * 4, (void**)&W.qs) && alloc(n * 4, (void**)&W.sq) && alloc(n * 4, (void**)&W.skey) &&
              alloc(kShards * segment_capacity(n) * 4, (void**)&W.aq) &&
              alloc(kShards * segment_capacity(n) * 4, (void**)&W.aqs) && alloc(kRayBins * 4, (void**)&W.sbins);
    if (!ok) {
        free_wave(L);
        return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (wave buffers)");
    }
    W.seg_cap = segment_capacity(n);
    L.cap_paths = n;
    return LUCID_OK;
}

// sh_a / sh_e / sh_t of lane L for n paths (traceVolumeMIS with media).
int ensure_vol_wave(lucid_ctx* ctx, Lane& L, int64_t n) {
    if (n <= L.vcap) return LUCID_OK;
    HIP_OK(hipStreamSynchronize(L.stream));
    for (void* q : L.vallocs) (void)hipFree(q);
    L.vallocs.clear();
    L.vcap = 0;
    for (int k = 0; k < 3; ++k) {
        void* q = nullptr;
        if (hipMalloc(&q, (size_t)n * 16) != hipSuccess) {
            for (void* r : L.vallocs) (void)hipFree(r);
            L.vallocs.clear();
            return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (volume NEE factors)");
        }
        L.vallocs.push_back(q);
    }
    L.vcap = n;
    return LUCID_OK;
}

// Creates the streams and pacing events of lanes 1..kLanes-1 (default
// priority, non-blocking).
int ensure_lanes(lucid_ctx* ctx) {
    for (int i = ctx->lanes_ready; i < kLanes; ++i) {
        Lane& L = ctx->lanes[i];
        HIP_OK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreateWithFlags(&L.wave_ev[0], hipEventDisableTiming));
        HIP_OK(hipEventCreateWithFlags(&L.wave_ev[1], hipEventDisableTiming));
        ctx->lanes_ready = i + 1;
    }
    return LUCID_OK;
}

void free_diag_wave(lucid_ctx* ctx) {
    for (void* p : ctx->diag_allocs) (void)hipFree(p);
    ctx->diag_allocs.clear();
    ctx->diag_cap_paths = 0;
    ctx->diag_slots = 0;
}
// Per-path diagnostic storage: 16 vertices x 32 B, `slots` records x 32 B,
// one count word.
int ensure_diag_wave(lucid_ctx* ctx, int64_t n, int32_t slots) {
    if (n <= ctx->diag_cap_paths && slots <= ctx->diag_slots) return LUCID_OK;
    free_diag_wave(ctx);
    n = std::max(n, ctx->diag_cap_paths);
    slots = std::max(slots, ctx->diag_slots);
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        ctx->diag_allocs.push_back(p);
        return p;
    };
    void* a = alloc((size_t)n * diag::kMaxDepth * 16);
    void* b = alloc((size_t)n * diag::kMaxDepth * 16);
    void* c = alloc((size_t)n * (size_t)slots * 16);
    void* d = alloc((size_t)n * (size_t)slots * 16);
    void* e = alloc((size_t)n * 4);
    if (!a || !b || !c || !d || !e) {
        free_diag_wave(ctx);
        return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (diagnostic path storage)");
    }
    ctx->diag_cap_paths = n;
    ctx->diag_slots = slots;
    return LUCID_OK;
}
constexpr int kGuideRing = 1 << 21, kGuideBatchCap = 65536;
int ensure_guide(lucid_ctx* ctx) {
    if (ctx->gstate) return LUCID_OK;
    ctx->gstate = guide::create(kGuideRing, kGuideBatchCap);
    if (!ctx->gstate) return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (guiding network)");
    guide::reset(ctx->gstate, ctx->gcfg.seed, ctx->gcfg.alpha, ctx->stream);
    ctx->gsteps = 0;
    return LUCID_OK;
}

void free_diag_film(lucid_ctx* ctx) {
    if (ctx->dfilm) (void)hipFree(ctx->dfilm);
    if (ctx->dstrat) (void)hipFree(ctx->dstrat);
    ctx->dfilm = nullptr; ctx->dstrat = nullptr;
    ctx->dfilm_cap = ctx->dstrat_cap = 0;
    ctx->diag_valid = false;
}

template <class T>
int ensure_buf(lucid_ctx* ctx, T** p, int64_t* cap, int64_t n) {
    if (n <= *cap) return LUCID_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed");
    *cap = n;
    return LUCID_OK;
}

__global__ void k_zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0) *p = 0ull;
}

__global__ void k_zero_counters(int32_t* cnt, unsigned long long* rays) {
    if (threadIdx.x < 16) cnt[threadIdx.x] = 0;
    if (threadIdx.x < 8) rays[threadIdx.x] = 0ull;
    for (int j = threadIdx.x; j < 5 * kShards; j += blockDim.x) cnt[kFetchBase + j * kLineInts] = 0;
}
// zeroes the closest-hit fetch counters (a second query over the same queue)
__global__ void k_zero_fetch(int32_t* cnt) {
    for (int j = threadIdx.x; j < kShards; j += blockDim.x) cnt[kFetchBase + j * kLineInts] = 0;
}

inline unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Records a start event of a timed launch of `kind`; returns its slot or -1.
int kev_begin(lucid_ctx* ctx, int kind, hipStream_t st) {
    if (!(ctx->profiling & LUCID_PROFILE_KERNELS)) return -1;
    size_t i = ctx->kev_used;
    while (ctx->kev.size() < 2 * (i + 1)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        ctx->kev.push_back(e);
    }
    if (ctx->kev_kind.size() < i + 1) ctx->kev_kind.resize(i + 1);
    ctx->kev_kind[i] = kind;
    ctx->kev_used = i + 1;
    (void)hipEventRecord(ctx->kev[2 * i], st);
    return (int)i;
}
void kev_end(lucid_ctx* ctx, int slot, hipStream_t st) {
    if (slot >= 0) (void)hipEventRecord(ctx->kev[2 * (size_t)slot + 1], st);
}

int validate(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr) {
    if (!ctx) return LUCID_ERR_INVALID;
    if (!scene || !cam || !pr) return fail(ctx, LUCID_ERR_INVALID, "null scene, camera or params");
    if (pr->tile_w <= 0 || pr->tile_h <= 0 || pr->full_w <= 0 || pr->full_h <= 0)
        return fail(ctx, LUCID_ERR_INVALID, "tile and image sizes must be positive");
    if ((int64_t)pr->tile_w * pr->tile_h > (1LL << 28)) return fail(ctx, LUCID_ERR_INVALID, "tile too large");
    if (pr->samples < 0 || pr->max_depth < 0) return fail(ctx, LUCID_ERR_INVALID, "negative samples or depth");
    if (pr->algorithm < 0 || pr->algorithm > LUCID_ALGO_GUIDED_MIS) return fail(ctx, LUCID_ERR_INVALID, "unknown algorithm");
    if (scene->device != ctx->device) return fail(ctx, LUCID_ERR_INVALID, "scene belongs to another device");
    return LUCID_OK;
}

// Runs the waves on lane L.  mode 0: film (render_tile), 1: per-sample
// radiance into samples_dev, 2: primary hits only, 3: session pass (the
// samples are added to `accum`, which is not cleared, with jitter on every
// sample as in a render_tile call of more than one sample; no diagnostics).
struct RunOut {
    int32_t samples_done = 0;  // samples of the call whose waves were issued
    float* samples_dev = nullptr;
    int32_t *mesh = nullptr, *tri = nullptr;
    float *u = nullptr, *v = nullptr, *t = nullptr;
};

// Wave buffers carved from caller-provided device memory instead of the
// lane's allocations (the no-allocation render): W's path arrays hold
// cap_paths paths, film holds the tile.
struct FixedWave {
    WaveDev W;
    float4* film;
    int64_t cap_paths;
};

int run_waves(lucid_ctx* ctx, Lane& L, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
              int mode, hipStream_t st, RunOut* ro, bool* cancelled, float4* accum = nullptr, int active_lanes = 1,
              const std::atomic<int>* cancel_flag = nullptr, const int32_t* s0_dev = nullptr,
              bool capturing = false, const FixedWave* fixed = nullptr) {
    *cancelled = false;
    if (!cancel_flag) cancel_flag = &ctx->cancel;
    const int64_t P = (int64_t)pr->tile_w * pr->tile_h;
    const int S = pr->samples;
    // Wavefront batch: as many paths as fit a quarter of free device memory,
    // at most 2^26.  Each depth iteration costs one launch per kernel whose
    // duration has a floor set by the slowest wavefront's traversal, so few
    // large batches beat many small ones.
    int64_t target = fixed ? fixed->cap_paths : ctx->wave_paths;
    if (target <= 0) {
        size_t free_b = 0, total_b = 0;
        target = (int64_t)1 << 26;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) {
            const int64_t fit = (int64_t)(free_b / 4 / kWaveBytesPerPath) / active_lanes + L.cap_paths;
            target = std::max<int64_t>(std::min<int64_t>(target, fit), (int64_t)1 << 16);
        }
    }
    // Diagnostic renders (render_tile with diagnostics enabled) also keep
    // per-path vertices and records: smaller batches, at most ~2 GB of them.
    const bool diag = mode == 0 && ctx->diag_on;
    // Guided renders keep 40 network outputs per path: batches of <= 2^25.
    const bool guided = !diag && pr->algorithm == LUCID_ALGO_GUIDED_MIS;
    if (guided) target = std::min<int64_t>(target, (int64_t)1 << 25);
    const int32_t slots = std::max(1, pr->max_depth);
    if (diag) {
        const int64_t per_path = (int64_t)diag::kMaxDepth * 32 + (int64_t)slots * 32 + 4;
        target = std::min<int64_t>(target, std::max<int64_t>((int64_t)1 << 16, ((int64_t)2 << 30) / per_path));
    }
    int64_t B = std::max<int64_t>(1, std::min<int64_t>(S, target / P));
    int rc = LUCID_OK;
    float4* film = accum;
    if (fixed) {
        if (P * B > fixed->cap_paths || diag || guided) return fail(ctx, LUCID_ERR_INVALID, "fixed wave misused");
        film = fixed->film;
    } else {
        rc = ensure_wave(ctx, L, P * B);
        if (rc) return rc;
        if (mode != 3) {
            rc = ensure_buf(ctx, &ctx->film, &ctx->film_cap, P);
            if (rc) return rc;
            film = ctx->film;
        }
    }
    SceneDev SD = scene->dev;
    WaveDev W = fixed ? fixed->W : L.W;
    W.sh_a = W.sh_e = W.sh_t = nullptr;
    const int32_t dsub = ctx->diag_sub;
    const int32_t dsw = pr->tile_w / dsub, dsh = pr->tile_h / dsub;
    if (diag) {
        rc = ensure_diag_wave(ctx, P * B, slots);
        if (rc) return rc;
        W.dv0 = static_cast<float4*>(ctx->diag_allocs[0]);
        W.dv1 = static_cast<float4*>(ctx->diag_allocs[1]);
        W.dr0 = static_cast<float4*>(ctx->diag_allocs[2]);
        W.dr1 = static_cast<float4*>(ctx->diag_allocs[3]);
        W.dcnt = static_cast<int32_t*>(ctx->diag_allocs[4]);
        W.dstride = ctx->diag_cap_paths;
        // the film is rebuilt for this tile (pybind_renderer.hpp:206-208)
        const int64_t fp = std::max<int64_t>(1, (int64_t)dsw * dsh);
        rc = ensure_buf(ctx, &ctx->dfilm, &ctx->dfilm_cap, fp);
        if (rc) return rc;
        rc = ensure_buf(ctx, &ctx->dstrat, &ctx->dstrat_cap, fp * 28);
        if (rc) return rc;
        if (!ctx->dpattern) {
            diag::PixelData pat;
            diag::pixel_init(pat);
            HIP_OK(hipMalloc((void**)&ctx->dpattern, sizeof(pat)));
            HIP_OK(hipMemcpy(ctx->dpattern, &pat, sizeof(pat), hipMemcpyHostToDevice));
            HIP_OK(hipMalloc((void**)&ctx->dtotal, sizeof(unsigned long long)));
        }
        const int64_t words = (int64_t)dsw * dsh * (int64_t)(sizeof(diag::PixelData) / 8);
        const int64_t nst = (int64_t)dsw * dsh * 28;
        const int64_t nin = std::max<int64_t>(std::max(words, nst), 1);
        k_diag_init<<<blocks(nin, 256), 256, 0, st>>>(reinterpret_cast<uint64_t*>(ctx->dfilm), ctx->dpattern, words,
                                                    ctx->dstrat, nst);
        k_zero_u64<<<1, 64, 0, st>>>(ctx->dtotal);
        ctx->diag_tw = pr->tile_w;
        ctx->diag_th = pr->tile_h;
        ctx->diag_valid = true;
        // the diagnostic tracers use intersectScene directly, without
        // volume-boundary skipping (path_tracer.hpp:893, 1314)
        SD.any_vbo = 0;
    } else {
        W.dv0 = W.dv1 = W.dr0 = W.dr1 = nullptr;
        W.dcnt = nullptr;
        W.dstride = 0;
    }
    guide::GuideDev G;
    std::memset(&G, 0, sizeof(G));
    W.gout = nullptr;
    W.gspare = 0;
    W.tslot = nullptr;
    W.tstride = 1;
    W.tnrec = 0;
    const bool train = guided && ctx->gcfg.train_steps > 0;
    if (guided) {
        rc = ensure_guide(ctx);
        if (rc) return rc;
        rc = ensure_buf(ctx, &ctx->gout, &ctx->gout_cap, (P * B + 1) * guide::kGoutStride);
        if (rc) return rc;
        W.gout = ctx->gout;
        W.gspare = P * B;
        if (train) {
            W.tstride = std::max(1, ctx->gcfg.record_stride);
            const int64_t nrec = (P * B + W.tstride - 1) / W.tstride;
            rc = ensure_buf(ctx, &ctx->tslot, &ctx->tslot_cap, nrec * kTrainSlots * 5);
            if (rc) return rc;
            W.tnrec = nrec;
            W.tslot = ctx->tslot;
        }
        // the guide samples only once it has been trained
        G = guide::dev(ctx->gstate, scene->bmin, scene->bmax, ctx->gsteps > 0 ? ctx->gcfg.alpha : 0.0f,
                       ctx->gcfg.learn_selection != 0);
    }
    // LDS stack sized to the scene's BVH depth; persistent grids of a few
    // workgroups per CU walk the live queues (grid-stride).
    const int sd = std::min(kStackMax, scene->host.max_depth + 2);
    const size_t lds = (size_t)sd * kBlock * sizeof(int32_t);
    // the launch shape of this render, read once (callers hold ctx->mu)
    const int32_t traverse_groups = ctx->traverse_groups, shade_groups = ctx->shade_groups;
    const int32_t refill_lanes = ctx->refill, ray_order = ctx->ray_order, shade_order = ctx->shade_order;
    const int32_t walk_min = ctx->walk_min, leaf_tris = ctx->leaf_tris;
    // traverse_groups > 0: persistent traversal with ray refill on that many
    // workgroups per CU; 0: one lane per queued ray
    const bool refill = traverse_groups > 0;
    // Batched session passes run up to four lanes at once: each gets half
    // the groups (measured best on 32 sessions).  Small waves keep at least
    // kMinRaysPerLane rays per lane, so no wavefront is left near-idle.
    const int tgroups = active_lanes > 1 ? std::max(4, traverse_groups / 2) : traverse_groups;
    const int64_t cap_trav = refill ? std::min<int64_t>((int64_t)ctx->num_cu * tgroups,
                                                        blocks(P * B, kBlock * kMinRaysPerLane))
                                    : INT64_MAX;
    const int64_t cap_shade = shade_groups > 0 ? (int64_t)ctx->num_cu * shade_groups : INT64_MAX;
    // the guided shading kernel runs with its own group count, whatever
    // shape the plain kernels were tuned to
    const int64_t cap_shade_guided = shade_groups > 0 ? (int64_t)ctx->num_cu * kGuidedShadeGroups : INT64_MAX;
    W.count_work = (ctx->profiling & LUCID_PROFILE_WORK) ? 1 : 0;
    W.stop = (mode == 0 && cancel_flag == &ctx->cancel && !capturing) ? ctx->cancel_dev : nullptr;
    if (mode != 3) {  // a session batch zeroes its lanes' counters once
        ctx->kev_used = 0;
        k_zero_counters<<<1, 64, 0, st>>>(W.cnt, W.rays);
        k_film_clear<<<blocks(P, 256), 256, 0, st>>>(film, (int32_t)P);
    }
    CameraBasis cb = camera_basis(*cam, pr->full_w, pr->full_h);
    CamDev cd;
    std::memcpy(cd.pos, cb.pos, 12); std::memcpy(cd.fwd, cb.fwd, 12);
    std::memcpy(cd.right, cb.right, 12); std::memcpy(cd.up, cb.up, 12);
    cd.aspect = cb.aspect; cd.scale = cb.scale;
    cd.fw = pr->full_w; cd.fh = pr->full_h; cd.tx = pr->tile_x; cd.ty = pr->tile_y;
    cd.tw = pr->tile_w; cd.th = pr->tile_h;
    cd.samples_in_call = mode == 3 ? std::max(S, 2) : S;
    cd.P = (int32_t)P;
    cd.max_depth = pr->max_depth;
    cd.s0p = s0_dev;
    uint64_t waves = 0;
    const int max_depth = mode == 2 ? 1 : pr->max_depth;
    // traceVolumeSimple / traceVolumeMIS reduce exactly to traceSimple /
    // traceMIS on a scene without participating media (no distance-sampling
    // draws, unit transmittance).
    int algo = pr->algorithm;
    if ((algo == LUCID_ALGO_VOLUME_SIMPLE || algo == LUCID_ALGO_VOLUME_MIS) && !scene->host.any_volume)
        algo = algo == LUCID_ALGO_VOLUME_SIMPLE ? LUCID_ALGO_SIMPLE : LUCID_ALGO_MIS;
    // with diagnostics on, "simple" and "mis" record through their
    // diagnostic tracers and every other algorithm through the NEE one
    // (pybind_renderer.hpp:296-307)
    if (diag)
        algo = pr->algorithm == LUCID_ALGO_SIMPLE ? kAlgoDiagSimple
               : (pr->algorithm == LUCID_ALGO_MIS || pr->algorithm == LUCID_ALGO_GUIDED_MIS) ? kAlgoDiagMis
                                                                                               : kAlgoDiagNee;
    // traceVolumeMIS through gridded media marches its NEE transmittance in
    // k_vol_march, between shading and the shadow rays
    const bool march = algo == LUCID_ALGO_VOLUME_MIS && SD.n_vols > 0;
    if (march && fixed) {
        if (!fixed->W.sh_a) return fail(ctx, LUCID_ERR_INVALID, "fixed wave without volume factors");
        W.sh_a = fixed->W.sh_a;
        W.sh_e = fixed->W.sh_e;
        W.sh_t = fixed->W.sh_t;
    } else if (march) {
        rc = ensure_vol_wave(ctx, L, P * B);
        if (rc) return rc;
        W.sh_a = static_cast<float4*>(L.vallocs[0]);
        W.sh_e = static_cast<float4*>(L.vallocs[1]);
        W.sh_t = static_cast<float4*>(L.vallocs[2]);
    }
    const bool bin = ray_order == 1 || (ray_order == 2 && P * B >= ((int64_t)1 << 24));
    // ray order 3: only the shadow rays are regrouped (the closest-hit
    // traversal, which walks nodes until most lanes hold a leaf, gains little
    // from it, the shadow rays about their sort's cost twice over)
    const bool bin_bounce = bin, bin_shadow = bin || ray_order == 3;
    BinBox bin_box;
    for (int k = 0; k < 3; ++k) {
        bin_box.lo[k] = scene->bmin[k];
        const float ext = scene->bmax[k] - scene->bmin[k];
        bin_box.sc[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    for (int s0 = 0; s0 < S; s0 += (int)B) {
        // Pacing: wait for the wave before the previous one, so at most two
        // waves are queued and a cancel takes effect within one wave.
        if (waves >= 2) (void)hipEventSynchronize(L.wave_ev[waves & 1]);
        if (cancel_flag->load(std::memory_order_relaxed)) {
            *cancelled = true;
            break;
        }
        int b = (int)std::min<int64_t>(B, S - s0);
        int64_t n = P * b;
        cd.s0 = pr->sample_offset + s0;
        cd.B = b;
        k_raygen<<<blocks(n, kBlock), kBlock, 0, st>>>(W, cd, (int32_t)n);
        int q = 0;
        for (int depth = 0; depth < max_depth; ++depth) {
            const unsigned g = (unsigned)std::min<int64_t>(blocks(n, kBlock), cap_trav);
            const unsigned gs = (unsigned)std::min<int64_t>(blocks(n, kBlock), cap_shade);
            const unsigned gsg = (unsigned)std::min<int64_t>(blocks(n, kBlock), cap_shade_guided);
            // the closest-hit pass reads the live queue through Wd
            WaveDev Wd = W;
            if (bin_bounce && depth > 0) {
                const int kb = kev_begin(ctx, 5, st);
                bin_rays<kBinBounce>(W, BinSrc{W.q[q], W.cnt + q, W.ray_o, W.ray_d}, bin_box, n, ctx->num_cu, st);
                kev_end(ctx, kb, st);
                Wd.q[q] = W.sq;
            }
            int ks = kev_begin(ctx, 0, st);
            if (refill)
                k_trace_refill<0><<<g, kBlock, lds, st>>>(SD, Wd, q, sd, refill_lanes, walk_min, leaf_tris);
            else
                k_intersect<<<g, kBlock, lds, st>>>(SD, Wd, q, sd);
            kev_end(ctx, ks, st);
            if (mode == 2) break;
            // the shading pass reads the live queue through Wsh
            WaveDev Wsh = W;
            if (shade_order == 1) {
                const int kb = kev_begin(ctx, 5, st);
                class_order(SD, W, W.q[q], W.cnt + q, W.sq, n, ctx->num_cu, st);
                kev_end(ctx, kb, st);
                Wsh.q[q] = W.sq;
            }
            // the last bounce samples no next direction: its NEE keeps the
            // plain weights, as the plain tracer's last bounce does
            const bool guide_here = guided && G.alpha > 0.0f && depth + 1 < max_depth &&
                                    (ctx->gcfg.max_depth <= 0 || depth < ctx->gcfg.max_depth);
            if (guide_here) {
                ks = kev_begin(ctx, 3, st);
                const unsigned gg =
                    (unsigned)std::min<int64_t>(blocks(n, kInferWaves * 64), (int64_t)ctx->num_cu * kInferGroupsPerCU);
                k_guide_infer<<<gg, kInferWaves * 64, 0, st>>>(SD, Wsh, G, q);
                kev_end(ctx, ks, st);
            }
            ks = kev_begin(ctx, 1, st);
            switch (algo) {
                case LUCID_ALGO_SIMPLE:
                    launch_shade<LUCID_ALGO_SIMPLE>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                case LUCID_ALGO_MIS:
                    launch_shade<LUCID_ALGO_MIS>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                case LUCID_ALGO_VOLUME_SIMPLE:
                    launch_shade<LUCID_ALGO_VOLUME_SIMPLE>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                case LUCID_ALGO_VOLUME_MIS:
                    launch_shade<LUCID_ALGO_VOLUME_MIS>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                case LUCID_ALGO_GUIDED_MIS: {
                    // Past the guided bounces (depth < gcfg.max_depth) a path
                    // is a plain MIS path: the bounce after the last guided
                    // one still weighs its emitter hits against that bounce's
                    // guide density and closes its training slot, every later
                    // one shades with the plain kernel, which draws the same
                    // numbers and computes the same values as the guided
                    // kernel with the guide off (rho = 1, no slots left open).
                    const int gd = ctx->gcfg.max_depth;
                    if (gd > 0 && depth > gd) {
                        launch_shade<LUCID_ALGO_MIS>(gs, st, SD, Wsh, G, q, max_depth);
                        break;
                    }
                    guide::GuideDev Gd = G;
                    if (!guide_here) Gd.alpha = 0.0f;  // this bounce samples the BSDF only
                    launch_shade<LUCID_ALGO_GUIDED_MIS>(gsg, st, SD, Wsh, Gd, q, max_depth);
                    break;
                }
                case kAlgoDiagSimple:
                    launch_shade<kAlgoDiagSimple>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                case kAlgoDiagNee:
                    launch_shade<kAlgoDiagNee>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                case kAlgoDiagMis:
                    launch_shade<kAlgoDiagMis>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
                default:
                    launch_shade<LUCID_ALGO_NEE>(gs, st, SD, Wsh, G, q, max_depth);
                    break;
            }
            concat_queues(W, 1 - q, n, st);
            if (march) {
                const unsigned gm = (unsigned)std::min<int64_t>(blocks(n, kBlock), (int64_t)ctx->num_cu * 16);
                k_vol_march<<<gm, kBlock, 0, st>>>(SD, W, refill_lanes);
            }
            kev_end(ctx, ks, st);
            // the shadow pass reads the NEE queue through Ws
            WaveDev Ws = W;
            if (bin_shadow && depth > 0) {
                const int kb = kev_begin(ctx, 5, st);
                bin_rays<kBinShadow>(W, BinSrc{W.qs, W.cnt + 2, W.sh_o, W.sh_d}, bin_box, n, ctx->num_cu, st);
                kev_end(ctx, kb, st);
                Ws.qs = W.sq;
            }
            ks = kev_begin(ctx, 2, st);
            if (refill && SD.any_vbo)
                k_trace_refill<2><<<g, kBlock, lds, st>>>(SD, Ws, 0, sd, refill_lanes, walk_min, leaf_tris);
            else if (refill)
                k_trace_refill<1><<<g, kBlock, lds, st>>>(SD, Ws, 0, sd, refill_lanes, walk_min, leaf_tris);
            else
                k_shadow<<<g, kBlock, lds, st>>>(SD, Ws, sd);
            kev_end(ctx, ks, st);
            q = 1 - q;
        }
        if (train) {  // timed with the training
            const int ks = kev_begin(ctx, 4, st);
            k_guide_collect<<<blocks((n + W.tstride - 1) / W.tstride, 256), 256, 0, st>>>(
                W, (int32_t)n, guide::ring(ctx->gstate), guide::ring_total(ctx->gstate),
                guide::ring_capacity(ctx->gstate));
            kev_end(ctx, ks, st);
        }
        if (mode == 0 || mode == 3) {
            k_film<<<blocks(P, 256), 256, 0, st>>>(W, film, (int32_t)P, b);
            if (diag && dsw > 0 && dsh > 0)
                k_diag_film<<<blocks((int64_t)dsw * dsh, 64), 64, 0, st>>>(
                    W, ctx->dfilm, reinterpret_cast<diag::StrategyWelford*>(ctx->dstrat), ctx->dtotal, pr->tile_w,
                    dsub, dsw, dsh, (int32_t)P, b);
        } else if (mode == 1) {
            k_samples_out<<<blocks(n, 256), 256, 0, st>>>(W, ro->samples_dev, (int32_t)P, b, s0, S);
        } else {
            k_hits_out<<<blocks(n, 256), 256, 0, st>>>(SD, W, ro->mesh, ro->tri, ro->u, ro->v, ro->t, (int32_t)P, b,
                                                       S, s0);
        }
        if (!capturing) (void)hipEventRecord(L.wave_ev[waves & 1], st);
        ++waves;
        if (ro) ro->samples_done = s0 + b;
    }
    // online training after the render (its time is part of the render's)
    if (train && !*cancelled) {
        unsigned long long recorded = 0;
        HIP_OK(hipMemcpyAsync(&recorded, guide::ring_total(ctx->gstate), 8, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (recorded > 0) {  // nothing to learn from before the first sample
            const int ks = kev_begin(ctx, 4, st);
            guide::train(ctx->gstate, G, ctx->gcfg.train_steps, ctx->gcfg.batch, ctx->gcfg.learning_rate, st);
            kev_end(ctx, ks, st);
            ctx->gsteps += ctx->gcfg.train_steps;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, LUCID_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    if (mode == 3) {
        ctx->stats.paths += (uint64_t)P * S;
        ctx->stats.waves += waves;
    } else {
        ctx->stats.paths = (uint64_t)P * S;
        ctx->stats.waves = waves;
    }
    return LUCID_OK;
}

// Reads the ray / work counters of lanes [0, nlanes) into ctx->stats.
int finish_stats(lucid_ctx* ctx, int nlanes = 1) {
    unsigned long long rays[6] = {0, 0, 0, 0, 0, 0};
    int32_t cnt[4] = {0, 0, 0, 0};
    for (int i = 0; i < nlanes; ++i) {
        Lane& L = ctx->lanes[i];
        if (!L.W.rays) continue;
        unsigned long long r[6];
        int32_t c[4];
        HIP_OK(hipMemcpyAsync(r, L.W.rays, sizeof(r), hipMemcpyDeviceToHost, L.stream));
        HIP_OK(hipMemcpyAsync(c, L.W.cnt, sizeof(c), hipMemcpyDeviceToHost, L.stream));
        HIP_OK(hipStreamSynchronize(L.stream));
        for (int k = 0; k < 6; ++k) rays[k] += r[k];
        cnt[3] += c[3];
    }
    ctx->stats.rays_closest = rays[0];
    ctx->stats.rays_shadow = rays[1];
    ctx->stats.stack_overflows = (uint64_t)cnt[3];
    ctx->stats.nodes_closest = rays[2];
    ctx->stats.tris_closest = rays[3];
    ctx->stats.nodes_shadow = rays[4];
    ctx->stats.tris_shadow = rays[5];
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->stats.ms_render = ms;
    double kms[6] = {0, 0, 0, 0, 0, 0};
    uint64_t kn[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < ctx->kev_used; ++i) {
        float e = 0.0f;
        if (hipEventElapsedTime(&e, ctx->kev[2 * i], ctx->kev[2 * i + 1]) == hipSuccess) {
            kms[ctx->kev_kind[i]] += e;
            kn[ctx->kev_kind[i]] += 1;
        }
    }
    ctx->stats.ms_intersect = kms[0]; ctx->stats.ms_shade = kms[1]; ctx->stats.ms_shadow = kms[2];
    ctx->stats.launches_intersect = kn[0]; ctx->stats.launches_shade = kn[1]; ctx->stats.launches_shadow = kn[2];
    ctx->stats.ms_guide_infer = kms[3]; ctx->stats.ms_guide_train = kms[4];
    ctx->stats.ms_ray_binning = kms[5];
    if (cnt[3] != 0) return fail(ctx, LUCID_ERR_DEVICE, "BVH traversal stack overflow");
    return LUCID_OK;
}

}  // namespace

// ================================================================ C ABI ==
extern "C" {

int lucid_create(int32_t device, lucid_ctx** out) {
    if (!out) return LUCID_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return LUCID_ERR_DEVICE;
    if (device < 0 || device >= n) return LUCID_ERR_INVALID;
    lucid_ctx* ctx = new (std::nothrow) lucid_ctx();
    if (!ctx) return LUCID_ERR_NOMEM;
    ctx->device = device;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
        ctx->num_cu = ncu;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->lanes[0].wave_ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->lanes[0].wave_ev[1], hipEventDisableTiming) != hipSuccess) {
        delete ctx;
        return LUCID_ERR_DEVICE;
    }
    ctx->lanes[0].stream = ctx->stream;
    void* dp = nullptr;
    // coherent (uncached on the device): every kernel read sees the host's
    // latest write, whichever XCD it runs on
    if (hipHostMalloc((void**)&ctx->cancel_host, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(&dp, ctx->cancel_host, 0) != hipSuccess) {
        if (ctx->cancel_host) (void)hipHostFree(ctx->cancel_host);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return LUCID_ERR_DEVICE;
    }
    __atomic_store_n(ctx->cancel_host, 0, __ATOMIC_SEQ_CST);
    ctx->cancel_dev = static_cast<const int32_t*>(dp);
    if (device < kMaxDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch[device]) {  // once per device, with the device current
            void* sp = nullptr;
            if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess && sp)
                g_scratch[device] = sp;
            else
                (void)hipGetLastError();  // no-allocation renders then report the failure
        }
        ctx->scratch = g_scratch[device];
    }
    *out = ctx;
    return LUCID_OK;
}

}  // extern "C"
namespace {
void free_arena(lucid_ctx* ctx);
}  // namespace
extern "C" {

void lucid_destroy(lucid_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < ctx->lanes_ready; ++i) {
        (void)hipStreamSynchronize(ctx->lanes[i].stream);
        free_wave(ctx->lanes[i]);
        if (ctx->lanes[i].ctr) (void)hipFree(ctx->lanes[i].ctr);
        for (void* q : ctx->lanes[i].vallocs) (void)hipFree(q);
    }
    guide::destroy(ctx->gstate);
    if (ctx->gout) (void)hipFree(ctx->gout);
    if (ctx->tslot) (void)hipFree(ctx->tslot);
    free_diag_wave(ctx);
    free_diag_film(ctx);
    if (ctx->dpattern) (void)hipFree(ctx->dpattern);
    if (ctx->dtotal) (void)hipFree(ctx->dtotal);
    if (ctx->film) (void)hipFree(ctx->film);
    if (ctx->out_dev) (void)hipFree(ctx->out_dev);
    free_arena(ctx);
    if (ctx->fses) (void)hipFree(ctx->fses);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    for (int i = 0; i < ctx->lanes_ready; ++i) {
        (void)hipEventDestroy(ctx->lanes[i].wave_ev[0]);
        (void)hipEventDestroy(ctx->lanes[i].wave_ev[1]);
        if (i > 0) (void)hipStreamDestroy(ctx->lanes[i].stream);
    }
    for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->cancel_host) (void)hipHostFree(ctx->cancel_host);
    delete ctx;
}

const char* lucid_last_error(const lucid_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int lucid_scene_create(lucid_ctx* ctx, const lucid_mesh_desc* meshes, int32_t num_meshes,
                       const lucid_light_desc* lights, int32_t num_lights, const float env_color[3],
                       float env_strength, lucid_scene** out) {
    if (!ctx || !out || !env_color) return fail(ctx, LUCID_ERR_INVALID, "null argument");
    *out = nullptr;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    lucid_scene* s = new (std::nothrow) lucid_scene();
    if (!s) return fail(ctx, LUCID_ERR_NOMEM, "out of host memory");
    s->device = ctx->device;
    std::string err;
    const bool device_build = ctx->scene_build == 1;
    if (!pack_scene(meshes, num_meshes, lights, num_lights, env_color, env_strength, s->host, err, !device_build)) {
        delete s;
        return fail(ctx, LUCID_ERR_INVALID, err);
    }
    SceneDev& D = s->dev;
    std::memset(&D, 0, sizeof(D));
    GNode root_node;
    std::memset(&root_node, 0, sizeof(root_node));
    int rc = LUCID_OK;
    if (device_build) {
        build::DeviceGeometry g;
        const bool ok = build::build_geometry(meshes, num_meshes, s->host, ctx->stream, g, err);
        s->allocs.insert(s->allocs.end(), g.allocs.begin(), g.allocs.end());
        if (!ok) {
            lucid_scene_release(s);
            return fail(ctx, LUCID_ERR_DEVICE, err);
        }
        D.nodes = g.nodes;
        D.tri = g.tri;
        D.fn = g.fn;
        D.vn = g.vn;
        D.tuv = g.tuv;
        s->host.root = g.root;
        s->host.max_depth = g.max_depth;
        s->n_nodes = g.n_nodes;
        s->built_on_device = 1;
        root_node = g.root_node;
    }
    const PackedScene& ps = s->host;
    if (ps.max_depth + 2 > kStackMax) {
        lucid_scene_release(s);
        return fail(ctx, LUCID_ERR_INVALID, "BVH deeper than the traversal stack");
    }
    auto up4 = [&](const std::vector<float>& v, const float4** dst) {
        const float* p = nullptr;
        int rc = upload(ctx, s, v, &p);
        *dst = reinterpret_cast<const float4*>(p);
        return rc;
    };
    if (!device_build) {
        rc = upload(ctx, s, ps.nodes, &D.nodes);
        // triangles interleaved: a leaf's triangles are one contiguous block
        std::vector<float> tri3 = interleave_triangles(ps);
        if (!rc) rc = up4(tri3, &D.tri);
        if (!rc) rc = up4(ps.tri_fn, &D.fn);
        if (!rc) rc = up4(ps.tri_vn, &D.vn);
        if (!rc) {
            const float* uv = nullptr;
            rc = upload(ctx, s, ps.tri_uv, &uv);
            D.tuv = reinterpret_cast<const float2*>(uv);
        }
        s->n_nodes = (int64_t)ps.nodes.size();
        if (!ps.nodes.empty()) root_node = ps.nodes[ps.root];
    }
    if (!rc) rc = upload(ctx, s, ps.mats, &D.mats);
    if (!rc) rc = upload(ctx, s, ps.node_ops, &D.prog);
    if (!rc) rc = upload(ctx, s, ps.lights, &D.lights);
    if (!rc) rc = upload(ctx, s, ps.cdf, &D.cdf);
    if (!rc) rc = upload(ctx, s, ps.native_light_ids, &D.native_ids);
    if (!rc) rc = upload(ctx, s, ps.volumes, &D.vols);  // replaced below for FP16 grids
    if (!rc && ctx->grid_bits == 16 && !ps.grid.empty()) {
        // FP16 storage: round to nearest even, and take each medium's
        // majorant from the rounded values it will be tracked against
        std::vector<__half> g16(ps.grid.size());
        for (size_t i = 0; i < g16.size(); ++i) {
            if (!(std::fabs(ps.grid[i]) <= 65504.0f)) {
                lucid_scene_release(s);
                return fail(ctx, LUCID_ERR_INVALID, "density grid value outside the FP16 range");
            }
            g16[i] = __float2half_rn(ps.grid[i]);
        }
        std::vector<GVolume> vols = ps.volumes;
        for (GVolume& v : vols) {
            if (!v.has_grid) continue;
            const size_t n = (size_t)v.dims[0] * v.dims[1] * v.dims[2];
            float mx = 0.0f;
            for (size_t i = 0; i < n; ++i) mx = std::max(mx, __half2float(g16[(size_t)v.grid_offset + i]));
            v.grid_max = mx;
        }
        const GVolume* dv = nullptr;
        rc = upload(ctx, s, vols, &dv);
        if (!rc) {
            D.vols = dv;
            s->host.volumes = vols;
            rc = upload(ctx, s, g16, &D.grid16);
        }
        s->grid_bytes = g16.size() * sizeof(__half);
    } else if (!rc) {
        rc = upload(ctx, s, ps.grid, &D.grid);
        s->grid_bytes = ps.grid.size() * sizeof(float);
    }
    if (rc) {
        lucid_scene_release(s);
        return rc;
    }
    D.n_lights = (int32_t)ps.lights.size();
    D.n_native = (int32_t)ps.native_light_ids.size();
    D.n_vols = (int32_t)ps.volumes.size();
    D.root = ps.root;
    D.total_area = ps.total_area;
    std::memcpy(D.env, ps.env, sizeof(D.env));
    D.native_start = ps.native_start;
    // scene box = union of the root's child boxes
    if (s->n_nodes > 0) {
        const GNode& r = root_node;
        for (int c = 0; c < 3; ++c) {
            s->bmin[c] = r.b[c];
            s->bmax[c] = r.b[3 + c];
            if (r.child[1] != kEmptyChild) {
                s->bmin[c] = std::min(s->bmin[c], r.b[6 + c]);
                s->bmax[c] = std::max(s->bmax[c], r.b[9 + c]);
            }
        }
    }
    D.any_vbo = 0;
    for (const GMat& m : ps.mats)
        if (m.flags & MF_VOLUME_BOUNDARY) D.any_vbo = 1;
    *out = s;
    return LUCID_OK;
}

void lucid_scene_retain(lucid_scene* s) {
    if (s) s->refs.fetch_add(1);
}
void lucid_scene_release(lucid_scene* s) {
    if (!s) return;
    if (s->refs.fetch_sub(1) == 1) {
        (void)hipSetDevice(s->device);
        (void)hipDeviceSynchronize();
        for (void* p : s->allocs) (void)hipFree(p);
        delete s;
    }
}
int32_t lucid_scene_num_triangles(const lucid_scene* s) { return s ? s->host.num_triangles : 0; }

int lucid_render_tile_device(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                             const lucid_render_params* pr, float* out_rgba_dev, void* stream) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!out_rgba_dev) return fail(ctx, LUCID_ERR_INVALID, "null output");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    // a cancel that came before the render is dropped, as render_tile does
    ctx->cancel.store(0);
    __atomic_store_n(ctx->cancel_host, 0, __ATOMIC_SEQ_CST);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    HIP_OK(hipEventRecord(ctx->ev0, st));
    bool cancelled = false;
    rc = run_waves(ctx, ctx->lanes[0], scene, cam, pr, 0, st, nullptr, &cancelled);
    if (rc) return rc;
    CameraBasis cb = camera_basis(*cam, pr->full_w, pr->full_h);
    int64_t P = (int64_t)pr->tile_w * pr->tile_h;
    k_output<<<blocks(P, 256), 256, 0, st>>>(ctx->film, out_rgba_dev, pr->tile_w, pr->tile_h, cb.sensitivity,
                                             cancelled ? 0 : 1, ctx->lanes[0].W.cnt + 7);
    HIP_OK(hipGetLastError());
    HIP_OK(hipEventRecord(ctx->ev1, st));
    return cancelled ? LUCID_ERR_CANCELLED : LUCID_OK;
}

int lucid_render_tile_device_noalloc(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                                     const lucid_render_params* pr, float* out_rgba_dev, void* stream) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!out_rgba_dev) return fail(ctx, LUCID_ERR_INVALID, "null output");
    if (pr->algorithm == LUCID_ALGO_GUIDED_MIS)
        return fail(ctx, LUCID_ERR_INVALID, "no-allocation renders do not train the guiding network");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->diag_on) return fail(ctx, LUCID_ERR_INVALID, "no-allocation renders do not record diagnostics");
    if (!ctx->scratch) return fail(ctx, LUCID_ERR_DEVICE, "no-allocation scratch unavailable on this device");
    HIP_OK(hipSetDevice(ctx->device));
    // layout, each piece 256-byte aligned: counters (cnt, rays), ray-bin
    // counters, the tile's film, then the wave's path arrays for as many
    // paths as fit (whole samples of the tile)
    const int64_t P = (int64_t)pr->tile_w * pr->tile_h;
    const bool vol = pr->algorithm == LUCID_ALGO_VOLUME_MIS;
    const int64_t per_path = kWaveBytesPerPath + (vol ? 48 : 0);
    const auto up = [](int64_t b) { return (b + 255) & ~(int64_t)255; };
    // the append segments hold up to kShards * kBlock entries more than the paths
    const int64_t head = up(kCounterBytes) + up((int64_t)kRayBins * 4) + up(P * 16) + 2 * up(kShards * kBlock * 4);
    const int64_t slack = 20 * 256;  // alignment of the path arrays
    const int64_t avail = (int64_t)kScratchBytes - head - slack;
    const int64_t cap = avail > 0 ? avail / per_path : 0;
    if (cap < P) {
        const int64_t max_px = ((int64_t)kScratchBytes - up(kCounterBytes) - up((int64_t)kRayBins * 4) -
                                2 * up(kShards * kBlock * 4) - slack) / (per_path + 16);
        return fail(ctx, LUCID_ERR_INVALID,
                    "tile of " + std::to_string(P) + " pixels exceeds the no-allocation scratch (" +
                        std::to_string(kScratchBytes >> 20) + " MiB: at most " + std::to_string(max_px) +
                        " pixels for this algorithm)");
    }
    unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
    int64_t off = 0;
    const auto take = [&](int64_t bytes) {
        void* q = base + off;
        off += up(bytes);
        return q;
    };
    FixedWave fw{};
    WaveDev& W = fw.W;
    W.cnt = static_cast<int32_t*>(take(kCounterBytes));
    W.rays = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(W.cnt) + 64);
    W.sbins = static_cast<uint32_t*>(take((int64_t)kRayBins * 4));
    fw.film = static_cast<float4*>(take(P * 16));
    W.ray_o = static_cast<float4*>(take(cap * 16));
    W.ray_d = static_cast<float4*>(take(cap * 16));
    W.thr = static_cast<float4*>(take(cap * 16));
    W.rad = static_cast<float4*>(take(cap * 16));
    W.rng = static_cast<ulonglong2*>(take(cap * 16));
    W.meta = static_cast<int4*>(take(cap * 16));
    W.hit = static_cast<float4*>(take(cap * 16));
    W.sh_o = static_cast<float4*>(take(cap * 16));
    W.sh_d = static_cast<float4*>(take(cap * 16));
    W.sh_c = static_cast<float4*>(take(cap * 16));
    W.q[0] = static_cast<int32_t*>(take(cap * 4));
    W.q[1] = static_cast<int32_t*>(take(cap * 4));
    W.qs = static_cast<int32_t*>(take(cap * 4));
    W.sq = static_cast<int32_t*>(take(cap * 4));
    W.skey = static_cast<uint32_t*>(take(cap * 4));
    W.seg_cap = segment_capacity(cap);
    W.aq = static_cast<int32_t*>(take(kShards * W.seg_cap * 4));
    W.aqs = static_cast<int32_t*>(take(kShards * W.seg_cap * 4));
    if (vol) {
        W.sh_a = static_cast<float4*>(take(cap * 16));
        W.sh_e = static_cast<float4*>(take(cap * 16));
        W.sh_t = static_cast<float4*>(take(cap * 16));
    }
    if (off > (int64_t)kScratchBytes) return fail(ctx, LUCID_ERR_INVALID, "no-allocation scratch layout overflow");
    fw.cap_paths = cap;
    // every buffer is written by a kernel before it is read (counters and
    // film are cleared by the render's first kernels); no per-launch events
    const int32_t prof = ctx->profiling;
    ctx->profiling = 0;
    ctx->cancel.store(0);
    __atomic_store_n(ctx->cancel_host, 0, __ATOMIC_SEQ_CST);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    bool cancelled = false;
    rc = run_waves(ctx, ctx->lanes[0], scene, cam, pr, 0, st, nullptr, &cancelled, nullptr, 1, nullptr, nullptr, false,
                   &fw);
    ctx->profiling = prof;
    if (rc) return rc;
    CameraBasis cb = camera_basis(*cam, pr->full_w, pr->full_h);
    k_output<<<blocks(P, 256), 256, 0, st>>>(fw.film, out_rgba_dev, pr->tile_w, pr->tile_h, cb.sensitivity,
                                             cancelled ? 0 : 1, W.cnt + 7);
    HIP_OK(hipGetLastError());
    return cancelled ? LUCID_ERR_CANCELLED : LUCID_OK;
}

int lucid_synchronize(lucid_ctx* ctx) {
    if (!ctx) return LUCID_ERR_INVALID;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return finish_stats(ctx);
}

int lucid_render_tile(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
                      float* out_rgba) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!out_rgba) return fail(ctx, LUCID_ERR_INVALID, "null output");
    int64_t n = (int64_t)pr->tile_w * pr->tile_h * 4;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        rc = ensure_buf(ctx, &ctx->out_dev, &ctx->out_cap, n);
        if (rc) return rc;
    }
    int rrc = lucid_render_tile_device(ctx, scene, cam, pr, ctx->out_dev, nullptr);
    if (rrc != LUCID_OK && rrc != LUCID_ERR_CANCELLED) return rrc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int32_t latch = 0;
    HIP_OK(hipMemcpyAsync(out_rgba, ctx->out_dev, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(&latch, ctx->lanes[0].W.cnt + 7, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (latch) rrc = LUCID_ERR_CANCELLED;  // the device saw the cancel: partial, unflipped
    rc = finish_stats(ctx);
    return rc ? rc : rrc;
}

int lucid_render_debug(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
                       int32_t mode, float* out_rgba) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!out_rgba) return fail(ctx, LUCID_ERR_INVALID, "null output");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    ctx->cancel.store(0);
    const int64_t P = (int64_t)pr->tile_w * pr->tile_h;
    Lane& L = ctx->lanes[0];
    rc = ensure_wave(ctx, L, P);
    if (!rc) rc = ensure_buf(ctx, &ctx->out_dev, &ctx->out_cap, P * 4);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    SceneDev SD = scene->dev;
    SD.any_vbo = 0;  // intersectScene itself
    WaveDev W = L.W;
    W.dcnt = nullptr;
    W.tslot = nullptr;
    W.tstride = 1;
    W.gout = nullptr;
    W.count_work = 0;
    W.stop = nullptr;
    CameraBasis cb = camera_basis(*cam, pr->full_w, pr->full_h);
    CamDev cd;
    std::memcpy(cd.pos, cb.pos, 12); std::memcpy(cd.fwd, cb.fwd, 12);
    std::memcpy(cd.right, cb.right, 12); std::memcpy(cd.up, cb.up, 12);
    cd.aspect = cb.aspect; cd.scale = cb.scale;
    cd.fw = pr->full_w; cd.fh = pr->full_h; cd.tx = pr->tile_x; cd.ty = pr->tile_y;
    cd.tw = pr->tile_w; cd.th = pr->tile_h;
    cd.samples_in_call = 1;  // no jitter
    cd.s0 = 0;
    cd.P = (int32_t)P;
    cd.B = 1;
    cd.max_depth = 1;
    cd.s0p = nullptr;
    const int sd = std::min(kStackMax, scene->host.max_depth + 2);
    const size_t lds = (size_t)sd * kBlock * sizeof(int32_t);
    const unsigned g = (unsigned)std::min<int64_t>(blocks(P, kBlock), (int64_t)ctx->num_cu * 16);
    k_zero_counters<<<1, 64, 0, st>>>(W.cnt, W.rays);
    k_raygen<<<blocks(P, kBlock), kBlock, 0, st>>>(W, cd, (int32_t)P);
    k_trace_refill<0><<<g, kBlock, lds, st>>>(SD, W, 0, sd, 48, ctx->walk_min, ctx->leaf_tris);
    if (mode == kDebugVolume) {
        k_debug_vol_prep<<<blocks(P, 256), 256, 0, st>>>(SD, W, (int32_t)P);
        k_zero_fetch<<<1, 64, 0, st>>>(W.cnt);  // fetch counters of the second query
        k_trace_refill<0><<<g, kBlock, lds, st>>>(SD, W, 1, sd, 48, ctx->walk_min, ctx->leaf_tris);
    }
    k_debug_out<<<blocks(P, 256), 256, 0, st>>>(SD, W, ctx->out_dev, pr->tile_w, pr->tile_h, mode);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out_rgba, ctx->out_dev, (size_t)P * 16, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    return LUCID_OK;
}

int lucid_render_samples(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
                         float* out_rgb) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!out_rgb) return fail(ctx, LUCID_ERR_INVALID, "null output");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    ctx->cancel.store(0);
    size_t n = (size_t)pr->tile_w * pr->tile_h * pr->samples * 3;
    float* dev = nullptr;
    if (n && hipMalloc(&dev, n * sizeof(float)) != hipSuccess) return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed");
    RunOut ro;
    ro.samples_dev = dev;
    bool cancelled = false;
    HIP_OK(hipEventRecord(ctx->ev0, ctx->stream));
    rc = run_waves(ctx, ctx->lanes[0], scene, cam, pr, 1, ctx->stream, &ro, &cancelled);
    if (!rc) {
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        if (n) (void)hipMemcpyAsync(out_rgb, dev, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, LUCID_ERR_DEVICE, "synchronize failed");
    }
    if (!rc) rc = finish_stats(ctx);
    if (dev) (void)hipFree(dev);
    if (!rc && cancelled) rc = LUCID_ERR_CANCELLED;
    return rc;
}

int lucid_primary_hits(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
                       int32_t* mesh, int32_t* tri, float* u, float* v, float* t) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!mesh || !tri || !u || !v || !t) return fail(ctx, LUCID_ERR_INVALID, "null output");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    ctx->cancel.store(0);
    size_t n = (size_t)pr->tile_w * pr->tile_h * pr->samples;
    RunOut ro;
    void* buf = nullptr;
    if (n && hipMalloc(&buf, n * 20) != hipSuccess) return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed");
    ro.mesh = (int32_t*)buf;
    ro.tri = ro.mesh + n;
    ro.u = (float*)(ro.tri + n);
    ro.v = ro.u + n;
    ro.t = ro.v + n;
    bool cancelled = false;
    HIP_OK(hipEventRecord(ctx->ev0, ctx->stream));
    rc = run_waves(ctx, ctx->lanes[0], scene, cam, pr, 2, ctx->stream, &ro, &cancelled);
    if (!rc) {
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        if (n) {
            (void)hipMemcpyAsync(mesh, ro.mesh, n * 4, hipMemcpyDeviceToHost, ctx->stream);
            (void)hipMemcpyAsync(tri, ro.tri, n * 4, hipMemcpyDeviceToHost, ctx->stream);
            (void)hipMemcpyAsync(u, ro.u, n * 4, hipMemcpyDeviceToHost, ctx->stream);
            (void)hipMemcpyAsync(v, ro.v, n * 4, hipMemcpyDeviceToHost, ctx->stream);
            (void)hipMemcpyAsync(t, ro.t, n * 4, hipMemcpyDeviceToHost, ctx->stream);
        }
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) rc = fail(ctx, LUCID_ERR_DEVICE, "synchronize failed");
    }
    if (!rc) rc = finish_stats(ctx);
    if (buf) (void)hipFree(buf);
    return rc;
}

void lucid_cancel(lucid_ctx* ctx) {
    if (!ctx) return;
    ctx->cancel.store(1);
    if (ctx->cancel_host) __atomic_store_n(ctx->cancel_host, 1, __ATOMIC_SEQ_CST);
}
int32_t lucid_is_cancelled(const lucid_ctx* ctx) { return ctx ? ctx->cancel.load() : 0; }
void lucid_reset_cancel(lucid_ctx* ctx) {
    if (!ctx) return;
    ctx->cancel.store(0);
    if (ctx->cancel_host) __atomic_store_n(ctx->cancel_host, 0, __ATOMIC_SEQ_CST);
}

int lucid_get_stats(const lucid_ctx* ctx, lucid_stats* out) {
    if (!ctx || !out) return LUCID_ERR_INVALID;
    *out = ctx->stats;
    return LUCID_OK;
}

int lucid_debug_host_trace(const lucid_mesh_desc* meshes, int32_t num_meshes, int32_t n_rays, const float* o,
                           const float* d, const float* tmin, const float* tmax, int32_t any_hit, int32_t* mesh,
                           int32_t* tri, float* u, float* v, float* t) {
    if (n_rays < 0 || (n_rays > 0 && (!o || !d || !tmin || !tmax || !mesh || !tri || !u || !v || !t)))
        return LUCID_ERR_INVALID;
    PackedScene ps;
    std::string err;
    const float env[3] = {0.0f, 0.0f, 0.0f};
    if (!pack_scene(meshes, num_meshes, nullptr, 0, env, 1.0f, ps, err)) return LUCID_ERR_INVALID;
    SceneDev S;
    std::memset(&S, 0, sizeof(S));
    S.nodes = ps.nodes.data();
    const std::vector<float> tri3 = interleave_triangles(ps);
    S.tri = reinterpret_cast<const float4*>(tri3.data());
    S.root = ps.root;
    const int sd = std::min(kStackMax, ps.max_depth + 2);
    std::vector<int32_t> stack((size_t)sd * kBlock);
    HostBounds hb{(int32_t)ps.nodes.size(), ps.num_triangles, false};
    int32_t ovf = 0;
    for (int32_t i = 0; i < n_rays; ++i) {
        uint32_t nn = 0, nt = 0;
        TraceHit h;
        const v3f ro = mk(o[3 * i], o[3 * i + 1], o[3 * i + 2]), rd = mk(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        bool hit = any_hit ? traverse<true>(S, ro, rd, tmin[i], tmax[i], stack.data(), sd, &ovf, nn, nt, h, &hb)
                           : traverse<false>(S, ro, rd, tmin[i], tmax[i], stack.data(), sd, &ovf, nn, nt, h, &hb);
        if (hb.violated || ovf) return LUCID_ERR_DEVICE;
        if (hit) {
            int32_t mi, ti;
            std::memcpy(&mi, &ps.tri_v0[4 * (size_t)h.prim + 3], 4);
            std::memcpy(&ti, &ps.tri_e1[4 * (size_t)h.prim + 3], 4);
            mesh[i] = mi; tri[i] = ti; u[i] = h.u; v[i] = h.v; t[i] = h.t;
        } else {
            mesh[i] = -1; tri[i] = -1; u[i] = 0.0f; v[i] = 0.0f; t[i] = 0.0f;
        }
    }
    return LUCID_OK;
}

int lucid_set_profiling(lucid_ctx* ctx, int32_t flags) {
    if (!ctx || flags < 0 || flags > 3) return LUCID_ERR_INVALID;
    ctx->profiling = flags;
    return LUCID_OK;
}

int lucid_set_grid_format(lucid_ctx* ctx, int32_t bits) {
    if (!ctx || (bits != 16 && bits != 32)) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->grid_bits = bits;
    return LUCID_OK;
}

int64_t lucid_scene_grid_bytes(const lucid_scene* s) { return s ? (int64_t)s->grid_bytes : 0; }

int lucid_set_scene_build(lucid_ctx* ctx, int32_t mode) {
    if (!ctx || (mode != 0 && mode != 1)) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->scene_build = mode;
    return LUCID_OK;
}

int lucid_scene_read_geometry(lucid_ctx* ctx, lucid_scene* s, void* nodes, float* tri, float* fn, float* vn,
                              float* uv, int64_t* n_nodes, int32_t* stack_depth, int32_t* built_on_device) {
    if (!ctx || !s) return fail(ctx, LUCID_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (n_nodes) *n_nodes = s->n_nodes;
    if (stack_depth) *stack_depth = s->host.max_depth;
    if (built_on_device) *built_on_device = s->built_on_device;
    if (!nodes) return LUCID_OK;
    HIP_OK(hipSetDevice(s->device));
    const size_t nt = (size_t)s->host.num_triangles;
    const SceneDev& D = s->dev;
    HIP_OK(hipDeviceSynchronize());
    if (s->n_nodes) HIP_OK(hipMemcpy(nodes, D.nodes, (size_t)s->n_nodes * sizeof(GNode), hipMemcpyDeviceToHost));
    if (nt && tri) HIP_OK(hipMemcpy(tri, D.tri, nt * 48, hipMemcpyDeviceToHost));
    if (nt && fn) HIP_OK(hipMemcpy(fn, D.fn, nt * 16, hipMemcpyDeviceToHost));
    if (vn) {
        if (nt && D.vn) HIP_OK(hipMemcpy(vn, D.vn, nt * 48, hipMemcpyDeviceToHost));
        else std::memset(vn, 0, nt * 48);
    }
    if (uv) {
        if (nt && D.tuv) HIP_OK(hipMemcpy(uv, D.tuv, nt * 24, hipMemcpyDeviceToHost));
        else std::memset(uv, 0, nt * 24);
    }
    return LUCID_OK;
}

int lucid_set_wave_paths(lucid_ctx* ctx, int64_t paths) {
    if (!ctx || paths < 0) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->wave_paths = paths;
    return LUCID_OK;
}

// Device time of one render of `pr` on lane 0 (output discarded).
static int timed_render(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
                        double* ms) {
    bool cancelled = false;
    HIP_OK(hipEventRecord(ctx->ev0, ctx->stream));
    int rc = run_waves(ctx, ctx->lanes[0], scene, cam, pr, 0, ctx->stream, nullptr, &cancelled);
    if (rc) return rc;
    HIP_OK(hipEventRecord(ctx->ev1, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    float e = 0.0f;
    HIP_OK(hipEventElapsedTime(&e, ctx->ev0, ctx->ev1));
    *ms = e;
    return cancelled ? fail(ctx, LUCID_ERR_CANCELLED, "autotune cancelled") : LUCID_OK;
}

int lucid_autotune(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* pr,
                   lucid_launch_shape* best) {
    int rc = validate(ctx, scene, cam, pr);
    if (rc) return rc;
    if (!best) return fail(ctx, LUCID_ERR_INVALID, "null output");
    if (ctx->diag_on) return fail(ctx, LUCID_ERR_INVALID, "autotune with diagnostics enabled");
    lucid_render_params tp = *pr;
    // guided renders would train the network: tune on the plain tracer
    if (tp.algorithm == LUCID_ALGO_GUIDED_MIS) tp.algorithm = LUCID_ALGO_MIS;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    ctx->cancel.store(0);
    const int32_t prof = ctx->profiling;
    ctx->profiling = 0;
    int32_t so = 0;  // shading order, searched after the ray order
    int32_t wk = ctx->walk_min;  // node-walk threshold, searched after the refill threshold
    int32_t lt = ctx->leaf_tris;  // triangles per leaf phase, searched after it
    auto measure = [&](int32_t ro, int32_t tg, int32_t rf, int32_t sg, double* ms) {
        ctx->walk_min = wk;
        ctx->leaf_tris = lt;
        ctx->ray_order = ro;
        ctx->traverse_groups = tg;
        ctx->refill = rf;
        ctx->shade_groups = sg;
        ctx->shade_order = so;
        double a = 0.0, b = 0.0;
        int r = timed_render(ctx, scene, cam, &tp, &a);  // first run: allocation, caches
        if (!r) r = timed_render(ctx, scene, cam, &tp, &b);
        *ms = std::min(a, b);
        return r;
    };
    // coordinate descent from the current shape over measured-good ranges
    int32_t ro = 0, tg = ctx->traverse_groups, rf = ctx->refill, sg = ctx->shade_groups;
    double t_best = 0.0;
    rc = measure(ro, tg, rf, sg, &t_best);
    for (int32_t v : {1, 3}) {  // every ray binned, shadow rays binned
        double t = 0.0;
        if (!rc && !(rc = measure(v, tg, rf, sg, &t)) && t < t_best) { t_best = t; ro = v; }
    }
    {
        double t = 0.0;
        so = 1;
        if (!rc && !(rc = measure(ro, tg, rf, sg, &t)) && t < t_best) t_best = t;
        else so = 0;
    }
    const int32_t tgs[] = {8, 12, 16, 24}, rfs[] = {32, 48, 60}, sgs[] = {4, 8, 16};
    for (int32_t v : tgs) {
        double t = 0.0;
        if (!rc && v != tg && !(rc = measure(ro, v, rf, sg, &t)) && t < t_best) { t_best = t; tg = v; }
    }
    for (int32_t v : rfs) {
        double t = 0.0;
        if (!rc && v != rf && !(rc = measure(ro, tg, v, sg, &t)) && t < t_best) { t_best = t; rf = v; }
    }
    {
        const int32_t wks[] = {0, 8, 16, 32};
        const int32_t first = wk;
        int32_t best_wk = wk;
        for (int32_t v : wks) {
            if (rc || v == first) continue;
            wk = v;
            double t = 0.0;
            if (!(rc = measure(ro, tg, rf, sg, &t)) && t < t_best) { t_best = t; best_wk = v; }
        }
        wk = best_wk;
        const int32_t lts[] = {1, 2, 4};
        const int32_t first_lt = lt;
        int32_t best_lt = lt;
        for (int32_t v : lts) {
            if (rc || v == first_lt) continue;
            lt = v;
            double t = 0.0;
            if (!(rc = measure(ro, tg, rf, sg, &t)) && t < t_best) { t_best = t; best_lt = v; }
        }
        lt = best_lt;
    }
    for (int32_t v : sgs) {
        double t = 0.0;
        if (!rc && v != sg && !(rc = measure(ro, tg, rf, v, &t)) && t < t_best) { t_best = t; sg = v; }
    }
    // wavefront batch size (the queue length every launch works through):
    // the automatic one wave, or the pass split into 2 or 4 waves of whole
    // samples, when those still hold >= 2^20 paths
    const int64_t P = (int64_t)tp.tile_w * tp.tile_h, total = P * tp.samples;
    const int64_t wave_saved = ctx->wave_paths;
    int64_t wp = wave_saved;
    for (int split : {2, 4}) {
        const int64_t per = ((tp.samples + split - 1) / split) * P;
        if (rc || per >= total || per < ((int64_t)1 << 20)) continue;
        ctx->wave_paths = per;
        double t = 0.0;
        if (!(rc = measure(ro, tg, rf, sg, &t)) && t < t_best) { t_best = t; wp = per; }
    }
    ctx->wave_paths = wp;
    ctx->walk_min = wk;
    ctx->leaf_tris = lt;
    ctx->ray_order = ro;
    ctx->traverse_groups = tg;
    ctx->refill = rf;
    ctx->shade_groups = sg;
    ctx->shade_order = so;
    ctx->profiling = prof;
    if (rc) return rc;
    rc = finish_stats(ctx);
    best->traverse_groups = tg;
    best->refill = rf;
    best->shade_groups = sg;
    best->ray_order = ro;
    best->shade_order = so;
    best->ms = t_best;
    best->wave_paths = wp;
    best->walk_min = wk;
    best->leaf_tris = lt;
    return rc;
}

int lucid_set_traversal_shape(lucid_ctx* ctx, int32_t walk_lanes, int32_t leaf_tris) {
    if (!ctx || walk_lanes < 0 || walk_lanes > 63 || leaf_tris < 1 || leaf_tris > 4) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->walk_min = walk_lanes;
    ctx->leaf_tris = leaf_tris;
    return LUCID_OK;
}

int lucid_set_refill_threshold(lucid_ctx* ctx, int32_t idle_lanes) {
    if (!ctx || idle_lanes < 1 || idle_lanes > 64) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->refill = idle_lanes;
    return LUCID_OK;
}

int lucid_set_session_fusion(lucid_ctx* ctx, int32_t enabled) {
    if (!ctx || enabled < 0 || enabled > 1) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->session_fusion = enabled;
    return LUCID_OK;
}

int lucid_set_session_graphs(lucid_ctx* ctx, int32_t on) {
    if (!ctx || on < 0 || on > 1) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->session_graphs = on;
    return LUCID_OK;
}

int lucid_set_ray_order(lucid_ctx* ctx, int32_t order) {
    if (!ctx || order < 0 || order > 3) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->ray_order = order;
    return LUCID_OK;
}

int lucid_set_shade_order(lucid_ctx* ctx, int32_t order) {
    if (!ctx || order < 0 || order > 1) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->shade_order = order;
    return LUCID_OK;
}

int lucid_set_launch_shape(lucid_ctx* ctx, int32_t traverse_groups_per_cu, int32_t shade_groups_per_cu) {
    if (!ctx || traverse_groups_per_cu < 0 || traverse_groups_per_cu > 64 || shade_groups_per_cu < 0 ||
        shade_groups_per_cu > 64)
        return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->traverse_groups = traverse_groups_per_cu;
    ctx->shade_groups = shade_groups_per_cu;
    return LUCID_OK;
}


int lucid_set_diagnostics(lucid_ctx* ctx, const lucid_diag_config* cfg) {
    if (!ctx) return LUCID_ERR_INVALID;
    if (cfg && cfg->enabled && cfg->subsample < 1) return fail(ctx, LUCID_ERR_INVALID, "subsample factor must be >= 1");
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->diag_on = (cfg && cfg->enabled) ? 1 : 0;
    ctx->diag_sub = (cfg && cfg->enabled) ? cfg->subsample : 1;
    // enable_diagnostics / disable_diagnostics drop the film
    // (pybind_renderer.hpp:893-908)
    free_diag_film(ctx);
    if (!ctx->diag_on) free_diag_wave(ctx);
    return LUCID_OK;
}

int lucid_diag_get_info(lucid_ctx* ctx, lucid_diag_info* out) {
    if (!ctx || !out) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(out, 0, sizeof(*out));
    out->pixel_bytes = (int64_t)sizeof(diag::PixelData);
    out->subsample = ctx->diag_sub;
    if (!ctx->diag_valid) return LUCID_OK;
    HIP_OK(hipSetDevice(ctx->device));
    unsigned long long tot = 0;
    HIP_OK(hipMemcpyAsync(&tot, ctx->dtotal, sizeof(tot), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    out->valid = 1;
    out->width = ctx->diag_tw;
    out->height = ctx->diag_th;
    out->sampled_width = ctx->diag_tw / ctx->diag_sub;
    out->sampled_height = ctx->diag_th / ctx->diag_sub;
    out->total_recorded = (int64_t)tot;
    return LUCID_OK;
}

namespace {
int diag_range(lucid_ctx* ctx, int64_t first, int64_t count) {
    if (!ctx->diag_valid) return fail(ctx, LUCID_ERR_INVALID, "no diagnostic film (render a tile with diagnostics on)");
    const int64_t n = (int64_t)(ctx->diag_tw / ctx->diag_sub) * (ctx->diag_th / ctx->diag_sub);
    if (first < 0 || count < 0 || first + count > n) return fail(ctx, LUCID_ERR_INVALID, "film pixel range out of bounds");
    return LUCID_OK;
}
}  // namespace

int lucid_diag_read_pixels(lucid_ctx* ctx, int64_t first, int64_t count, void* dst) {
    if (!ctx || (count > 0 && !dst)) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = diag_range(ctx, first, count);
    if (rc || count == 0) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipMemcpyAsync(dst, ctx->dfilm + first, (size_t)count * sizeof(diag::PixelData), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}

int lucid_diag_write_pixels(lucid_ctx* ctx, int64_t first, int64_t count, const void* src) {
    if (!ctx || (count > 0 && !src)) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = diag_range(ctx, first, count);
    if (rc || count == 0) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipMemcpyAsync(ctx->dfilm + first, src, (size_t)count * sizeof(diag::PixelData), hipMemcpyHostToDevice,
                          ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}

int lucid_diag_read_strategy(lucid_ctx* ctx, int64_t first, int64_t count, double* dst) {
    if (!ctx || (count > 0 && !dst)) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = diag_range(ctx, first, count);
    if (rc || count == 0) return rc;
    HIP_OK(hipSetDevice(ctx->device));
    HIP_OK(hipMemcpyAsync(dst, ctx->dstrat + first * 28, (size_t)count * 28 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}

int lucid_diag_clear(lucid_ctx* ctx) {
    if (!ctx) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->diag_valid) return LUCID_OK;
    HIP_OK(hipSetDevice(ctx->device));
    const int64_t px = (int64_t)(ctx->diag_tw / ctx->diag_sub) * (ctx->diag_th / ctx->diag_sub);
    const int64_t words = px * (int64_t)(sizeof(diag::PixelData) / 8), nst = px * 28;
    const int64_t nin = std::max<int64_t>(std::max(words, nst), 1);
    k_diag_init<<<blocks(nin, 256), 256, 0, ctx->stream>>>(reinterpret_cast<uint64_t*>(ctx->dfilm), ctx->dpattern,
                                                         words, ctx->dstrat, nst);
    k_zero_u64<<<1, 64, 0, ctx->stream>>>(ctx->dtotal);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}


int lucid_set_guiding(lucid_ctx* ctx, const lucid_guiding_config* cfg) {
    if (!ctx || !cfg) return LUCID_ERR_INVALID;
    if (!(cfg->alpha >= 0.0f && cfg->alpha <= 1.0f) || cfg->train_steps < 0 || cfg->batch < 16 ||
        cfg->batch > 65536 || !(cfg->learning_rate > 0.0f) || cfg->record_stride < 1)
        return fail(ctx, LUCID_ERR_INVALID, "guiding: alpha in [0,1], train_steps >= 0, batch in [16, 65536], "
                                            "learning_rate > 0, record_stride >= 1");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    ctx->gcfg = *cfg;
    int rc = ensure_guide(ctx);
    if (rc) return rc;
    guide::reset(ctx->gstate, ctx->gcfg.seed, ctx->gcfg.alpha, ctx->stream);
    ctx->gsteps = 0;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}

int lucid_guiding_reset(lucid_ctx* ctx) {
    if (!ctx) return LUCID_ERR_INVALID;
    lucid_guiding_config c = ctx->gcfg;
    return lucid_set_guiding(ctx, &c);
}

int lucid_guiding_get_stats(lucid_ctx* ctx, lucid_guiding_stats* out) {
    if (!ctx || !out) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::memset(out, 0, sizeof(*out));
    out->num_params = guide::kParams;
    out->train_steps = ctx->gsteps;
    out->active = ctx->gsteps > 0 ? 1 : 0;
    if (!ctx->gstate) return LUCID_OK;
    HIP_OK(hipSetDevice(ctx->device));
    unsigned long long tot = 0;
    HIP_OK(hipMemcpyAsync(&tot, guide::ring_total(ctx->gstate), 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    out->samples_total = (int64_t)tot;
    out->last_loss = guide::last_loss(ctx->gstate, ctx->stream);
    return LUCID_OK;
}

int lucid_guiding_get_params(lucid_ctx* ctx, float* params, int32_t n) {
    if (!ctx || !params || n != guide::kParams) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    int rc = ensure_guide(ctx);
    if (rc) return rc;
    return guide::get_params(ctx->gstate, params, ctx->stream) ? fail(ctx, LUCID_ERR_DEVICE, "guiding: copy failed")
                                                               : LUCID_OK;
}

int lucid_guiding_set_params(lucid_ctx* ctx, const float* params, int32_t n) {
    if (!ctx || !params || n != guide::kParams) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    int rc = ensure_guide(ctx);
    if (rc) return rc;
    if (guide::set_params(ctx->gstate, params, ctx->stream)) return fail(ctx, LUCID_ERR_DEVICE, "guiding: copy failed");
    ctx->gsteps = std::max<int64_t>(ctx->gsteps, 1);  // explicit weights count as trained
    return LUCID_OK;
}

int lucid_guiding_debug_infer(lucid_ctx* ctx, const float* params, const float* samples, int32_t n,
                              const float bmin[3], const float bmax[3], float* out) {
    if (!ctx || !params || !samples || !bmin || !bmax || !out || n < 1 || n > kGuideBatchCap) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    int rc = ensure_guide(ctx);
    if (rc) return rc;
    rc = guide::debug_infer(ctx->gstate, params, reinterpret_cast<const guide::Sample*>(samples), n, bmin, bmax, out,
                            ctx->stream);
    if (rc) return fail(ctx, rc == 1 ? LUCID_ERR_INVALID : LUCID_ERR_DEVICE, "guiding debug inference failed");
    guide::reset(ctx->gstate, ctx->gcfg.seed, ctx->gcfg.alpha, ctx->stream);
    ctx->gsteps = 0;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}

// Test hook of the render-path inference (k_guide_infer): n rays are traced
// by the closest-hit kernel, then the network runs over the queue `order`
// (m path indices, in that order) exactly as a guided bounce does.
int lucid_guiding_debug_render_infer(lucid_ctx* ctx, lucid_scene* scene, const float* params, int32_t n,
                                     const float* rays, const int32_t* order, int32_t m, float* out, float* frame,
                                     float* box) {
    if (!ctx || !scene || !params || !rays || !order || !out || !frame || !box || n < 1 || m < 0 || m > n ||
        n > (1 << 24))
        return LUCID_ERR_INVALID;
    for (int32_t i = 0; i < m; ++i)
        if (order[i] < 0 || order[i] >= n) return fail(ctx, LUCID_ERR_INVALID, "queue entry out of range");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    if (scene->device != ctx->device) return fail(ctx, LUCID_ERR_INVALID, "scene belongs to another device");
    int rc = ensure_guide(ctx);
    if (!rc) rc = ensure_wave(ctx, ctx->lanes[0], n);
    if (!rc) rc = ensure_buf(ctx, &ctx->gout, &ctx->gout_cap, ((int64_t)n + 1) * guide::kGoutStride);
    if (rc) return rc;
    if (guide::set_params(ctx->gstate, params, ctx->stream)) return fail(ctx, LUCID_ERR_DEVICE, "guiding: copy failed");
    ctx->gsteps = std::max<int64_t>(ctx->gsteps, 1);
    hipStream_t st = ctx->stream;
    WaveDev W = ctx->lanes[0].W;
    W.gout = ctx->gout;
    W.gspare = n;
    W.dcnt = nullptr;
    W.count_work = 0;
    std::vector<float4> o4((size_t)n), d4((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        o4[i] = make_float4(rays[6 * (size_t)i], rays[6 * (size_t)i + 1], rays[6 * (size_t)i + 2], 0.0f);
        d4[i] = make_float4(rays[6 * (size_t)i + 3], rays[6 * (size_t)i + 4], rays[6 * (size_t)i + 5], 0.0f);
    }
    HIP_OK(hipMemcpyAsync(W.ray_o, o4.data(), (size_t)n * 16, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(W.ray_d, d4.data(), (size_t)n * 16, hipMemcpyHostToDevice, st));
    k_zero_counters<<<1, 64, 0, st>>>(W.cnt, W.rays);
    // every path traced (queue 1 = 0..n-1), then the network over `order` (queue 0)
    std::vector<int32_t> all((size_t)n);
    for (int32_t i = 0; i < n; ++i) all[i] = i;
    HIP_OK(hipMemcpyAsync(W.q[1], all.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
    k_set_i32<<<1, 64, 0, st>>>(W.cnt + 1, n);
    const int sd = std::min(kStackMax, scene->host.max_depth + 2);
    const unsigned g = (unsigned)std::min<int64_t>(blocks(n, kBlock), (int64_t)ctx->num_cu * 16);
    k_trace_refill<0><<<g, kBlock, (size_t)sd * kBlock * sizeof(int32_t), st>>>(scene->dev, W, 1, sd, 48,
                                                                                  ctx->walk_min, ctx->leaf_tris);
    if (m > 0) HIP_OK(hipMemcpyAsync(W.q[0], order, (size_t)m * 4, hipMemcpyHostToDevice, st));
    k_set_i32<<<1, 64, 0, st>>>(W.cnt, m);
    // outputs start as NaN: entries the kernel must not write stay NaN
    HIP_OK(hipMemsetAsync(W.gout, 0xff, (size_t)n * guide::kGoutStride * sizeof(__half), st));
    const guide::GuideDev G = guide::dev(ctx->gstate, scene->bmin, scene->bmax, ctx->gcfg.alpha,
                                         ctx->gcfg.learn_selection != 0);
    const unsigned gg = (unsigned)std::min<int64_t>(blocks(std::max<int32_t>(m, 1), kInferWaves * 64),
                                                    (int64_t)ctx->num_cu * kInferGroupsPerCU);
    k_guide_infer<<<gg, kInferWaves * 64, 0, st>>>(scene->dev, W, G, 0);
    k_debug_frame<<<blocks(n, 256), 256, 0, st>>>(scene->dev, W, n);
    HIP_OK(hipGetLastError());
    std::vector<__half> h((size_t)n * guide::kGoutStride);
    std::vector<float4> hit((size_t)n);
    HIP_OK(hipMemcpyAsync(h.data(), W.gout, h.size() * sizeof(__half), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(hit.data(), W.hit, (size_t)n * 16, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(o4.data(), W.sh_o, (size_t)n * 16, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(d4.data(), W.sh_d, (size_t)n * 16, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    for (int32_t i = 0; i < n; ++i)
        for (int f = 0; f < guide::kOutTrained; ++f)
            out[(size_t)i * guide::kOutTrained + f] = __half2float(h[(size_t)i * guide::kGoutStride + f]);
    for (int32_t i = 0; i < n; ++i) {
        int32_t prim;
        std::memcpy(&prim, &hit[i].y, 4);
        const bool hit_any = prim >= 0;
        const float nan = std::nanf("");
        frame[6 * (size_t)i + 0] = hit_any ? o4[i].x : nan;
        frame[6 * (size_t)i + 1] = hit_any ? o4[i].y : nan;
        frame[6 * (size_t)i + 2] = hit_any ? o4[i].z : nan;
        frame[6 * (size_t)i + 3] = hit_any ? d4[i].x : nan;
        frame[6 * (size_t)i + 4] = hit_any ? d4[i].y : nan;
        frame[6 * (size_t)i + 5] = hit_any ? d4[i].z : nan;
    }
    for (int k = 0; k < 3; ++k) {
        box[k] = scene->bmin[k];
        box[3 + k] = scene->bmax[k];
    }
    return LUCID_OK;
}

int lucid_guiding_debug_step(lucid_ctx* ctx, const float* params, const float* samples, int32_t n,
                             const float bmin[3], const float bmax[3], float* out, float* grad, float* loss) {
    if (!ctx || !params || !samples || !bmin || !bmax || !out || !grad || !loss || n < 1 || n > kGuideBatchCap)
        return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    int rc = ensure_guide(ctx);
    if (rc) return rc;
    rc = guide::debug_step(ctx->gstate, params, reinterpret_cast<const guide::Sample*>(samples), n, bmin, bmax, out,
                           grad, loss, ctx->stream);
    if (rc) return fail(ctx, rc == 1 ? LUCID_ERR_INVALID : LUCID_ERR_DEVICE, "guiding debug step failed");
    // the network under training was overwritten: start afresh
    guide::reset(ctx->gstate, ctx->gcfg.seed, ctx->gcfg.alpha, ctx->stream);
    ctx->gsteps = 0;
    HIP_OK(hipStreamSynchronize(ctx->stream));
    return LUCID_OK;
}

}  // extern "C"

// ============================================================= sessions ==
// Progressive accumulation (the reference's async viewport accumulator,
// src/gpu/path_tracer.cpp:825-1296, and the sample_offset-chained
// render_tile calls of its render sessions) and batched passes over many
// sessions (one pass of every session, spread over the context's lanes).
namespace {

int session_clear(lucid_session* s) {
    lucid_ctx* ctx = s->ctx;
    const int64_t P = (int64_t)s->w * s->h;
    k_film_clear<<<blocks(P, 256), 256, 0, ctx->stream>>>(s->accum, (int32_t)P);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(ctx->stream));
    s->samples = 0;
    return LUCID_OK;
}

// Adds up to `spp` samples to the session on lane L (caller holds ctx->mu);
// the samples of every wave issued are counted, also when cancelled.
// What a captured session pass depends on besides its first sample index:
// a replay is valid only while every field is unchanged.
struct PassKey {
    const void* lane;
    hipStream_t stream;
    int64_t cap_paths, vcap, wave_paths;
    WaveDev W;
    const lucid_scene* scene;
    lucid_camera cam;
    const float4* accum;
    int32_t w, h, max_depth, algorithm, spp, active_lanes;
    int32_t traverse_groups, refill, shade_groups, ray_order, shade_order, num_cu, walk_min, leaf_tris;
};
std::vector<uint8_t> pass_key(const lucid_session* s, const Lane& L, int32_t spp, int active_lanes) {
    PassKey k;
    std::memset(&k, 0, sizeof(k));
    const lucid_ctx* ctx = s->ctx;
    k.lane = &L; k.stream = L.stream; k.cap_paths = L.cap_paths; k.vcap = L.vcap; k.wave_paths = ctx->wave_paths;
    k.W = L.W; k.scene = s->scene; k.cam = s->cam; k.accum = s->accum;
    k.w = s->w; k.h = s->h; k.max_depth = s->max_depth; k.algorithm = s->algorithm; k.spp = spp;
    k.active_lanes = active_lanes;
    k.traverse_groups = ctx->traverse_groups; k.refill = ctx->refill; k.shade_groups = ctx->shade_groups;
    k.ray_order = ctx->ray_order; k.shade_order = ctx->shade_order; k.num_cu = ctx->num_cu;
    k.walk_min = ctx->walk_min;
    k.leaf_tris = ctx->leaf_tris;
    const uint8_t* b = reinterpret_cast<const uint8_t*>(&k);
    return std::vector<uint8_t>(b, b + sizeof(k));
}

void drop_graph(lucid_session* s) {
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
    s->gseen = false;
    s->gsig.clear();
}

// A plain session pass (MIS, NEE, simple, volume) of one wave is launched as
// one graph: the first pass with a new key runs launch by launch, the second
// is captured, and later ones replay it with the new first sample index.
// The kernels are those of run_waves, so results are the same bit for bit.
int session_pass(lucid_session* s, Lane& L, int32_t spp, int active_lanes, const std::atomic<int>* cancel_flag,
                 bool* cancelled) {
    lucid_ctx* ctx = s->ctx;
    if (s->samples + spp > INT32_MAX) return fail(ctx, LUCID_ERR_INVALID, "session sample count overflows");
    lucid_render_params pr{0, 0, s->w, s->h, s->w, s->h, spp, (int32_t)s->samples, s->max_depth, s->algorithm};
    const int64_t P = (int64_t)s->w * s->h;
    // one wave for sure: the lane already holds P * spp paths (see run_waves' batch size)
    const bool graphable = ctx->session_graphs && s->algorithm != LUCID_ALGO_GUIDED_MIS && !ctx->diag_on &&
                           ctx->profiling == 0 && P * spp <= L.cap_paths && P * spp <= ((int64_t)1 << 26) &&
                           (ctx->wave_paths <= 0 || ctx->wave_paths >= P * spp);
    std::vector<uint8_t> key;
    if (graphable) {
        key = pass_key(s, L, spp, active_lanes);
        if (key != s->gsig) drop_graph(s);
    }
    if (graphable && s->gseen && !s->gexec) {
        if (!s->s0_dev) HIP_OK(hipMalloc((void**)&s->s0_dev, 64 * sizeof(int32_t)));
        if (cancel_flag->load(std::memory_order_relaxed)) {
            *cancelled = true;
            return LUCID_OK;
        }
        HIP_OK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeRelaxed));
        RunOut ro;
        const int rc = run_waves(ctx, L, s->scene, &s->cam, &pr, 3, L.stream, &ro, cancelled, s->accum,
                                 active_lanes, cancel_flag, s->s0_dev, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(L.stream, &g);
        hipGraphExec_t ex = nullptr;
        const bool ok = !rc && e == hipSuccess && g && !*cancelled && ro.samples_done == spp &&
                        hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        if (!ok) {
            (void)hipGetLastError();
            drop_graph(s);
            ctx->session_graphs = 0;  // capture failed: plain launches from now on
            if (rc) return rc;
            if (*cancelled) return LUCID_OK;
            return session_pass(s, L, spp, active_lanes, cancel_flag, cancelled);
        }
        s->gexec = ex;
        s->gsig = key;
        // run_waves counted this pass's paths while it was captured
        k_set_i32<<<1, 64, 0, L.stream>>>(s->s0_dev, (int32_t)s->samples);
        HIP_OK(hipGraphLaunch(s->gexec, L.stream));
        s->samples += spp;
        return LUCID_OK;
    }
    if (graphable && s->gexec) {
        if (cancel_flag->load(std::memory_order_relaxed)) {
            *cancelled = true;
            return LUCID_OK;
        }
        k_set_i32<<<1, 64, 0, L.stream>>>(s->s0_dev, (int32_t)s->samples);
        HIP_OK(hipGraphLaunch(s->gexec, L.stream));
        s->samples += spp;
        ctx->stats.paths += (uint64_t)P * spp;
        ctx->stats.waves += 1;
        return LUCID_OK;
    }
    RunOut ro;
    int rc = run_waves(ctx, L, s->scene, &s->cam, &pr, 3, L.stream, &ro, cancelled, s->accum, active_lanes,
                       cancel_flag);
    if (!rc) s->samples += ro.samples_done;
    if (!rc && graphable && !*cancelled) {
        s->gsig = key;
        s->gseen = true;
    }
    return rc;
}

// Averages the accumulator into the snapshot and bumps the revision.
int take_snapshot(lucid_session* s) {
    lucid_ctx* ctx = s->ctx;
    const int64_t P = (int64_t)s->w * s->h;
    int64_t n;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        n = s->samples;
        if (n > 0) {
            CameraBasis cb = camera_basis(s->cam, s->w, s->h);
            k_session_out<<<blocks(P, 256), 256, 0, ctx->stream>>>(s->accum, s->snap_dev, s->w, s->h,
                                                                  1.0f / (float)n, cb.sensitivity);
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(s->snap_pinned, s->snap_dev, P * 4 * sizeof(float), hipMemcpyDeviceToHost,
                                  ctx->stream));
            HIP_OK(hipStreamSynchronize(ctx->stream));
        }
    }
    {
        std::lock_guard<std::mutex> lock(s->smu);
        if (n > 0) s->snap.assign(s->snap_pinned, s->snap_pinned + P * 4);
        s->snap_samples = n;
    }
    s->revision.fetch_add(1, std::memory_order_release);
    return LUCID_OK;
}

// The async worker: passes of growing sample count (1, 2, 4, ... while a
// pass stays under ~20 ms, at most 256) with a snapshot after each, until
// stopped.  A reset clears the accumulator at the top of the next pass.
void session_worker(lucid_session* s) {
    lucid_ctx* ctx = s->ctx;
    int32_t spp = 1;
    while (!s->stop.load(std::memory_order_acquire)) {
        if (s->reset_pending.exchange(false, std::memory_order_acq_rel)) {
            {
                std::lock_guard<std::mutex> lock(s->pmu);
                s->cam = s->pend_cam;
                s->max_depth = s->pend_depth;
            }
            int rc;
            {
                std::lock_guard<std::mutex> lock(ctx->mu);
                (void)hipSetDevice(ctx->device);
                rc = session_clear(s);
            }
            if (rc) {
                s->err = ctx->err;
                break;
            }
            {
                std::lock_guard<std::mutex> lock(s->smu);
                s->snap_samples = 0;
            }
            spp = 1;
        }
        const auto t0 = std::chrono::steady_clock::now();
        int rc;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            (void)hipSetDevice(ctx->device);
            bool cancelled = false;
            rc = session_pass(s, ctx->lanes[0], spp, 1, &s->stop, &cancelled);
            if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
                rc = fail(ctx, LUCID_ERR_DEVICE, "async pass: synchronize failed");
        }
        if (rc) {
            s->err = ctx->err;
            s->failed.store(true, std::memory_order_release);
            break;
        }
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms < 20.0 && spp < 256) spp *= 2;
        else if (ms > 80.0 && spp > 1) spp /= 2;
        // a snapshot of the old camera is not published after a reset request
        if (!s->reset_pending.load(std::memory_order_acquire) && take_snapshot(s)) {
            s->err = ctx->err;
            s->failed.store(true, std::memory_order_release);
            break;
        }
    }
    if (!s->failed.load() && !s->reset_pending.load(std::memory_order_acquire)) (void)take_snapshot(s);
}

void session_stop(lucid_session* s) {
    if (!s->worker.joinable()) {
        s->running.store(false);
        return;
    }
    s->stop.store(1, std::memory_order_release);
    s->worker.join();
    s->running.store(false);
}

void free_arena(lucid_ctx* ctx) {
    if (!ctx->arena) return;
    for (void* q : ctx->arena->allocs) (void)hipFree(q);
    delete ctx->arena;
    ctx->arena = nullptr;
}

// A pass over these sessions can run fused: persistent traversal, one
// surface algorithm for all (MIS, NEE or simple), scenes with geometry and
// without participating media or volume-boundary meshes.
bool fusable(const lucid_ctx* ctx, lucid_session* const* ss, int32_t n) {
    if (!ctx->session_fusion || n < 2 || ctx->traverse_groups <= 0 || ctx->diag_on) return false;
    const int32_t algo = ss[0]->algorithm;
    if (algo != LUCID_ALGO_MIS && algo != LUCID_ALGO_NEE && algo != LUCID_ALGO_SIMPLE) return false;
    int64_t P = 0;
    for (int32_t i = 0; i < n; ++i) {
        const lucid_session* s = ss[i];
        if (s->algorithm != algo || s->scene->n_nodes <= 0 || s->scene->dev.any_vbo || s->scene->dev.n_vols)
            return false;
        P += (int64_t)s->w * s->h;
    }
    return P > 0 && P < ((int64_t)1 << 30);
}

// the scenes of `ss` merged (or the cached merge of the same scenes);
// scene_of[i] = the merged scene index of session i
int ensure_arena(lucid_ctx* ctx, lucid_session* const* ss, int32_t n, std::vector<int>& scene_of) {
    std::vector<uint64_t> key;
    std::vector<lucid_scene*> scenes;
    scene_of.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        lucid_scene* sc = ss[i]->scene;
        int k = 0;
        while (k < (int)key.size() && key[k] != sc->serial) ++k;
        if (k == (int)key.size()) {
            key.push_back(sc->serial);
            scenes.push_back(sc);
        }
        scene_of[i] = k;
    }
    if (ctx->arena && ctx->arena->key == key) return LUCID_OK;
    free_arena(ctx);
    FusedArena* A = new FusedArena();
    ctx->arena = A;
    A->key = key;
    const int ns = (int)scenes.size();
    int64_t N = 0, T = 0, M = 0, G = 0, Lt = 0, Nt = 0;
    bool any_vn = false, any_uv = false;
    for (lucid_scene* sc : scenes) {
        A->node_off.push_back((int32_t)N);
        A->tri_off.push_back((int32_t)T);
        A->mat_off.push_back((int32_t)M);
        A->prog_off.push_back((int32_t)G);
        A->light_off.push_back((int32_t)Lt);
        A->native_off.push_back((int32_t)Nt);
        N += sc->n_nodes;
        T += sc->host.num_triangles;
        M += (int64_t)sc->host.mats.size();
        G += (int64_t)sc->host.node_ops.size();
        Lt += (int64_t)sc->host.lights.size();
        Nt += (int64_t)sc->host.native_light_ids.size();
        any_vn = any_vn || sc->dev.vn;
        any_uv = any_uv || sc->dev.tuv;
        A->sd = std::max(A->sd, std::min(kStackMax, sc->host.max_depth + 2));
    }
    if (N + 1 >= INT32_MAX || T >= ((int64_t)1 << 28)) return fail(ctx, LUCID_ERR_INVALID, "fused pass: scenes too large");
    const auto get = [&](int64_t bytes, void** out) {
        *out = nullptr;
        if (bytes <= 0) return LUCID_OK;
        if (hipMalloc(out, (size_t)bytes) != hipSuccess) return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (fused scenes)");
        A->allocs.push_back(*out);
        return LUCID_OK;
    };
    void *nodes, *tri, *fn, *vn, *uv, *mats, *prog, *lights, *cdf, *nat;
    int rc = get(N * 64, &nodes);
    if (!rc) rc = get(T * 48, &tri);
    if (!rc) rc = get(T * 16, &fn);
    if (!rc) rc = get(any_vn ? T * 48 : 0, &vn);
    if (!rc) rc = get(any_uv ? T * 24 : 0, &uv);
    if (!rc) rc = get(M * (int64_t)sizeof(GMat), &mats);
    if (!rc) rc = get(G * (int64_t)sizeof(GNodeOp), &prog);
    if (!rc) rc = get(Lt * (int64_t)sizeof(GLight), &lights);
    if (!rc) rc = get(Lt * 4, &cdf);
    if (!rc) rc = get(Nt * 4, &nat);
    if (rc) {
        free_arena(ctx);
        return rc;
    }
    hipStream_t st = ctx->stream;
    const auto cp = [&](void* dst, int64_t off_bytes, const void* src, int64_t bytes) {
        if (bytes <= 0) return LUCID_OK;
        if (!src) return hipMemsetAsync(static_cast<char*>(dst) + off_bytes, 0, (size_t)bytes, st) == hipSuccess
                             ? LUCID_OK
                             : fail(ctx, LUCID_ERR_DEVICE, "fused pass: clear failed");
        return hipMemcpyAsync(static_cast<char*>(dst) + off_bytes, src, (size_t)bytes, hipMemcpyDeviceToDevice, st) ==
                       hipSuccess
                   ? LUCID_OK
                   : fail(ctx, LUCID_ERR_DEVICE, "fused pass: copy failed");
    };
    for (int k = 0; k < ns && !rc; ++k) {
        const lucid_scene* sc = scenes[k];
        const SceneDev& D = sc->dev;
        const int64_t t = sc->host.num_triangles;
        rc = cp(nodes, (int64_t)A->node_off[k] * 64, D.nodes, sc->n_nodes * 64);
        if (!rc) rc = cp(tri, (int64_t)A->tri_off[k] * 48, D.tri, t * 48);
        if (!rc) rc = cp(fn, (int64_t)A->tri_off[k] * 16, D.fn, t * 16);
        if (!rc && any_vn) rc = cp(vn, (int64_t)A->tri_off[k] * 48, D.vn, t * 48);
        if (!rc && any_uv) rc = cp(uv, (int64_t)A->tri_off[k] * 24, D.tuv, t * 24);
        if (!rc) rc = cp(mats, (int64_t)A->mat_off[k] * (int64_t)sizeof(GMat), D.mats,
                         (int64_t)sc->host.mats.size() * (int64_t)sizeof(GMat));
        if (!rc) rc = cp(prog, (int64_t)A->prog_off[k] * (int64_t)sizeof(GNodeOp), D.prog,
                         (int64_t)sc->host.node_ops.size() * (int64_t)sizeof(GNodeOp));
        if (!rc) rc = cp(lights, (int64_t)A->light_off[k] * (int64_t)sizeof(GLight), D.lights,
                         (int64_t)sc->host.lights.size() * (int64_t)sizeof(GLight));
        if (!rc) rc = cp(cdf, (int64_t)A->light_off[k] * 4, D.cdf, (int64_t)sc->host.lights.size() * 4);
        if (!rc) rc = cp(nat, (int64_t)A->native_off[k] * 4, D.native_ids, (int64_t)sc->host.native_light_ids.size() * 4);
        if (!rc && sc->n_nodes > 0)
            k_rebase_nodes<<<blocks(sc->n_nodes, 256), 256, 0, st>>>(static_cast<GNode*>(nodes) + A->node_off[k],
                                                                     sc->n_nodes, A->node_off[k], A->tri_off[k]);
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(ctx, LUCID_ERR_DEVICE, "fused pass: merge failed");
    if (rc) {
        free_arena(ctx);
        return rc;
    }
    SceneDev& D = A->dev;
    std::memset(&D, 0, sizeof(D));
    D.nodes = static_cast<const GNode*>(nodes);
    D.tri = static_cast<const float4*>(tri);
    D.fn = static_cast<const float4*>(fn);
    D.vn = static_cast<const float4*>(vn);
    D.tuv = static_cast<const float2*>(uv);
    D.mats = static_cast<const GMat*>(mats);
    D.prog = G > 0 ? static_cast<const GNodeOp*>(prog) : nullptr;
    D.lights = static_cast<const GLight*>(lights);
    D.cdf = static_cast<const float*>(cdf);
    D.native_ids = static_cast<const int32_t*>(nat);
    return LUCID_OK;
}

// One pass of `spp` samples over fusable sessions (see fusable): waves of
// whole samples of every session's pixels, each bounce one traversal, one
// shading and one shadow launch for all of them.
int fused_sessions_pass(lucid_ctx* ctx, lucid_session* const* ss, int32_t n, int32_t spp, bool* cancelled) {
    *cancelled = false;
    std::vector<int> scene_of;
    int rc = ensure_arena(ctx, ss, n, scene_of);
    if (rc) return rc;
    const FusedArena& A = *ctx->arena;
    std::vector<FusedSession> hs((size_t)n);
    int64_t P = 0;
    int32_t max_depth = 0;
    for (int32_t i = 0; i < n; ++i) {
        const lucid_session* s = ss[i];
        if (s->samples + spp > INT32_MAX) return fail(ctx, LUCID_ERR_INVALID, "session sample count overflows");
        FusedSession& f = hs[i];
        std::memset(&f, 0, sizeof(f));
        const CameraBasis cb = camera_basis(s->cam, s->w, s->h);
        std::memcpy(f.cam.pos, cb.pos, 12); std::memcpy(f.cam.fwd, cb.fwd, 12);
        std::memcpy(f.cam.right, cb.right, 12); std::memcpy(f.cam.up, cb.up, 12);
        f.cam.aspect = cb.aspect; f.cam.scale = cb.scale;
        f.cam.fw = s->w; f.cam.fh = s->h; f.cam.tx = 0; f.cam.ty = 0; f.cam.tw = s->w; f.cam.th = s->h;
        f.cam.samples_in_call = std::max(spp, 2);  // a session's samples are always jittered
        f.cam.s0 = (int32_t)s->samples;
        f.cam.P = s->w * s->h;
        f.cam.max_depth = s->max_depth;
        f.cam.s0p = nullptr;
        f.accum = s->accum;
        f.pix_off = P;
        const int k = scene_of[i];
        const SceneDev& D = s->scene->dev;
        f.root = A.node_off[k] + D.root;
        f.max_depth = s->max_depth;
        f.mat_off = A.mat_off[k]; f.prog_off = A.prog_off[k];
        f.light_off = A.light_off[k]; f.native_off = A.native_off[k];
        f.n_lights = D.n_lights; f.n_native = D.n_native; f.native_start = D.native_start;
        f.total_area = D.total_area;
        f.env[0] = D.env[0]; f.env[1] = D.env[1]; f.env[2] = D.env[2];
        P += (int64_t)s->w * s->h;
        max_depth = std::max(max_depth, s->max_depth);
    }
    rc = ensure_buf(ctx, &ctx->fses, &ctx->fses_cap, (int64_t)n);
    if (rc) return rc;
    Lane& L = ctx->lanes[0];
    hipStream_t st = L.stream;
    HIP_OK(hipMemcpyAsync(ctx->fses, hs.data(), hs.size() * sizeof(FusedSession), hipMemcpyHostToDevice, st));
    // waves of whole samples: as in run_waves, at most 2^26 paths and a
    // quarter of free memory
    int64_t target = ctx->wave_paths;
    if (target <= 0) {
        size_t free_b = 0, total_b = 0;
        target = (int64_t)1 << 26;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0) {
            const int64_t fit = (int64_t)(free_b / 4 / kWaveBytesPerPath) + L.cap_paths;
            target = std::max<int64_t>(std::min<int64_t>(target, fit), (int64_t)1 << 16);
        }
    }
    const int64_t B = std::max<int64_t>(1, std::min<int64_t>(spp, target / P));
    if (P * B >= INT32_MAX) return fail(ctx, LUCID_ERR_INVALID, "fused pass: wave too large");
    rc = ensure_wave(ctx, L, P * B);
    if (rc) return rc;
    WaveDev W = L.W;
    W.sh_a = W.sh_e = W.sh_t = nullptr;
    W.dv0 = W.dv1 = W.dr0 = W.dr1 = nullptr;
    W.dcnt = nullptr;
    W.dstride = 0;
    W.gout = nullptr;
    W.gspare = 0;
    W.tslot = nullptr;
    W.tstride = 1;
    W.tnrec = 0;
    W.stop = nullptr;
    W.count_work = (ctx->profiling & LUCID_PROFILE_WORK) ? 1 : 0;
    FusedDev F;
    F.ses = ctx->fses;
    F.n_ses = n;
    F.P = P;
    const SceneDev& SD = A.dev;
    const int sd = A.sd;
    const size_t lds = (size_t)sd * kBlock * sizeof(int32_t);
    const int64_t cap_trav = (int64_t)ctx->num_cu * ctx->traverse_groups;
    const int64_t cap_shade = ctx->shade_groups > 0 ? (int64_t)ctx->num_cu * ctx->shade_groups : INT64_MAX;
    const int32_t refill = ctx->refill, walk_min = ctx->walk_min, leaf_tris = ctx->leaf_tris;
    const int32_t algo = ss[0]->algorithm;
    int64_t done = 0;
    uint64_t waves = 0;
    for (int64_t s0 = 0; s0 < spp; s0 += B) {
        if (waves >= 2) (void)hipEventSynchronize(L.wave_ev[waves & 1]);
        if (ctx->cancel.load(std::memory_order_relaxed)) {
            *cancelled = true;
            break;
        }
        const int64_t b = std::min<int64_t>(B, spp - s0);
        const int64_t np = P * b;
        k_raygen_fused<<<blocks(np, kBlock), kBlock, 0, st>>>(W, F, (int32_t)s0, np);
        int q = 0;
        for (int depth = 0; depth < max_depth; ++depth) {
            const unsigned g = (unsigned)std::min<int64_t>(blocks(np, kBlock), cap_trav);
            const unsigned gs = (unsigned)std::min<int64_t>(blocks(np, kBlock), cap_shade);
            int ks = kev_begin(ctx, 0, st);
            k_trace_refill<0, true><<<g, kBlock, lds, st>>>(SD, W, q, sd, refill, walk_min, leaf_tris);
            kev_end(ctx, ks, st);
            ks = kev_begin(ctx, 1, st);
            if (algo == LUCID_ALGO_MIS)
                launch_shade_fused<LUCID_ALGO_MIS>(gs, st, SD, W, F, q);
            else if (algo == LUCID_ALGO_NEE)
                launch_shade_fused<LUCID_ALGO_NEE>(gs, st, SD, W, F, q);
            else
                launch_shade_fused<LUCID_ALGO_SIMPLE>(gs, st, SD, W, F, q);
            concat_queues(W, 1 - q, np, st);
            kev_end(ctx, ks, st);
            ks = kev_begin(ctx, 2, st);
            k_trace_refill<1, true><<<g, kBlock, lds, st>>>(SD, W, 0, sd, refill, walk_min, leaf_tris);
            kev_end(ctx, ks, st);
            q = 1 - q;
        }
        k_film_fused<<<blocks(P, 256), 256, 0, st>>>(W, F, (int32_t)b);
        (void)hipEventRecord(L.wave_ev[waves & 1], st);
        ++waves;
        done += b;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, LUCID_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    for (int32_t i = 0; i < n; ++i) ss[i]->samples += done;
    ctx->stats.paths += (uint64_t)P * (uint64_t)done;
    ctx->stats.waves += waves;
    return LUCID_OK;
}

}  // namespace

extern "C" {

int lucid_session_create(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, int32_t width, int32_t height,
                         int32_t max_depth, int32_t algorithm, lucid_session** out) {
    if (!ctx || !out) return LUCID_ERR_INVALID;
    *out = nullptr;
    if (!scene || !cam) return fail(ctx, LUCID_ERR_INVALID, "null scene or camera");
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1LL << 28))
        return fail(ctx, LUCID_ERR_INVALID, "session size must be positive and at most 2^28 pixels");
    if (max_depth < 0) return fail(ctx, LUCID_ERR_INVALID, "negative depth");
    if (algorithm < 0 || algorithm > LUCID_ALGO_GUIDED_MIS) return fail(ctx, LUCID_ERR_INVALID, "unknown algorithm");
    if (scene->device != ctx->device) return fail(ctx, LUCID_ERR_INVALID, "scene belongs to another device");
    lucid_session* s = new (std::nothrow) lucid_session();
    if (!s) return fail(ctx, LUCID_ERR_NOMEM, "out of host memory");
    s->ctx = ctx;
    s->cam = *cam;
    s->w = width;
    s->h = height;
    s->max_depth = max_depth;
    s->algorithm = algorithm;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = LUCID_OK;
    if (hipSetDevice(ctx->device) != hipSuccess ||
        hipMalloc((void**)&s->accum, (size_t)width * height * sizeof(float4)) != hipSuccess)
        rc = fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (session accumulator)");
    if (!rc) rc = session_clear(s);
    if (rc) {
        if (s->accum) (void)hipFree(s->accum);
        delete s;
        return rc;
    }
    lucid_scene_retain(scene);
    s->scene = scene;
    *out = s;
    return LUCID_OK;
}

void lucid_session_destroy(lucid_session* s) {
    if (!s) return;
    session_stop(s);
    lucid_ctx* ctx = s->ctx;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        if (s->accum) (void)hipFree(s->accum);
        if (s->snap_dev) (void)hipFree(s->snap_dev);
        if (s->snap_pinned) (void)hipHostFree(s->snap_pinned);
        drop_graph(s);
        if (s->s0_dev) (void)hipFree(s->s0_dev);
    }
    lucid_scene_release(s->scene);
    delete s;
}

int lucid_session_set_camera(lucid_session* s, const lucid_camera* cam, int32_t max_depth) {
    if (!s || !cam || max_depth < 0) return LUCID_ERR_INVALID;
    lucid_ctx* ctx = s->ctx;
    if (s->running.load()) return fail(ctx, LUCID_ERR_INVALID, "session is accumulating asynchronously");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    s->cam = *cam;
    s->max_depth = max_depth;
    return session_clear(s);
}

int lucid_session_clear(lucid_session* s) {
    if (!s) return LUCID_ERR_INVALID;
    lucid_ctx* ctx = s->ctx;
    if (s->running.load()) return fail(ctx, LUCID_ERR_INVALID, "session is accumulating asynchronously");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    return session_clear(s);
}

int64_t lucid_session_samples(const lucid_session* s) { return s ? s->samples : 0; }

int lucid_sessions_render(lucid_ctx* ctx, lucid_session* const* sessions, int32_t n, int32_t samples) {
    if (!ctx) return LUCID_ERR_INVALID;
    if (n < 0 || (n > 0 && !sessions)) return fail(ctx, LUCID_ERR_INVALID, "null session list");
    if (samples < 1) return fail(ctx, LUCID_ERR_INVALID, "samples must be positive");
    for (int32_t i = 0; i < n; ++i) {
        if (!sessions[i] || sessions[i]->ctx != ctx) return fail(ctx, LUCID_ERR_INVALID, "session of another context");
        if (sessions[i]->running.load())
            return fail(ctx, LUCID_ERR_INVALID, "session is accumulating asynchronously");
        for (int32_t j = 0; j < i; ++j)
            if (sessions[j] == sessions[i]) return fail(ctx, LUCID_ERR_INVALID, "session listed twice");
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    ctx->cancel.store(0);
    int rc = ensure_lanes(ctx);
    if (rc) return rc;
    // Longest-processing-time assignment of the sessions to lanes; guided
    // sessions share the context's network buffers and all run on lane 0.
    std::vector<int32_t> order(n);
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    auto work = [&](int32_t i) { return (int64_t)sessions[i]->w * sessions[i]->h; };
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return work(a) > work(b); });
    std::vector<int32_t> lane_of(n);
    int64_t load[kLanes] = {0, 0, 0, 0};
    for (int32_t i : order) {
        int l = 0;
        if (sessions[i]->algorithm != LUCID_ALGO_GUIDED_MIS)
            for (int k = 1; k < kLanes; ++k)
                if (load[k] < load[l]) l = k;
        lane_of[i] = l;
        load[l] += work(i);
    }
    int nl = 1;
    for (int k = 0; k < kLanes; ++k)
        if (load[k] > 0) nl = k + 1;
    ctx->stats = lucid_stats{};
    ctx->kev_used = 0;
    HIP_OK(hipEventRecord(ctx->ev0, ctx->stream));
    if (fusable(ctx, sessions, n)) {
        // the sessions' graphs describe lane passes: a fused pass makes them stale
        for (int32_t i = 0; i < n; ++i) drop_graph(sessions[i]);
        rc = ensure_counters(ctx, ctx->lanes[0]);
        if (rc) return rc;
        k_zero_counters<<<1, 64, 0, ctx->stream>>>(ctx->lanes[0].W.cnt, ctx->lanes[0].W.rays);
        bool cancelled = false;
        rc = fused_sessions_pass(ctx, sessions, n, samples, &cancelled);
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc)
            rc = fail(ctx, LUCID_ERR_DEVICE, "fused session pass: synchronize failed");
        if (rc) return rc;
        rc = finish_stats(ctx, 1);
        if (rc) return rc;
        return cancelled ? LUCID_ERR_CANCELLED : LUCID_OK;
    }
    for (int k = 0; k < nl; ++k) {
        Lane& L = ctx->lanes[k];
        rc = ensure_counters(ctx, L);
        if (rc) return rc;
        if (k > 0) HIP_OK(hipStreamWaitEvent(L.stream, ctx->ev0, 0));
        k_zero_counters<<<1, 64, 0, L.stream>>>(L.W.cnt, L.W.rays);
    }
    // issue round-robin over the lanes so that every lane starts early
    std::vector<std::vector<int32_t>> per_lane(kLanes);
    for (int32_t i : order) per_lane[lane_of[i]].push_back(i);
    bool cancelled = false;
    for (size_t r = 0; !rc && !cancelled; ++r) {
        bool any = false;
        for (int k = 0; k < nl && !rc && !cancelled; ++k) {
            if (r >= per_lane[k].size()) continue;
            any = true;
            rc = session_pass(sessions[per_lane[k][r]], ctx->lanes[k], samples, nl, &ctx->cancel, &cancelled);
        }
        if (!any) break;
    }
    for (int k = 1; k < nl; ++k) {
        (void)hipEventRecord(ctx->lanes[k].wave_ev[0], ctx->lanes[k].stream);
        (void)hipStreamWaitEvent(ctx->stream, ctx->lanes[k].wave_ev[0], 0);
    }
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc)
        rc = fail(ctx, LUCID_ERR_DEVICE, "session pass: synchronize failed");
    if (rc) return rc;
    rc = finish_stats(ctx, nl);
    if (rc) return rc;
    return cancelled ? LUCID_ERR_CANCELLED : LUCID_OK;
}

int lucid_session_read(lucid_session* s, float* out_rgba, int32_t average, int64_t* samples) {
    if (!s || !out_rgba) return LUCID_ERR_INVALID;
    lucid_ctx* ctx = s->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    const int64_t P = (int64_t)s->w * s->h;
    int rc = ensure_buf(ctx, &ctx->out_dev, &ctx->out_cap, P * 4);
    if (rc) return rc;
    const int64_t n = s->samples;
    CameraBasis cb = camera_basis(s->cam, s->w, s->h);
    const float inv = average ? (n > 0 ? 1.0f / (float)n : 0.0f) : 1.0f;
    k_session_out<<<blocks(P, 256), 256, 0, ctx->stream>>>(s->accum, ctx->out_dev, s->w, s->h, inv, cb.sensitivity);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out_rgba, ctx->out_dev, P * 4 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    if (samples) *samples = n;
    return LUCID_OK;
}

int lucid_session_start_async(lucid_session* s) {
    if (!s) return LUCID_ERR_INVALID;
    lucid_ctx* ctx = s->ctx;
    session_stop(s);
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        const size_t bytes = (size_t)s->w * s->h * 4 * sizeof(float);
        if (!s->snap_dev && hipMalloc((void**)&s->snap_dev, bytes) != hipSuccess)
            return fail(ctx, LUCID_ERR_NOMEM, "hipMalloc failed (session snapshot)");
        if (!s->snap_pinned && hipHostMalloc((void**)&s->snap_pinned, bytes, hipHostMallocDefault) != hipSuccess)
            return fail(ctx, LUCID_ERR_NOMEM, "hipHostMalloc failed (session snapshot)");
        int rc = session_clear(s);
        if (rc) return rc;
    }
    {
        std::lock_guard<std::mutex> lock(s->smu);
        s->snap_samples = 0;
    }
    s->err.clear();
    s->failed.store(false);
    s->reset_pending.store(false);
    s->stop.store(0);
    s->revision.store(0);
    s->running.store(true);
    try {
        s->worker = std::thread(session_worker, s);
    } catch (...) {
        s->running.store(false);
        return fail(ctx, LUCID_ERR_DEVICE, "cannot start the session worker thread");
    }
    return LUCID_OK;
}

int lucid_session_stop_async(lucid_session* s) {
    if (!s) return LUCID_ERR_INVALID;
    session_stop(s);
    return LUCID_OK;
}

int lucid_session_reset_async(lucid_session* s, const lucid_camera* cam, int32_t width, int32_t height,
                              int32_t max_depth) {
    if (!s || !cam || max_depth < 0) return LUCID_ERR_INVALID;
    if (!s->running.load()) return LUCID_OK;
    if (width != s->w || height != s->h) {
        std::lock_guard<std::mutex> lock(s->ctx->mu);  // the worker may be writing the error text
        return fail(s->ctx, LUCID_ERR_INVALID, "reset with other dimensions: stop and start the session instead");
    }
    {
        std::lock_guard<std::mutex> lock(s->pmu);
        s->pend_cam = *cam;
        s->pend_depth = max_depth;
    }
    s->reset_pending.store(true, std::memory_order_release);
    return LUCID_OK;
}

int lucid_session_poll_async(lucid_session* s, float* out_rgba, int64_t* samples, uint32_t* revision) {
    if (!s) return LUCID_ERR_INVALID;
    std::lock_guard<std::mutex> lock(s->smu);
    const int64_t n = s->running.load() ? s->snap_samples : 0;
    if (samples) *samples = n;
    if (revision) *revision = s->revision.load(std::memory_order_acquire);
    if (out_rgba && n > 0) std::memcpy(out_rgba, s->snap.data(), s->snap.size() * sizeof(float));
    return LUCID_OK;
}

int32_t lucid_session_async_running(const lucid_session* s) { return s && s->running.load() ? 1 : 0; }

const char* lucid_session_error(const lucid_session* s) {
    if (!s) return "null session";
    return s->failed.load(std::memory_order_acquire) ? s->err.c_str() : "";
}

}  // extern "C"
