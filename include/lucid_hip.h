/*
 * lucid_hip.h -- C ABI of the MI355X path-tracing library (liblucid_hip.so).
 *
 * Plain pointers, sizes and POD structs only.  The host owns every host
 * buffer passed in or out; the library owns device memory.  No exception
 * crosses this boundary: every call returns a status code and the last error
 * text of a context is available from lucid_last_error().
 *
 * Reference interfaces each entry point replaces (paths relative to
 * LucidRenderer/cpp_renderer/ of daikiad/blender-lucid-renderer):
 *   lucid_scene_create   PyRenderer::loadSceneFromJsonString, the mesh /
 *                        native-light / environment conversion
 *                        (include/pybind_renderer.hpp:1258-1746) plus
 *                        finalizeMeshBounds + BVH::build
 *                        (include/core/scene.hpp:65-227) and the GPU scene
 *                        packer pack_scene_for_path_tracer
 *                        (src/gpu/path_tracer.cpp:162-469)
 *   lucid_render_tile    PyRenderer::render_tile
 *                        (include/pybind_renderer.hpp:185-361) and the Dawn
 *                        dispatch PathTracer::render (src/gpu/path_tracer.cpp:591-822)
 *   lucid_cancel & co.   PyRenderer::cancel / is_cancelled / reset_cancel
 *                        (include/pybind_renderer.hpp:86-102)
 *   lucid_primary_hits   intersectScene on the camera rays of render_tile
 *                        (include/geometry/bvh.hpp:155-177), exposed for the
 *                        bit-exact primary-hit check
 *   lucid_render_samples traceMIS & co. per sample (include/integrator/
 *                        path_tracer.hpp:128-835), per-sample radiance out
 *   lucid_material_eval  getMaterialParams / getEmission with the node-tree
 *                        evaluator (include/integrator/path_tracer.hpp:45-72,
 *                        src/node_evaluator.cpp:168-740), on the host
 */
#ifndef LUCID_HIP_H
#define LUCID_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LUCID_OK 0
#define LUCID_ERR_INVALID 1   /* bad argument */
#define LUCID_ERR_DEVICE 2    /* HIP runtime / device failure */
#define LUCID_ERR_CANCELLED 3 /* render stopped by lucid_cancel */
#define LUCID_ERR_NOMEM 4     /* device allocation failed */

/* integrators: the algorithm strings of PyRenderer::set_algorithm
   (include/pybind_renderer.hpp:301-327); unknown strings mean NEE */
#define LUCID_ALGO_SIMPLE 0
#define LUCID_ALGO_NEE 1
#define LUCID_ALGO_MIS 2
#define LUCID_ALGO_VOLUME_SIMPLE 3
#define LUCID_ALGO_VOLUME_MIS 4
/* Extension, no counterpart in the reference: traceMIS with learned path
   guiding at opaque non-delta surfaces (lucid_set_guiding). */
#define LUCID_ALGO_GUIDED_MIS 5

/* native light types (include/light/light.hpp:22-28) and area shapes (:30-35) */
#define LUCID_LIGHT_POINT 1
#define LUCID_LIGHT_SUN 2
#define LUCID_LIGHT_SPOT 3
#define LUCID_LIGHT_AREA 4
#define LUCID_SHAPE_SQUARE 0
#define LUCID_SHAPE_RECTANGLE 1
#define LUCID_SHAPE_DISK 2
#define LUCID_SHAPE_ELLIPSE 3

typedef struct lucid_ctx lucid_ctx;
typedef struct lucid_scene lucid_scene;

/* Node-tree socket values (include/core/material.hpp:36-70, parsed by
   parseSocketValue, src/node_evaluator.cpp:47-69): a JSON number is FLOAT, a
   3-array VEC3, a 4-array VEC4; booleans, strings, other arrays and absent
   values are NONE (the evaluator reads none of them). */
#define LUCID_SOCK_NONE 0
#define LUCID_SOCK_FLOAT 1
#define LUCID_SOCK_VEC3 2
#define LUCID_SOCK_VEC4 3

/* One socket of a node (NodeSocket, include/core/material.hpp:75-84). The
   caller resolves the tree's "links" as parseNodeTree does
   (src/node_evaluator.cpp:87-96, the last link to a socket wins): an input
   linked from (node, socket) carries both names, an unlinked one NULL. */
typedef struct lucid_node_socket {
    const char* name;
    int32_t kind;              /* LUCID_SOCK_* */
    float value[4];            /* FLOAT: value[0]; VEC3: [0..2]; VEC4: [0..3] */
    const char* linked_node;   /* NULL when not linked */
    const char* linked_socket;
} lucid_node_socket;

/* One node of a material's "node_tree" (MaterialNode, material.hpp:89-105). */
typedef struct lucid_node {
    const char* name;
    const char* type;          /* Blender bl_idname, e.g. "ShaderNodeBsdfPrincipled" */
    const lucid_node_socket* inputs;
    int32_t num_inputs;
    const lucid_node_socket* outputs;
    int32_t num_outputs;
} lucid_node;

/* One mesh of the scene JSON "meshes" array (pybind_renderer.hpp:1266-1507). */
typedef struct lucid_mesh_desc {
    int32_t num_vertices;
    const float* vertices;    /* 3 floats per vertex */
    int32_t num_triangles;
    const int32_t* indices;   /* 3 per triangle */
    const float* tri_normals; /* 9 floats per triangle ("triangle_normals") or NULL */
    int32_t num_tri_normals;  /* triangles [0, num_tri_normals) carry normals */
    int32_t smooth;           /* "smooth" */
    float base_color[3];
    float metallic, roughness;
    float emission[3];
    float transmission, ior;
    int32_t has_volume;       /* "volume" object present */
    float vol_color[3], vol_absorption[3];
    float vol_density, vol_anisotropy;
    int32_t grid_dims[3];     /* heterogeneous grid, x fastest */
    float grid_min[3], grid_max[3];
    const float* grid;        /* dims product floats, or NULL */
    int32_t surface_linked;   /* Material Output Surface linked (0 + volume => boundary-only) */
    /* Node-tree material (pybind_renderer.hpp:1381-1384): when use_nodes is
       set and num_nodes > 0 the tree replaces base_color, metallic,
       roughness, transmission, ior and emission as getMaterialParams /
       getEmission do (include/integrator/path_tracer.hpp:45-72). */
    int32_t use_nodes;        /* "use_nodes" */
    const lucid_node* nodes;  /* "node_tree" "nodes", in JSON order */
    int32_t num_nodes;
    const float* tri_uvs;     /* "triangle_uvs": 6 floats (u0 v0 u1 v1 u2 v2) per triangle, or NULL */
    int32_t num_tri_uvs;      /* triangles [0, num_tri_uvs) carry UVs; the rest read uv (0, 0) */
} lucid_mesh_desc;

/* One entry of "native_lights" (pybind_renderer.hpp:1544-1746); the library
   applies the reference's watt -> radiance conversions. */
typedef struct lucid_light_desc {
    int32_t type;             /* LUCID_LIGHT_* */
    float position[3];
    int32_t has_direction;
    float direction[3];
    float color[3];
    float energy;
    float radius;
    float spot_size;
    float spot_blend;
    float size, size_y;
    int32_t shape;            /* LUCID_SHAPE_* */
    int32_t has_right, has_up;
    float right[3], up[3];
} lucid_light_desc;

/* set_camera arguments (pybind_renderer.hpp:131-160) plus the film exposure. */
typedef struct lucid_camera {
    float pos[3];
    float dir[3];
    float up[3];
    float fov_deg;
    float exposure_ev;
} lucid_camera;

/* render_tile arguments (pybind_renderer.hpp:185-191). */
typedef struct lucid_render_params {
    int32_t tile_x, tile_y, tile_w, tile_h;
    int32_t full_w, full_h;
    int32_t samples;
    int32_t sample_offset;
    int32_t max_depth;
    int32_t algorithm;        /* LUCID_ALGO_* */
} lucid_render_params;

/* Counters of the last render on a context. */
typedef struct lucid_stats {
    uint64_t rays_closest;    /* camera + extension closest-hit queries */
    uint64_t rays_shadow;     /* shadow (any-hit) queries */
    uint64_t paths;           /* camera paths traced */
    uint64_t waves;           /* wavefront batches */
    uint64_t stack_overflows; /* traversal stack overflows (must stay 0) */
    double ms_render;         /* device time of the last render (HIP events) */
    /* with LUCID_PROFILE_WORK: BVH nodes (64 B) and triangles (48 B) fetched */
    uint64_t nodes_closest, tris_closest, nodes_shadow, tris_shadow;
    /* with LUCID_PROFILE_KERNELS: summed HIP-event time and launch count of
       the intersect, shade and shadow kernels */
    double ms_intersect, ms_shade, ms_shadow;
    uint64_t launches_intersect, launches_shade, launches_shadow;
    /* with LUCID_PROFILE_KERNELS, guided renders: guiding-network inference
       and the training after the render */
    double ms_guide_infer, ms_guide_train;
    /* with LUCID_PROFILE_KERNELS: regrouping the ray queues (lucid_set_ray_order) */
    double ms_ray_binning;
} lucid_stats;

#define LUCID_PROFILE_WORK 1    /* count node / triangle fetches */
#define LUCID_PROFILE_KERNELS 2 /* time every kernel launch with HIP events */

/* Context: one device, one in-flight render. */
int lucid_create(int32_t device, lucid_ctx** out);
void lucid_destroy(lucid_ctx* ctx);
const char* lucid_last_error(const lucid_ctx* ctx);

/* Scene: packed and uploaded once; refcounted (create holds one reference). */
int lucid_scene_create(lucid_ctx* ctx, const lucid_mesh_desc* meshes, int32_t num_meshes,
                       const lucid_light_desc* lights, int32_t num_lights, const float env_color[3],
                       float env_strength, lucid_scene** out);
void lucid_scene_retain(lucid_scene* scene);
void lucid_scene_release(lucid_scene* scene);
int32_t lucid_scene_num_triangles(const lucid_scene* scene);

/* render_tile: writes tile_w*tile_h*4 floats to out_rgba (host memory):
   RGBA, bottom row first, RGB = max(0, 2^EV * sum over `samples` of the
   path radiance), alpha = 1.  On LUCID_ERR_CANCELLED the partial sums are
   written top row first, as the reference's early return does. */
int lucid_render_tile(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                      const lucid_render_params* params, float* out_rgba);

/* Same, with the output left in device memory (out_rgba_dev, tile_w*tile_h*4
   floats) and no host copy; `stream` is a hipStream_t or NULL for the
   library's own stream.  Returns after the work is queued. */
int lucid_render_tile_device(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                             const lucid_render_params* params, float* out_rgba_dev, void* stream);
int lucid_synchronize(lucid_ctx* ctx);

/* No counterpart in the reference.  lucid_render_tile_device for callers
   that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc): it calls no hipMalloc,
   hipFree or allocating library routine.  Its counters, film and wave
   buffers live in a 256 MiB device array of the library that every GPU the
   library runs on holds until the process ends; the scene, the camera and
   the output are the caller's.  Same output as lucid_render_tile_device,
   bit for bit.  Limits: tile_w * tile_h at most about 1.31M pixels (1.06M
   for LUCID_ALGO_VOLUME_MIS), beyond which LUCID_ERR_INVALID says so; not
   for LUCID_ALGO_GUIDED_MIS or with diagnostics enabled.  All calls on one
   GPU share the array: calls that overlap on one GPU (other contexts,
   other streams) race. */
int lucid_render_tile_device_noalloc(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                                     const lucid_render_params* params, float* out_rgba_dev, void* stream);

/* Per-sample radiance (before exposure and clamp) of every pixel and sample:
   out_rgb[((py*tile_w + px)*samples + s)*3 + c], tile order (top row first). */
int lucid_render_samples(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                         const lucid_render_params* params, float* out_rgb);

/* Debug render of a tile (PyRenderer::render_debug,
   include/pybind_renderer.hpp:371-439, and render_debug_gpu :449-520): one
   unjittered camera ray per pixel, its closest hit by intersectScene from
   t = 0, shown per `mode` as
     LUCID_DEBUG_NORMAL   shading normal * 0.5 + 0.5 (traceNormal,
                          include/renderer.hpp:58-65; a miss gives 0.5),
     LUCID_DEBUG_ALBEDO   the material's albedo at the hit (traceAlbedo :67-78),
     LUCID_DEBUG_EMISSION its emission (traceEmission :80-93),
     LUCID_DEBUG_VOLUME   a medium's thickness map (traceVolume :109-190);
   any other mode paints every pixel magenta.  tile_w * tile_h * 4 floats,
   RGBA, bottom row first, alpha 1.  params' samples, offset, depth and
   algorithm are ignored. */
#define LUCID_DEBUG_NORMAL 0
#define LUCID_DEBUG_ALBEDO 1
#define LUCID_DEBUG_EMISSION 2
#define LUCID_DEBUG_VOLUME 3
int lucid_render_debug(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                       const lucid_render_params* params, int32_t mode, float* out_rgba);

/* Closest hit of every (jittered) camera ray of a tile: mesh index, original
   triangle index, raw barycentric u, v and t (t = 0, ids = -1 on a miss). */
int lucid_primary_hits(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam,
                       const lucid_render_params* params, int32_t* mesh, int32_t* tri, float* u, float* v,
                       float* t);

/* getMaterialParams + getEmission (include/integrator/path_tracer.hpp:45-72,
   with the node tree of src/node_evaluator.cpp:168-740) of one mesh's
   material at `count` texture coordinates (uv[2i], uv[2i+1]), through the
   same compiled node programs the device runs.  Writes 10 floats per
   coordinate: albedo RGB, emission RGB, metallic, roughness (clamped to
   >= 0.01), transmission, ior.  Host only: no device is needed.  On
   LUCID_ERR_INVALID the reason is written to err (err_len bytes). */
int lucid_material_eval(const lucid_mesh_desc* mesh, int32_t count, const float* uv, float* out, char* err,
                        int32_t err_len);

/* Cooperative cancel, callable from any thread (PyRenderer::cancel,
   include/pybind_renderer.hpp:225-245, which polls per 8x8 micro-tile).
   The host polls it between wavefront batches; render_tile's kernels read it
   from mapped host memory once per bounce, and a set flag ends every path at
   its next shading, so a cancel lands within one bounce.  The film then
   keeps the batches completed before the cancel was seen, and the rows are
   left unflipped (top row first), as the reference's early return leaves
   them.  render entry points clear the flag on entry. */
void lucid_cancel(lucid_ctx* ctx);
int32_t lucid_is_cancelled(const lucid_ctx* ctx);
void lucid_reset_cancel(lucid_ctx* ctx);

int lucid_get_stats(const lucid_ctx* ctx, lucid_stats* out);

/* Paths per wavefront batch (0 = automatic: up to 2^26 paths, bounded by a
   quarter of free device memory at 188 B per path). */
int lucid_set_wave_paths(lucid_ctx* ctx, int64_t paths);

/* Launch shape of the traversal (closest and shadow) and shading kernels.
   Traversal: N > 0 workgroups per compute unit run persistently, each lane
   tracing one ray at a time and wavefronts fetching new rays for idle lanes
   (see lucid_set_refill_threshold); 0 launches one lane per queued ray.
   Shading: persistent grid-stride over the queue with N workgroups per CU,
   or 0 for one lane per path.  Defaults: 16 and 8.  A tuning knob with no
   counterpart in the reference (its Dawn backend fixes workgroup counts per
   dispatch). */
int lucid_set_launch_shape(lucid_ctx* ctx, int32_t traverse_groups_per_cu, int32_t shade_groups_per_cu);
/* Autotuner (no counterpart in the reference): times launch shapes on
   renders of `params` (discarded; guided requests are tuned on plain MIS)
   by coordinate descent -- the ray order (0, 1 and 3), traversal groups
   per CU in {8, 12, 16, 24}, then
   the refill threshold in {32, 48, 60}, then the node-walk threshold in
   {0, 8, 16, 32} and the triangles per leaf phase in {1, 2, 4}, then
   shading groups per CU in {4, 8, 16}, then the wavefront batch (one wave, or the pass split into 2
   or 4 waves of whole samples, each of at least 2^20 paths:
   lucid_set_wave_paths), two renders each -- and keeps the fastest on the
   context.  Launch shapes do not change results.  Not with diagnostics
   enabled. */
typedef struct lucid_launch_shape {
    int32_t traverse_groups, refill, shade_groups;
    int32_t ray_order;  /* lucid_set_ray_order: 0, 1 or 3 */
    int32_t shade_order;  /* lucid_set_shade_order: 0 or 1 */
    double ms;          /* device time of the chosen shape's render */
    int64_t wave_paths; /* lucid_set_wave_paths: 0 automatic, or paths per batch */
    int32_t walk_min;   /* lucid_set_traversal_shape: walk_lanes */
    int32_t leaf_tris;  /* lucid_set_traversal_shape: leaf_tris */
} lucid_launch_shape;
int lucid_autotune(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, const lucid_render_params* params,
                   lucid_launch_shape* best);

/* Shading order (no counterpart in the reference; the regrouping by
   material the reference's megakernel cannot do): 1 regroups each bounce's
   live queue by the shading class of its hits (miss, emitter, dielectric,
   metal, partly transmissive, glass, node-tree material) with wavefront
   ballots and prefix sums before the shading launch, so that a wavefront's
   paths take the same branches; 0 (default) shades in compaction order.
   Results do not depend on it.  The autotuner measures both. */
int lucid_set_shade_order(lucid_ctx* ctx, int32_t order);

/* Session passes as graphs (no counterpart in the reference): 1 (default)
   launches a plain (not guided) one-wave session pass as one captured HIP
   graph once the same pass (session, lane, size, samples per pass, launch
   shape) has run twice, replaying it with the new first sample index; 0
   launches every kernel on its own.  Results do not depend on it. */
int lucid_set_session_graphs(lucid_ctx* ctx, int32_t on);

/* Fused session passes (no counterpart in the reference, which renders one
   session at a time behind gpu_mutex_, include/pybind_renderer.hpp:1118):
   with 1 (default) a lucid_sessions_render call whose sessions all use the
   same surface algorithm (MIS, NEE or simple) on scenes with geometry and
   without participating media or volume-boundary meshes runs as one fused
   pass: the sessions' scenes merged into one set of device arrays (cached
   for the next pass over the same scenes), every bounce one traversal, one
   shading and one shadow launch over the paths of all sessions.  Other
   batches, and every batch with 0, spread their sessions over the stream
   lanes.  Results do not depend on it: each session's sums equal those of
   its own render_tile bit for bit. */
int lucid_set_session_fusion(lucid_ctx* ctx, int32_t enabled);

/* Ray order (no counterpart in the reference): 1 regroups, from the second
   bounce on, the bounce rays of each depth by direction octant and origin
   cell (8x8x8 over the scene box) before the closest-hit pass, and the
   shadow rays by origin cell (16x16x16) before the shadow pass; 0 keeps the
   order in which shading queued them; 2 (default) regroups in waves of at
   least 2^24 paths only; 3 regroups the shadow rays only.  Results do not
   depend on it. */
int lucid_set_ray_order(lucid_ctx* ctx, int32_t order);

/* Persistent traversal (no counterpart in the reference): a wavefront walks
   BVH nodes while more than `walk_lanes` of its lanes still look for a leaf
   (0..63; default 16), then tests up to `leaf_tris` triangles (1..4; default
   4) of the leaf each lane holds; a lane keeps the rest of its leaf for the
   next leaf phase.  Results do not depend on either. */
int lucid_set_traversal_shape(lucid_ctx* ctx, int32_t walk_lanes, int32_t leaf_tris);
/* Persistent traversal: a wavefront fetches new rays once this many of its
   64 lanes are idle (1..64, default 48).  A tuning knob, no counterpart in
   the reference. */
int lucid_set_refill_threshold(lucid_ctx* ctx, int32_t idle_lanes);

/* ------------------------------------------------------------------------
   DiagnosticFilm (diagnostics/diagnostic_film.hpp, pixel_path_data.hpp).
   With diagnostics enabled, lucid_render_tile / lucid_render_tile_device
   trace through the diagnostic tracers (traceSimpleWithDiagnostics for
   "simple", traceMISWithDiagnostics for "mis", traceNEEWithDiagnostics for
   every other algorithm; path_tracer.hpp:880-1514) and rebuild a film sized
   to the tile (pybind_renderer.hpp:206-208, 296-317).  Film pixel (fx, fy)
   holds PixelPathData<16, 4> in the reference's byte layout (9576 bytes) and
   an FP64 Welford {n, mean[3], M2[3]} per sampling strategy (BSDF, NEE,
   MIS_BSDF, MIS_NEE) over each path's per-strategy contribution.  Records of
   tile pixels (fx*sub+dx, fy*sub+dy) are applied sample by sample, the
   block's pixels in row-major order within a sample.
   ------------------------------------------------------------------------ */
typedef struct lucid_diag_config {
    int32_t enabled;    /* 0: render_tile records nothing */
    int32_t subsample;  /* PathRecordingConfig::subsample_factor, >= 1 */
} lucid_diag_config;

typedef struct lucid_diag_info {
    int32_t valid;                         /* a diagnostic render built the film */
    int32_t width, height;                 /* tile the film was built for */
    int32_t sampled_width, sampled_height; /* width / subsample, height / subsample */
    int32_t subsample;
    int64_t total_recorded;                /* DiagnosticFilm::total_paths_recorded */
    int64_t pixel_bytes;                   /* bytes per film pixel (9576) */
} lucid_diag_info;

/* PyRenderer::enable_diagnostics / disable_diagnostics
   (pybind_renderer.hpp:893-908): cfg NULL or enabled = 0 disables; both drop
   the current film. */
int lucid_set_diagnostics(lucid_ctx* ctx, const lucid_diag_config* cfg);
int lucid_diag_get_info(lucid_ctx* ctx, lucid_diag_info* out);
/* Film pixels [first, first + count) in row-major sampled order, as
   PixelPathData<16, 4> bytes (DiagnosticFilm::pixel_data, :118-127).  Write
   stores host-side edits back (PyRenderer::get_pixel_diagnostic sorts a
   pixel's groups in place, pybind_renderer.hpp:1022-1031). */
int lucid_diag_read_pixels(lucid_ctx* ctx, int64_t first, int64_t count, void* dst);
int lucid_diag_write_pixels(lucid_ctx* ctx, int64_t first, int64_t count, const void* src);
/* Per-strategy FP64 Welford of film pixels [first, first + count): 28
   doubles per pixel (4 strategies x {n, mean[3], M2[3]}).  No counterpart in
   the reference (its film keeps float statistics only). */
int lucid_diag_read_strategy(lucid_ctx* ctx, int64_t first, int64_t count, double* dst);
/* DiagnosticFilm::clear (diagnostic_film.hpp:305-312). */
int lucid_diag_clear(lucid_ctx* ctx);

/* ------------------------------------------------------------------------
   Learned path guiding (extension; no counterpart in the reference).
   LUCID_ALGO_GUIDED_MIS renders traceMIS where the bounce direction at
   opaque non-delta surfaces comes from the one-sample mixture
     pdf = alpha * p_guide + (1 - alpha) * p_bsdf
   of the BSDF and an MLP (64 frequency-encoded inputs, 3 x 64 ReLU, 8 von
   Mises-Fisher lobes; BF16 weights, FP32 accumulation on MFMA), with NEE
   weighted against the same mixture.  The network is trained online: every
   guided render records, for 1 path in record_stride, each bounce's
   (position, normal, outgoing, incoming direction, lum(W_bsdf * L_in)), and
   after the render runs train_steps Adam steps of batch `batch` on
   -sum w log p_guide(wi) / sum w.  Until the first training step the guided
   render samples the BSDF only (alpha = 0).  The state lives in the context.
   ------------------------------------------------------------------------ */
typedef struct lucid_guiding_config {
    float alpha;           /* guide-sampling probability, 0..1 */
    int32_t train_steps;   /* Adam steps after each guided render (0: frozen) */
    int32_t batch;         /* training batch, <= 65536 */
    float learning_rate;
    int32_t record_stride; /* 1 path in N (hashed within each block of N)
                              records training samples, >= 1; default 32 */
    uint32_t seed;         /* network initialisation */
    int32_t max_depth;     /* bounces guided: depth < max_depth (0: all) */
    int32_t learn_selection; /* 1: the guide-sampling probability is a network
                                output per surface point, trained with the
                                one-sample mixture's loss and starting at alpha
                                everywhere; 0: alpha everywhere */
} lucid_guiding_config;

typedef struct lucid_guiding_stats {
    int64_t samples_total;  /* training samples recorded so far */
    int64_t train_steps;    /* Adam steps taken so far */
    float last_loss;        /* loss of the last step */
    int32_t active;         /* 1 once the network has been trained */
    int32_t num_params;     /* FP32 master parameters */
} lucid_guiding_stats;

/* Sets the configuration and starts from a fresh network. */
int lucid_set_guiding(lucid_ctx* ctx, const lucid_guiding_config* cfg);
/* Fresh network with the current configuration (e.g. after a scene change). */
int lucid_guiding_reset(lucid_ctx* ctx);
int lucid_guiding_get_stats(lucid_ctx* ctx, lucid_guiding_stats* out);
/* FP32 master parameters (num_params floats): per layer l, W_l [in][out]
   row-major then b_l; layers 64x64, 64x64, 64x64, 64x48 (column 40: the
   selection logit; columns 41..47 0). */
int lucid_guiding_get_params(lucid_ctx* ctx, float* params, int32_t n);
int lucid_guiding_set_params(lucid_ctx* ctx, const float* params, int32_t n);
/* One forward / backward pass of the training kernels over exactly n
   samples (16 floats each: position, normal, outgoing, incoming direction,
   weight, the BSDF sampler's density of the incoming direction, 2 unused)
   with `params`, positions mapped through the box [bmin, bmax], selection
   learning on: raw outputs (n x 40), gradient (num_params) and loss.  For
   the gradient check against the CPU reference. */
int lucid_guiding_debug_step(lucid_ctx* ctx, const float* params, const float* samples, int32_t n,
                             const float bmin[3], const float bmax[3], float* out, float* grad, float* loss);
/* The rendering-side inference kernel's forward pass (4 MFMA tiles per
   wavefront) on the same sample format: raw outputs (n x 40).  For tests. */
int lucid_guiding_debug_infer(lucid_ctx* ctx, const float* params, const float* samples, int32_t n,
                              const float bmin[3], const float bmax[3], float* out);
/* The inference kernel of guided renders (k_guide_infer), run as a guided
   bounce runs it: n rays (6 floats each: origin, direction) are traced by
   the closest-hit kernel, then the network with `params` runs over the queue
   `order` (m path indices, any order, each < n).  out: n x 41 raw outputs
   (the 40 mixture outputs and the selection logit; FP16 widened;
   NaN for paths the kernel must not write: not queued, or
   missed), frame: n x 6 (hit point, shading normal; NaN on a miss), box: the
   scene box the features are normalised by (min[3], max[3]).  For tests. */
int lucid_guiding_debug_render_infer(lucid_ctx* ctx, lucid_scene* scene, const float* params, int32_t n,
                                     const float* rays, const int32_t* order, int32_t m, float* out, float* frame,
                                     float* box);

/* Runs the device BVH traversal code on the CPU over a scene packed from
   `meshes` (no device needed), checking every node and triangle index before
   use: closest hit (any_hit = 0) or occlusion (any_hit = 1) of n_rays rays.
   Outputs as lucid_primary_hits.  Returns LUCID_ERR_DEVICE if an index would
   leave its array or the stack would overflow.  For tests of the traversal;
   it replaces intersectScene (include/geometry/bvh.hpp:155-177) on the host. */
int lucid_debug_host_trace(const lucid_mesh_desc* meshes, int32_t num_meshes, int32_t n_rays, const float* o,
                           const float* d, const float* tmin, const float* tmax, int32_t any_hit, int32_t* mesh,
                           int32_t* tri, float* u, float* v, float* t);

/* The device BVH build's selection (scene_build.hip select_nth) run on the
   host: permutes keys and idx together exactly as std::nth_element (GNU
   libstdc++ introselect) orders n elements around position nth by key, as
   the reference's BVH build calls it (include/core/scene.hpp:138-142).
   depth_limit < 0: the library's own partition budget, 2 floor(log2 n);
   >= 0: that budget (the heap-select fallback at 0).  For tests. */
int lucid_debug_select_nth(float* keys, int32_t* idx, int32_t n, int32_t nth, int32_t depth_limit);

/* Storage of the density grids of scenes created after this call: 32 (FP32,
   the reference's values, default) or 16 (FP16, round to nearest even; half
   the device bytes, arithmetic still FP32; each medium's delta-tracking
   majorant is taken from the rounded values).  FP16 refuses grids with
   values beyond +-65504.  Replaces the FP32 grid upload of the reference's
   GPU packer (src/gpu/path_tracer.cpp:162-469, volume_grid_buf) and the
   float grid of sample_grid_trilinear (include/volume/transmittance.hpp:298-352). */
int lucid_set_grid_format(lucid_ctx* ctx, int32_t bits);
/* Device bytes of a scene's density grids. */
int64_t lucid_scene_grid_bytes(const lucid_scene* scene);

/* Where scenes created from now on build their per-mesh BVHs, node array and
   triangle records: 1 on the device (the default), 0 on the host.  Both give
   the same bytes: the reference's median-split trees (core/scene.hpp:65-152,
   std::nth_element order) under a top-level tree over the mesh roots. */
int lucid_set_scene_build(lucid_ctx* ctx, int32_t mode);
/* The scene's traversal data in host memory, for tests of the two builds:
   n_nodes nodes of 64 bytes (pass nodes = NULL to read only the counts),
   then per triangle in BVH order tri (12 floats: v0 | mesh, e1 | triangle,
   e2 | light index, ids as int32 bits), fn (4 floats: face normal | smooth
   flag), vn (12 floats: 3 vertex normals; zeros without smooth meshes) and
   uv (6 floats; zeros without node materials).  Any of tri / fn / vn / uv
   may be NULL.  stack_depth: the deepest node path; built_on_device: 1 for a
   device build. */
int lucid_scene_read_geometry(lucid_ctx* ctx, lucid_scene* scene, void* nodes, float* tri, float* fn, float* vn,
                              float* uv, int64_t* n_nodes, int32_t* stack_depth, int32_t* built_on_device);

/* Instrumentation flags (LUCID_PROFILE_*), 0 = off (default). */
int lucid_set_profiling(lucid_ctx* ctx, int32_t flags);

/* ---------------------------------------------------------- sessions --
   A session is a progressive accumulation of one full frame (width x
   height) of one scene and camera: a device radiance sum and its sample
   count.  Every pass adds samples [n, n + k) with the per-pixel seeds of
   render_tile and jitter on every sample, so a session that has taken N
   samples in any number of passes holds exactly the sum of one
   render_tile(0, 0, w, h, w, h, N, 0, max_depth) call (N > 1).  Sessions
   retain their scene; the library owns their device memory. */
typedef struct lucid_session lucid_session;

/* Replaces the per-session state of the reference's render sessions
   (P:render_session.py:530-608 chaining render_tile with sample_offset =
   running total) and of its async accumulator
   (src/gpu/path_tracer.cpp:838-884, created by start_async :1106-1245). */
int lucid_session_create(lucid_ctx* ctx, lucid_scene* scene, const lucid_camera* cam, int32_t width, int32_t height,
                         int32_t max_depth, int32_t algorithm, lucid_session** out);
/* Stops a running worker and frees the session (stop_async + the
   AsyncState destructor, src/gpu/path_tracer.cpp:1247-1254). */
void lucid_session_destroy(lucid_session* s);
/* New camera and depth; clears the accumulator.  Not while async. */
int lucid_session_set_camera(lucid_session* s, const lucid_camera* cam, int32_t max_depth);
int lucid_session_clear(lucid_session* s);
int64_t lucid_session_samples(const lucid_session* s);

/* One pass over many sessions: adds `samples` samples to each.  Sessions
   are spread over the context's streams (longest first) so that their
   launches overlap; guided sessions run one after another on the first.
   Replaces the per-session render_tile calls of the reference's session
   layer (P:render_session.py:103 runs one worker per session, serialised
   by gpu_mutex_, include/pybind_renderer.hpp:1118).  Honours lucid_cancel:
   returns LUCID_ERR_CANCELLED with every issued wave counted. */
int lucid_sessions_render(lucid_ctx* ctx, lucid_session* const* sessions, int32_t n, int32_t samples);

/* Reads the accumulator as render_tile's layout (rows bottom first, alpha
   1): average != 0 gives (sum / samples) * 2^EV clamped at 0 (what
   render_poll_async returns), average == 0 the sum * 2^EV clamped at 0
   (what render_tile returns). */
int lucid_session_read(lucid_session* s, float* out_rgba, int32_t average, int64_t* samples);

/* Async accumulation on a worker thread: PyRenderer::render_start_async
   (include/pybind_renderer.hpp:689-753; the worker of
   src/gpu/path_tracer.cpp:887-955).  Clears the accumulator and starts
   passes of 1, 2, 4, ... samples (grown while a pass takes < 20 ms, at most
   256), publishing an averaged snapshot after every pass.  Cancel of the
   context does not stop it; lucid_session_stop_async does. */
int lucid_session_start_async(lucid_session* s);
/* render_stop_async (pybind_renderer.hpp:787-794): stops after the current
   wave and publishes a final snapshot. */
int lucid_session_stop_async(lucid_session* s);
/* render_reset_async (pybind_renderer.hpp:800-834, reset_async
   src/gpu/path_tracer.cpp:1256-1276): new camera and depth, the
   accumulator is cleared before the next pass and no snapshot of the old
   camera is published after this call.  No-op when not running;
   LUCID_ERR_INVALID when the size differs (stop and start instead). */
int lucid_session_reset_async(lucid_session* s, const lucid_camera* cam, int32_t width, int32_t height,
                              int32_t max_depth);
/* render_poll_async / render_snapshot_revision_async
   (pybind_renderer.hpp:755-785, 848-855): the latest snapshot's sample
   count (0 = none yet, or not running), its pixels into out_rgba when it
   is not NULL and the count is > 0 (width * height * 4 floats, rows bottom
   first), and the monotonic snapshot revision. */
int lucid_session_poll_async(lucid_session* s, float* out_rgba, int64_t* samples, uint32_t* revision);
/* is_render_async_running (pybind_renderer.hpp:836-843). */
int32_t lucid_session_async_running(const lucid_session* s);
/* Why the async worker stopped on its own ("" if it did not). */
const char* lucid_session_error(const lucid_session* s);

#ifdef __cplusplus
}
#endif

#endif /* LUCID_HIP_H */
