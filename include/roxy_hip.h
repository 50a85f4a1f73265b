/*
 * roxy_hip.h — C ABI of the MI355X (gfx950) path-tracing backend.
 *
 * Every entry point returns an int status (RX_OK == 0) and never throws
 * across the ABI; rx_last_error() returns a thread-local message for the
 * last failing call on the calling thread.  All pointers named `d_*` or
 * typed as device structures are HIP device pointers that the caller owns
 * (the library borrows them for the duration of the call).  `stream` is a
 * hipStream_t passed as void*; NULL means the null stream.  Work is
 * stream-ordered: a call returns once its kernels are enqueued unless it
 * says otherwise.
 *
 * Reference interface each entry point replaces (GuyontheCouch006/roxy-dcc):
 *   rx_bvh_*         rendering/taichi/bvh_builder.py:315-558 (GPUBVHBuilder.build,
 *                    _morton_order, _build_morton_node) and :560-655 (upload)
 *   rx_closest_hit   rendering/taichi/intersect.py:249-309 (scene_intersect,
 *                    bvh_intersect :81-152)
 *   rx_any_hit       rendering/taichi/intersect.py:222-246 (scene_occluded,
 *                    bvh_occluded :155-219)
 *   rx_render        rendering/taichi/trace.py:239-283 (render_kernel) driven by
 *                    rendering/taichi_renderer.py:140-156 (one launch per sample)
 *   rx_denoise       rendering/denoise.py:23-104 (edge_aware_denoise)
 *   rx_display       rendering/denoise.py:107-108 (linear_to_display)
 *   rx_flatten_indexed  scene/mesh.py:397-430 (indexed_triangle_arrays) via
 *                    rendering/taichi/extractor.py:174-190
 */
#ifndef ROXY_HIP_H
#define ROXY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RX_ABI_VERSION 6

/* status codes */
#define RX_OK 0
#define RX_ERR_INVALID 1     /* bad argument (maps to ValueError) */
#define RX_ERR_CAPACITY 2    /* capacity exceeded (maps to AssertionError) */
#define RX_ERR_HIP 3         /* a HIP runtime call failed */
#define RX_ERR_WORKSPACE 4   /* workspace too small */

/* Primitive kinds written by rx_closest_hit (`kind` output). */
#define RX_KIND_MISS 0
#define RX_KIND_PRIM 1   /* analytic primitive slot (id = slot index) */
#define RX_KIND_BVH 2    /* BVH triangle (id = index in BVH leaf order) */

/* Analytic primitive types: rendering/taichi/intersect.py:265-282 */
#define RX_PRIM_SPHERE 0
#define RX_PRIM_PLANE 1
#define RX_PRIM_CUBE 2
#define RX_PRIM_TRIANGLE 4

/* Material types: scene/materials.py:60,88,117,163,183 */
#define RX_MAT_DIFFUSE 0
#define RX_MAT_METAL 1
#define RX_MAT_DIELECTRIC 2
#define RX_MAT_EMISSIVE 3
#define RX_MAT_GLOSSY 4

/* Largest leaf the packed node link can describe, and largest triangle
 * count (the packed link keeps 23 bits of tri_start). */
#define RX_MAX_LEAF_TRIS 255
#define RX_MAX_TRIANGLES (1 << 23)
/* Deepest BVH (root depth 0) the traversal stacks support. */
#define RX_MAX_BVH_DEPTH 63

/*
 * BVH2 node, 64 bytes, one per reference node in the reference's preorder
 * numbering (node i here is node i of GPUBVHBuilder.nodes).  Each record
 * holds the boxes of BOTH children, so one 64-byte fetch answers both
 * child slab tests of rendering/taichi/intersect.py:121-126.
 *   box[0..5]  left child  min xyz, max xyz
 *   box[6..11] right child min xyz, max xyz
 *   link[c]    child c: its node index if it is interior, otherwise
 *              ~((tri_start << 8) | tri_count) (negative) for a leaf
 *   node[c]    child c's node index (interior or leaf), for export
 * A leaf's own record has box[0..5] = its own box, link = {~leaf, -1},
 * node = {-1, -1}.
 */
typedef struct rx_node {
    float box[12];
    int32_t link[2];
    int32_t node[2];
} rx_node;

/* Root record of a BVH: the root's own box (tested first, as
 * intersect.py:95-100 does) and its link (0, or ~leaf if the root is a
 * leaf).  depth = deepest node depth (root = 0). */
typedef struct rx_bvh_header {
    float bmin[3];
    int32_t root_link;
    float bmax[3];
    int32_t depth;
} rx_bvh_header;

/* Triangle geometry in BVH leaf order, 48 bytes: v0, e1 = v1 - v0,
 * e2 = v2 - v0 (the first two lines of intersect.py:57-78, done once). */
typedef struct rx_tri {
    float v0[3];
    float e1[3];
    float e2[3];
    float pad[3];
} rx_tri;

/* Shading data in BVH leaf order, 48 bytes: per-vertex normals
 * (fields.py:62-64), palette material index (fields.py:71), has_uv. */
typedef struct rx_tri_shade {
    float n0[3];
    float n1[3];
    float n2[3];
    int32_t mat;
    int32_t has_uv;
    int32_t pad;
} rx_tri_shade;

/* Analytic primitive slot, 48 bytes (fields.py:22-33).
 *   sphere:   a = center xyz, radius
 *   plane:    a = normal xyz, offset   (n.p + offset = 0)
 *   cube:     a = center xyz, 0; b = half extents xyz, 0
 *   triangle: v0 = a.xyz, v1 = (a.w, b.x, b.y), v2 = (b.z, b.w, c.x) */
typedef struct rx_prim {
    float a[4];
    float b[4];
    float c[3];
    int32_t type;
} rx_prim;

/* Material record, 32 bytes (fields.py:35-40 and :73-79). */
typedef struct rx_material {
    int32_t type;
    float albedo[3];
    float roughness;
    float ior;
    float emission;
    int32_t texture;
} rx_material;

/* Emissive-sphere light, 32 bytes (fields.py:43-47). */
typedef struct rx_light {
    float center[3];
    float radius;
    float albedo[3];
    float emission;
} rx_light;

/* Image texture of a scene's texture cache (fields.py:82-88): u8 RGB
 * texels, row-major rows of 3*width bytes, starting `offset` bytes into
 * the scene's texel pool; flip_v as ImageTexture.flip_v.  32 bytes. */
typedef struct rx_texture {
    int32_t width;
    int32_t height;
    int32_t flip_v;
    int32_t pad0;
    long long offset;
    long long pad1;
} rx_texture;

/* Scene view: device pointers plus host-side sizes.  Every record array must
 * be 16-byte aligned (records are read as 16-byte words); the entry points
 * return RX_ERR_INVALID otherwise.  tri_uv (6 floats per triangle in leaf
 * order: uv0, uv1, uv2), textures and texels are needed only when a
 * material has a texture (material.texture >= 0, trace.py:129-136).
 * flags: RX_SCENE_FINITE_BOXES when the caller knows every triangle vertex
 * (so every BVH box coordinate) is finite; rx_render may then trace the
 * tree from a copy without child links (see rx_render_workspace_size_for_
 * scene).  Set it only when it holds: a NaN box coordinate would be read as
 * a leaf mark in that copy.  0 is always safe. */
#define RX_SCENE_FINITE_BOXES 1
typedef struct rx_scene {
    const rx_node* nodes;
    const rx_bvh_header* header;
    const rx_tri* tris;
    const rx_tri_shade* tri_shade;
    const rx_material* materials;
    const rx_prim* prims;
    const rx_material* prim_materials;
    const rx_light* lights;
    const float* tri_uv;
    const rx_texture* textures;
    const unsigned char* texels;
    int32_t n_nodes;
    int32_t n_tris;
    int32_t n_materials;
    int32_t n_prims;
    int32_t n_lights;
    int32_t bvh_depth;
    int32_t n_textures;
    int32_t flags;
} rx_scene;

/* Pinhole camera (rendering/taichi/camera.py:23-30). */
typedef struct rx_camera {
    float pos[3];
    float fwd[3];
    float right[3];
    float up[3];
    float fov_tan;
    float aspect;
    int32_t width;
    int32_t height;
} rx_camera;

/* Flags for rx_render_params.flags */
#define RX_FLAG_PROFILE 1        /* time each stage with HIP events (synchronises) */
#define RX_FLAG_COUNT_TRAFFIC 2  /* count node pops / triangle / primitive tests */
#define RX_FLAG_PRIM_PREPASS 4   /* closest-primitive scan as its own pass over each ray
                                    queue (automatic above 256 analytic primitives) */
#define RX_FLAG_KEEP_QUEUE_ORDER 32 /* rx_render: trace every ray queue in path order
                                       (primary rays sample-major, secondary rays as shade
                                       wrote them) instead of pixel-major primaries and
                                       secondaries sorted by direction and origin (same
                                       results, slower) */
#define RX_FLAG_ADAPTIVE 16      /* adaptive stop: rendering/ray_tracer.py:531-541 (see
                                    adaptive_* below) */
#define RX_FLAG_SMALL_BLOCKS 8   /* traversal in 64-thread blocks with the deep stack (the
                                    shape used anyway for BVHs deeper than 23).  The batch
                                    workspace query sees it; for rx_render size the
                                    workspace with bvh_depth = RX_MAX_BVH_DEPTH. */

/* Integrator settings (render_kernel arguments, trace.py:239-254). */
typedef struct rx_render_params {
    int32_t max_depth;
    int32_t use_sky;
    int32_t direct_light_mode;       /* 0 = one (random light), 1 = all */
    int32_t direct_light_max_depth;
    float sample_clamp;              /* <= 0 disables the clamp */
    float bg[3];
    uint32_t seed_lo;
    uint32_t seed_hi;
    int32_t flags;
    int32_t samples_per_pass;        /* samples traced together per wavefront pass */
    int32_t trav_blocks_per_cu;      /* resident blocks per CU of the persistent traversal
                                        kernels: 0 = the register budget's maximum (8 blocks
                                        of 256, or 32 of 64 threads), else 1 .. that maximum */
    /* With RX_FLAG_ADAPTIVE (rendering/ray_tracer.py:531-541): after sample n
     * (n = samples rendered, n >= 2) the error is the mean over pixels and
     * channels of |mean after n - mean after n-1| (linear colour); the job
     * stops after sample n when n < its requested total, n >= min_samples,
     * (n - min_samples) % check_interval == 0 and error <= threshold.  A pass
     * whose samples run past the stop accumulates only up to it. */
    float adaptive_threshold;
    int32_t adaptive_min_samples;
    int32_t adaptive_check_interval;
} rx_render_params;

/* Per-pixel running means written by rx_render (fields.py:14-19). */
typedef struct rx_accum {
    float* color;    /* H*W*3 linear radiance */
    float* normal;   /* H*W*3 first-hit normal */
    float* albedo;   /* H*W*3 first-hit albedo */
    float* depth;    /* H*W   first-hit t */
    float* display;  /* H*W*3 sqrt(min(color, 1)) */
} rx_accum;

/* Device-side counters (64-bit).  rays = extension rays + shadow rays cast,
 * counted as trace.py:78-79,150-151 does. */
typedef struct rx_counters {
    unsigned long long rays;
    unsigned long long shadow_rays;
    unsigned long long node_pops;      /* closest-hit traversal only */
    unsigned long long tri_tests;      /* closest-hit traversal only */
    unsigned long long prim_tests;     /* closest-hit traversal only */
    unsigned long long shadow_node_pops;
    unsigned long long shadow_tri_tests;
    unsigned long long shadow_prim_tests;
} rx_counters;

/* Host-side stage timings filled when RX_FLAG_PROFILE is set. */
typedef struct rx_profile {
    float ms_generate;
    float ms_extend;
    float ms_shade;
    float ms_connect;
    float ms_accumulate;
    int32_t n_extend;
    int32_t n_connect;
    int32_t n_passes;
    float ms_prim_scan;     /* primitive pre-pass (scenes with many primitives) */
    int32_t n_prim_scan;
    float ms_reorder;       /* secondary-ray reordering (keys, sort, gather) */
    int32_t n_reorder;      /* its launches (three per reordered bounce) */
} rx_profile;

/* ---- library ---------------------------------------------------------- */
int rx_version(void);
const char* rx_last_error(void);

/* ---- BVH build: rendering/taichi/bvh_builder.py:412-486 ----------------
 * Node count and depth of the count-median Morton tree over n_tris
 * triangles with leaves of at most max_leaf (closed form: the topology of
 * _build_morton_node depends only on the count). */
int rx_bvh_morton_shape(int n_tris, int max_leaf, int* n_nodes, int* depth);
int rx_bvh_workspace_size(int n_tris, int max_leaf, size_t* bytes);

/* Reference-layout node arrays (fields.py:50-56); every pointer is device. */
typedef struct rx_bvh_ref {
    float* aabb_min;    /* n_nodes*3 */
    float* aabb_max;    /* n_nodes*3 */
    int32_t* left;      /* n_nodes, -1 at leaves */
    int32_t* right;     /* n_nodes, -1 at leaves */
    int32_t* tri_start; /* n_nodes, -1 at interior nodes */
    int32_t* tri_count; /* n_nodes, 0 at interior nodes */
} rx_bvh_ref;

/* GPU Morton build (bvh_builder.py:347-368, :416-486).  d_v0/d_v1/d_v2 are
 * n_tris*3 floats.  Writes the reference-layout arrays into *ref (device
 * arrays the caller sized with rx_bvh_morton_shape), the packed 64-byte
 * nodes, the header and the leaf order (_finish_ordered_arrays'
 * ordered_tris).  Synchronous on `stream` only. */
int rx_bvh_build_morton(const float* d_v0, const float* d_v1, const float* d_v2,
                        int n_tris, int max_leaf,
                        void* d_workspace, size_t workspace_bytes,
                        const rx_bvh_ref* ref, rx_node* d_nodes,
                        rx_bvh_header* d_header, int32_t* d_order,
                        void* stream);

/* Pack reference-layout node arrays (built elsewhere, e.g. by the median
 * builders) into the 64-byte nodes and the header.  `depth` is the tree's
 * deepest node depth. */
int rx_bvh_pack(const rx_bvh_ref* ref, int n_nodes, int depth,
                rx_node* d_nodes, rx_bvh_header* d_header, void* stream);

/* core/bvh.py:2-48 (BVHNode, the CPU tracer's per-mesh tree) built on the
 * GPU and flattened in preorder into the same layouts.  Boxes are the union
 * of Triangle.local_bounds (scene/mesh.py:68-76, fp64 vertex min/max -/+
 * 1e-4), rounded to fp32 when stored; a node with more than max_leaf
 * triangles at the root, or 4 below it (the recursion does not forward
 * max_leaf_size, core/bvh.py:12-13), at depth <= 20 sorts them stably by
 * the fp64 centroid ((v0 + v1) + v2) * (1 / 3) along the longest box axis
 * (ties: x, then y) and splits at len // 2.  Vertices are n_tris*3 doubles
 * (the reference's Python floats), finite.  The shape depends only on the
 * count.  rx_bvh_build_object_median returns once the build has finished
 * on `stream`. */
int rx_bvh_object_median_shape(int n_tris, int max_leaf, int* n_nodes, int* depth);
int rx_bvh_object_median_workspace_size(int n_tris, int max_leaf, size_t* bytes);
int rx_bvh_build_object_median(const double* d_v0, const double* d_v1, const double* d_v2,
                               int n_tris, int max_leaf,
                               void* d_workspace, size_t workspace_bytes,
                               const rx_bvh_ref* ref, rx_node* d_nodes,
                               rx_bvh_header* d_header, int32_t* d_order,
                               void* stream);

/* Gather triangles into leaf order (bvh_builder.py:527-544) and precompute
 * edges.  d_n* may be NULL (face normal is used); d_mat_idx may be NULL
 * (material 0); d_has_uv may be NULL. */
int rx_bvh_pack_triangles(const int32_t* d_order, int n_tris,
                          const float* d_v0, const float* d_v1, const float* d_v2,
                          const float* d_n0, const float* d_n1, const float* d_n2,
                          const int32_t* d_mat_idx, const int32_t* d_has_uv,
                          rx_tri* d_tris, rx_tri_shade* d_shade, void* stream);
/* Gather per-corner UVs (n x 2 each) into leaf order, 6 floats per
 * triangle (the _bvh_uv* fields of fields.py:65-67 in BVH order). */
int rx_bvh_pack_uvs(const int32_t* d_order, int n_tris, const float* d_uv0, const float* d_uv1,
                    const float* d_uv2, float* d_tri_uv, void* stream);

/* ---- queries: rendering/taichi/intersect.py ------------------------------
 * d_rays: 6*n floats, SoA (ox[n], oy[n], oz[n], dx[n], dy[n], dz[n]).
 * Outputs are n-element device arrays; any output pointer may be NULL. */
int rx_closest_hit(const rx_scene* scene, const float* d_rays, int n,
                   float* d_t, int32_t* d_id, int32_t* d_kind,
                   float* d_u, float* d_v, rx_counters* d_counters, void* stream);

int rx_any_hit(const rx_scene* scene, const float* d_rays, const float* d_tmax,
               int n, int32_t* d_occluded, rx_counters* d_counters, void* stream);

/* ---- integrator: rendering/taichi/trace.py:139-283 ----------------------
 * Wavefront path tracer: generate, extend, shade, connect, accumulate.
 * Renders samples [first_sample, first_sample + n_samples) and folds them
 * into *accum in sample order with trace.py's running mean.  The
 * workspace must hold rx_render_workspace_size() bytes for the same
 * width, height, params->samples_per_pass and scene->bvh_depth.  *profile is written (and
 * the call synchronises) only when RX_FLAG_PROFILE is set. */
int rx_render_workspace_size(int width, int height, int samples_per_pass,
                             int bvh_depth, size_t* bytes);
/* The same workspace with room past it for the eye-relative copies of the
 * scene that rx_render's camera-ray traversal reads when they fit: node
 * boxes and triangle vertices minus the camera position, with the terms of
 * the intersection tests built from them alone (64 bytes per node, 48 per
 * triangle, 32 per analytic primitive), and with 1..256 analytic
 * primitives 32 bytes per 8x8 pixel tile: the primitives the tile's camera
 * rays can hit, the only ones they then test.  When the scene has
 * RX_SCENE_FINITE_BOXES, a tree of depth d >= 1 and at most 4 * n_nodes
 * heap slots (2^(d+1) - 1), the room holds instead of the node copy a
 * link-free copy of the tree that every traversal of the call reads: the
 * interior nodes' child boxes in heap order (root slot 0, children of slot
 * h at 2h+1 and 2h+2; 64 bytes per interior slot, 2^d - 1 of them, in
 * world space and again minus the camera position), a leaf child's box
 * marked on its x axis in a way the slab test does not see, and 4 bytes
 * per heap slot for the leaves' triangle ranges.  Same images, ray counts,
 * node pops and triangle tests either way (primitive tests can only drop);
 * a workspace of rx_render_workspace_size() bytes renders without them.
 * No counterpart in the reference (render_kernel recomputes these per ray
 * and scans every primitive, rendering/taichi/intersect.py:28-78,259-287). */
int rx_render_workspace_size_for_scene(const rx_scene* scene, int width, int height,
                                       int samples_per_pass, size_t* bytes);
int rx_render(const rx_scene* scene, const rx_camera* camera,
              const rx_render_params* params, int first_sample, int n_samples,
              void* d_workspace, size_t workspace_bytes,
              const rx_accum* accum, rx_counters* d_counters,
              rx_profile* profile, void* stream);

/* rx_render without a caller workspace, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The workspace is a
 * library-owned device array of RX_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends; a pass needing more (rx_render_workspace_size)
 * fails with RX_ERR_CAPACITY, so use fewer samples_per_pass.  No HIP
 * allocation is made; RX_FLAG_PROFILE is rejected.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
#define RX_NOALLOC_SCRATCH_BYTES (768ull << 20)
int rx_render_noalloc(const rx_scene* scene, const rx_camera* camera,
                      const rx_render_params* params, int first_sample, int n_samples,
                      const rx_accum* accum, rx_counters* d_counters, void* stream);

/* ---- batches: many scenes in one wavefront ------------------------------
 * One job = one scene, camera, settings and accumulation target.  A batch
 * traces a pass of every unfinished job together: each pass renders
 * min(samples_left, samples_per_pass) samples of every job, in the
 * job's sample order, so a job's result does not depend on the others in
 * the batch.  first_sample/samples_left/rays advance on the device, so a
 * pass can be captured once in a HIP graph and replayed.  No reference
 * counterpart: render_kernel renders one scene per launch
 * (rendering/taichi/trace.py:239); this is the batched form of it. */
typedef struct rx_job {
    rx_scene scene;
    rx_camera camera;
    rx_render_params params;   /* params.samples_per_pass = samples per pass */
    rx_accum accum;
    int32_t first_sample;      /* next sample index (advanced per pass) */
    int32_t samples_left;      /* samples still to render (decremented per pass) */
    int32_t path_offset;       /* set by the library: first path slot */
    int32_t pixel_offset;      /* set by the library: first pixel slot */
    unsigned long long rays;   /* rays cast for this job (accumulated) */
    float adaptive_error;      /* last adaptive error (RX_FLAG_ADAPTIVE), -1 before one exists */
    int32_t adaptive_stopped;  /* 1 once the adaptive rule stopped the job */
} rx_job;

typedef struct rx_batch rx_batch;

int rx_batch_workspace_size(const rx_job* jobs, int n_jobs, size_t* bytes);
/* Copies the job records into the workspace and prepares the passes.
 * use_graph != 0 captures one pass in a HIP graph on the first run and
 * replays it.  The workspace and every pointer in the jobs must stay
 * valid until rx_batch_destroy. */
int rx_batch_create(const rx_job* jobs, int n_jobs, void* d_workspace, size_t workspace_bytes,
                    int use_graph, rx_batch** out);
int rx_batch_run(rx_batch* batch, int passes, rx_counters* d_counters, rx_profile* profile, void* stream);
/* Node counts of the pass captured by rx_batch_run (kernels, memory fills,
 * any other node), or -1 each before a capture.  No reference counterpart:
 * lets a caller check that a replayed pass holds only kernel nodes. */
int rx_batch_graph_nodes(const rx_batch* batch, int* kernels, int* memsets, int* others);
/* Copies the job records (progress and ray counts) back; synchronises `stream`. */
int rx_batch_read_jobs(rx_batch* batch, rx_job* jobs_out, void* stream);
int rx_batch_destroy(rx_batch* batch);

/* ---- ray-queue sort ------------------------------------------------------
 * The stable LSD radix sort rx_render uses to reorder secondary ray queues
 * (no reference counterpart: render_kernel traces pixels in place).  Sorts
 * the first *d_count pairs (read on the device, clamped to cap) of
 * (d_keys, d_vals) by the low key_bits of the keys; the sorted pairs end
 * in (d_keys, d_vals), (d_keys2, d_vals2) are the second buffers (cap
 * entries each).  Counts below min_n are left as they are.  Enqueues only
 * kernels (no memory fills), so it can be captured in a HIP graph. */
int rx_radix_sort_scratch_size(long long cap, size_t* bytes);
int rx_radix_sort_pairs(uint32_t* d_keys, uint32_t* d_vals, uint32_t* d_keys2, uint32_t* d_vals2,
                        const int32_t* d_count, long long cap, int min_n, int key_bits, void* d_scratch,
                        size_t scratch_bytes, void* stream);

/* ---- denoise: rendering/denoise.py:23-108 -------------------------------
 * Guided bilateral filter over a (2*radius+1)^2 window whose taps are
 * `dilation` pixels apart (dilation 1 is the reference filter), reflect
 * padding.  d_normal/d_albedo/d_depth may be NULL (guide skipped).  d_out
 * must not alias d_color. */
int rx_denoise(const float* d_color, const float* d_normal, const float* d_albedo,
               const float* d_depth, int width, int height, int radius, int dilation,
               float sigma_color, float sigma_normal, float sigma_albedo,
               float sigma_depth, float amount, float* d_out, void* stream);

/* sqrt(min(max(x, 0), 1)) elementwise over n floats. */
int rx_display(const float* d_linear, int n, float* d_out, void* stream);

/* ---- scene flattening: scene/mesh.py:397-430 (indexed_triangle_arrays,
 * with the transforms of :547-568) as rendering/taichi/extractor.py:174-190
 * calls it for an IndexedMesh object.  HOST pointers: the caller's numpy
 * arrays.  Corner positions through `matrix` (row-major 4x4, perspective
 * divide where |w| > 1e-12), per-vertex normals (or the face normal of the
 * untransformed corners where a triangle has no normal indices) through
 * the 3x3 part of `normal_matrix`, then normalised; uvs where all three
 * indices are >= 0, else zeros.  fp64 in a fixed order, fp32 out; a NULL
 * matrix is the identity.  threads <= 0 uses every host core. */
int rx_flatten_indexed(const double* positions, int n_positions, const int32_t* tri_pos_idx, int n_tris,
                       const double* normals, int n_normals, const int32_t* tri_normal_idx,
                       const double* uvs, int n_uvs, const int32_t* tri_uv_idx,
                       const double* matrix, const double* normal_matrix,
                       float* v0, float* v1, float* v2, float* n0, float* n1, float* n2,
                       float* uv0, float* uv1, float* uv2, int32_t* has_uv, int threads);

#ifdef __cplusplus
}
#endif

#endif /* ROXY_HIP_H */
