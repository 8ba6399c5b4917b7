/*
 * crisoptix.h -- C ABI of the MI355X-native ReSTIR PT / Conditional ReSTIR frame loop.
 *
 * This is the drop-in boundary that replaces the OptiX launch path of CrisOptix:
 * every entry point below names the reference interface it stands in for.
 * All functions return 0 on success and a negative cx_status on failure; they never
 * throw across the ABI.  cx_last_error() returns a message for the last failure on a
 * context.  Device pointers are plain `void*`/typed pointers into HIP device memory;
 * host pointers are marked `host`.  No torch or glm types appear in any signature.
 *
 * Threading: one cx_ctx per host thread (the reference is single-threaded on stream 0,
 * src/Renderer.cpp:259-371).  Work is enqueued on the caller's hipStream_t (passed as
 * `void*`, NULL = the default stream); nothing inside a frame synchronises the device.
 */
#ifndef CRISOPTIX_H
#define CRISOPTIX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CX_ABI_VERSION 10

typedef enum cx_status {
    CX_OK = 0,
    CX_ERR_INVALID_ARG = -1,
    CX_ERR_HIP = -2,
    CX_ERR_OUT_OF_MEMORY = -3,
    CX_ERR_UNSUPPORTED = -4,
    CX_ERR_NO_DEVICE = -5,
    CX_ERR_BUILD = -6
} cx_status;

/* RendererType (include/graph/restir/settings.h:5-10). */
typedef enum cx_renderer_type {
    CX_PATH_TRACER = 0,
    CX_RESTIR_PT = 1,
    CX_CONDITIONAL_RESTIR = 2
} cx_renderer_type;

/* PrefixRadiusType (include/graph/restir/settings.h:15-19). */
typedef enum cx_radius_type {
    CX_RADIUS_CONSTANT = 0,
    CX_RADIUS_PATH_LENGTH = 1
} cx_radius_type;

/* The six raygen programs of a frame (src/Renderer.cpp:91-136). */
typedef enum cx_pass {
    CX_PASS_PREFIX_GEN_TEMP = 0,    /* shader/prefix_gen_temp_reuse.cu:100 */
    CX_PASS_PREFIX_SPATIAL = 1,     /* shader/prefix_spatial_reuse.cu:136 */
    CX_PASS_SUFFIX_GEN_TEMP = 2,    /* shader/suffix_gen_temp_reuse.cu:114 */
    CX_PASS_SUFFIX_SPATIAL = 3,     /* shader/suffix_spatial_reuse.cu:147 */
    CX_PASS_PREFIX_STORE_ENTRIES = 4, /* shader/prefix_store_entries.cu:9 */
    CX_PASS_ENTRIES_BUILD = 5,      /* PrefixAccelStruct::Rebuild, src/PrefixAccelStruct.cpp:14-26 */
    CX_PASS_FINAL_GATHER = 6,       /* shader/final_gather.cu:323 */
    CX_PASS_COUNT = 7
} cx_pass;

/*
 * Per-frame parameter block.  Mirrors LaunchParams (include/graph/LaunchParams.h:11-36)
 * and RestirParams (include/graph/restir/RestirParams.h:11-59) with every buffer view and
 * traversable handle removed: buffers are owned by the renderer, and the block is passed
 * by value as a kernel argument (no __constant__ upload, so it is graph-capturable).
 * `windowRadius` is WINDOW_RADIUS (include/graph/restir/path_gen.h:12), a parameter here.
 * `random` and `frameIdx` are maintained by cx_renderer_frame() exactly like
 * Renderer::LaunchFrame (src/Renderer.cpp:212-215); the values passed in are ignored there.
 */
typedef struct cx_launch_params {
    uint32_t random;
    uint32_t width;
    uint32_t height;
    uint32_t windowRadius;
    uint64_t frameIdx;
    float neeProb;
    int32_t enableAccum;
    int32_t maxPathLen;
    int32_t rendererType;          /* cx_renderer_type */
    /* RestirParams */
    float reconMinRoughness;
    float reconMinDistance;
    int32_t spatialRoundIdx;
    int32_t prefixLen;
    int32_t prefixEnableTemporal;
    int32_t prefixEnableSpatial;
    int32_t prefixSpatialCount;
    int32_t prefixSpatialRounds;
    int32_t suffixEnableTemporal;
    int32_t suffixEnableSpatial;
    int32_t suffixSpatialCount;
    int32_t suffixSpatialRounds;
    int32_t gatherN;
    int32_t gatherM;
    float gatherRadius;
    int32_t gatherRadiusType;      /* cx_radius_type */
    int32_t trackPrefixStats;
    int32_t pad0;
} cx_launch_params;

/* CameraData (include/graph/Camera.h:6-12) without the unused view matrices. */
typedef struct cx_camera {
    float pos[3];
    float U[3];
    float V[3];
    float W[3];
} cx_camera;

/* Material constants (include/model/Material.h:16-33, src/Material.cpp:3-24); the texture
 * bindings of a material are in cx_material_tex. */
typedef struct cx_material {
    float diffColor[3];
    float specF0[3];
    float emissiveColor[3];
    float roughness;
} cx_material;

/* SpecTexUsage (include/model/Material.h:10-14).  OCC_ROUGH_METAL reads roughness from the
 * green and metalness from the blue channel of the specular value (texture or constant) and
 * sets the specular colour to 1 before the metal mix (shader/brdf.cu:127-138). */
typedef enum cx_spec_tex_usage {
    CX_SPEC_TEX_COLOR = 0,
    CX_SPEC_TEX_OCC_ROUGH_METAL = 1
} cx_spec_tex_usage;

/* Texture bindings of one material (MaterialSbtData::hasDiffTex/diffTex, hasSpecTex/specTex,
 * hasRoughTex/roughTex, specTexUsage; include/model/Material.h:23-32): an index into
 * cx_scene_desc.textures, or -1 for none. */
typedef struct cx_material_tex {
    int32_t diffTex;
    int32_t specTex;
    int32_t roughTex;
    uint32_t specTexUsage;         /* cx_spec_tex_usage */
} cx_material_tex;

/* An RGBA8 texture (src/Texture.cpp:7-76: stbi_load(..., STBI_rgb_alpha), rows in file order,
 * sampled with wrap addressing, bilinear filtering, normalized coordinates and normalized-float
 * reads).  Sampling at (u, v): x = frac(u) * width - 0.5, y = frac(v) * height - 0.5; texels
 * floor(x), floor(x) + 1 (and likewise in y) wrapped modulo the size; weights
 * alpha = frac(x), beta = frac(y), each rounded to a multiple of 1/256 (the 8 fractional bits
 * the CUDA programming guide gives for linear filtering); the value is the exact weighted sum
 * of the 8-bit texels rounded once to float: sum(w_i * texel_i) / (65536 * 255). */
/* Where cx_scene_build builds the triangle BVH.
 *   HOST:   binned SAH (32 bins) on the host, collapsed to the requested width, uploaded.
 *   DEVICE: on the GPU from the uploaded vertices: 30-bit Morton sort of the triangle
 *           centroids, PLOC clustering (Meister and Bittner 2018, radius 64), SAH leaf
 *           formation, collapse to the requested width; deterministic.  A tree deeper than
 *           the traversal stack is replaced by the host build
 *           (cx_scene_info.bvhBuilder then reports HOST). */
typedef enum cx_bvh_builder {
    CX_BVH_BUILD_HOST = 0,
    CX_BVH_BUILD_DEVICE = 1,
} cx_bvh_builder;

typedef struct cx_texture {
    const uint8_t* rgba;           /* host, [height][width][4] */
    uint32_t width;
    uint32_t height;
} cx_texture;

/*
 * Scene description, host memory.  The reference's Model/Mesh/ModelInstance hierarchy
 * (src/Scene.cpp:73-174) is flattened into one instance with an identity transform,
 * because rehydration always reads transforms[0] (include/graph/Interaction.h:143).
 * Each triangle carries the dense id of its material (the MeshSbtData callable indices of
 * include/model/Mesh.h:10-16 become material ids).  Emitters are the explicit emitters of
 * Scene::BuildEmitterTable (src/Scene.cpp:176-204): world-space triangles and one colour
 * each, sampled by NEE.  An emitter's triangles are traced only if they are also listed as
 * geometry, as the reference's scenes do (implicit emitters are mesh instances,
 * src/Scene.cpp:184-197; an explicit Emitter wraps the mesh of a traced instance,
 * src/main.cpp:164-167).  triMask then keeps them out of occlusion queries: a shadow ray aimed
 * at a point of a traced light would otherwise find the light itself at t = tmax.
 */
/* Visibility of a triangle (cx_scene_desc.triMask), the per-instance visibilityMask of an
 * OptiX IAS reduced to two bits: closest-hit (surface) rays see triangles with bit 0 set,
 * occlusion rays those with bit 1 set (the reference traces with mask 1 and gives every
 * instance 1, include/graph/trace.h:25-88). */
enum {
    CX_VIS_SURFACE = 1,
    CX_VIS_OCCLUSION = 2,
    CX_VIS_ALL = 3
};
typedef struct cx_scene_desc {
    const float* positions;        /* host, [numVertices][3], finite (CX_ERR_INVALID_ARG otherwise) */
    const float* normals;          /* host, [numVertices][3] */
    uint32_t numVertices;
    const uint32_t* indices;       /* host, [numTriangles][3] */
    const uint32_t* triMaterial;   /* host, [numTriangles] */
    uint32_t numTriangles;
    const cx_material* materials;  /* host, [numMaterials] */
    uint32_t numMaterials;
    const float* emitterFaces;     /* host, [numEmitterFaces][3 vertices][3] world space */
    const uint32_t* emitterFaceOffset; /* host, [numEmitters + 1] prefix offsets into faces */
    const float* emitterColor;     /* host, [numEmitters][3] */
    uint32_t numEmitters;
    uint32_t bvhMaxLeafSize;       /* most triangles per leaf, 1..15; 0 = default (2) */
    uint32_t bvhWidth;             /* children per BVH node: 2, 4 or 8; 0 = default (4).  A wide
                                      BVH whose worst-case traversal stack would not fit falls
                                      back to 2 (cx_scene_info.bvhWidth reports the width used) */
    const float* uvs;              /* host, [numVertices][2] texture coordinates (Vertex::uv,
                                      src/Model.cpp:79-92), or NULL: all (0, 0) */
    const cx_material_tex* materialTex; /* host, [numMaterials], or NULL: no textures and
                                      CX_SPEC_TEX_COLOR for every material */
    const cx_texture* textures;    /* host, [numTextures] */
    uint32_t numTextures;
    uint32_t bvhBuilder;           /* cx_bvh_builder: where the BVH is built (the traversal's
                                      answers are the same for either tree) */
    const uint8_t* triMask;        /* host, [numTriangles] CX_VIS_* bits of each triangle, or NULL:
                                      every triangle CX_VIS_ALL.  Emitter geometry is given
                                      CX_VIS_SURFACE: surface rays hit lights, shadow rays ignore
                                      them */
} cx_scene_desc;

typedef struct cx_scene_info {
    uint32_t numTriangles;
    uint32_t numVertices;
    uint32_t numNodes;
    uint32_t maxDepth;
    uint32_t numEmitters;
    uint32_t numEmitterFaces;
    float sceneSize;               /* Scene::GetSceneSize, src/Scene.cpp:48-71 */
    float boundsMin[3];
    float boundsMax[3];
    uint64_t deviceBytes;
    uint32_t bvhWidth;             /* width the traversal uses */
    uint32_t numWideNodes;
    uint32_t stackBound;           /* worst-case traversal stack entries */
    uint32_t numTextures;
    uint32_t bvhBuilder;           /* cx_bvh_builder used */
    float buildMs;                 /* wall time of the BVH build and its upload (host build) or of
                                      the device build, ms */
} cx_scene_info;

/* Per-pass timing of the last frame, ms (src/Renderer.cpp:396-402), plus ray counts. */
typedef struct cx_frame_stats {
    float passMs[CX_PASS_COUNT];
    float totalMs;
    uint64_t closestRays;
    uint64_t shadowRays;
    uint64_t passClosestRays[CX_PASS_COUNT];
    uint64_t passShadowRays[CX_PASS_COUNT];
    /* PrefixAccelStruct::Stats (include/graph/restir/PrefixAccelStruct.h:32-44) */
    uint32_t minNeighCount;
    uint32_t maxNeighCount;
    uint64_t totalNeighCount;
    uint32_t numEntries;
    /* Ray traversal launches of the wavefront mode (the dominant kernel): their count in the
       frame and their summed duration from HIP events around each launch (timing enabled). */
    uint32_t traceLaunches;
    float traceMs;
    /* Nonzero if the wavefront queues ever failed since the renderer was created: bit 0, an
       append found its queue stripe full (the queues are sized so that it cannot; a diagnostic of
       that sizing); bit 1, a tail kernel ended a job that was still running at its step bound
       (the bound is the longest trace chain of the machine, so it cannot either).  A frame with
       either bit set has dropped work; bench.py refuses its numbers. */
    uint32_t queueOverflow;
    /* Rays traced in place by the tail kernels (cx_renderer_set_wave_rounds), included in the
       counts above; the traversal launches answered the rest. */
    uint64_t tailClosestRays;
    uint64_t tailShadowRays;
} cx_frame_stats;

typedef struct cx_ctx cx_ctx;
typedef struct cx_scene cx_scene;
typedef struct cx_renderer cx_renderer;
typedef struct cx_entries cx_entries;

/* ---- context (replaces InitCuda/InitOptix, src/main.cpp:35-65) ---- */
int cx_ctx_create(int device, cx_ctx** out);
int cx_ctx_destroy(cx_ctx* ctx);
const char* cx_last_error(const cx_ctx* ctx);
int cx_abi_version(void);

/* ---- scene: replaces Mesh::BuildAccel, Model::BuildAccel, Scene::BuildAccel and
 *      Scene::BuildEmitterTable (src/Mesh.cpp:54-135, src/Model.cpp:376-457,
 *      src/Scene.cpp:73-204).  Builds a binary BVH (desc->bvhBuilder: binned SAH on the
 *      host, or PLOC on the device), collapses it to desc->bvhWidth children per node and
 *      keeps it in device memory.  Synchronous. ---- */
int cx_scene_build(cx_ctx* ctx, const cx_scene_desc* desc, cx_scene** out);
int cx_scene_destroy(cx_scene* scene);
int cx_scene_get_info(const cx_scene* scene, cx_scene_info* info);

/* ---- traversal: stands in for optixTrace with the surface / occlusion ray flags
 *      (include/graph/trace.h:25-88, shader/mesh_hit.cu:17-69, shader/miss.cu:8-26).
 *      rays: device, [n][8] = origin xyz, direction xyz, tmin, tmax.
 *      Closest hit writes primitive index (0xFFFFFFFF on a miss), t and the OptiX
 *      barycentrics (u weights vertex 1, v weights vertex 2) over the triangles with
 *      CX_VIS_SURFACE.  Occlusion writes 1 if any triangle with CX_VIS_OCCLUSION is hit in
 *      [tmin, tmax], else 0. ---- */
int cx_trace_closest(cx_ctx* ctx, const cx_scene* scene, const float* d_rays, uint32_t n,
                     uint32_t* d_prim, float* d_t, float* d_u, float* d_v, void* stream);
int cx_trace_occluded(cx_ctx* ctx, const cx_scene* scene, const float* d_rays, uint32_t n,
                      uint32_t* d_occluded, void* stream);

/* ---- host traversal over the same node formats, for checking a build on a machine without a
 *      GPU: builds the BVH of `desc` at `bvhWidth` (as cx_scene_build) on the host and answers
 *      closest-hit (anyHit = 0: prim, t, u, v) or occlusion (anyHit = 1: occluded in prim)
 *      queries with the device traversal code compiled for the CPU.  All pointers are host
 *      memory.  *widthUsed receives the width actually traversed (a wide BVH whose stack
 *      bound does not fit falls back to 2).  No counterpart in the reference. ---- */
int cx_bvh_host_trace(const cx_scene_desc* desc, const float* rays, uint32_t n, int anyHit, uint32_t* prim,
                      float* t, float* u, float* v, uint32_t* widthUsed);

/* ---- prefix-entry search structure: replaces PrefixAccelStruct::Rebuild/BuildGas/
 *      BuildTlas (src/PrefixAccelStruct.cpp:14-221) and the point query of
 *      __intersection__prefix_entry (shader/prefix_search_hit.cu:9-73).
 *      Entry i is a candidate for query q iff valid[i] != 0 and |q - pos[i]| <= radius[i],
 *      where radius[i] is the half-width of the entry's box, ((p+r)-(p-r))/2 in float.
 *      The k nearest candidates by (distance, index) are returned in ascending order. ---- */
int cx_entries_create(cx_ctx* ctx, uint32_t capacity, cx_entries** out);
int cx_entries_destroy(cx_entries* entries);
/* d_pos: device [n][3]; d_radius: device [n] (the requested radius r, before the box
 * round trip); d_valid: device [n] bytes.  sceneMin/sceneMax: host, bounds used for the
 * Morton quantisation (entries outside are clamped, which only affects speed). */
int cx_entries_build(cx_ctx* ctx, cx_entries* entries, const float* d_pos, const float* d_radius,
                     const uint8_t* d_valid, uint32_t n, const float* sceneMin,
                     const float* sceneMax, void* stream);
/* d_query: device [nq][3]; outputs: d_idx/d_dist device [nq][k], d_count device [nq]. */
int cx_entries_query(cx_ctx* ctx, const cx_entries* entries, const float* d_query, uint32_t nq,
                     uint32_t k, uint32_t* d_idx, float* d_dist, uint32_t* d_count, void* stream);
/* No counterpart in the reference: cx_entries_build followed by cx_entries_query in one call that
 * allocates no device memory, for callers that cannot allocate (code under hipGraph capture,
 * serving loops, harnesses that forbid hipMalloc).  The entry records, sort buffers and tree
 * live in a fixed device array of the library (about 8.4 MB per device, held until the process
 * ends) whose address cx_ctx_create takes on its device; the context's device must be current.
 * Arguments mean what they mean for cx_entries_build / cx_entries_query, and the outputs are
 * theirs bit for bit.  n above 65536 (a 256 x 256 frame's entries) returns CX_ERR_INVALID_ARG.
 * Every call on one device shares that array: calls that overlap on one device (on two
 * streams) race, and each call overwrites what the previous one left. */
int cx_entries_search_static(cx_ctx* ctx, const float* d_pos, const float* d_radius, const uint8_t* d_valid,
                             uint32_t n, const float* sceneMin, const float* sceneMax, const float* d_query,
                             uint32_t nq, uint32_t k, uint32_t* d_idx, float* d_dist, uint32_t* d_count,
                             void* stream);

/* ---- renderer: Renderer(w, h, ctx, Camera&, const Scene&) and LaunchFrame
 *      (include/graph/Renderer.h:11-21, src/Renderer.cpp:13-406). ---- */
/* Defaults of src/Renderer.cpp:34-58 (reconMinDistance = 0.01 * sceneSize). */
int cx_default_params(const cx_scene* scene, uint32_t width, uint32_t height,
                      cx_launch_params* out);
int cx_renderer_create(cx_ctx* ctx, const cx_scene* scene, uint32_t width, uint32_t height,
                       cx_renderer** out);
int cx_renderer_destroy(cx_renderer* r);
int cx_renderer_set_params(cx_renderer* r, const cx_launch_params* params);
int cx_renderer_get_params(const cx_renderer* r, cx_launch_params* params);
/* Seeds the per-frame random of src/Renderer.cpp:215 (std::random_device there). */
int cx_renderer_seed(cx_renderer* r, uint64_t seed);
/* Resets all per-pixel state to the values uploaded by the constructor
 * (src/Renderer.cpp:155-199) and the front/back indices to 0/0. */
int cx_renderer_reset(cx_renderer* r, void* stream);
/* One LaunchFrame: d_output is caller-owned device memory, [height][width][3] floats.
 * cameraChanged mirrors Camera::HasChanged (src/Renderer.cpp:213). */
int cx_renderer_frame(cx_renderer* r, const cx_camera* cam, int cameraChanged, float* d_output,
                      void* stream);
/* One optixLaunch of a single raygen program with the renderer's current state and the
 * parameter block `params` (frame fields included), for callers that drive the phase
 * order themselves.  Buffer indices are those of the last cx_renderer_frame or
 * cx_renderer_begin_frame. */
int cx_renderer_begin_frame(cx_renderer* r, const cx_camera* cam, int cameraChanged,
                            float* d_output);
int cx_pass_launch(cx_renderer* r, cx_pass pass, const cx_launch_params* params, void* stream);
/*
 * Execution strategy of the passes; both give identical results.  No counterpart in the
 * reference, whose passes are single OptiX raygen launches.
 *   CX_EXEC_WAVEFRONT (default): path generation and shifts run as state machines whose
 *     rays are answered by separate traversal kernels, several launches per pass, with no
 *     host synchronisation inside a frame.
 *   CX_EXEC_MEGAKERNEL: one kernel per pass, traversal inline.
 */
typedef enum cx_exec_mode {
    CX_EXEC_WAVEFRONT = 0,
    CX_EXEC_MEGAKERNEL = 1
} cx_exec_mode;
int cx_renderer_set_exec_mode(cx_renderer* r, int mode);
int cx_renderer_get_exec_mode(const cx_renderer* r);
/* ---- multi-view batch (no counterpart in the reference, which renders one view) ----
 * Enqueues one frame of each of n renderers (views of any scenes and resolutions) and
 * returns: view i runs cx_renderer_frame(renderers[i], &cams[i], cameraChanged[i], outs[i])
 * on streams[i], or, when streams is NULL, on the i-th stream of the context's pool, so the
 * views' kernels overlap on the GPU.  Each view's output is its plain frame's, bit for bit.
 * cx_batch_sync waits for the pool's streams.  The HIP runtime spreads streams over
 * GPU_MAX_HW_QUEUES hardware queues (default 4); with 16 views, 16 queues overlap them best,
 * and the final-phase side stream (cx_renderer_set_overlap) is best turned off there. */
int cx_batch_frame(cx_ctx* ctx, cx_renderer* const* renderers, const cx_camera* cams, const int* cameraChanged,
                   float* const* outs, uint32_t n, void* const* streams);
int cx_batch_sync(cx_ctx* ctx);

/* Wavefront mode: the grid-stride traversal and step kernels launch at most maxBlocks blocks
 * of 128 threads (default 8192); a tuning knob (crisoptix_amd/tune.py), results unchanged. */
int cx_renderer_set_queue_grid(cx_renderer* r, uint32_t maxBlocks);

/* Wavefront mode: each path or shift machine runs at most `rounds` rounds of one traversal
 * launch and one step launch, then finishes its remaining jobs in one kernel that traces each
 * job's rays in place (the same traversal, so the same answers and ray counts).  `rounds` applies
 * to every machine; 0 runs every round as traversal + step; CX_WAVE_ROUNDS_DEFAULT (the initial
 * setting) gives suffix walks 3 rounds and the other machines 2.  Results are identical for
 * every value; a tuning knob (crisoptix_amd/tune.py). */
#define CX_WAVE_ROUNDS_DEFAULT 0xFFFFFFFFu
int cx_renderer_set_wave_rounds(cx_renderer* r, uint32_t rounds);

/* Device-random frames (off by default).  With enable != 0, each cx_renderer_frame draws its
 * frame random number (the value src/Renderer.cpp:220 takes from the host generator) from a PCG32
 * whose state lives in device memory, seeded here with `seed` (the host generator's seeding), and
 * reads the camera, that number and the frame index from a device parameter block written by a
 * one-thread kernel at the start of the frame.  A frame enqueued inside a caller's stream capture
 * (a hipGraph of a whole optimisation step, say) then draws a new random number on every replay.
 * No counterpart in the reference; for callers that record frames into their own graphs.  The
 * adjoint of a frame reads the same block.  cx_pass_launch keeps the host generator.  A frame
 * enqueued on a stream that the caller is capturing keeps to that stream (no side-stream fork)
 * and must not allocate: run one frame with the same parameters eagerly first; graph mode and
 * timing are refused there (CX_ERR_INVALID_ARG). */
int cx_renderer_set_device_random(cx_renderer* r, int enable, uint64_t seed);

/* Final-phase overlap (on by default; wavefront Conditional ReSTIR): cx_renderer_frame runs the
 * prefix-entry store and the entry-tree build (src/Renderer.cpp:351-366) on a renderer-owned
 * stream beside the final gather's canonical prefix and suffix walks, which do not read them,
 * and joins it before the gather's neighbour query.  Results are identical either way; with the
 * overlap on, the per-pass timings of the store and build passes are part of the final
 * gather's. */
int cx_renderer_set_overlap(cx_renderer* r, int enable);

/* Graph mode (no counterpart in the reference, which launches each pass with optixLaunch):
 * cx_renderer_frame records the frame's launches once into a hipGraph per front-buffer parity
 * and then replays it with one hipGraphLaunch per frame; the camera, the frame's random number
 * and its index reach the replay through a small device buffer written by a one-thread kernel
 * ahead of it.  A graph is recaptured when the parameters (other than random and frameIdx), the
 * output buffer, the execution mode, the queue grid, the round bounds or a buffer allocation
 * change.  Images are identical to the eager frame's.  Frames with timing enabled run eagerly.
 * A frame on the null stream is captured and replayed on a renderer-owned blocking stream. */
int cx_renderer_set_graph_mode(cx_renderer* r, int enable);
/* Number of frame graphs captured so far (a diagnostic for the recapture rules). */
int cx_renderer_graph_captures(const cx_renderer* r);

/* Timings are gathered with hipEvents when enabled (off by default). */
int cx_renderer_enable_timing(cx_renderer* r, int enable);
int cx_renderer_get_stats(cx_renderer* r, cx_frame_stats* stats);
/* Device pointers of per-pixel state, for tests and checkpointing.  Records are chunked
 * structure-of-arrays: chunk c of pixel p is the 16-byte word at base + (c * P + p) * 16
 * (field order in crisoptix_amd/csrc/cx_restir_dev.h). */
typedef struct cx_state_view {
    void* prefixRes[2];      /* 9 chunks per pixel: PrefixPath + reservoir weight/confidence */
    void* suffixRes[2];      /* 7 chunks per pixel: SuffixPath + reservoir weight/confidence */
    void* gbuffer;           /* [P] {rng lo, rng hi, primaryIntValid, canonical-buffer flags} */
    void* neighIdx;          /* [P][k] uint32 neighbour pixel index (PrefixNeighbor::pixelIdx) */
    void* neighDist;         /* [P][k] float distance (PrefixNeighbor::distance) */
    void* neighCount;        /* [P] uint32 */
    void* entryPos;          /* [P] float4 {pos.xyz, radius} of the prefix entries */
    void* entryValid;        /* [P] uint8 */
    uint32_t frontBufferIdx;
    uint32_t backBufferIdx;
    uint32_t prefixChunks;
    uint32_t suffixChunks;
    uint32_t k;
    uint32_t pad0;
} cx_state_view;
int cx_renderer_state(cx_renderer* r, cx_state_view* view);

/* ---- image output: ToneMapping (src/kernel/tonemap.cu:7-39).  d_hdr: device [n] float
 *      linear values (n = 3 * pixels), d_ldr: device [n] bytes, sRGB-encoded, clamped to
 *      [0, 1] and truncated (LinearToSrgb); NaN encodes as 0. ---- */
int cx_tonemap(cx_ctx* ctx, const float* d_hdr, uint8_t* d_ldr, uint32_t n, void* stream);

/* ---- differentiable rendering (no counterpart in the reference, which has none) ----
 * Material and emitter parameters live in two device tables.  The sampling table drives every
 * direction, hit and discrete decision of the frame loop (and is what cx_scene_build
 * uploads); the evaluation table gives the BSDF values, emitter colours and derivatives of
 * cx_renderer_adjoint.  A forward pass sets both; a finite-difference check perturbs only the
 * evaluation table, which replays the same paths at other parameter values. */
enum {
    CX_TABLE_SAMPLING = 1,
    CX_TABLE_EVAL = 2,
};
/* Sets the constants of every material; texture bindings stay as the scene was built. */
int cx_scene_set_materials(cx_scene* scene, const cx_material* materials, uint32_t n, int tables);
/* colors: host, [n][3] emitter radiance (Emitter colour, include/model/Emitter.h:13-55) */
/* Material rows from device memory (`rows`: n cx_material in device memory, enqueued on
 * `stream`); like cx_scene_set_materials otherwise.  For optimisation loops whose parameters stay
 * on the device.  No counterpart in the reference. */
int cx_scene_set_materials_device(cx_scene* scene, const float* rows, uint32_t n, int tables, void* stream);
int cx_scene_set_emitter_colors(cx_scene* scene, const float* colors, uint32_t n, int tables);

/* Path-replay adjoint of the last frame's estimate (ReSTIR PT or the path tracer), with no
 * tape: each pixel re-traces its reservoir-selected path from the reservoir record (the shift
 * that produced it, prefix_reuse.h:10-99, and the source's post-reconnection vertices from its
 * PCG32 state) and accumulates W * dL/dI . df/dtheta with W and the path held fixed (the path
 * tracer: dL/dI . df/dtheta / p).  Call after cx_renderer_frame and before the next one.
 *   d_dLdI     device [P][3] float, dL/d(pixel estimate); may be null when no gradient is asked
 *   d_gAlbedo  device [materials][3] double, accumulated (+=), or null; zero for a material
 *              whose diffuse colour comes from a texture (texels are not parameters)
 *   d_gRough   device [materials] double, accumulated, or null; zero for a material whose
 *              roughness comes from a texture or from CX_SPEC_TEX_OCC_ROUGH_METAL
 *   d_gEmit    device [emitters][3] double, accumulated, or null
 *   d_replay   device [P][3] float, the replayed estimate W * f at the evaluation table, or null
 *   d_status   device [3] uint32, accumulated: pixels replayed, pixels whose replayed f differs
 *              in any bit from the record's (zero when the two tables are equal), pixels
 *              skipped (the replay left the recorded path), or null */
int cx_renderer_adjoint(cx_renderer* r, const float* d_dLdI, double* d_gAlbedo, double* d_gRough, double* d_gEmit,
                        float* d_replay, uint32_t* d_status, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CRISOPTIX_H */
