#ifndef HVVR_GPU_H
#define HVVR_GPU_H
/*
 * Thin C ABI over the gfx950 (MI355X) HIP implementation of HVVR's per-frame render
 * path. Everything the reference splits between its CPU traversal and its CUDA
 * kernels runs on the GPU behind these entry points: per-tile frustum setup, block
 * and tile frustum-BVH traversal, triangle-list intersection, GBuffer compression,
 * GGX resolve, remap, and the scene refit.
 *
 * Plain pointers and sizes only. Struct arrays passed as `const void*` use the
 * reference's byte layouts (ShadingVertex 32 B, PrecomputedTriangleShade 16 B,
 * BVHNode 128 B, SimpleMaterial 64 B, LightUnion 56 B, DirectionalBeam 36 B,
 * SimpleRayFrustum 96 B, matrices column-major float[16]).
 *
 * Errors: every function returns HVVR_GPU_OK (0) or a negative HVVR_GPU_ERR_* code;
 * hvvr_gpu_last_error() returns a description of the last failure on the calling
 * thread. (The reference has no error returns: cutilSafeCall prints and exits,
 * raycaster/cuda_util.h:18-29.)
 *
 * Threading: a context, and the scenes and views created from it, are used from one
 * host thread at a time. Work is issued asynchronously on the context's HIP stream;
 * read-back functions synchronise that stream.
 */
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HVVR_GPU_OK = 0,
    HVVR_GPU_ERR_INVALID = -1,   /* bad argument */
    HVVR_GPU_ERR_HIP = -2,       /* a HIP runtime call failed */
    HVVR_GPU_ERR_NO_DEVICE = -3, /* no GPU / no gfx950 device */
    HVVR_GPU_ERR_OVERFLOW = -4,  /* a traversal stack or triangle-list budget was exceeded */
    HVVR_GPU_ERR_STATE = -5      /* call out of order (e.g. render before a scene update) */
};

typedef struct hvvr_gpu_context hvvr_gpu_context;
typedef struct hvvr_gpu_scene hvvr_gpu_scene;
typedef struct hvvr_gpu_view hvvr_gpu_view;

const char* hvvr_gpu_last_error(void);
/* Library version string, e.g. "hvvr_gpu 0.1 gfx950". Never fails. */
const char* hvvr_gpu_version(void);

/* ---- context: replaces GPUContext::cudaInit / cleanup (raycaster/gpu_context.h:27-42) ---- */
int hvvr_gpu_context_create(int device, hvvr_gpu_context** out);
int hvvr_gpu_context_destroy(hvvr_gpu_context* ctx);
int hvvr_gpu_context_synchronize(hvvr_gpu_context* ctx);
/* 1 = every render records per-stage HIP events on the stream (no synchronisation), read
 * back with hvvr_gpu_view_stage_times, and times its standard tile-list kernel's dispatch
 * (hvvr_gpu_view_tile_list_kernel_time); 2 = only the dispatch timing, which adds no packet
 * to the stream and leaves the frame's own timing alone; 0 = off. Other values fail. */
int hvvr_gpu_context_set_profiling(hvvr_gpu_context* ctx, int enable);

/* ---- scene: replaces GPUSceneState (raycaster/gpu_scene_state.h:31-73) ---- */
typedef struct {
    const void* vertices;               /* ShadingVertex[vertexCount], model space */
    const uint32_t* vertexToTransform;  /* [vertexCount], index into transforms */
    uint32_t vertexCount;
    const void* triangles;              /* PrecomputedTriangleShade[triangleCount], global vertex indices */
    uint32_t triangleCount;             /* <= 2^26 (GBuffer sort-key limit, intersect.cu:52-71) */
    const void* nodes;                  /* BVHNode[nodeCount]: topology only, bounds are computed here */
    uint32_t nodeCount;
    const void* materials;              /* SimpleMaterial[materialCount] */
    uint32_t materialCount;
    const float* transforms;            /* [transformCount][16] model-to-world */
    uint32_t transformCount;
} hvvr_gpu_scene_desc;

/* GPUSceneState::setGeometry + updateMaterials (gpu_scene_state.cu:469-599, :83-85).
 * Uploads the scene and derives the refit plan; bounds are computed by the next update. */
int hvvr_gpu_scene_create(hvvr_gpu_context* ctx, const hvvr_gpu_scene_desc* desc, hvvr_gpu_scene** out);
int hvvr_gpu_scene_destroy(hvvr_gpu_scene* scene);
/* GPUSceneState::animate (gpu_scene_state.cu:601-606): new model-to-world matrices. */
int hvvr_gpu_scene_set_transforms(hvvr_gpu_scene* scene, const float* transforms, uint32_t count);
/* Replace model-space vertices [first, first+count) (per-vertex animation; no reference
 * counterpart, the reference only animates per-model transforms). */
int hvvr_gpu_scene_set_vertices(hvvr_gpu_scene* scene, uint32_t first, uint32_t count, const void* vertices);
/* GPUSceneState::updateLighting (gpu_scene_state.cu:47-81): LightUnion[count]; lights past
 * 1 directional / 4 point / 4 spot are dropped and counted in *dropped (may be NULL). */
int hvvr_gpu_scene_set_lights(hvvr_gpu_scene* scene, const void* lights, uint32_t count, uint32_t* dropped);
/* GPUSceneState::update (gpu_scene_state.cu:377-451): vertex transform, triangle precompute
 * and bottom-up BVH refit, issued on the context stream when the scene is dirty. */
int hvvr_gpu_scene_update(hvvr_gpu_scene* scene);
/* CreateTexture (texture.cu:210-272) for every texture of the scene at once, replacing any
 * earlier set; texture i is the index materials name (SimpleMaterial texture fields). Level
 * 0 is converted (8-bit formats kept as RGBA8, the rest as RGBA32F; a one-channel texture
 * reads (r, 0, 0, 1)) and the mip chain built on the device. Shading samples them with
 * wrap addressing, trilinear filtering and up to 8 anisotropic probes
 * (hvvr_amd/csrc/gpu/texture.h has the exact arithmetic). Synchronous. */
typedef struct {
    uint32_t format; /* TextureFormat: 1 r8g8b8a8_unorm_srgb, 2 r8g8b8a8_unorm, 3 r16g16b16a16_unorm,
                        4 r32g32b32a32_float, 5 r16g16b16a16_float, 6 r11g11b10_float, 7 r32_float */
    uint32_t width, height, strideElements;
    const void* data; /* strideElements x height texels of the format */
} hvvr_gpu_texture_desc;
int hvvr_gpu_scene_set_textures(hvvr_gpu_scene* scene, uint32_t count, const hvvr_gpu_texture_desc* textures);
/* info[4] = width, height, mip levels, stored kind (0 RGBA8 unorm, 1 RGBA8 sRGB, 2 RGBA32F);
 * *bytes = size of the texture's stored chain: level l at the sum of the earlier levels'
 * sizes, each rounded up to 16 bytes */
int hvvr_gpu_scene_texture_info(hvvr_gpu_scene* scene, uint32_t index, uint32_t info[4], uint64_t* bytes);
int hvvr_gpu_scene_read_texture(hvvr_gpu_scene* scene, uint32_t index, void* chain);
/* GPUSceneState::fetchUpdatedBVH (gpu_scene_state.cu:454-467) and friends: synchronous reads. */
int hvvr_gpu_scene_read_nodes(hvvr_gpu_scene* scene, void* nodes);          /* BVHNode[nodeCount] */
int hvvr_gpu_scene_read_triangles(hvvr_gpu_scene* scene, void* triangles);  /* PrecomputedTriangleIntersect[] */
int hvvr_gpu_scene_read_vertices(hvvr_gpu_scene* scene, void* vertices);    /* world-space ShadingVertex[] */

/* ---- view (one camera): replaces GPUCamera (raycaster/gpu_camera.h:100-228) ---- */
typedef struct {
    uint32_t msaaRate;              /* 1, 2, 4, 8, 16 or 32 (COLOR_MODE_MSAA_RATE, raycaster_common.h:27) */
    uint32_t sampleCount;           /* padded sample count, a multiple of 8192 */
    uint32_t validSampleCount;
    const void* beams;              /* DirectionalBeam[sampleCount], camera space */
    const void* tileFrusta;         /* SimpleRayFrustum[sampleCount/128], camera space */
    const void* blockFrusta;        /* SimpleRayFrustum[sampleCount/8192], camera space */
    const int32_t* remap;           /* [width*height*split], sample index or -1; NULL = no remap
                                       (a foveated view, whose image is its reconstruction) */
    uint32_t width, height, split;  /* output image; split = colour samples per pixel (1..3) */
    float lensRadius, focalDistance;/* ThinLens: radius 0 = pinhole; radius > 0 = depth of field
                                       (IntersectTriangleThreadDoF, prim_tests.h:247-392), focal > 0 */
} hvvr_gpu_view_desc;

/* GPUCamera::updateConfig (gpu_camera.cu:149-192). */
int hvvr_gpu_view_create(hvvr_gpu_context* ctx, const hvvr_gpu_view_desc* desc, hvvr_gpu_view** out);

/* ---- sample setup on the device: Camera::setSamples (camera.cpp:21-64, 135-156) ---- */
typedef struct {
    uint32_t msaaRate;
    uint32_t width, height, split;  /* render target the remap addresses; colour samples per pixel */
    float sampleToCamera[9];        /* column-major 3x3: Camera::getSampleToCamera (camera.cpp:167-171) */
    float lensRadius, focalDistance;/* ThinLens */
    float fovXDegrees, fovYDegrees; /* fovXDegrees > 0: rays from a spherical section of this size
                                       (Camera::setSphericalWarpSettings, camera.cpp:206-224) */
} hvvr_gpu_sample_setup_desc;
/* A view whose sample hierarchy is built on the device from samples (Sample[count], 28 B
 * each, host memory): naiveXYCluster (sample_clustering.h:17-49) as stable sorts, then beams,
 * padded tile and block bounds, frusta (sample_hierarchy.cpp:46-190) and the remap
 * (camera.cpp:57-63); equal to the host setup bit for bit for the perspective mapping.
 * boundsOut (may be NULL): float[4] sample bounds, lower x y, upper x y (camera.cpp:36-42). */
int hvvr_gpu_view_create_from_samples(hvvr_gpu_context* ctx, const hvvr_gpu_sample_setup_desc* desc,
                                      const void* samples, uint32_t count, hvvr_gpu_view** out, float* boundsOut);
/* The same for the width x height grid (getGridSamples, samples.cpp:21-42) generated on the
 * device, with a radial barrel distortion r' = r (1 + k1 r^2 + k2 r^4) when k1 or k2 is
 * nonzero. split must be 1. */
int hvvr_gpu_view_create_grid(hvvr_gpu_context* ctx, const hvvr_gpu_sample_setup_desc* desc, float k1, float k2,
                              hvvr_gpu_view** out, float* boundsOut);
/* A view's sample hierarchy: DirectionalBeam[sampleCount], SimpleRayFrustum[tileCount],
 * SimpleRayFrustum[blockCount], int32[width*height*split] (any may be NULL). Synchronous. */
int hvvr_gpu_view_read_sample_setup(hvvr_gpu_view* view, void* beams, void* tileFrusta, void* blockFrusta,
                                    int32_t* remap);
/* shape[8] = msaaRate, sampleCount, validSampleCount, tileCount, blockCount, width, height, split */
int hvvr_gpu_view_get_shape(hvvr_gpu_view* view, uint32_t shape[8]);
int hvvr_gpu_view_destroy(hvvr_gpu_view* view);
/* Raycaster::intersectAndResolveBeamBatch + GPUCamera::intersectShadeResolve + remap
 * (render.cpp:418-451, gpu_camera.cu:271-293, remap.cu:79-124): one frame of one view.
 * cameraToWorld: float[16]; jitter: float[2] (frame jitter, used when the subsample
 * spin is enabled, see hvvr_gpu_view_set_jitter_spin). Asynchronous. */
int hvvr_gpu_view_render(hvvr_gpu_view* view, hvvr_gpu_scene* scene, const float* cameraToWorld,
                         const float* jitter);
/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that forbid hipMalloc): hvvr_gpu_view_render with
 * the tile-list kernels' per-frame scratch (the queued configurations' record scratch and
 * the five tile queues) taken from one 16 MiB device array that the library holds per GPU
 * for the life of the process, instead of the view's own buffers. Allocates nothing; the
 * results are those of hvvr_gpu_view_render. Views of more than 65,536 tiles (8,388,608
 * samples) are rejected (HVVR_GPU_ERR_INVALID). Every call on one GPU uses the same array:
 * calls whose work overlaps on the device (views on different streams) race, so serialise
 * them (one stream, or wait for the previous call's frame). */
int hvvr_gpu_view_render_shared_scratch(hvvr_gpu_view* view, hvvr_gpu_scene* scene, const float* cameraToWorld,
                                        const float* jitter);
/* Several views of one scene in one submission (batched multi-view). */
int hvvr_gpu_render_views(hvvr_gpu_scene* scene, uint32_t count, hvvr_gpu_view* const* views,
                          const float* cameraToWorld /* [count][16] */, const float* jitter /* [count][2] or NULL */);
/* Batches: many views of one scene with each kernel of the frame launched once for all
 * of them (the reference renders its cameras one after another, render.cpp:204-219). The
 * views must be pinhole grid views (no depth of field, not foveated) of the batch's
 * context and outlive the batch; any sizes, MSAA rates and shading precisions mix. After
 * a batch render every view reads back as after its own render. */
typedef struct hvvr_gpu_batch hvvr_gpu_batch;
typedef struct {
    uint32_t largeGroups, hugeGroups, giantGroups; /* persistent workgroups of the queued tile-list configurations */
    uint32_t walkGroups;                           /* ... and of the LDS-stack walk */
    uint32_t xcdSwizzle;                           /* 1: each XCD takes a contiguous run of tiles */
    uint32_t streams;                              /* 1..4 chains the views are dealt to, each on its own
                                                      stream (forked and joined inside the frame) */
} hvvr_gpu_launch_shape;
int hvvr_gpu_batch_create(hvvr_gpu_context* ctx, hvvr_gpu_view* const* views, uint32_t count, hvvr_gpu_batch** out);
int hvvr_gpu_batch_destroy(hvvr_gpu_batch* batch);
/* 1: capture the frame (argument upload, optional refit, the kernel chain) as a HIP graph
 * on the next render and replay it on later ones (captured again when the views' MSAA
 * rates, shading precisions or sizes, the scene, the refit setting or the launch shape
 * change). */
int hvvr_gpu_batch_set_graph(hvvr_gpu_batch* batch, int enable);
/* 1: every frame refits the scene's BVH first, inside the batch's chain (animated
 * vertices or transforms set through hvvr_gpu_scene_set_*). 0 (default): the scene is
 * updated as for single-view renders, when dirty. */
int hvvr_gpu_batch_set_refit(hvvr_gpu_batch* batch, int everyFrame);
/* The launch shape changes how work is scheduled, never a result. */
int hvvr_gpu_batch_set_launch_shape(hvvr_gpu_batch* batch, const hvvr_gpu_launch_shape* shape);
int hvvr_gpu_batch_get_launch_shape(hvvr_gpu_batch* batch, hvvr_gpu_launch_shape* shape);
/* One frame of every view; cameraToWorld: [count][16]. Asynchronous; a render waits for
 * the batch's previous frame. */
int hvvr_gpu_batch_render(hvvr_gpu_batch* batch, hvvr_gpu_scene* scene, const float* cameraToWorld);
/* Waits for the last frame; *gpuMs (may be NULL) = its device time. */
int hvvr_gpu_batch_synchronize(hvvr_gpu_batch* batch, float* gpuMs);
/* Tries launch shapes one field at a time (coordinate descent), framesPerTrial frames
 * each (median device time), keeps the fastest and leaves it set; a candidate whose
 * per-view work counters differ from the first fails the call. */
int hvvr_gpu_batch_autotune(hvvr_gpu_batch* batch, hvvr_gpu_scene* scene, const float* cameraToWorld,
                            uint32_t framesPerTrial, hvvr_gpu_launch_shape* best, float* bestMs);

/* Tile traversal keeps its distance-sorted stack in registers, up to `entries` (1..64,
 * default 64; it holds only nodes that still have work). Tiles that need more (or more
 * than 768 visited nodes, or that meet a NaN distance) go to larger configurations (2048,
 * 4608 and 10240 nodes; 256, 512 and 1024 stack entries, scaled down with a limit below
 * 64) and what none holds to a walk with the reference's full-depth stack (4096 entries,
 * TileFrame::stackSize, traverse_avx.h:281) in LDS. The result is the same either way;
 * the limit only moves work between the paths. */
int hvvr_gpu_view_set_traversal_stack_limit(hvvr_gpu_view* view, uint32_t entries);
/* 32 (default): FP32 shading, the oracle's arithmetic. 16: the GGX colour terms of each
 * light in packed FP16 (the scalar D / visibility / Fresnel weights and the radiance sum
 * stay FP32); error stated against the FP64 oracle. */
int hvvr_gpu_view_set_shading_precision(hvvr_gpu_view* view, int bits);
/* Per-subsample depth and barycentrics. The reference keeps each subsample's nearest hit
 * distance in IntersectSamples' sampleTMax (intersect.cu:117-168) and computes barycentrics
 * with calcUVW (prim_tests.h:233-245, thin lens :366-392) only at shading time; neither is
 * read back. 1 = the following frames also write, for every subsample, the ray parameter
 * t of its nearest hit and the barycentrics of that hit on the same ray (allocates
 * sampleCount * msaaRate * 16 bytes on first use). */
int hvvr_gpu_view_set_hit_output(hvvr_gpu_view* view, int enable);
/* JITTER_SAMPLES (kernel_constants.h:15): 1 = rotate the subsample spiral by jitter.x*6.28 */
int hvvr_gpu_view_set_jitter_spin(hvvr_gpu_view* view, int enable);

/* ---- polar foveated rendering (RayCasterGPUMode::GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT,
 * render.cpp:454-488, gpu_foveated.cu, remap.cu:127-158) ----
 * A foveated view is created with eye-space polar beams and frusta (hvvr_host_foveated_*)
 * and remap NULL; width x height is the reconstruction target. Each frame traces the
 * polar samples, scatters them into the polar image (ring x azimuth), and reconstructs
 * the screen image with a bicubic polar lookup, a reprojected and neighbourhood-clamped
 * temporal blend, and eccentricity-dependent contrast enhancement. */
typedef struct {
    const uint32_t* polarRemap;     /* [validSampleCount][2]: polar texel (column, ring) of each sample */
    uint32_t samplesPerRing, ringCount; /* samplesPerRing * ringCount == validSampleCount */
    float eccentricityMap[6];       /* m, w0, S, 1/log(1+m), 1/w0, 1/m (hvvr_host_foveated_params) */
    float maxEccentricity;          /* radians */
    float taaAlpha;                 /* TemporalFilterSettings::alpha (gpu_camera.h:41-45), 0.05 */
    float taaStddevMultiplier;      /* TemporalFilterSettings::stddevMultiplier, 1 */
    int contrastEnhancement;        /* ContrastEnhancementSettings::enable (gpu_camera.cu:180-181), 1 */
    float contrastFe;               /* ContrastEnhancementSettings::f_e, 1 */
} hvvr_gpu_foveated_desc;
int hvvr_gpu_view_set_foveated(hvvr_gpu_view* view, const hvvr_gpu_foveated_desc* desc);
/* Per-frame transforms, as hvvr_host_foveated_frame computes them. */
typedef struct hvvr_gpu_foveated_frame {
    float eyeToWorld[16];          /* eye (beam) space to world; the frame's camera-to-world */
    float cullPlanes[16];          /* 4 world-space planes (normal xyz, dist); tiles wholly on the
                                      positive side of one are skipped; dist[0] = +inf: no culling */
    float sampleToEye[9];          /* screen sample space to eye space (column-major 3x3) */
    float eyeToSamplePrevious[16]; /* eye space to the previous frame's sample space */
} hvvr_gpu_foveated_frame;
/* One foveated frame (asynchronous). The reconstruction is read with
 * hvvr_gpu_view_read_image. Stage [6] of the stage times covers the polar remap and the
 * reconstruction. */
int hvvr_gpu_view_render_foveated(hvvr_gpu_view* view, hvvr_gpu_scene* scene, const hvvr_gpu_foveated_frame* frame,
                                  const float* jitter);
enum {
    HVVR_GPU_FOVEATED_POLAR_COLOR = 0, /* uint32 RGBA8 sRGB [ringCount][samplesPerRing] */
    HVVR_GPU_FOVEATED_POLAR_DEPTH = 1, /* float [ringCount][samplesPerRing] t-max (inf = miss) */
    HVVR_GPU_FOVEATED_MOMENT1 = 2,     /* uint64 RGBA16 unorm [ringCount][samplesPerRing] */
    HVVR_GPU_FOVEATED_MOMENT2 = 3,
    HVVR_GPU_FOVEATED_FILTERED = 4,    /* uint32 [height][width]: this frame's temporal filter output */
    HVVR_GPU_FOVEATED_PREVIOUS = 5,    /* uint32 [height][width]: the previous frame's */
    HVVR_GPU_FOVEATED_BLUR1 = 6,       /* uint32 [height][width]: contrast blur, first pass (along y) */
    HVVR_GPU_FOVEATED_BLUR2 = 7,       /* uint32 [height][width]: second pass (along x) */
    HVVR_GPU_FOVEATED_TMAX = 8         /* float [sampleCount] per-sample t-max */
};
/* Synchronous read of a foveated intermediate of the last frame; bytes must match. */
int hvvr_gpu_view_read_foveated(hvvr_gpu_view* view, int buffer, void* dst, uint64_t bytes);

/* ---- overlapping views (the serving loop) ----
 * With its own stream a view renders on a private non-blocking HIP stream, ordered after
 * the scene's last update, so several views' frames overlap on the device. Image reads
 * can then be queued behind the frame and polled. */
int hvvr_gpu_view_set_own_stream(hvvr_gpu_view* view, int enable);
/* Queue a copy of the last frame's image to host memory (pinned memory from
 * hvvr_gpu_host_alloc for a truly asynchronous copy); needs the view's own stream. */
int hvvr_gpu_view_read_image_async(hvvr_gpu_view* view, uint32_t* image, uint32_t strideInPixels);
/* 1 when everything queued on the view's own stream up to its last
 * hvvr_gpu_view_read_image_async has finished, 0 when not yet, < 0 on error. */
int hvvr_gpu_view_query(hvvr_gpu_view* view);
/* Device memory the view holds, in bytes; and what a view of this description (and
 * foveated description, or NULL) would hold, before creating it. */
int hvvr_gpu_view_device_bytes(hvvr_gpu_view* view, uint64_t* bytes);
int hvvr_gpu_view_estimate_bytes(const hvvr_gpu_view_desc* desc, const hvvr_gpu_foveated_desc* fov, uint64_t* bytes);
/* Triangle-list storage. The reference holds a fixed MAX_TRI_INDICES_TO_INTERSECT
 * (raycaster/magic_constants.h:19) and stops with a message when a frame needs more; a
 * view here starts at 1024 entries per tile and can grow. After a frame has finished
 * (waits for it): *entries = the list entries it needed, *capacity = what the view holds,
 * *overflowed = 1 when the frame overflowed its lists or a tile's traversal stack (its
 * reads then fail with HVVR_GPU_ERR_OVERFLOW). *entries > *capacity: reserve *entries and
 * render the frame again; otherwise the overflow was a traversal stack's, which more list
 * storage does not cure. */
int hvvr_gpu_view_list_demand(hvvr_gpu_view* view, uint64_t* entries, uint64_t* capacity, int* overflowed);
/* Grows the list storage to `entries` (at most 2^32 - 1; never shrinks); waits for the
 * view's work first. */
int hvvr_gpu_view_reserve_lists(hvvr_gpu_view* view, uint64_t entries);
/* Undoes the reconstruction history's advance of the last foveated frame, so that the
 * same frame can be rendered again (after its lists grew) from the same history. */
int hvvr_gpu_view_foveated_rewind(hvvr_gpu_view* view);
/* Page-locked host memory (NULL on failure). */
void* hvvr_gpu_host_alloc(uint64_t bytes);
void hvvr_gpu_host_free(void* ptr);

/* Synchronous reads of the last rendered frame. */
int hvvr_gpu_view_read_image(hvvr_gpu_view* view, uint32_t* image, uint32_t strideInPixels);
int hvvr_gpu_view_read_sample_colors(hvvr_gpu_view* view, uint32_t* colors); /* [sampleCount] RGBA8 */
/* GBuffer in logical order: for sample s, slots s*A .. s*A+A-1 as (triIndex, sampleMask),
 * ascending triangle id, misses (0xffffffff) last, unused slots (0xffffffff, 0). */
int hvvr_gpu_view_read_gbuffer(hvvr_gpu_view* view, uint32_t* triIndex, uint32_t* sampleMask);
/* hits[(sample * A + i) * 4 + k], k = 0: t of subsample i's nearest hit (the ray is
 * origin + t * (centre + du * alpha.x + dv * alpha.y), unnormalised; FLT_MAX on a miss);
 * k = 1..3: barycentrics (u, v, w) of that hit, weights of the triangle's vertices 0, 1, 2
 * (0 on a miss). The last frame must have been rendered with the hit output on. */
int hvvr_gpu_view_read_hits(hvvr_gpu_view* view, float* hits);
/* Per-tile triangle lists of the last frame in tile order: counts[tileCount]; lists
 * concatenated into triIndices (capacity `capacity` entries); *total = list entries. */
/* GPUCamera::dumpRays + DumpRaysKernel (gpu_camera.cu, intersect.cu:407-458): every
 * subsample's world-space ray for camera-to-world `cameraToWorld` (float[16]), as SimpleRay
 * {float3 origin, float3 direction} (24 B), sample s's subsample k at s*msaaRate + k. Sample
 * order, or with scanlineOrder the image order of the remap (locations without a sample
 * get NaN rays). origin = the camera position (the reference's thin-lens origin is a TODO,
 * so depth of field is not applied). capacity: rays the buffer holds (validSampleCount
 * x msaaRate needed). Synchronous. */
int hvvr_gpu_view_dump_rays(hvvr_gpu_view* view, const float* cameraToWorld, int scanlineOrder, void* rays,
                            uint64_t capacity);
int hvvr_gpu_view_read_tile_lists(hvvr_gpu_view* view, uint32_t* counts, uint32_t* triIndices, uint64_t capacity,
                                  uint64_t* total);
typedef struct {
    uint64_t listEntries;      /* triangle-list entries over all tiles */
    uint64_t survivorPairs;    /* sum over tiles of 128 x triangles surviving the tile culls */
    uint32_t occupiedTiles;
    uint32_t tileCount;
    uint32_t maxTileStack;     /* deepest tile traversal stack seen */
    uint32_t overflow;         /* nonzero if a budget overflowed */
    uint64_t nodeVisits;       /* BVH nodes tested by the tile traversal (4 child boxes each) */
    uint32_t largeTiles;       /* tiles the standard LDS configuration passed to the large one */
    uint32_t hugeTiles;        /* ... the large one to the huge one */
    uint32_t giantTiles;       /* ... the huge one to the giant one */
    uint32_t walkTiles;        /* ... the colossal one to the LDS-stack walk */
    uint64_t testedPairs;      /* (sample, survivor) pairs the intersect tested: survivorPairs less
                                  the wave-level tests skipped because no ray of the wave could
                                  reach the triangle in front of its current hits */
    uint32_t colossalTiles;    /* tiles the giant configuration passed to the colossal one */
} hvvr_gpu_view_stats;
int hvvr_gpu_view_get_stats(hvvr_gpu_view* view, hvvr_gpu_view_stats* stats);
/* Per-stage GPU milliseconds of the view's last frame rendered with profiling enabled
 * (waits for that frame's last event):
 * [0] frustum setup, [1] block traversal, [2] tile lists, [3] tile lists of tiles that
 * did not fit in LDS, [4] intersect, [5] resolve, [6] remap (foveated: polar remap and
 * reconstruction). Returns the count. */
int hvvr_gpu_view_stage_times(hvvr_gpu_view* view, float* ms, int maxCount);
/* No counterpart in the reference. GPU milliseconds of the view's standard tile-list kernel
 * (its one launch over every tile, without the queued configurations) in the last frame
 * rendered with profiling level 1 or 2, from HIP events the kernel's dispatch records
 * (waits for them). */
int hvvr_gpu_view_tile_list_kernel_time(hvvr_gpu_view* view, float* ms);

#ifdef __cplusplus
}
#endif
#endif
