#ifndef HVVR_HOST_H
#define HVVR_HOST_H
/*
 * C interface to the host-side (CPU, one-off) half of the library: seeded scene
 * generation, BVH construction, scene flattening and camera sample setup. It has no
 * counterpart in the reference, which is a static C++ library; it exists so test
 * harnesses and benchmarks written in other languages can build the exact inputs the
 * C++ API (include/hvvr/raycaster.h) builds, and hand them to the GPU C ABI
 * (include/hvvr_gpu.h).
 *
 * All functions return 0 on success and a negative code on failure;
 * hvvr_host_last_error() describes the last failure on the calling thread.
 */
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hvvr_host_scene hvvr_host_scene;
typedef struct hvvr_host_view hvvr_host_view;

enum {
    HVVR_SCENE_HEIGHTFIELD_BOXES = 1, /* 64x64 heightfield + 256 boxes, 11,264 triangles */
    HVVR_SCENE_ROOM = 2               /* meshCount x 256-triangle meshes in a 30x15x30 m room */
};
enum {
    HVVR_SCENE_VERTS = 0,           /* ShadingVertex[verts], model space */
    HVVR_SCENE_VERTEX_TRANSFORM = 1,/* uint32[verts] */
    HVVR_SCENE_TRI_SHADE = 2,       /* PrecomputedTriangleShade[tris] */
    HVVR_SCENE_NODES = 3,           /* BVHNode[nodes], topology only */
    HVVR_SCENE_MATERIALS = 4,       /* SimpleMaterial[materials] */
    HVVR_SCENE_TRANSFORMS = 5,      /* float[16] column-major per model */
    HVVR_SCENE_LIGHTS = 6           /* LightUnion[lights], 56 bytes each */
};

const char* hvvr_host_last_error(void);

int hvvr_host_scene_create(int kind, uint64_t seed, uint32_t meshCount, hvvr_host_scene** out);
/* Build a scene from raw meshes: meshes[i] uses vertices [vertexStart[i], vertexStart[i+1])
 * and triangles [triStart[i], triStart[i+1]) with mesh-local vertex indices; material of each
 * triangle indexes materials[] of the scene (materials are not duplicated per mesh). */
int hvvr_host_scene_create_from_meshes(uint32_t meshCount, const uint32_t* vertexStart, const void* verts,
                                       const uint32_t* triStart, const uint32_t* triIndices,
                                       const uint32_t* triMaterial, const float* transforms16,
                                       uint32_t materialCount, const void* materials, uint32_t lightCount,
                                       const void* lights, hvvr_host_scene** out);
/* Scene files in the reference's binary format (include/hvvr/model_import.h,
 * samples_shared/model_import_bin.cpp): save writes the seeded scene of
 * hvvr_host_scene_create (its meshes after GenerateTopology, their transforms, the
 * lights); load flattens a file's meshes with their stored topology, transforms and lights
 * (textures are checked and skipped: shading here is untextured). */
int hvvr_host_scene_save_bin(int kind, uint64_t seed, uint32_t meshCount, const char* path);
int hvvr_host_scene_load_bin(const char* path, hvvr_host_scene** out);
void hvvr_host_scene_destroy(hvvr_host_scene* scene);
/* counts[0..6] = verts, tris, nodes, materials, transforms, lights, models */
int hvvr_host_scene_counts(const hvvr_host_scene* scene, uint64_t counts[7]);
int hvvr_host_scene_copy(const hvvr_host_scene* scene, int array, void* dst, uint64_t bytes);

enum {
    HVVR_VIEW_PRESET_HEIGHTFIELD = 1, /* pinhole 640x480, 60 deg vertical FOV */
    HVVR_VIEW_PRESET_ROOM_EYE = 2     /* one barrel-distorted eye of a stereo pair */
};
enum {
    HVVR_VIEW_BEAMS = 0,        /* DirectionalBeam[sampleCount], camera space */
    HVVR_VIEW_TILE_FRUSTA = 1,  /* SimpleRayFrustum[blockCount*64], camera space */
    HVVR_VIEW_BLOCK_FRUSTA = 2, /* SimpleRayFrustum[blockCount], camera space */
    HVVR_VIEW_REMAP = 3         /* int32[width*height*split] */
};
int hvvr_host_view_create_preset(int preset, uint32_t eye, uint32_t width, uint32_t height, hvvr_host_view** out);
/* a preset's camera: params[8] = viewport lower x y, upper x y, barrel k1 k2, lens radius,
 * focal distance */
int hvvr_host_view_preset_params(int preset, uint32_t eye, uint32_t width, uint32_t height, float params[8]);
/* Camera::getSampleToCamera of a viewport (lower x y, upper x y), column-major 3x3
 * (camera.cpp:167-171) */
int hvvr_host_sample_to_camera(const float viewportLowerUpper[4], float out[9]);
/* samples: Sample[count] (28 bytes each: pixel x,y u32; position x,y f32; minor axis x,y f32;
 * major axis length f32) */
int hvvr_host_view_create_from_samples(const void* samples, uint32_t count, uint32_t split, uint32_t width,
                                       uint32_t height, const float viewportLowerUpper[4], float lensRadius,
                                       float focalDistance, hvvr_host_view** out);
/* As hvvr_host_view_create_from_samples, with rays from a spherical section of
 * fovXDegrees x fovYDegrees when fovXDegrees > 0 (Camera::setSphericalWarpSettings,
 * camera.cpp:206-224; sample_hierarchy.cpp:27-35, 68-80, 144-150); the viewport still sets
 * the sample-to-camera matrix the view reports. */
int hvvr_host_view_create_spherical(const void* samples, uint32_t count, uint32_t split, uint32_t width,
                                    uint32_t height, const float viewportLowerUpper[4], float fovXDegrees,
                                    float fovYDegrees, float lensRadius, float focalDistance, hvvr_host_view** out);
void hvvr_host_view_destroy(hvvr_host_view* view);
/* counts[0..6] = validSampleCount, sampleCount, blockCount, width, height, split, tileCount */
int hvvr_host_view_counts(const hvvr_host_view* view, uint64_t counts[7]);
int hvvr_host_view_copy(const hvvr_host_view* view, int array, void* dst, uint64_t bytes);
/* camera-to-world matrix (column-major float[16]) and lens (radius, focal distance) of a preset */
int hvvr_host_view_camera(const hvvr_host_view* view, float cameraToWorld[16], float lens[2]);

/* Sample space of a view: sampleToCamera (column-major 3x3) and the bounds of the sample
 * positions (lower x, y, upper x, y), camera.cpp:21-42. Either output may be NULL. */
int hvvr_host_view_sample_space(const hvvr_host_view* view, float sampleToCamera[9], float boundsLowerUpper[4]);

/* ---- polar foveated sampling (raycaster/foveated.cpp, samples.cpp:44-142, frusta.cu:26-121) ---- */
typedef struct hvvr_host_foveated hvvr_host_foveated;
enum {
    HVVR_FOVEATED_BEAMS = 0,        /* DirectionalBeam[sampleCount], eye space, clustered */
    HVVR_FOVEATED_TILE_FRUSTA = 1,  /* SimpleRayFrustum[blockCount*64], eye space */
    HVVR_FOVEATED_BLOCK_FRUSTA = 2, /* SimpleRayFrustum[blockCount], eye space */
    HVVR_FOVEATED_POLAR_REMAP = 3   /* uint32[validSampleCount][2]: polar texel (column, ring) */
};
/* pattern: degreeTrackingError, maxFOVDegrees, marSlope, fovealMARDegrees
 * (RayCasterSpecification::FoveatedSamplePattern, raycaster_spec.h:29-34) */
int hvvr_host_foveated_create(const float pattern[4], hvvr_host_foveated** out);
void hvvr_host_foveated_destroy(hvvr_host_foveated* f);
/* counts[0..5] = validSampleCount, sampleCount, blockCount, samplesPerRing, ringCount, tileCount */
int hvvr_host_foveated_counts(const hvvr_host_foveated* f, uint64_t counts[6]);
int hvvr_host_foveated_copy(const hvvr_host_foveated* f, int array, void* dst, uint64_t bytes);
/* eccentricity map (m, w0, S, 1/log(1+m), 1/w0, 1/m; radians) and the maximum eccentricity */
int hvvr_host_foveated_params(const hvvr_host_foveated* f, float emap[6], float* maxEccentricity);
/* One frame's transforms (render.cpp:454-488): from the camera-to-world matrix, the gaze
 * direction in camera space and the camera's sample space, out[57] = eyeToWorld[16],
 * cullPlanes[4][4] (world-space normal, dist), sampleToEye[9], eyeToSamplePrevious[16].
 * history[25] carries worldToEyePrevious[16] (zero at first) and
 * eyePreviousToSamplePrevious[9] (identity at first) from frame to frame. */
int hvvr_host_foveated_frame(const float cameraToWorld[16], const float eyeDir[3], const float sampleToCamera[9],
                             const float boundsLowerUpper[4], float history[25], float out[57]);

/* Grid samples (Sample[width*height]) with optional barrel distortion. */
int hvvr_host_grid_samples(uint32_t width, uint32_t height, float k1, float k2, void* dst);
/* Spiral subsample offsets: out[2*i], out[2*i+1] for i < msaaRate. */
int hvvr_host_subsample_offsets(uint32_t msaaRate, float spinAngle, float* out);
/* Depth-of-field lens table: 4*128*msaaRate float2 entries. */
int hvvr_host_lens_table(uint32_t msaaRate, float* out);
/* Matrix of a transform (translation xyz, quaternion xyzw, scale). */
int hvvr_host_transform_matrix(const float translation[3], const float rotation[4], float scale, float out[16]);

#ifdef __cplusplus
}
#endif
#endif
