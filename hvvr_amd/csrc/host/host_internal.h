#pragma once
// Host-side (CPU, one-off per scene / per camera configuration) building blocks:
// BVH construction, scene flattening, sample hierarchy setup and the seeded scene
// generators used by the benchmarks and tests.

#include "hvvr/raycaster.h"
#include "hvvr/types.h"
#include "../common/foveated.h"
#include "../common/subsample.h"

#include <stdint.h>
#include <string>
#include <vector>

namespace hvvr {

// SAH BVH2 -> BVH4 -> depth-first topology; reorders triangles and vertices
// (shared/bvh.cpp:542-556). Returns false (with a message) on malformed meshes.
bool GenerateTopology(MeshData& mesh, std::string* error);

// One model of a scene: mesh + model-to-world transform.
struct SceneModel {
    const MeshData* mesh;
    matrix4x4 modelToWorld;
};

// Flattened scene in the global layout the GPU consumes (raycaster/scene_update.cpp:22-143).
struct FlatScene {
    std::vector<ShadingVertex> verts;         // model space
    std::vector<uint32_t> vertexToTransform;  // model index per vertex
    std::vector<PrecomputedTriangleShade> triShade;
    std::vector<BVHNode> nodes;               // topology filled, bounds zero
    std::vector<SimpleMaterial> materials;
    std::vector<matrix4x4> transforms;        // model to world, one per model
    std::vector<LightUnion> lights;
};
void buildScene(const std::vector<SceneModel>& models, FlatScene& out);

// ---------------------------------------------------------------------------------
// samples (raycaster/samples.h:17-67)
// ---------------------------------------------------------------------------------
static_assert(sizeof(Sample) == 28, "Sample layout");

// Radial barrel distortion r' = r (1 + k1 r^2 + k2 r^4) applied to sample positions in
// centred sample space, extents scaled by the local radial magnification.
void applyBarrelDistortion(std::vector<Sample>& samples, float k1, float k2);

// Camera-space sample hierarchy ready for rendering (raycaster/camera.h:28-47).
using SampleSetup = SampleData;
// camera.cpp:315-319
matrix3x3 sampleToCamera(vector2 viewportLower, vector2 viewportUpper);
// camera.cpp:21-64 + sample_hierarchy.cpp:27-190: the perspective mapping, or the
// spherical section when settings.type says so
void buildSampleSetup(const Sample* samples, uint32_t count, uint32_t split, uint32_t width, uint32_t height,
                      const Sample2Dto3DMappingSettings& settings, SampleSetup& out);

// ---------------------------------------------------------------------------------
// scene and ray dumps (raycaster/render.cpp:28-201); binary forms append 'b' to the name
// ---------------------------------------------------------------------------------
void dumpRaysToRFF(const std::string& filename, const std::vector<SimpleRay>& rays, bool binary);
void dumpSceneToOFF(const std::string& filename, const std::vector<const Model*>& models, bool binary);
void dumpBVH(const std::string& filename, const std::vector<const Model*>& models, bool binary);

// ---------------------------------------------------------------------------------
// polar foveated path (raycaster/foveated.cpp, samples.cpp:44-142, frusta.cu:26-121)
// ---------------------------------------------------------------------------------
// Eye-space polar samples, ring-major: sample r * samplesPerRing + i is azimuth column i
// of ring r. False (with a message) for a pattern outside the supported range.
bool polarFoveatedSamples(const FoveatedSamplePattern& p, EccentricityMap& emap, uint32_t& samplesPerRing,
                          uint32_t& ringCount, std::vector<DirectionalBeam>& out, std::string* error);

struct FoveatedSetup {
    FoveatedSamplePattern pattern;
    EccentricityMap emap;
    float maxEccentricity = 0.0f; // radians
    uint32_t samplesPerRing = 0, ringCount = 0;
    uint32_t validSampleCount = 0, sampleCount = 0, blockCount = 0;
    std::vector<DirectionalBeam> beams;          // [sampleCount], eye space, clustered; padding repeats the last
    std::vector<SimpleRayFrustum> tileFrusta;    // [blockCount * 64], eye space; degenerate past the samples
    std::vector<SimpleRayFrustum> blockFrusta;   // [blockCount]
    std::vector<vector2ui> polarRemap;           // [validSampleCount] polar texel (column, ring) of each sample
};
bool buildFoveatedSetup(const FoveatedSamplePattern& p, FoveatedSetup& out, std::string* error);

// previous-frame state of a foveated camera (camera.h:121-122)
struct FoveatedHistory {
    matrix4x4 worldToEyePrevious = matrix4x4(vector4(0.0f), vector4(0.0f), vector4(0.0f), vector4(0.0f));
    matrix3x3 eyePreviousToSamplePrevious = matrix3x3::identity();
};
struct FoveatedFrame {
    matrix4x4 eyeToWorld;          // beam (eye) space to world
    vector4 cullPlanes[4];         // world space (normal, dist): tiles wholly outside one are culled
    matrix3x3 sampleToEye;         // screen sample space to eye space
    matrix4x4 eyeToSamplePrevious; // eye space to the previous frame's sample space
};
// One frame's transforms from the camera, the gaze direction (camera space) and the
// camera's sample bounds; advances `history`.
void foveatedFrame(const matrix4x4& cameraToWorld, const vector3& eyeDir, const matrix3x3& sampleToCam,
                   vector2 boundsLower, vector2 boundsUpper, FoveatedHistory& history, FoveatedFrame& out);

// ---------------------------------------------------------------------------------
// seeded generators
// ---------------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed, uint64_t stream = 0x9e3779b97f4a7c15ull) {
        state = 0;
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    // uniform double in [0,1) with 53 random bits
    double uniform() {
        uint64_t hi = next(), lo = next();
        return double(((hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
    }
    float uniformf(float lo, float hi) { return float(lo + (hi - lo) * uniform()); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300)
            u1 = 1e-300;
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

struct SceneGenResult {
    std::vector<MeshData> meshes;
    std::vector<matrix4x4> modelToWorld;
    std::vector<transform> transforms; // the same placements as transforms (what .bin files store)
    std::vector<LightUnion> lights;
};
// 64x64-cell heightfield (8192 triangles) + 256 boxes (3072 triangles), per-triangle GGX
// materials, one directional light.
void generateHeightfieldBoxesScene(uint64_t seed, SceneGenResult& out);
// meshCount seeded meshes of 256 triangles (subdivided boxes and cylinders) in a
// 30 x 15 x 30 m room, 4 point lights + 1 directional light.
void generateRoomScene(uint64_t seed, uint32_t meshCount, SceneGenResult& out);

// Camera presets for the benchmark scenes.
struct CameraPreset {
    uint32_t width, height;
    vector2 viewportLower, viewportUpper;
    float barrelK1, barrelK2;
    transform cameraToWorld;
    ThinLens lens;
};
// pinhole 640x480, 60 degree vertical FOV looking at the heightfield
CameraPreset heightfieldCamera();
// one eye of a 1080x1200 stereo pair with barrel distortion (k1 0.22, k2 0.24), eye = 0 or 1
CameraPreset roomEyeCamera(uint32_t eye, uint32_t width, uint32_t height);

} // namespace hvvr

// sets the message hvvr_host_last_error() returns on this thread
void hvvrSetHostError(const std::string& msg);

// the host C API's handles (include/hvvr_host.h), shared with the serving loop
struct hvvr_host_view {
    hvvr::SampleSetup setup;
    hvvr::matrix4x4 cameraToWorld = hvvr::matrix4x4::identity();
    hvvr::ThinLens lens = {0.0f, 1.0f};
    hvvr::matrix3x3 sampleToCam = hvvr::matrix3x3::identity();
};

struct hvvr_host_foveated {
    hvvr::FoveatedSetup setup;
};
