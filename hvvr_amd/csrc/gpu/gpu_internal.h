#pragma once
// Internal state behind the C ABI (include/hvvr_gpu.h) and the kernel launchers.

#include "../../../include/hvvr_gpu.h"
#include "hvvr/types.h"
#include "texture.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace hvvr {

void setLastError(const std::string& msg);

#define HVVR_HIP_CHECK(expr)                                                                    \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            setLastError(std::string(#expr) + ": " + hipGetErrorString(_e));                    \
            return HVVR_GPU_ERR_HIP;                                                            \
        }                                                                                       \
    } while (0)

// Owning device allocation.
template <class T>
struct DeviceBuffer {
    T* ptr = nullptr;
    size_t count = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    hipError_t resize(size_t n) {
        if (n == count && ptr)
            return hipSuccess;
        release();
        count = n;
        if (n == 0)
            return hipSuccess;
        return hipMalloc((void**)&ptr, n * sizeof(T));
    }
    size_t bytes() const { return count * sizeof(T); }
};

enum { kStageCount = 7 };
// streams a batched render spreads independent views over (all at default priority)
enum { kViewStreams = 4 };

struct RefitLevel {
    uint32_t start, count;
};

// world-space planes (normal, dist); a tile frustum wholly on the positive side of one
// is culled (render.cpp:271-279, 397-416)
struct CullPlanes {
    vector4 planes[4];
    int enabled;
};

// polar foveated reconstruction of one view (raycaster/gpu_foveated.cu)
struct FoveatedState {
    uint32_t samplesPerRing = 0, ringCount = 0; // polar image: width = azimuth columns, height = rings
    float emap[6] = {};                         // EccentricityMap (m, w0, S, 1/log(1+m), 1/w0, 1/m)
    float maxEccentricity = 0.0f;
    float taaAlpha = 0.05f, taaStddevMultiplier = 1.0f; // TemporalFilterSettings (gpu_camera.h:41-45)
    bool contrast = true;                                // ContrastEnhancementSettings (gpu_camera.cu:180-181)
    float contrastFe = 1.0f;
    uint32_t historyIndex = 0;                           // which history image this frame writes
    DeviceBuffer<vector2ui> polarRemap;                  // [validSampleCount]
    DeviceBuffer<uint32_t> polarColor;                   // [rings][samplesPerRing] RGBA8 sRGB
    DeviceBuffer<float> polarDepth;                      // [rings][samplesPerRing]
    DeviceBuffer<uint64_t> moment1, moment2;             // [rings][samplesPerRing] RGBA16 unorm
    DeviceBuffer<uint32_t> history[2];                   // [height][width] RGBA8 sRGB, this and the last frame
    DeviceBuffer<uint32_t> filtered[2];                  // contrast enhancement: after the first and second pass
    DeviceBuffer<float> srgbToLinear;                    // [256] sRGB byte -> linear
    DeviceBuffer<uint8_t> srgbRoundTrip;                 // [256] byte -> linear -> byte (the polar remap)
};

} // namespace hvvr

struct hvvr_gpu_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // 1: views record per-stage HIP events on the stream and time their standard tile-list
    // kernel's dispatch; 2: only the dispatch timing; 0: neither
    int profiling = 0;
    // batched renders: views fan out from `stream` onto these and join back
    hipStream_t viewStreams[hvvr::kViewStreams] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t joins[hvvr::kViewStreams] = {};
    // device address of the library's shared scratch on this context's GPU
    // (hvvr_gpu_view_render_shared_scratch)
    uint32_t* sharedScratch = nullptr;
};

struct hvvr_gpu_scene {
    hvvr_gpu_context* ctx = nullptr;
    uint32_t vertexCount = 0, triangleCount = 0, nodeCount = 0, materialCount = 0, transformCount = 0;
    hvvr::DeviceBuffer<hvvr::ShadingVertex> vertsModel, vertsWorld;
    hvvr::DeviceBuffer<uint32_t> vertexToTransform;
    hvvr::DeviceBuffer<hvvr::matrix4x4> transforms;
    hvvr::DeviceBuffer<hvvr::PrecomputedTriangleShade> triShade;
    hvvr::DeviceBuffer<hvvr::PrecomputedTriangleIntersect> triIntersect;
    hvvr::DeviceBuffer<hvvr::BVHNode> nodes;
    hvvr::DeviceBuffer<hvvr::SimpleMaterial> materials;
    hvvr::DeviceBuffer<uint32_t> triToNodeChild;
    hvvr::DeviceBuffer<uint32_t> taskChild, taskNodeChild;
    std::vector<hvvr::RefitLevel> levels; // deepest first
    hvvr::LightingEnvironment lighting = {};
    // textures (hvvr_gpu_scene_set_textures): every texture's mip chain in one buffer
    uint32_t textureCount = 0;
    hvvr::DeviceBuffer<uint8_t> texels;
    hvvr::DeviceBuffer<hvvr::TexInfo> texInfo;
    hvvr::DeviceBuffer<float> texLut; // [512]: TextureSet::srgbLut, then lodLut
    std::vector<hvvr::TexInfo> texInfoHost;
    std::vector<uint64_t> texBase, texBytes; // each texture's chain: start and size in `texels`
    bool dirty = true;
    bool everUpdated = false;
    hipEvent_t ready = nullptr; // recorded on the context stream after each update
    ~hvvr_gpu_scene() {
        if (ready)
            (void)hipEventDestroy(ready);
    }
};

struct hvvr_gpu_view {
    hvvr_gpu_context* ctx = nullptr;
    uint32_t msaaRate = 1;
    uint32_t sampleCount = 0, validSampleCount = 0, tileCount = 0, blockCount = 0;
    uint32_t width = 0, height = 0, split = 1;
    hvvr::ThinLens lens = {0.0f, 1.0f};
    bool dof = false;                                   // lens.radius > 0
    bool jitterSpin = false;
    bool halfShading = false;                           // GGX colour terms in packed FP16
    hvvr::vector2 offsets[32];
    hvvr::matrix4x4 lastCameraToWorld = hvvr::matrix4x4::identity();

    hvvr::DeviceBuffer<hvvr::DirectionalBeam> beams;
    hvvr::DeviceBuffer<hvvr::SimpleRayFrustum> tileFrustaCam, blockFrustaCam, tileFrustaWorld;
    hvvr::DeviceBuffer<hvvr::FrustumPre> tilePre, blockPre;
    hvvr::DeviceBuffer<hvvr::StackEntry> blockStacks;   // [blockCount][64]
    hvvr::DeviceBuffer<uint32_t> blockStackSize;        // [blockCount]
    hvvr::DeviceBuffer<uint32_t> recScratch;            // record node indices and emission order of the
                                                        // queued tile-list configurations, per workgroup
    hvvr::DeviceBuffer<uint32_t> tileLevel;             // [tileCount] tile-list configuration hint
    hvvr::DeviceBuffer<uint32_t> tileWork;              // [4][tileCount] per tile, last frame: BVH nodes
                                                        // tested by the tile traversal, its deepest stack,
                                                        // triangles surviving intersect's tile cull,
                                                        // (wave, survivor) tests intersect ran
    const hvvr_gpu_scene* hintScene = nullptr;          // the scene the hints were last written for
    hvvr::DeviceBuffer<uint32_t> tileQueues;            // [5][tileCount] tiles queued for the large, huge,
                                                        // giant and colossal LDS configurations and the walk
    hvvr::DeviceBuffer<uint32_t> tileCounts;            // [tileCount] list length
    hvvr::DeviceBuffer<uint32_t> tileOffsets;           // [tileCount] list start in triIndices
    hvvr::DeviceBuffer<uint32_t> triIndices;            // list storage, tiles in allocation order
    hvvr::DeviceBuffer<hvvr::GBufferSubsample> gbuffer; // [sampleCount * A], wave-interleaved
    hvvr::DeviceBuffer<uint32_t> sampleResults;         // [sampleCount]
    hvvr::DeviceBuffer<int32_t> remap;                  // [width*height*split]
    hvvr::DeviceBuffer<uint32_t> image;                 // [width*height]
    hvvr::DeviceBuffer<hvvr::vector2> lensTable;        // depth of field: [4][A][128] lens positions
    hvvr::DeviceBuffer<uint32_t> occRank;               // depth of field: rank among occupied tiles
    hvvr::DeviceBuffer<uint64_t> counters;              // [0] survivors, [1] overflow flag, [2] max tile stack,
                                                        // [3] list entries allocated, [4] node visits,
                                                        // [5..8] tiles queued for the large, huge and
                                                        // giant configurations and for the walk; [0], [2]
                                                        // and [4] are reduced from tileWork on request
    uint64_t listCapacity = 0;
    uint32_t waveStackLimit = 64; // register stack entries per tile traversal (<= 64)
    // foveated path (hvvr_gpu_view_set_foveated): cull planes and the t-max buffer apply to
    // foveated renders only
    bool foveated = false;
    hvvr::CullPlanes cull = {};
    hvvr::DeviceBuffer<float> tMax; // [sampleCount]
    // per-subsample hit output (hvvr_gpu_view_set_hit_output): [sampleCount * A]
    // (t, u, v, w), written by intersect while enabled
    bool hitOutput = false;
    bool hitsValid = false; // the last frame wrote `hits`
    hvvr::DeviceBuffer<hvvr::vector4> hits;
    hvvr::FoveatedState fov;
    hipEvent_t stageEvents[hvvr::kStageCount + 1] = {};
    bool stagesRecorded = false;
    // start and end of the standard tile-list kernel's dispatch (profiling levels 1 and 2):
    // recorded by the dispatch itself, no extra packets on the stream
    hipEvent_t kernelEvents[2] = {};
    bool kernelTimed = false;
    bool rendered = false;
    // own-stream mode (hvvr_gpu_view_set_own_stream): renders and reads go to `ownStream`,
    // after the scene's `ready` event; `done` marks the end of the last queued work
    hipStream_t ownStream = nullptr;
    hipEvent_t done = nullptr;
    hipStream_t lastStream = nullptr; // the stream of the last render
    ~hvvr_gpu_view() {
        for (auto& ev : stageEvents)
            if (ev)
                (void)hipEventDestroy(ev);
        for (auto& ev : kernelEvents)
            if (ev)
                (void)hipEventDestroy(ev);
        if (done)
            (void)hipEventDestroy(done);
        if (ownStream)
            (void)hipStreamDestroy(ownStream);
    }
};

namespace hvvr {

// Per-frame scratch of the tile-list kernels taken from outside the view (the shared
// scratch of hvvr_gpu_view_render_shared_scratch): the queued configurations' record
// scratch, `groups[L]` persistent workgroups' worth for the large, huge, giant and
// colossal configurations (L = 0..3), and the five tile queues, [5][tileCount].
struct TileListScratch {
    uint32_t* rec;
    uint32_t* queues;
    uint32_t groups[4];
};
// the persistent workgroups of the large, huge, giant and colossal configurations whose
// record scratch fits `words` 32-bit words (at most as many as a view's own scratch holds)
void tileListGroupsFor(size_t words, uint32_t groups[4]);

// hvvr_gpu_view_create; with `deviceSetup` the sample hierarchy is allocated, not uploaded
int makeView(hvvr_gpu_context* ctx, const hvvr_gpu_view_desc* d, bool deviceSetup, hvvr_gpu_view** out);

// kernel launchers (all asynchronous on `stream`)
void launchSceneRefit(hvvr_gpu_scene* s, hipStream_t stream);
void launchFrustumSetup(hvvr_gpu_view* v, const matrix4x4& m, hipStream_t stream);
void launchBlockTraverse(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream);
// `scratch` null: the view's own record scratch and tile queues; `timing` non-null: the
// standard kernel's dispatch records its start and end in timing[0] and timing[1]
void launchTileLists(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream,
                     const TileListScratch* scratch = nullptr, hipEvent_t* timing = nullptr);
void launchTileListFallback(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream,
                            const TileListScratch* scratch = nullptr);
void launchOccupiedRank(hvvr_gpu_view* v, hipStream_t stream);
// counters[0], [2] and [4] from the view's per-tile work of its last frame
void launchTileWorkReduce(hvvr_gpu_view* v, hipStream_t stream);
void launchIntersect(hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m, hipStream_t stream);
void launchResolve(hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m, hipStream_t stream);
void launchRemap(hvvr_gpu_view* v, hipStream_t stream);
// foveated reconstruction: polar remap, moments, reprojection + temporal filter, contrast
// enhancement; `sampleToEye` and `eyeToSamplePrevious` as in hvvr_gpu_foveated_frame
void launchFoveatedReconstruct(hvvr_gpu_view* v, const matrix3x3& sampleToEye, const matrix4x4& eyeToSamplePrevious,
                               hipStream_t stream);

} // namespace hvvr
