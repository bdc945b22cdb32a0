#pragma once
// Kernel arguments of the per-frame chain, one struct per kernel, and the launchers that
// run a kernel over many views at once.
//
// A single view's launch passes its struct by value (kernel arguments). A batch keeps one
// struct per view in device memory, rewritten from page-locked host memory every frame
// (so a captured graph replays with new cameras), and launches each kernel once for all
// views: grid-indexed kernels find their view from the prefix of per-view work units
// (`start`, viewCount + 1 entries, start[viewCount] = all units), persistent kernels walk
// every view's queue in turn.

#include "gpu_internal.h"

namespace hvvr {

struct FrustumSetupArgs {
    const SimpleRayFrustum* tileCam;
    const SimpleRayFrustum* blockCam;
    SimpleRayFrustum* tileWorld;
    FrustumPre* tilePre;
    FrustumPre* blockPre;
    uint32_t tileCount, blockCount;
    matrix4x4 cameraToWorld;
    CullPlanes cull;
    // a single view's frame counters, zeroed by the first workgroup (the later kernels of
    // the frame run after this one on the same stream); batches clear theirs separately
    unsigned long long* counters;
    uint32_t counterCount;
};

struct BlockTraverseArgs {
    const BVHNode* nodes;
    const FrustumPre* blockPre;
    StackEntry* stacks;
    uint32_t* stackSize;
};

struct TileListArgs {
    const BVHNode* nodes;
    const FrustumPre* tilePre;
    const StackEntry* blockStacks;
    const uint32_t* blockStackSize;
    uint32_t* counts;
    uint32_t* offsets;
    uint32_t* triOut;
    uint64_t capacity;
    const uint32_t* queue;   // the tiles this launch takes (queued kernels)
    uint32_t* nextQueue;     // where a tile this configuration cannot hold goes
    uint32_t queueCounter;   // counters[] slot holding the length of `queue`
    uint32_t nextCounter;    // and of `nextQueue`
    unsigned long long* counters;
    uint32_t stackLimit;
    uint32_t* recScratch; // per-workgroup record scratch of the queued configurations
    // per tile, the queued configuration (1..4, 5 = the walk) whose capacity the tile's
    // last list on this scene needed; the queued kernels send a tile with a higher hint
    // straight to that queue (every configuration builds the same list, so a hint only
    // saves work)
    uint32_t* tileLevel;
    uint32_t useHints; // 0 on a view's first frame of a scene
    uint32_t tileCount;
    // per tile: [tile] BVH nodes tested, [tileCount + tile] deepest stack (plain stores:
    // one atomic per tile on a shared counter serialises every tile of the launch at one
    // memory channel)
    uint32_t* tileWork;
};

struct TileWalkArgs {
    const BVHNode* nodes;
    const FrustumPre* tilePre;
    const StackEntry* blockStacks;
    const uint32_t* blockStackSize;
    const uint32_t* queue;
    uint32_t* counts;
    uint32_t* offsets;
    uint32_t* triOut;
    uint64_t capacity;
    unsigned long long* counters;
    uint32_t* tileLevel;
    uint32_t* tileWork; // as in TileListArgs
    uint32_t tileCount;
    uint32_t stackLimit; // the view's register-stack limit (sets the configurations' capacity)
};

struct IntersectArgs {
    GBufferSubsample* gbuffer;
    const DirectionalBeam* beams;
    const uint32_t* tileCounts;
    const uint32_t* tileOffsets;
    const uint32_t* triIndices;
    const SimpleRayFrustum* worldTiles;
    const PrecomputedTriangleIntersect* tris;
    uint32_t* tileSurvivors; // per tile: list entries surviving the tile cull
    uint32_t* tileTested;    // per tile: (wave, survivor) tests run (the rest were skipped as occluded)
    // optional per-subsample output [sampleCount * A]: (ray parameter of the nearest hit,
    // FLT_MAX on a miss; barycentrics u, v, w of that hit, 0 on a miss); null = not written
    vector4* hits;
    uint64_t listCapacity;
    matrix4x4 cameraToWorld;
    vector2 offsets[32];
    // depth of field: lens-position table [4][A][128], rank of each tile among the
    // tiles with a non-empty list (its table row is rank % 4), thin lens
    const vector2* lensTable;
    const uint32_t* occRank;
    float lensRadius, focalDistance;
};

struct ResolveArgs {
    uint32_t* sampleResults;
    const GBufferSubsample* gbuffer;
    const DirectionalBeam* beams;
    const uint32_t* tileCounts;
    const PrecomputedTriangleIntersect* tris;
    const PrecomputedTriangleShade* triShade;
    const ShadingVertex* verts;
    const SimpleMaterial* materials;
    matrix4x4 cameraToWorld;
    vector2 offsets[32];
    LightingEnvironment env;
    // depth of field (see IntersectArgs)
    const vector2* lensTable;
    const uint32_t* occRank;
    float lensRadius, focalDistance;
    // foveated path: per sample, the view-axis distance of subsample 0's hit (infinity on
    // a miss; ResolveKernel's TMaxBuffer, resolve.cu:386-457); null = not written
    float* tMax;
    TextureSet tex; // the scene's textures (count 0: none)
};

struct RemapArgs {
    const uint32_t* src;
    const int32_t* remap;
    uint32_t* image;
    uint32_t width, height, split;
};

// one view's arguments (host)
FrustumSetupArgs frustumSetupArgs(const hvvr_gpu_view* v, const matrix4x4& m);
BlockTraverseArgs blockTraverseArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s);
// level 0..4: the standard, large, huge, giant and colossal tile-list configurations
TileListArgs tileListArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, int level);
TileWalkArgs tileWalkArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s);
IntersectArgs intersectArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m);
ResolveArgs resolveArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m);
RemapArgs remapArgs(const hvvr_gpu_view* v);

// work units of one view in each grid-indexed kernel
inline uint32_t frustumSetupUnits(const hvvr_gpu_view* v) { return (v->tileCount + v->blockCount + 127) / 128; }
inline uint32_t remapUnits(const hvvr_gpu_view* v) { return ((v->width + 15) / 16) * ((v->height + 7) / 8); }

// Launch shape of a batch: streams, persistent workgroups of the queued tile-list
// configurations and the walk, and whether grid-indexed tile kernels map consecutive
// workgroups to one XCD (workgroups are dealt to the eight XCDs in turn; with the swizzle,
// XCD x takes a contiguous run of tiles, which share BVH nodes and triangles in its L2).
// None of these changes a result.
struct LaunchShape {
    uint32_t largeGroups = 2560, hugeGroups = 1024, giantGroups = 512, walkGroups = 1024;
    uint32_t xcdSwizzle = 0;
    uint32_t streams = 1; // batches: chains the views are dealt to, each on its own stream
};
// the most persistent workgroups each queued configuration launches (its record scratch
// is sized for them): as many as the chip holds at once; the colossal configuration, which
// takes a CU's whole LDS and a few tiles of a dense scene, a quarter of the CUs
enum : uint32_t { kMaxLargeGroups = 2560, kMaxHugeGroups = 1024, kMaxGiantGroups = 512, kMaxColossalGroups = 64 };
// 32-bit words of record scratch one view (or one batch chain) needs
size_t tileListScratchWords();

// batch launchers; `all` = device array of per-view arguments
void launchClearCountersBatch(uint64_t* const* counters, uint32_t viewCount, uint32_t slots, hipStream_t stream);
void launchFrustumSetupBatch(const FrustumSetupArgs* all, const uint32_t* start, uint32_t viewCount, uint32_t units,
                             hipStream_t stream);
void launchBlockTraverseBatch(const BlockTraverseArgs* all, const uint32_t* start, uint32_t viewCount,
                              uint32_t units, hipStream_t stream);
// all[level * viewCount + view]: the five configurations' arguments
void launchTileListsBatch(const TileListArgs* all, const uint32_t* start, uint32_t viewCount, uint32_t units,
                          const LaunchShape& shape, hipStream_t stream);
void launchTileWalkBatch(const TileWalkArgs* all, uint32_t viewCount, uint32_t maxTiles, const LaunchShape& shape,
                         hipStream_t stream);
void launchIntersectBatch(int msaaRate, const IntersectArgs* all, const uint32_t* start, uint32_t viewCount,
                          uint32_t units, const LaunchShape& shape, hipStream_t stream);
// tex: the scene has textures (the resolve variant that reads them)
void launchResolveBatch(int msaaRate, bool half, bool tex, const ResolveArgs* all, const uint32_t* start,
                        uint32_t viewCount, uint32_t units, const LaunchShape& shape, hipStream_t stream);
void launchRemapBatch(const RemapArgs* all, const uint32_t* start, uint32_t viewCount, uint32_t units,
                      hipStream_t stream);

#ifdef __HIPCC__
// One view's arguments from a batch's argument array, read as constant memory: invariant
// scalar loads, which the register allocator repeats where it needs a field instead of
// keeping the whole struct live (read through a plain pointer, the batch tile-list kernel
// took 121 VGPRs and 4 waves per SIMD where its single-view twin takes 92 and 5).
// (Address spaces exist only in the device compilation; the host pass, which never runs
// this, reads the plain pointer.)
template <class T>
__device__ inline T batchArgs(const T* all, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    using ConstT = const __attribute__((address_space(4))) T;
    return ((ConstT*)all)[v];
#else
    return all[v];
#endif
}

// the view whose work units [start[v], start[v+1]) hold `unit` (wave-uniform search)
__device__ inline uint32_t batchView(const uint32_t* __restrict__ start, uint32_t viewCount, uint32_t unit) {
    uint32_t lo = 0, hi = viewCount;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= unit)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// workgroup b of n, dealt round-robin to 8 XCDs, as the unit it takes: with the swizzle,
// XCD x takes units [x * n / 8, (x + 1) * n / 8) in order
__device__ inline uint32_t xcdUnit(uint32_t b, uint32_t n, uint32_t swizzle) {
    if (!swizzle)
        return b;
    const uint32_t xcd = b & 7u, k = b >> 3;
    const uint32_t q = n >> 3, r = n & 7u; // XCD x holds q + (x < r) units
    const uint32_t first = xcd * q + (xcd < r ? xcd : r);
    return first + k;
}
#endif

} // namespace hvvr
