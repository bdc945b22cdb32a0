// Batched multi-view rendering (include/hvvr_gpu.h, hvvr_gpu_batch_*).
//
// The reference renders its enabled cameras one after another (render.cpp:204-219), each
// camera's frame a sequence of kernel launches. A batch renders many views of one scene
// with every kernel of the frame launched once for all of them: frustum setup, block
// traversal, the four tile-list configurations and the walk, then intersect and resolve
// once per (MSAA rate, shading precision) group, and the remap. Each view's kernel
// arguments live in device memory (launch_args.h), filled every frame from one
// page-locked block with one copy, so the frame can be captured once as a HIP graph
// (copy, optional BVH refit, counter clear, the chain) and replayed with new cameras.
// The counters are cleared by a kernel, not a memset, so the captured graph holds none.
// The launch shape (persistent workgroup counts, XCD tile swizzle) is the autotuner's
// to pick; none of it changes a result.

#include "gpu_internal.h"
#include "launch_args.h"

#include <algorithm>
#include <map>
#include <string.h>

using namespace hvvr;

namespace hvvr {

enum : uint32_t { kCounterSlots = 16 };

__global__ void clearCountersBatchKernel(uint64_t* const* __restrict__ counters, uint32_t n, uint32_t slots) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * slots)
        counters[i / slots][i % slots] = 0ull;
}

void launchClearCountersBatch(uint64_t* const* counters, uint32_t viewCount, uint32_t slots, hipStream_t stream) {
    const uint32_t n = viewCount * slots;
    clearCountersBatchKernel<<<(n + 255) / 256, 256, 0, stream>>>(counters, viewCount, slots);
}

} // namespace hvvr

namespace {

struct Group {
    uint32_t msaaRate = 1;
    bool half = false;
    std::vector<uint32_t> members; // view indices
    uint32_t units = 0;            // tiles
    size_t intersectOff = 0, resolveOff = 0, startOff = 0;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

} // namespace

struct hvvr_gpu_batch {
    hvvr_gpu_context* ctx = nullptr;
    std::vector<hvvr_gpu_view*> views;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr, sceneFence = nullptr;
    LaunchShape shape;
    bool graphEnabled = false, refitEveryFrame = false, inFlight = false;
    // the frame's structure: groups, units and the argument-block layout
    std::vector<Group> groups;
    std::vector<uint32_t> remapMembers;
    uint32_t uFrustum = 0, uBlock = 0, uTile = 0, uRemap = 0, maxTiles = 0;
    size_t offCounters = 0, offFrustum = 0, offBlock = 0, offTile = 0, offWalk = 0, offRemap = 0;
    size_t sFrustum = 0, sBlock = 0, sTile = 0, sRemap = 0, argBytes = 0;
    std::vector<uint64_t> signature; // what the layout and a captured graph were built for
    uint8_t* hostArgs = nullptr;     // page-locked
    DeviceBuffer<uint8_t> devArgs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graphSignature;

    ~hvvr_gpu_batch() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        for (hvvr_gpu_view* v : views)
            if (v->lastStream == stream)
                v->lastStream = nullptr;
        dropGraph();
        if (hostArgs)
            (void)hipHostFree(hostArgs);
        if (start)
            (void)hipEventDestroy(start);
        if (stop)
            (void)hipEventDestroy(stop);
        if (sceneFence)
            (void)hipEventDestroy(sceneFence);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
    void dropGraph() {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graphSignature.clear();
    }
};

namespace {

int batchFail(const std::string& msg, int code = HVVR_GPU_ERR_INVALID) {
    setLastError(msg);
    return code;
}

#define BATCH_HIP(expr)                                                                                           \
    do {                                                                                                          \
        hipError_t _e = (expr);                                                                                   \
        if (_e != hipSuccess)                                                                                     \
            return batchFail(std::string(#expr) + ": " + hipGetErrorString(_e), HVVR_GPU_ERR_HIP);                \
    } while (0)

// everything the launches depend on besides the per-frame arguments
std::vector<uint64_t> frameSignature(const hvvr_gpu_batch* b, const hvvr_gpu_scene* s) {
    std::vector<uint64_t> sig = {uint64_t(uintptr_t(s)), uint64_t(b->refitEveryFrame), b->shape.largeGroups,
                                 b->shape.hugeGroups,    b->shape.giantGroups,          b->shape.walkGroups,
                                 b->shape.xcdSwizzle};
    for (const hvvr_gpu_view* v : b->views) {
        sig.push_back(v->msaaRate);
        sig.push_back(v->halfShading ? 1 : 0);
        sig.push_back(v->remap.ptr ? 1 : 0);
        sig.push_back(v->tileCount);
    }
    return sig;
}

// groups, unit counts and the argument-block layout
int planLayout(hvvr_gpu_batch* b) {
    const uint32_t n = uint32_t(b->views.size());
    std::map<std::pair<uint32_t, bool>, Group> byKind;
    b->remapMembers.clear();
    b->uFrustum = b->uBlock = b->uTile = b->uRemap = b->maxTiles = 0;
    for (uint32_t i = 0; i < n; i++) {
        const hvvr_gpu_view* v = b->views[i];
        Group& g = byKind[{v->msaaRate, v->halfShading}];
        g.msaaRate = v->msaaRate;
        g.half = v->halfShading;
        g.members.push_back(i);
        g.units += v->tileCount;
        b->uFrustum += frustumSetupUnits(v);
        b->uBlock += v->blockCount;
        b->uTile += v->tileCount;
        b->maxTiles = std::max(b->maxTiles, v->tileCount);
        if (v->remap.ptr) {
            b->remapMembers.push_back(i);
            b->uRemap += remapUnits(v);
        }
    }
    b->groups.clear();
    for (auto& kv : byKind)
        b->groups.push_back(kv.second);
    size_t at = 0;
    auto take = [&](size_t bytes) {
        size_t off = at;
        at = align256(at + bytes);
        return off;
    };
    b->offCounters = take(n * sizeof(uint64_t*));
    b->offFrustum = take(n * sizeof(FrustumSetupArgs));
    b->offBlock = take(n * sizeof(BlockTraverseArgs));
    b->offTile = take(4 * n * sizeof(TileListArgs));
    b->offWalk = take(n * sizeof(TileWalkArgs));
    for (Group& g : b->groups) {
        g.intersectOff = take(g.members.size() * sizeof(IntersectArgs));
        g.resolveOff = take(g.members.size() * sizeof(ResolveArgs));
    }
    b->offRemap = take(b->remapMembers.size() * sizeof(RemapArgs) + 1);
    b->sFrustum = take((n + 1) * 4);
    b->sBlock = take((n + 1) * 4);
    b->sTile = take((n + 1) * 4);
    for (Group& g : b->groups)
        g.startOff = take((g.members.size() + 1) * 4);
    b->sRemap = take((b->remapMembers.size() + 1) * 4);
    if (at > b->argBytes || !b->hostArgs) {
        if (b->hostArgs)
            (void)hipHostFree(b->hostArgs);
        b->hostArgs = nullptr;
        BATCH_HIP(hipHostMalloc((void**)&b->hostArgs, at, hipHostMallocDefault));
        hipError_t e = b->devArgs.resize(at);
        if (e != hipSuccess)
            return batchFail(std::string("batch argument block: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
        b->argBytes = at;
        b->dropGraph(); // its copy node names the old buffers
    }
    return HVVR_GPU_OK;
}

// this frame's arguments into the page-locked block
void fillArgs(hvvr_gpu_batch* b, const hvvr_gpu_scene* s, const matrix4x4* m) {
    const uint32_t n = uint32_t(b->views.size());
    uint8_t* h = b->hostArgs;
    auto counters = reinterpret_cast<uint64_t**>(h + b->offCounters);
    auto frustum = reinterpret_cast<FrustumSetupArgs*>(h + b->offFrustum);
    auto block = reinterpret_cast<BlockTraverseArgs*>(h + b->offBlock);
    auto tile = reinterpret_cast<TileListArgs*>(h + b->offTile);
    auto walk = reinterpret_cast<TileWalkArgs*>(h + b->offWalk);
    auto sF = reinterpret_cast<uint32_t*>(h + b->sFrustum);
    auto sB = reinterpret_cast<uint32_t*>(h + b->sBlock);
    auto sT = reinterpret_cast<uint32_t*>(h + b->sTile);
    sF[0] = sB[0] = sT[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        hvvr_gpu_view* v = b->views[i];
        v->cull.enabled = 0;
        counters[i] = v->counters.ptr;
        frustum[i] = frustumSetupArgs(v, m[i]);
        block[i] = blockTraverseArgs(v, s);
        for (int level = 0; level < 4; level++)
            tile[level * n + i] = tileListArgs(v, s, level);
        walk[i] = tileWalkArgs(v, s);
        sF[i + 1] = sF[i] + frustumSetupUnits(v);
        sB[i + 1] = sB[i] + v->blockCount;
        sT[i + 1] = sT[i] + v->tileCount;
    }
    for (const Group& g : b->groups) {
        auto ia = reinterpret_cast<IntersectArgs*>(h + g.intersectOff);
        auto ra = reinterpret_cast<ResolveArgs*>(h + g.resolveOff);
        auto st = reinterpret_cast<uint32_t*>(h + g.startOff);
        st[0] = 0;
        for (size_t k = 0; k < g.members.size(); k++) {
            const hvvr_gpu_view* v = b->views[g.members[k]];
            ia[k] = intersectArgs(v, s, m[g.members[k]]);
            ra[k] = resolveArgs(v, s, m[g.members[k]]);
            st[k + 1] = st[k] + v->tileCount;
        }
    }
    auto rm = reinterpret_cast<RemapArgs*>(h + b->offRemap);
    auto sR = reinterpret_cast<uint32_t*>(h + b->sRemap);
    sR[0] = 0;
    for (size_t k = 0; k < b->remapMembers.size(); k++) {
        const hvvr_gpu_view* v = b->views[b->remapMembers[k]];
        rm[k] = remapArgs(v);
        sR[k + 1] = sR[k] + remapUnits(v);
    }
}

// the frame's work on the batch stream (eagerly, or into a capture)
void enqueueFrame(hvvr_gpu_batch* b, hvvr_gpu_scene* s) {
    const uint32_t n = uint32_t(b->views.size());
    hipStream_t st = b->stream;
    uint8_t* d = b->devArgs.ptr;
    (void)hipMemcpyAsync(d, b->hostArgs, b->argBytes, hipMemcpyHostToDevice, st);
    if (b->refitEveryFrame)
        launchSceneRefit(s, st);
    launchClearCountersBatch(reinterpret_cast<uint64_t* const*>(d + b->offCounters), n, kCounterSlots, st);
    auto u32 = [&](size_t off) { return reinterpret_cast<const uint32_t*>(d + off); };
    launchFrustumSetupBatch(reinterpret_cast<const FrustumSetupArgs*>(d + b->offFrustum), u32(b->sFrustum), n,
                            b->uFrustum, st);
    launchBlockTraverseBatch(reinterpret_cast<const BlockTraverseArgs*>(d + b->offBlock), u32(b->sBlock), n,
                             b->uBlock, st);
    launchTileListsBatch(reinterpret_cast<const TileListArgs*>(d + b->offTile), u32(b->sTile), n, b->uTile, b->shape,
                         st);
    launchTileWalkBatch(reinterpret_cast<const TileWalkArgs*>(d + b->offWalk), n, b->maxTiles, b->shape, st);
    for (const Group& g : b->groups) {
        const uint32_t gn = uint32_t(g.members.size());
        launchIntersectBatch(int(g.msaaRate), reinterpret_cast<const IntersectArgs*>(d + g.intersectOff),
                             u32(g.startOff), gn, g.units, b->shape, st);
        launchResolveBatch(int(g.msaaRate), g.half, reinterpret_cast<const ResolveArgs*>(d + g.resolveOff),
                           u32(g.startOff), gn, g.units, b->shape, st);
    }
    if (!b->remapMembers.empty())
        launchRemapBatch(reinterpret_cast<const RemapArgs*>(d + b->offRemap), u32(b->sRemap),
                         uint32_t(b->remapMembers.size()), b->uRemap, st);
}

int renderBatch(hvvr_gpu_batch* b, hvvr_gpu_scene* s, const float* m16) {
    if (!b || !s || !m16)
        return batchFail("null argument");
    if (s->ctx != b->ctx)
        return batchFail("scene and batch belong to different contexts");
    for (const hvvr_gpu_view* v : b->views)
        if (v->dof || v->foveated)
            return batchFail("batched views must be pinhole grid views (no depth of field, not foveated)");
    BATCH_HIP(hipSetDevice(b->ctx->device));
    // the previous frame has read the argument block before it is rewritten
    if (b->inFlight)
        BATCH_HIP(hipStreamSynchronize(b->stream));
    b->inFlight = false;
    // work already queued on the context stream (scene uploads and updates, single-view
    // renders of this scene) comes first
    BATCH_HIP(hipEventRecord(b->sceneFence, b->ctx->stream));
    BATCH_HIP(hipStreamWaitEvent(b->stream, b->sceneFence, 0));
    if (!b->refitEveryFrame && (!s->everUpdated || s->dirty)) {
        int rc = hvvr_gpu_scene_update(s);
        if (rc)
            return rc;
        BATCH_HIP(hipStreamWaitEvent(b->stream, s->ready, 0));
    }
    std::vector<uint64_t> sig = frameSignature(b, s);
    if (sig != b->signature) {
        int rc = planLayout(b);
        if (rc)
            return rc;
        b->signature = sig;
    }
    std::vector<matrix4x4> m(b->views.size());
    memcpy(m.data(), m16, m.size() * sizeof(matrix4x4));
    fillArgs(b, s, m.data());
    BATCH_HIP(hipEventRecord(b->start, b->stream));
    if (b->graphEnabled) {
        if (!b->exec || b->graphSignature != sig) {
            b->dropGraph();
            BATCH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            enqueueFrame(b, s);
            hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
            if (e != hipSuccess)
                return batchFail(std::string("graph capture: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
            e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                b->dropGraph();
                return batchFail(std::string("graph instantiation: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
            }
            b->graphSignature = sig;
        }
        BATCH_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
        enqueueFrame(b, s);
        BATCH_HIP(hipGetLastError());
    }
    BATCH_HIP(hipEventRecord(b->stop, b->stream));
    if (b->refitEveryFrame) {
        s->dirty = false;
        s->everUpdated = true;
        BATCH_HIP(hipEventRecord(s->ready, b->stream));
    }
    for (size_t i = 0; i < b->views.size(); i++) {
        hvvr_gpu_view* v = b->views[i];
        v->lastCameraToWorld = m[i];
        v->rendered = true;
        v->stagesRecorded = false;
        v->lastStream = b->stream;
    }
    b->inFlight = true;
    return HVVR_GPU_OK;
}

int syncBatch(hvvr_gpu_batch* b, float* ms) {
    BATCH_HIP(hipStreamSynchronize(b->stream));
    b->inFlight = false;
    if (ms)
        BATCH_HIP(hipEventElapsedTime(ms, b->start, b->stop));
    return HVVR_GPU_OK;
}

} // namespace

extern "C" {

int hvvr_gpu_batch_create(hvvr_gpu_context* ctx, hvvr_gpu_view* const* views, uint32_t count, hvvr_gpu_batch** out) {
    if (!ctx || !views || !out || count == 0)
        return batchFail("null argument or empty batch");
    for (uint32_t i = 0; i < count; i++) {
        if (!views[i] || views[i]->ctx != ctx)
            return batchFail("every batched view must belong to the batch's context");
        for (uint32_t j = 0; j < i; j++)
            if (views[j] == views[i])
                return batchFail("a view appears twice in the batch");
    }
    hvvr_gpu_batch* b = new hvvr_gpu_batch();
    b->ctx = ctx;
    b->views.assign(views, views + count);
    bool ok = hipSetDevice(ctx->device) == hipSuccess &&
              hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&b->start) == hipSuccess && hipEventCreate(&b->stop) == hipSuccess &&
              hipEventCreateWithFlags(&b->sceneFence, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete b;
        return batchFail("batch stream or event creation failed", HVVR_GPU_ERR_HIP);
    }
    *out = b;
    return HVVR_GPU_OK;
}

int hvvr_gpu_batch_destroy(hvvr_gpu_batch* b) {
    delete b;
    return HVVR_GPU_OK;
}

int hvvr_gpu_batch_set_graph(hvvr_gpu_batch* b, int enable) {
    if (!b)
        return batchFail("null batch");
    if (b->inFlight)
        BATCH_HIP(hipStreamSynchronize(b->stream));
    b->inFlight = false;
    b->graphEnabled = enable != 0;
    if (!b->graphEnabled)
        b->dropGraph();
    return HVVR_GPU_OK;
}

int hvvr_gpu_batch_set_refit(hvvr_gpu_batch* b, int everyFrame) {
    if (!b)
        return batchFail("null batch");
    b->refitEveryFrame = everyFrame != 0;
    return HVVR_GPU_OK;
}

int hvvr_gpu_batch_set_launch_shape(hvvr_gpu_batch* b, const hvvr_gpu_launch_shape* shape) {
    if (!b || !shape)
        return batchFail("null argument");
    auto inRange = [](uint32_t x) { return x >= 1 && x <= 65536; };
    if (!inRange(shape->largeGroups) || !inRange(shape->hugeGroups) || !inRange(shape->giantGroups) ||
        !inRange(shape->walkGroups) || shape->xcdSwizzle > 1)
        return batchFail("launch shape: workgroup counts in [1, 65536], xcdSwizzle 0 or 1");
    b->shape.largeGroups = shape->largeGroups;
    b->shape.hugeGroups = shape->hugeGroups;
    b->shape.giantGroups = shape->giantGroups;
    b->shape.walkGroups = shape->walkGroups;
    b->shape.xcdSwizzle = shape->xcdSwizzle;
    return HVVR_GPU_OK;
}

int hvvr_gpu_batch_get_launch_shape(hvvr_gpu_batch* b, hvvr_gpu_launch_shape* shape) {
    if (!b || !shape)
        return batchFail("null argument");
    shape->largeGroups = b->shape.largeGroups;
    shape->hugeGroups = b->shape.hugeGroups;
    shape->giantGroups = b->shape.giantGroups;
    shape->walkGroups = b->shape.walkGroups;
    shape->xcdSwizzle = b->shape.xcdSwizzle;
    return HVVR_GPU_OK;
}

int hvvr_gpu_batch_render(hvvr_gpu_batch* b, hvvr_gpu_scene* s, const float* cameraToWorld) {
    return renderBatch(b, s, cameraToWorld);
}

int hvvr_gpu_batch_synchronize(hvvr_gpu_batch* b, float* gpuMs) {
    if (!b)
        return batchFail("null batch");
    return syncBatch(b, gpuMs);
}

int hvvr_gpu_batch_autotune(hvvr_gpu_batch* b, hvvr_gpu_scene* s, const float* cameraToWorld,
                            uint32_t framesPerTrial, hvvr_gpu_launch_shape* best, float* bestMs) {
    if (!b || !s || !cameraToWorld || framesPerTrial == 0)
        return batchFail("null argument or no frames per trial");
    // per-view work counters of the current shape: every candidate must reproduce them
    // (the shape only schedules work)
    auto measure = [&](float& ms, std::vector<uint64_t>& work) -> int {
        std::vector<float> t;
        for (uint32_t f = 0; f < framesPerTrial + 1; f++) {
            int rc = renderBatch(b, s, cameraToWorld);
            if (rc)
                return rc;
            float x = 0.0f;
            rc = syncBatch(b, &x);
            if (rc)
                return rc;
            if (f > 0) // the first frame may capture a graph
                t.push_back(x);
        }
        std::sort(t.begin(), t.end());
        ms = t[t.size() / 2];
        work.clear();
        for (const hvvr_gpu_view* v : b->views) {
            uint64_t c[5];
            BATCH_HIP(hipMemcpy(c, v->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
            work.push_back(c[0]); // survivor pairs
            work.push_back(c[3]); // list entries
            work.push_back(c[4]); // node visits
        }
        return HVVR_GPU_OK;
    };
    LaunchShape cur = b->shape;
    float curMs = 0.0f;
    std::vector<uint64_t> baseWork, work;
    int rc = measure(curMs, baseWork);
    if (rc)
        return rc;
    // coordinate descent over the shape's fields
    const std::vector<uint32_t> largeC = {320, 640, 1280, 2560}, hugeC = {256, 512, 1024}, giantC = {128, 256, 512},
                                walkC = {256, 1024, 2048}, swizzleC = {0, 1};
    struct Field {
        uint32_t LaunchShape::*member;
        const std::vector<uint32_t>* values;
    };
    const Field fields[] = {{&LaunchShape::xcdSwizzle, &swizzleC}, {&LaunchShape::largeGroups, &largeC},
                            {&LaunchShape::hugeGroups, &hugeC},    {&LaunchShape::giantGroups, &giantC},
                            {&LaunchShape::walkGroups, &walkC}};
    for (const Field& f : fields) {
        for (uint32_t value : *f.values) {
            if (cur.*(f.member) == value)
                continue;
            LaunchShape trial = cur;
            trial.*(f.member) = value;
            b->shape = trial;
            float ms = 0.0f;
            rc = measure(ms, work);
            if (rc) {
                b->shape = cur;
                return rc;
            }
            if (work != baseWork) {
                b->shape = cur;
                return batchFail("autotune: a launch shape changed the frame's work counters", HVVR_GPU_ERR_STATE);
            }
            if (ms < curMs) {
                cur = trial;
                curMs = ms;
            }
        }
    }
    b->shape = cur;
    if (best)
        rc = hvvr_gpu_batch_get_launch_shape(b, best);
    if (bestMs)
        *bestMs = curMs;
    return rc;
}

} // extern "C"
