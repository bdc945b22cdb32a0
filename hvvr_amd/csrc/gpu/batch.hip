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
// The views can be dealt to up to four chains, each on its own stream (forked from the
// batch's stream and joined back, inside the graph too), so one chain's intersect
// overlaps another's traversal. The launch shape (streams, persistent workgroup counts,
// XCD tile swizzle) is the autotuner's to pick; none of it changes a result.

#include "gpu_internal.h"
#include "launch_args.h"

#include <algorithm>
#include <map>
#include <memory>
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

namespace {

// The views of one stream's chain: their frame structure and argument block.
struct Chain {
    std::vector<hvvr_gpu_view*> views;
    hipStream_t stream = nullptr; // null: the batch's own stream
    std::vector<Group> groups;
    std::vector<uint32_t> remapMembers;
    uint32_t uFrustum = 0, uBlock = 0, uTile = 0, uRemap = 0, maxTiles = 0;
    size_t offCounters = 0, offFrustum = 0, offBlock = 0, offTile = 0, offWalk = 0, offRemap = 0;
    size_t sFrustum = 0, sBlock = 0, sTile = 0, sRemap = 0, argBytes = 0;
    std::vector<uint32_t> index; // position of each view in the batch (its camera)
    uint8_t* hostArgs = nullptr; // page-locked
    DeviceBuffer<uint8_t> devArgs;
    hipEvent_t joined = nullptr;
};

} // namespace

struct hvvr_gpu_batch {
    hvvr_gpu_context* ctx = nullptr;
    std::vector<hvvr_gpu_view*> views;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, stop = nullptr, sceneFence = nullptr, fork = nullptr;
    LaunchShape shape;
    bool graphEnabled = false, refitEveryFrame = false, inFlight = false;
    std::vector<std::unique_ptr<Chain>> chains;
    std::vector<hipStream_t> extraStreams; // chain streams beyond the batch's own
    std::vector<uint64_t> signature;      // what the chains and the captured graphs were built for
    // One graph per chain, replayed on the chain's stream (the chains run side by side;
    // a single graph holding them all ran its branches one after the other), and one for
    // the refit that precedes them.
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    hipGraph_t refitGraph = nullptr;
    hipGraphExec_t refitExec = nullptr;
    std::vector<uint64_t> graphSignature;

    ~hvvr_gpu_batch() {
        if (stream)
            (void)hipStreamSynchronize(stream);
        for (hvvr_gpu_view* v : views)
            if (v->lastStream == stream)
                v->lastStream = nullptr;
        dropGraph();
        dropChains();
        for (hipStream_t st : extraStreams)
            (void)hipStreamDestroy(st);
        if (start)
            (void)hipEventDestroy(start);
        if (stop)
            (void)hipEventDestroy(stop);
        if (sceneFence)
            (void)hipEventDestroy(sceneFence);
        if (fork)
            (void)hipEventDestroy(fork);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
    void dropGraph() {
        for (hipGraphExec_t e : execs)
            if (e)
                (void)hipGraphExecDestroy(e);
        for (hipGraph_t g : graphs)
            if (g)
                (void)hipGraphDestroy(g);
        if (refitExec)
            (void)hipGraphExecDestroy(refitExec);
        if (refitGraph)
            (void)hipGraphDestroy(refitGraph);
        execs.clear();
        graphs.clear();
        refitExec = nullptr;
        refitGraph = nullptr;
        graphSignature.clear();
    }
    void dropChains() {
        for (auto& c : chains) {
            if (c->hostArgs)
                (void)hipHostFree(c->hostArgs);
            if (c->joined)
                (void)hipEventDestroy(c->joined);
        }
        chains.clear();
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
    // (whether the scene has textures picks the resolve variant a captured graph holds)
    std::vector<uint64_t> sig = {uint64_t(uintptr_t(s)), uint64_t(s->textureCount > 0), uint64_t(b->refitEveryFrame),
                                 b->shape.largeGroups,
                                 b->shape.hugeGroups,    b->shape.giantGroups,          b->shape.walkGroups,
                                 b->shape.xcdSwizzle,    b->shape.streams};
    for (const hvvr_gpu_view* v : b->views) {
        sig.push_back(v->msaaRate);
        sig.push_back(v->halfShading ? 1 : 0);
        sig.push_back(v->remap.ptr ? 1 : 0);
        sig.push_back(v->tileCount);
    }
    return sig;
}

// one chain's groups, unit counts and argument-block layout
int planChain(Chain& c) {
    const uint32_t n = uint32_t(c.views.size());
    std::map<std::pair<uint32_t, bool>, Group> byKind;
    c.remapMembers.clear();
    c.uFrustum = c.uBlock = c.uTile = c.uRemap = c.maxTiles = 0;
    for (uint32_t i = 0; i < n; i++) {
        const hvvr_gpu_view* v = c.views[i];
        Group& g = byKind[{v->msaaRate, v->halfShading}];
        g.msaaRate = v->msaaRate;
        g.half = v->halfShading;
        g.members.push_back(i);
        g.units += v->tileCount;
        c.uFrustum += frustumSetupUnits(v);
        c.uBlock += v->blockCount;
        c.uTile += v->tileCount;
        c.maxTiles = std::max(c.maxTiles, v->tileCount);
        if (v->remap.ptr) {
            c.remapMembers.push_back(i);
            c.uRemap += remapUnits(v);
        }
    }
    c.groups.clear();
    for (auto& kv : byKind)
        c.groups.push_back(kv.second);
    size_t at = 0;
    auto take = [&](size_t bytes) {
        size_t off = at;
        at = align256(at + bytes);
        return off;
    };
    c.offCounters = take(n * sizeof(uint64_t*));
    c.offFrustum = take(n * sizeof(FrustumSetupArgs));
    c.offBlock = take(n * sizeof(BlockTraverseArgs));
    c.offTile = take(5 * n * sizeof(TileListArgs));
    c.offWalk = take(n * sizeof(TileWalkArgs));
    for (Group& g : c.groups) {
        g.intersectOff = take(g.members.size() * sizeof(IntersectArgs));
        g.resolveOff = take(g.members.size() * sizeof(ResolveArgs));
    }
    c.offRemap = take(c.remapMembers.size() * sizeof(RemapArgs) + 1);
    c.sFrustum = take((n + 1) * 4);
    c.sBlock = take((n + 1) * 4);
    c.sTile = take((n + 1) * 4);
    for (Group& g : c.groups)
        g.startOff = take((g.members.size() + 1) * 4);
    c.sRemap = take((c.remapMembers.size() + 1) * 4);
    BATCH_HIP(hipHostMalloc((void**)&c.hostArgs, at, hipHostMallocDefault));
    hipError_t e = c.devArgs.resize(at);
    if (e != hipSuccess)
        return batchFail(std::string("batch argument block: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
    c.argBytes = at;
    BATCH_HIP(hipEventCreateWithFlags(&c.joined, hipEventDisableTiming));
    return HVVR_GPU_OK;
}

// the views dealt to streamCount chains (round-robin, so large and small views spread),
// each chain planned; chain 0 runs on the batch's stream
int planChains(hvvr_gpu_batch* b) {
    b->dropGraph(); // its nodes name the old chains' buffers
    b->dropChains();
    const uint32_t k = std::min<uint32_t>(b->shape.streams, uint32_t(b->views.size()));
    while (b->extraStreams.size() + 1 < k) {
        hipStream_t st = nullptr;
        BATCH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b->extraStreams.push_back(st);
    }
    for (uint32_t c = 0; c < k; c++) {
        auto chain = std::make_unique<Chain>();
        chain->stream = c == 0 ? b->stream : b->extraStreams[c - 1];
        for (uint32_t i = c; i < b->views.size(); i += k) {
            chain->views.push_back(b->views[i]);
            chain->index.push_back(i);
        }
        int rc = planChain(*chain);
        b->chains.push_back(std::move(chain));
        if (rc)
            return rc;
    }
    return HVVR_GPU_OK;
}

// this frame's arguments of one chain into its page-locked block
void fillArgs(Chain& c, const hvvr_gpu_scene* s, const matrix4x4* m) {
    const uint32_t n = uint32_t(c.views.size());
    uint8_t* h = c.hostArgs;
    auto counters = reinterpret_cast<uint64_t**>(h + c.offCounters);
    auto frustum = reinterpret_cast<FrustumSetupArgs*>(h + c.offFrustum);
    auto block = reinterpret_cast<BlockTraverseArgs*>(h + c.offBlock);
    auto tile = reinterpret_cast<TileListArgs*>(h + c.offTile);
    auto walk = reinterpret_cast<TileWalkArgs*>(h + c.offWalk);
    auto sF = reinterpret_cast<uint32_t*>(h + c.sFrustum);
    auto sB = reinterpret_cast<uint32_t*>(h + c.sBlock);
    auto sT = reinterpret_cast<uint32_t*>(h + c.sTile);
    sF[0] = sB[0] = sT[0] = 0;
    for (uint32_t i = 0; i < n; i++) {
        hvvr_gpu_view* v = c.views[i];
        const matrix4x4& mi = m[c.index[i]];
        v->cull.enabled = 0;
        counters[i] = v->counters.ptr;
        frustum[i] = frustumSetupArgs(v, mi);
        block[i] = blockTraverseArgs(v, s);
        for (int level = 0; level < 5; level++) {
            tile[level * n + i] = tileListArgs(v, s, level);
            // the chain's persistent workgroups take every view's queue in turn: one
            // scratch (its first view's) serves them
            tile[level * n + i].recScratch = c.views[0]->recScratch.ptr;
        }
        walk[i] = tileWalkArgs(v, s);
        sF[i + 1] = sF[i] + frustumSetupUnits(v);
        sB[i + 1] = sB[i] + v->blockCount;
        sT[i + 1] = sT[i] + v->tileCount;
    }
    for (const Group& g : c.groups) {
        auto ia = reinterpret_cast<IntersectArgs*>(h + g.intersectOff);
        auto ra = reinterpret_cast<ResolveArgs*>(h + g.resolveOff);
        auto st = reinterpret_cast<uint32_t*>(h + g.startOff);
        st[0] = 0;
        for (size_t k = 0; k < g.members.size(); k++) {
            const hvvr_gpu_view* v = c.views[g.members[k]];
            const matrix4x4& mi = m[c.index[g.members[k]]];
            ia[k] = intersectArgs(v, s, mi);
            ra[k] = resolveArgs(v, s, mi);
            st[k + 1] = st[k] + v->tileCount;
        }
    }
    auto rm = reinterpret_cast<RemapArgs*>(h + c.offRemap);
    auto sR = reinterpret_cast<uint32_t*>(h + c.sRemap);
    sR[0] = 0;
    for (size_t k = 0; k < c.remapMembers.size(); k++) {
        const hvvr_gpu_view* v = c.views[c.remapMembers[k]];
        rm[k] = remapArgs(v);
        sR[k + 1] = sR[k] + remapUnits(v);
    }
}

// one chain's work on its stream
void enqueueChain(const hvvr_gpu_batch* b, Chain& c, bool tex) {
    const uint32_t n = uint32_t(c.views.size());
    hipStream_t st = c.stream;
    uint8_t* d = c.devArgs.ptr;
    (void)hipMemcpyAsync(d, c.hostArgs, c.argBytes, hipMemcpyHostToDevice, st);
    launchClearCountersBatch(reinterpret_cast<uint64_t* const*>(d + c.offCounters), n, kCounterSlots, st);
    auto u32 = [&](size_t off) { return reinterpret_cast<const uint32_t*>(d + off); };
    launchFrustumSetupBatch(reinterpret_cast<const FrustumSetupArgs*>(d + c.offFrustum), u32(c.sFrustum), n,
                            c.uFrustum, st);
    launchBlockTraverseBatch(reinterpret_cast<const BlockTraverseArgs*>(d + c.offBlock), u32(c.sBlock), n, c.uBlock,
                             st);
    launchTileListsBatch(reinterpret_cast<const TileListArgs*>(d + c.offTile), u32(c.sTile), n, c.uTile, b->shape,
                         st);
    launchTileWalkBatch(reinterpret_cast<const TileWalkArgs*>(d + c.offWalk), n, c.maxTiles, b->shape, st);
    for (const Group& g : c.groups) {
        const uint32_t gn = uint32_t(g.members.size());
        launchIntersectBatch(int(g.msaaRate), reinterpret_cast<const IntersectArgs*>(d + g.intersectOff),
                             u32(g.startOff), gn, g.units, b->shape, st);
        launchResolveBatch(int(g.msaaRate), g.half, tex, reinterpret_cast<const ResolveArgs*>(d + g.resolveOff),
                           u32(g.startOff), gn, g.units, b->shape, st);
    }
    if (!c.remapMembers.empty())
        launchRemapBatch(reinterpret_cast<const RemapArgs*>(d + c.offRemap), u32(c.sRemap),
                         uint32_t(c.remapMembers.size()), c.uRemap, st);
}

// the frame on the batch's stream: the optional refit, then the chains, forked onto their
// streams and joined back; `graphs`: each chain (and the refit) as its captured graph
void enqueueFrame(hvvr_gpu_batch* b, hvvr_gpu_scene* s, bool graphs) {
    if (b->refitEveryFrame) {
        if (graphs)
            (void)hipGraphLaunch(b->refitExec, b->stream);
        else
            launchSceneRefit(s, b->stream);
    }
    if (b->chains.size() > 1)
        (void)hipEventRecord(b->fork, b->stream);
    for (size_t k = 1; k < b->chains.size(); k++)
        (void)hipStreamWaitEvent(b->chains[k]->stream, b->fork, 0);
    for (size_t k = 0; k < b->chains.size(); k++) {
        Chain& c = *b->chains[k];
        if (graphs)
            (void)hipGraphLaunch(b->execs[k], c.stream ? c.stream : b->stream);
        else
            enqueueChain(b, c, s->textureCount > 0);
    }
    for (size_t k = 1; k < b->chains.size(); k++) {
        (void)hipEventRecord(b->chains[k]->joined, b->chains[k]->stream);
        (void)hipStreamWaitEvent(b->stream, b->chains[k]->joined, 0);
    }
}

// captures the refit and each chain as a graph of its own
int captureGraphs(hvvr_gpu_batch* b, hvvr_gpu_scene* s) {
    auto capture = [&](hipStream_t st, auto&& body, hipGraph_t& g, hipGraphExec_t& e) -> int {
        BATCH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        body();
        hipError_t err = hipStreamEndCapture(st, &g);
        if (err != hipSuccess)
            return batchFail(std::string("graph capture: ") + hipGetErrorString(err), HVVR_GPU_ERR_HIP);
        err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        if (err != hipSuccess)
            return batchFail(std::string("graph instantiation: ") + hipGetErrorString(err), HVVR_GPU_ERR_HIP);
        return HVVR_GPU_OK;
    };
    if (b->refitEveryFrame) {
        int rc = capture(b->stream, [&] { launchSceneRefit(s, b->stream); }, b->refitGraph, b->refitExec);
        if (rc)
            return rc;
    }
    b->graphs.assign(b->chains.size(), nullptr);
    b->execs.assign(b->chains.size(), nullptr);
    for (size_t k = 0; k < b->chains.size(); k++) {
        Chain& c = *b->chains[k];
        hipStream_t st = c.stream ? c.stream : b->stream;
        int rc = capture(st, [&] { enqueueChain(b, c, s->textureCount > 0); }, b->graphs[k], b->execs[k]);
        if (rc)
            return rc;
    }
    return HVVR_GPU_OK;
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
    // the previous frame has read the argument blocks before they are rewritten
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
        int rc = planChains(b);
        if (rc)
            return rc;
        b->signature = sig;
    }
    std::vector<matrix4x4> m(b->views.size());
    memcpy(m.data(), m16, m.size() * sizeof(matrix4x4));
    for (auto& c : b->chains)
        fillArgs(*c, s, m.data());
    BATCH_HIP(hipEventRecord(b->start, b->stream));
    if (b->graphEnabled) {
        if (b->execs.empty() || b->graphSignature != sig) {
            b->dropGraph();
            int rc = captureGraphs(b, s);
            if (rc) {
                b->dropGraph();
                return rc;
            }
            b->graphSignature = sig;
        }
        enqueueFrame(b, s, true);
    } else {
        enqueueFrame(b, s, false);
    }
    BATCH_HIP(hipGetLastError());
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
        v->hitsValid = v->hitOutput;
        v->stagesRecorded = false;
        v->lastStream = b->stream;
        v->hintScene = s;
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
              hipEventCreateWithFlags(&b->sceneFence, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
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
        !inRange(shape->walkGroups) || shape->xcdSwizzle > 1 || shape->streams < 1 || shape->streams > 4)
        return batchFail("launch shape: workgroup counts in [1, 65536], xcdSwizzle 0 or 1, streams 1..4");
    b->shape.largeGroups = shape->largeGroups;
    b->shape.hugeGroups = shape->hugeGroups;
    b->shape.giantGroups = shape->giantGroups;
    b->shape.walkGroups = shape->walkGroups;
    b->shape.xcdSwizzle = shape->xcdSwizzle;
    b->shape.streams = shape->streams;
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
    shape->streams = b->shape.streams;
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
        // a frame leaves survivors, the deepest stack and node visits per tile; fold them
        // into each view's counters (the batch frame itself clears those slots)
        for (hvvr_gpu_view* v : b->views)
            launchTileWorkReduce(v, b->stream);
        BATCH_HIP(hipStreamSynchronize(b->stream));
        work.clear();
        for (const hvvr_gpu_view* v : b->views) {
            uint64_t c[5];
            BATCH_HIP(hipMemcpy(c, v->counters.ptr, sizeof(c), hipMemcpyDeviceToHost));
            work.push_back(c[0]); // survivor pairs
            work.push_back(c[2]); // deepest tile stack
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
    const std::vector<uint32_t> largeC = {640, 1280, 2560}, hugeC = {256, 512, 1024}, giantC = {128, 256, 512},
                                walkC = {256, 1024, 2048}, swizzleC = {0, 1}, streamsC = {1, 2, 4};
    struct Field {
        uint32_t LaunchShape::*member;
        const std::vector<uint32_t>* values;
    };
    const Field fields[] = {{&LaunchShape::streams, &streamsC},        {&LaunchShape::xcdSwizzle, &swizzleC},
                            {&LaunchShape::largeGroups, &largeC},
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
