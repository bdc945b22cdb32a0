// C ABI over the gfx950 render path (include/hvvr_gpu.h): resource management,
// the refit plan derived from the node topology, and per-frame kernel sequencing.

#include "../common/subsample.h"
#include "gpu_internal.h"
#include "launch_args.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string.h>
#include <vector>

// Shared scratch of hvvr_gpu_view_render_shared_scratch, one copy per GPU for the life of
// the process: the record scratch of the queued tile-list configurations
// (kSharedRecordWords words, as many persistent workgroups as it holds), then the five tile
// queues, [5][tileCount] words. Kernels and host copies use the device address from
// hipGetSymbolAddress (sharedScratchOf below), never this name.
enum : uint32_t {
    kSharedRecordWords = 59u << 16,                       // 14.75 MiB
    kSharedMaxTiles = 1u << 16,                           // 8,388,608 samples
    kSharedScratchWords = kSharedRecordWords + 5u * kSharedMaxTiles // 16 MiB in all
};
static __device__ __align__(256) uint32_t scratch[kSharedScratchWords];

namespace {
enum { kMaxDevices = 64 };
std::mutex gScratchLock;
uint32_t* gScratchPtr[kMaxDevices] = {}; // device address of `scratch` on each GPU
} // namespace

namespace hvvr {

thread_local std::string gLastError;
void setLastError(const std::string& msg) {
    gLastError = msg;
}

static int fail(const std::string& msg, int code = HVVR_GPU_ERR_INVALID) {
    setLastError(msg);
    return code;
}

template <class T>
static hipError_t upload(DeviceBuffer<T>& buf, const void* src, size_t count, hipStream_t stream) {
    hipError_t e = buf.resize(count);
    if (e != hipSuccess || count == 0)
        return e;
    e = hipMemcpyAsync(buf.ptr, src, count * sizeof(T), hipMemcpyHostToDevice, stream);
    if (e != hipSuccess)
        return e;
    return hipStreamSynchronize(stream);
}

// Refit plan (raycaster/gpu_scene_state.cu:506-589): leaf slot of every triangle, and
// child->parent tasks for internal nodes grouped by parent depth, deepest first.
static int buildRefitPlan(const BVHNode* nodes, uint32_t nodeCount, uint32_t triCount, std::vector<uint32_t>& triToNode,
                          std::vector<uint32_t>& taskChild, std::vector<uint32_t>& taskNodeChild,
                          std::vector<RefitLevel>& levels) {
    if (nodeCount == 0)
        return fail("scene has no BVH nodes");
    if (nodeCount >= (1u << 28))
        return fail("scene has too many BVH nodes (limit 2^28)");
    triToNode.assign(triCount, 0xffffffffu);
    std::vector<uint32_t> depth(nodeCount, 0);
    for (uint32_t n = 0; n < nodeCount; n++) {
        for (int k = 0; k < 4; k++) {
            if ((nodes[n].leafMask >> k) & 1u)
                continue;
            uint64_t child = uint64_t(n) + nodes[n].data[1 + k];
            if (child <= n || child >= nodeCount)
                return fail("BVH child offset out of range");
            depth[child] = depth[n] + 1;
        }
    }
    struct Task {
        uint32_t child, nodeChild, parentDepth;
    };
    std::vector<Task> tasks;
    for (uint32_t n = 0; n < nodeCount; n++) {
        for (int k = 0; k < 4; k++) {
            uint32_t nodeChild = (n << 2) | uint32_t(k);
            if (!((nodes[n].leafMask >> k) & 1u)) {
                tasks.push_back({n + nodes[n].data[1 + k], nodeChild, depth[n]});
            } else {
                uint32_t start = nodes[n].data[k], end = nodes[n].data[k + 1];
                if (end < start || end > triCount)
                    return fail("BVH leaf triangle range out of range");
                for (uint32_t t = start; t < end; t++)
                    triToNode[t] = nodeChild;
            }
        }
    }
    for (uint32_t t = 0; t < triCount; t++)
        if (triToNode[t] == 0xffffffffu)
            return fail("triangle not referenced by any BVH leaf");
    std::stable_sort(tasks.begin(), tasks.end(), [](const Task& a, const Task& b) {
        if (a.parentDepth != b.parentDepth)
            return a.parentDepth > b.parentDepth;
        return a.child < b.child;
    });
    taskChild.clear();
    taskNodeChild.clear();
    levels.clear();
    for (size_t i = 0; i < tasks.size(); i++) {
        if (i == 0 || tasks[i].parentDepth != tasks[i - 1].parentDepth)
            levels.push_back({uint32_t(i), 0});
        levels.back().count++;
        taskChild.push_back(tasks[i].child);
        taskNodeChild.push_back(tasks[i].nodeChild);
    }
    return HVVR_GPU_OK;
}

static bool validRate(uint32_t a) {
    return a == 1 || a == 2 || a == 4 || a == 8 || a == 16 || a == 32;
}

template <class T>
static int readBack(hvvr_gpu_scene* s, DeviceBuffer<T>& buf, void* dst) {
    if (!s || !dst)
        return fail("null argument");
    if (!s->everUpdated)
        return fail("scene has not been updated yet", HVVR_GPU_ERR_STATE);
    HVVR_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    if (buf.count)
        HVVR_HIP_CHECK(hipMemcpy(dst, buf.ptr, buf.bytes(), hipMemcpyDeviceToHost));
    return HVVR_GPU_OK;
}
} // namespace hvvr

using namespace hvvr;

extern "C" {

const char* hvvr_gpu_last_error(void) {
    return gLastError.c_str();
}

const char* hvvr_gpu_version(void) {
    return "hvvr_gpu 0.1 gfx950";
}

int hvvr_gpu_context_create(int device, hvvr_gpu_context** out) {
    if (!out)
        return fail("null output");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail("no HIP device available", HVVR_GPU_ERR_NO_DEVICE);
    if (device < 0 || device >= count)
        return fail("device index out of range", HVVR_GPU_ERR_NO_DEVICE);
    hipDeviceProp_t prop;
    HVVR_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(std::string("device is ") + prop.gcnArchName + ", this build targets gfx950",
                    HVVR_GPU_ERR_NO_DEVICE);
    HVVR_HIP_CHECK(hipSetDevice(device));
    if (device >= kMaxDevices)
        return fail("device index beyond the library's device table", HVVR_GPU_ERR_NO_DEVICE);
    uint32_t* shared = nullptr;
    {
        // the shared scratch's address on this device, looked up once per device with the
        // device current (each GPU holds its own copy of the array)
        std::lock_guard<std::mutex> lock(gScratchLock);
        if (!gScratchPtr[device]) {
            void* p = nullptr;
            hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
            gScratchPtr[device] = static_cast<uint32_t*>(p);
        }
        shared = gScratchPtr[device];
    }
    hvvr_gpu_context* c = new hvvr_gpu_context();
    c->device = device;
    c->sharedScratch = shared;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    for (auto& vs : c->viewStreams)
        ok = ok && hipStreamCreateWithFlags(&vs, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) == hipSuccess;
    for (auto& j : c->joins)
        ok = ok && hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        hvvr_gpu_context_destroy(c);
        return fail("stream or event creation failed", HVVR_GPU_ERR_HIP);
    }
    *out = c;
    return HVVR_GPU_OK;
}

int hvvr_gpu_context_destroy(hvvr_gpu_context* c) {
    if (!c)
        return HVVR_GPU_OK;
    (void)hipSetDevice(c->device);
    for (auto& vs : c->viewStreams)
        if (vs) {
            (void)hipStreamSynchronize(vs);
            (void)hipStreamDestroy(vs);
        }
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    if (c->fork)
        (void)hipEventDestroy(c->fork);
    for (auto& j : c->joins)
        if (j)
            (void)hipEventDestroy(j);
    delete c;
    return HVVR_GPU_OK;
}

int hvvr_gpu_context_synchronize(hvvr_gpu_context* c) {
    if (!c)
        return fail("null context");
    HVVR_HIP_CHECK(hipStreamSynchronize(c->stream));
    return HVVR_GPU_OK;
}

int hvvr_gpu_context_set_profiling(hvvr_gpu_context* c, int enable) {
    if (!c)
        return fail("null context");
    if (enable < 0 || enable > 2)
        return fail("profiling level must be 0, 1 or 2");
    c->profiling = enable;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_create(hvvr_gpu_context* ctx, const hvvr_gpu_scene_desc* d, hvvr_gpu_scene** out) {
    if (!ctx || !d || !out)
        return fail("null argument");
    // GBuffer sort keys are (id << 5 | subsample), decoded with a sign-filling shift
    // (intersect.cu:67,71): ids must stay below 2^26 to round-trip
    if (d->triangleCount > (1u << 26))
        return fail("triangleCount must be at most 2^26");
    if ((d->vertexCount && (!d->vertices || !d->vertexToTransform)) || (d->triangleCount && !d->triangles) ||
        !d->nodes || !d->nodeCount || (d->materialCount && !d->materials) ||
        (d->transformCount && !d->transforms))
        return fail("scene description has null arrays");
    const BVHNode* nodes = static_cast<const BVHNode*>(d->nodes);
    const PrecomputedTriangleShade* tris = static_cast<const PrecomputedTriangleShade*>(d->triangles);
    for (uint32_t t = 0; t < d->triangleCount; t++) {
        for (int k = 0; k < 3; k++)
            if (tris[t].indices[k] >= d->vertexCount)
                return fail("triangle vertex index out of range");
        if (tris[t].material >= d->materialCount)
            return fail("triangle material index out of range");
    }
    for (uint32_t v = 0; v < d->vertexCount; v++)
        if (d->vertexToTransform[v] >= d->transformCount)
            return fail("vertex transform index out of range");
    std::vector<uint32_t> triToNode, taskChild, taskNodeChild;
    std::vector<RefitLevel> levels;
    int rc = buildRefitPlan(nodes, d->nodeCount, d->triangleCount, triToNode, taskChild, taskNodeChild, levels);
    if (rc)
        return rc;

    hvvr_gpu_scene* s = new hvvr_gpu_scene();
    if (hipEventCreateWithFlags(&s->ready, hipEventDisableTiming) != hipSuccess) {
        delete s;
        return fail("scene event creation failed", HVVR_GPU_ERR_HIP);
    }
    s->ctx = ctx;
    s->vertexCount = d->vertexCount;
    s->triangleCount = d->triangleCount;
    s->nodeCount = d->nodeCount;
    s->materialCount = d->materialCount;
    s->transformCount = d->transformCount;
    s->levels = levels;
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess)
            e = r;
    };
    step(upload(s->vertsModel, d->vertices, d->vertexCount, st));
    step(s->vertsWorld.resize(d->vertexCount));
    step(upload(s->vertexToTransform, d->vertexToTransform, d->vertexCount, st));
    step(upload(s->transforms, d->transforms, d->transformCount, st));
    step(upload(s->triShade, d->triangles, d->triangleCount, st));
    step(s->triIntersect.resize(d->triangleCount));
    step(upload(s->nodes, d->nodes, d->nodeCount, st));
    step(upload(s->materials, d->materials, d->materialCount, st));
    step(upload(s->triToNodeChild, triToNode.data(), triToNode.size(), st));
    step(upload(s->taskChild, taskChild.data(), taskChild.size(), st));
    step(upload(s->taskNodeChild, taskNodeChild.data(), taskNodeChild.size(), st));
    if (e != hipSuccess) {
        delete s;
        return fail(std::string("scene upload failed: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
    }
    *out = s;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_destroy(hvvr_gpu_scene* s) {
    if (s) {
        (void)hipStreamSynchronize(s->ctx->stream);
        delete s;
    }
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_set_transforms(hvvr_gpu_scene* s, const float* transforms, uint32_t count) {
    if (!s || (!transforms && count))
        return fail("null argument");
    if (count != s->transformCount)
        return fail("transform count differs from the scene's");
    if (count)
        HVVR_HIP_CHECK(hipMemcpyAsync(s->transforms.ptr, transforms, count * sizeof(matrix4x4),
                                      hipMemcpyHostToDevice, s->ctx->stream));
    HVVR_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    s->dirty = true;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_set_vertices(hvvr_gpu_scene* s, uint32_t first, uint32_t count, const void* verts) {
    if (!s || (!verts && count))
        return fail("null argument");
    if (uint64_t(first) + count > s->vertexCount)
        return fail("vertex range out of bounds");
    if (count)
        HVVR_HIP_CHECK(hipMemcpyAsync(s->vertsModel.ptr + first, verts, size_t(count) * sizeof(ShadingVertex),
                                      hipMemcpyHostToDevice, s->ctx->stream));
    HVVR_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    s->dirty = true;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_set_lights(hvvr_gpu_scene* s, const void* lights, uint32_t count, uint32_t* dropped) {
    if (!s || (!lights && count))
        return fail("null argument");
    LightingEnvironment env;
    memset(&env, 0, sizeof(env));
    uint32_t drop = 0;
    const LightUnion* L = static_cast<const LightUnion*>(lights);
    for (uint32_t i = 0; i < count; i++) {
        switch (L[i].type) {
            case LightTypeDirectional:
                if (env.directionalLightCount < MAX_DIRECTIONAL_LIGHTS)
                    env.directionalLights[env.directionalLightCount++] = L[i].directional;
                else
                    drop++;
                break;
            case LightTypePoint:
                if (env.pointLightCount < MAX_POINT_LIGHTS)
                    env.pointLights[env.pointLightCount++] = L[i].point;
                else
                    drop++;
                break;
            case LightTypeSpot:
                if (env.spotLightCount < MAX_SPOT_LIGHTS)
                    env.spotLights[env.spotLightCount++] = L[i].spot;
                else
                    drop++;
                break;
            default:
                drop++;
                break;
        }
    }
    s->lighting = env;
    if (dropped)
        *dropped = drop;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_update(hvvr_gpu_scene* s) {
    if (!s)
        return fail("null scene");
    if (!s->dirty)
        return HVVR_GPU_OK;
    HVVR_HIP_CHECK(hipSetDevice(s->ctx->device));
    launchSceneRefit(s, s->ctx->stream);
    HVVR_HIP_CHECK(hipGetLastError());
    HVVR_HIP_CHECK(hipEventRecord(s->ready, s->ctx->stream));
    s->dirty = false;
    s->everUpdated = true;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_read_nodes(hvvr_gpu_scene* s, void* dst) {
    return readBack(s, s->nodes, dst);
}
int hvvr_gpu_scene_read_triangles(hvvr_gpu_scene* s, void* dst) {
    return readBack(s, s->triIntersect, dst);
}
int hvvr_gpu_scene_read_vertices(hvvr_gpu_scene* s, void* dst) {
    return readBack(s, s->vertsWorld, dst);
}

int hvvr_gpu_view_create(hvvr_gpu_context* ctx, const hvvr_gpu_view_desc* d, hvvr_gpu_view** out) {
    return makeView(ctx, d, false, out);
}

} // extern "C"

// A view and its buffers. The sample hierarchy (beams, camera-space frusta, remap) is
// uploaded from d's host arrays, or with `deviceSetup` only allocated (the remap for
// width x height x split entries) for the device sample setup to fill.
int hvvr::makeView(hvvr_gpu_context* ctx, const hvvr_gpu_view_desc* d, bool deviceSetup, hvvr_gpu_view** out) {
    if (!ctx || !d || !out)
        return fail("null argument");
    if (!validRate(d->msaaRate))
        return fail("msaaRate must be 1, 2, 4, 8, 16 or 32");
    if (d->sampleCount == 0 || d->sampleCount % BLOCK_SIZE != 0)
        return fail("sampleCount must be a positive multiple of 8192");
    if (d->validSampleCount > d->sampleCount)
        return fail("validSampleCount exceeds sampleCount");
    if (!deviceSetup && (!d->beams || !d->tileFrusta || !d->blockFrusta))
        return fail("view description has null arrays");
    if (d->split < 1 || d->split > 3 || d->width == 0 || d->height == 0)
        return fail("invalid image size or split");
    if (!(d->lensRadius >= 0.0f) || !std::isfinite(d->lensRadius))
        return fail("lensRadius must be finite and >= 0");
    if (d->lensRadius > 0.0f && (!(d->focalDistance > 0.0f) || !std::isfinite(d->focalDistance)))
        return fail("depth of field needs a finite focalDistance > 0");
    uint64_t remapCount = d->remap || deviceSetup ? uint64_t(d->width) * d->height * d->split : 0;
    for (uint64_t i = 0; d->remap && i < remapCount; i++)
        if (d->remap[i] >= int32_t(d->sampleCount))
            return fail("remap entry out of range");

    hvvr_gpu_view* v = new hvvr_gpu_view();
    v->ctx = ctx;
    v->msaaRate = d->msaaRate;
    v->sampleCount = d->sampleCount;
    v->validSampleCount = d->validSampleCount;
    v->tileCount = d->sampleCount / TILE_SIZE;
    v->blockCount = d->sampleCount / BLOCK_SIZE;
    v->width = d->width;
    v->height = d->height;
    v->split = d->split;
    v->lens = {d->lensRadius, d->focalDistance};
    v->dof = d->lensRadius > 0.0f;
    subsampleOffsets(v->msaaRate, 0.0f, v->offsets);
    hipStream_t st = ctx->stream;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess)
            e = r;
    };
    if (deviceSetup) {
        step(v->beams.resize(v->sampleCount));
        step(v->tileFrustaCam.resize(v->tileCount));
        step(v->blockFrustaCam.resize(v->blockCount));
        step(v->remap.resize(remapCount));
    } else {
        step(upload(v->beams, d->beams, v->sampleCount, st));
        step(upload(v->tileFrustaCam, d->tileFrusta, v->tileCount, st));
        step(upload(v->blockFrustaCam, d->blockFrusta, v->blockCount, st));
        step(upload(v->remap, d->remap, remapCount, st));
    }
    step(v->tileFrustaWorld.resize(v->tileCount));
    step(v->tilePre.resize(v->tileCount));
    step(v->blockPre.resize(v->blockCount));
    step(v->blockStacks.resize(size_t(v->blockCount) * BLOCK_STACK_SIZE));
    step(v->blockStackSize.resize(v->blockCount));
    step(v->tileCounts.resize(v->tileCount));
    step(v->tileQueues.resize(size_t(5) * v->tileCount));
    step(v->recScratch.resize(tileListScratchWords()));
    step(v->tileLevel.resize(v->tileCount));
    step(v->tileWork.resize(size_t(4) * v->tileCount));
    step(v->tileOffsets.resize(v->tileCount));
    v->listCapacity = std::max<uint64_t>(uint64_t(v->tileCount) * 1024, 1u << 20);
    step(v->triIndices.resize(v->listCapacity));
    step(v->gbuffer.resize(size_t(v->sampleCount) * v->msaaRate));
    step(v->sampleResults.resize(v->sampleCount));
    step(v->image.resize(size_t(v->width) * v->height));
    std::vector<vector2> table; // lives until the stream synchronisation below
    if (v->dof) {
        lensPositionTable(v->msaaRate, table);
        step(upload(v->lensTable, table.data(), table.size(), st));
        step(v->occRank.resize(v->tileCount));
    }
    step(v->counters.resize(16));
    // the per-stage events only time the stages: no system-scope fence, whose cache
    // writeback and invalidation at every stage boundary would slow the frame they time
    for (auto& ev : v->stageEvents)
        step(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
    for (auto& ev : v->kernelEvents)
        step(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
    if (e == hipSuccess)
        e = hipMemsetAsync(v->image.ptr, 0, v->image.bytes(), st);
    // no hints yet (0), and zero work: neither may hold what a freed allocation left
    if (e == hipSuccess)
        e = hipMemsetAsync(v->tileLevel.ptr, 0, v->tileLevel.bytes(), st);
    if (e == hipSuccess)
        e = hipMemsetAsync(v->tileWork.ptr, 0, v->tileWork.bytes(), st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        delete v;
        return fail(std::string("view allocation failed: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
    }
    *out = v;
    return HVVR_GPU_OK;
}

extern "C" {

int hvvr_gpu_view_destroy(hvvr_gpu_view* v) {
    if (v) {
        (void)hipStreamSynchronize(v->ctx->stream);
        delete v;
    }
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_set_traversal_stack_limit(hvvr_gpu_view* v, uint32_t entries) {
    if (!v)
        return fail("null view");
    if (entries < 1 || entries > 64)
        return fail("traversal stack limit must be in [1, 64]");
    v->waveStackLimit = entries;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_set_shading_precision(hvvr_gpu_view* v, int bits) {
    if (!v)
        return fail("null view");
    if (bits != 16 && bits != 32)
        return fail("shading precision must be 16 or 32 bits");
    v->halfShading = bits == 16;
    return HVVR_GPU_OK;
}

static int viewWait(hvvr_gpu_view* v);

int hvvr_gpu_view_set_hit_output(hvvr_gpu_view* v, int enable) {
    if (!v)
        return fail("null view");
    if (enable && !v->hits.ptr) {
        if (v->rendered) {
            int rc = viewWait(v);
            if (rc)
                return rc;
        }
        HVVR_HIP_CHECK(hipSetDevice(v->ctx->device));
        HVVR_HIP_CHECK(v->hits.resize(size_t(v->sampleCount) * v->msaaRate));
    }
    v->hitOutput = enable != 0;
    v->hitsValid = false;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_set_jitter_spin(hvvr_gpu_view* v, int enable) {
    if (!v)
        return fail("null view");
    v->jitterSpin = enable != 0;
    if (!v->jitterSpin)
        subsampleOffsets(v->msaaRate, 0.0f, v->offsets);
    return HVVR_GPU_OK;
}

// one frame of one view; `fov` (a foveated frame) switches on the screen cull planes, the
// t-max output and the polar reconstruction in place of the remap
static int renderOne(hvvr_gpu_view* v, hvvr_gpu_scene* s, const float* m16, const float* jitter, hipStream_t st,
                     const hvvr_gpu_foveated_frame* fov = nullptr, const TileListScratch* scratch = nullptr) {
    hvvr_gpu_context* c = v->ctx;
    matrix4x4 m;
    memcpy(&m, m16, sizeof(m));
    v->cull.enabled = 0;
    if (fov) {
        memcpy(v->cull.planes, fov->cullPlanes, sizeof(v->cull.planes));
        v->cull.enabled = std::isinf(fov->cullPlanes[3]) ? 0 : 1;
    }
    v->lastCameraToWorld = m;
    if (v->jitterSpin && jitter)
        subsampleOffsets(v->msaaRate, jitter[0] * 6.28f, v->offsets);
    const bool prof = c->profiling == 1;
    const bool timed = c->profiling != 0;
    auto mark = [&](int i) {
        if (prof)
            (void)hipEventRecord(v->stageEvents[i], st);
    };
    mark(0); // the frame's counters are zeroed by the frustum setup kernel
    launchFrustumSetup(v, m, st);
    mark(1);
    launchBlockTraverse(v, s, st);
    mark(2);
    launchTileLists(v, s, st, scratch, timed ? v->kernelEvents : nullptr);
    mark(3);
    launchTileListFallback(v, s, st, scratch);
    if (v->dof)
        launchOccupiedRank(v, st);
    mark(4);
    launchIntersect(v, s, m, st);
    mark(5);
    launchResolve(v, s, m, st);
    mark(6);
    if (fov) {
        matrix3x3 sampleToEye;
        matrix4x4 eyeToSamplePrevious;
        memcpy(&sampleToEye, fov->sampleToEye, sizeof(sampleToEye));
        memcpy(&eyeToSamplePrevious, fov->eyeToSamplePrevious, sizeof(eyeToSamplePrevious));
        launchFoveatedReconstruct(v, sampleToEye, eyeToSamplePrevious, st);
    } else if (v->remap.ptr) {
        launchRemap(v, st);
    }
    mark(7);
    HVVR_HIP_CHECK(hipGetLastError());
    v->stagesRecorded = prof;
    v->kernelTimed = timed;
    v->rendered = true;
    v->hitsValid = v->hitOutput;
    v->hintScene = s; // this frame writes every tile's configuration hint
    v->lastStream = st;
    return HVVR_GPU_OK;
}

// the stream a single-view render goes to: the view's own (after the scene's last
// update) or the context's
static hipStream_t renderStream(hvvr_gpu_view* v, hvvr_gpu_scene* s) {
    if (!v->ownStream)
        return v->ctx->stream;
    (void)hipStreamWaitEvent(v->ownStream, s->ready, 0);
    return v->ownStream;
}

int hvvr_gpu_view_render(hvvr_gpu_view* v, hvvr_gpu_scene* s, const float* m16, const float* jitter) {
    if (!v || !s || !m16)
        return fail("null argument");
    if (v->ctx != s->ctx)
        return fail("view and scene belong to different contexts");
    if (!s->everUpdated || s->dirty) {
        int rc = hvvr_gpu_scene_update(s);
        if (rc)
            return rc;
    }
    return renderOne(v, s, m16, jitter, renderStream(v, s));
}

int hvvr_gpu_view_render_shared_scratch(hvvr_gpu_view* v, hvvr_gpu_scene* s, const float* m16, const float* jitter) {
    if (!v || !s || !m16)
        return fail("null argument");
    if (v->ctx != s->ctx)
        return fail("view and scene belong to different contexts");
    if (!v->ctx->sharedScratch)
        return fail("the context has no shared scratch", HVVR_GPU_ERR_STATE);
    if (v->tileCount > kSharedMaxTiles)
        return fail("view of " + std::to_string(v->tileCount) + " tiles: the shared scratch holds the tile queues of " +
                    std::to_string(kSharedMaxTiles) + " tiles (" + std::to_string(kSharedMaxTiles * TILE_SIZE) +
                    " samples) at most");
    if (!s->everUpdated || s->dirty) {
        int rc = hvvr_gpu_scene_update(s); // kernels only: allocates nothing
        if (rc)
            return rc;
    }
    // Nothing in the shared scratch has to start at zero: a queue entry is read only
    // below its length, which the frame's frustum setup zeroes in the view's counters,
    // and a workgroup writes its record scratch for a tile before reading it.
    TileListScratch sc;
    sc.rec = v->ctx->sharedScratch;
    sc.queues = v->ctx->sharedScratch + kSharedRecordWords;
    tileListGroupsFor(kSharedRecordWords, sc.groups);
    return renderOne(v, s, m16, jitter, renderStream(v, s), nullptr, &sc);
}

static int viewSync(hvvr_gpu_view* v);

// the sRGB byte -> linear table the reconstruction samples through (sRgbToLinear,
// util_graphics.h:16-18), and byte -> linear -> byte (linearToSRgb, :26-28, truncated)
static void srgbTables(std::vector<float>& toLinear, std::vector<uint8_t>& roundTrip) {
    toLinear.resize(256);
    roundTrip.resize(256);
    for (int k = 0; k < 256; k++) {
        float x = float(k) / 255.0f;
        float lin = (x <= 0.04045f) ? x * (1.0f / 12.92f) : powf((x + .055f) * (1.0f / 1.055f), 2.4f);
        float enc = (lin <= 0.0031308f) ? (lin * 12.92f) : (powf(lin, 1.0f / 2.4f) * 1.055f - 0.055f);
        float b = enc * 255.0f;
        toLinear[k] = lin;
        roundTrip[k] = uint8_t(b > 0.0f ? (b < 255.0f ? uint32_t(b) : 255u) : 0u);
    }
}

int hvvr_gpu_view_set_foveated(hvvr_gpu_view* v, const hvvr_gpu_foveated_desc* d) {
    if (!v || !d || !d->polarRemap)
        return fail("null argument");
    if (v->remap.ptr)
        return fail("a foveated view is created without a remap");
    if (d->samplesPerRing == 0 || d->ringCount == 0 ||
        uint64_t(d->samplesPerRing) * d->ringCount != uint64_t(v->validSampleCount))
        return fail("samplesPerRing * ringCount must equal the view's validSampleCount");
    if (!(d->maxEccentricity > 0.0f) || !std::isfinite(d->maxEccentricity))
        return fail("maxEccentricity must be finite and > 0");
    for (int k = 0; k < 6; k++)
        if (!std::isfinite(d->eccentricityMap[k]))
            return fail("eccentricity map values must be finite");
    if (!(d->taaAlpha >= 0.0f && d->taaAlpha <= 1.0f) || !std::isfinite(d->taaStddevMultiplier) ||
        !std::isfinite(d->contrastFe))
        return fail("invalid temporal filter or contrast settings");
    // every polar texel is written by exactly one sample
    std::vector<uint8_t> seen(size_t(d->samplesPerRing) * d->ringCount, 0);
    for (uint32_t i = 0; i < v->validSampleCount; i++) {
        uint32_t x = d->polarRemap[2 * i], y = d->polarRemap[2 * i + 1];
        if (x >= d->samplesPerRing || y >= d->ringCount)
            return fail("polar remap entry out of range");
        uint8_t& s = seen[size_t(y) * d->samplesPerRing + x];
        if (s)
            return fail("polar remap maps two samples to one texel");
        s = 1;
    }
    FoveatedState& f = v->fov;
    f.samplesPerRing = d->samplesPerRing;
    f.ringCount = d->ringCount;
    memcpy(f.emap, d->eccentricityMap, sizeof(f.emap));
    f.maxEccentricity = d->maxEccentricity;
    f.taaAlpha = d->taaAlpha;
    f.taaStddevMultiplier = d->taaStddevMultiplier;
    f.contrast = d->contrastEnhancement != 0;
    f.contrastFe = d->contrastFe;
    f.historyIndex = 0;
    const size_t polar = size_t(d->samplesPerRing) * d->ringCount;
    const size_t pixels = size_t(v->width) * v->height;
    std::vector<float> toLinear;
    std::vector<uint8_t> roundTrip;
    srgbTables(toLinear, roundTrip);
    hipStream_t st = v->ctx->stream;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess)
            e = r;
    };
    step(upload(f.polarRemap, d->polarRemap, v->validSampleCount, st));
    step(upload(f.srgbToLinear, toLinear.data(), 256, st));
    step(upload(f.srgbRoundTrip, roundTrip.data(), 256, st));
    step(f.polarColor.resize(polar));
    step(f.polarDepth.resize(polar));
    step(f.moment1.resize(polar));
    step(f.moment2.resize(polar));
    for (int k = 0; k < 2; k++) {
        step(f.history[k].resize(pixels));
        step(f.filtered[k].resize(pixels));
    }
    step(v->tMax.resize(v->sampleCount));
    for (int k = 0; k < 2; k++) {
        if (e == hipSuccess)
            e = hipMemsetAsync(f.history[k].ptr, 0, f.history[k].bytes(), st);
        if (e == hipSuccess)
            e = hipMemsetAsync(f.filtered[k].ptr, 0, f.filtered[k].bytes(), st);
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        v->foveated = false;
        return fail(std::string("foveated allocation failed: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
    }
    v->foveated = true;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_render_foveated(hvvr_gpu_view* v, hvvr_gpu_scene* s, const hvvr_gpu_foveated_frame* fr,
                                  const float* jitter) {
    if (!v || !s || !fr)
        return fail("null argument");
    if (!v->foveated)
        return fail("view is not foveated: call hvvr_gpu_view_set_foveated first", HVVR_GPU_ERR_STATE);
    if (v->ctx != s->ctx)
        return fail("view and scene belong to different contexts");
    if (!s->everUpdated || s->dirty) {
        int rc = hvvr_gpu_scene_update(s);
        if (rc)
            return rc;
    }
    return renderOne(v, s, fr->eyeToWorld, jitter, renderStream(v, s), fr);
}

int hvvr_gpu_view_read_foveated(hvvr_gpu_view* v, int buffer, void* dst, uint64_t bytes) {
    if (!dst)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc)
        return rc;
    if (!v->foveated)
        return fail("view is not foveated", HVVR_GPU_ERR_STATE);
    FoveatedState& f = v->fov;
    // the last frame wrote history[historyIndex ^ 1] (the index has advanced since)
    const uint32_t last = f.historyIndex ^ 1u;
    const void* src = nullptr;
    uint64_t n = 0;
    switch (buffer) {
        case HVVR_GPU_FOVEATED_POLAR_COLOR: src = f.polarColor.ptr; n = f.polarColor.bytes(); break;
        case HVVR_GPU_FOVEATED_POLAR_DEPTH: src = f.polarDepth.ptr; n = f.polarDepth.bytes(); break;
        case HVVR_GPU_FOVEATED_MOMENT1: src = f.moment1.ptr; n = f.moment1.bytes(); break;
        case HVVR_GPU_FOVEATED_MOMENT2: src = f.moment2.ptr; n = f.moment2.bytes(); break;
        case HVVR_GPU_FOVEATED_FILTERED: src = f.history[last].ptr; n = f.history[last].bytes(); break;
        case HVVR_GPU_FOVEATED_PREVIOUS: src = f.history[last ^ 1u].ptr; n = f.history[last ^ 1u].bytes(); break;
        case HVVR_GPU_FOVEATED_BLUR1: src = f.filtered[0].ptr; n = f.filtered[0].bytes(); break;
        case HVVR_GPU_FOVEATED_BLUR2: src = f.filtered[1].ptr; n = f.filtered[1].bytes(); break;
        case HVVR_GPU_FOVEATED_TMAX: src = v->tMax.ptr; n = v->tMax.bytes(); break;
        default: return fail("unknown foveated buffer");
    }
    if (bytes != n)
        return fail("foveated buffer size mismatch: expected " + std::to_string(n) + " bytes");
    HVVR_HIP_CHECK(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
    return HVVR_GPU_OK;
}

int hvvr_gpu_render_views(hvvr_gpu_scene* s, uint32_t count, hvvr_gpu_view* const* views, const float* m16,
                          const float* jitter) {
    if (!s || (count && (!views || !m16)))
        return fail("null argument");
    for (uint32_t i = 0; i < count; i++) {
        if (!views[i])
            return fail("null view");
        if (views[i]->ctx != s->ctx)
            return fail("view and scene belong to different contexts");
        if (views[i]->foveated)
            return fail("foveated views render with hvvr_gpu_view_render_foveated");
    }
    if (count == 0)
        return HVVR_GPU_OK;
    hvvr_gpu_context* c = s->ctx;
    if (!s->everUpdated || s->dirty) {
        int rc = hvvr_gpu_scene_update(s);
        if (rc)
            return rc;
    }
    if (count == 1)
        return renderOne(views[0], s, m16, jitter, c->stream);
    // Independent views run on separate streams so one view's latency-bound traversal
    // overlaps another's compute-bound intersection; a view listed twice stays on one
    // stream. Everything forks after the scene update and joins back before later work
    // on the context stream.
    const uint32_t nStreams = count < uint32_t(kViewStreams) ? count : uint32_t(kViewStreams);
    HVVR_HIP_CHECK(hipEventRecord(c->fork, c->stream));
    for (uint32_t k = 0; k < nStreams; k++)
        HVVR_HIP_CHECK(hipStreamWaitEvent(c->viewStreams[k], c->fork, 0));
    for (uint32_t i = 0; i < count; i++) {
        uint32_t first = i;
        for (uint32_t j = 0; j < i; j++)
            if (views[j] == views[i]) {
                first = j;
                break;
            }
        int rc = renderOne(views[i], s, m16 + 16 * size_t(i), jitter ? jitter + 2 * size_t(i) : nullptr,
                           c->viewStreams[first % nStreams]);
        if (rc)
            return rc;
    }
    for (uint32_t k = 0; k < nStreams; k++) {
        HVVR_HIP_CHECK(hipEventRecord(c->joins[k], c->viewStreams[k]));
        HVVR_HIP_CHECK(hipStreamWaitEvent(c->stream, c->joins[k], 0));
    }
    return HVVR_GPU_OK;
}

static int viewSync(hvvr_gpu_view* v) {
    if (!v)
        return fail("null view");
    if (!v->rendered)
        return fail("view has not been rendered", HVVR_GPU_ERR_STATE);
    hipStream_t st = v->lastStream ? v->lastStream : v->ctx->stream;
    HVVR_HIP_CHECK(hipStreamSynchronize(st));
    if (st != v->ctx->stream)
        HVVR_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
    uint64_t counters[4];
    HVVR_HIP_CHECK(hipMemcpyAsync(counters, v->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost, st));
    HVVR_HIP_CHECK(hipStreamSynchronize(st));
    if (counters[1])
        return fail("tile traversal exceeded its stack or triangle-list budget", HVVR_GPU_ERR_OVERFLOW);
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_set_own_stream(hvvr_gpu_view* v, int enable) {
    if (!v)
        return fail("null view");
    if (enable && !v->ownStream) {
        HVVR_HIP_CHECK(hipSetDevice(v->ctx->device));
        HVVR_HIP_CHECK(hipStreamCreateWithFlags(&v->ownStream, hipStreamNonBlocking));
        HVVR_HIP_CHECK(hipEventCreateWithFlags(&v->done, hipEventDisableTiming));
    } else if (!enable && v->ownStream) {
        HVVR_HIP_CHECK(hipStreamSynchronize(v->ownStream));
        (void)hipEventDestroy(v->done);
        (void)hipStreamDestroy(v->ownStream);
        v->done = nullptr;
        v->ownStream = nullptr;
        v->lastStream = nullptr;
    }
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_read_image_async(hvvr_gpu_view* v, uint32_t* image, uint32_t stride) {
    if (!v || !image)
        return fail("null argument");
    if (!v->ownStream)
        return fail("asynchronous reads need the view's own stream", HVVR_GPU_ERR_STATE);
    if (!v->rendered)
        return fail("view has not been rendered", HVVR_GPU_ERR_STATE);
    if (stride < v->width)
        return fail("stride smaller than image width");
    HVVR_HIP_CHECK(hipMemcpy2DAsync(image, size_t(stride) * 4, v->image.ptr, size_t(v->width) * 4,
                                    size_t(v->width) * 4, v->height, hipMemcpyDeviceToHost, v->ownStream));
    HVVR_HIP_CHECK(hipEventRecord(v->done, v->ownStream));
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_query(hvvr_gpu_view* v) {
    if (!v)
        return fail("null view");
    if (!v->ownStream)
        return fail("query needs the view's own stream", HVVR_GPU_ERR_STATE);
    hipError_t e = hipEventQuery(v->done);
    if (e == hipSuccess)
        return 1;
    if (e == hipErrorNotReady)
        return 0;
    return fail(std::string("hipEventQuery: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
}

// waits for the view's last frame (its own stream and the context's)
static int viewWait(hvvr_gpu_view* v) {
    hipStream_t st = v->lastStream ? v->lastStream : v->ctx->stream;
    HVVR_HIP_CHECK(hipStreamSynchronize(st));
    if (st != v->ctx->stream)
        HVVR_HIP_CHECK(hipStreamSynchronize(v->ctx->stream));
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_list_demand(hvvr_gpu_view* v, uint64_t* entries, uint64_t* capacity, int* overflowed) {
    if (!v || !entries || !capacity || !overflowed)
        return fail("null argument");
    if (!v->rendered)
        return fail("view has not been rendered", HVVR_GPU_ERR_STATE);
    int rc = viewWait(v);
    if (rc)
        return rc;
    uint64_t counters[4];
    HVVR_HIP_CHECK(hipMemcpy(counters, v->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost));
    *entries = counters[3];
    *capacity = v->listCapacity;
    *overflowed = counters[1] ? 1 : 0;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_reserve_lists(hvvr_gpu_view* v, uint64_t entries) {
    if (!v)
        return fail("null view");
    if (entries > 0xffffffffull)
        return fail("list storage is indexed by 32-bit offsets: at most 2^32 - 1 entries");
    if (entries <= v->listCapacity)
        return HVVR_GPU_OK;
    if (v->rendered) {
        int rc = viewWait(v);
        if (rc)
            return rc;
    }
    HVVR_HIP_CHECK(hipSetDevice(v->ctx->device));
    hipError_t e = v->triIndices.resize(entries);
    if (e != hipSuccess) {
        // keep a usable view: back to the capacity it had
        (void)v->triIndices.resize(v->listCapacity);
        return fail(std::string("list storage allocation failed: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP);
    }
    v->listCapacity = entries;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_foveated_rewind(hvvr_gpu_view* v) {
    if (!v)
        return fail("null view");
    if (!v->foveated || !v->rendered)
        return fail("view has no foveated frame to rewind", HVVR_GPU_ERR_STATE);
    int rc = viewWait(v);
    if (rc)
        return rc;
    v->fov.historyIndex ^= 1u;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_estimate_bytes(const hvvr_gpu_view_desc* d, const hvvr_gpu_foveated_desc* f, uint64_t* bytes) {
    if (!d || !bytes)
        return fail("null argument");
    if (d->sampleCount == 0 || d->sampleCount % BLOCK_SIZE != 0 || !validRate(d->msaaRate))
        return fail("invalid view description");
    // the allocations of hvvr_gpu_view_create and hvvr_gpu_view_set_foveated
    const uint64_t S = d->sampleCount, T = S / TILE_SIZE, B = S / BLOCK_SIZE, A = d->msaaRate;
    const uint64_t pixels = uint64_t(d->width) * d->height;
    const uint64_t listCapacity = std::max<uint64_t>(T * 1024, 1u << 20);
    uint64_t b = S * sizeof(DirectionalBeam) + 2 * T * sizeof(SimpleRayFrustum) + B * sizeof(SimpleRayFrustum) +
                 T * sizeof(FrustumPre) + B * sizeof(FrustumPre) + B * BLOCK_STACK_SIZE * sizeof(StackEntry) +
                 B * 4 + 12 * T * 4 + uint64_t(tileListScratchWords()) * 4 + listCapacity * 4 +
                 S * A * sizeof(GBufferSubsample) + S * 4 + pixels * 4 + 16 * 8;
    if (d->remap)
        b += pixels * d->split * 4;
    if (d->lensRadius > 0.0f)
        b += 4ull * TILE_SIZE * A * sizeof(vector2) + T * 4; // lens table rows, occupied-tile ranks
    if (f) {
        const uint64_t polar = uint64_t(f->samplesPerRing) * f->ringCount;
        b += uint64_t(d->validSampleCount) * sizeof(vector2ui) + polar * (4 + 4 + 8 + 8) + 4 * pixels * 4 + 256 * 4 +
             256 + S * 4;
    }
    *bytes = b;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_device_bytes(hvvr_gpu_view* v, uint64_t* bytes) {
    if (!v || !bytes)
        return fail("null argument");
    uint64_t b = v->beams.bytes() + v->tileFrustaCam.bytes() + v->blockFrustaCam.bytes() +
                 v->tileFrustaWorld.bytes() + v->tilePre.bytes() + v->blockPre.bytes() + v->blockStacks.bytes() +
                 v->blockStackSize.bytes() + v->tileQueues.bytes() + v->recScratch.bytes() + v->tileLevel.bytes() +
                 v->tileWork.bytes() + v->tileCounts.bytes() + v->tileOffsets.bytes() + v->triIndices.bytes() +
                 v->gbuffer.bytes() + v->sampleResults.bytes() + v->remap.bytes() + v->image.bytes() +
                 v->lensTable.bytes() + v->occRank.bytes() + v->counters.bytes() + v->tMax.bytes() +
                 v->hits.bytes();
    const FoveatedState& f = v->fov;
    b += f.polarRemap.bytes() + f.polarColor.bytes() + f.polarDepth.bytes() + f.moment1.bytes() +
         f.moment2.bytes() + f.history[0].bytes() + f.history[1].bytes() + f.filtered[0].bytes() +
         f.filtered[1].bytes() + f.srgbToLinear.bytes() + f.srgbRoundTrip.bytes();
    *bytes = b;
    return HVVR_GPU_OK;
}

void* hvvr_gpu_host_alloc(uint64_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        setLastError("hipHostMalloc failed");
        return nullptr;
    }
    return p;
}

void hvvr_gpu_host_free(void* p) {
    if (p)
        (void)hipHostFree(p);
}

int hvvr_gpu_view_read_image(hvvr_gpu_view* v, uint32_t* image, uint32_t stride) {
    if (!image)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc)
        return rc;
    if (stride < v->width)
        return fail("stride smaller than image width");
    HVVR_HIP_CHECK(hipMemcpy2D(image, size_t(stride) * 4, v->image.ptr, size_t(v->width) * 4, size_t(v->width) * 4,
                               v->height, hipMemcpyDeviceToHost));
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_read_sample_colors(hvvr_gpu_view* v, uint32_t* colors) {
    if (!colors)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc)
        return rc;
    HVVR_HIP_CHECK(hipMemcpy(colors, v->sampleResults.ptr, v->sampleResults.bytes(), hipMemcpyDeviceToHost));
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_read_gbuffer(hvvr_gpu_view* v, uint32_t* triIndex, uint32_t* sampleMask) {
    if (!triIndex || !sampleMask)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc)
        return rc;
    const uint32_t A = v->msaaRate;
    std::vector<GBufferSubsample> raw(v->gbuffer.count);
    std::vector<uint32_t> counts(v->tileCount);
    HVVR_HIP_CHECK(hipMemcpy(raw.data(), v->gbuffer.ptr, v->gbuffer.bytes(), hipMemcpyDeviceToHost));
    HVVR_HIP_CHECK(hipMemcpy(counts.data(), v->tileCounts.ptr, counts.size() * 4, hipMemcpyDeviceToHost));
    const uint32_t maskAll = A < 32 ? ((1u << A) - 1u) : 0xffffffffu;
    for (uint32_t s = 0; s < v->sampleCount; s++) {
        uint32_t* ot = triIndex + size_t(s) * A;
        uint32_t* om = sampleMask + size_t(s) * A;
        for (uint32_t k = 0; k < A; k++) {
            ot[k] = badTriIndex;
            om[k] = 0;
        }
        if (counts[s / TILE_SIZE] == 0) { // empty tile: every subsample missed
            om[0] = maskAll;
            continue;
        }
        const GBufferSubsample* base = raw.data() + size_t(s / 64) * 64 * A;
        uint32_t lane = s % 64, combined = 0;
        for (uint32_t k = 0; k < A; k++) {
            if (combined == maskAll)
                break;
            GBufferSubsample g = base[k * 64 + lane];
            ot[k] = g.triIndex;
            om[k] = g.sampleMask;
            combined |= g.sampleMask;
            if (g.triIndex == badTriIndex)
                break;
        }
    }
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_read_hits(hvvr_gpu_view* v, float* hits) {
    if (!hits)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc)
        return rc;
    if (!v->hitsValid)
        return fail("the last frame was rendered without the hit output (hvvr_gpu_view_set_hit_output)",
                    HVVR_GPU_ERR_STATE);
    HVVR_HIP_CHECK(hipMemcpy(hits, v->hits.ptr, v->hits.bytes(), hipMemcpyDeviceToHost));
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_read_tile_lists(hvvr_gpu_view* v, uint32_t* counts, uint32_t* tris, uint64_t capacity,
                                  uint64_t* total) {
    if (!counts)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc)
        return rc;
    uint64_t t = 0;
    HVVR_HIP_CHECK(hipMemcpy(counts, v->tileCounts.ptr, size_t(v->tileCount) * 4, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < v->tileCount; i++)
        t += counts[i];
    if (total)
        *total = t;
    if (tris) {
        if (capacity < t)
            return fail("tile list capacity too small", HVVR_GPU_ERR_OVERFLOW);
        // lists are stored in allocation order: gather them into tile order
        uint64_t allocated = 0;
        HVVR_HIP_CHECK(hipMemcpy(&allocated, v->counters.ptr + 3, 8, hipMemcpyDeviceToHost));
        std::vector<uint32_t> offsets(v->tileCount), all(allocated);
        HVVR_HIP_CHECK(hipMemcpy(offsets.data(), v->tileOffsets.ptr, offsets.size() * 4, hipMemcpyDeviceToHost));
        if (allocated)
            HVVR_HIP_CHECK(hipMemcpy(all.data(), v->triIndices.ptr, size_t(allocated) * 4, hipMemcpyDeviceToHost));
        uint64_t at = 0;
        for (uint32_t i = 0; i < v->tileCount; i++) {
            if (uint64_t(offsets[i]) + counts[i] > allocated)
                return fail("tile list outside the allocated range", HVVR_GPU_ERR_STATE);
            if (counts[i])
                memcpy(tris + at, all.data() + offsets[i], size_t(counts[i]) * 4);
            at += counts[i];
        }
    }
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_get_stats(hvvr_gpu_view* v, hvvr_gpu_view_stats* st) {
    if (!st)
        return fail("null argument");
    int rc = viewSync(v);
    if (rc && rc != HVVR_GPU_ERR_OVERFLOW)
        return rc;
    uint64_t counters[16];
    hipStream_t stream = v->lastStream ? v->lastStream : v->ctx->stream;
    launchTileWorkReduce(v, stream);
    HVVR_HIP_CHECK(hipGetLastError());
    HVVR_HIP_CHECK(hipMemcpyAsync(counters, v->counters.ptr, sizeof(counters), hipMemcpyDeviceToHost, stream));
    HVVR_HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<uint32_t> counts(v->tileCount);
    HVVR_HIP_CHECK(hipMemcpy(counts.data(), v->tileCounts.ptr, counts.size() * 4, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (uint32_t c : counts)
        total += c;
    memset(st, 0, sizeof(*st));
    st->listEntries = total;
    st->survivorPairs = counters[0];
    st->overflow = counters[1] ? 1u : 0u;
    st->maxTileStack = uint32_t(counters[2]);
    st->nodeVisits = counters[4];
    st->largeTiles = uint32_t(counters[5]);
    st->hugeTiles = uint32_t(counters[6]);
    st->giantTiles = uint32_t(counters[7]);
    st->colossalTiles = uint32_t(counters[8]);
    st->walkTiles = uint32_t(counters[9]);
    st->testedPairs = counters[10];
    st->tileCount = v->tileCount;
    for (uint32_t c : counts)
        st->occupiedTiles += c ? 1u : 0u;
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_stage_times(hvvr_gpu_view* v, float* ms, int maxCount) {
    if (!v || !ms)
        return fail("null argument");
    if (!v->stagesRecorded)
        return fail("no profiled frame: enable profiling on the context before rendering", HVVR_GPU_ERR_STATE);
    HVVR_HIP_CHECK(hipEventSynchronize(v->stageEvents[kStageCount]));
    int n = std::min(maxCount, int(kStageCount));
    for (int i = 0; i < n; i++)
        HVVR_HIP_CHECK(hipEventElapsedTime(&ms[i], v->stageEvents[i], v->stageEvents[i + 1]));
    return n;
}

int hvvr_gpu_view_tile_list_kernel_time(hvvr_gpu_view* v, float* ms) {
    if (!v || !ms)
        return fail("null argument");
    if (!v->kernelTimed)
        return fail("no profiled frame: enable profiling on the context before rendering", HVVR_GPU_ERR_STATE);
    HVVR_HIP_CHECK(hipEventSynchronize(v->kernelEvents[1]));
    HVVR_HIP_CHECK(hipEventElapsedTime(ms, v->kernelEvents[0], v->kernelEvents[1]));
    return HVVR_GPU_OK;
}

} // extern "C"
