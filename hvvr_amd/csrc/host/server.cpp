// Frame-request serving loop (include/hvvr_server.h) over the GPU C ABI.
//
// Scenes stay resident (the caller's hvvr_gpu_scene handles). Each request names a view
// shape; the server keeps a pool of GPU views per shape, each with its own HIP stream and
// a page-locked staging image. A request is admitted when a slot is free and a view of
// its shape is idle or fits the device budget (idle views of other shapes are evicted,
// least recently used first); its frame and the image copy are queued on the view's
// stream, so admitted requests overlap on the device. poll() retires finished requests
// and admits queued ones: requests that arrive while others run join as slots free up.

#include "host_internal.h"

#include "hvvr_gpu.h"
#include "hvvr_server.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <string.h>
#include <thread>
#include <unordered_map>

using namespace hvvr;

namespace {

double nowMs() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int serverFail(const std::string& msg, int code = HVVR_GPU_ERR_INVALID) {
    hvvrSetHostError(msg);
    return code;
}
// a failed GPU call: its message, for hvvr_host_last_error
int gpuFail(const char* what, int code) {
    hvvrSetHostError(std::string(what) + ": " + hvvr_gpu_last_error());
    return code;
}

struct Shape {
    bool foveated = false;
    uint32_t msaa = 1;
    uint32_t width = 0, height = 0;
    // grid shape: a host sample setup, or a grid built on the device per view
    SampleSetup setup;
    ThinLens lens = {0.0f, 1.0f};
    bool deviceGrid = false;
    hvvr_gpu_sample_setup_desc grid = {};
    float k1 = 0.0f, k2 = 0.0f;
    // foveated shape
    FoveatedSetup fov;
    matrix3x3 screenSampleToCam = matrix3x3::identity();
    vector2 boundsLower, boundsUpper;
    uint64_t bytes = 0;    // device bytes of one new view (hvvr_gpu_view_estimate_bytes, then
                           // what a view grown to listHint holds)
    uint64_t listHint = 0; // list entries a view of this shape has needed (0 = the default)
};

struct PooledView {
    hvvr_gpu_view* view = nullptr;
    uint32_t shape = 0;
    uint32_t* staging = nullptr; // page-locked width x height image
    uint64_t bytes = 0;
    double lastUsed = 0.0;
    bool busy = false;
    const hvvr_gpu_scene* lastScene = nullptr; // the scene of its last frame
};

struct Request {
    uint64_t id;
    hvvr_gpu_scene* scene;
    uint32_t shape;
    matrix4x4 cameraToWorld;
    vector3 eyeDir;
    uint32_t* image;
    double submitted = 0.0, completed = 0.0;
    PooledView* pv = nullptr;
    uint32_t launches = 0;
};

} // namespace

struct hvvr_server {
    hvvr_gpu_context* ctx = nullptr;
    uint64_t budget = 0, used = 0, peakUsed = 0;
    uint32_t maxInFlight = 1, peakInFlight = 0;
    int shadingBits = 32;
    uint32_t viewsCreated = 0, viewsEvicted = 0;
    uint64_t nextId = 1, submitted = 0;
    std::vector<Shape> shapes;
    std::vector<PooledView*> pool;
    std::deque<Request*> queue;
    std::vector<Request*> inFlight;
    std::unordered_map<uint64_t, double> done; // request id -> latency (ms)
    std::unordered_map<uint64_t, std::string> failed; // request id -> why
    std::vector<double> latencies;
    uint32_t listGrowths = 0;

    ~hvvr_server() {
        for (Request* r : inFlight)
            delete r;
        for (Request* r : queue)
            delete r;
        for (PooledView* pv : pool)
            release(pv);
    }

    void release(PooledView* pv) {
        if (pv->view)
            hvvr_gpu_view_destroy(pv->view);
        hvvr_gpu_host_free(pv->staging);
        delete pv;
    }

    int createView(uint32_t shapeId, PooledView*& out) {
        Shape& sh = shapes[shapeId];
        hvvr_gpu_view_desc d = {};
        d.msaaRate = sh.msaa;
        d.split = 1;
        d.width = sh.width;
        d.height = sh.height;
        if (sh.foveated) {
            d.sampleCount = sh.fov.sampleCount;
            d.validSampleCount = sh.fov.validSampleCount;
            d.beams = sh.fov.beams.data();
            d.tileFrusta = sh.fov.tileFrusta.data();
            d.blockFrusta = sh.fov.blockFrusta.data();
            d.remap = nullptr;
        } else {
            d.sampleCount = sh.setup.sampleCount;
            d.validSampleCount = sh.setup.validSampleCount;
            d.beams = sh.setup.beams.data();
            d.tileFrusta = sh.setup.tileFrusta.data();
            d.blockFrusta = sh.setup.blockFrusta.data();
            d.remap = sh.setup.remap.data();
            d.split = sh.setup.splitColorSamples;
            d.lensRadius = sh.lens.radius;
            d.focalDistance = sh.lens.focalDistance;
        }
        PooledView* pv = new PooledView();
        pv->shape = shapeId;
        int rc = sh.deviceGrid ? hvvr_gpu_view_create_grid(ctx, &sh.grid, sh.k1, sh.k2, &pv->view, nullptr)
                               : hvvr_gpu_view_create(ctx, &d, &pv->view);
        if (!rc && sh.foveated) {
            hvvr_gpu_foveated_desc f = {};
            f.polarRemap = (const uint32_t*)sh.fov.polarRemap.data();
            f.samplesPerRing = sh.fov.samplesPerRing;
            f.ringCount = sh.fov.ringCount;
            const EccentricityMap& e = sh.fov.emap;
            float em[6] = {e.m, e.w0, e.S, e.invLogA, e.invW0, e.invM};
            memcpy(f.eccentricityMap, em, sizeof(em));
            f.maxEccentricity = sh.fov.maxEccentricity;
            f.taaAlpha = 0.05f;
            f.taaStddevMultiplier = 1.0f;
            f.contrastEnhancement = 1;
            f.contrastFe = 1.0f;
            rc = hvvr_gpu_view_set_foveated(pv->view, &f);
        }
        if (!rc)
            rc = hvvr_gpu_view_set_own_stream(pv->view, 1);
        if (!rc)
            rc = hvvr_gpu_view_set_shading_precision(pv->view, shadingBits);
        if (!rc && sh.listHint)
            rc = hvvr_gpu_view_reserve_lists(pv->view, sh.listHint);
        if (!rc)
            rc = hvvr_gpu_view_device_bytes(pv->view, &pv->bytes);
        if (!rc) {
            pv->staging = (uint32_t*)hvvr_gpu_host_alloc(uint64_t(sh.width) * sh.height * 4);
            if (!pv->staging)
                rc = HVVR_GPU_ERR_HIP;
        }
        if (rc) {
            release(pv);
            return gpuFail("view creation", rc);
        }
        viewsCreated++;
        used += pv->bytes;
        peakUsed = std::max(peakUsed, used);
        pool.push_back(pv);
        out = pv;
        return HVVR_GPU_OK;
    }

    // an idle view of the shape, or a new one if the budget allows after evicting idle
    // views of other shapes (least recently used first); null when none fits yet. An idle
    // view whose last frame was of the same scene comes first: a view keeps per-tile
    // hints of the tile-list configuration each tile needed on its last scene, and a
    // dense scene's frame without them walks its deep tiles through every configuration.
    int acquire(uint32_t shapeId, const hvvr_gpu_scene* scene, PooledView*& out) {
        out = nullptr;
        for (PooledView* pv : pool)
            if (!pv->busy && pv->shape == shapeId && (!out || pv->lastScene == scene)) {
                out = pv;
                if (pv->lastScene == scene)
                    break;
            }
        if (out)
            return HVVR_GPU_OK;
        const uint64_t need = shapes[shapeId].bytes;
        while (used + need > budget) {
            PooledView* victim = nullptr;
            for (PooledView* pv : pool)
                if (!pv->busy && (!victim || pv->lastUsed < victim->lastUsed))
                    victim = pv;
            if (!victim)
                return HVVR_GPU_OK; // wait for a request to finish
            used -= victim->bytes;
            pool.erase(std::find(pool.begin(), pool.end(), victim));
            release(victim);
            viewsEvicted++;
        }
        return createView(shapeId, out);
    }

    int launch(Request* r) {
        PooledView* pv = r->pv;
        const Shape& sh = shapes[r->shape];
        r->launches++;
        int rc;
        if (sh.foveated) {
            FoveatedHistory h; // every request is a first frame: no history to reproject
            FoveatedFrame fr;
            foveatedFrame(r->cameraToWorld, r->eyeDir, sh.screenSampleToCam, sh.boundsLower, sh.boundsUpper, h, fr);
            hvvr_gpu_foveated_frame f;
            memcpy(f.eyeToWorld, &fr.eyeToWorld, sizeof(f.eyeToWorld));
            memcpy(f.cullPlanes, fr.cullPlanes, sizeof(f.cullPlanes));
            memcpy(f.sampleToEye, &fr.sampleToEye, sizeof(f.sampleToEye));
            memcpy(f.eyeToSamplePrevious, &fr.eyeToSamplePrevious, sizeof(f.eyeToSamplePrevious));
            rc = hvvr_gpu_view_render_foveated(pv->view, r->scene, &f, nullptr);
        } else {
            rc = hvvr_gpu_view_render(pv->view, r->scene, (const float*)&r->cameraToWorld, nullptr);
        }
        if (!rc)
            rc = hvvr_gpu_view_read_image_async(pv->view, pv->staging, sh.width);
        return rc ? gpuFail("render", rc) : HVVR_GPU_OK;
    }

    // Grows a view's list storage to `entries` (4 bytes each) within the budget, evicting
    // idle views of any shape first. Returns 1 (nothing changed) when even then the growth
    // does not fit.
    int growLists(PooledView* pv, uint64_t entries, uint64_t capacity) {
        Shape& sh = shapes[pv->shape];
        const uint64_t growth = entries > capacity ? (entries - capacity) * 4u : 0u;
        while (used + growth > budget) {
            PooledView* victim = nullptr;
            for (PooledView* v : pool)
                if (!v->busy && v != pv && (!victim || v->lastUsed < victim->lastUsed))
                    victim = v;
            if (!victim)
                return 1;
            used -= victim->bytes;
            pool.erase(std::find(pool.begin(), pool.end(), victim));
            release(victim);
            viewsEvicted++;
        }
        int rc = hvvr_gpu_view_reserve_lists(pv->view, entries);
        uint64_t bytes = 0;
        if (!rc)
            rc = hvvr_gpu_view_device_bytes(pv->view, &bytes);
        if (rc)
            return gpuFail("list storage", rc);
        used = used - pv->bytes + bytes;
        peakUsed = std::max(peakUsed, used);
        pv->bytes = bytes;
        if (entries > sh.listHint) {
            // views of one shape differ only in list storage: new ones start at this size
            sh.listHint = entries;
            sh.bytes = bytes;
        }
        listGrowths++;
        return HVVR_GPU_OK;
    }

    int poll() {
        int completed = 0;
        for (size_t i = 0; i < inFlight.size();) {
            Request* r = inFlight[i];
            int q = hvvr_gpu_view_query(r->pv->view);
            if (q < 0)
                return gpuFail("query", q);
            if (q == 0) {
                i++;
                continue;
            }
            Shape& sh = shapes[r->shape];
            uint64_t need = 0, cap = 0;
            int overflowed = 0;
            int rc = hvvr_gpu_view_list_demand(r->pv->view, &need, &cap, &overflowed);
            if (rc)
                return gpuFail("list demand", rc);
            const char* failure = nullptr;
            if (overflowed && need > cap && r->launches == 1) {
                // the lists did not fit: grow the view's storage (and that of the shape's
                // new views) and render the frame again; list indices are 32-bit
                const uint64_t want = need + need / 4;
                if (need > 0xffffffffull) {
                    failure = "the frame's triangle lists exceed 2^32 - 1 entries";
                } else {
                    rc = growLists(r->pv, std::min<uint64_t>(want, 0xffffffffull), cap);
                    if (rc < 0)
                        return rc;
                    if (rc == 0) {
                        rc = launch(r);
                        if (rc)
                            return rc;
                        i++;
                        continue;
                    }
                    failure = "the frame's triangle lists exceed the server's device budget";
                }
            }
            r->completed = nowMs();
            r->pv->busy = false;
            r->pv->lastUsed = r->completed;
            if (failure) {
                failed[r->id] = failure;
            } else if (overflowed) {
                failed[r->id] = need > cap ? "the frame's triangle lists overflowed their storage again"
                                           : "the frame overflowed a tile's traversal stack";
            } else {
                memcpy(r->image, r->pv->staging, size_t(sh.width) * sh.height * 4);
                double lat = r->completed - r->submitted;
                done[r->id] = lat;
                latencies.push_back(lat);
            }
            delete r;
            inFlight.erase(inFlight.begin() + long(i));
            completed++;
        }
        while (!queue.empty() && inFlight.size() < maxInFlight) {
            Request* r = queue.front();
            PooledView* pv = nullptr;
            int rc = acquire(r->shape, r->scene, pv);
            if (rc)
                return rc;
            if (!pv)
                break; // memory frees up as requests finish
            pv->busy = true;
            pv->lastUsed = nowMs();
            pv->lastScene = r->scene;
            r->pv = pv;
            queue.pop_front();
            rc = launch(r);
            if (rc) {
                pv->busy = false;
                delete r;
                return rc;
            }
            inFlight.push_back(r);
            peakInFlight = std::max(peakInFlight, uint32_t(inFlight.size()));
        }
        return completed;
    }
};

extern "C" {

int hvvr_server_create(hvvr_gpu_context* ctx, const hvvr_server_desc* d, hvvr_server** out) {
    if (!ctx || !d || !out || d->maxInFlight < 1 || d->maxInFlight > 64 || d->deviceBudgetBytes == 0)
        return serverFail("invalid server description");
    if (d->shadingPrecision != 0 && d->shadingPrecision != 16 && d->shadingPrecision != 32)
        return serverFail("shading precision must be 16 or 32");
    hvvr_server* s = new hvvr_server();
    s->shadingBits = d->shadingPrecision ? int(d->shadingPrecision) : 32;
    s->ctx = ctx;
    s->budget = d->deviceBudgetBytes;
    s->maxInFlight = d->maxInFlight;
    *out = s;
    return HVVR_GPU_OK;
}

int hvvr_server_destroy(hvvr_server* s) {
    if (!s)
        return HVVR_GPU_OK;
    while (!s->inFlight.empty()) {
        if (s->poll() < 0)
            break;
        std::this_thread::yield();
    }
    delete s;
    return HVVR_GPU_OK;
}

int hvvr_server_add_shape(hvvr_server* s, const hvvr_host_view* v, uint32_t msaaRate, uint32_t* id) {
    if (!s || !v || !id || v->setup.sampleCount == 0)
        return serverFail("invalid shape");
    Shape sh;
    sh.msaa = msaaRate;
    sh.width = v->setup.width;
    sh.height = v->setup.height;
    sh.setup = v->setup;
    sh.lens = v->lens;
    hvvr_gpu_view_desc d = {};
    d.msaaRate = msaaRate;
    d.sampleCount = sh.setup.sampleCount;
    d.validSampleCount = sh.setup.validSampleCount;
    d.remap = sh.setup.remap.data();
    d.width = sh.width;
    d.height = sh.height;
    d.split = sh.setup.splitColorSamples;
    d.lensRadius = sh.lens.radius;
    int rc = hvvr_gpu_view_estimate_bytes(&d, nullptr, &sh.bytes);
    if (rc)
        return gpuFail("shape", rc);
    if (sh.bytes > s->budget)
        return serverFail("a view of this shape exceeds the server's device budget", HVVR_GPU_ERR_OVERFLOW);
    s->shapes.push_back(std::move(sh));
    *id = uint32_t(s->shapes.size() - 1);
    return HVVR_GPU_OK;
}

int hvvr_server_add_grid_shape(hvvr_server* s, const hvvr_gpu_sample_setup_desc* g, float k1, float k2,
                               uint32_t* id) {
    if (!s || !g || !id || g->width == 0 || g->height == 0 || g->split != 1)
        return serverFail("invalid grid shape");
    Shape sh;
    sh.deviceGrid = true;
    sh.grid = *g;
    sh.k1 = k1;
    sh.k2 = k2;
    sh.msaa = g->msaaRate;
    sh.width = g->width;
    sh.height = g->height;
    sh.lens = {g->lensRadius, g->focalDistance};
    const uint64_t count = uint64_t(g->width) * g->height;
    hvvr_gpu_view_desc d = {};
    d.msaaRate = g->msaaRate;
    d.sampleCount = uint32_t((count + BLOCK_SIZE - 1) / BLOCK_SIZE * BLOCK_SIZE);
    d.validSampleCount = uint32_t(count);
    d.width = g->width;
    d.height = g->height;
    d.split = 1;
    d.lensRadius = g->lensRadius;
    int rc = hvvr_gpu_view_estimate_bytes(&d, nullptr, &sh.bytes);
    if (rc)
        return gpuFail("shape", rc);
    sh.bytes += count * 4; // the remap (the estimate counts it when the description carries one)
    if (sh.bytes > s->budget)
        return serverFail("a view of this shape exceeds the server's device budget", HVVR_GPU_ERR_OVERFLOW);
    s->shapes.push_back(std::move(sh));
    *id = uint32_t(s->shapes.size() - 1);
    return HVVR_GPU_OK;
}

int hvvr_server_add_foveated_shape(hvvr_server* s, const hvvr_host_foveated* f, const hvvr_host_view* screen,
                                   uint32_t msaaRate, uint32_t* id) {
    if (!s || !f || !screen || !id || screen->setup.width == 0 || screen->setup.height == 0)
        return serverFail("invalid foveated shape");
    Shape sh;
    sh.foveated = true;
    sh.msaa = msaaRate;
    sh.width = screen->setup.width;
    sh.height = screen->setup.height;
    sh.fov = f->setup;
    sh.screenSampleToCam = screen->sampleToCam;
    sh.boundsLower = screen->setup.boundsLower;
    sh.boundsUpper = screen->setup.boundsUpper;
    hvvr_gpu_view_desc d = {};
    d.msaaRate = msaaRate;
    d.sampleCount = sh.fov.sampleCount;
    d.validSampleCount = sh.fov.validSampleCount;
    d.width = sh.width;
    d.height = sh.height;
    d.split = 1;
    hvvr_gpu_foveated_desc fd = {};
    fd.samplesPerRing = sh.fov.samplesPerRing;
    fd.ringCount = sh.fov.ringCount;
    int rc = hvvr_gpu_view_estimate_bytes(&d, &fd, &sh.bytes);
    if (rc)
        return gpuFail("shape", rc);
    if (sh.bytes > s->budget)
        return serverFail("a view of this shape exceeds the server's device budget", HVVR_GPU_ERR_OVERFLOW);
    s->shapes.push_back(std::move(sh));
    *id = uint32_t(s->shapes.size() - 1);
    return HVVR_GPU_OK;
}

int hvvr_server_shape_bytes(hvvr_server* s, uint32_t shapeId, uint64_t* bytes) {
    if (!s || !bytes || shapeId >= s->shapes.size())
        return serverFail("invalid shape");
    *bytes = s->shapes[shapeId].bytes;
    return HVVR_GPU_OK;
}

int hvvr_server_submit(hvvr_server* s, hvvr_gpu_scene* scene, uint32_t shapeId, const float m[16],
                       const float* eyeDir, uint32_t* image, uint64_t* requestId) {
    if (!s || !scene || !m || !image || shapeId >= s->shapes.size())
        return serverFail("invalid request");
    Request* r = new Request();
    r->id = s->nextId++;
    r->scene = scene;
    r->shape = shapeId;
    memcpy(&r->cameraToWorld, m, 64);
    r->eyeDir = eyeDir ? vector3(eyeDir[0], eyeDir[1], eyeDir[2]) : vector3(0.0f, 0.0f, -1.0f);
    r->image = image;
    r->submitted = nowMs();
    s->queue.push_back(r);
    s->submitted++;
    if (requestId)
        *requestId = r->id;
    return HVVR_GPU_OK;
}

int hvvr_server_poll(hvvr_server* s) {
    if (!s)
        return serverFail("null server");
    return s->poll();
}

int hvvr_server_drain(hvvr_server* s) {
    if (!s)
        return serverFail("null server");
    while (!s->queue.empty() || !s->inFlight.empty()) {
        int rc = s->poll();
        if (rc < 0)
            return rc;
        if (rc == 0)
            std::this_thread::yield();
    }
    return HVVR_GPU_OK;
}

int hvvr_server_request_done(hvvr_server* s, uint64_t id, double* latencyMs) {
    if (!s)
        return serverFail("null server");
    auto f = s->failed.find(id);
    if (f != s->failed.end())
        return serverFail("request " + std::to_string(id) + " failed: " + f->second, HVVR_GPU_ERR_OVERFLOW);
    auto it = s->done.find(id);
    if (it == s->done.end())
        return 0;
    if (latencyMs)
        *latencyMs = it->second;
    return 1;
}

int hvvr_server_get_stats(hvvr_server* s, hvvr_server_stats* st) {
    if (!s || !st)
        return serverFail("null argument");
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->latencies.size();
    st->queued = s->queue.size();
    st->inFlight = s->inFlight.size();
    st->deviceBytes = s->used;
    st->peakDeviceBytes = s->peakUsed;
    st->viewsCreated = s->viewsCreated;
    st->viewsEvicted = s->viewsEvicted;
    st->peakInFlight = s->peakInFlight;
    st->failed = s->failed.size();
    st->listGrowths = s->listGrowths;
    if (!s->latencies.empty()) {
        std::vector<double> l = s->latencies;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) {
            size_t k = size_t(p * double(l.size() - 1) + 0.5);
            return l[std::min(k, l.size() - 1)];
        };
        st->p50Ms = pct(0.50);
        st->p99Ms = pct(0.99);
        double sum = 0.0;
        for (double x : l)
            sum += x;
        st->meanMs = sum / double(l.size());
    }
    return HVVR_GPU_OK;
}

} // extern "C"
