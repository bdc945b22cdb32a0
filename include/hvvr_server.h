#ifndef HVVR_SERVER_H
#define HVVR_SERVER_H
/*
 * Frame-request serving loop over the GPU C ABI. It has no counterpart in the reference,
 * which renders a fixed set of cameras per frame (render.cpp:204-219); it is the serving
 * layer of this build: scenes stay resident on the device, view requests of uneven
 * resolution, MSAA rate and sampling (grid or polar foveated) are admitted by their HBM
 * footprint, and requests that arrive while others are on the GPU are admitted as slots
 * and memory free up, each on its own HIP stream, so their frames overlap on the device.
 *
 * Single-threaded: the caller submits requests and polls; nothing runs behind its back.
 * Errors: 0 or a negative HVVR_GPU_ERR_* code; hvvr_host_last_error() describes it.
 */
#include "hvvr_gpu.h"
#include "hvvr_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hvvr_server hvvr_server;

typedef struct {
    uint64_t deviceBudgetBytes; /* device memory the server may hold in views (scenes are the caller's) */
    uint32_t maxInFlight;       /* requests on the device at once, 1..64 */
    uint32_t shadingPrecision;  /* 32 (default when 0) or 16: hvvr_gpu_view_set_shading_precision */
} hvvr_server_desc;

int hvvr_server_create(hvvr_gpu_context* ctx, const hvvr_server_desc* desc, hvvr_server** out);
/* Waits for the requests in flight, then frees every view. */
int hvvr_server_destroy(hvvr_server* server);

/* A view shape requests name: the sample setup of a host view (width x height image) at an
 * MSAA rate. The server copies what it needs; the host view may be destroyed afterwards. */
int hvvr_server_add_shape(hvvr_server* server, const hvvr_host_view* view, uint32_t msaaRate, uint32_t* shapeId);
/* A grid shape whose views build their sample hierarchy on the device
 * (hvvr_gpu_view_create_grid: width x height grid samples, barrel distortion k1 k2, the
 * descriptor's viewport mapping and lens): adding it does no host-side sample setup. */
int hvvr_server_add_grid_shape(hvvr_server* server, const hvvr_gpu_sample_setup_desc* grid, float k1, float k2,
                               uint32_t* shapeId);
/* A polar foveated shape: the polar sample set, reconstructed into `screen`'s image, whose
 * sample space and bounds give the eye transforms and cull planes. */
int hvvr_server_add_foveated_shape(hvvr_server* server, const hvvr_host_foveated* fov, const hvvr_host_view* screen,
                                   uint32_t msaaRate, uint32_t* shapeId);
/* Device bytes one view of the shape holds. */
int hvvr_server_shape_bytes(hvvr_server* server, uint32_t shapeId, uint64_t* bytes);

/* Queue one frame: scene (resident, updated), shape, camera-to-world matrix, gaze direction
 * in camera space (foveated shapes; NULL = straight ahead) and the host image it is
 * written to (width x height RGBA8 of the shape), valid until the request completes.
 * Requests are admitted first come, first served. */
int hvvr_server_submit(hvvr_server* server, hvvr_gpu_scene* scene, uint32_t shapeId, const float cameraToWorld[16],
                       const float* eyeDir, uint32_t* image, uint64_t* requestId);
/* Completes the finished requests (their images are written), then admits queued ones
 * while a slot is free and their views fit the budget (idle views of other shapes are
 * evicted, oldest first). A request takes an idle view of its shape whose last frame was
 * of the same scene when there is one (it keeps that scene's per-tile traversal hints).
 * Returns the number completed, or < 0 on error. */
int hvvr_server_poll(hvvr_server* server);
/* Polls until nothing is queued or in flight. */
int hvvr_server_drain(hvvr_server* server);
/* 1 once the request has completed, 0 if not yet; *latencyMs (may be NULL) = submit to
 * image written, in milliseconds. A request whose frame overflowed a tile's traversal
 * stack fails: HVVR_GPU_ERR_OVERFLOW (its image is not written). A frame whose triangle
 * lists overflowed is rendered again once, after the view's list storage has grown to
 * what it needed (and new views of its shape start at that size). */
int hvvr_server_request_done(hvvr_server* server, uint64_t requestId, double* latencyMs);

typedef struct {
    uint64_t submitted, completed, queued, inFlight;
    uint64_t deviceBytes;        /* held in views now */
    uint64_t peakDeviceBytes;
    uint32_t viewsCreated, viewsEvicted, peakInFlight;
    double p50Ms, p99Ms, meanMs; /* submit-to-image latency over completed requests */
    uint64_t failed;             /* requests that failed (not in completed) */
    uint32_t listGrowths;        /* frames rendered again after their list storage grew */
} hvvr_server_stats;
int hvvr_server_get_stats(hvvr_server* server, hvvr_server_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
