// Ray dump of one view (hvvr_gpu_view_dump_rays): every subsample's world-space ray, as
// DumpRaysKernel does (raycaster/intersect.cu:407-441, GPUCamera::dumpRays,
// gpu_camera.cu dumpRays), for cross-checks with external raytracers through the RFF
// files the host writes (raycaster/render.cpp:28-67).
//
// One thread per output sample: the sample's beam in world space (the camera rotation
// applied to centre ray and differentials, gpu_samples.h:78-82), then per subsample
// direction = normalize(centre + du * a.x + dv * a.y) from the view's subsample offsets
// and origin = the camera position (the reference leaves the thin-lens origin as a TODO, so
// depth-of-field views dump pinhole rays too). In scanline order output slot i reads the
// sample the remap gives image location i; a location with no sample gets NaN rays (the
// reference leaves them unwritten), which the RFF writers drop.

#include "gpu_internal.h"

#include <algorithm>
#include <string.h>

namespace hvvr {

struct DumpRaysArgs {
    const DirectionalBeam* beams;
    const int32_t* remap; // null: sample order
    SimpleRay* out;
    uint32_t count;       // output samples: valid samples, or remap entries in scanline order
    uint32_t sampleCount; // valid samples
    uint32_t msaaRate;
    matrix4x4 cameraToWorld;
    vector2 offsets[32];
};

__global__ __launch_bounds__(128) void dumpRaysKernel(DumpRaysArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count)
        return;
    uint32_t sample = i;
    if (a.remap) {
        const int32_t s = a.remap[i];
        sample = s < 0 ? 0xffffffffu : uint32_t(s);
    }
    SimpleRay* out = a.out + size_t(i) * a.msaaRate;
    if (sample >= a.sampleCount) {
        const float nan = __builtin_nanf("");
        for (uint32_t k = 0; k < a.msaaRate; k++)
            out[k] = {vector3(nan), vector3(nan)};
        return;
    }
    const matrix3x3 rot = a.cameraToWorld.upper3x3();
    const DirectionalBeam b = a.beams[sample];
    const vector3 c = rot * b.centerRay, du = rot * b.du, dv = rot * b.dv;
    const vector3 origin = a.cameraToWorld.m3.xyz();
    for (uint32_t k = 0; k < a.msaaRate; k++) {
        const vector2 o = a.offsets[k];
        out[k] = {origin, normalize(c + du * o.x + dv * o.y)};
    }
}

} // namespace hvvr

using namespace hvvr;

extern "C" int hvvr_gpu_view_dump_rays(hvvr_gpu_view* v, const float* cameraToWorld, int scanlineOrder, void* rays,
                                       uint64_t capacity) {
    if (!v || !cameraToWorld || !rays) {
        setLastError("null argument");
        return HVVR_GPU_ERR_INVALID;
    }
    if (scanlineOrder && !v->remap.ptr) {
        setLastError("scanline order needs the view's remap (a foveated view has none)");
        return HVVR_GPU_ERR_INVALID;
    }
    DumpRaysArgs a;
    a.beams = v->beams.ptr;
    a.remap = scanlineOrder ? v->remap.ptr : nullptr;
    a.sampleCount = v->validSampleCount;
    a.count = scanlineOrder ? uint32_t(std::min<size_t>(v->remap.count, v->validSampleCount)) : v->validSampleCount;
    a.msaaRate = v->msaaRate;
    memcpy(&a.cameraToWorld, cameraToWorld, sizeof(a.cameraToWorld));
    memcpy(a.offsets, v->offsets, sizeof(a.offsets));
    const uint64_t n = uint64_t(a.count) * a.msaaRate;
    if (capacity < n) {
        setLastError("ray buffer holds " + std::to_string(capacity) + " rays, the view dumps " + std::to_string(n));
        return HVVR_GPU_ERR_INVALID;
    }
    HVVR_HIP_CHECK(hipSetDevice(v->ctx->device));
    hipStream_t st = v->ctx->stream;
    DeviceBuffer<SimpleRay> out;
    HVVR_HIP_CHECK(out.resize(n ? n : 1));
    a.out = out.ptr;
    if (a.count)
        dumpRaysKernel<<<(a.count + 127) / 128, 128, 0, st>>>(a);
    HVVR_HIP_CHECK(hipGetLastError());
    HVVR_HIP_CHECK(hipMemcpyAsync(rays, out.ptr, n * sizeof(SimpleRay), hipMemcpyDeviceToHost, st));
    HVVR_HIP_CHECK(hipStreamSynchronize(st));
    return HVVR_GPU_OK;
}
