// Camera sample setup on the device: what Camera::setSamples builds (camera.cpp:21-64,
// sample_clustering.h:17-49, sample_hierarchy.cpp:46-190), so a new view shape (size,
// distortion, lens) needs no host-side setup.
//
//   1. samples: uploaded (Sample[] from the caller), or a width x height grid generated
//      here with the optional radial barrel distortion (samples.cpp:21-42);
//   2. naiveXYCluster: four rounds of stable sorts of the sample order, the whole set by
//      (y, x), then runs of 8192 x xStrata by (x, y), blocks of 8192 by (y, x), rows of
//      128 x blockXStrata by (x, y). Each round sorts 128-bit keys (run index, the primary
//      and secondary coordinates as order-preserving integers, the sample's position before
//      the round) with a bitonic network: the keys are unique, so the one sorted order is
//      the stable sort's, bit for bit the host's std::stable_sort;
//   3. per slot (padded to a multiple of 8192 with the last sample) the camera-space beam;
//      per tile of 128 slots the padded 2-D bounds and the fitted frustum; per block the
//      merged bounds of its 64 tiles and its frustum;
//   4. the image-location -> sample remap (the last sample of a location wins, as the
//      host's loop leaves it).
// The arithmetic is common/sample_setup.h, shared with the host setup: beams, frusta and
// remap equal the host's bit for bit for the perspective mapping (the spherical section's
// sinf / cosf may differ in the last bit between host and device libraries).

#include "../common/sample_setup.h"
#include "gpu_internal.h"

#include <algorithm>
#include <string.h>

namespace hvvr {

namespace {

__device__ inline uint32_t orderKey(float f) {
    // order-preserving integer image of a float (-0 taken as +0, as float comparison does)
    uint32_t u = f == 0.0f ? 0u : __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline bool keyGreater(const uint4& a, const uint4& b) {
    if (a.x != b.x)
        return a.x > b.x;
    if (a.y != b.y)
        return a.y > b.y;
    if (a.z != b.z)
        return a.z > b.z;
    return a.w > b.w;
}

__global__ void gridSamplesKernel(RawSample* out, uint32_t width, uint32_t height, float k1, float k2, int barrel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= width * height)
        return;
    RawSample s = gridSample(i % width, i / width, width, height);
    if (barrel)
        barrelDistort(s, width, height, k1, k2);
    out[i] = s;
}

__global__ void iotaKernel(uint32_t* ord, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        ord[i] = i;
}

// keys of one clustering round: run index of the position, primary, secondary, position
__global__ void clusterKeysKernel(uint4* keys, const uint32_t* ord, const RawSample* samples, uint32_t n,
                                  uint32_t padded, uint32_t runLength, int byY) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= padded)
        return;
    if (i >= n) {
        keys[i] = make_uint4(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu);
        return;
    }
    const RawSample s = samples[ord[i]];
    const float p = byY ? s.y : s.x, q = byY ? s.x : s.y;
    keys[i] = make_uint4(runLength ? i / runLength : 0u, orderKey(p), orderKey(q), i);
}

// one compare-exchange step (distance j) of the bitonic merge of size k, in global memory
__global__ void bitonicGlobalKernel(uint4* keys, uint32_t j, uint32_t k, uint32_t padded) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t l = i ^ j;
    if (i >= padded || l <= i)
        return;
    const uint4 a = keys[i], b = keys[l];
    const bool ascending = (i & k) == 0;
    if (keyGreater(a, b) == ascending) {
        keys[i] = b;
        keys[l] = a;
    }
}

enum : uint32_t { kLocalSort = 2048 };
// the steps of distance <= jStart of merges of size k (every merge size up to kLocalSort
// when kFull), inside aligned chunks of kLocalSort keys held in LDS
__global__ __launch_bounds__(1024) void bitonicLocalKernel(uint4* keys, uint32_t k, uint32_t jStart, int kFull) {
    __shared__ uint4 sh[kLocalSort];
    const uint32_t base = blockIdx.x * kLocalSort, t = threadIdx.x;
    sh[t] = keys[base + t];
    sh[t + 1024] = keys[base + t + 1024];
    __syncthreads();
    const uint32_t kFirst = kFull ? 2u : k, kLast = kFull ? kLocalSort : k;
    for (uint32_t kk = kFirst; kk <= kLast; kk <<= 1) {
        for (uint32_t j = kFull ? kk >> 1 : jStart; j > 0; j >>= 1) {
            // thread t takes the pair whose lower index has bit j clear
            const uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
            const uint32_t l = i | j;
            const bool ascending = ((base + i) & kk) == 0;
            const uint4 a = sh[i], b = sh[l];
            if (keyGreater(a, b) == ascending) {
                sh[i] = b;
                sh[l] = a;
            }
            __syncthreads();
        }
    }
    keys[base + t] = sh[t];
    keys[base + t + 1024] = sh[t + 1024];
}

__global__ void clusterGatherKernel(uint32_t* ordOut, const uint32_t* ordIn, const uint4* keys, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        ordOut[i] = ordIn[keys[i].w];
}

__global__ void beamsKernel(DirectionalBeam* beams, const RawSample* samples, const uint32_t* ord, uint32_t n,
                            uint32_t sampleCount, SampleMapping m) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= sampleCount)
        return;
    beams[slot] = sampleBeam(samples[ord[min(slot, n - 1u)]], m);
}

// per tile: the bounds of its 128 slots in slot order, and its frustum
__global__ void tileFrustaKernel(SimpleRayFrustum* frusta, Bounds2D* bounds, const RawSample* samples,
                                 const uint32_t* ord, uint32_t n, uint32_t tileCount, SampleMapping m) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tileCount)
        return;
    Bounds2D b;
    b.setEmpty();
    for (uint32_t i = 0; i < TILE_SIZE; i++)
        b.mergeSample(samples[ord[min(t * TILE_SIZE + i, n - 1u)]]);
    bounds[t] = b;
    frusta[t] = frustumFrom2D(b, m);
}

__global__ void blockFrustaKernel(SimpleRayFrustum* frusta, const Bounds2D* tileBounds, uint32_t blockCount,
                                  SampleMapping m) {
    const uint32_t blk = blockIdx.x * blockDim.x + threadIdx.x;
    if (blk >= blockCount)
        return;
    Bounds2D b;
    b.setEmpty();
    for (uint32_t t = 0; t < TILES_PER_BLOCK; t++)
        b.merge(tileBounds[blk * TILES_PER_BLOCK + t]);
    frusta[blk] = frustumFrom2D(b, m);
}

__global__ void fillKernel(int32_t* p, uint64_t n, int32_t value) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = value;
}

__global__ void remapScatterKernel(int32_t* remap, uint64_t remapCount, const RawSample* samples, const uint32_t* ord,
                                   uint32_t n, uint32_t width, uint32_t split) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t src = ord[i];
    const RawSample s = samples[src];
    const uint64_t idx = (uint64_t(s.pixelY) * width + s.pixelX) * split + src % split;
    if (idx < remapCount)
        atomicMax(remap + idx, int32_t(i));
}

// lower = min(1, positions), upper = max(0, positions) (camera.cpp:36-42)
__global__ __launch_bounds__(1024) void sampleBoundsKernel(const RawSample* samples, uint32_t n, float* out) {
    __shared__ float sh[4][1024];
    const uint32_t t = threadIdx.x;
    float lx = 1.0f, ly = 1.0f, ux = 0.0f, uy = 0.0f;
    for (uint32_t i = t; i < n; i += 1024) {
        const RawSample s = samples[i];
        lx = fminf(lx, s.x);
        ly = fminf(ly, s.y);
        ux = fmaxf(ux, s.x);
        uy = fmaxf(uy, s.y);
    }
    sh[0][t] = lx;
    sh[1][t] = ly;
    sh[2][t] = ux;
    sh[3][t] = uy;
    __syncthreads();
    for (uint32_t h = 512; h > 0; h >>= 1) {
        if (t < h) {
            sh[0][t] = fminf(sh[0][t], sh[0][t + h]);
            sh[1][t] = fminf(sh[1][t], sh[1][t + h]);
            sh[2][t] = fmaxf(sh[2][t], sh[2][t + h]);
            sh[3][t] = fmaxf(sh[3][t], sh[3][t + h]);
        }
        __syncthreads();
    }
    if (t < 4)
        out[t] = sh[t][0];
}

inline uint32_t groups(uint64_t n, uint32_t size) { return uint32_t((n + size - 1) / size); }

// sorts keys[0, padded) ascending (padded: a power of two >= kLocalSort)
void bitonicSort(uint4* keys, uint32_t padded, hipStream_t st) {
    const uint32_t chunks = padded / kLocalSort;
    bitonicLocalKernel<<<chunks, 1024, 0, st>>>(keys, 0, 0, 1);
    for (uint32_t k = kLocalSort * 2; k <= padded; k <<= 1) {
        for (uint32_t j = k >> 1; j >= kLocalSort; j >>= 1)
            bitonicGlobalKernel<<<groups(padded, 256), 256, 0, st>>>(keys, j, k, padded);
        bitonicLocalKernel<<<chunks, 1024, 0, st>>>(keys, k, kLocalSort >> 1, 0);
    }
}

int setupFail(const std::string& msg, int code = HVVR_GPU_ERR_INVALID) {
    setLastError(msg);
    return code;
}

// Builds a view from samples already on the device (RawSample[count]); frees nothing it
// did not allocate.
int buildDeviceSetup(hvvr_gpu_context* ctx, const hvvr_gpu_sample_setup_desc* d, const RawSample* samples,
                     uint32_t count, hvvr_gpu_view** out, float* boundsOut) {
    const uint32_t blockCount = (count + BLOCK_SIZE - 1) / BLOCK_SIZE;
    hvvr_gpu_view_desc vd = {};
    vd.msaaRate = d->msaaRate;
    vd.sampleCount = blockCount * BLOCK_SIZE;
    vd.validSampleCount = count;
    vd.width = d->width;
    vd.height = d->height;
    vd.split = d->split;
    vd.lensRadius = d->lensRadius;
    vd.focalDistance = d->focalDistance;
    hvvr_gpu_view* v = nullptr;
    int rc = makeView(ctx, &vd, true, &v);
    if (rc)
        return rc;
    SampleMapping m;
    memcpy(&m.sampleToCamera, d->sampleToCamera, sizeof(m.sampleToCamera));
    m.lens = {d->lensRadius, d->focalDistance};
    m.fovXDegrees = d->fovXDegrees;
    m.fovYDegrees = d->fovYDegrees;
    m.spherical = d->fovXDegrees > 0.0f ? 1 : 0;

    uint32_t padded = kLocalSort;
    while (padded < count)
        padded <<= 1;
    hipStream_t st = ctx->stream;
    DeviceBuffer<uint32_t> ord[2];
    DeviceBuffer<uint4> keys;
    DeviceBuffer<Bounds2D> tileBounds;
    DeviceBuffer<float> bounds;
    auto cleanup = [&](int code) {
        delete v;
        return code;
    };
    if (ord[0].resize(count) || ord[1].resize(count) || keys.resize(padded) || tileBounds.resize(v->tileCount) ||
        bounds.resize(4))
        return cleanup(setupFail("sample setup: device allocation failed", HVVR_GPU_ERR_HIP));
    iotaKernel<<<groups(count, 256), 256, 0, st>>>(ord[0].ptr, count);
    // naiveXYCluster (sample_clustering.h:17-49)
    const ClusterStrata cs = clusterStrata(blockCount);
    const struct {
        uint64_t run;
        int byY;
    } rounds[4] = {{0, 1},
                   {uint64_t(BLOCK_SIZE) * cs.xStrata, 0},
                   {BLOCK_SIZE, 1},
                   {uint64_t(TILE_SIZE) * cs.blockXStrata, 0}};
    int cur = 0;
    for (const auto& r : rounds) {
        const uint32_t run = r.run >= count ? 0u : uint32_t(r.run); // one run holding everything
        clusterKeysKernel<<<groups(padded, 256), 256, 0, st>>>(keys.ptr, ord[cur].ptr, samples, count, padded, run,
                                                               r.byY);
        bitonicSort(keys.ptr, padded, st);
        clusterGatherKernel<<<groups(count, 256), 256, 0, st>>>(ord[cur ^ 1].ptr, ord[cur].ptr, keys.ptr, count);
        cur ^= 1;
    }
    const uint32_t* sorted = ord[cur].ptr;
    beamsKernel<<<groups(v->sampleCount, 256), 256, 0, st>>>(v->beams.ptr, samples, sorted, count, v->sampleCount, m);
    tileFrustaKernel<<<groups(v->tileCount, 128), 128, 0, st>>>(v->tileFrustaCam.ptr, tileBounds.ptr, samples, sorted,
                                                                count, v->tileCount, m);
    blockFrustaKernel<<<groups(v->blockCount, 64), 64, 0, st>>>(v->blockFrustaCam.ptr, tileBounds.ptr, v->blockCount,
                                                                m);
    const uint64_t remapCount = v->remap.count;
    fillKernel<<<groups(remapCount, 256), 256, 0, st>>>(v->remap.ptr, remapCount, -1);
    remapScatterKernel<<<groups(count, 256), 256, 0, st>>>(v->remap.ptr, remapCount, samples, sorted, count, d->width,
                                                           d->split);
    sampleBoundsKernel<<<1, 1024, 0, st>>>(samples, count, bounds.ptr);
    float b[4];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(b, bounds.ptr, sizeof(b), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return cleanup(setupFail(std::string("sample setup: ") + hipGetErrorString(e), HVVR_GPU_ERR_HIP));
    if (boundsOut)
        memcpy(boundsOut, b, sizeof(b));
    *out = v;
    return HVVR_GPU_OK;
}

int checkSetupDesc(hvvr_gpu_context* ctx, const hvvr_gpu_sample_setup_desc* d, uint64_t count, hvvr_gpu_view** out) {
    if (!ctx || !d || !out)
        return setupFail("null argument");
    if (count == 0 || count > (1u << 28))
        return setupFail("sample count must be in 1 .. 2^28");
    if (d->split < 1 || d->split > 3 || d->width == 0 || d->height == 0)
        return setupFail("invalid image size or split");
    if (!(d->fovXDegrees >= 0.0f && d->fovYDegrees >= 0.0f && d->fovXDegrees <= 360.0f && d->fovYDegrees <= 180.0f))
        return setupFail("spherical section: angles in [0, 360] x [0, 180] degrees");
    return HVVR_GPU_OK;
}

} // namespace

} // namespace hvvr

using namespace hvvr;

extern "C" {

int hvvr_gpu_view_create_from_samples(hvvr_gpu_context* ctx, const hvvr_gpu_sample_setup_desc* d,
                                      const void* samples, uint32_t count, hvvr_gpu_view** out, float* boundsOut) {
    int rc = checkSetupDesc(ctx, d, count, out);
    if (rc)
        return rc;
    if (!samples)
        return setupFail("null samples");
    HVVR_HIP_CHECK(hipSetDevice(ctx->device));
    DeviceBuffer<RawSample> dev;
    HVVR_HIP_CHECK(dev.resize(count));
    HVVR_HIP_CHECK(hipMemcpyAsync(dev.ptr, samples, size_t(count) * sizeof(RawSample), hipMemcpyHostToDevice,
                                  ctx->stream));
    return buildDeviceSetup(ctx, d, dev.ptr, count, out, boundsOut);
}

int hvvr_gpu_view_create_grid(hvvr_gpu_context* ctx, const hvvr_gpu_sample_setup_desc* d, float k1, float k2,
                              hvvr_gpu_view** out, float* boundsOut) {
    const uint64_t count = d ? uint64_t(d->width) * d->height : 0;
    int rc = checkSetupDesc(ctx, d, count, out);
    if (rc)
        return rc;
    if (d->split != 1)
        return setupFail("grid samples have one colour sample per pixel (split 1)");
    HVVR_HIP_CHECK(hipSetDevice(ctx->device));
    DeviceBuffer<RawSample> dev;
    HVVR_HIP_CHECK(dev.resize(count));
    const int barrel = k1 != 0.0f || k2 != 0.0f;
    gridSamplesKernel<<<groups(count, 256), 256, 0, ctx->stream>>>(dev.ptr, d->width, d->height, k1, k2, barrel);
    HVVR_HIP_CHECK(hipGetLastError());
    return buildDeviceSetup(ctx, d, dev.ptr, uint32_t(count), out, boundsOut);
}

int hvvr_gpu_view_read_sample_setup(hvvr_gpu_view* v, void* beams, void* tileFrusta, void* blockFrusta,
                                    int32_t* remap) {
    if (!v)
        return setupFail("null view");
    HVVR_HIP_CHECK(hipSetDevice(v->ctx->device));
    hipStream_t st = v->ctx->stream;
    if (beams)
        HVVR_HIP_CHECK(hipMemcpyAsync(beams, v->beams.ptr, v->beams.bytes(), hipMemcpyDeviceToHost, st));
    if (tileFrusta)
        HVVR_HIP_CHECK(
            hipMemcpyAsync(tileFrusta, v->tileFrustaCam.ptr, v->tileFrustaCam.bytes(), hipMemcpyDeviceToHost, st));
    if (blockFrusta)
        HVVR_HIP_CHECK(
            hipMemcpyAsync(blockFrusta, v->blockFrustaCam.ptr, v->blockFrustaCam.bytes(), hipMemcpyDeviceToHost, st));
    if (remap && v->remap.ptr)
        HVVR_HIP_CHECK(hipMemcpyAsync(remap, v->remap.ptr, v->remap.bytes(), hipMemcpyDeviceToHost, st));
    HVVR_HIP_CHECK(hipStreamSynchronize(st));
    return HVVR_GPU_OK;
}

int hvvr_gpu_view_get_shape(hvvr_gpu_view* v, uint32_t shape[8]) {
    if (!v || !shape)
        return setupFail("null argument");
    const uint32_t s[8] = {v->msaaRate, v->sampleCount, v->validSampleCount, v->tileCount,
                           v->blockCount, v->width,       v->height,           v->split};
    memcpy(shape, s, sizeof(s));
    return HVVR_GPU_OK;
}

} // extern "C"
