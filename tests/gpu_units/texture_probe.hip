// GPU probe of the texture sampler (hvvr_amd/csrc/gpu/texture.h): texSampleGrad and texLod
// on caller-given queries, one thread per query, for tests/test_texture.py to compare with
// the oracle. Test infrastructure only.

#include <hip/hip_runtime.h>

#include "texture.h"

using namespace hvvr;

namespace {

// q[6 i ..]: u, v, ddx.x, ddx.y, ddy.x, ddy.y
__global__ void sampleProbe(const uint8_t* texels, const TexInfo* info, const float* tables, const float* q, uint32_t n,
                            float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* p = q + 6 * size_t(i);
    const vector4 c = texSampleGrad(texels, *info, tables, tables + 256, vector2(p[0], p[1]), vector2(p[2], p[3]),
                                    vector2(p[4], p[5]));
    out[4 * size_t(i) + 0] = c.x;
    out[4 * size_t(i) + 1] = c.y;
    out[4 * size_t(i) + 2] = c.z;
    out[4 * size_t(i) + 3] = c.w;
}

__global__ void lodProbe(const float* tables, const float* x, uint32_t n, uint32_t levels, uint32_t* level, float* f) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    texLod(x[i], levels, tables + 256, level[i], f[i]);
}

template <class T>
int upload(T** d, const void* h, size_t bytes) {
    if (hipMalloc(reinterpret_cast<void**>(d), bytes ? bytes : 4) != hipSuccess)
        return -1;
    return (bytes && hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) ? -2 : 0;
}

} // namespace

// chain: the texture's stored levels, info.offset relative to it; tables: [512] sRGB table
// then mip-fraction thresholds
extern "C" int texture_probe_sample(const uint8_t* chain, uint64_t chainBytes, const void* info, const float* tables,
                                    const float* queries, uint32_t n, float* out) {
    const TexInfo* hi = static_cast<const TexInfo*>(info);
    if (hi->levels == 0 || hi->levels > kMaxTexLevels || n == 0)
        return -5;
    uint8_t* dTex = nullptr;
    TexInfo* dInfo = nullptr;
    float *dTab = nullptr, *dQ = nullptr, *dOut = nullptr;
    int rc = upload(&dTex, chain, chainBytes);
    if (!rc)
        rc = upload(&dInfo, info, sizeof(TexInfo));
    if (!rc)
        rc = upload(&dTab, tables, 512 * sizeof(float));
    if (!rc)
        rc = upload(&dQ, queries, size_t(n) * 6 * sizeof(float));
    if (!rc && hipMalloc(reinterpret_cast<void**>(&dOut), size_t(n) * 16) != hipSuccess)
        rc = -1;
    if (!rc) {
        sampleProbe<<<(n + 127) / 128, 128>>>(dTex, dInfo, dTab, dQ, n, dOut);
        if (hipDeviceSynchronize() != hipSuccess)
            rc = -3;
    }
    if (!rc && hipMemcpy(out, dOut, size_t(n) * 16, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -4;
    (void)hipFree(dTex);
    (void)hipFree(dInfo);
    (void)hipFree(dTab);
    (void)hipFree(dQ);
    (void)hipFree(dOut);
    return rc;
}

extern "C" int texture_probe_lod(const float* tables, const float* x, uint32_t n, uint32_t levels, uint32_t* level,
                                 float* f) {
    if (n == 0 || levels == 0)
        return -5;
    float *dTab = nullptr, *dX = nullptr, *dF = nullptr;
    uint32_t* dL = nullptr;
    int rc = upload(&dTab, tables, 512 * sizeof(float));
    if (!rc)
        rc = upload(&dX, x, size_t(n) * 4);
    if (!rc && (hipMalloc(reinterpret_cast<void**>(&dL), size_t(n) * 4) != hipSuccess ||
                hipMalloc(reinterpret_cast<void**>(&dF), size_t(n) * 4) != hipSuccess))
        rc = -1;
    if (!rc) {
        lodProbe<<<(n + 127) / 128, 128>>>(dTab, dX, n, levels, dL, dF);
        if (hipDeviceSynchronize() != hipSuccess)
            rc = -3;
    }
    if (!rc && (hipMemcpy(level, dL, size_t(n) * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(f, dF, size_t(n) * 4, hipMemcpyDeviceToHost) != hipSuccess))
        rc = -4;
    (void)hipFree(dTab);
    (void)hipFree(dX);
    (void)hipFree(dL);
    (void)hipFree(dF);
    return rc;
}
