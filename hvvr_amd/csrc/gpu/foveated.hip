// Polar foveated reconstruction (raycaster/gpu_foveated.cu, remap.cu:127-158):
//   polarRemapKernel      scatter of the shaded samples and their t-max into the polar
//                         image (azimuth x ring), RemapPolarFoveatedKernel
//   momentsKernel         3x3 mean and mean square of the polar image per texel, stored
//                         as 16-bit unorm, ComputeMoments
//   reprojectKernel       per screen pixel: bicubic polar lookup, reprojection into the
//                         previous frame, neighbourhood-clamped temporal blend,
//                         FoveatedPolarToScreenSpaceKernel
//   eccentricityBlurKernel separable 11-tap blur whose falloff sharpens with eccentricity,
//                         SeparableFilterUsingEccentricity (run twice)
//   contrastKernel        contrast enhancement against the blurred image,
//                         FinishConstrastEnhancement
//
// The reference samples CUDA textures. Here every image is a plain array in HBM and the
// texture unit's behaviour is restated: normalized coordinates, x = u*N - 0.5, texel
// floor(x) and the next, the blend weight frac(x) held with 8 fractional bits (so fetches
// at texel centres return the texel); wrap (u mod 1) across the polar azimuth, clamp at
// every other edge; sRGB texels decoded to linear per texel before the blend. The decode
// is a 256-entry table computed on the host with the sRGB formula (util_graphics.h:16-18).
// Each block stages it in LDS. The images are small (the polar image of a 315k-sample
// pattern is 1.2 MB), so their reads are served from L2.

#include "../common/foveated.h"
#include "device_common.h"
#include "gpu_internal.h"

namespace hvvr {

namespace {

enum : uint32_t { kFovBlockX = 16, kFovBlockY = 16 };

// float -> u32 as v_cvt_u32_f32 converts: NaN and negatives give 0, saturating at 2^32-1
__device__ inline uint32_t cvtU32(float x) {
    return x > 0.0f ? (x < 4294967040.0f ? uint32_t(x) : 0xffffffffu) : 0u;
}

// float -> i32 for texel indices: NaN gives 0, the range is limited far beyond any image
__device__ inline int texelIndex(float fx) {
    return (fx >= -1.0e9f && fx <= 1.0e9f) ? int(fx) : 0;
}

// ToColor4Unorm8SRgb (shading_helpers.h:193-200): rgb sRGB-encoded, every channel
// scaled by 255 and truncated, alpha linear
__device__ inline uint32_t toColor4Unorm8SRgb(const vector4& c) {
    uint32_t r = cvtU32(linearToSRgb(c.x) * 255.0f);
    uint32_t g = cvtU32(linearToSRgb(c.y) * 255.0f);
    uint32_t b = cvtU32(linearToSRgb(c.z) * 255.0f);
    uint32_t a = cvtU32(c.w * 255.0f);
    return r | (g << 8) | (b << 16) | (a << 24);
}

// ToColor4Unorm16 (shading_helpers.h:201-208)
__device__ inline uint64_t toColor4Unorm16(const vector4& c) {
    const float C = 65535.0f;
    uint64_t r = cvtU32(c.x * C), g = cvtU32(c.y * C), b = cvtU32(c.z * C), a = cvtU32(c.w * C);
    return r | (g << 16) | (b << 32) | (a << 48);
}

__device__ inline vector4 decodeSRgb(uint32_t c, const float* lut) {
    return vector4(lut[c & 0xffu], lut[(c >> 8) & 0xffu], lut[(c >> 16) & 0xffu], float(c >> 24) / 255.0f);
}

__device__ inline vector4 decodeUnorm16(uint64_t c) {
    return vector4(float(uint32_t(c & 0xffffu)) / 65535.0f, float(uint32_t((c >> 16) & 0xffffu)) / 65535.0f,
                   float(uint32_t((c >> 32) & 0xffffu)) / 65535.0f, float(uint32_t(c >> 48)) / 65535.0f);
}

// one axis of a linear-filtered fetch: the two texels and the weight of the second
struct Axis {
    uint32_t i0, i1;
    float a;
};
__device__ inline Axis linearAxis(float u, uint32_t n, bool wrap) {
    if (wrap)
        u = u - floorf(u);
    const float x = u * float(n) - 0.5f;
    const float fx = floorf(x);
    Axis r;
    r.a = floorf((x - fx) * 256.0f + 0.5f) * (1.0f / 256.0f);
    int i = texelIndex(fx);
    const int ni = int(n);
    if (wrap) {
        int m = i % ni;
        m = m < 0 ? m + ni : m;
        r.i0 = uint32_t(m);
        r.i1 = uint32_t(m + 1 == ni ? 0 : m + 1);
    } else {
        r.i0 = uint32_t(i < 0 ? 0 : (i >= ni ? ni - 1 : i));
        r.i1 = uint32_t(i + 1 < 0 ? 0 : (i + 1 >= ni ? ni - 1 : i + 1));
    }
    return r;
}
__device__ inline uint32_t pointAxis(float u, uint32_t n, bool wrap) {
    if (wrap)
        u = u - floorf(u);
    int i = texelIndex(floorf(u * float(n)));
    const int ni = int(n);
    return uint32_t(i < 0 ? 0 : (i >= ni ? ni - 1 : i));
}

// (1-a)(1-b) T00 + a(1-b) T10 + (1-a) b T01 + a b T11
__device__ inline vector4 blend(const vector4& t00, const vector4& t10, const vector4& t01, const vector4& t11, float a,
                                float b) {
    return t00 * ((1.0f - a) * (1.0f - b)) + t10 * (a * (1.0f - b)) + t01 * ((1.0f - a) * b) + t11 * (a * b);
}

struct Image8 {
    const uint32_t* p;
    uint32_t w, h;
    bool wrapX;
};
__device__ inline vector4 sample8(const Image8& t, float u, float v, const float* lut) {
    const Axis x = linearAxis(u, t.w, t.wrapX), y = linearAxis(v, t.h, false);
    const uint32_t* r0 = t.p + size_t(y.i0) * t.w;
    const uint32_t* r1 = t.p + size_t(y.i1) * t.w;
    return blend(decodeSRgb(r0[x.i0], lut), decodeSRgb(r0[x.i1], lut), decodeSRgb(r1[x.i0], lut),
                 decodeSRgb(r1[x.i1], lut), x.a, y.a);
}
__device__ inline vector4 sample16(const uint64_t* p, uint32_t w, uint32_t h, float u, float v) {
    const Axis x = linearAxis(u, w, true), y = linearAxis(v, h, false);
    const uint64_t* r0 = p + size_t(y.i0) * w;
    const uint64_t* r1 = p + size_t(y.i1) * w;
    return blend(decodeUnorm16(r0[x.i0]), decodeUnorm16(r0[x.i1]), decodeUnorm16(r1[x.i0]), decodeUnorm16(r1[x.i1]),
                 x.a, y.a);
}

// bicubicFast (gpu_foveated.cu:37-71): 4-tap cubic B-spline from bilinear taps
__device__ inline vector4 bicubic(const Image8& t, vector2 coord, const float* lut) {
    const vector2 pix = coord * vector2(float(t.w), float(t.h));
    const vector2 center = vector2(floorf(pix.x - 0.5f), floorf(pix.y - 0.5f)) + vector2(0.5f, 0.5f);
    const vector2 iDim(1.0f / float(t.w), 1.0f / float(t.h));
    const vector2 one(1.0f, 1.0f);
    const vector2 f = pix - center;
    const vector2 f2 = f * f;
    const vector2 f3 = f2 * f;
    const vector2 omf2 = (one - f) * (one - f);
    const vector2 omf3 = omf2 * (one - f);
    const float sixth = 1.0f / 6.0f;
    const vector2 w0 = omf3 * sixth;
    const vector2 w1 = ((one * (4.0f / 6.0f)) + (f3 * 0.5f)) - f2;
    const vector2 w3 = f3 * sixth;
    const vector2 w2 = ((one - w0) - w1) - w3;
    const vector2 s0 = w0 + w1;
    const vector2 s1 = w2 + w3;
    const vector2 f0 = w1 / (w0 + w1);
    const vector2 f1 = w3 / (w2 + w3);
    const vector2 t0 = ((center - one) + f0) * iDim;
    const vector2 t1 = ((center + one) + f1) * iDim;
    return sample8(t, t0.x, t0.y, lut) * s0.x * s0.y + sample8(t, t1.x, t0.y, lut) * s1.x * s0.y +
           sample8(t, t0.x, t1.y, lut) * s0.x * s1.y + sample8(t, t1.x, t1.y, lut) * s1.x * s1.y;
}

// clamp(x, lo, hi) = min(hi, max(lo, x)) with a < b ? a : b selects (util.h:49-56): a NaN
// bound leaves x unchanged
__device__ inline float clampSel(float x, float lo, float hi) {
    float m = lo > x ? lo : x;
    return hi < m ? hi : m;
}
__device__ inline float lerpf(float a, float b, float t) {
    return a + (b - a) * t;
}

// pixel (i, j) of a w x h image -> eye-space direction -> eccentricity
__device__ inline vector3 pixelDirection(uint32_t i, uint32_t j, uint32_t w, uint32_t h, const matrix3x3& sampleToEye) {
    const vector2 nc((float(i) + 0.5f) / float(w), (float(j) + 0.5f) / float(h));
    return normalize(sampleToEye * vector3(nc.x, nc.y, 1.0f));
}

__device__ inline void loadLut(float* lds, const float* lut) {
    const uint32_t t = threadIdx.y * blockDim.x + threadIdx.x;
    if (t < 256)
        lds[t] = lut[t];
    __syncthreads();
}

__global__ void polarRemapKernel(const uint32_t* __restrict__ results, const float* __restrict__ tMax,
                                 const vector2ui* __restrict__ remap, uint32_t count,
                                 const uint8_t* __restrict__ roundTrip, uint32_t* __restrict__ polarColor,
                                 float* __restrict__ polarDepth, uint32_t width) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    const vector2ui xy = remap[i];
    const uint32_t c = results[i];
    // FromColor4Unorm8SRgb then ToColor4Unorm8SRgb: rgb per byte through the table
    const uint32_t a = cvtU32((float(c >> 24) / 255.0f) * 255.0f);
    const uint32_t out = uint32_t(roundTrip[c & 0xffu]) | (uint32_t(roundTrip[(c >> 8) & 0xffu]) << 8) |
                         (uint32_t(roundTrip[(c >> 16) & 0xffu]) << 16) | (a << 24);
    const size_t texel = size_t(xy.y) * width + xy.x;
    polarColor[texel] = out;
    polarDepth[texel] = tMax[i];
}

__global__ __launch_bounds__(256) void momentsKernel(const uint32_t* __restrict__ polarColor, uint32_t w, uint32_t h,
                                                     const float* __restrict__ lut, uint64_t* __restrict__ m1Out,
                                                     uint64_t* __restrict__ m2Out) {
    __shared__ float slut[256];
    loadLut(slut, lut);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= w || j >= h)
        return;
    const Image8 raw = {polarColor, w, h, true};
    const vector2 invDim(1.0f / float(w), 1.0f / float(h));
    const vector2 coord(invDim.x * float(i), invDim.y * float(j));
    const float offsets[3] = {-1.0f, 0.0f, 1.0f};
    vector4 m1(0.0f), m2(0.0f);
    for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) {
            const vector4 c = sample8(raw, coord.x + offsets[x] * invDim.x, coord.y + offsets[y] * invDim.y, slut);
            m1 = m1 + c;
            m2 = m2 + vector4(c.x * c.x, c.y * c.y, c.z * c.z, c.w * c.w);
        }
    const float weight = 1.0f / 9.0f;
    m1 = m1 * weight;
    m2 = m2 * weight;
    const size_t texel = size_t(j) * w + i;
    m1Out[texel] = toColor4Unorm16(m1);
    m2Out[texel] = toColor4Unorm16(m2);
}

struct ReprojectArgs {
    const uint32_t* polarColor;
    const float* polarDepth;
    const uint64_t *moment1, *moment2;
    const uint32_t* previous; // last frame's filtered image
    uint32_t* history;        // this frame's filtered image
    uint32_t* image;          // the output, when no contrast pass follows (else null)
    uint32_t polarW, polarH, width, height;
    matrix3x3 sampleToEye;
    matrix4x4 eyeToSamplePrevious;
    EccentricityToTexCoord eToTex;
    float alpha, stddevMultiplier;
    const float* lut;
};

__global__ __launch_bounds__(256) void reprojectKernel(ReprojectArgs a) {
    __shared__ float slut[256];
    loadLut(slut, a.lut);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= a.width || j >= a.height)
        return;
    const vector3 dir = pixelDirection(i, j, a.width, a.height, a.sampleToEye);
    float theta, ecc;
    eyeSpaceDirectionToAngularEyeCoord(dir, theta, ecc);
    const vector2 coord = angularEyeCoordToPolarTextureCoord(theta, ecc, a.eToTex);
    const Image8 raw = {a.polarColor, a.polarW, a.polarH, true};
    const vector4 newValue = bicubic(raw, coord, slut);
    vector4 result = newValue;
    const float tValue =
        a.polarDepth[size_t(pointAxis(coord.y, a.polarH, false)) * a.polarW + pointAxis(coord.x, a.polarW, true)];
    if (tValue < INFINITY) {
        const vector3 eyePos = angularEyeCoordToDirection(theta, ecc) * tValue;
        const vector4 prev = a.eyeToSamplePrevious * vector4(eyePos, 1.0f);
        const float rz = 1.0f / prev.z;
        const vector2 old(prev.x * rz, prev.y * rz);
        float alpha = a.alpha;
        vector4 oldValue = newValue;
        if (old.x > 0 && old.y > 0 && old.x < 1 && old.y < 1 && tValue > 0) {
            const Image8 prevImg = {a.previous, a.width, a.height, false};
            oldValue = sample8(prevImg, old.x, old.y, slut);
        }
        // clampToNeighborhood (gpu_foveated.cu:150-162)
        const vector4 m1 = sample16(a.moment1, a.polarW, a.polarH, coord.x, coord.y);
        const vector4 m2 = sample16(a.moment2, a.polarW, a.polarH, coord.x, coord.y);
        const vector4 var = m2 - vector4(m1.x * m1.x, m1.y * m1.y, m1.z * m1.z, m1.w * m1.w);
        const vector4 sd(sqrtf(var.x), sqrtf(var.y), sqrtf(var.z), sqrtf(var.w));
        const vector4 lo = m1 - sd * a.stddevMultiplier;
        const vector4 hi = m1 + sd * a.stddevMultiplier;
        vector4 clamped(clampSel(oldValue.x, lo.x, hi.x), clampSel(oldValue.y, lo.y, hi.y),
                        clampSel(oldValue.z, lo.z, hi.z), clampSel(oldValue.w, lo.w, hi.w));
        // blend weight 0.5 in the fovea easing to `alpha` between 20% and 35% of the max
        // eccentricity; the clamp fades out toward the periphery
        float nE = ecc * a.eToTex.invMaxEccentricity;
        const float mn = 0.2f, mx = 0.35f;
        const float t = clampSel((nE - mn) / (mx - mn), 0.0f, 1.0f);
        alpha = lerpf(0.5f, alpha, t);
        nE = sqrtf(sqrtf(sqrtf(sqrtf(nE))));
        clamped = vector4(lerpf(clamped.x, oldValue.x, nE), lerpf(clamped.y, oldValue.y, nE),
                          lerpf(clamped.z, oldValue.z, nE), lerpf(clamped.w, oldValue.w, nE));
        result = newValue * alpha + clamped * (1.0f - alpha);
    }
    const uint32_t px = toColor4Unorm8SRgb(result);
    const size_t pix = size_t(j) * a.width + i;
    a.history[pix] = px;
    if (a.image)
        a.image[pix] = px;
}

__global__ __launch_bounds__(256) void eccentricityBlurKernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                              uint32_t w, uint32_t h, int stepX, int stepY,
                                                              matrix3x3 sampleToEye, const float* __restrict__ lut) {
    __shared__ float slut[256];
    loadLut(slut, lut);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= w || j >= h)
        return;
    const float ecc = acosf(-pixelDirection(i, j, w, h, sampleToEye).z);
    const int radius = 5;
    const float radiusF = 5.1f; // float(filterRadius + 0.1)
    const float sharp = sqrtf(ecc);
    vector4 sum(0.0f);
    float wsum = 0.0f;
    for (int r = -radius; r <= radius; r++) {
        int x = int(i) + r * stepX, y = int(j) + r * stepY;
        x = x < 0 ? 0 : (x > int(w) - 1 ? int(w) - 1 : x);
        y = y < 0 ? 0 : (y > int(h) - 1 ? int(h) - 1 : y);
        const float normDist = fabsf(float(r)) / radiusF;
        const float weight = powf(1.0f - normDist, sharp);
        sum = sum + decodeSRgb(in[size_t(y) * w + x], slut) * weight;
        wsum = wsum + weight;
    }
    const float inv = 1.0f / wsum;
    out[size_t(j) * w + i] = toColor4Unorm8SRgb(sum * inv);
}

__global__ __launch_bounds__(256) void contrastKernel(const uint32_t* __restrict__ unfiltered,
                                                      const uint32_t* __restrict__ filtered, uint32_t* __restrict__ image,
                                                      uint32_t w, uint32_t h, matrix3x3 sampleToEye,
                                                      float invMaxEccentricity, float fe,
                                                      const float* __restrict__ lut) {
    __shared__ float slut[256];
    loadLut(slut, lut);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= w || j >= h)
        return;
    const float ecc = acosf(-pixelDirection(i, j, w, h, sampleToEye).z);
    float sigma = 8.0f;
    const float t = ecc * invMaxEccentricity;
    const float c = clampSel(t * t, 0.0f, 1.0f);
    sigma *= (0.001f > c ? 0.001f : c);
    const size_t pix = size_t(j) * w + i;
    const vector4 p = decodeSRgb(unfiltered[pix], slut);
    const vector4 mean = decodeSRgb(filtered[pix], slut);
    // enhanceContrast (gpu_foveated.cu:296-304)
    const float scale = 1.0f + sigma * fe;
    const vector4 s = mean + (p - mean) * scale;
    const vector4 r(clampSel(s.x, 0.0f, 1.0f), clampSel(s.y, 0.0f, 1.0f), clampSel(s.z, 0.0f, 1.0f),
                    clampSel(s.w, 0.0f, 1.0f));
    image[pix] = toColor4Unorm8SRgb(r);
}

dim3 grid2(uint32_t w, uint32_t h) {
    return dim3((w + kFovBlockX - 1) / kFovBlockX, (h + kFovBlockY - 1) / kFovBlockY);
}

} // namespace

void launchFoveatedReconstruct(hvvr_gpu_view* v, const matrix3x3& sampleToEye, const matrix4x4& eyeToSamplePrevious,
                               hipStream_t stream) {
    FoveatedState& f = v->fov;
    const uint32_t cur = f.historyIndex, prev = cur ^ 1u;
    const dim3 block(kFovBlockX, kFovBlockY);
    if (v->validSampleCount)
        polarRemapKernel<<<(v->validSampleCount + 255) / 256, 256, 0, stream>>>(
            v->sampleResults.ptr, v->tMax.ptr, f.polarRemap.ptr, v->validSampleCount, f.srgbRoundTrip.ptr,
            f.polarColor.ptr, f.polarDepth.ptr, f.samplesPerRing);
    momentsKernel<<<grid2(f.samplesPerRing, f.ringCount), block, 0, stream>>>(
        f.polarColor.ptr, f.samplesPerRing, f.ringCount, f.srgbToLinear.ptr, f.moment1.ptr, f.moment2.ptr);

    ReprojectArgs a;
    a.polarColor = f.polarColor.ptr;
    a.polarDepth = f.polarDepth.ptr;
    a.moment1 = f.moment1.ptr;
    a.moment2 = f.moment2.ptr;
    a.previous = f.history[prev].ptr;
    a.history = f.history[cur].ptr;
    a.image = f.contrast ? nullptr : v->image.ptr;
    a.polarW = f.samplesPerRing;
    a.polarH = f.ringCount;
    a.width = v->width;
    a.height = v->height;
    a.sampleToEye = sampleToEye;
    a.eyeToSamplePrevious = eyeToSamplePrevious;
    a.eToTex.emap.m = f.emap[0];
    a.eToTex.emap.w0 = f.emap[1];
    a.eToTex.emap.S = f.emap[2];
    a.eToTex.emap.invLogA = f.emap[3];
    a.eToTex.emap.invW0 = f.emap[4];
    a.eToTex.emap.invM = f.emap[5];
    a.eToTex.texMapSize = float(f.ringCount);
    a.eToTex.invTexMapSize = 1.0f / float(f.ringCount);
    a.eToTex.invMaxEccentricity = 1.0f / f.maxEccentricity;
    a.alpha = f.taaAlpha;
    a.stddevMultiplier = f.taaStddevMultiplier;
    a.lut = f.srgbToLinear.ptr;
    const dim3 g = grid2(v->width, v->height);
    reprojectKernel<<<g, block, 0, stream>>>(a);
    if (f.contrast) {
        // the first pass steps along y, the second along x (gpu_foveated.cu:371-377)
        eccentricityBlurKernel<<<g, block, 0, stream>>>(f.history[cur].ptr, f.filtered[0].ptr, v->width, v->height, 0,
                                                        1, sampleToEye, f.srgbToLinear.ptr);
        eccentricityBlurKernel<<<g, block, 0, stream>>>(f.filtered[0].ptr, f.filtered[1].ptr, v->width, v->height, 1,
                                                        0, sampleToEye, f.srgbToLinear.ptr);
        contrastKernel<<<g, block, 0, stream>>>(f.history[cur].ptr, f.filtered[1].ptr, v->image.ptr, v->width,
                                                v->height, sampleToEye, a.eToTex.invMaxEccentricity, f.contrastFe,
                                                f.srgbToLinear.ptr);
    }
    f.historyIndex = prev;
}

} // namespace hvvr
