/* Oracle for the polar foveated path (TEST INFRASTRUCTURE ONLY; see hvvr_oracle.h).
 *
 *   t-max of subsample 0       raycaster/resolve.cu:386-457
 *   sRGB tables                shared/util_graphics.h:16-28
 *   polar remap                raycaster/remap.cu:127-141
 *   moments                    raycaster/gpu_foveated.cu:136-186
 *   reprojection + temporal    raycaster/gpu_foveated.cu:37-71 (bicubicFast), 88-130, 188-251
 *   eccentricity blur          raycaster/gpu_foveated.cu:253-294
 *   contrast enhancement       raycaster/gpu_foveated.cu:296-333
 *   eccentricity map           raycaster/foveated.h:86-111
 *
 * The reference reads CUDA textures; their documented behaviour is restated here: with
 * normalized coordinates a linear fetch takes texels floor(u*N - 0.5) and the next with the
 * weight frac(u*N - 0.5) held in 8 fractional bits (so a fetch at a texel centre returns the
 * texel), u wraps (mod 1) on the polar azimuth and clamps at every other edge, sRGB texels
 * are decoded to linear before the blend, 16-bit unorm texels read as v/65535. */
#include "or_dof.h"
#include "or_math.h"

#include <stdlib.h>

static const float kPi = 3.141592653589f;

/* float -> u32 conversion with the GPU's saturation (NaN, negatives -> 0) */
static uint32_t cvt_u32(float x) {
    return x > 0.0f ? (x < 4294967040.0f ? (uint32_t)x : 0xffffffffu) : 0u;
}
static int texel_index(float fx) {
    return (fx >= -1.0e9f && fx <= 1.0e9f) ? (int)fx : 0;
}

typedef struct { float x, y, z, w; } ov4;
static ov4 v4(float x, float y, float z, float w) { ov4 r = {x, y, z, w}; return r; }
static ov4 add4(ov4 a, ov4 b) { return v4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static ov4 sub4(ov4 a, ov4 b) { return v4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static ov4 scl4(ov4 a, float s) { return v4(a.x * s, a.y * s, a.z * s, a.w * s); }

static float srgb_to_linear(float x) {
    return (x <= 0.04045f) ? x * (1.0f / 12.92f) : powf((x + .055f) * (1.0f / 1.055f), 2.4f);
}
static float linear_to_srgb(float x) {
    return (x <= 0.0031308f) ? (x * 12.92f) : (powf(x, 1.0f / 2.4f) * 1.055f - 0.055f);
}

void or_srgb_tables(float toLinear[256], uint8_t roundTrip[256]) {
    for (int k = 0; k < 256; k++) {
        float lin = srgb_to_linear((float)k / 255.0f);
        float b = linear_to_srgb(lin) * 255.0f;
        toLinear[k] = lin;
        roundTrip[k] = (uint8_t)(b > 0.0f ? (b < 255.0f ? (uint32_t)b : 255u) : 0u);
    }
}

static float g_lut[256];
static uint8_t g_rt[256];
static int g_lut_ready = 0;
static void ensure_lut(void) {
    if (!g_lut_ready) {
        or_srgb_tables(g_lut, g_rt);
        g_lut_ready = 1;
    }
}

/* ToColor4Unorm8SRgb (shading_helpers.h:193-200) */
static uint32_t to_unorm8_srgb(ov4 c) {
    uint32_t r = cvt_u32(linear_to_srgb(c.x) * 255.0f);
    uint32_t g = cvt_u32(linear_to_srgb(c.y) * 255.0f);
    uint32_t b = cvt_u32(linear_to_srgb(c.z) * 255.0f);
    uint32_t a = cvt_u32(c.w * 255.0f);
    return r | (g << 8) | (b << 16) | (a << 24);
}
/* ToColor4Unorm16 (shading_helpers.h:201-208) */
static uint64_t to_unorm16(ov4 c) {
    const float C = 65535.0f;
    uint64_t r = cvt_u32(c.x * C), g = cvt_u32(c.y * C), b = cvt_u32(c.z * C), a = cvt_u32(c.w * C);
    return r | (g << 16) | (b << 32) | (a << 48);
}
static ov4 decode_srgb(uint32_t c) {
    return v4(g_lut[c & 0xffu], g_lut[(c >> 8) & 0xffu], g_lut[(c >> 16) & 0xffu], (float)(c >> 24) / 255.0f);
}
static ov4 decode16(uint64_t c) {
    return v4((float)(uint32_t)(c & 0xffffu) / 65535.0f, (float)(uint32_t)((c >> 16) & 0xffffu) / 65535.0f,
              (float)(uint32_t)((c >> 32) & 0xffffu) / 65535.0f, (float)(uint32_t)(c >> 48) / 65535.0f);
}

typedef struct { uint32_t i0, i1; float a; } axis_t;
static axis_t linear_axis(float u, uint32_t n, int wrap) {
    if (wrap) u = u - floorf(u);
    float x = u * (float)n - 0.5f;
    float fx = floorf(x);
    axis_t r;
    r.a = floorf((x - fx) * 256.0f + 0.5f) * (1.0f / 256.0f);
    int i = texel_index(fx), ni = (int)n;
    if (wrap) {
        int m = i % ni;
        m = m < 0 ? m + ni : m;
        r.i0 = (uint32_t)m;
        r.i1 = (uint32_t)(m + 1 == ni ? 0 : m + 1);
    } else {
        r.i0 = (uint32_t)(i < 0 ? 0 : (i >= ni ? ni - 1 : i));
        r.i1 = (uint32_t)(i + 1 < 0 ? 0 : (i + 1 >= ni ? ni - 1 : i + 1));
    }
    return r;
}
static uint32_t point_axis(float u, uint32_t n, int wrap) {
    if (wrap) u = u - floorf(u);
    int i = texel_index(floorf(u * (float)n)), ni = (int)n;
    return (uint32_t)(i < 0 ? 0 : (i >= ni ? ni - 1 : i));
}
static ov4 blend(ov4 t00, ov4 t10, ov4 t01, ov4 t11, float a, float b) {
    ov4 r = scl4(t00, (1.0f - a) * (1.0f - b));
    r = add4(r, scl4(t10, a * (1.0f - b)));
    r = add4(r, scl4(t01, (1.0f - a) * b));
    return add4(r, scl4(t11, a * b));
}
static ov4 sample8(const uint32_t* p, uint32_t w, uint32_t h, int wrapX, float u, float v) {
    axis_t x = linear_axis(u, w, wrapX), y = linear_axis(v, h, 0);
    const uint32_t* r0 = p + (uint64_t)y.i0 * w;
    const uint32_t* r1 = p + (uint64_t)y.i1 * w;
    return blend(decode_srgb(r0[x.i0]), decode_srgb(r0[x.i1]), decode_srgb(r1[x.i0]), decode_srgb(r1[x.i1]), x.a, y.a);
}
static ov4 sample16(const uint64_t* p, uint32_t w, uint32_t h, float u, float v) {
    axis_t x = linear_axis(u, w, 1), y = linear_axis(v, h, 0);
    const uint64_t* r0 = p + (uint64_t)y.i0 * w;
    const uint64_t* r1 = p + (uint64_t)y.i1 * w;
    return blend(decode16(r0[x.i0]), decode16(r0[x.i1]), decode16(r1[x.i0]), decode16(r1[x.i1]), x.a, y.a);
}

/* bicubicFast: 4-tap B-spline from bilinear taps, per axis */
static ov4 bicubic(const uint32_t* p, uint32_t w, uint32_t h, float cu, float cv) {
    float pix[2] = {cu * (float)w, cv * (float)h};
    float dim[2] = {(float)w, (float)h};
    float t0[2], t1[2], s0[2], s1[2];
    for (int k = 0; k < 2; k++) {
        float center = floorf(pix[k] - 0.5f) + 0.5f;
        float iDim = 1.0f / dim[k];
        float f = pix[k] - center;
        float f2 = f * f, f3 = f2 * f;
        float omf2 = (1.0f - f) * (1.0f - f);
        float omf3 = omf2 * (1.0f - f);
        float sixth = 1.0f / 6.0f;
        float w0 = omf3 * sixth;
        float w1 = ((1.0f * (4.0f / 6.0f)) + (f3 * 0.5f)) - f2;
        float w3 = f3 * sixth;
        float w2 = ((1.0f - w0) - w1) - w3;
        s0[k] = w0 + w1;
        s1[k] = w2 + w3;
        float f0 = w1 / (w0 + w1), f1 = w3 / (w2 + w3);
        t0[k] = ((center - 1.0f) + f0) * iDim;
        t1[k] = ((center + 1.0f) + f1) * iDim;
    }
    ov4 r = scl4(scl4(sample8(p, w, h, 1, t0[0], t0[1]), s0[0]), s0[1]);
    r = add4(r, scl4(scl4(sample8(p, w, h, 1, t1[0], t0[1]), s1[0]), s0[1]));
    r = add4(r, scl4(scl4(sample8(p, w, h, 1, t0[0], t1[1]), s0[0]), s1[1]));
    return add4(r, scl4(scl4(sample8(p, w, h, 1, t1[0], t1[1]), s1[0]), s1[1]));
}

/* clamp = min(hi, max(lo, x)) with a < b ? a : b selects (util.h:49-56) */
static float clamp_sel(float x, float lo, float hi) {
    float m = lo > x ? lo : x;
    return hi < m ? hi : m;
}
static float lerpf(float a, float b, float t) { return a + (b - a) * t; }

float or_emap_apply(const float e[6], float i) {
    float e0 = i * e[1];
    float x = (e0 - e[2]) * e[4];
    float gx = (e[1] * (powf(1.0f + e[0], x) - 1.0f)) * e[5];
    float e1 = gx + e[2];
    return (e0 < e[2]) ? e0 : e1;
}
float or_emap_apply_inverse(const float e[6], float E) {
    float i0 = E * e[4];
    float sDivW = e[2] * e[4];
    float x = logf((e[0] * (i0 - sDivW)) + 1.0f) * e[3];
    float i1 = x + sDivW;
    return (E < e[2]) ? i0 : i1;
}

/* column-major 3x3 times v */
static ov3 m33mul(const float m[9], ov3 v) {
    return v3(m[0] * v.x + m[3] * v.y + m[6] * v.z, m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
static ov3 pixel_dir(uint32_t i, uint32_t j, uint32_t w, uint32_t h, const float s2e[9]) {
    float nx = ((float)i + 0.5f) / (float)w, ny = ((float)j + 0.5f) / (float)h;
    return norm3(m33mul(s2e, v3(nx, ny, 1.0f)));
}

void or_polar_remap(const uint32_t* results, const float* tMax, const uint32_t* remap, uint32_t count,
                    uint32_t width, uint32_t* polarColor, float* polarDepth) {
    ensure_lut();
    for (uint32_t i = 0; i < count; i++) {
        uint32_t c = results[i];
        uint32_t a = cvt_u32(((float)(c >> 24) / 255.0f) * 255.0f);
        uint32_t out = (uint32_t)g_rt[c & 0xffu] | ((uint32_t)g_rt[(c >> 8) & 0xffu] << 8) |
                       ((uint32_t)g_rt[(c >> 16) & 0xffu] << 16) | (a << 24);
        uint64_t texel = (uint64_t)remap[2 * i + 1] * width + remap[2 * i];
        polarColor[texel] = out;
        polarDepth[texel] = tMax[i];
    }
}

void or_polar_moments(const uint32_t* pc, uint32_t w, uint32_t h, uint64_t* m1Out, uint64_t* m2Out) {
    ensure_lut();
    const float offs[3] = {-1.0f, 0.0f, 1.0f};
    float ix = 1.0f / (float)w, iy = 1.0f / (float)h;
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            float cx = ix * (float)i, cy = iy * (float)j;
            ov4 m1 = v4(0, 0, 0, 0), m2 = v4(0, 0, 0, 0);
            for (int x = 0; x < 3; x++)
                for (int y = 0; y < 3; y++) {
                    ov4 c = sample8(pc, w, h, 1, cx + offs[x] * ix, cy + offs[y] * iy);
                    m1 = add4(m1, c);
                    m2 = add4(m2, v4(c.x * c.x, c.y * c.y, c.z * c.z, c.w * c.w));
                }
            const float weight = 1.0f / 9.0f;
            m1 = scl4(m1, weight);
            m2 = scl4(m2, weight);
            m1Out[(uint64_t)j * w + i] = to_unorm16(m1);
            m2Out[(uint64_t)j * w + i] = to_unorm16(m2);
        }
}

void or_foveated_reproject(const o_reproject* a, uint32_t* out) {
    ensure_lut();
    const float* e = a->emap;
    const float invTex = 1.0f / (float)a->polarH;
    const float invMaxE = 1.0f / a->maxEccentricity;
    for (uint32_t j = 0; j < a->height; j++)
        for (uint32_t i = 0; i < a->width; i++) {
            ov3 dir = pixel_dir(i, j, a->width, a->height, a->sampleToEye);
            float ecc = acosf(-dir.z);
            float theta = -atan2f(dir.y, dir.x);
            float cu = (theta + kPi) / (2.0f * kPi);
            float cv = (or_emap_apply_inverse(e, ecc) + 0.5f) * invTex;
            ov4 newValue = bicubic(a->polarColor, a->polarW, a->polarH, cu, cv);
            ov4 result = newValue;
            float t = a->polarDepth[(uint64_t)point_axis(cv, a->polarH, 0) * a->polarW + point_axis(cu, a->polarW, 1)];
            if (t < INFINITY) {
                /* angularEyeCoordToDirection */
                float z = -cosf(ecc);
                float xyLen = sqrtf(1.0f - z * z);
                ov3 eyeDir = v3(cosf(-theta) * xyLen, sinf(-theta) * xyLen, z);
                ov3 pos = scale3(eyeDir, t);
                float prev[4];
                m4mul(a->eyeToSamplePrevious, pos.x, pos.y, pos.z, 1.0f, prev);
                float rz = 1.0f / prev[2];
                float ox = prev[0] * rz, oy = prev[1] * rz;
                float alpha = a->alpha;
                ov4 oldValue = newValue;
                if (ox > 0 && oy > 0 && ox < 1 && oy < 1 && t > 0)
                    oldValue = sample8(a->previous, a->width, a->height, 0, ox, oy);
                ov4 m1 = sample16(a->moment1, a->polarW, a->polarH, cu, cv);
                ov4 m2 = sample16(a->moment2, a->polarW, a->polarH, cu, cv);
                ov4 var = sub4(m2, v4(m1.x * m1.x, m1.y * m1.y, m1.z * m1.z, m1.w * m1.w));
                ov4 sd = v4(sqrtf(var.x), sqrtf(var.y), sqrtf(var.z), sqrtf(var.w));
                ov4 lo = sub4(m1, scl4(sd, a->stddevMultiplier));
                ov4 hi = add4(m1, scl4(sd, a->stddevMultiplier));
                ov4 cl = v4(clamp_sel(oldValue.x, lo.x, hi.x), clamp_sel(oldValue.y, lo.y, hi.y),
                            clamp_sel(oldValue.z, lo.z, hi.z), clamp_sel(oldValue.w, lo.w, hi.w));
                float nE = ecc * invMaxE;
                const float mn = 0.2f, mx = 0.35f;
                float tt = clamp_sel((nE - mn) / (mx - mn), 0.0f, 1.0f);
                alpha = lerpf(0.5f, alpha, tt);
                nE = sqrtf(sqrtf(sqrtf(sqrtf(nE))));
                cl = v4(lerpf(cl.x, oldValue.x, nE), lerpf(cl.y, oldValue.y, nE), lerpf(cl.z, oldValue.z, nE),
                        lerpf(cl.w, oldValue.w, nE));
                result = add4(scl4(newValue, alpha), scl4(cl, 1.0f - alpha));
            }
            out[(uint64_t)j * a->width + i] = to_unorm8_srgb(result);
        }
}

void or_foveated_blur(const uint32_t* in, uint32_t* out, uint32_t w, uint32_t h, int stepX, int stepY,
                      const float s2e[9]) {
    ensure_lut();
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            float ecc = acosf(-pixel_dir(i, j, w, h, s2e).z);
            float sharp = sqrtf(ecc);
            ov4 sum = v4(0, 0, 0, 0);
            float wsum = 0.0f;
            for (int r = -5; r <= 5; r++) {
                int x = (int)i + r * stepX, y = (int)j + r * stepY;
                x = x < 0 ? 0 : (x > (int)w - 1 ? (int)w - 1 : x);
                y = y < 0 ? 0 : (y > (int)h - 1 ? (int)h - 1 : y);
                float normDist = fabsf((float)r) / 5.1f;
                float weight = powf(1.0f - normDist, sharp);
                sum = add4(sum, scl4(decode_srgb(in[(uint64_t)y * w + x]), weight));
                wsum = wsum + weight;
            }
            out[(uint64_t)j * w + i] = to_unorm8_srgb(scl4(sum, 1.0f / wsum));
        }
}

void or_foveated_contrast(const uint32_t* unfiltered, const uint32_t* filtered, uint32_t* out, uint32_t w, uint32_t h,
                          const float s2e[9], float maxEccentricity, float fe) {
    ensure_lut();
    const float invMaxE = 1.0f / maxEccentricity;
    for (uint32_t j = 0; j < h; j++)
        for (uint32_t i = 0; i < w; i++) {
            float ecc = acosf(-pixel_dir(i, j, w, h, s2e).z);
            float sigma = 8.0f;
            float t = ecc * invMaxE;
            float c = clamp_sel(t * t, 0.0f, 1.0f);
            sigma *= (0.001f > c ? 0.001f : c);
            uint64_t pix = (uint64_t)j * w + i;
            ov4 p = decode_srgb(unfiltered[pix]), mean = decode_srgb(filtered[pix]);
            float scale = 1.0f + sigma * fe;
            ov4 s = add4(mean, scl4(sub4(p, mean), scale));
            out[pix] = to_unorm8_srgb(v4(clamp_sel(s.x, 0.0f, 1.0f), clamp_sel(s.y, 0.0f, 1.0f),
                                         clamp_sel(s.z, 0.0f, 1.0f), clamp_sel(s.w, 0.0f, 1.0f)));
        }
}

/* barycentrics at subsample offset (cx, cy): IntersectTriangleThread::calcUVW */
static void pinhole_uvw(const o_tri* T, ov3 cam, ov3 dir, ov3 du, ov3 dv, float cx, float cy, float* b) {
    ov3 ndir = neg3(dir), ndu = neg3(du), ndv = neg3(dv);
    ov3 normal = cross3(T->e0, T->e1);
    ov3 v0o = sub3(cam, T->v0);
    float denomC = dot3(ndir, normal);
    ov3 eC = cross3(ndir, v0o);
    float vC = dot3(T->e1, eC), wC = -dot3(T->e0, eC);
    float dD0 = dot3(ndu, normal), dD1 = dot3(ndv, normal);
    ov3 cu = cross3(ndu, v0o), cv = cross3(ndv, v0o);
    float dV0 = dot3(T->e1, cu), dV1 = dot3(T->e1, cv);
    float dW0 = -dot3(T->e0, cu), dW1 = -dot3(T->e0, cv);
    float denom = fmaf(dD1, cy, fmaf(dD0, cx, denomC));
    float v = fmaf(dV1, cy, fmaf(dV0, cx, vC));
    float w = fmaf(dW1, cy, fmaf(dW0, cx, wC));
    float u = (denom - v) - w;
    float dinv = 1.0f / denom;
    b[0] = u * dinv;
    b[1] = v * dinv;
    b[2] = w * dinv;
}

/* IntersectTriangleThreadDoF::calcUVW (prim_tests.h:366-392) after IntersectTriangleTile<true>::setup
 * (:263-291) at lens position lensUV and direction offset dirUV */
static void dof_uvw(const o_tri* T, ov3 cam, const dof_tile* tileLens, ov3 lcfc, ov2 lensUV, ov2 dirUV, float* b) {
    ov3 lu = tileLens->lensU, lv = tileLens->lensV;
    ov3 normal = cross3(T->e0, T->e1);
    ov3 v0ToLens = sub3(cam, T->v0);
    float dDA0 = dot3(neg3(lu), normal), dDA1 = dot3(neg3(lv), normal);
    float denomC = dot3(neg3(lcfc), normal);
    float denom = fmaf(dDA1, dirUV.y, fmaf(dDA0, dirUV.x, denomC));
    ov3 o = axpy2(v0ToLens, lensUV.x, lu, lensUV.y, lv);
    ov3 rd = axpy2(lcfc, dirUV.x, lu, dirUV.y, lv);
    ov3 e = cross3_fma(neg3(rd), o);
    float v = dot3_fma(T->e1, e);
    float w = -dot3_fma(T->e0, e);
    float u = (denom - v) - w;
    float dinv = 1.0f / denom;
    b[0] = u * dinv;
    b[1] = v * dinv;
    b[2] = w * dinv;
}

int or_hits_view(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                 const uint32_t* triIds, const o_tri* tris, const o_dof* dof, const uint32_t* counts, float* uvwOut) {
    if (A < 1 || A > 32 || !counts) return -1;
    ov3 cam = v3(m[12], m[13], m[14]);
    uint32_t rank = 0, rankTile = 0, row = 0;
    dof_tile tileLens;
    memset(&tileLens, 0, sizeof(tileLens));
    for (uint32_t s = 0; s < sampleCount; s++) {
        uint32_t tile = s / 128;
        if (dof && s % 128 == 0 && counts[tile]) {
            const o_beam* mid = &beams[(uint64_t)tile * 128 + 64];
            dof_tile_load(m, m3mul(m, mid->du), m3mul(m, mid->dv), dof->radius, dof->focalDistance, &tileLens);
            row = dof_row(counts, tile, &rank, &rankTile);
        }
        ov3 dir = m3mul(m, beams[s].c);
        for (uint32_t i = 0; i < A; i++) {
            float* b = uvwOut + ((uint64_t)s * A + i) * 3;
            uint32_t tri = triIds[(uint64_t)s * A + i];
            b[0] = b[1] = b[2] = 0.0f;
            if (counts[tile] == 0 || tri == 0xffffffffu) continue;
            if (dof) {
                ov2 lensUV, dirUV;
                dof_sample_uvs(dof->lensTable, row, A, i, s % 128, offsets[i], tileLens.focalToLensScale, &lensUV,
                               &dirUV);
                dof_uvw(&tris[tri], cam, &tileLens, dof_lens_to_focal(m, dir, dof->focalDistance), lensUV, dirUV, b);
            } else {
                pinhole_uvw(&tris[tri], cam, dir, m3mul(m, beams[s].du), m3mul(m, beams[s].dv), offsets[i].x,
                            offsets[i].y, b);
            }
        }
    }
    return 0;
}

int or_tmax_view(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                 const uint32_t* triIds, const o_tri* tris, const o_dof* dof, const uint32_t* counts, float* tMaxOut) {
    if (A < 1 || A > 32 || !counts) return -1;
    ov3 cam = v3(m[12], m[13], m[14]);
    ov3 look = v3(-m[8], -m[9], -m[10]);
    uint32_t rank = 0, rankTile = 0;
    dof_tile tileLens;
    memset(&tileLens, 0, sizeof(tileLens));
    uint32_t row = 0;
    for (uint32_t s = 0; s < sampleCount; s++) {
        uint32_t tile = s / 128;
        if (dof && s % 128 == 0 && counts[tile]) {
            const o_beam* mid = &beams[(uint64_t)tile * 128 + 64];
            dof_tile_load(m, m3mul(m, mid->du), m3mul(m, mid->dv), dof->radius, dof->focalDistance, &tileLens);
            row = dof_row(counts, tile, &rank, &rankTile);
        }
        uint32_t tri = triIds[(uint64_t)s * A];
        if (counts[tile] == 0 || tri == 0xffffffffu) {
            tMaxOut[s] = INFINITY;
            continue;
        }
        const o_tri* T = &tris[tri];
        ov3 dir = m3mul(m, beams[s].c);
        float b[3];
        if (dof) {
            ov2 lensUV, dirUV;
            dof_sample_uvs(dof->lensTable, row, A, 0, s % 128, offsets[0], tileLens.focalToLensScale, &lensUV, &dirUV);
            dof_uvw(T, cam, &tileLens, dof_lens_to_focal(m, dir, dof->focalDistance), lensUV, dirUV, b);
        } else {
            pinhole_uvw(T, cam, dir, m3mul(m, beams[s].du), m3mul(m, beams[s].dv), offsets[0].x, offsets[0].y, b);
        }
        ov3 v1 = add3(T->v0, T->e0), v2 = add3(T->v0, T->e1);
        ov3 pos = add3(add3(scale3(T->v0, b[0]), scale3(v1, b[1])), scale3(v2, b[2]));
        tMaxOut[s] = dot3(sub3(pos, cam), look);
    }
    return 0;
}
