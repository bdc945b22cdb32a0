/* Scene textures for the resolve: storage, mip generation and tex2DGrad sampling, restated
 * on the CPU. TEST INFRASTRUCTURE ONLY.
 *
 *   texture atlas, CreateTexture      raycaster/texture.cu:210-272
 *   d_mipmap (mip level L+1 from L)   raycaster/texture.cu:78-107
 *   ReadTexture -> tex2DGrad          raycaster/shading.h:45-47, used by GGXShade :99-127
 *   sRGBToLinear (sRGB8 decode)       shared/util_graphics.h:16-18
 *
 * The reference samples through the CUDA texture unit (wrap addressing, linear filtering,
 * linear mip filtering, 8x anisotropy, sRGB decode before filtering), whose results are not
 * published bit for bit. The product restates it in arithmetic a CPU can repeat; this file
 * is that restatement written again from its definition:
 *   bilinear   x = u * size - 0.5, i0 = floor(x), weight of i0 + 1 = frac(x) rounded to a
 *              multiple of 1/256 (the unit's 8-bit fraction); wrap: u - floor(u) first and
 *              indices modulo size; clamp (mip generation): indices limited to [0, size)
 *   level      lambda = log2(P / N) clamped to [0, levels - 1] with P the longer of the two
 *              UV gradients in texels and N = min(ceil(Pmax / Pmin), 8) probes
 *              (EXT_texture_filter_anisotropic); its fraction rounded to 1/256; computed as
 *              binary exponent + count of thresholds 2^((j + 0.5) / 256) at or below the
 *              mantissa, which is floor(256 log2(m) + 0.5) (counted linearly here)
 *   probes     N trilinear samples at uv + major * ((i + 0.5) / N - 0.5), averaged
 *   mips       mean of four clamp-addressed bilinear reads of level L at (x, y), (x+1, y),
 *              (x+1, y+1), (x, y+1) over the new size, 8-bit kinds in undecoded unorm,
 *              * 255, min 255, truncated; float kinds keep the float mean
 * Pins: tests/test_texture.py (constant textures, bilinear closed forms on 2x2 textures,
 * level and fraction against log2 in double, anisotropic probes of a linear ramp, wrap
 * periodicity, sRGB decode values, 2x2 -> 1x1 and constant-chain mips).
 */
#include "or_math.h"

enum { K_UNORM8 = 0, K_SRGB8 = 1, K_FLOAT4 = 2 };

static const uint8_t* g_texels;
static const o_texinfo* g_info;
static uint32_t g_count;
static float g_tables[512];
static int g_tables_ready;

void or_texture_tables(float out[512]) {
    for (int k = 0; k < 256; k++) {
        float x = (float)k / 255.0f;
        out[k] = (x <= 0.04045f) ? x * (1.0f / 12.92f) : powf((x + .055f) * (1.0f / 1.055f), 2.4f);
    }
    for (int j = 0; j < 256; j++) out[256 + j] = (float)exp2((j + 0.5) / 256.0);
}

static const float* tables(void) {
    if (!g_tables_ready) {
        or_texture_tables(g_tables);
        g_tables_ready = 1;
    }
    return g_tables;
}

static uint32_t level_size(uint32_t n, uint32_t l) {
    uint32_t s = n >> l;
    return s ? s : 1u;
}

uint32_t or_texture_levels(uint32_t width, uint32_t height) {
    uint32_t s = width > height ? width : height, n = 0;
    while (s) { s /= 2; n++; }
    return n;
}

uint64_t or_texture_layout(o_texinfo* t) {
    uint64_t elem = t->kind == K_FLOAT4 ? 16 : 4, at = 0;
    for (uint32_t l = 0; l < t->levels; l++) {
        t->offset[l] = at;
        at += (uint64_t)level_size(t->width, l) * level_size(t->height, l) * elem;
        at = (at + 15) & ~(uint64_t)15;
    }
    return at;
}

/* one texel as four floats; sRGB bytes decoded only when `decode` */
static void fetch(const uint8_t* base, const o_texinfo* t, uint32_t l, uint32_t x, uint32_t y, int decode,
                  float out[4]) {
    size_t i = (size_t)y * level_size(t->width, l) + x;
    if (t->kind == K_FLOAT4) {
        memcpy(out, base + t->offset[l] + 16 * i, 16);
        return;
    }
    const uint8_t* p = base + t->offset[l] + 4 * i;
    const float* lut = tables();
    for (int c = 0; c < 4; c++)
        out[c] = (decode && t->kind == K_SRGB8 && c < 3) ? lut[p[c]] : (float)p[c] / 255.0f;
}

static float frac8(float f) { return floorf(f * 256.0f + 0.5f) * (1.0f / 256.0f); }

/* the two texel indices of one axis and the weight of the second */
static void axis(float u, uint32_t n, int wrap, uint32_t* i0, uint32_t* i1, float* a) {
    if (wrap) u = u - floorf(u);
    float x = u * (float)n - 0.5f, fx = floorf(x);
    *a = frac8(x - fx);
    int i = (fx >= -1.0e9f && fx <= 1.0e9f) ? (int)fx : 0, ni = (int)n;
    if (wrap) {
        int m = ((i % ni) + ni) % ni;
        *i0 = (uint32_t)m;
        *i1 = (uint32_t)((m + 1) % ni);
    } else {
        *i0 = (uint32_t)(i < 0 ? 0 : (i >= ni ? ni - 1 : i));
        *i1 = (uint32_t)(i + 1 < 0 ? 0 : (i + 1 >= ni ? ni - 1 : i + 1));
    }
}

static void bilinear(const uint8_t* base, const o_texinfo* t, uint32_t l, float u, float v, int wrap, int decode,
                     float out[4]) {
    uint32_t x0, x1, y0, y1;
    float a, b;
    axis(u, level_size(t->width, l), wrap, &x0, &x1, &a);
    axis(v, level_size(t->height, l), wrap, &y0, &y1, &b);
    float t00[4], t10[4], t01[4], t11[4];
    fetch(base, t, l, x0, y0, decode, t00);
    fetch(base, t, l, x1, y0, decode, t10);
    fetch(base, t, l, x0, y1, decode, t01);
    fetch(base, t, l, x1, y1, decode, t11);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    for (int c = 0; c < 4; c++) out[c] = t00[c] * w00 + t10[c] * w10 + t01[c] * w01 + t11[c] * w11;
}

void or_texture_lod(float x, uint32_t levels, uint32_t* level, float* f) {
    *level = 0;
    *f = 0.0f;
    if (!(x >= 1.0f)) return;
    int e;
    float m = 2.0f * frexpf(x, &e); /* x = m * 2^(e-1), m in [1, 2); inf: e undefined */
    if (isinf(x) || (uint32_t)(e - 1) >= levels - 1u) {
        *level = levels - 1u;
        return;
    }
    const float* thr = tables() + 256;
    uint32_t k = 0;
    for (int j = 0; j < 256; j++) k += thr[j] <= m;
    *level = (uint32_t)(e - 1);
    *f = (float)k * (1.0f / 256.0f);
}

void or_texture_sample(const uint8_t* base, const o_texinfo* t, float u, float v, float dxu, float dxv, float dyu,
                       float dyv, float out[4]) {
    float W = (float)t->width, Hh = (float)t->height;
    float px = dxu * W, py = dxv * Hh, qx = dyu * W, qy = dyv * Hh;
    float lx = sqrtf(px * px + py * py), ly = sqrtf(qx * qx + qy * qy);
    float pMax = lx >= ly ? lx : ly, pMin = lx >= ly ? ly : lx;
    float mu = lx >= ly ? dxu : dyu, mv = lx >= ly ? dxv : dyv;
    uint32_t n = 1;
    if (pMax > 0.0f) {
        float ratio = pMin > 0.0f ? ceilf(pMax / pMin) : 8.0f;
        n = ratio < 8.0f ? (uint32_t)ratio : 8u;
        if (n < 1u) n = 1u;
    }
    uint32_t l0, l1;
    float f;
    or_texture_lod(pMax / (float)n, t->levels, &l0, &f);
    l1 = l0 + 1u < t->levels ? l0 + 1u : l0;
    float sum[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < n; i++) {
        float s = ((float)i + 0.5f) / (float)n - 0.5f;
        float pu = u + mu * s, pv = v + mv * s;
        float c0[4], c1[4];
        bilinear(base, t, l0, pu, pv, 1, 1, c0);
        bilinear(base, t, l1, pu, pv, 1, 1, c1);
        for (int c = 0; c < 4; c++) sum[c] = sum[c] + (c0[c] * (1.0f - f) + c1[c] * f);
    }
    float inv = 1.0f / (float)n;
    for (int c = 0; c < 4; c++) out[c] = sum[c] * inv;
}

void or_texture_build_mips(uint8_t* base, const o_texinfo* t) {
    for (uint32_t l = 0; l + 1 < t->levels; l++) {
        uint32_t w = level_size(t->width, l + 1), h = level_size(t->height, l + 1);
        float px = (float)(1.0 / (double)w), py = (float)(1.0 / (double)h);
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                float u0 = (float)x * px, u1 = (float)(x + 1) * px, v0 = (float)y * py, v1 = (float)(y + 1) * py;
                float a[4], b[4], c[4], d[4], s[4];
                bilinear(base, t, l, u0, v0, 0, 0, a);
                bilinear(base, t, l, u1, v0, 0, 0, b);
                bilinear(base, t, l, u1, v1, 0, 0, c);
                bilinear(base, t, l, u0, v1, 0, 0, d);
                for (int k = 0; k < 4; k++) s[k] = (((a[k] + b[k]) + c[k]) + d[k]) / 4.0f;
                size_t i = (size_t)y * w + x;
                if (t->kind == K_FLOAT4) {
                    memcpy(base + t->offset[l + 1] + 16 * i, s, 16);
                    continue;
                }
                uint8_t* o = base + t->offset[l + 1] + 4 * i;
                for (int k = 0; k < 4; k++) {
                    float q = s[k] * 255.0f;
                    q = q < 255.0f ? q : 255.0f;
                    o[k] = q > 0.0f ? (uint8_t)q : 0;
                }
            }
    }
}

void or_set_textures(const uint8_t* texels, const o_texinfo* info, uint32_t count) {
    g_texels = count ? texels : 0;
    g_info = count ? info : 0;
    g_count = count;
}

int or_texture_read(uint32_t id, float u, float v, float dxu, float dxv, float dyu, float dyv, float out[4]) {
    if (id == 0xffffu || id >= g_count) return 0;
    or_texture_sample(g_texels, &g_info[id], u, v, dxu, dxv, dyu, dyv, out);
    return 1;
}

int or_textures_present(uint32_t id) { return id != 0xffffu && id < g_count; }
