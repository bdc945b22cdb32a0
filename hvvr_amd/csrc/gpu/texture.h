#pragma once
// Scene textures in HBM and their sampling in the resolve (the reference's texture atlas,
// raycaster/texture.cu:78-272, read through tex2DGrad in GGXShade, shading.h:45-47, 99-127).
//
// Storage: each texture is a mip chain, level L of max(1, w >> L) x max(1, h >> L) texels,
// levels concatenated. 8-bit formats keep their RGBA8 bytes (kind unorm8 or srgb8); every
// other format is stored as RGBA32F (converted once at upload).
//
// The reference samples through the CUDA texture unit: wrap addressing, linear filtering,
// linear mip filtering and up to 8x anisotropy, sRGB decoded per texel before filtering.
// That is restated here in arithmetic a CPU can repeat: the texture unit's bilinear weights
// with its documented 8-bit fraction (x = u * size - 0.5, weight rounded to 1/256), the mip
// fraction likewise, and anisotropy by the published EXT_texture_filter_anisotropic method
// (N = min(ceil(Pmax / Pmin), 8) trilinear probes along the major gradient axis at level
// log2(Pmax / N)); the hardware's own anisotropic footprint is not published. Every step is
// exact or correctly rounded (no log2: see texLod), so a CPU restatement compiled without
// contraction gives the same bits.
//
// Mip levels are built as d_mipmap does (texture.cu:78-107): texel (x, y) of level L + 1 is
// the mean of four bilinear, clamp-addressed reads of level L at normalized coordinates
// (x, y), (x + 1, y), (x + 1, y + 1), (x, y + 1) over the new size, in the raw (undecoded)
// unorm values, scaled by 255, clamped and truncated to bytes. The reference writes bytes
// for every format; float textures here average in float instead.

#include "hvvr/types.h"

namespace hvvr {

enum : uint32_t { TexUnorm8 = 0, TexSrgb8 = 1, TexFloat4 = 2 };
enum : uint32_t { kMaxTexLevels = 16, kMaxAnisotropy = 8 };

struct TexInfo {
    uint32_t width, height, levels, kind;
    uint64_t offset[kMaxTexLevels]; // byte offset of each level in the texel buffer
};

struct TextureSet {
    const uint8_t* texels;
    const TexInfo* info;
    const float* srgbLut; // [256] sRGB byte -> linear (util_graphics.h:16-18)
    const float* lodLut;  // [256] the mip-fraction thresholds of texLod
    uint32_t count;
};

// lodLut[j] = 2^((j + 0.5) / 256) rounded to float, j = 0..255 (computed once on the host)

HVVR_HD uint32_t texLevelSize(uint32_t n, uint32_t level) {
    const uint32_t s = n >> level;
    return s ? s : 1u;
}

// float -> texel index: NaN gives 0, the range is limited far beyond any texture
HVVR_HD int texIndex(float f) { return (f >= -1.0e9f && f <= 1.0e9f) ? int(f) : 0; }

// the texture unit's 8-bit fraction: rounded to a multiple of 1/256
HVVR_HD float texFraction(float f) { return floorf(f * 256.0f + 0.5f) * (1.0f / 256.0f); }

struct TexAxis {
    uint32_t i0, i1;
    float a; // weight of i1
};

HVVR_HD TexAxis texAxis(float u, uint32_t n, bool wrap) {
    if (wrap)
        u = u - floorf(u);
    const float x = u * float(n) - 0.5f;
    const float fx = floorf(x);
    TexAxis r;
    r.a = texFraction(x - fx);
    const int i = texIndex(fx), ni = int(n);
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

// one texel of level `level`; `decode`: sRGB bytes through the table (sampling does, mip
// generation does not)
HVVR_HD vector4 texFetch(const uint8_t* texels, const TexInfo& t, uint32_t level, uint32_t x, uint32_t y,
                         const float* srgbLut, bool decode) {
    const size_t i = size_t(y) * texLevelSize(t.width, level) + x;
    if (t.kind == TexFloat4) {
        const float* p = reinterpret_cast<const float*>(texels + t.offset[level]) + 4 * i;
        return vector4(p[0], p[1], p[2], p[3]);
    }
    const uint32_t c = reinterpret_cast<const uint32_t*>(texels + t.offset[level])[i];
    const uint32_t r = c & 0xffu, g = (c >> 8) & 0xffu, b = (c >> 16) & 0xffu, a = c >> 24;
    if (decode && t.kind == TexSrgb8)
        return vector4(srgbLut[r], srgbLut[g], srgbLut[b], float(a) / 255.0f);
    return vector4(float(r) / 255.0f, float(g) / 255.0f, float(b) / 255.0f, float(a) / 255.0f);
}

// (1-a)(1-b) T00 + a(1-b) T10 + (1-a) b T01 + a b T11
HVVR_HD vector4 texBlend(const vector4& t00, const vector4& t10, const vector4& t01, const vector4& t11, float a,
                         float b) {
    return t00 * ((1.0f - a) * (1.0f - b)) + t10 * (a * (1.0f - b)) + t01 * ((1.0f - a) * b) + t11 * (a * b);
}

HVVR_HD vector4 texBilinear(const uint8_t* texels, const TexInfo& t, uint32_t level, float u, float v, bool wrap,
                            const float* srgbLut, bool decode) {
    const TexAxis ax = texAxis(u, texLevelSize(t.width, level), wrap);
    const TexAxis ay = texAxis(v, texLevelSize(t.height, level), wrap);
    return texBlend(texFetch(texels, t, level, ax.i0, ay.i0, srgbLut, decode),
                    texFetch(texels, t, level, ax.i1, ay.i0, srgbLut, decode),
                    texFetch(texels, t, level, ax.i0, ay.i1, srgbLut, decode),
                    texFetch(texels, t, level, ax.i1, ay.i1, srgbLut, decode), ax.a, ay.a);
}

// Level of detail lambda = log2(x) clamped to [0, levels - 1], as the level floor(lambda)
// and the mip fraction rounded to 1/256 (texFraction of lambda - floor(lambda)). For x >= 1
// the level is x's binary exponent and round(256 log2(m)) of its mantissa m in [1, 2) is the
// count of thresholds lodLut[j] <= m: no log2, whose last bit the CPU's and the GPU's
// libraries could round differently. NaN and x < 1 give level 0; the top level takes f = 0.
HVVR_HD void texLod(float x, uint32_t levels, const float* lodLut, uint32_t& level, float& f) {
    level = 0;
    f = 0.0f;
    if (!(x >= 1.0f))
        return;
    const uint32_t bits = __builtin_bit_cast(uint32_t, x);
    const uint32_t e = (bits >> 23) - 127u; // x >= 1: normal or infinite
    if (e >= levels - 1u) {
        level = levels - 1u;
        return;
    }
    const float m = __builtin_bit_cast(float, (bits & 0x7fffffu) | 0x3f800000u);
    uint32_t k = 0; // thresholds <= m among lodLut[0..254], then lodLut[255]
    for (uint32_t step = 128; step; step >>= 1)
        if (lodLut[k + step - 1u] <= m)
            k += step;
    k += lodLut[k] <= m ? 1u : 0u;
    level = e;
    f = float(k) * (1.0f / 256.0f);
}

// tex2DGrad(tex, uv, ddx, ddy) as restated above
HVVR_HD vector4 texSampleGrad(const uint8_t* texels, const TexInfo& t, const float* srgbLut, const float* lodLut,
                              vector2 uv, vector2 ddx, vector2 ddy) {
    const float W = float(t.width), H = float(t.height);
    const float px = ddx.x * W, py = ddx.y * H, qx = ddy.x * W, qy = ddy.y * H;
    const float lx = sqrtf(px * px + py * py), ly = sqrtf(qx * qx + qy * qy);
    const float pMax = lx >= ly ? lx : ly, pMin = lx >= ly ? ly : lx;
    const vector2 major = lx >= ly ? ddx : ddy;
    uint32_t n = 1;
    if (pMax > 0.0f) {
        const float ratio = pMin > 0.0f ? ceilf(pMax / pMin) : float(kMaxAnisotropy);
        n = ratio < float(kMaxAnisotropy) ? uint32_t(ratio) : kMaxAnisotropy;
        n = n < 1u ? 1u : n;
    }
    uint32_t level0;
    float f;
    texLod(pMax / float(n), t.levels, lodLut, level0, f);
    const uint32_t level1 = level0 + 1u < t.levels ? level0 + 1u : level0;
    vector4 sum(0.0f);
    for (uint32_t i = 0; i < n; i++) {
        const float s = (float(i) + 0.5f) / float(n) - 0.5f;
        const vector2 p = uv + major * s;
        const vector4 c0 = texBilinear(texels, t, level0, p.x, p.y, true, srgbLut, true);
        const vector4 c1 = texBilinear(texels, t, level1, p.x, p.y, true, srgbLut, true);
        sum = sum + (c0 * (1.0f - f) + c1 * f);
    }
    return sum * (1.0f / float(n));
}

// texel (x, y) of level `level + 1` from level `level` (d_mipmap, texture.cu:78-107)
HVVR_HD void texMipTexel(const uint8_t* texels, const TexInfo& t, uint32_t level, uint32_t x, uint32_t y,
                         uint8_t* out) {
    const uint32_t w = texLevelSize(t.width, level + 1u), h = texLevelSize(t.height, level + 1u);
    const float px = float(1.0 / double(w)), py = float(1.0 / double(h));
    const float u0 = float(x) * px, u1 = float(x + 1u) * px, v0 = float(y) * py, v1 = float(y + 1u) * py;
    vector4 c = texBilinear(texels, t, level, u0, v0, false, nullptr, false) +
                texBilinear(texels, t, level, u1, v0, false, nullptr, false) +
                texBilinear(texels, t, level, u1, v1, false, nullptr, false) +
                texBilinear(texels, t, level, u0, v1, false, nullptr, false);
    const size_t i = size_t(y) * w + x;
    if (t.kind == TexFloat4) {
        c = c / 4.0f;
        float* p = reinterpret_cast<float*>(out + t.offset[level + 1u]) + 4 * i;
        p[0] = c.x;
        p[1] = c.y;
        p[2] = c.z;
        p[3] = c.w;
        return;
    }
    c = c / 4.0f;
    c = c * 255.0f;
    const float ch[4] = {c.x < 255.0f ? c.x : 255.0f, c.y < 255.0f ? c.y : 255.0f, c.z < 255.0f ? c.z : 255.0f,
                         c.w < 255.0f ? c.w : 255.0f};
    uint32_t v = 0;
    for (int k = 0; k < 4; k++)
        v |= (ch[k] > 0.0f ? uint32_t(ch[k]) : 0u) << (8 * k);
    reinterpret_cast<uint32_t*>(out + t.offset[level + 1u])[i] = v;
}

HVVR_HD uint32_t texLevelCount(uint32_t width, uint32_t height) {
    uint32_t s = width > height ? width : height, levels = 0;
    while (s) {
        s /= 2;
        levels++;
    }
    return levels;
}

} // namespace hvvr
