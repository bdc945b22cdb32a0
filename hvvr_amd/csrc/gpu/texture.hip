// Scene textures (include/hvvr_gpu.h, hvvr_gpu_scene_set_textures): the reference's
// CreateTexture (raycaster/texture.cu:210-272) for a whole scene at once. Level 0 is
// converted on the host into the stored kinds of texture.h (RGBA8 bytes for the 8-bit
// formats, RGBA32F for the rest), uploaded into one buffer with every texture's chain, and
// the mip levels are built on the device, one kernel per level (d_mipmap, texture.cu:78-107,
// restated in texture.h).

#include "gpu_internal.h"

#include <math.h>
#include <string.h>

namespace hvvr {

namespace {

enum : uint32_t { // TextureFormat (texture.h:16-26)
    FmtNone = 0,
    FmtSrgb8 = 1,
    FmtUnorm8 = 2,
    FmtUnorm16 = 3,
    FmtFloat32x4 = 4,
    FmtFloat16x4 = 5,
    FmtR11G11B10 = 6,
    FmtR32 = 7
};

uint32_t bytesPerTexel(uint32_t format) {
    switch (format) {
        case FmtSrgb8:
        case FmtUnorm8:
        case FmtR11G11B10:
        case FmtR32:
            return 4;
        case FmtUnorm16:
        case FmtFloat16x4:
            return 8;
        case FmtFloat32x4:
            return 16;
        default:
            return 0;
    }
}

float halfToFloatHost(uint16_t h) {
    const uint32_t sign = uint32_t(h >> 15) << 31, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    float f;
    if (e == 0)
        f = ldexpf(float(m), -24);
    else if (e == 31)
        f = m ? NAN : INFINITY;
    else
        f = ldexpf(float(m | 0x400u), int(e) - 25);
    uint32_t bits;
    memcpy(&bits, &f, 4);
    bits |= sign;
    memcpy(&f, &bits, 4);
    return f;
}

// an unsigned float of `mbits` mantissa and 5 exponent bits (the packed 11/11/10 format)
float smallFloat(uint32_t v, uint32_t mbits) {
    const uint32_t e = v >> mbits, m = v & ((1u << mbits) - 1u);
    if (e == 0)
        return ldexpf(float(m), -14 - int(mbits));
    if (e == 31)
        return m ? NAN : INFINITY;
    return ldexpf(float(m | (1u << mbits)), int(e) - 15 - int(mbits));
}

// texel (x, y) of the caller's level 0 as RGBA32F
void toFloat4(uint32_t format, const uint8_t* p, float out[4]) {
    if (format == FmtUnorm16) {
        uint16_t c[4];
        memcpy(c, p, 8);
        for (int k = 0; k < 4; k++)
            out[k] = float(c[k]) / 65535.0f;
    } else if (format == FmtFloat32x4) {
        memcpy(out, p, 16);
    } else if (format == FmtFloat16x4) {
        uint16_t c[4];
        memcpy(c, p, 8);
        for (int k = 0; k < 4; k++)
            out[k] = halfToFloatHost(c[k]);
    } else if (format == FmtR11G11B10) {
        uint32_t c;
        memcpy(&c, p, 4);
        out[0] = smallFloat(c & 0x7ffu, 6);
        out[1] = smallFloat((c >> 11) & 0x7ffu, 6);
        out[2] = smallFloat(c >> 22, 5);
        out[3] = 1.0f;
    } else { // FmtR32: a one-channel texture reads (r, 0, 0, 1)
        memcpy(out, p, 4);
        out[1] = out[2] = 0.0f;
        out[3] = 1.0f;
    }
}

__global__ void mipLevelKernel(uint8_t* texels, const TexInfo* info, uint32_t texture, uint32_t level) {
    const TexInfo t = info[texture];
    const uint32_t w = texLevelSize(t.width, level + 1u), h = texLevelSize(t.height, level + 1u);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h)
        return;
    texMipTexel(texels, t, level, i % w, i / w, texels);
}

int texFail(const std::string& msg, int code = HVVR_GPU_ERR_INVALID) {
    setLastError(msg);
    return code;
}

} // namespace

// [0, 256): sRGB byte -> linear (sRGBToLinear, shared/util_graphics.h:16-18, on k / 255);
// [256, 512): texLod's thresholds 2^((j + 0.5) / 256)
void textureTables(float out[512]) {
    for (int k = 0; k < 256; k++) {
        const float x = float(k) / 255.0f;
        out[k] = (x <= 0.04045f) ? x * (1.0f / 12.92f) : powf((x + .055f) * (1.0f / 1.055f), 2.4f);
    }
    for (int j = 0; j < 256; j++)
        out[256 + j] = float(exp2((j + 0.5) / 256.0));
}

// bytes of a texture's chain and its level offsets (from 0)
uint64_t textureChainLayout(TexInfo& t) {
    const uint64_t elem = t.kind == TexFloat4 ? 16 : 4;
    uint64_t at = 0;
    for (uint32_t l = 0; l < t.levels; l++) {
        t.offset[l] = at;
        at += uint64_t(texLevelSize(t.width, l)) * texLevelSize(t.height, l) * elem;
        at = (at + 15) & ~uint64_t(15);
    }
    return at;
}

} // namespace hvvr

using namespace hvvr;

extern "C" {

int hvvr_gpu_scene_set_textures(hvvr_gpu_scene* s, uint32_t count, const hvvr_gpu_texture_desc* textures) {
    if (!s || (count && !textures))
        return texFail("null argument");
    if (count >= SimpleMaterial::badTextureIndex)
        return texFail("too many textures: index 65535 is reserved for 'no texture'");
    std::vector<TexInfo> info(count);
    std::vector<uint64_t> base(count), bytes(count);
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; i++) {
        const hvvr_gpu_texture_desc& d = textures[i];
        const uint32_t bpt = bytesPerTexel(d.format);
        if (!bpt || d.width == 0 || d.height == 0 || d.strideElements < d.width || !d.data)
            return texFail("texture " + std::to_string(i) + ": unknown format, empty size, stride below width or no data");
        TexInfo& t = info[i];
        t.width = d.width;
        t.height = d.height;
        t.levels = texLevelCount(d.width, d.height);
        if (t.levels > kMaxTexLevels)
            return texFail("texture " + std::to_string(i) + " is larger than 32768 texels on a side");
        t.kind = d.format == FmtSrgb8 ? TexSrgb8 : (d.format == FmtUnorm8 ? TexUnorm8 : TexFloat4);
        base[i] = total;
        bytes[i] = textureChainLayout(t);
        total += (bytes[i] + 255) & ~uint64_t(255);
    }
    // level 0 of every texture, in the stored kind, at its place in the buffer
    std::vector<uint8_t> host(total);
    for (uint32_t i = 0; i < count; i++) {
        const hvvr_gpu_texture_desc& d = textures[i];
        const TexInfo& t = info[i];
        const uint8_t* src = static_cast<const uint8_t*>(d.data);
        const uint32_t bpt = bytesPerTexel(d.format);
        uint8_t* dst = host.data() + base[i];
        for (uint32_t y = 0; y < t.height; y++) {
            const uint8_t* row = src + size_t(y) * d.strideElements * bpt;
            if (t.kind != TexFloat4) {
                memcpy(dst + size_t(y) * t.width * 4, row, size_t(t.width) * 4);
                continue;
            }
            for (uint32_t x = 0; x < t.width; x++)
                toFloat4(d.format, row + size_t(x) * bpt, reinterpret_cast<float*>(dst) + 4 * (size_t(y) * t.width + x));
        }
    }
    // offsets within the scene's buffer
    for (uint32_t i = 0; i < count; i++)
        for (uint32_t l = 0; l < info[i].levels; l++)
            info[i].offset[l] += base[i];
    float lut[512];
    textureTables(lut);
    HVVR_HIP_CHECK(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    HVVR_HIP_CHECK(hipStreamSynchronize(st)); // frames in flight may read the old textures
    HVVR_HIP_CHECK(s->texels.resize(total ? total : 256));
    HVVR_HIP_CHECK(s->texInfo.resize(count ? count : 1));
    HVVR_HIP_CHECK(s->texLut.resize(512));
    if (total)
        HVVR_HIP_CHECK(hipMemcpyAsync(s->texels.ptr, host.data(), total, hipMemcpyHostToDevice, st));
    if (count)
        HVVR_HIP_CHECK(hipMemcpyAsync(s->texInfo.ptr, info.data(), count * sizeof(TexInfo), hipMemcpyHostToDevice, st));
    HVVR_HIP_CHECK(hipMemcpyAsync(s->texLut.ptr, lut, sizeof(lut), hipMemcpyHostToDevice, st));
    for (uint32_t i = 0; i < count; i++)
        for (uint32_t l = 0; l + 1 < info[i].levels; l++) {
            const uint32_t n = texLevelSize(info[i].width, l + 1) * texLevelSize(info[i].height, l + 1);
            mipLevelKernel<<<(n + 255) / 256, 256, 0, st>>>(s->texels.ptr, s->texInfo.ptr, i, l);
        }
    HVVR_HIP_CHECK(hipGetLastError());
    HVVR_HIP_CHECK(hipStreamSynchronize(st));
    s->textureCount = count;
    s->texInfoHost = info;
    s->texBase = base;
    s->texBytes = bytes;
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_texture_info(hvvr_gpu_scene* s, uint32_t index, uint32_t info[4], uint64_t* bytes) {
    if (!s || !info)
        return texFail("null argument");
    if (index >= s->textureCount)
        return texFail("texture index out of range");
    const TexInfo& t = s->texInfoHost[index];
    info[0] = t.width;
    info[1] = t.height;
    info[2] = t.levels;
    info[3] = t.kind;
    if (bytes)
        *bytes = s->texBytes[index];
    return HVVR_GPU_OK;
}

int hvvr_gpu_scene_read_texture(hvvr_gpu_scene* s, uint32_t index, void* dst) {
    if (!s || !dst)
        return texFail("null argument");
    if (index >= s->textureCount)
        return texFail("texture index out of range");
    HVVR_HIP_CHECK(hipSetDevice(s->ctx->device));
    HVVR_HIP_CHECK(hipMemcpyAsync(dst, s->texels.ptr + s->texBase[index], s->texBytes[index], hipMemcpyDeviceToHost,
                                  s->ctx->stream));
    HVVR_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    return HVVR_GPU_OK;
}

} // extern "C"
