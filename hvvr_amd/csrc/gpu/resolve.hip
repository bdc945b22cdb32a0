// Resolve (MSAA GGX shading, ACES tone map, sRGB8) and remap for one view.
//
//   resolveKernel  ResolveKernel + ShadeMSAA (raycaster/resolve.cu:161-284, 324-384) and
//                  ClearEmptyKernel (resolve.cu:527-542) in one launch: one 128-thread
//                  workgroup per tile, one thread per sample; tiles whose triangle list
//                  was empty write the clear colour 0xff000000 without touching the
//                  GBuffer. Per unique triangle of a sample: centroid of its subsample
//                  offsets, barycentrics (calcUVW, prim_tests.h:233-245), vertex
//                  interpolation (shading.h:19-36), GGXShade (shading.h:84-161) weighted
//                  by the covered subsample count.
//   remapKernel    RemapKernel (raycaster/remap.cu:35-77): sample order -> image order,
//                  split-colour merge for 2 (pentile) or 3 channels.
//
// Textures (texture.h): a material naming textures (SimpleMaterial's diffuse / specular /
// glossy fields) reads them with the UV derivatives the reference takes for tex2DGrad
// (resolve.cu:258-275): barycentrics at the centroid and at the centroid moved by one unit
// along each subsample axis, the interpolated UVs' differences. Untextured materials skip
// that work and shade with their constant colours.

#include "device_common.h"
#include "dof.h"
#include "gpu_internal.h"
#include "launch_args.h"
#include "uvw.h"

namespace hvvr {

// The shading's divisions, normalisations and powers use the hardware's approximate forms
// (v_rcp_f32, v_rsq_f32, v_log_f32 / v_exp_f32, about 1 ulp): the reference builds its
// kernels with fast math (raycaster.vcxproj: FastMath), where they are approximate too.
// Colours stay within 1/255 of the oracle's correctly rounded FP32 (and of FP64); the
// intersection and the depth and barycentric readbacks keep exact divisions.
__device__ inline float rcpFast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline vector3 normalizeFast(const vector3& v) { return v * __builtin_amdgcn_rsqf(dot(v, v)); }
// linearToSRgb (util_graphics.h) with the power as exp2(log2(x) / 2.4)
__device__ inline float linearToSRgbFast(float x) {
    return (x <= 0.0031308f) ? (x * 12.92f)
                             : (__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 2.4f)) * 1.055f - 0.055f);
}

__device__ inline vector3 ggxSpecular(vector3 N, vector3 L, vector3 V, float roughness, vector3 F0) {
    // GGX_SchlickFresnel_SchlickSmithGGX (shading_helpers.h:89-112)
    float alpha = roughness * roughness;
    vector3 H = normalizeFast(L + V);
    float alpha2 = alpha * alpha;
    float NdotH = saturate(dot(N, H));
    float q = NdotH * NdotH * (alpha2 - 1.0f) + 1.0f;
    float D = alpha2 * rcpFast(3.141592653589f * (q * q));
    float LdotH = saturate(dot(L, H));
    float x = 1.0f - LdotH;
    float x2 = x * x;
    float p5 = x2 * x2 * x;
    vector3 F = F0 + (vector3(1.0f) - F0) * p5;
    float NdotL = saturate(dot(N, L));
    float NdotV = saturate(dot(N, V));
    float k = alpha * 0.5f;
    float vis = rcpFast(NdotL * (1.0f - k) + k) * rcpFast(NdotV * (1.0f - k) + k);
    return ((F * D) * vis) * 0.25f;
}

__device__ inline vector3 pbrBrdf(vector3 L, vector3 N, vector3 V, vector3 base, float roughness, vector3 F0,
                                  vector3 lightColor) {
    // PBR_BRDF1 (shading_helpers.h:115-122)
    vector3 spec = ggxSpecular(N, L, V, roughness, F0);
    float NdotL = saturate(dot(N, L));
    vector3 diffuse = ((base * lightColor) * (vector3(1.0f) - spec)) * NdotL;
    return diffuse + lightColor * spec;
}

// A material's inputs to GGXShade (shading.h:84-161): the constants, or where the
// material names a texture, its value at the sample (shading.h:99-127): the emissive
// model's colour from the diffuse slot, else base colour, F0 and roughness = 1 - glossy.x.
struct SurfaceInputs {
    uint32_t model_;
    vector3 emissiveColor_; // emissive model
    vector3 emissive_;      // added to the radiance of the GGX model (never textured)
    vector3 base_, F0_;
    float roughness_;
    __device__ uint32_t model() const { return model_; }
    __device__ vector3 emissiveColor() const { return emissiveColor_; }
    __device__ vector3 emissive() const { return emissive_; }
    __device__ vector3 base() const { return base_; }
    __device__ vector3 F0() const { return F0_; }
    __device__ float roughness() const { return roughness_; }
};
// an untextured material's inputs, read where the shading uses them
struct MaterialInputs {
    const SimpleMaterial& m;
    __device__ uint32_t model() const { return uint32_t(m.textureIDsAndShadingModel & 0xffffu); }
    __device__ vector3 emissiveColor() const { return m.emissive.xyz(); }
    __device__ vector3 emissive() const { return m.emissive.xyz(); }
    __device__ vector3 base() const { return m.diffuse.xyz(); }
    __device__ vector3 F0() const { return m.specular.xyz(); }
    __device__ float roughness() const { return 1.0f - m.glossiness; }
};

__device__ inline bool textureIndexValid(uint32_t id, const TextureSet& tex) {
    return id != SimpleMaterial::badTextureIndex && id < tex.count;
}

// whether shading this material reads a texture
__device__ inline bool materialTextured(const SimpleMaterial& mat, const TextureSet& tex) {
    const uint64_t code = mat.textureIDsAndShadingModel;
    const uint32_t mask = SimpleMaterial::textureMask;
    const uint32_t diffuse = uint32_t(code >> SimpleMaterial::diffuseBitShift) & mask;
    if (uint32_t(code & 0xffffu) == ShadingModelEmissive)
        return textureIndexValid(diffuse, tex);
    return textureIndexValid(diffuse, tex) ||
           textureIndexValid(uint32_t(code >> SimpleMaterial::specularBitShift) & mask, tex) ||
           textureIndexValid(uint32_t(code >> SimpleMaterial::glossyBitShift) & mask, tex);
}

__device__ inline SurfaceInputs surfaceInputs(const SimpleMaterial& mat, const TextureSet& tex, vector2 uv,
                                              vector2 ddx, vector2 ddy) {
    SurfaceInputs in;
    const uint64_t code = mat.textureIDsAndShadingModel;
    in.model_ = uint32_t(code & 0xffffu);
    in.emissiveColor_ = mat.emissive.xyz();
    in.emissive_ = mat.emissive.xyz();
    in.base_ = mat.diffuse.xyz();
    in.F0_ = mat.specular.xyz();
    in.roughness_ = 1.0f - mat.glossiness;
    const uint32_t mask = SimpleMaterial::textureMask;
    const uint32_t diffuse = uint32_t(code >> SimpleMaterial::diffuseBitShift) & mask;
    const uint32_t specular = uint32_t(code >> SimpleMaterial::specularBitShift) & mask;
    const uint32_t glossy = uint32_t(code >> SimpleMaterial::glossyBitShift) & mask;
    auto read = [&](uint32_t id) { return texSampleGrad(tex.texels, tex.info[id], tex.srgbLut, tex.lodLut, uv, ddx, ddy); };
    if (in.model_ == ShadingModelEmissive) {
        if (textureIndexValid(diffuse, tex))
            in.emissiveColor_ = read(diffuse).xyz();
        return in;
    }
    if (textureIndexValid(diffuse, tex))
        in.base_ = read(diffuse).xyz();
    if (textureIndexValid(specular, tex))
        in.F0_ = read(specular).xyz();
    if (textureIndexValid(glossy, tex))
        in.roughness_ = 1.0f - read(glossy).x;
    return in;
}

template <class In>
__device__ inline vector3 ggxShade(const In& in, vector3 pos, vector3 nrm, vector3 camPos,
                                   const LightingEnvironment& env) {
    // GGXShade (shading.h:84-161)
    vector3 normal = normalizeFast(nrm);
    if (in.model() == ShadingModelEmissive)
        return in.emissiveColor();
    vector3 base = in.base();
    vector3 V = normalizeFast(camPos - pos);
    vector3 F0 = in.F0();
    float roughness = in.roughness();
    vector3 rad = in.emissive();
    rad += base * vector3(0.15f, 0.15f, 0.165f);
    for (int i = 0; i < MAX_DIRECTIONAL_LIGHTS; i++) {
        if (i >= env.directionalLightCount)
            break;
        rad += pbrBrdf(env.directionalLights[i].Direction, normal, V, base, roughness, F0,
                       env.directionalLights[i].Power);
    }
    for (int i = 0; i < MAX_POINT_LIGHTS; i++) {
        if (i >= env.pointLightCount)
            break;
        const LightPoint& l = env.pointLights[i];
        vector3 delta = l.Position - pos;
        const float d2 = dot(delta, delta), inv = __builtin_amdgcn_rsqf(d2);
        float r = d2 * inv;
        vector3 L = delta * inv;
        float fdist = saturate((l.FalloffEnd - r) * l.FalloffScale);
        rad += pbrBrdf(L, normal, V, base, roughness, F0, l.Color * fdist);
    }
    for (int i = 0; i < MAX_SPOT_LIGHTS; i++) {
        if (i >= env.spotLightCount)
            break;
        const LightSpot& l = env.spotLights[i];
        vector3 delta = l.Position - pos;
        const float d2 = dot(delta, delta), inv = __builtin_amdgcn_rsqf(d2);
        float r = d2 * inv;
        vector3 L = delta * inv;
        float fdist = saturate((l.FalloffEnd - r) * l.FalloffScale);
        vector3 lit = pbrBrdf(L, normal, V, base, roughness, F0, l.Color * fdist);
        float cosA = dot(l.Direction, -L);
        float fang = saturate((cosA - l.CosOuterAngle) * l.CosAngleScale);
        rad += lit * fang;
    }
    return vector3(fmaxv(rad.x, 0.0f), fmaxv(rad.y, 0.0f), fmaxv(rad.z, 0.0f));
}

// ---- reduced-precision shading: the colour terms of each light in packed FP16 ----
// PBR_BRDF1 = diffuse + lightColor * spec with diffuse = base * lightColor * (1 - spec) *
// NdotL and spec = F * s, s = D * vis / 4, which is lightColor * (base NdotL) +
// lightColor * F * (1 - base NdotL) * s. The scalar terms (D, vis, Fresnel weight) keep
// FP32: D * vis spans many decades (up to ~1e7 for a glossy highlight seen edge-on). The
// two colour vectors, bounded by the light colour, are formed with packed FP16 operations
// (v_pk_mul_f16 / v_pk_add_f16); the radiance sum stays FP32. The scalar terms use the
// hardware reciprocal and reciprocal square root, as the FP32 path does.
typedef _Float16 half4v __attribute__((ext_vector_type(4)));

typedef float float4v __attribute__((ext_vector_type(4)));

__device__ inline half4v toHalf4(const vector3& v) {
    return __builtin_convertvector((float4v{v.x, v.y, v.z, 0.0f}), half4v);
}

__device__ inline void pbrBrdfHalf(const vector3& L, const vector3& N, const vector3& V, half4v base,
                                   float roughness, half4v F0, half4v lightColor, float4v& rad) {
    float alpha = roughness * roughness;
    vector3 H = normalizeFast(L + V);
    float alpha2 = alpha * alpha;
    float NdotH = saturate(dot(N, H));
    float q = NdotH * NdotH * (alpha2 - 1.0f) + 1.0f;
    float D = alpha2 * rcpFast(3.141592653589f * (q * q));
    float LdotH = saturate(dot(L, H));
    float x = 1.0f - LdotH;
    float x2 = x * x;
    float p5 = x2 * x2 * x;
    float NdotL = saturate(dot(N, L));
    float NdotV = saturate(dot(N, V));
    float k = alpha * 0.5f;
    float vis = rcpFast(NdotL * (1.0f - k) + k) * rcpFast(NdotV * (1.0f - k) + k);
    float sc = (D * vis) * 0.25f;
    const half4v one = (_Float16)1.0f;
    const half4v F = F0 + (one - F0) * (_Float16)p5;
    const half4v a = base * (_Float16)NdotL;
    const half4v u = lightColor * a;
    const half4v w = (lightColor * F) * (one - a);
    rad += __builtin_convertvector(u, float4v) + __builtin_convertvector(w, float4v) * sc;
}

template <class In>
__device__ inline vector3 ggxShadeHalf(const In& in, vector3 pos, vector3 nrm, vector3 camPos,
                                       const LightingEnvironment& env) {
    vector3 normal = normalizeFast(nrm);
    if (in.model() == ShadingModelEmissive)
        return in.emissiveColor();
    const half4v base = toHalf4(in.base());
    vector3 V = normalizeFast(camPos - pos);
    const half4v F0 = toHalf4(in.F0());
    float roughness = in.roughness();
    vector3 amb = in.emissive() + in.base() * vector3(0.15f, 0.15f, 0.165f);
    float4v rad = {amb.x, amb.y, amb.z, 0.0f};
    for (int i = 0; i < MAX_DIRECTIONAL_LIGHTS; i++) {
        if (i >= env.directionalLightCount)
            break;
        pbrBrdfHalf(env.directionalLights[i].Direction, normal, V, base, roughness, F0,
                    toHalf4(env.directionalLights[i].Power), rad);
    }
    for (int i = 0; i < MAX_POINT_LIGHTS; i++) {
        if (i >= env.pointLightCount)
            break;
        const LightPoint& l = env.pointLights[i];
        vector3 delta = l.Position - pos;
        const float d2 = dot(delta, delta), inv = __builtin_amdgcn_rsqf(d2);
        float r = d2 * inv;
        vector3 L = delta * inv;
        float fdist = saturate((l.FalloffEnd - r) * l.FalloffScale);
        pbrBrdfHalf(L, normal, V, base, roughness, F0, toHalf4(l.Color * fdist), rad);
    }
    for (int i = 0; i < MAX_SPOT_LIGHTS; i++) {
        if (i >= env.spotLightCount)
            break;
        const LightSpot& l = env.spotLights[i];
        vector3 delta = l.Position - pos;
        const float d2 = dot(delta, delta), inv = __builtin_amdgcn_rsqf(d2);
        float r = d2 * inv;
        vector3 L = delta * inv;
        float fdist = saturate((l.FalloffEnd - r) * l.FalloffScale);
        float cosA = dot(l.Direction, -L);
        float fang = saturate((cosA - l.CosOuterAngle) * l.CosAngleScale);
        pbrBrdfHalf(L, normal, V, base, roughness, F0, toHalf4(l.Color * (fdist * fang)), rad);
    }
    return vector3(fmaxv(rad.x, 0.0f), fmaxv(rad.y, 0.0f), fmaxv(rad.z, 0.0f));
}

__device__ inline float acesFilm(float x) {
    // ACESFilm (shading_helpers.h:158-170)
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return saturate((x * (a * x + b)) * rcpFast(x * (c * x + d) + e));
}

// Tex: the scene has textures (a variant of its own: the texture reads' registers would
// cost untextured scenes occupancy)
template <int A, bool DoF, bool Half, bool Tex>
__device__ inline void resolveBody(const ResolveArgs& args, uint32_t tile) {
    __shared__ DofTile dofTile;
    const uint32_t tid = threadIdx.x;
    const uint32_t sample = tile * TILE_SIZE + tid;
    if (args.tileCounts[tile] == 0) {
        // ClearEmptyKernel (resolve.cu:527-542)
        args.sampleResults[sample] = 0xff000000u;
        if (args.tMax)
            args.tMax[sample] = INFINITY;
        return;
    }
    const uint32_t lane = tid & 63u;
    const GBufferSubsample* waveBase = args.gbuffer + size_t(sample >> 6) * 64 * A;
    const matrix4x4& m = args.cameraToWorld;
    const matrix3x3 rot = m.upper3x3();
    const vector3 cam = m.m3.xyz();
    DirectionalBeam beam = args.beams[sample];
    const vector3 ndir = -(rot * beam.centerRay);
    const vector3 ndu = -(rot * beam.du);
    const vector3 ndv = -(rot * beam.dv);
    const uint32_t maskAll = A < 32 ? ((1u << A) - 1u) : 0xffffffffu;
    // depth of field: tile lens frame from sample 64 (ResolveKernel, resolve.cu:358-366),
    // per-sample focal offset, the lens-table row of this tile
    vector3 lcfc(0.0f);
    uint32_t lensBase = 0;
    if constexpr (DoF) {
        if (tid == 64)
            dofTile = dofTileLoad(rot, beam, args.lensRadius, args.focalDistance);
        lcfc = dofLensToFocal(rot, rot * beam.centerRay, args.focalDistance);
        lensBase = (args.occRank[tile] % DOF_LENS_TABLE_ROWS) * TILE_SIZE * A + tid;
        __syncthreads();
    }

    vector3 result(0.0f);
    uint32_t combined = 0;
    uint32_t tri0 = badTriIndex; // the triangle subsample 0 hit
    for (int slot = 0; slot < A; slot++) {
        if (combined == maskAll)
            break;
        GBufferSubsample g = waveBase[slot * 64 + lane];
        if (g.triIndex == badTriIndex)
            break;
        uint32_t mask = g.sampleMask;
        combined |= mask;
        if (mask & 1u)
            tri0 = g.triIndex;
        float inv = 1.0f / float(__popc(mask));
        float cx = 0.0f, cy = 0.0f;
        float lx = 0.0f, ly = 0.0f, dx = 0.0f, dy = 0.0f;
        for (uint32_t bits = mask; bits; bits &= bits - 1) {
            int i = __ffs(bits) - 1;
            vector2 o = args.offsets[i];
            cx = cx + o.x * inv;
            cy = cy + o.y * inv;
            if constexpr (DoF) {
                // centroid lens position and direction offset (resolve.cu:212-222)
                const vector2 lensUV = args.lensTable[lensBase + i * TILE_SIZE];
                const vector2 dirUV = dofDirUV(o, dofTile.focalToLensScale, lensUV);
                lx = lx + lensUV.x * inv;
                ly = ly + lensUV.y * inv;
                dx = dx + dirUV.x * inv;
                dy = dy + dirUV.y * inv;
            }
        }
        PrecomputedTriangleIntersect T = args.tris[g.triIndex];
        float bx, by, bz;
        if constexpr (DoF)
            dofUVW(T, cam, dofTile, lcfc, lx, ly, dx, dy, bx, by, bz); // at the centroid
        else
            pinholeUVW(T, cam, ndir, ndu, ndv, cx, cy, bx, by, bz);
        PrecomputedTriangleShade S = args.triShade[g.triIndex];
        ShadingVertex V0 = args.verts[S.indices[0]], V1 = args.verts[S.indices[1]], V2 = args.verts[S.indices[2]];
        vector3 pos = V0.pos * bx + V1.pos * by + V2.pos * bz;
        vector3 n0(halfToFloat(V0.normal[0]), halfToFloat(V0.normal[1]), halfToFloat(V0.normal[2]));
        vector3 n1(halfToFloat(V1.normal[0]), halfToFloat(V1.normal[1]), halfToFloat(V1.normal[2]));
        vector3 n2(halfToFloat(V2.normal[0]), halfToFloat(V2.normal[1]), halfToFloat(V2.normal[2]));
        vector3 nrm = n0 * bx + n1 * by + n2 * bz;
        const SimpleMaterial& mat = args.materials[S.material];
        const bool textured = Tex && materialTextured(mat, args.tex);
        vector2 uv(0.0f), dUVdX(0.0f), dUVdY(0.0f);
        if (textured) {
            // UV derivatives (resolve.cu:258-275): barycentrics one unit along each
            // subsample axis from the centroid (for depth of field, along each axis of the
            // focal-plane offset)
            float xb0, xb1, xb2, yb0, yb1, yb2;
            if constexpr (DoF) {
                dofUVW(T, cam, dofTile, lcfc, lx, ly, dx + 1.0f, dy, xb0, xb1, xb2);
                dofUVW(T, cam, dofTile, lcfc, lx, ly, dx, dy + 1.0f, yb0, yb1, yb2);
            } else {
                pinholeUVW(T, cam, ndir, ndu, ndv, cx + 1.0f, cy, xb0, xb1, xb2);
                pinholeUVW(T, cam, ndir, ndu, ndv, cx, cy + 1.0f, yb0, yb1, yb2);
            }
            uv = V0.uv * bx + V1.uv * by + V2.uv * bz;
            dUVdX = (V0.uv * xb0 + V1.uv * xb1 + V2.uv * xb2) - uv;
            dUVdY = (V0.uv * yb0 + V1.uv * yb1 + V2.uv * yb2) - uv;
        }
        vector3 c;
        if (textured) {
            const SurfaceInputs in = surfaceInputs(mat, args.tex, uv, dUVdX, dUVdY);
            c = Half ? ggxShadeHalf(in, pos, nrm, cam, args.env) : ggxShade(in, pos, nrm, cam, args.env);
        } else {
            const MaterialInputs in{mat};
            c = Half ? ggxShadeHalf(in, pos, nrm, cam, args.env) : ggxShade(in, pos, nrm, cam, args.env);
        }
        result += c * float(__popc(mask));
    }
    result = result * (1.0f / float(A));
    float r = linearToSRgbFast(acesFilm(result.x));
    float gch = linearToSRgbFast(acesFilm(result.y));
    float b = linearToSRgbFast(acesFilm(result.z));
    uint32_t ri = uint32_t(r * 255.0f), gi = uint32_t(gch * 255.0f), bi = uint32_t(b * 255.0f);
    args.sampleResults[sample] = ri | (gi << 8) | (bi << 16) | 0xff000000u;

    if (args.tMax) {
        // hit point of subsample 0 from its barycentrics, projected on the view axis
        float t = INFINITY;
        if (tri0 != badTriIndex) {
            PrecomputedTriangleIntersect T = args.tris[tri0];
            float bx, by, bz;
            const vector2 a0 = args.offsets[0];
            if constexpr (DoF) {
                const vector2 lensUV = args.lensTable[lensBase];
                const vector2 dirUV = dofDirUV(a0, dofTile.focalToLensScale, lensUV);
                dofUVW(T, cam, dofTile, lcfc, lensUV.x, lensUV.y, dirUV.x, dirUV.y, bx, by, bz);
            } else {
                pinholeUVW(T, cam, ndir, ndu, ndv, a0.x, a0.y, bx, by, bz);
            }
            const vector3 v1 = T.v0 + T.edge0, v2 = T.v0 + T.edge1;
            const vector3 pos = T.v0 * bx + v1 * by + v2 * bz;
            t = dot(pos - cam, -m.m2.xyz());
        }
        args.tMax[sample] = t;
    }
}

template <int A, bool DoF, bool Half, bool Tex>
__global__ __launch_bounds__(128) void resolveKernel(ResolveArgs args) {
    resolveBody<A, DoF, Half, Tex>(args, blockIdx.x);
}

// every tile of the views whose units `start` lists (pinhole views of one MSAA rate and
// shading precision)
template <int A, bool Half, bool Tex>
__global__ __launch_bounds__(128) void resolveBatchKernel(const ResolveArgs* __restrict__ all,
                                                          const uint32_t* __restrict__ start, uint32_t n,
                                                          uint32_t swizzle) {
    const uint32_t unit = xcdUnit(blockIdx.x, gridDim.x, swizzle);
    const uint32_t v = batchView(start, n, unit);
    resolveBody<A, false, Half, Tex>(all[v], unit - start[v]);
}

__device__ inline void remapBody(const RemapArgs& a, uint32_t x, uint32_t y) {
    const uint32_t* __restrict__ src = a.src;
    const int32_t* __restrict__ remap = a.remap;
    uint32_t* __restrict__ image = a.image;
    const uint32_t width = a.width, height = a.height, split = a.split;
    if (x >= width || y >= height)
        return;
    size_t pix = size_t(y) * width + x;
    if (split == 1) {
        int32_t o = remap[pix];
        if (o >= 0)
            image[pix] = src[o];
        return;
    }
    uint32_t c[3] = {0, 0, 0};
    for (uint32_t ch = 0; ch < split; ch++) {
        int32_t o = remap[pix * split + ch];
        if (o >= 0)
            c[ch] = src[o];
    }
    if (split == 2) // mergeSplitColorsPentile (remap.cu:25-31)
        image[pix] = (c[1] & 0xffu) | (c[0] & 0xff00u) | (c[1] & 0xff0000u) | 0xff000000u;
    else // mergeSplitColors (remap.cu:18-23)
        image[pix] = (c[0] & 0xffu) | (c[1] & 0xff00u) | (c[2] & 0xff0000u) | 0xff000000u;
}

__global__ void remapKernel(RemapArgs a) {
    remapBody(a, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y);
}

// 16x8 pixel blocks of every view, row-major per view
__global__ __launch_bounds__(128) void remapBatchKernel(const RemapArgs* __restrict__ all,
                                                        const uint32_t* __restrict__ start, uint32_t n) {
    const uint32_t v = batchView(start, n, blockIdx.x);
    const RemapArgs a = batchArgs(all, v);
    const uint32_t local = blockIdx.x - start[v], bx = (a.width + 15) / 16;
    remapBody(a, (local % bx) * 16 + (threadIdx.x & 15u), (local / bx) * 8 + (threadIdx.x >> 4));
}

ResolveArgs resolveArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m) {
    ResolveArgs a;
    a.sampleResults = v->sampleResults.ptr;
    a.gbuffer = v->gbuffer.ptr;
    a.beams = v->beams.ptr;
    a.tileCounts = v->tileCounts.ptr;
    a.tris = s->triIntersect.ptr;
    a.triShade = s->triShade.ptr;
    a.verts = s->vertsWorld.ptr;
    a.materials = s->materials.ptr;
    a.cameraToWorld = m;
    for (int i = 0; i < 32; i++)
        a.offsets[i] = v->offsets[i];
    a.env = s->lighting;
    a.lensTable = v->lensTable.ptr;
    a.occRank = v->occRank.ptr;
    a.lensRadius = v->lens.radius;
    a.focalDistance = v->lens.focalDistance;
    a.tMax = v->foveated ? v->tMax.ptr : nullptr;
    a.tex = {s->texels.ptr, s->texInfo.ptr, s->texLut.ptr, s->texLut.ptr + 256, s->textureCount};
    return a;
}

RemapArgs remapArgs(const hvvr_gpu_view* v) {
    RemapArgs a;
    a.src = v->sampleResults.ptr;
    a.remap = v->remap.ptr;
    a.image = v->image.ptr;
    a.width = v->width;
    a.height = v->height;
    a.split = v->split;
    return a;
}

void launchResolve(hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m, hipStream_t stream) {
    const ResolveArgs a = resolveArgs(v, s, m);
    dim3 grid(v->tileCount), block(128);
    const bool dof = v->dof;
    const bool half = v->halfShading;
    const bool tex = s->textureCount > 0;
    switch (v->msaaRate) {
#define HVVR_RESOLVE_TEX(A, D, F)                                                     \
    if (tex)                                                                          \
        resolveKernel<A, D, F, true><<<grid, block, 0, stream>>>(a);                  \
    else                                                                              \
        resolveKernel<A, D, F, false><<<grid, block, 0, stream>>>(a);
#define HVVR_RESOLVE_CASE(A)                                                          \
    case A:                                                                           \
        if (dof && half) {                                                            \
            HVVR_RESOLVE_TEX(A, true, true)                                           \
        } else if (dof) {                                                             \
            HVVR_RESOLVE_TEX(A, true, false)                                          \
        } else if (half) {                                                            \
            HVVR_RESOLVE_TEX(A, false, true)                                          \
        } else {                                                                      \
            HVVR_RESOLVE_TEX(A, false, false)                                         \
        }                                                                             \
        break;
        HVVR_RESOLVE_CASE(1)
        HVVR_RESOLVE_CASE(2)
        HVVR_RESOLVE_CASE(4)
        HVVR_RESOLVE_CASE(8)
        HVVR_RESOLVE_CASE(16)
        HVVR_RESOLVE_CASE(32)
#undef HVVR_RESOLVE_CASE
#undef HVVR_RESOLVE_TEX
        default: break;
    }
}

void launchRemap(hvvr_gpu_view* v, hipStream_t stream) {
    dim3 block(16, 8);
    dim3 grid((v->width + 15) / 16, (v->height + 7) / 8);
    remapKernel<<<grid, block, 0, stream>>>(remapArgs(v));
}

void launchResolveBatch(int msaaRate, bool half, bool tex, const ResolveArgs* all, const uint32_t* start,
                        uint32_t viewCount, uint32_t units, const LaunchShape& shape, hipStream_t stream) {
    switch (msaaRate) {
#define HVVR_RESOLVE_BATCH_KERNEL(A, F, T) \
    resolveBatchKernel<A, F, T><<<units, 128, 0, stream>>>(all, start, viewCount, shape.xcdSwizzle)
#define HVVR_RESOLVE_BATCH_CASE(A)                                                                            \
    case A:                                                                                                   \
        if (half && tex)                                                                                      \
            HVVR_RESOLVE_BATCH_KERNEL(A, true, true);                                                         \
        else if (half)                                                                                        \
            HVVR_RESOLVE_BATCH_KERNEL(A, true, false);                                                        \
        else if (tex)                                                                                         \
            HVVR_RESOLVE_BATCH_KERNEL(A, false, true);                                                        \
        else                                                                                                  \
            HVVR_RESOLVE_BATCH_KERNEL(A, false, false);                                                       \
        break;
        HVVR_RESOLVE_BATCH_CASE(1)
        HVVR_RESOLVE_BATCH_CASE(2)
        HVVR_RESOLVE_BATCH_CASE(4)
        HVVR_RESOLVE_BATCH_CASE(8)
        HVVR_RESOLVE_BATCH_CASE(16)
        HVVR_RESOLVE_BATCH_CASE(32)
#undef HVVR_RESOLVE_BATCH_CASE
#undef HVVR_RESOLVE_BATCH_KERNEL
        default: break;
    }
}

void launchRemapBatch(const RemapArgs* all, const uint32_t* start, uint32_t viewCount, uint32_t units,
                      hipStream_t stream) {
    remapBatchKernel<<<units, 128, 0, stream>>>(all, start, viewCount);
}

} // namespace hvvr
