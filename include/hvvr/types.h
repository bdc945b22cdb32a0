#pragma once
// Plain-old-data layouts and scalar math shared by the host C++ library and the
// gfx950 HIP kernels. Every struct here is byte-compatible with the reference's
// layout of the same name (cited per struct), so scene data, beams and frusta can
// cross the C ABI as raw arrays.
//
// Arithmetic contract (applies to every file that includes this header):
//   * all code is compiled with -ffp-contract=off: a*b+c is two roundings unless
//     written as fmaf();
//   * fmaf() appears exactly where the reference fuses (its _mm256_fmadd_ps calls
//     and the "n FMA" annotations in libraries/hvvr/raycaster/prim_tests.h);
//   * division and sqrt are IEEE correctly rounded (HIP default for f32);
//   * denormals are kept (no flush) on both the CPU and the GPU side.

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HVVR_HD __host__ __device__ inline
#define HVVR_DEVICE __device__ inline
#else
#define HVVR_HD inline
#endif

namespace hvvr {

// ---------------------------------------------------------------------------------
// vectors (libraries/hvvr/shared/vector_math.h)
// ---------------------------------------------------------------------------------
struct vector2 {
    float x, y;
    vector2() = default;
    HVVR_HD vector2(float _x, float _y) : x(_x), y(_y) {}
    HVVR_HD explicit vector2(float v) : x(v), y(v) {}
    HVVR_HD vector2 operator+(const vector2& v) const { return vector2(x + v.x, y + v.y); }
    HVVR_HD vector2 operator-(const vector2& v) const { return vector2(x - v.x, y - v.y); }
    HVVR_HD vector2 operator*(const vector2& v) const { return vector2(x * v.x, y * v.y); }
    HVVR_HD vector2 operator*(float a) const { return vector2(x * a, y * a); }
    HVVR_HD vector2 operator/(const vector2& v) const { return vector2(x / v.x, y / v.y); }
    HVVR_HD vector2 operator-() const { return vector2(-x, -y); }
    HVVR_HD vector2& operator+=(const vector2& v) { x += v.x; y += v.y; return *this; }
};
HVVR_HD vector2 operator*(float a, const vector2& v) { return vector2(a * v.x, a * v.y); }

struct vector3 {
    float x, y, z;
    vector3() = default;
    HVVR_HD vector3(float _x, float _y, float _z) : x(_x), y(_y), z(_z) {}
    HVVR_HD explicit vector3(float v) : x(v), y(v), z(v) {}
    HVVR_HD explicit vector3(const vector2& v, float _z) : x(v.x), y(v.y), z(_z) {}
    HVVR_HD vector3 operator+(const vector3& v) const { return vector3(x + v.x, y + v.y, z + v.z); }
    HVVR_HD vector3 operator-(const vector3& v) const { return vector3(x - v.x, y - v.y, z - v.z); }
    HVVR_HD vector3 operator*(const vector3& v) const { return vector3(x * v.x, y * v.y, z * v.z); }
    HVVR_HD vector3 operator/(const vector3& v) const { return vector3(x / v.x, y / v.y, z / v.z); }
    HVVR_HD vector3 operator*(float a) const { return vector3(x * a, y * a, z * a); }
    HVVR_HD vector3 operator/(float a) const { return vector3(x / a, y / a, z / a); }
    HVVR_HD vector3 operator-() const { return vector3(-x, -y, -z); }
    HVVR_HD vector3& operator+=(const vector3& v) { x += v.x; y += v.y; z += v.z; return *this; }
    HVVR_HD float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
HVVR_HD vector3 operator*(float a, const vector3& v) { return vector3(a * v.x, a * v.y, a * v.z); }

struct vector4 {
    float x, y, z, w;
    vector4() = default;
    HVVR_HD vector4(float _x, float _y, float _z, float _w) : x(_x), y(_y), z(_z), w(_w) {}
    HVVR_HD explicit vector4(float v) : x(v), y(v), z(v), w(v) {}
    HVVR_HD vector4(const vector3& v, float _w) : x(v.x), y(v.y), z(v.z), w(_w) {}
    HVVR_HD vector3 xyz() const { return vector3(x, y, z); }
    HVVR_HD vector4 operator+(const vector4& v) const { return vector4(x + v.x, y + v.y, z + v.z, w + v.w); }
    HVVR_HD vector4 operator-(const vector4& v) const { return vector4(x - v.x, y - v.y, z - v.z, w - v.w); }
    HVVR_HD vector4 operator*(float a) const { return vector4(x * a, y * a, z * a, w * a); }
    HVVR_HD vector4 operator/(float a) const { return vector4(x / a, y / a, z / a, w / a); }
};

struct vector2ui {
    uint32_t x, y;
};
struct vector2i {
    int32_t x, y;
};

HVVR_HD float dot(const vector3& a, const vector3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
HVVR_HD float dot(const vector2& a, const vector2& b) { return a.x * b.x + a.y * b.y; }
HVVR_HD vector3 cross(const vector3& a, const vector3& b) {
    return vector3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
HVVR_HD float length(const vector3& v) { return sqrtf(dot(v, v)); }
HVVR_HD float length(const vector2& v) { return sqrtf(dot(v, v)); }
HVVR_HD vector3 normalize(const vector3& v) { return v / length(v); }
HVVR_HD float fminv(float a, float b) { return a < b ? a : b; }
HVVR_HD float fmaxv(float a, float b) { return a > b ? a : b; }
HVVR_HD vector3 vmin(const vector3& a, const vector3& b) {
    return vector3(fminv(a.x, b.x), fminv(a.y, b.y), fminv(a.z, b.z));
}
HVVR_HD vector3 vmax(const vector3& a, const vector3& b) {
    return vector3(fmaxv(a.x, b.x), fmaxv(a.y, b.y), fmaxv(a.z, b.z));
}
// __saturatef semantics: clamp to [0,1], NaN -> 0
HVVR_HD float saturate(float x) { return x > 0.0f ? (x < 1.0f ? x : 1.0f) : 0.0f; }

// ---------------------------------------------------------------------------------
// quaternion / transform / column-major matrices (vector_math.h:770-1340)
// ---------------------------------------------------------------------------------
struct quaternion {
    float x, y, z, w;
    HVVR_HD quaternion() : x(0), y(0), z(0), w(1) {}
    HVVR_HD quaternion(float _x, float _y, float _z, float _w) : x(_x), y(_y), z(_z), w(_w) {}
    HVVR_HD static quaternion identity() { return quaternion(0, 0, 0, 1); }
    static quaternion fromAxisAngle(const vector3& axis, float radians) {
        float c = cosf(radians * 0.5f);
        float s = sinf(radians * 0.5f);
        return quaternion(axis.x * s, axis.y * s, axis.z * s, c);
    }
    HVVR_HD quaternion operator*(const quaternion& q) const {
        return quaternion(w * q.x + x * q.w + y * q.z - z * q.y, w * q.y - x * q.z + y * q.w + z * q.x,
                          w * q.z + x * q.y - y * q.x + z * q.w, w * q.w - (x * q.x + y * q.y + z * q.z));
    }
    HVVR_HD quaternion conj() const { return quaternion(-x, -y, -z, w); }
    HVVR_HD vector3 operator*(const vector3& d) const {
        quaternion r = (*this) * quaternion(d.x, d.y, d.z, 0.0f) * conj();
        return vector3(r.x, r.y, r.z);
    }
};

struct transform {
    quaternion rotation;
    vector3 translation;
    float scale;
    HVVR_HD transform() : rotation(), translation(0.0f), scale(1.0f) {}
    HVVR_HD transform(const vector3& t, const quaternion& r, float s) : rotation(r), translation(t), scale(s) {}
    HVVR_HD static transform identity() { return transform(vector3(0.0f), quaternion::identity(), 1.0f); }
};

struct matrix3x3 {
    vector3 m0, m1, m2; // columns
    matrix3x3() = default;
    HVVR_HD matrix3x3(const vector3& a, const vector3& b, const vector3& c) : m0(a), m1(b), m2(c) {}
    HVVR_HD static matrix3x3 identity() {
        return matrix3x3(vector3(1, 0, 0), vector3(0, 1, 0), vector3(0, 0, 1));
    }
    // vector_math.h:1008-1025
    HVVR_HD explicit matrix3x3(const quaternion& q) {
        float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
        float xx = q.x * x2, yy = q.y * y2, zz = q.z * z2;
        float xy = q.x * y2, yz = q.y * z2, xz = q.z * x2;
        float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
        m0 = vector3(1.0f - (yy + zz), xy + wz, xz - wy);
        m1 = vector3(xy - wz, 1.0f - (xx + zz), yz + wx);
        m2 = vector3(xz + wy, yz - wx, 1.0f - (xx + yy));
    }
    HVVR_HD vector3 operator*(const vector3& v) const {
        return vector3(m0.x * v.x + m1.x * v.y + m2.x * v.z, m0.y * v.x + m1.y * v.y + m2.y * v.z,
                       m0.z * v.x + m1.z * v.y + m2.z * v.z);
    }
    HVVR_HD matrix3x3 operator*(const matrix3x3& m) const {
        return matrix3x3((*this) * m.m0, (*this) * m.m1, (*this) * m.m2);
    }
    HVVR_HD matrix3x3 operator*(float s) const { return matrix3x3(m0 * s, m1 * s, m2 * s); }
    // vector_math.h:1072-1077
    static matrix3x3 rotationFromZAxis(const vector3& z) {
        vector3 y = (fabsf(z.y) > 0.85f) ? vector3(-1.0f, 0.0f, 0.0f) : vector3(0.0f, 1.0f, 0.0f);
        vector3 x = normalize(cross(y, z));
        y = normalize(cross(z, x));
        return matrix3x3(x, y, z);
    }
};

HVVR_HD matrix3x3 invert(const matrix3x3& m) {
    const vector3& x = m.m0;
    const vector3& y = m.m1;
    const vector3& z = m.m2;
    float det = x.x * (y.y * z.z - z.y * y.z) - x.y * (y.x * z.z - y.z * z.x) + x.z * (y.x * z.y - y.y * z.x);
    float s = 1.0f / det;
    return matrix3x3(vector3(y.y * z.z - z.y * y.z, x.z * z.y - x.y * z.z, x.y * y.z - x.z * y.y) * s,
                     vector3(y.z * z.x - y.x * z.z, x.x * z.z - x.z * z.x, y.x * x.z - x.x * y.z) * s,
                     vector3(y.x * z.y - z.x * y.y, z.x * x.y - x.x * z.y, x.x * y.y - y.x * x.y) * s);
}

struct matrix4x4 {
    vector4 m0, m1, m2, m3; // columns
    matrix4x4() = default;
    HVVR_HD matrix4x4(const vector4& a, const vector4& b, const vector4& c, const vector4& d)
        : m0(a), m1(b), m2(c), m3(d) {}
    HVVR_HD matrix4x4(const matrix3x3& m, const vector3& t)
        : m0(m.m0, 0.0f), m1(m.m1, 0.0f), m2(m.m2, 0.0f), m3(t, 1.0f) {}
    HVVR_HD static matrix4x4 identity() { return matrix4x4(matrix3x3::identity(), vector3(0.0f)); }
    // vector_math.h:1161: matrix3x3(t.rotation) * t.scale, t.translation
    HVVR_HD explicit matrix4x4(const transform& t) : matrix4x4(matrix3x3(t.rotation) * t.scale, t.translation) {}
    HVVR_HD matrix3x3 upper3x3() const { return matrix3x3(m0.xyz(), m1.xyz(), m2.xyz()); }
    HVVR_HD vector4 operator*(const vector4& v) const {
        return vector4(m0.x * v.x + m1.x * v.y + m2.x * v.z + m3.x * v.w,
                       m0.y * v.x + m1.y * v.y + m2.y * v.z + m3.y * v.w,
                       m0.z * v.x + m1.z * v.y + m2.z * v.z + m3.z * v.w,
                       m0.w * v.x + m1.w * v.y + m2.w * v.z + m3.w * v.w);
    }
    HVVR_HD matrix4x4 operator*(const matrix4x4& m) const {
        return matrix4x4((*this) * m.m0, (*this) * m.m1, (*this) * m.m2, (*this) * m.m3);
    }
};

// ---------------------------------------------------------------------------------
// scene layouts
// ---------------------------------------------------------------------------------
// IEEE binary16 bits, round-to-nearest-even conversion (F16C _cvtss_sh(x, 0) and
// CUDA __float2half both round to nearest even).
HVVR_HD uint16_t floatToHalfRNE(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    uint32_t x = v.u;
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t absx = x & 0x7fffffffu;
    if (absx >= 0x7f800000u) // inf or nan
        return (uint16_t)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u : 0u));
    if (absx >= 0x477ff000u) // rounds to >= 65520 -> inf
        return (uint16_t)(sign | 0x7c00u);
    if (absx < 0x38800000u) { // result is a half denormal or zero
        if (absx < 0x33000000u) // < 2^-25 rounds to zero
            return (uint16_t)sign;
        uint32_t mant = (absx & 0x7fffffu) | 0x800000u;
        int shift = 126 - (int)(absx >> 23); // 14..24
        uint32_t half = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (half & 1u)))
            half++;
        return (uint16_t)(sign | half);
    }
    uint32_t h = ((absx >> 13) - (112u << 10));
    uint32_t rem = absx & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u)))
        h++;
    return (uint16_t)(sign | h);
}
HVVR_HD float halfToFloat(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    uint32_t out;
    if (e == 0) {
        if (m == 0) {
            out = sign;
        } else { // denormal half -> normal float
            int shift = 0;
            while ((m & 0x400u) == 0) {
                m <<= 1;
                shift++;
            }
            m &= 0x3ffu;
            out = sign | ((uint32_t)(113 - shift) << 23) | (m << 13);
        }
    } else if (e == 31) {
        out = sign | 0x7f800000u | (m << 13);
    } else {
        out = sign | ((e + 112u) << 23) | (m << 13);
    }
    union { uint32_t u; float f; } v;
    v.u = out;
    return v.f;
}

// ShadingVertex (graphics_types.h:83-88), 32 bytes
struct ShadingVertex {
    vector3 pos;
    uint16_t normal[4]; // vector4h
    vector2 uv;
    uint32_t pad0;
};
static_assert(sizeof(ShadingVertex) == 32, "ShadingVertex layout");

// PrecomputedTriangleIntersect / Shade (graphics_types.h:203-211)
struct PrecomputedTriangleIntersect {
    vector3 v0;
    vector3 edge0;
    vector3 edge1;
};
static_assert(sizeof(PrecomputedTriangleIntersect) == 36, "PrecomputedTriangleIntersect layout");
struct PrecomputedTriangleShade {
    uint32_t indices[3];
    uint32_t material;
};

// BVHNode (raycaster/bvh_node.h:17-37), 128 bytes. bounds[0..5] = xMax, xNegMin,
// yMax, yNegMin, zMax, zNegMin (4 children each). data[] is the reference's BoxData
// union: data[0] = leaf.triIndex[0]; data[1+k] = children.offset[k] (internal child k)
// or leaf.triIndex[1+k] (leaf child k).
struct alignas(32) BVHNode {
    float bounds[6][4];
    uint32_t leafMask;
    uint32_t data[5];
    uint32_t pad[2];
};
static_assert(sizeof(BVHNode) == 128, "BVHNode layout");

// TopologyNode (graphics_types.h:126-151)
struct TopologyNode {
    uint32_t leafMask;
    uint32_t data[7];
    bool isLeaf(int i) const { return (leafMask >> i) & 1u; }
};

// SimpleMaterial (raycaster/material.h:22-49), 64 bytes
enum : uint32_t { ShadingModelNone = 0, ShadingModelPhong = 1, ShadingModelEmissive = 2 };
struct SimpleMaterial {
    enum : uint32_t { textureIndexBitCount = 16 };
    enum : uint32_t { textureMask = (1u << textureIndexBitCount) - 1u };
    enum : uint32_t { badTextureIndex = textureMask };
    enum : uint32_t { diffuseBitShift = 48, specularBitShift = 32, glossyBitShift = 16 };
    static uint64_t buildShadingCode(uint32_t shadingModel, uint32_t diffuseID, uint32_t specularID,
                                     uint32_t glossinessID) {
        return uint64_t(shadingModel) | (uint64_t(diffuseID) << diffuseBitShift) |
               (uint64_t(glossinessID) << glossyBitShift) | (uint64_t(specularID) << specularBitShift);
    }
    vector4 emissive;
    vector4 diffuse;
    vector4 specular;
    float glossiness;
    float opacity;
    uint64_t textureIDsAndShadingModel;
};
static_assert(sizeof(SimpleMaterial) == 64, "SimpleMaterial layout");

// lights (raycaster/light.h:19-59)
enum : uint32_t { LightTypeNone = 0, LightTypeDirectional = 1, LightTypePoint = 2, LightTypeSpot = 3 };
struct LightDirectional {
    vector3 Direction;
    vector3 Power;
};
struct LightPoint {
    vector3 Position;
    float FalloffEnd;
    vector3 Color;
    float FalloffScale;
};
struct LightSpot {
    vector3 Position;
    float FalloffEnd;
    vector3 Direction;
    float FalloffScale;
    vector3 Color;
    float CosOuterAngle;
    float CosAngleScale;
};
struct LightUnion {
    uint32_t type;
    union {
        LightDirectional directional;
        LightPoint point;
        LightSpot spot;
    };
    LightUnion() : type(LightTypeNone) {}
};
static_assert(sizeof(LightUnion) == 56, "LightUnion layout");

enum { MAX_DIRECTIONAL_LIGHTS = 1, MAX_POINT_LIGHTS = 4, MAX_SPOT_LIGHTS = 4 };
// LightingEnvironment (raycaster/lighting_env.h:19-25)
struct LightingEnvironment {
    LightDirectional directionalLights[MAX_DIRECTIONAL_LIGHTS];
    LightPoint pointLights[MAX_POINT_LIGHTS];
    LightSpot spotLights[MAX_SPOT_LIGHTS];
    int32_t directionalLightCount;
    int32_t pointLightCount;
    int32_t spotLightCount;
};
static_assert(sizeof(LightingEnvironment) == 372, "LightingEnvironment layout");

// ---------------------------------------------------------------------------------
// camera / sample layouts
// ---------------------------------------------------------------------------------
// tiling constants (raycaster/raycaster_common.h:17-20)
enum : uint32_t { TILE_SIZE = 128, TILES_PER_BLOCK = 64, BLOCK_SIZE = TILE_SIZE * TILES_PER_BLOCK };

struct ThinLens {
    float radius;
    float focalDistance;
};
// DirectionalBeam (raycaster/samples.h:63-67)
struct DirectionalBeam {
    vector3 centerRay;
    vector3 du;
    vector3 dv;
};
static_assert(sizeof(DirectionalBeam) == 36, "DirectionalBeam layout");
// SimpleRayFrustum (graphics_types.h:182-185)
// SimpleRay (graphics_types.h:175-178): one ray of a ray dump
struct SimpleRay {
    vector3 origin;
    vector3 direction;
};
static_assert(sizeof(SimpleRay) == 24, "SimpleRay layout");

struct SimpleRayFrustum {
    vector3 origins[4];
    vector3 directions[4];
};
static_assert(sizeof(SimpleRayFrustum) == 96, "SimpleRayFrustum layout");

// RaycasterGBufferSubsample (raycaster/gpu_camera.h:31-39)
struct GBufferSubsample {
    uint32_t triIndex;
    uint32_t sampleMask;
};
enum : uint32_t { badTriIndex = 0xffffffffu };

// TileTriRange (raycaster/gpu_camera.h:47-50)
struct TileTriRange {
    uint32_t start;
    uint32_t end;
};

// A world-space frustum pre-swizzled for the child-AABB tests. This is this
// library's layout of the data traverse::avx::Frustum precomputes
// (raycaster/traverse_avx.cpp:22-211); the arithmetic that fills it is restated in
// frustum_setup() below.
struct FrustumPre {
    float planeNeg[4][4]; // [plane] = (x<0?x:0, y<0?y:0, z<0?z:0, w), multiplies the max lanes
    float planePos[4][4]; // [plane] = (x>0?-x:0, ...), multiplies the negated-min lanes
    float projLo[3];      // child passes axis a if bounds[2a][c] >= projLo[a]
    float projHi[3];      //                     and bounds[2a+1][c] >= projHi[a]
    float edge[12][3];
    float edgeMin[12];
    float edgeMax[12];
    int32_t distanceAxis; // index into BVHNode::bounds used for the distance estimate
    int32_t degenerate;   // plane[0].w == -inf: the frustum culls everything
};
static_assert(sizeof(FrustumPre) == 400, "FrustumPre layout");

// one entry of a block's sorted stack (the reference's BlockFrame::sortedStack)
struct StackEntry {
    float tMin;
    uint32_t nodeMask; // node index in bits 0..27, leaf mask in bits 28..31
};
enum : uint32_t { STACK_NODE_MASK = 0x0fffffffu };
enum : uint32_t { BLOCK_STACK_SIZE = 64 };
enum : uint32_t { TILE_STACK_SIZE = 4096 };

} // namespace hvvr
