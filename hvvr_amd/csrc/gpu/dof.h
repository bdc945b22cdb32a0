#pragma once
// Depth-of-field (thin lens) helpers shared by the intersect and resolve kernels.
//
//   TileData<true>::load                 raycaster/tile_data.h:37-60
//   lensCenterToFocalCenter              raycaster/intersect.cu:127-129, resolve.cu:353-355
//   GetSampleUVsDoF                      raycaster/gpu_samples.h:84-103
//   IntersectTriangleTile<true>::setup   raycaster/prim_tests.h:263-291
//   IntersectTriangleThreadDoF           raycaster/prim_tests.h:295-392
//
// FMAs sit where prim_tests.h annotates them ("2 FMA", "6 FMA", "3 MUL 3 FMA", "1 MUL 2
// FMA"); everything else is plain float arithmetic, as in the CPU oracle (or_dof.h).

#include "hvvr/types.h"

namespace hvvr {

// rows of the lens-position table (DOF_LENS_POS_LOOKUP_TABLE_TILES, kernel_constants.h:23)
enum : uint32_t { DOF_LENS_TABLE_ROWS = 4 };

struct DofTile {
    vector3 lensU, lensV;
    vector2 focalToLensScale;
};

// TileData<true>::load from the camera-space beam of the tile's sample 64. The reference
// applies the camera rotation to that beam twice (GetDirectionalSample3D, then
// tile_data.h:43-46); kept for parity. rsqrtf is 1/sqrtf here.
__device__ inline DofTile dofTileLoad(const matrix3x3& rot, const DirectionalBeam& beam, float radius, float focal) {
    const vector3 focalU = rot * (rot * beam.du);
    const vector3 focalV = rot * (rot * beam.dv);
    const float focalUMagInv = 1.0f / sqrtf(dot(focalU, focalU));
    const float focalVMagInv = 1.0f / sqrtf(dot(focalV, focalV));
    const float focalUMag = 1.0f / focalUMagInv, focalVMag = 1.0f / focalVMagInv;
    DofTile t;
    t.lensU = focalU * (radius * focalUMagInv);
    t.lensV = focalV * (radius * focalVMagInv);
    const float s = focal / radius;
    t.focalToLensScale = vector2(s * focalUMag, s * focalVMag);
    return t;
}

// (focalDistance / zed) * normalize(centerRay), zed = dot(R * (0,0,-1), dir); dir world space
__device__ inline vector3 dofLensToFocal(const matrix3x3& rot, const vector3& dirWorld, float focal) {
    const vector3 centerDir = normalize(dirWorld);
    const float zed = dot(rot * vector3(0.0f, 0.0f, -1.0f), centerDir);
    return centerDir * (focal / zed);
}

__device__ inline vector3 crossFma(const vector3& a, const vector3& b) {
    return vector3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ inline float dotFma(const vector3& a, const vector3& b) {
    return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x));
}
// a + s*u + t*v per component
__device__ inline vector3 axpy2(const vector3& a, float s, const vector3& u, float t, const vector3& v) {
    return vector3(fmaf(t, v.x, fmaf(s, u.x, a.x)), fmaf(t, v.y, fmaf(s, u.y, a.y)), fmaf(t, v.z, fmaf(s, u.z, a.z)));
}

// GetSampleUVsDoF: dirUV = aaOffset * focalToLensScale - lensUV
__device__ inline vector2 dofDirUV(vector2 aa, vector2 scale, vector2 lensUV) {
    return vector2(fmaf(aa.x, scale.x, -lensUV.x), fmaf(aa.y, scale.y, -lensUV.y));
}

// IntersectTriangleTile<true> (plus the triangle normal, which the per-sample
// constructor recomputes from the same edges)
struct TriTileDoF {
    vector3 e0, e1, normal, v0ToLensCenter;
    float tCenter, tU, tV;
    vector2 dDenomDAlpha;
};

// 0 = all rays behind the plane (cull), 1 = partial, 2 = all in front
__device__ inline int dofTriSetup(const PrecomputedTriangleIntersect& tri, const vector3& lensCenter,
                                  const vector3& lensU, const vector3& lensV, TriTileDoF& d) {
    d.e0 = tri.edge0;
    d.e1 = tri.edge1;
    d.normal = cross(d.e0, d.e1);
    d.v0ToLensCenter = lensCenter - tri.v0;
    d.tCenter = dot(d.v0ToLensCenter, d.normal);
    d.tU = dot(lensU, d.normal);
    d.tV = dot(lensV, d.normal);
    const float tExtent = fabsf(d.tU) + fabsf(d.tV);
    const float tMax = d.tCenter + tExtent;
    d.dDenomDAlpha = vector2(0.0f, 0.0f);
    if (tMax < 0.0f)
        return 0;
    d.dDenomDAlpha = vector2(dot(-lensU, d.normal), dot(-lensV, d.normal));
    const float tMin = d.tCenter - tExtent;
    return tMin < 0.0f ? 1 : 2;
}

// IntersectTriangleThreadDoF::test; denomCenter = dot(-lensCenterToFocalCenter, normal)
__device__ inline bool dofTest(const TriTileDoF& d, float denomCenter, const vector3& lcfc, const vector3& lensU,
                               const vector3& lensV, vector2 lensUV, vector2 dirUV, float& depth) {
    const float t = fmaf(d.tV, lensUV.y, fmaf(d.tU, lensUV.x, d.tCenter));
    const float denom = fmaf(d.dDenomDAlpha.y, dirUV.y, fmaf(d.dDenomDAlpha.x, dirUV.x, denomCenter));
    const float depthDelta = fmaf(depth, denom, -t);
    const vector3 o = axpy2(d.v0ToLensCenter, lensUV.x, lensU, lensUV.y, lensV);
    const vector3 rd = axpy2(lcfc, dirUV.x, lensU, dirUV.y, lensV);
    const vector3 e = crossFma(-rd, o);
    const float v = dotFma(d.e1, e);
    const float w = -dotFma(d.e0, e);
    const float u = (denom - v) - w;
    const int32_t test = int32_t(__float_as_uint(t) | __float_as_uint(depthDelta) | __float_as_uint(u) |
                                 __float_as_uint(v) | __float_as_uint(w));
    if (test < 0)
        return false;
    depth = t * (1.0f / denom);
    return true;
}

} // namespace hvvr
