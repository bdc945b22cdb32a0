#pragma once
// Barycentrics of a subsample ray on a triangle, shared by the resolve (shading at the
// centroid of a triangle's subsamples, the foveated t-max) and the intersect kernel's
// per-subsample hit output. Same float operations as the intersect's own test, so the
// barycentrics and the depth belong to the ray the hit was found on.

#include "dof.h"
#include "hvvr/types.h"

namespace hvvr {

// IntersectTriangleThread::calcUVW (prim_tests.h:233-245): barycentrics of the ray at
// subsample offset alpha; ndir/ndu/ndv are the negated world-space beam vectors
__device__ inline void pinholeUVW(const PrecomputedTriangleIntersect& T, const vector3& cam, const vector3& ndir,
                                  const vector3& ndu, const vector3& ndv, float cx, float cy, float& bx, float& by,
                                  float& bz) {
    vector3 normal = cross(T.edge0, T.edge1);
    vector3 v0o = cam - T.v0;
    float denomC = dot(ndir, normal);
    vector3 eC = cross(ndir, v0o);
    float vC = dot(T.edge1, eC);
    float wC = -dot(T.edge0, eC);
    float dD0 = dot(ndu, normal), dD1 = dot(ndv, normal);
    vector3 cu = cross(ndu, v0o), cv = cross(ndv, v0o);
    float dV0 = dot(T.edge1, cu), dV1 = dot(T.edge1, cv);
    float dW0 = -dot(T.edge0, cu), dW1 = -dot(T.edge0, cv);
    float denom = fmaf(dD1, cy, fmaf(dD0, cx, denomC));
    float v = fmaf(dV1, cy, fmaf(dV0, cx, vC));
    float w = fmaf(dW1, cy, fmaf(dW0, cx, wC));
    float u = (denom - v) - w;
    float dinv = 1.0f / denom;
    bx = u * dinv;
    by = v * dinv;
    bz = w * dinv;
}

// IntersectTriangleTile<true>::setup + IntersectTriangleThreadDoF::calcUVW
// (prim_tests.h:263-291, 366-392) at lens position (lx, ly) and direction offset (dx, dy)
__device__ inline void dofUVW(const PrecomputedTriangleIntersect& T, const vector3& cam, const DofTile& tileLens,
                              const vector3& lcfc, float lx, float ly, float dx, float dy, float& bx, float& by,
                              float& bz) {
    const vector3 lu = tileLens.lensU, lv = tileLens.lensV;
    const vector3 normal = cross(T.edge0, T.edge1);
    const vector3 v0ToLens = cam - T.v0;
    const float dDA0 = dot(-lu, normal), dDA1 = dot(-lv, normal);
    const float denomC = dot(-lcfc, normal);
    const float denom = fmaf(dDA1, dy, fmaf(dDA0, dx, denomC));
    const vector3 o = axpy2(v0ToLens, lx, lu, ly, lv);
    const vector3 rd = axpy2(lcfc, dx, lu, dy, lv);
    const vector3 e = crossFma(-rd, o);
    const float v = dotFma(T.edge1, e);
    const float w = -dotFma(T.edge0, e);
    const float u = (denom - v) - w;
    const float dinv = 1.0f / denom;
    bx = u * dinv;
    by = v * dinv;
    bz = w * dinv;
}

} // namespace hvvr
