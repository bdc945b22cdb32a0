/* Depth-of-field helpers shared by the oracle's intersect and resolve (test
 * infrastructure only). Float arithmetic with FMAs where prim_tests.h annotates them.
 *
 *   TileData<true>::load                 raycaster/tile_data.h:37-60
 *   lensCenterToFocalCenter              raycaster/intersect.cu:127-129, resolve.cu:353-355
 *   GetSampleUVsDoF                      raycaster/gpu_samples.h:84-103
 *   IntersectTriangleTile<true>::setup   raycaster/prim_tests.h:263-291
 *   IntersectTriangleThreadDoF           raycaster/prim_tests.h:295-392
 */
#ifndef OR_DOF_H
#define OR_DOF_H
#include "or_math.h"

/* "3 MUL 3 FMA" cross product and "1 MUL 2 FMA" dot product of the DoF test */
static inline ov3 cross3_fma(ov3 a, ov3 b) {
    return v3(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
static inline float dot3_fma(ov3 a, ov3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
/* a + s*u + t*v per component, "6 FMA" */
static inline ov3 axpy2(ov3 a, float s, ov3 u, float t, ov3 v) {
    return v3(fmaf(t, v.x, fmaf(s, u.x, a.x)), fmaf(t, v.y, fmaf(s, u.y, a.y)), fmaf(t, v.z, fmaf(s, u.z, a.z)));
}

typedef struct {
    ov3 lensU, lensV;
    ov2 focalToLensScale;
} dof_tile;

/* TileData<true>::load from the world-space beam of the tile's sample 64. The reference
 * rotates that beam by the camera rotation once more (tile_data.h:43-46 applied to the
 * output of GetDirectionalSample3D); kept as is. rsqrtf is 1/sqrtf here. */
static inline void dof_tile_load(const float m[16], ov3 duWorld, ov3 dvWorld, float radius, float focal,
                                 dof_tile* t) {
    ov3 focalU = m3mul(m, duWorld), focalV = m3mul(m, dvWorld);
    float focalUMagInv = 1.0f / sqrtf(dot3(focalU, focalU));
    float focalVMagInv = 1.0f / sqrtf(dot3(focalV, focalV));
    float focalUMag = 1.0f / focalUMagInv, focalVMag = 1.0f / focalVMagInv;
    t->lensU = scale3(focalU, radius * focalUMagInv);
    t->lensV = scale3(focalV, radius * focalVMagInv);
    float s = focal / radius;
    t->focalToLensScale.x = s * focalUMag;
    t->focalToLensScale.y = s * focalVMag;
}

/* (focalDistance / zed) * normalize(centerRay), zed = dot(R * (0,0,-1), dir) */
static inline ov3 dof_lens_to_focal(const float m[16], ov3 dirWorld, float focal) {
    ov3 centerDir = norm3(dirWorld);
    float zed = dot3(m3mul(m, v3(0.0f, 0.0f, -1.0f)), centerDir);
    return scale3(centerDir, focal / zed);
}

/* GetSampleUVsDoF: lens position from the table, direction offset "2 FMA" */
static inline void dof_sample_uvs(const ov2* table, uint32_t row, uint32_t A, uint32_t sub, uint32_t s, ov2 aa,
                                  ov2 scale, ov2* lensUV, ov2* dirUV) {
    *lensUV = table[(uint64_t)row * 128 * A + (uint64_t)sub * 128 + s];
    dirUV->x = fmaf(aa.x, scale.x, -lensUV->x);
    dirUV->y = fmaf(aa.y, scale.y, -lensUV->y);
}

typedef struct {
    ov3 e0, e1, v0ToLensCenter;
    float tCenter, tU, tV;
    ov2 dDenomDAlpha;
} dof_tri;

/* IntersectTriangleTile<true>::setup: 0 = all out, 1 = partial, 2 = all in */
static inline int dof_tri_setup(const o_tri* tri, ov3 lensCenter, ov3 lensU, ov3 lensV, dof_tri* d) {
    d->dDenomDAlpha.x = d->dDenomDAlpha.y = 0.0f;
    d->e0 = tri->e0;
    d->e1 = tri->e1;
    ov3 normal = cross3(d->e0, d->e1);
    d->v0ToLensCenter = sub3(lensCenter, tri->v0);
    d->tCenter = dot3(d->v0ToLensCenter, normal);
    d->tU = dot3(lensU, normal);
    d->tV = dot3(lensV, normal);
    float tExtent = fabsf(d->tU) + fabsf(d->tV);
    float tMax = d->tCenter + tExtent;
    if (tMax < 0.0f) return 0;
    d->dDenomDAlpha.x = dot3(neg3(lensU), normal);
    d->dDenomDAlpha.y = dot3(neg3(lensV), normal);
    float tMin = d->tCenter - tExtent;
    return tMin < 0.0f ? 1 : 2;
}

/* IntersectTriangleThreadDoF::test */
static inline int dof_test(const dof_tri* d, float denomCenter, ov3 lcfc, ov3 lensU, ov3 lensV, ov2 lensUV,
                           ov2 dirUV, float* depth) {
    float t = fmaf(d->tV, lensUV.y, fmaf(d->tU, lensUV.x, d->tCenter));
    float denom = fmaf(d->dDenomDAlpha.y, dirUV.y, fmaf(d->dDenomDAlpha.x, dirUV.x, denomCenter));
    float depthDelta = fmaf(*depth, denom, -t);
    ov3 v0ToRayOrigin = axpy2(d->v0ToLensCenter, lensUV.x, lensU, lensUV.y, lensV);
    ov3 rayDir = axpy2(lcfc, dirUV.x, lensU, dirUV.y, lensV);
    ov3 e = cross3_fma(neg3(rayDir), v0ToRayOrigin);
    float v = dot3_fma(d->e1, e);
    float w = -dot3_fma(d->e0, e);
    float u = (denom - v) - w;
    int32_t test = (int32_t)(fbits(t) | fbits(depthDelta) | fbits(u) | fbits(v) | fbits(w));
    if (test < 0) return 0;
    *depth = t * (1.0f / denom);
    return 1;
}

/* rank of each tile among the tiles with a non-empty list (the reference's compacted
 * tile index), mod 4: the lens-table row */
static inline uint32_t dof_row(const uint32_t* counts, uint32_t tile, uint32_t* rankCache, uint32_t* cacheTile) {
    /* incremental: callers walk tiles in ascending order */
    while (*cacheTile < tile) {
        if (counts[*cacheTile]) (*rankCache)++;
        (*cacheTile)++;
    }
    return *rankCache % 4u;
}

/* per-sample depth-of-field context for the resolve */
typedef struct {
    dof_tile tile;
    ov3 lensToFocal;
    const ov2* table;
    uint32_t row, sample;
} dof_sample_ctx;

#endif
