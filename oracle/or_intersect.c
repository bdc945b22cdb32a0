/* Oracle: tile intersection and GBuffer compression (test infrastructure only).
 *
 *   IntersectKernel / IntersectTile / IntersectSamples   raycaster/intersect.cu:117-342
 *   FrustumPlanes                                        raycaster/prim_tests.h:18-50
 *   TestRayTriangle<false>, TestTriangleFrustaUVW        raycaster/prim_tests.h:53-116
 *   GetDifferentials                                     raycaster/prim_tests.h:118-134
 *   IntersectTriangleTile<false>::setup                  raycaster/prim_tests.h:148-163
 *   IntersectTriangleThread ctor / test                  raycaster/prim_tests.h:175-231
 *   TileData<false>::load, GetDirectionalSample3D        raycaster/tile_data.h:20-29, gpu_samples.h:287-291
 *   EmitGBuffer + sortBitonic                            raycaster/intersect.cu:28-89, sort.h:40-74
 *
 * The reference processes a tile's list in chunks of 128 entries: each chunk is culled
 * (one triangle per thread), survivors are compacted in list order, then every sample
 * tests every survivor in order. Chunking does not change the order in which a sample
 * meets the survivors, so this restatement culls the whole list, then tests it in order.
 * FMAs are placed where prim_tests.h:195-229 annotates them ("2 FMA", "1 FMA").
 * With depth of field (prim_tests.h:247-392, or_dof.h) the tile cull keeps partially
 * behind-the-lens triangles without the UVW test (intersect.cu:211-236).
 */
#include "or_dof.h"
#include <stdlib.h>

#define BAD_TRI 0xffffffffu

typedef struct {
    float t;
    ov3 e0, e1, v0ToRayOrigin;
} tri_tile;

/* FrustumPlanes ctor */
static void frustum_planes(const o_sfrustum* r, float p[4][4]) {
    for (int n = 0; n < 4; n++) {
        ov3 dir = cross3(r->d[n], r->d[(n + 1) % 4]);
        p[n][0] = dir.x;
        p[n][1] = dir.y;
        p[n][2] = dir.z;
        p[n][3] = dot3(r->o[n], dir);
    }
}

/* FrustumPlanes::test: false when all three vertices are outside one plane */
static int planes_test(const float p[4][4], const o_tri* t) {
    for (int n = 0; n < 4; n++) {
        ov3 dir = v3(p[n][0], p[n][1], p[n][2]);
        float d0 = dot3(dir, t->v0);
        float d1 = d0 + dot3(dir, t->e0);
        float d2 = d0 + dot3(dir, t->e1);
        if (d0 > p[n][3] && d1 > p[n][3] && d2 > p[n][3]) return 0;
    }
    return 1;
}

/* TestRayTriangle<false>: (u, v, w, t) */
static void ray_tri(ov3 o, ov3 dir, const o_tri* tri, float out[4]) {
    ov3 n = cross3(tri->e0, tri->e1);
    float denom = dot3(neg3(dir), n);
    ov3 v0ToRayOrigin = sub3(o, tri->v0);
    float t = dot3(v0ToRayOrigin, n);
    ov3 e = cross3(neg3(dir), v0ToRayOrigin);
    float v = dot3(tri->e1, e);
    float w = -dot3(tri->e0, e);
    float ood = 1.0f / denom;
    t *= ood;
    v *= ood;
    w *= ood;
    float u = 1.0f - v - w;
    out[0] = u;
    out[1] = v;
    out[2] = w;
    out[3] = t;
}

/* TestTriangleFrustaUVW: 0 = all out (reject), otherwise keep */
static int frusta_uvw_test(const o_sfrustum* r, const o_tri* tri) {
    float uvw[4][4];
    for (int n = 0; n < 4; n++) ray_tri(r->o[n], r->d[n], tri, uvw[n]);
    float uMin = FLT_MAX, uMax = -FLT_MAX, vMin = FLT_MAX, vMax = -FLT_MAX, wMin = FLT_MAX, wMax = -FLT_MAX;
    for (int n = 0; n < 4; n++) {
        uMin = fminf(uMin, uvw[n][0]);
        uMax = fmaxf(uMax, uvw[n][0]);
        vMin = fminf(vMin, uvw[n][1]);
        vMax = fmaxf(vMax, uvw[n][1]);
        wMin = fminf(wMin, uvw[n][2]);
        wMax = fmaxf(wMax, uvw[n][2]);
    }
    if (uMax < 0 || uMin > 1 || vMax < 0 || vMin > 1 || wMax < 0 || wMin > 1) return 0;
    return 1;
}

/* IntersectTriangleTile<false>::setup: 0 when the ray origin is behind the plane */
static int tile_setup(const o_tri* tri, ov3 rayOrigin, tri_tile* tt) {
    tt->e0 = tri->e0;
    tt->e1 = tri->e1;
    ov3 normal = cross3(tt->e0, tt->e1);
    tt->v0ToRayOrigin = sub3(rayOrigin, tri->v0);
    tt->t = dot3(tt->v0ToRayOrigin, normal);
    return !(tt->t < 0.0f);
}

typedef struct {
    float denomCenter, vCenter, wCenter;
    float dD[2], dV[2], dW[2];
} tri_thread;

/* IntersectTriangleThread ctor + GetDifferentials */
static void thread_setup(const tri_tile* tt, ov3 dir, ov3 du, ov3 dv, tri_thread* th) {
    ov3 normal = cross3(tt->e0, tt->e1);
    th->denomCenter = dot3(neg3(dir), normal);
    ov3 eCenter = cross3(neg3(dir), tt->v0ToRayOrigin);
    th->vCenter = dot3(tt->e1, eCenter);
    th->wCenter = -dot3(tt->e0, eCenter);
    ov3 n2 = cross3(tt->e0, tt->e1);
    th->dD[0] = dot3(neg3(du), n2);
    th->dD[1] = dot3(neg3(dv), n2);
    ov3 cu = cross3(neg3(du), tt->v0ToRayOrigin);
    ov3 cv = cross3(neg3(dv), tt->v0ToRayOrigin);
    th->dV[0] = dot3(tt->e1, cu);
    th->dV[1] = dot3(tt->e1, cv);
    th->dW[0] = -dot3(tt->e0, cu);
    th->dW[1] = -dot3(tt->e0, cv);
}

/* IntersectTriangleThread::test */
static int thread_test(const tri_thread* th, float t, ov2 a, float* depth) {
    float denom = fmaf(th->dD[1], a.y, fmaf(th->dD[0], a.x, th->denomCenter));
    float depthDelta = fmaf(*depth, denom, -t);
    float v = fmaf(th->dV[1], a.y, fmaf(th->dV[0], a.x, th->vCenter));
    float w = fmaf(th->dW[1], a.y, fmaf(th->dW[0], a.x, th->wCenter));
    float u = (denom - v) - w;
    int32_t test = (int32_t)(fbits(depthDelta) | fbits(u) | fbits(v) | fbits(w));
    if (test < 0) return 0;
    *depth = t * (1.0f / denom);
    return 1;
}

int or_intersect_view(const o_beam* beams, uint32_t sampleCount, const o_sfrustum* worldTiles, uint32_t tileCount,
                      const float m[16], const ov2* offsets, uint32_t A, const uint32_t* counts,
                      const uint32_t* lists, const o_tri* tris, uint32_t* triOut, float* depthOut,
                      uint64_t* survivorsOut) {
    return or_intersect_view_dof(beams, sampleCount, worldTiles, tileCount, m, offsets, A, counts, lists, tris, NULL,
                                 triOut, depthOut, survivorsOut);
}

int or_intersect_view_dof(const o_beam* beams, uint32_t sampleCount, const o_sfrustum* worldTiles,
                          uint32_t tileCount, const float m[16], const ov2* offsets, uint32_t A,
                          const uint32_t* counts, const uint32_t* lists, const o_tri* tris, const o_dof* dof,
                          uint32_t* triOut, float* depthOut, uint64_t* survivorsOut) {
    if (A < 1 || A > 32 || sampleCount < (uint64_t)tileCount * 128) return -1;
    if (dof && (!(dof->radius > 0.0f) || !dof->lensTable)) return -1;
    ov3 rayOrigin = v3(m[12], m[13], m[14]);
    uint64_t survivors = 0;
    uint64_t listOffset = 0;
    uint32_t* surv = (uint32_t*)malloc(sizeof(uint32_t) * 1);
    uint64_t survCap = 1;
    uint32_t rank = 0, rankTile = 0;
    for (uint32_t tile = 0; tile < tileCount; tile++) {
        uint32_t n = counts[tile];
        const uint32_t* list = lists + listOffset;
        listOffset += n;
        for (uint32_t s = 0; s < 128; s++)
            for (uint32_t i = 0; i < A; i++) {
                triOut[((uint64_t)tile * 128 + s) * A + i] = BAD_TRI;
                depthOut[((uint64_t)tile * 128 + s) * A + i] = FLT_MAX;
            }
        if (n == 0) continue;
        uint32_t row = dof ? dof_row(counts, tile, &rank, &rankTile) : 0;
        dof_tile dt;
        memset(&dt, 0, sizeof(dt));
        if (dof) {
            const o_beam* mid = &beams[(uint64_t)tile * 128 + 64];
            dof_tile_load(m, m3mul(m, mid->du), m3mul(m, mid->dv), dof->radius, dof->focalDistance, &dt);
        }
        float planes[4][4];
        frustum_planes(&worldTiles[tile], planes);
        if (n > survCap) {
            survCap = n;
            surv = (uint32_t*)realloc(surv, sizeof(uint32_t) * survCap);
        }
        uint32_t ns = 0;
        for (uint32_t k = 0; k < n; k++) {
            const o_tri* tri = &tris[list[k]];
            if (!planes_test(planes, tri)) continue;
            if (dof) {
                dof_tri d;
                int r = dof_tri_setup(tri, rayOrigin, dt.lensU, dt.lensV, &d);
                if (r == 0) continue;
                if (r == 2 && !frusta_uvw_test(&worldTiles[tile], tri)) continue;
            } else {
                tri_tile tt;
                if (!tile_setup(tri, rayOrigin, &tt)) continue;
                if (!frusta_uvw_test(&worldTiles[tile], tri)) continue;
            }
            surv[ns++] = list[k];
        }
        survivors += (uint64_t)ns * 128;
        for (uint32_t s = 0; s < 128; s++) {
            uint64_t sample = (uint64_t)tile * 128 + s;
            ov3 dir = m3mul(m, beams[sample].c);
            ov3 du = m3mul(m, beams[sample].du);
            ov3 dv = m3mul(m, beams[sample].dv);
            float depth[32];
            uint32_t hit[32];
            for (uint32_t i = 0; i < A; i++) {
                depth[i] = FLT_MAX;
                hit[i] = BAD_TRI;
            }
            ov3 lcfc = dof ? dof_lens_to_focal(m, dir, dof->focalDistance) : v3(0, 0, 0);
            for (uint32_t k = 0; k < ns; k++) {
                if (dof) {
                    dof_tri d;
                    dof_tri_setup(&tris[surv[k]], rayOrigin, dt.lensU, dt.lensV, &d);
                    ov3 normal = cross3(d.e0, d.e1);
                    float denomCenter = dot3(neg3(lcfc), normal);
                    for (uint32_t i = 0; i < A; i++) {
                        ov2 lensUV, dirUV;
                        dof_sample_uvs(dof->lensTable, row, A, i, s, offsets[i], dt.focalToLensScale, &lensUV, &dirUV);
                        if (dof_test(&d, denomCenter, lcfc, dt.lensU, dt.lensV, lensUV, dirUV, &depth[i]))
                            hit[i] = surv[k];
                    }
                } else {
                    tri_tile tt;
                    tile_setup(&tris[surv[k]], rayOrigin, &tt);
                    tri_thread th;
                    thread_setup(&tt, dir, du, dv, &th);
                    for (uint32_t i = 0; i < A; i++)
                        if (thread_test(&th, tt.t, offsets[i], &depth[i])) hit[i] = surv[k];
                }
            }
            for (uint32_t i = 0; i < A; i++) {
                triOut[sample * A + i] = hit[i];
                depthOut[sample * A + i] = depth[i];
            }
        }
    }
    free(surv);
    if (survivorsOut) *survivorsOut = survivors;
    return 0;
}

int or_dof_rays(const o_beam* beams, uint32_t sampleCount, const float m[16], const ov2* offsets, uint32_t A,
                const o_dof* dof, const uint32_t* counts, ov3* originsOut, ov3* dirsOut) {
    if (A < 1 || A > 32 || !dof || !(dof->radius > 0.0f) || !dof->lensTable) return -1;
    ov3 lensCenter = v3(m[12], m[13], m[14]);
    uint32_t rank = 0, rankTile = 0;
    for (uint32_t tile = 0; tile < sampleCount / 128; tile++) {
        int occupied = counts[tile] != 0;
        uint32_t row = occupied ? dof_row(counts, tile, &rank, &rankTile) : 0;
        dof_tile dt;
        memset(&dt, 0, sizeof(dt));
        if (occupied) {
            const o_beam* mid = &beams[(uint64_t)tile * 128 + 64];
            dof_tile_load(m, m3mul(m, mid->du), m3mul(m, mid->dv), dof->radius, dof->focalDistance, &dt);
        }
        for (uint32_t s = 0; s < 128; s++) {
            uint64_t sample = (uint64_t)tile * 128 + s;
            ov3 lcfc = dof_lens_to_focal(m, m3mul(m, beams[sample].c), dof->focalDistance);
            for (uint32_t i = 0; i < A; i++) {
                ov3 o = v3(0, 0, 0), d = v3(0, 0, 0);
                if (occupied) {
                    ov2 lensUV, dirUV;
                    dof_sample_uvs(dof->lensTable, row, A, i, s, offsets[i], dt.focalToLensScale, &lensUV, &dirUV);
                    o = axpy2(lensCenter, lensUV.x, dt.lensU, lensUV.y, dt.lensV);
                    d = axpy2(lcfc, dirUV.x, dt.lensU, dirUV.y, dt.lensV);
                }
                originsOut[sample * A + i] = o;
                dirsOut[sample * A + i] = d;
            }
        }
    }
    return 0;
}

uint32_t or_emit_gbuffer(const uint32_t* sampleTri, uint32_t A, uint32_t* slotTri, uint32_t* slotMask) {
    /* keys (tri << 5) | subsample, ascending; the miss id sorts last */
    uint32_t keys[32];
    for (uint32_t i = 0; i < A; i++) keys[i] = (sampleTri[i] << 5) | i;
    for (uint32_t i = 1; i < A; i++) { /* insertion sort: same order as a bitonic sort of distinct keys */
        uint32_t k = keys[i];
        int j = (int)i - 1;
        while (j >= 0 && keys[j] > k) {
            keys[j + 1] = keys[j];
            j--;
        }
        keys[j + 1] = k;
    }
    uint32_t count = 0;
    uint32_t curTri = (uint32_t)((int32_t)keys[0] >> 5);
    uint32_t curMask = 1u << (keys[0] & 31u);
    for (uint32_t i = 1; i < A; i++) {
        uint32_t t = (uint32_t)((int32_t)keys[i] >> 5);
        uint32_t bit = 1u << (keys[i] & 31u);
        if (t == curTri) {
            curMask |= bit;
        } else {
            slotTri[count] = curTri;
            slotMask[count] = curMask;
            count++;
            curTri = t;
            curMask = bit;
        }
    }
    slotTri[count] = curTri;
    slotMask[count] = curMask;
    return count + 1;
}
