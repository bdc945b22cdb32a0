// Triangle-list intersection for one view (raycaster/intersect.cu:117-342 with the
// primitive tests of raycaster/prim_tests.h:18-231), one 128-thread workgroup per tile
// (= two wave64s), one thread per sample, A subsamples per thread.
//
// Per tile the list is processed in chunks of 128 entries. Each thread culls one
// triangle against the tile (4 frustum planes, back-face/behind-origin, 4-corner
// barycentric range); survivors are compacted into LDS in list order with a 64-bit
// wave ballot + popcount and a two-wave prefix; then every thread tests its A
// subsamples against every survivor in order, keeping the nearest hit. The per-tile
// triangle setup (edges, normal, origin offset, plane distance t) is computed once
// by the culling thread and shared through LDS, so the per-sample loop reads it with
// LDS broadcasts. FMAs are placed as annotated in prim_tests.h:195-229.
//
// Output: the compressed GBuffer (EmitGBuffer, intersect.cu:28-89) laid out per
// wave64: slot k of lane l of wave w at gbuffer[w*64*A + k*64 + l]. If every lane of
// the wave saw a single triangle id across its subsamples, only slot 0 is written.
// Tiles with an empty list also write their (all-miss) GBuffer, so every GBuffer slot
// the resolve pass can reach was written this frame.

#include "device_common.h"
#include "dof.h"
#include "gpu_internal.h"
#include "launch_args.h"
#include "uvw.h"

#include <type_traits>

namespace hvvr {


struct TriTile {
    float t;
    vector3 e0, e1, normal, v0ToRayOrigin;
};

template <int N>
__device__ inline void sortKeysAscending(uint32_t* k) {
    // bitonic network (raycaster/sort.h:40-74) over distinct keys; fully unrolled so the
    // keys stay in registers
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size / 2; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < N / 2; i++) {
                int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
                int hi = lo | stride;
                bool ascending = (lo & size) == 0;
                uint32_t a = k[lo], b = k[hi];
                bool swap = ascending ? (a > b) : (a < b);
                k[lo] = swap ? b : a;
                k[hi] = swap ? a : b;
            }
        }
    }
}

template <int A>
__device__ inline void emitGBuffer(const uint32_t* hit, GBufferSubsample* waveBase, uint32_t lane) {
    bool oneTri = true;
#pragma unroll
    for (int i = 1; i < A; i++)
        oneTri = oneTri && (hit[i] == hit[0]);
    const uint32_t maskAll = A < 32 ? ((1u << A) - 1u) : 0xffffffffu;
    if (__ballot(oneTri) == ~0ull) {
        GBufferSubsample g;
        g.triIndex = hit[0];
        g.sampleMask = maskAll;
        waveBase[lane] = g;
        return;
    }
    uint32_t keys[A];
#pragma unroll
    for (int i = 0; i < A; i++)
        keys[i] = (hit[i] << 5) | uint32_t(i);
    sortKeysAscending<A>(keys);
    GBufferSubsample g;
    g.triIndex = uint32_t(int32_t(keys[0]) >> 5);
    g.sampleMask = 1u << (keys[0] & 31u);
    uint32_t slot = 0;
#pragma unroll
    for (int i = 1; i < A; i++) {
        uint32_t t = uint32_t(int32_t(keys[i]) >> 5);
        uint32_t bit = 1u << (keys[i] & 31u);
        if (t == g.triIndex) {
            g.sampleMask |= bit;
        } else {
            waveBase[slot * 64 + lane] = g;
            slot++;
            g.triIndex = t;
            g.sampleMask = bit;
        }
    }
    waveBase[slot * 64 + lane] = g;
}

// TestTriangleFrustaUVW (prim_tests.h:89-116): false when all four corner rays of the tile
// miss the triangle on the same side of one barycentric coordinate
__device__ inline bool frustaUVWPass(const SimpleRayFrustum& frustum, const PrecomputedTriangleIntersect& tri) {
    float uMin = FLT_MAX, uMax = -FLT_MAX, vMin = FLT_MAX, vMax = -FLT_MAX;
    float wMin = FLT_MAX, wMax = -FLT_MAX;
    for (int n = 0; n < 4; n++) {
        vector3 rdir = frustum.directions[n];
        vector3 o = frustum.origins[n];
        vector3 nn = cross(tri.edge0, tri.edge1);
        float denom = dot(-rdir, nn);
        vector3 v0o = o - tri.v0;
        float tr = dot(v0o, nn);
        vector3 e = cross(-rdir, v0o);
        float v = dot(tri.edge1, e);
        float w = -dot(tri.edge0, e);
        float ood = 1.0f / denom;
        tr *= ood;
        v *= ood;
        w *= ood;
        float u = 1.0f - v - w;
        (void)tr;
        uMin = fminf(uMin, u);
        uMax = fmaxf(uMax, u);
        vMin = fminf(vMin, v);
        vMax = fmaxf(vMax, v);
        wMin = fminf(wMin, w);
        wMax = fmaxf(wMax, w);
    }
    return !(uMax < 0 || uMin > 1 || vMax < 0 || vMin > 1 || wMax < 0 || wMin > 1);
}

// ---------------------------------------------------------------------------------
// Occlusion skip (pinhole path). The lists arrive roughly front to back (the tile walk's
// distance order), so after the first triangles most of a wave's subsamples already hold a
// near hit. A survivor is skipped for a whole wave when no subsample ray of the wave can
// pass its depth test, i.e. when the test's computed `depth * denom - t` is negative for
// every subsample. This changes no result (the test would have rejected it), only the
// work. Two sufficient conditions, with `far` the wave's largest current depth:
//  - plane: t > far * (upper bound of every computed denom over the wave): then
//    depth * denom <= far * denom < t for each subsample (denom < 0 gives depth * denom <= 0);
//  - vertices: a ray the test accepts hits the triangle's plane at a point whose
//    barycentrics are at least -E / den (E: bound on the rounding of the computed u, v, w;
//    den: lower bound of the true denominators over the wave), so its ray parameter is at
//    least (min over vertices of (v - origin) . f - (sum E / den) * edge spread) / max (D . f)
//    for any direction f with D . f > 0; when that exceeds far (with the rounding of t and
//    denom), the depth test fails.
// The ray of subsample a is D = -(ndir + a.x ndu + a.y ndv) in exact arithmetic on the
// lane's computed vectors; the bounds use L1 norms (never below the Euclidean ones) and
// leave a relative margin of 1e-5 (more than a hundred roundings) on every quantity, so
// each holds for the computed values the test compares, not only for exact geometry.
struct WaveBundle {
    vector3 dc, f;      // lane 0's centre ray, and it normalised
    float r;            // bound on |D - dc| (L1) over the wave's subsample rays
    float dsum;         // bound on |ndir| + |a.x||ndu| + |a.y||ndv| (L1)
    float dcL1;
    float dfMax, dfMin; // bounds on D . f
};

constexpr float kOccSlack = 1e-5f;
// survivors per occlusion test, each test with the wave's farthest hit refreshed. On a
// 1080x1200 A = 4 eye of the room scene the intersect takes 0.110 ms with 8, 0.1155 ms
// with 16 and 0.121 ms with 32 (fewer skips), 0.111-0.115 ms with 6 and 4 (the extra tests
// cost more than the extra skips save).
constexpr uint32_t kOccBatch = 8;
constexpr float kUlp = 0x1p-24f;

__device__ inline float l1(const vector3& v) { return fabsf(v.x) + fabsf(v.y) + fabsf(v.z); }

// The wave's largest value, wave-uniform; every lane active. Within each row of 16 lanes,
// four DPP row shifts leave the row's maximum in its last lane; row_bcast:15 and :31 carry
// rows 0-2 into lane 63, which readlane broadcasts. (A butterfly of shuffles took six LDS
// round trips per call.)
template <int Ctrl, int RowMask>
__device__ inline float dppMaxStep(float v) {
    const int b = int(fbits(v));
    return fmaxf(v, bitsf(uint32_t(__builtin_amdgcn_update_dpp(b, b, Ctrl, RowMask, 0xf, false))));
}
__device__ inline float waveMaxF(float v) {
    v = dppMaxStep<0x111, 0xf>(v); // row_shr:1
    v = dppMaxStep<0x112, 0xf>(v); // row_shr:2
    v = dppMaxStep<0x114, 0xf>(v); // row_shr:4
    v = dppMaxStep<0x118, 0xf>(v); // row_shr:8
    v = dppMaxStep<0x142, 0xa>(v); // row_bcast:15 into rows 1 and 3
    v = dppMaxStep<0x143, 0xc>(v); // row_bcast:31 into rows 2 and 3
    return bitsf(uint32_t(__builtin_amdgcn_readlane(int(fbits(v)), 63)));
}

// every lane of the wave active
template <int A>
__device__ inline WaveBundle waveBundle(const vector3& ndir, const vector3& ndu, const vector3& ndv,
                                        const vector2* offsets) {
    WaveBundle w;
    w.dc = vector3(-bitsf(__builtin_amdgcn_readfirstlane(fbits(ndir.x))),
                   -bitsf(__builtin_amdgcn_readfirstlane(fbits(ndir.y))),
                   -bitsf(__builtin_amdgcn_readfirstlane(fbits(ndir.z))));
    const float len = sqrtf(dot(w.dc, w.dc));
    w.f = vector3(w.dc.x / len, w.dc.y / len, w.dc.z / len);
    const float l1dir = l1(ndir), l1du = l1(ndu), l1dv = l1(ndv);
    float r = 0.0f, dsum = 0.0f, dmax = -FLT_MAX, dmin = FLT_MAX;
#pragma unroll
    for (int i = 0; i < A; i++) {
        const vector2 a = offsets[i];
        const vector3 D = -(ndir + ndu * a.x + ndv * a.y);
        dsum = fmaxf(dsum, l1dir + fabsf(a.x) * l1du + fabsf(a.y) * l1dv);
        r = fmaxf(r, l1(D - w.dc));
        const float df = dot(D, w.f);
        dmax = fmaxf(dmax, df);
        dmin = fminf(dmin, df);
    }
    w.dsum = waveMaxF(dsum) * (1.0f + kOccSlack);
    w.r = (waveMaxF(r) + 8.0f * kUlp * w.dsum) * (1.0f + kOccSlack);
    dmax = waveMaxF(dmax);
    dmin = -waveMaxF(-dmin);
    const float eDf = 16.0f * kUlp * w.dsum; // the dot's rounding and D's own
    w.dfMax = dmax + eDf + kOccSlack * fabsf(dmax);
    w.dfMin = dmin - eDf - kOccSlack * fabsf(dmin);
    w.dcL1 = l1(w.dc) * (1.0f + kOccSlack);
    return w;
}

// true when the triangle's depth test fails for every subsample of the wave
__device__ inline bool occludedForWave(const TriTile& c, const WaveBundle& w, float far) {
    const float nL1 = l1(c.normal) * (1.0f + kOccSlack);
    const float denC = -dot(w.dc, c.normal);
    // computed denom vs its exact value, and denC's own rounding
    const float eDen = 64.0f * kUlp * (w.dsum + w.dcL1) * nL1 + kOccSlack * fabsf(denC);
    const float spr = w.r * nL1 * (1.0f + kOccSlack);
    // plane bound
    const float hiDen = denC + spr + eDen;
    const float denMax = hiDen + kOccSlack * fabsf(hiDen);
    if (c.t > 0.0f) {
        const float lim = far * denMax;
        if (c.t > lim + kOccSlack * fabsf(lim))
            return true;
    }
    // vertex bound
    const float loDen = denC - spr - eDen;
    const float denMin = loDen - kOccSlack * fabsf(loDen);
    if (!(denMin > 0.0f) || !(w.dfMin > 0.0f))
        return false;
    const float ovL1 = l1(c.v0ToRayOrigin), e0L1 = l1(c.e0), e1L1 = l1(c.e1);
    // The test's u, v, w use the edges and its denom and t the computed normal; for a
    // triangle whose normal is not small against its edges (not a sliver) the computed
    // normal's rounding (at most 4 ulp of |e0| |e1| per component) stays within 32 ulp of
    // it, which the error terms here cover.
    if (!(8.0f * nL1 >= e0L1 * e1L1))
        return false;
    const float spread = (e0L1 + e1L1) * (1.0f + kOccSlack);
    const float g0 = -dot(c.v0ToRayOrigin, w.f);
    const float g1 = g0 + dot(c.e0, w.f), g2 = g0 + dot(c.e1, w.f);
    const float gmin = fminf(g0, fminf(g1, g2)) - 16.0f * kUlp * (ovL1 + spread) - kOccSlack * fabsf(g0);
    // rounding of the computed u, v, w (cross and dot products, the fma chains, u's subtractions)
    const float eBary = 128.0f * kUlp * w.dsum * (ovL1 * fmaxf(e0L1, e1L1) + nL1);
    const float num = gmin - (eBary / denMin) * spread * (1.0f + kOccSlack);
    if (!(num > 0.0f))
        return false;
    const float sLow = num / w.dfMax * (1.0f - kOccSlack);
    const float eT = 64.0f * kUlp * ovL1 * nL1; // rounding of t, and of the normal in it
    const float rhs = (sLow - eT / denMin) * (1.0f - kOccSlack);
    const float lhs = far * (1.0f + eDen / denMin) * (1.0f + kOccSlack);
    return rhs > 0.0f && lhs < rhs;
}

template <int A, bool DoF>
__device__ inline void intersectBody(const IntersectArgs& args, uint32_t tile) {
    using Cached = typename std::conditional<DoF, TriTileDoF, TriTile>::type;
    __shared__ Cached cache[128];
    __shared__ DofTile dofTile;
    __shared__ uint32_t cacheIndex[128];
    __shared__ float planes[4][4];
    __shared__ SimpleRayFrustum frustum;
    __shared__ uint32_t waveSurvivors[2];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint32_t listCount = args.tileCounts[tile];
    const uint32_t listStart = args.tileOffsets[tile];
    // a list that did not fit the list budget is skipped (the traversal raised the
    // overflow flag, and the frame reports an error)
    if (uint64_t(listStart) + listCount > args.listCapacity)
        listCount = 0;
    const uint32_t sample = tile * TILE_SIZE + tid;

    const matrix4x4& m = args.cameraToWorld;
    const matrix3x3 rot = m.upper3x3();
    const vector3 rayOrigin = m.m3.xyz();
    DirectionalBeam beam = args.beams[sample];
    const vector3 ndir = -(rot * beam.centerRay);
    const vector3 ndu = -(rot * beam.du);
    const vector3 ndv = -(rot * beam.dv);

    // depth of field: tile lens frame from sample 64 (IntersectTile, intersect.cu:182-189),
    // per-sample focal offset, the lens-table row of this tile
    vector3 lcfc(0.0f);
    uint32_t lensBase = 0;
    if constexpr (DoF) {
        if (tid == 64)
            dofTile = dofTileLoad(rot, beam, args.lensRadius, args.focalDistance);
        lcfc = dofLensToFocal(rot, rot * beam.centerRay, args.focalDistance);
        lensBase = (args.occRank[tile] % DOF_LENS_TABLE_ROWS) * TILE_SIZE * A + tid;
    }

    if (tid == 0) {
        SimpleRayFrustum fr = args.worldTiles[tile];
        frustum = fr;
        for (int n = 0; n < 4; n++) {
            vector3 dir = cross(fr.directions[n], fr.directions[(n + 1) % 4]);
            planes[n][0] = dir.x;
            planes[n][1] = dir.y;
            planes[n][2] = dir.z;
            planes[n][3] = dot(fr.origins[n], dir);
        }
    }

    // each wave's ray bundle for the occlusion skip, kept in LDS (read every 16 survivors)
    __shared__ WaveBundle bundles[2];
    if constexpr (!DoF) {
        const WaveBundle b = waveBundle<A>(ndir, ndu, ndv, args.offsets);
        if (lane == 0)
            bundles[wave] = b;
    }

    float depth[A];
    uint32_t hit[A];
#pragma unroll
    for (int i = 0; i < A; i++) {
        depth[i] = FLT_MAX;
        hit[i] = badTriIndex;
    }
    __syncthreads();

    uint32_t survivorsSum = 0;
    // survivors this wave skipped (wave-uniform: a scalar register)
    uint32_t skippedSum = 0;
    for (uint32_t chunk = 0; chunk < listCount; chunk += 128) {
        // ---- cull one list entry per thread ----
        bool keep = false;
        uint32_t triIndex = badTriIndex;
        Cached tt;
        if (chunk + tid < listCount) {
            triIndex = args.triIndices[listStart + chunk + tid];
            PrecomputedTriangleIntersect tri = args.tris[triIndex];
            // FrustumPlanes::test (prim_tests.h:30-49)
            bool inside = true;
            for (int n = 0; n < 4; n++) {
                vector3 dir(planes[n][0], planes[n][1], planes[n][2]);
                float d0 = dot(dir, tri.v0);
                float d1 = d0 + dot(dir, tri.edge0);
                float d2 = d0 + dot(dir, tri.edge1);
                if (d0 > planes[n][3] && d1 > planes[n][3] && d2 > planes[n][3])
                    inside = false;
            }
            if (inside) {
                if constexpr (DoF) {
                    // IntersectTriangleTile<true>::setup; the corner-ray UVW test only when
                    // every lens position sees the triangle's front (intersect.cu:211-236)
                    int r = dofTriSetup(tri, rayOrigin, dofTile.lensU, dofTile.lensV, tt);
                    keep = r == 1 || (r == 2 && frustaUVWPass(frustum, tri));
                } else {
                    // IntersectTriangleTile<false>::setup (prim_tests.h:148-163)
                    tt.e0 = tri.edge0;
                    tt.e1 = tri.edge1;
                    tt.normal = cross(tt.e0, tt.e1);
                    tt.v0ToRayOrigin = rayOrigin - tri.v0;
                    tt.t = dot(tt.v0ToRayOrigin, tt.normal);
                    if (!(tt.t < 0.0f))
                        keep = frustaUVWPass(frustum, tri);
                }
            }
        }
        // ---- ordered compaction into LDS ----
        uint64_t ballot = __ballot(keep);
        uint32_t prefix = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0)
            waveSurvivors[wave] = uint32_t(__popcll(ballot));
        __syncthreads();
        uint32_t total = waveSurvivors[0] + waveSurvivors[1];
        if (keep) {
            uint32_t slot = (wave ? waveSurvivors[0] : 0u) + prefix;
            cache[slot] = tt;
            cacheIndex[slot] = triIndex;
        }
        __syncthreads();
        survivorsSum += total;

        // ---- every sample against every survivor, in list order ----
        if constexpr (DoF) {
            for (uint32_t j = 0; j < total; j++) {
                const TriTileDoF& c = cache[j];
                const uint32_t idx = cacheIndex[j];
                // IntersectTriangleThreadDoF (prim_tests.h:295-355)
                const float denomCenter = dot(-lcfc, c.normal);
#pragma unroll
                for (int i = 0; i < A; i++) {
                    const vector2 lensUV = args.lensTable[lensBase + i * TILE_SIZE];
                    const vector2 dirUV = dofDirUV(args.offsets[i], dofTile.focalToLensScale, lensUV);
                    if (dofTest(c, denomCenter, lcfc, dofTile.lensU, dofTile.lensV, lensUV, dirUV, depth[i]))
                        hit[i] = idx;
                }
            }
        } else
        for (uint32_t j0 = 0; j0 < total; j0 += kOccBatch) {
          // the wave's farthest current hit, and the next survivors it cannot reach
          float far = depth[0];
#pragma unroll
          for (int i = 1; i < A; i++)
              far = fmaxf(far, depth[i]);
          far = waveMaxF(far);
          const uint32_t n = total - j0 < kOccBatch ? total - j0 : kOccBatch;
          uint32_t todo = (1u << n) - 1u;
          if (far < FLT_MAX) {
              const bool skip = lane < n && occludedForWave(cache[j0 + lane], bundles[wave], far);
              const uint32_t skipMask = uint32_t(__ballot(skip));
              todo &= ~skipMask;
              skippedSum += uint32_t(__builtin_popcount(skipMask));
          }
          for (; todo; todo &= todo - 1u) {
            const uint32_t j = j0 + uint32_t(__builtin_ctz(todo));
            const TriTile& c = cache[j];
            const uint32_t idx = cacheIndex[j];
            // IntersectTriangleThread ctor + GetDifferentials (prim_tests.h:118-134, 175-188)
            float denomCenter = dot(ndir, c.normal);
            vector3 eCenter = cross(ndir, c.v0ToRayOrigin);
            float vCenter = dot(c.e1, eCenter);
            float wCenter = -dot(c.e0, eCenter);
            float dD0 = dot(ndu, c.normal);
            float dD1 = dot(ndv, c.normal);
            vector3 cu = cross(ndu, c.v0ToRayOrigin);
            vector3 cv = cross(ndv, c.v0ToRayOrigin);
            float dV0 = dot(c.e1, cu), dV1 = dot(c.e1, cv);
            float dW0 = -dot(c.e0, cu), dW1 = -dot(c.e0, cv);
            const float t = c.t;
#pragma unroll
            for (int i = 0; i < A; i++) {
                // IntersectTriangleThread::test (prim_tests.h:192-231)
                const vector2 a = args.offsets[i];
                float denom = fmaf(dD1, a.y, fmaf(dD0, a.x, denomCenter));
                float depthDelta = fmaf(depth[i], denom, -t);
                float v = fmaf(dV1, a.y, fmaf(dV0, a.x, vCenter));
                float w = fmaf(dW1, a.y, fmaf(dW0, a.x, wCenter));
                float u = (denom - v) - w;
                int32_t test = int32_t(fbits(depthDelta) | fbits(u) | fbits(v) | fbits(w));
                if (test >= 0) {
                    depth[i] = t * (1.0f / denom);
                    hit[i] = idx;
                }
            }
          }
        }
        __syncthreads(); // before the next chunk overwrites the cache
    }
    if (lane == 0)
        waveSurvivors[wave] = skippedSum; // the loop's last read of it preceded a barrier
    __syncthreads();
    if (tid == 0) {
        args.tileSurvivors[tile] = survivorsSum;
        args.tileTested[tile] = 2u * survivorsSum - waveSurvivors[0] - waveSurvivors[1];
    }

    if (args.hits) {
        // per-subsample hit output: the nearest hit's ray parameter as the test left it,
        // and its barycentrics on that subsample's ray (calcUVW, prim_tests.h:233-245, or
        // its thin-lens twin :366-392), recomputed here from the winning triangle. The
        // first pass parks (depth, id) in the output so the second need not be unrolled.
        vector4* out = args.hits + size_t(sample) * A;
#pragma unroll
        for (int i = 0; i < A; i++)
            out[i] = vector4(depth[i], __uint_as_float(hit[i]), 0.0f, 0.0f);
        for (int i = 0; i < A; i++) {
            const vector4 h = out[i];
            const uint32_t id = __float_as_uint(h.y);
            float bx = 0.0f, by = 0.0f, bz = 0.0f;
            if (id != badTriIndex) {
                const PrecomputedTriangleIntersect T = args.tris[id];
                if constexpr (DoF) {
                    const vector2 lensUV = args.lensTable[lensBase + i * TILE_SIZE];
                    const vector2 dirUV = dofDirUV(args.offsets[i], dofTile.focalToLensScale, lensUV);
                    dofUVW(T, rayOrigin, dofTile, lcfc, lensUV.x, lensUV.y, dirUV.x, dirUV.y, bx, by, bz);
                } else {
                    pinholeUVW(T, rayOrigin, ndir, ndu, ndv, args.offsets[i].x, args.offsets[i].y, bx, by, bz);
                }
            }
            out[i] = vector4(h.x, bx, by, bz);
        }
    }

    GBufferSubsample* waveBase = args.gbuffer + size_t(sample >> 6) * 64 * A;
    emitGBuffer<A>(hit, waveBase, lane);
}

// Waves per SIMD the register allocator is held to. Left alone, it gives the 4x and 8x
// pinhole kernels 98 and 106 VGPRs (4 waves); held to 6 and 5 waves they take 77 and 89
// with no spills, and the 4x kernel runs 9% faster. Elsewhere (1) the natural allocation
// already avoids spills and a tighter one would not.
constexpr int intersectWaves(int A, bool DoF) { return DoF ? 1 : A == 4 ? 6 : A == 8 ? 5 : 1; }

template <int A, bool DoF>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(intersectWaves(A, DoF))))
void intersectKernel(IntersectArgs args) {
    intersectBody<A, DoF>(args, blockIdx.x);
}

// every tile of the views whose units `start` lists (pinhole views of one MSAA rate)
template <int A>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(intersectWaves(A, false))))
void intersectBatchKernel(const IntersectArgs* __restrict__ all,
                                                            const uint32_t* __restrict__ start, uint32_t n,
                                                            uint32_t swizzle) {
    const uint32_t unit = xcdUnit(blockIdx.x, gridDim.x, swizzle);
    const uint32_t v = batchView(start, n, unit);
    intersectBody<A, false>(batchArgs(all, v), unit - start[v]);
}

IntersectArgs intersectArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m) {
    IntersectArgs a;
    a.gbuffer = v->gbuffer.ptr;
    a.beams = v->beams.ptr;
    a.tileCounts = v->tileCounts.ptr;
    a.tileOffsets = v->tileOffsets.ptr;
    a.triIndices = v->triIndices.ptr;
    a.worldTiles = v->tileFrustaWorld.ptr;
    a.tris = s->triIntersect.ptr;
    a.tileSurvivors = v->tileWork.ptr + size_t(2) * v->tileCount;
    a.tileTested = v->tileWork.ptr + size_t(3) * v->tileCount;
    a.hits = v->hitOutput ? v->hits.ptr : nullptr;
    a.listCapacity = v->listCapacity;
    a.cameraToWorld = m;
    for (int i = 0; i < 32; i++)
        a.offsets[i] = v->offsets[i];
    a.lensTable = v->lensTable.ptr;
    a.occRank = v->occRank.ptr;
    a.lensRadius = v->lens.radius;
    a.focalDistance = v->lens.focalDistance;
    return a;
}

void launchIntersect(hvvr_gpu_view* v, const hvvr_gpu_scene* s, const matrix4x4& m, hipStream_t stream) {
    const IntersectArgs a = intersectArgs(v, s, m);
    dim3 grid(v->tileCount), block(128);
    const bool dof = v->dof;
    switch (v->msaaRate) {
#define HVVR_INTERSECT_CASE(A)                                                        \
    case A:                                                                           \
        if (dof)                                                                      \
            intersectKernel<A, true><<<grid, block, 0, stream>>>(a);                  \
        else                                                                          \
            intersectKernel<A, false><<<grid, block, 0, stream>>>(a);                 \
        break;
        HVVR_INTERSECT_CASE(1)
        HVVR_INTERSECT_CASE(2)
        HVVR_INTERSECT_CASE(4)
        HVVR_INTERSECT_CASE(8)
        HVVR_INTERSECT_CASE(16)
        HVVR_INTERSECT_CASE(32)
#undef HVVR_INTERSECT_CASE
        default: break;
    }
}

void launchIntersectBatch(int msaaRate, const IntersectArgs* all, const uint32_t* start, uint32_t viewCount,
                          uint32_t units, const LaunchShape& shape, hipStream_t stream) {
    switch (msaaRate) {
#define HVVR_INTERSECT_BATCH_CASE(A)                                                                          \
    case A:                                                                                                   \
        intersectBatchKernel<A><<<units, 128, 0, stream>>>(all, start, viewCount, shape.xcdSwizzle);          \
        break;
        HVVR_INTERSECT_BATCH_CASE(1)
        HVVR_INTERSECT_BATCH_CASE(2)
        HVVR_INTERSECT_BATCH_CASE(4)
        HVVR_INTERSECT_BATCH_CASE(8)
        HVVR_INTERSECT_BATCH_CASE(16)
        HVVR_INTERSECT_BATCH_CASE(32)
#undef HVVR_INTERSECT_BATCH_CASE
        default: break;
    }
}

} // namespace hvvr
