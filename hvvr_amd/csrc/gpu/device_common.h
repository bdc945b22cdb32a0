#pragma once
// Device-side helpers shared by the gfx950 kernels: frustum construction and the
// child-AABB tests used by block and tile traversal, plus small utilities.
//
// The arithmetic restates traverse::avx::Frustum (raycaster/traverse_avx.cpp:22-211)
// and testBVHNodeChildren/Refine (raycaster/traverse_avx.h:104-217) lane by lane: fmaf
// exactly where the reference uses _mm256_fmadd_ps, plain float ops elsewhere (the
// whole library builds with -ffp-contract=off).

#include "hvvr/types.h"

#include <float.h>

namespace hvvr {

__device__ inline uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ inline float bitsf(uint32_t u) { return __uint_as_float(u); }
// _mm256_min_ps / _mm256_max_ps: first operand only when it strictly wins
__device__ inline float minps(float a, float b) { return a < b ? a : b; }
__device__ inline float maxps(float a, float b) { return a > b ? a : b; }
// order-preserving integer image of a float (raycaster/sort.h:14-32)
__device__ inline uint32_t floatFlip(uint32_t f) {
    uint32_t mask = uint32_t(-int32_t(f >> 31)) | 0x80000000u;
    return f ^ mask;
}
__device__ inline uint32_t floatUnflip(uint32_t f) {
    uint32_t mask = ((f & 0x80000000u) == 0) ? 0xffffffffu : 0x80000000u;
    return f ^ mask;
}
// max in the FloatFlip order (total order, -0 < +0)
__device__ inline float maxTotal(float a, float b) {
    return floatFlip(fbits(a)) >= floatFlip(fbits(b)) ? a : b;
}
// linearToSRgb (shared/util_graphics.h:26-28)
__device__ inline float linearToSRgb(float x) {
    return (x <= 0.0031308f) ? (x * 12.92f) : (powf(x, 1.0f / 2.4f) * 1.055f - 0.055f);
}

// Frustum precompute from world-space corner rays (origins o, directions d).
__device__ inline void frustumPrecompute(const vector3 o[4], const vector3 d[4], FrustumPre& f) {
    int degenerate = isinf(o[0].x) ? 1 : 0;
    float plane[4][4];
    vector3 avg(0.0f, 0.0f, 0.0f);
    for (int p = 0; p < 4; p++) {
        if (degenerate) {
            plane[p][0] = plane[p][1] = plane[p][2] = 0.0f;
            plane[p][3] = -INFINITY;
        } else {
            vector3 n = cross(d[p], d[(p + 1) % 4]);
            plane[p][0] = n.x;
            plane[p][1] = n.y;
            plane[p][2] = n.z;
            plane[p][3] = dot(o[p], n);
        }
        avg += d[p];
    }
    f.degenerate = degenerate;
    if (fabsf(avg.x) > fabsf(avg.y)) {
        if (fabsf(avg.x) > fabsf(avg.z))
            f.distanceAxis = avg.x < 0 ? 0 : 1;
        else
            f.distanceAxis = avg.z < 0 ? 4 : 5;
    } else {
        if (fabsf(avg.y) > fabsf(avg.z))
            f.distanceAxis = avg.y < 0 ? 2 : 3;
        else
            f.distanceAxis = avg.z < 0 ? 4 : 5;
    }
    for (int p = 0; p < 4; p++) {
        for (int a = 0; a < 3; a++) {
            float c = plane[p][a];
            f.planeNeg[p][a] = c < 0.0f ? c : 0.0f;
            f.planePos[p][a] = c > 0.0f ? -c : 0.0f;
        }
        f.planeNeg[p][3] = plane[p][3];
        f.planePos[p][3] = 0.0f;
    }
    for (int a = 0; a < 3; a++) {
        float lo = 0.0f, hi = 0.0f;
        for (int v = 0; v < 4; v++) {
            float oa = o[v][a], da = d[v][a];
            float l = -oa + FLT_MAX * maxps(0.0f, -da);
            float h = oa + FLT_MAX * maxps(0.0f, da);
            lo = v == 0 ? l : maxps(lo, l);
            hi = v == 0 ? h : maxps(hi, h);
        }
        f.projLo[a] = -lo;
        f.projHi[a] = -hi;
    }
    const vector3 axes[3] = {vector3(1, 0, 0), vector3(0, 1, 0), vector3(0, 0, 1)};
    for (int v = 0; v < 4; v++)
        for (int a = 0; a < 3; a++) {
            vector3 e = cross(d[v], axes[a]);
            f.edge[v * 3 + a][0] = e.x;
            f.edge[v * 3 + a][1] = e.y;
            f.edge[v * 3 + a][2] = e.z;
        }
    const float eps = 0.000001f;
    for (int e = 0; e < 12; e++) {
        float ex = f.edge[e][0], ey = f.edge[e][1], ez = f.edge[e][2];
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int v = 0; v < 4; v++) {
            float d0 = fmaf(ex, o[v].x, fmaf(ey, o[v].y, ez * o[v].z));
            mn = minps(mn, d0);
            mx = maxps(mx, d0);
            float d1 = fmaf(ex, d[v].x, fmaf(ey, d[v].y, ez * d[v].z));
            mn = minps(d1 + eps, 0.0f) * FLT_MAX + mn;
            mx = maxps(d1 - eps, 0.0f) * FLT_MAX + mx;
        }
        f.edgeMin[e] = mn;
        f.edgeMax[e] = mx;
    }
}

// Coarse child test: 4 frustum planes (n-vertex) and the frustum's axis projections.
__device__ inline uint32_t testChildren(const FrustumPre& f, const BVHNode& n) {
    uint32_t planeMask = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        bool pass = true;
#pragma unroll
        for (int p = 0; p < 4; p++) {
            float dmax = fmaf(f.planeNeg[p][0], n.bounds[0][c],
                              fmaf(f.planeNeg[p][1], n.bounds[2][c], f.planeNeg[p][2] * n.bounds[4][c]));
            float dneg = fmaf(f.planePos[p][0], n.bounds[1][c],
                              fmaf(f.planePos[p][1], n.bounds[3][c], f.planePos[p][2] * n.bounds[5][c]));
            float dist = dmax + dneg;
            pass = pass && (dist <= f.planeNeg[p][3]);
        }
        planeMask |= pass ? (1u << c) : 0u;
    }
    if (!planeMask)
        return 0;
    uint32_t projMask = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        bool pass = true;
#pragma unroll
        for (int a = 0; a < 3; a++)
            pass = pass && (n.bounds[2 * a][c] >= f.projLo[a]) && (n.bounds[2 * a + 1][c] >= f.projHi[a]);
        projMask |= pass ? (1u << c) : 0u;
    }
    return planeMask & projMask;
}

// Refinement: the 12 frustum-edge x world-axis separating axes.
__device__ inline uint32_t testChildrenRefine(const FrustumPre& f, const BVHNode& n, uint32_t prev) {
    uint32_t result = prev;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        bool cull = false;
        for (int e = 0; e < 12; e++) {
            float ex = f.edge[e][0], ey = f.edge[e][1], ez = f.edge[e][2];
            float px = ex > 0.0f ? n.bounds[0][c] : -n.bounds[1][c];
            float py = ey > 0.0f ? n.bounds[2][c] : -n.bounds[3][c];
            float pz = ez > 0.0f ? n.bounds[4][c] : -n.bounds[5][c];
            float nx = ex > 0.0f ? -n.bounds[1][c] : n.bounds[0][c];
            float ny = ey > 0.0f ? -n.bounds[3][c] : n.bounds[2][c];
            float nz = ez > 0.0f ? -n.bounds[5][c] : n.bounds[4][c];
            float maxBox = fmaf(ex, px, fmaf(ey, py, ez * pz));
            float minBox = fmaf(ex, nx, fmaf(ey, ny, ez * nz));
            cull = cull || (maxBox < f.edgeMin[e]) || (f.edgeMax[e] < minBox);
        }
        if (cull)
            result &= ~(1u << c);
    }
    return result;
}

} // namespace hvvr
