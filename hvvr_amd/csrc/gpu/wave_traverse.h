#pragma once
// Wave64 building blocks shared by the traversal kernels.
//
// For a node test with one wave per frustum (block traversal): lane = (child c =
// lane & 3, role r = lane >> 2). A node's coarse test (4 frustum planes x 4 children in
// roles 0..3, 3 axis projections x 4 children in roles 4..6) and its refinement (12
// edge axes x 4 children in roles 0..11) each take one pass, and a 64-bit ballot folds
// the per-lane verdicts into the 4-bit child mask (testBVHNodeChildren/Refine,
// raycaster/traverse_avx.h:104-217). Every lane keeps the frustum coefficients of its
// role in registers, and each lane's node operands are loaded straight from the node,
// with no per-lane indexing into private arrays (which the compiler would place in
// scratch memory).
//
// For per-child tests (tile lists): one lane evaluates every plane, projection and
// edge axis of one child, so a wave tests 16 nodes at once.

#include "device_common.h"

namespace hvvr {

struct WaveFrustum {
    // coefficients of this lane's role
    float pn0, pn1, pn2, pw; // plane: planeNeg xyz, w
    float pp0, pp1, pp2;     // plane: planePos xyz
    float projLo, projHi;    // projection axis
    float ex, ey, ez, eMin, eMax;
    int axis;
    int degenerate;
};

__device__ inline void loadWaveFrustum(const FrustumPre& f, uint32_t role, WaveFrustum& w) {
    uint32_t p = role < 4 ? role : 0;
    w.pn0 = f.planeNeg[p][0];
    w.pn1 = f.planeNeg[p][1];
    w.pn2 = f.planeNeg[p][2];
    w.pw = f.planeNeg[p][3];
    w.pp0 = f.planePos[p][0];
    w.pp1 = f.planePos[p][1];
    w.pp2 = f.planePos[p][2];
    uint32_t a = (role >= 4 && role < 7) ? role - 4 : 0;
    w.projLo = f.projLo[a];
    w.projHi = f.projHi[a];
    uint32_t e = role < 12 ? role : 0;
    w.ex = f.edge[e][0];
    w.ey = f.edge[e][1];
    w.ez = f.edge[e][2];
    w.eMin = f.edgeMin[e];
    w.eMax = f.edgeMax[e];
    w.axis = f.distanceAxis;
    w.degenerate = f.degenerate;
}

// lanes where `b` holds (HIP's __ballot takes an int, which costs a select and a compare)
__device__ inline uint64_t ballot(bool b) {
    return __builtin_amdgcn_ballot_w64(b);
}

// fold a ballot of per-(role, child) "cull" bits into a 4-bit mask of culled children
__device__ inline uint32_t foldChildren(uint64_t b) {
    uint64_t x = b | (b >> 32);
    x |= x >> 16;
    x |= x >> 8;
    x |= x >> 4;
    return uint32_t(x) & 0xfu;
}

struct NodeLanes {
    float b0, b1, b2, b3, b4, b5; // bounds[k][child]
    float projMax, projNegMin;    // bounds[2a][child], bounds[2a+1][child] for this lane's projection axis a
};

// `projAxis` is the lane's projection axis (0 for lanes without one)
__device__ inline void loadNodeLanes(const BVHNode* nodes, uint32_t node, uint32_t c, uint32_t projAxis,
                                     NodeLanes& nl) {
    const float* base = &nodes[node].bounds[0][0];
    nl.b0 = base[0 * 4 + c];
    nl.b1 = base[1 * 4 + c];
    nl.b2 = base[2 * 4 + c];
    nl.b3 = base[3 * 4 + c];
    nl.b4 = base[4 * 4 + c];
    nl.b5 = base[5 * 4 + c];
    nl.projMax = base[(2 * projAxis) * 4 + c];
    nl.projNegMin = base[(2 * projAxis + 1) * 4 + c];
    // held in every lane here: otherwise the compiler moves these two loads into the
    // projection lanes' branch of the test, where they cost a memory round trip of their own
    asm volatile("" : "+v"(nl.projMax), "+v"(nl.projNegMin));
}

__device__ inline uint32_t laneProjAxis(uint32_t role) {
    return (role >= 4 && role < 7) ? role - 4 : 0u;
}

// testBVHNodeChildren (traverse_avx.h:104-164), one pass over 28 lanes. Every lane
// evaluates both tests and keeps its role's verdict: with selects instead of branches the
// projection operands load with the others, not behind a branch of their own.
__device__ inline uint32_t waveTestChildren(const WaveFrustum& w, const NodeLanes& n, uint32_t role) {
    const float dmax = fmaf(w.pn0, n.b0, fmaf(w.pn1, n.b2, w.pn2 * n.b4));
    const float dneg = fmaf(w.pp0, n.b1, fmaf(w.pp1, n.b3, w.pp2 * n.b5));
    const bool planeCull = !((dmax + dneg) <= w.pw);
    const bool projCull = !((n.projMax >= w.projLo) & (n.projNegMin >= w.projHi));
    const bool cull = role < 4 ? planeCull : (role < 7 ? projCull : false);
    return 0xfu & ~foldChildren(ballot(cull));
}

// testBVHNodeChildrenRefine (traverse_avx.h:166-217), one pass over 48 lanes
__device__ inline uint32_t waveTestRefine(const WaveFrustum& w, const NodeLanes& n, uint32_t role, uint32_t prev) {
    const float px = w.ex > 0.0f ? n.b0 : -n.b1;
    const float py = w.ey > 0.0f ? n.b2 : -n.b3;
    const float pz = w.ez > 0.0f ? n.b4 : -n.b5;
    const float nx = w.ex > 0.0f ? -n.b1 : n.b0;
    const float ny = w.ey > 0.0f ? -n.b3 : n.b2;
    const float nz = w.ez > 0.0f ? -n.b5 : n.b4;
    const float maxBox = fmaf(w.ex, px, fmaf(w.ey, py, w.ez * pz));
    const float minBox = fmaf(w.ex, nx, fmaf(w.ey, ny, w.ez * nz));
    const bool cull = (role < 12) & ((maxBox < w.eMin) | (w.eMax < minBox));
    return prev & ~foldChildren(ballot(cull));
}

__device__ inline uint32_t uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Orders this wave's LDS accesses: every lane's earlier LDS reads/writes complete before
// any later one (the compiler must not move LDS accesses of different addresses across).
__device__ inline void waveLdsSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---------------------------------------------------------------------------------
// per-child tests: one lane, one child, every axis of the frustum
// ---------------------------------------------------------------------------------
// Coarse test of one child box (testBVHNodeChildren, traverse_avx.h:104-164, for one
// child lane): the four frustum planes against the box's n-vertex, then the three
// world-axis projections. Same operations and rounding as testChildren(). Every test is
// evaluated and the verdicts are combined with `&`: a short-circuit chain compiles to
// one branch per test, each reloading its frustum coefficients with a scalar load and a
// wait, where this way the coefficients load in batches.
__device__ inline bool childCoarsePass(const FrustumPre* __restrict__ f, float b0, float b1, float b2, float b3,
                                       float b4, float b5) {
    bool pass = true;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        float dmax = fmaf(f->planeNeg[p][0], b0, fmaf(f->planeNeg[p][1], b2, f->planeNeg[p][2] * b4));
        float dneg = fmaf(f->planePos[p][0], b1, fmaf(f->planePos[p][1], b3, f->planePos[p][2] * b5));
        pass &= (dmax + dneg) <= f->planeNeg[p][3];
    }
    pass &= (b0 >= f->projLo[0]) & (b1 >= f->projHi[0]);
    pass &= (b2 >= f->projLo[1]) & (b3 >= f->projHi[1]);
    pass &= (b4 >= f->projLo[2]) & (b5 >= f->projHi[2]);
    return pass;
}

// Refinement of one child box (testBVHNodeChildrenRefine, traverse_avx.h:166-217, for
// one child lane): true if one of the 12 frustum-edge x world-axis separating axes
// separates the box from the frustum. Same operations as testChildrenRefine(), every
// axis evaluated (see childCoarsePass).
__device__ inline bool childRefineCull(const FrustumPre* __restrict__ f, float b0, float b1, float b2, float b3,
                                       float b4, float b5) {
    bool cull = false;
#pragma unroll
    for (int e = 0; e < 12; e++) {
        float ex = f->edge[e][0], ey = f->edge[e][1], ez = f->edge[e][2];
        float px = ex > 0.0f ? b0 : -b1;
        float py = ey > 0.0f ? b2 : -b3;
        float pz = ez > 0.0f ? b4 : -b5;
        float nx = ex > 0.0f ? -b1 : b0;
        float ny = ey > 0.0f ? -b3 : b2;
        float nz = ez > 0.0f ? -b5 : b4;
        float maxBox = fmaf(ex, px, fmaf(ey, py, ez * pz));
        float minBox = fmaf(ex, nx, fmaf(ey, ny, ez * nz));
        cull |= (maxBox < f->edgeMin[e]) | (f->edgeMax[e] < minBox);
    }
    return cull;
}

// ---------------------------------------------------------------------------------
// wave-wide integer helpers
// ---------------------------------------------------------------------------------
// number of set bits of `b` in lanes below this one
__device__ inline uint32_t lanesBelow(uint64_t b) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(b >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(b), 0u));
}

// Inclusive prefix sum over the wave; every lane must be active. Within each row of 16
// lanes, four DPP row shifts (by 1, 2, 4 and 8 lanes; a lane with no source lane adds the
// 0 it keeps); then row_bcast:15 adds lane 15's sum (row 0) to row 1 and lane 47's (row 2)
// to row 3, and row_bcast:31 adds lane 31's (rows 0 and 1) to rows 2 and 3. No LDS
// round trip, unlike a shuffle (ds_bpermute) per step.
__device__ inline uint32_t waveInclusiveScan(uint32_t v) {
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x111, 0xf, 0xf, false)); // row_shr:1
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x112, 0xf, 0xf, false)); // row_shr:2
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xf, false)); // row_shr:4
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xf, false)); // row_shr:8
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false)); // row_bcast:15
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false)); // row_bcast:31
    return v;
}

// sum over the wave, in every lane; every lane must be active
__device__ inline uint32_t waveSum(uint32_t v) {
    return uint32_t(__builtin_amdgcn_readlane(int(waveInclusiveScan(v)), 63));
}

} // namespace hvvr
