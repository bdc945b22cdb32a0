#pragma once
// Cooperative tile traversal: one wave64 walks one tile's frustum through the BVH.
//
// Same algorithm, same emission order as traverseTiles (raycaster/traverse_avx.h:
// 412-527), but laid out for a 64-wide wavefront instead of one AVX2 core:
//   * lane = (child c = lane & 3, role r = lane >> 2). A node's coarse test (4 frustum
//     planes x 4 children in roles 0..3, 3 axis projections x 4 children in roles 4..6)
//     and its refinement (12 edge axes x 4 children in roles 0..11) each take one pass,
//     and a 64-bit ballot folds the per-lane verdicts into the 4-bit child mask;
//   * every lane keeps the frustum coefficients of its role in registers for the whole
//     tile, and reads the six bounds of its child from the node (the node index is
//     wave-uniform, so leafMask/child offsets are scalar loads);
//   * the distance-sorted stack lives in LDS; an insertion compares up to 64 entries in
//     one ballot and shifts them with one LDS read + one LDS write per lane;
//   * all control flow is wave-uniform, so the wave never diverges.
// Stacks deeper than the LDS capacity set a per-tile overflow flag; the caller
// re-traverses those tiles with the global-memory thread-per-tile kernel.

#include "device_common.h"

namespace hvvr {

enum : uint32_t { WAVE_STACK_SIZE = 512 };

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

// fold a ballot of per-(role, child) "cull" bits into a 4-bit mask of culled children
__device__ inline uint32_t foldChildren(uint64_t b) {
    uint64_t x = b | (b >> 32);
    x |= x >> 16;
    x |= x >> 8;
    x |= x >> 4;
    return uint32_t(x) & 0xfu;
}

struct NodeLanes {
    float b[6]; // bounds[k][child] for k = 0..5
};

__device__ inline void loadNodeLanes(const BVHNode* nodes, uint32_t node, uint32_t c, NodeLanes& nl) {
    const float* base = &nodes[node].bounds[0][0];
#pragma unroll
    for (int k = 0; k < 6; k++)
        nl.b[k] = base[k * 4 + c];
}

// testBVHNodeChildren (traverse_avx.h:104-164), one pass over 28 lanes
__device__ inline uint32_t waveTestChildren(const WaveFrustum& w, const NodeLanes& n, uint32_t role) {
    bool cull = false;
    if (role < 4) {
        float dmax = fmaf(w.pn0, n.b[0], fmaf(w.pn1, n.b[2], w.pn2 * n.b[4]));
        float dneg = fmaf(w.pp0, n.b[1], fmaf(w.pp1, n.b[3], w.pp2 * n.b[5]));
        cull = !((dmax + dneg) <= w.pw);
    } else if (role < 7) {
        uint32_t a = role - 4;
        float mx = a == 0 ? n.b[0] : (a == 1 ? n.b[2] : n.b[4]);
        float ng = a == 0 ? n.b[1] : (a == 1 ? n.b[3] : n.b[5]);
        cull = !(mx >= w.projLo && ng >= w.projHi);
    }
    return 0xfu & ~foldChildren(__ballot(cull));
}

// testBVHNodeChildrenRefine (traverse_avx.h:166-217), one pass over 48 lanes
__device__ inline uint32_t waveTestRefine(const WaveFrustum& w, const NodeLanes& n, uint32_t role, uint32_t prev) {
    bool cull = false;
    if (role < 12) {
        float px = w.ex > 0.0f ? n.b[0] : -n.b[1];
        float py = w.ey > 0.0f ? n.b[2] : -n.b[3];
        float pz = w.ez > 0.0f ? n.b[4] : -n.b[5];
        float nx = w.ex > 0.0f ? -n.b[1] : n.b[0];
        float ny = w.ey > 0.0f ? -n.b[3] : n.b[2];
        float nz = w.ez > 0.0f ? -n.b[5] : n.b[4];
        float maxBox = fmaf(w.ex, px, fmaf(w.ey, py, w.ez * pz));
        float minBox = fmaf(w.ex, nx, fmaf(w.ey, ny, w.ez * nz));
        cull = (maxBox < w.eMin) || (w.eMax < minBox);
    }
    return prev & ~foldChildren(__ballot(cull));
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

// Walks one tile. `emit(t0, t1)` is called (wave-uniformly) for every leaf triangle range
// in emission order. Returns false if the stack would exceed `stackLimit` entries
// (results then incomplete). `stack` is the caller's LDS array of at least `stackLimit`
// (<= WAVE_STACK_SIZE) entries.
template <class Emit>
__device__ inline bool waveTraverseTile(const BVHNode* __restrict__ nodes, const FrustumPre& f,
                                        const StackEntry* __restrict__ blockStack, uint32_t blockStackSize,
                                        StackEntry* stack, uint32_t stackLimit, uint32_t lane, Emit&& emit,
                                        uint32_t* maxDepth) {
    const uint32_t c = lane & 3u, role = lane >> 2;
    WaveFrustum w;
    loadWaveFrustum(f, role, w);
    if (w.degenerate || blockStackSize == 0)
        return true;
    if (blockStackSize > stackLimit)
        return false;
    if (lane < blockStackSize)
        stack[lane] = blockStack[lane];
    waveLdsSync();
    uint32_t top = blockStackSize - 1;
    uint32_t deepest = top;
    StackEntry e = stack[top];
    uint32_t node = uniform(e.nodeMask & STACK_NODE_MASK);
    uint32_t entryMask = uniform(e.nodeMask >> 28);
    for (;;) {
        NodeLanes nl;
        loadNodeLanes(nodes, node, c, nl);
        const uint32_t leafMask = nodes[node].leafMask;
        uint32_t mask = waveTestChildren(w, nl, role);
        bool pop = true;
        if (mask) {
            uint32_t leaf = mask & leafMask;
            if (leaf) {
                mask = waveTestRefine(w, nl, role, mask);
                leaf = mask & leafMask;
                for (int k = 0; k < 4; k++) {
                    if (leaf & (1u << k)) {
                        uint32_t t0 = nodes[node].data[k], t1 = nodes[node].data[k + 1];
                        if (t1 > t0)
                            emit(t0, t1);
                    }
                }
            }
            mask &= ~leafMask;
            if (mask && entryMask == 0) {
                if ((mask & (mask - 1)) == 0) {
                    // single hit child: test it next without touching the stack
                    node = node + nodes[node].data[1 + __builtin_ctz(mask)];
                    continue;
                }
                bool first = true;
                for (int k = 0; k < 4; k++) {
                    if (!(mask & (1u << k)))
                        continue;
                    if (!first)
                        top++;
                    first = false;
                    if (top >= stackLimit)
                        return false;
                    const float dist = -nodes[node].bounds[w.axis][k];
                    // entries above the insertion point are the contiguous run from the top
                    // whose distance is smaller than `dist`; shift them up by one
                    uint32_t pos = top;
                    for (;;) {
                        int j = int(pos) - 1 - int(lane);
                        bool smaller = j >= 0 && stack[j >= 0 ? j : 0].tMin < dist;
                        uint64_t run = ~__ballot(smaller);
                        uint32_t n = run ? uint32_t(__builtin_ctzll(run)) : 64u;
                        StackEntry moved;
                        if (lane < n)
                            moved = stack[j];
                        waveLdsSync();
                        if (lane < n)
                            stack[j + 1] = moved;
                        waveLdsSync();
                        pos -= n;
                        if (n < 64)
                            break;
                    }
                    if (lane == 0) {
                        StackEntry ne;
                        ne.tMin = dist;
                        ne.nodeMask = (node + nodes[node].data[1 + k]) & STACK_NODE_MASK;
                        stack[pos] = ne;
                    }
                    waveLdsSync();
                }
                deepest = top > deepest ? top : deepest;
                e = stack[top];
                node = uniform(e.nodeMask & STACK_NODE_MASK);
                entryMask = uniform(e.nodeMask >> 28);
                pop = false;
            }
        }
        if (pop) {
            if (top == 0)
                break;
            top--;
            e = stack[top];
            node = uniform(e.nodeMask & STACK_NODE_MASK);
            entryMask = uniform(e.nodeMask >> 28);
        }
    }
    if (maxDepth)
        *maxDepth = deepest + 1;
    return true;
}

} // namespace hvvr
