// Per-frame frustum setup and frustum-BVH traversal on the GPU. The reference runs
// both on the CPU with AVX2 (raycaster/render.cpp:282-365, raycaster/traversal.cpp:
// 99-330, raycaster/traverse_avx.h:293-527); here they are kernels whose output, the
// per-tile triangle lists, never leaves HBM.
//
//   frustumSetupKernel        one thread per tile or block frustum: camera->world
//                             transform (traversal.h:38-47) + precompute (traverse_avx.cpp)
//   blockTraverseWaveKernel   one wave64 per block: cooperative node tests, the <= 64-entry
//                             cull stack in registers, then the distance sort of
//                             BlockFrame::sort as a parallel rank computation
//   tileListKernel            one wave64 per tile, in three phases (see below); writes the
//                             tile's list at an offset claimed with one atomic add, so no
//                             count pass and no scan are needed
//   tileListQueueKernel<Cfg>  the same for the tiles a smaller configuration queued (large,
//                             huge, giant, colossal: more records and deeper stacks)
//   tileWalkKernel            one wave64 per tile with a 4096-entry stack in LDS, for the
//                             tiles no configuration holds (count pass, allocation, emit
//                             pass)
#include <algorithm>

#include <hip/hip_ext.h>

#include "device_common.h"
#include "gpu_internal.h"
#include "launch_args.h"
#include "reg_stack.h"
#include "wave_traverse.h"

namespace hvvr {

// planeCullsFrustum (render.cpp:271-279): every corner ray, at its origin and 10^4 along
// it, lies on the positive side of the plane
__device__ inline bool planeCullsFrustum(const vector4& p, const vector3 o[4], const vector3 d[4]) {
    const vector3 n(p.x, p.y, p.z);
    bool allOut = true;
    for (int v = 0; v < 4; v++) {
        allOut = allOut && dot(n, o[v]) > p.w;
        allOut = allOut && dot(n, o[v] + d[v] * 10000.0f) > p.w;
    }
    return allOut;
}

__device__ inline void frustumSetupBody(const FrustumSetupArgs& a, uint32_t i) {
    const uint32_t tileCount = a.tileCount;
    if (i >= tileCount + a.blockCount)
        return;
    const matrix4x4& m = a.cameraToWorld;
    const CullPlanes& cull = a.cull;
    const SimpleRayFrustum* __restrict__ tileCam = a.tileCam;
    const SimpleRayFrustum* __restrict__ blockCam = a.blockCam;
    bool isTile = i < tileCount;
    const SimpleRayFrustum& src = isTile ? tileCam[i] : blockCam[i - tileCount];
    vector3 o[4], d[4];
    matrix3x3 r = m.upper3x3();
    // an infinite first origin marks a culled (degenerate) frustum (render.cpp:331-338)
    bool culled = isinf(src.origins[0].x);
    for (int v = 0; v < 4; v++) {
        vector4 t = m * vector4(src.origins[v], 1.0f);
        o[v] = culled ? vector3(INFINITY, INFINITY, INFINITY) : vector3(t.x, t.y, t.z);
        d[v] = culled ? vector3(0.0f, 0.0f, 0.0f) : r * src.directions[v];
    }
    // foveated path: tiles wholly outside the screen rectangle are culled the same way
    // (blockTransformAndCullTask, render.cpp:311-340)
    if (isTile && cull.enabled && !culled) {
        for (int k = 0; k < 4; k++)
            culled = culled || planeCullsFrustum(cull.planes[k], o, d);
        if (culled)
            for (int v = 0; v < 4; v++) {
                o[v] = vector3(INFINITY, INFINITY, INFINITY);
                d[v] = vector3(0.0f, 0.0f, 0.0f);
            }
    }
    FrustumPre f;
    frustumPrecompute(o, d, f);
    if (isTile) {
        SimpleRayFrustum w;
        for (int v = 0; v < 4; v++) {
            w.origins[v] = o[v];
            w.directions[v] = d[v];
        }
        a.tileWorld[i] = w;
        a.tilePre[i] = f;
    } else {
        a.blockPre[i - tileCount] = f;
    }
}

__global__ __launch_bounds__(128) void frustumSetupKernel(FrustumSetupArgs a) {
    if (blockIdx.x == 0 && threadIdx.x < a.counterCount)
        a.counters[threadIdx.x] = 0ull;
    frustumSetupBody(a, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(128) void frustumSetupBatchKernel(const FrustumSetupArgs* __restrict__ all,
                                                               const uint32_t* __restrict__ start, uint32_t n) {
    const uint32_t v = batchView(start, n, blockIdx.x);
    frustumSetupBody(batchArgs(all, v), (blockIdx.x - start[v]) * 128u + threadIdx.x);
}

struct CullEntry {
    uint32_t bits; // leaf mask (<= 0xf) or the float bits of a size estimate
    float tMinTemp;
    uint32_t node;
};

// DAZ view of raw float bits: the reference runs the traversal with DAZ set, under which
// a leaf entry's mask (stored in the float slot) compares as zero
__device__ inline float daz(uint32_t bits) {
    return (bits & 0x7f800000u) == 0 ? 0.0f : bitsf(bits);
}
__device__ inline float ftz(float x) {
    return (fbits(x) & 0x7f800000u) == 0 ? bitsf(fbits(x) & 0x80000000u) : x;
}

// traverseBlocks + BlockFrame::sort, one wave64 per block
__device__ inline void blockTraverseBody(const BlockTraverseArgs& a, CullEntry* cull, uint32_t b) {
    const BVHNode* __restrict__ nodes = a.nodes;
    const FrustumPre* __restrict__ blockPre = a.blockPre;
    StackEntry* __restrict__ stacks = a.stacks;
    uint32_t* __restrict__ stackSize = a.stackSize;
    const uint32_t lane = threadIdx.x, c = lane & 3u, role = lane >> 2;
    const uint32_t projAxis = laneProjAxis(role);
    WaveFrustum w;
    loadWaveFrustum(blockPre[b], role, w);
    if (w.degenerate) {
        if (lane == 0)
            stackSize[b] = 0;
        return;
    }
    const int axis = w.axis, sizeAxis = w.axis & ~1;
    // The stack in registers, bottom-up: entry q in lane q of (sBits, sT, sNode). The walk
    // stops at 60 entries and a step inserts at most 4 (a leaf entry and 3 children, or 4
    // children), so it never holds more than 62.
    uint32_t top = 0;
    uint32_t sBits = 0u, sNode = 0u;
    float sT = 0.0f;
    float tMin = 0.0f;
    uint32_t node = 0;
    // Inserts (bits, t, node) below the run of top entries whose (DAZ) key exceeds `key`:
    // one above the highest entry whose key does not, position 0 when there is none. One
    // ballot finds the position, two DPP shifts per register move the entries above it.
    auto insert = [&](float key, uint32_t eBits, float eT, uint32_t eNode) {
        const uint64_t notLarger = ballot(!(daz(sBits) > key)) & lanesBelowMask(top);
        const uint32_t p = notLarger ? 64u - uint32_t(__builtin_clzll(notLarger)) : 0u;
        const uint32_t ub = laneShiftUp(sBits, 0u), un = laneShiftUp(sNode, 0u);
        const float ut = bitsf(laneShiftUp(fbits(sT), 0u));
        sBits = lane < p ? sBits : (lane == p ? eBits : ub);
        sNode = lane < p ? sNode : (lane == p ? eNode : un);
        sT = lane < p ? sT : (lane == p ? eT : ut);
        top++;
    };
    for (;;) {
        // Child c's size estimate, distance and offset, loaded with the test operands: the
        // inserts take child k's from lane k, with no memory round trip each.
        const BVHNode* __restrict__ nd = nodes + node;
        const float sizeA = nd->bounds[sizeAxis][c], sizeB = nd->bounds[sizeAxis + 1][c];
        const float distC = -nd->bounds[axis][c];
        const uint32_t offC = nd->data[1 + c];
        const uint32_t leafMask = nd->leafMask;
        NodeLanes nl;
        loadNodeLanes(nodes, node, c, projAxis, nl);
        const float sizeC = ftz(sizeA + sizeB);
        uint32_t mask = waveTestChildren(w, nl, role);
        mask = waveTestRefine(w, nl, role, mask);
        if (mask) {
            uint32_t leaf = mask & leafMask;
            if (leaf) {
                mask &= ~leaf;
                insert(0.0f, leaf, tMin, node);
            }
            for (int k = 0; k < 4; k++) {
                if (!(mask & (1u << k)))
                    continue;
                const uint32_t size = readLane(fbits(sizeC), uint32_t(k));
                insert(daz(size), size, bitsf(readLane(fbits(distC), uint32_t(k))), node + readLane(offC, uint32_t(k)));
            }
        }
        if (top >= BLOCK_STACK_SIZE - 4 || top == 0 || readLane(sBits, top - 1u) <= 0xfu)
            break;
        top--;
        node = readLane(sNode, top);
        tMin = bitsf(readLane(fbits(sT), top));
    }
    if (lane < top) {
        cull[lane].bits = sBits;
        cull[lane].tMinTemp = sT;
        cull[lane].node = sNode;
    }
    waveLdsSync();
    // BlockFrame::sort as ranks: entry t ends above every entry with a larger distance
    // and above earlier entries with an equal one
    if (lane < top) {
        CullEntry me = cull[lane];
        uint32_t rank = 0;
        for (uint32_t s = 0; s < top; s++) {
            float o = cull[s].tMinTemp;
            rank += (o > me.tMinTemp || (s < lane && o == me.tMinTemp)) ? 1u : 0u;
        }
        StackEntry out;
        out.tMin = me.tMinTemp;
        uint32_t lm = me.bits <= 0xfu ? me.bits : 0u;
        out.nodeMask = (me.node & STACK_NODE_MASK) | (lm << 28);
        stacks[size_t(b) * BLOCK_STACK_SIZE + rank] = out;
    }
    if (lane == 0)
        stackSize[b] = top;
}

__global__ __launch_bounds__(64) void blockTraverseWaveKernel(BlockTraverseArgs a) {
    __shared__ CullEntry cull[BLOCK_STACK_SIZE + 4];
    blockTraverseBody(a, cull, blockIdx.x);
}

__global__ __launch_bounds__(64) void blockTraverseBatchKernel(const BlockTraverseArgs* __restrict__ all,
                                                               const uint32_t* __restrict__ start, uint32_t n) {
    __shared__ CullEntry cull[BLOCK_STACK_SIZE + 4];
    const uint32_t v = batchView(start, n, blockIdx.x);
    blockTraverseBody(batchArgs(all, v), cull, blockIdx.x - start[v]);
}

// ---------------------------------------------------------------------------------
// Tile lists: traverseTiles (traverse_avx.h:412-527), one wave64 per tile.
//
// The reference walks a tile's BVH cut depth-first from its block's sorted stack,
// keeping the stack sorted by distance and emitting each visited node's surviving leaf
// triangles in visit order. A node's test depends only on the tile frustum and the
// node, not on when the walk reaches it, so the walk is split into three phases:
//   A. tests: breadth-first over the nodes the walk will visit, 16 x Slots nodes per step
//      (lane = node slot x child, each lane testing one child against every frustum
//      axis; two node slots per lane group in the standard configuration, whose VGPRs
//      bound its occupancy, four in the queued ones, whose LDS does). Each visited node
//      becomes an LDS record: its final child mask (coarse test, refined when a leaf
//      child passes), its leaf mask, the leaf mask of its stack entry and the index of
//      its first internal child's record; the internal children that pass are appended
//      as new records, in child order, with their distance keys.
//   B. order: the reference's stack walk replayed over the records, with the stack in
//      one VGPR pair: pops, the single-child shortcut, the sorted insertion (one ballot
//      compares every entry), and the sequence of records whose leaves are emitted.
//   C. emit: the list length is summed, space is claimed with one atomic add, and the
//      triangle ranges of the emitted leaves are written in order.
// The dependent global loads per tile drop from one per visited node (~370 on the
// 262k-triangle room scene) to one per 32 nodes.
//
// Phase B leaves out dead records, those with no passing child: the reference pops them
// and does nothing else. Its stack stays sorted by key, ties newest on top (the block
// stack arrives sorted and every insert lands at its sorted position), so a pop always
// takes the least (key, newest) entry; an entry that changes nothing when popped can be
// dropped without moving any other pop. The same holds for a record none of whose
// subtree emits (a pass between phases A and B marks those dead too). A NaN key breaks
// that order (it inserts on top of anything), and the register stack's empty lanes hold
// -inf, so a tile with a NaN or infinite key, or with an unsorted block stack, goes to the
// global-stack kernel, as does a tile whose records or stack exceed the capacity.
//
// Five configurations share the code. The standard one (680 records, a one-register
// 64-entry stack, 8 KB of LDS) runs every tile; a tile it cannot hold is queued for
// the large one (2048 records, a four-register 256-entry stack, 16 KB of LDS), what that
// cannot hold for the huge one (4608 records, eight registers, 512 entries, 36 KB), and
// so on to the giant one (10240 records, sixteen registers, 1024 entries, 80 KB) and the
// colossal one (20480 records, the giant one's stack, all 160 KB of a CU's LDS); each
// takes its queue with persistent workgroups, as many as fit the chip at once (the
// colossal one with a few: it takes the odd tile of a dense scene seen through a wide
// tile frustum, whose records, not its stack, outgrow the giant one; a 32-register stack
// made its unrolled walk 20% slower). What none holds goes to the wave-per-tile walk with
// its stack in LDS.
// ---------------------------------------------------------------------------------
// Record meta: first child record (bits 0..19), child mask (20..23), leaf mask (24..27),
// stack-entry leaf mask (28..31).
__device__ inline uint32_t metaMask(uint32_t m) { return (m >> 20) & 0xfu; }
__device__ inline uint32_t metaLeaf(uint32_t m) { return (m >> 24) & 0xfu; }
__device__ inline uint32_t metaEntry(uint32_t m) { return m >> 28; }
__device__ inline uint32_t packMeta(uint32_t first, uint32_t mask, uint32_t leaf, uint32_t entry) {
    return first | (mask << 20) | (leaf << 24) | (entry << 28);
}

// The standard configuration keeps the record index in bits 10..19 of a stack entry's
// meta (records < 1024), so a pop needs no LDS read; the others stack record indices.
// The standard configuration keeps every record array in LDS. The others keep only the
// keys and metas the stack walk reads there; each record's node index and the emission
// order go to a per-workgroup region of global scratch (their persistent workgroups
// reuse it tile after tile), which halves their LDS and doubles the workgroups a CU holds
// (the walk is latency-bound: throughput follows the resident waves).
struct TileListSmall {
    enum : uint32_t { RecCap = 680, StackCap = REG_STACK_CAP, Packed = 1, LdsRecords = 1, Level = 0, Slots = 2 };
    using Stack = RegStack;
};
struct TileListLarge {
    enum : uint32_t { RecCap = 2048, StackCap = REG_STACK4_CAP, Packed = 0, LdsRecords = 0, Level = 1, Slots = 4 };
    using Stack = RegStack4;
};
struct TileListHuge {
    enum : uint32_t { RecCap = 4608, StackCap = REG_STACK8_CAP, Packed = 0, LdsRecords = 0, Level = 2, Slots = 4 };
    using Stack = RegStack8;
};
struct TileListGiant {
    enum : uint32_t { RecCap = 10240, StackCap = REG_STACK16_CAP, Packed = 0, LdsRecords = 0, Level = 3, Slots = 4 };
    using Stack = RegStack16;
};
struct TileListColossal {
    enum : uint32_t { RecCap = 20480, StackCap = REG_STACK16_CAP, Packed = 0, LdsRecords = 0, Level = 4, Slots = 4 };
    using Stack = RegStack16;
};

// 32-bit words of one workgroup's scratch region: node indices, then 16-bit order entries
template <class Cfg>
constexpr uint32_t recScratchWords() {
    return Cfg::LdsRecords ? 0u : ((Cfg::RecCap + Cfg::RecCap / 2 + 63u) & ~63u);
}

// (the standard configuration keeps its emission order in bits 10..19 of recMeta, which
// its stored metas leave free: entry i of the order in recMeta[i])
template <class Cfg, bool InLds = bool(Cfg::LdsRecords)>
struct TileListLdsRecords {
    uint32_t recNode[Cfg::RecCap];
};
template <class Cfg>
struct TileListLdsRecords<Cfg, false> {};

template <class Cfg>
struct TileListLds : TileListLdsRecords<Cfg> {
    uint32_t recKey[Cfg::RecCap];  // float bits of the record's distance key; in phase C, the
                                   // list length of emitted record i
    uint32_t recMeta[Cfg::RecCap]; // packMeta(first child record, child mask, leaf mask,
                                   // stack-entry leaf mask)
};

template <class Cfg>
__device__ inline uint32_t metaFirst(uint32_t m) {
    return Cfg::Packed ? (m & 0x3ffu) : (m & 0xfffffu);
}
// stack entry of record r with meta m, and back
template <class Cfg>
__device__ inline uint32_t stackEntry(uint32_t r, uint32_t m) {
    return Cfg::Packed ? ((m & ~(0x3ffu << 10)) | (r << 10)) : r;
}
template <class Cfg>
__device__ inline void fromStackEntry(uint32_t e, const TileListLds<Cfg>& sh, uint32_t& rec, uint32_t& meta) {
    if (Cfg::Packed) {
        rec = (e >> 10) & 0x3ffu;
        meta = e;
    } else {
        rec = e;
        meta = uniform(sh.recMeta[e]);
    }
}


__device__ inline void stackInit(RegStack& st, float t, uint32_t m, uint32_t size, uint32_t lane) {
    st.t = lane < size ? t : -INFINITY;
    st.m = m;
    st.size = size;
}
// One child box of a node, as one lane of a 4-lane node group loads it.
struct ChildBox {
    float b0, b1, b2, b3, b4, b5; // bounds[k][child]
    float key;                    // distance key: -bounds[distanceAxis][child]
    uint32_t leafMask;            // the node's leaf mask
    uint32_t childOffset;         // data[1 + child]
};

__device__ inline ChildBox loadChildBox(const BVHNode* __restrict__ nodes, uint32_t node, uint32_t c, uint32_t axis) {
    const BVHNode* __restrict__ nd = nodes + node;
    const float* bb = &nd->bounds[0][0];
    ChildBox x;
    x.b0 = bb[c];
    x.b1 = bb[4 + c];
    x.b2 = bb[8 + c];
    x.b3 = bb[12 + c];
    x.b4 = bb[16 + c];
    x.b5 = bb[20 + c];
    x.key = -bb[axis * 4 + c];
    x.leafMask = nd->leafMask;
    x.childOffset = nd->data[1 + c];
    return x;
}


// Builds one tile's list; false when the tile does not fit this configuration (its count
// and offset are then zero).
template <class Cfg>
__device__ bool tileListBody(const TileListArgs& g, TileListLds<Cfg>& sh, uint32_t tile, uint32_t lane) {
    using Stack = typename Cfg::Stack;
    const BVHNode* __restrict__ nodes = g.nodes;
    const uint32_t block = tile / TILES_PER_BLOCK;
    const FrustumPre* __restrict__ f = g.tilePre + tile;
    const uint32_t n0 = g.blockStackSize[block];
    if (n0 == 0 || f->degenerate) {
        if (lane == 0) {
            g.counts[tile] = 0u;
            g.offsets[tile] = 0u;
            g.tileWork[tile] = 0u;
            g.tileWork[g.tileCount + tile] = 0u;
        }
        return true;
    }
    if constexpr (Cfg::Level > 0) {
        // A tile whose last list on this scene needed a still bigger configuration goes
        // straight to that one's queue (queue h - 1, its length in counters[4 + h]). The
        // standard configuration takes every tile first and keeps no hints: its code stays
        // as short as it is. Hints are 0 (none) or 1..5; anything else is ignored.
        const uint32_t hint = g.useHints ? g.tileLevel[tile] : 0u;
        if (hint > uint32_t(Cfg::Level) && hint <= 5u) {
            if (lane == 0) {
                const unsigned long long slot = atomicAdd(&g.counters[4 + hint], 1ull);
                g.nextQueue[size_t(hint - 1u - Cfg::Level) * g.tileCount + slot] = tile;
            }
            return true;
        }
    }
    uint32_t* gNode = nullptr; // the record node indices and emission order outside LDS
    uint16_t* gOrder = nullptr;
    if constexpr (!Cfg::LdsRecords) {
        gNode = g.recScratch + size_t(blockIdx.x) * recScratchWords<Cfg>();
        gOrder = reinterpret_cast<uint16_t*>(gNode + Cfg::RecCap);
    }
    auto recNode = [&](uint32_t i) -> uint32_t& {
        if constexpr (Cfg::LdsRecords)
            return sh.recNode[i];
        else
            return gNode[i];
    };
    // the emission order: entry i is the i-th record whose leaves are emitted
    auto order = [&](uint32_t i) -> uint32_t {
        if constexpr (Cfg::LdsRecords)
            return (sh.recMeta[i] >> 10) & 0x3ffu;
        else
            return gOrder[i];
    };
    auto setOrder = [&](uint32_t i, uint32_t rec) {
        if constexpr (Cfg::LdsRecords)
            sh.recMeta[i] = (sh.recMeta[i] & ~(0x3ffu << 10)) | (rec << 10);
        else
            gOrder[i] = uint16_t(rec);
    };
    const uint32_t limit = g.stackLimit < uint32_t(Cfg::StackCap) ? g.stackLimit : uint32_t(Cfg::StackCap);
    bool overflow = n0 > limit;
    uint32_t deepest = n0; // deepest stack size seen
    uint32_t nEmit = 0;
    uint32_t nrec = n0; // records

    // ---- phase A: node tests, breadth-first ----
    if (!overflow) {
        const StackEntry* bs = g.blockStacks + size_t(block) * BLOCK_STACK_SIZE;
        bool nanKey = false;
        if (lane < n0) {
            StackEntry e = bs[lane];
            recNode(lane) = e.nodeMask & STACK_NODE_MASK;
            sh.recKey[lane] = fbits(e.tMin);
            sh.recMeta[lane] = packMeta(0u, 0u, 0u, e.nodeMask >> 28);
            nanKey = !(fabsf(e.tMin) <= FLT_MAX); // NaN or infinite
        }
        overflow = ballot(nanKey) != 0ull;
        waveLdsSync();
        const uint32_t s = lane >> 2, c = lane & 3u;
        const uint32_t axis = uint32_t(f->distanceAxis);
        const uint64_t groupBelow = (1ull << (4 * s)) - 1ull;
        constexpr uint32_t S = Cfg::Slots; // records per lane group and step
        for (uint32_t head = 0; head < nrec && !overflow;) {
            // this step tests records [head, end), S per lane group (so their node loads
            // overlap); the children it appends start at nrec
            const uint32_t end = head + 16 * S < nrec ? head + 16 * S : nrec;
            uint32_t recI[S], nodeR[S], entryR[S];
            bool valid[S];
#pragma unroll
            for (uint32_t k = 0; k < S; k++) {
                recI[k] = head + 16 * k + s;
                valid[k] = recI[k] < end;
                nodeR[k] = 0;
                entryR[k] = 0;
                if (valid[k]) {
                    nodeR[k] = recNode(recI[k]);
                    entryR[k] = metaEntry(sh.recMeta[recI[k]]);
                }
            }
            ChildBox X[S];
#pragma unroll
            for (uint32_t k = 0; k < S; k++)
                X[k] = loadChildBox(nodes, nodeR[k], c, axis);
            // (`&` rather than `&&` throughout: no branch per lane condition)
            bool pass[S], refine[S];
            uint64_t need[S];
            uint32_t nNeed = 0;
#pragma unroll
            for (uint32_t k = 0; k < S; k++) {
                pass[k] = valid[k] & childCoarsePass(f, X[k].b0, X[k].b1, X[k].b2, X[k].b3, X[k].b4, X[k].b5);
                const uint32_t m = uint32_t(ballot(pass[k]) >> (4 * s)) & 0xfu;
                // refine only nodes with a passing leaf child (traverse_avx.h:447-452)
                refine[k] = pass[k] & ((m & X[k].leafMask) != 0u);
                need[k] = ballot(refine[k]);
                nNeed += uint32_t(__builtin_popcountll(need[k]));
            }
            if (nNeed != 0u) {
                if (nNeed < 64u) {
                    // Few boxes need the 12-axis test (typically a sixth of the lanes): move
                    // them to lanes 0 .. nNeed - 1 with forward permutes (every other lane
                    // sends to lane 63, which then holds no needed box), test them in one
                    // pass, and read each verdict back from the ballot.
                    uint32_t d[S];
                    float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, g3 = 0.0f, g4 = 0.0f, g5 = 0.0f;
                    uint32_t base = 0;
#pragma unroll
                    for (uint32_t k = 0; k < S; k++) {
                        const uint32_t nk = uint32_t(__builtin_popcountll(need[k]));
                        d[k] = refine[k] ? base + lanesBelow(need[k]) : 63u;
                        const bool mine = lane >= base && lane < base + nk; // slot k's boxes land here
                        auto take = [&](float& gq, float v) {
                            const uint32_t x = uint32_t(__builtin_amdgcn_ds_permute(int(d[k] << 2), int(fbits(v))));
                            gq = mine ? bitsf(x) : gq;
                        };
                        take(g0, X[k].b0);
                        take(g1, X[k].b1);
                        take(g2, X[k].b2);
                        take(g3, X[k].b3);
                        take(g4, X[k].b4);
                        take(g5, X[k].b5);
                        base += nk;
                    }
                    const bool cull = childRefineCull(f, g0, g1, g2, g3, g4, g5);
                    const uint64_t culled = ballot(cull & (lane < nNeed));
#pragma unroll
                    for (uint32_t k = 0; k < S; k++)
                        pass[k] &= !(refine[k] & (((culled >> d[k]) & 1ull) != 0ull));
                } else {
#pragma unroll
                    for (uint32_t k = 0; k < S; k++) {
                        const bool cullK = childRefineCull(f, X[k].b0, X[k].b1, X[k].b2, X[k].b3, X[k].b4, X[k].b5);
                        pass[k] &= !(refine[k] & cullK);
                    }
                }
            }
            // internal children are pushed only from entries without a leaf mask
            // (traverse_avx.h:476-483); slot k's children follow those of slots below it
            uint32_t before = 0;
            bool badKey = false;
#pragma unroll
            for (uint32_t k = 0; k < S; k++) {
                const uint32_t maskK = uint32_t(ballot(pass[k]) >> (4 * s)) & 0xfu;
                const bool inner = pass[k] && ((X[k].leafMask >> c) & 1u) == 0u && entryR[k] == 0u;
                const uint64_t bi = ballot(inner);
                const uint32_t slot = nrec + before + lanesBelow(bi);
                if (inner && slot < Cfg::RecCap) {
                    recNode(slot) = nodeR[k] + X[k].childOffset;
                    sh.recKey[slot] = fbits(X[k].key);
                    sh.recMeta[slot] = 0u;
                }
                if (valid[k] && c == 0) {
                    const uint32_t first = nrec + before + uint32_t(__builtin_popcountll(bi & groupBelow));
                    sh.recMeta[recI[k]] = packMeta(first, maskK, X[k].leafMask, entryR[k]);
                }
                before += uint32_t(__builtin_popcountll(bi));
                badKey |= inner && !(fabsf(X[k].key) <= FLT_MAX);
            }
            nrec += before;
            head = end;
            waveLdsSync();
            if (nrec > Cfg::RecCap || ballot(badKey) != 0ull) {
                overflow = true;
                break;
            }
        }
    }

    // ---- records whose subtree emits nothing leave the walk ----
    // Only a record with a passing leaf child emits, and a visit pushes only the record's
    // own children, so a record none of whose subtree emits changes no emission when it is
    // visited: its entries are popped later without effect. Dropping it moves no other pop,
    // as for a record with no passing child (the stack stays sorted, and the order of the
    // remaining entries, ties included, is the same), and the walk already drops a record
    // whose mask is clear, so that is how such a record is marked. Children follow their
    // parent in record order: one pass from the last 64 records to the first, a chunk
    // repeated while a record in it loses its mask (parents and children in one chunk).
    if (!overflow) {
        for (int32_t b = int32_t((nrec - 1u) & ~63u); b >= 0; b -= 64) {
            const uint32_t r = uint32_t(b) + lane;
            const bool valid = r < nrec;
            const uint32_t meta = valid ? sh.recMeta[r] : 0u;
            const uint32_t mask = metaMask(meta), leaf = metaLeaf(meta);
            const uint32_t inner = metaEntry(meta) == 0u ? mask & ~leaf : 0u;
            const uint32_t first = metaFirst<Cfg>(meta);
            const uint32_t nInner = uint32_t(__builtin_popcount(inner));
            // emitting records keep their mask; the others keep it while a child is live
            bool open = valid && mask != 0u && (mask & leaf) == 0u;
            // only a record whose children start inside this chunk can change in a later round
            const bool local = open && first < uint32_t(b) + 64u;
            for (uint32_t round = 0; round < 64u; round++) {
                // the children's metas in one LDS round trip (no short-circuit chain)
                uint32_t any = 0;
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    if (open && j < nInner)
                        any |= sh.recMeta[first + j];
                const bool drop = open && metaMask(any) == 0u;
                waveLdsSync(); // every lane has read its children before any mask changes
                if (drop)
                    sh.recMeta[r] = meta & ~(0xfu << 20);
                waveLdsSync();
                open = open && !drop;
                if (ballot(drop) == 0ull || ballot(open && local) == 0ull)
                    break;
            }
        }
        waveLdsSync();
    }

    // ---- phase B: the reference's stack walk over the live records, stack in VGPRs ----
    if (!overflow) {
        Stack st;
        {
            // the block stack's live entries, compacted in order (entry q at position q)
            const bool valid = lane < n0;
            const uint32_t key = valid ? sh.recKey[lane] : 0u;
            const uint32_t meta = valid ? sh.recMeta[lane] : 0u;
            const uint64_t live = ballot(valid && metaMask(meta) != 0u);
            const uint32_t size = uint32_t(__builtin_popcountll(live));
            const uint32_t dest = ((live >> lane) & 1u) ? lanesBelow(live) : size + lanesBelow(~live);
            const float t = bitsf(uint32_t(__builtin_amdgcn_ds_permute(int(dest << 2), int(key))));
            const uint32_t m =
                uint32_t(__builtin_amdgcn_ds_permute(int(dest << 2), int(stackEntry<Cfg>(lane, meta))));
            stackInit(st, t, m, size, lane);
            const float above = __shfl_down(t, 1);
            if (ballot(lane + 1u < size && t < above) != 0ull)
                overflow = true;
            deepest = size;
        }
        uint32_t emitBuf = 0; // emission order, 64 records at a time
        uint32_t steps = 0;
        uint32_t curRec = 0, curMeta = 0;
        bool walking = !overflow && st.size != 0u;
        if (walking)
            fromStackEntry<Cfg>(stackPop(st, lane), sh, curRec, curMeta);
        while (walking) {
            // every record is reached at most once; the bound keeps the loop finite whatever
            // the records hold
            if (++steps > nrec) {
                overflow = true;
                break;
            }
            const uint32_t mask = metaMask(curMeta), leafMask = metaLeaf(curMeta);
            if (mask & leafMask) {
                emitBuf = writeLane(emitBuf, curRec, nEmit & 63u, lane);
                nEmit++;
                if ((nEmit & 63u) == 0u)
                    setOrder(nEmit - 64u + lane, emitBuf);
            }
            // internal children are followed only from entries without a leaf mask
            const uint32_t inner = metaEntry(curMeta) == 0u ? mask & ~leafMask : 0u;
            // Every path ends at the one pop below (or skips it), so the loop keeps a single
            // latch: no state variable or register copies between the ways a step can end.
            bool pop = true;
            if (inner != 0u) {
                // the hit children's records, contiguous from `first` in child order; lanes
                // past them read an in-range record and are masked off
                const uint32_t first = metaFirst<Cfg>(curMeta);
                const uint32_t nChildren = uint32_t(__builtin_popcount(inner));
                const uint32_t r = first + lane < uint32_t(Cfg::RecCap) ? first + lane : uint32_t(Cfg::RecCap) - 1u;
                const uint32_t childKey = sh.recKey[r];
                const uint32_t childMeta = sh.recMeta[r];
                const uint32_t liveBits =
                    uint32_t(ballot(metaMask(childMeta) != 0u)) & ((1u << nChildren) - 1u);
                if (nChildren == 1u) {
                    // single hit child: test it next without touching the stack
                    if (liveBits != 0u) {
                        curRec = first;
                        curMeta = readLane(childMeta, 0);
                        pop = false;
                    }
                } else if (liveBits != 0u) {
                    // children in child order, each inserted below the run of closer entries
                    // (traverse_avx.h:499-520), dead ones left out
                    const uint32_t nLive = uint32_t(__builtin_popcount(liveBits));
                    if (st.size + nLive - 1u >= limit) {
                        overflow = true;
                        break;
                    }
                    // The walk pops the top right after these inserts. When the closest child
                    // would land on top (ties go to the newest insert), visit it directly and
                    // insert only its siblings: the stack ends the same. The newest of the
                    // closest live children (the tile has no NaN keys): the least key over the
                    // four child lanes (two quad DPP steps), then the highest live lane
                    // holding it.
                    const bool live = ((uint64_t(liveBits) >> lane) & 1ull) != 0ull;
                    const float ck = live ? bitsf(childKey) : INFINITY;
                    float least = fminf(ck, bitsf(uint32_t(__builtin_amdgcn_update_dpp(
                                                0, int(fbits(ck)), 0xb1, 0xf, 0xf, false)))); // lane ^ 1
                    least = fminf(least, bitsf(uint32_t(__builtin_amdgcn_update_dpp(
                                             0, int(fbits(least)), 0x4e, 0xf, 0xf, false)))); // lane ^ 2
                    const uint32_t atLeast = uint32_t(ballot(ck == least)) & liveBits;
                    const uint32_t best = 31u - uint32_t(__builtin_clz(atLeast));
                    const float bestKey = bitsf(readLane(fbits(least), 0));
                    const bool take = st.size == 0u || bestKey <= stackTopKey(st);
                    const uint32_t entry = stackEntry<Cfg>(first + lane, childMeta);
                    // at most four siblings: straight-line code (a loop here keeps the
                    // compiler from holding the stack in the same registers on every path)
                    uint32_t sib = liveBits & ~(take ? 1u << best : 0u);
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        if (sib != 0u) {
                            const uint32_t j = uint32_t(__builtin_ctz(sib));
                            stackInsert(st, bitsf(readLane(childKey, j)), readLane(entry, j), lane);
                            sib &= sib - 1u;
                        }
                    }
                    const uint32_t depth = st.size + (take ? 1u : 0u);
                    deepest = depth > deepest ? depth : deepest;
                    if (take) {
                        curRec = first + best;
                        curMeta = readLane(childMeta, best);
                        pop = false;
                    }
                }
            }
            if (pop) {
                if (st.size == 0u)
                    break;
                fromStackEntry<Cfg>(stackPop(st, lane), sh, curRec, curMeta);
            }
        }
        if (!overflow && (nEmit & 63u) != 0u && lane < (nEmit & 63u))
            setOrder((nEmit & ~63u) + lane, emitBuf);
        waveLdsSync();
    }
    if (overflow) {
        if (lane == 0) {
            g.counts[tile] = 0u;
            g.offsets[tile] = 0u;
        }
        return false;
    }

    // ---- phase C: emit ----
    // The first 64 emitted records keep their node's leaf ranges and prefix in registers
    // from the length pass to the write pass; later ones (dense tiles) are read again.
    uint32_t total = 0;
    uint32_t d0[5] = {0u, 0u, 0u, 0u, 0u};
    uint32_t leaf0 = 0, len0 = 0, incl0 = 0;
    for (uint32_t base = 0; base < nEmit; base += 64) {
        const uint32_t i = base + lane;
        uint32_t len = 0;
        if (i < nEmit) {
            const uint32_t rec = order(i);
            const uint32_t meta = sh.recMeta[rec];
            const uint32_t leaf = metaMask(meta) & metaLeaf(meta);
            const uint32_t* d = nodes[recNode(rec)].data;
            uint32_t dk[5];
#pragma unroll
            for (int k = 0; k < 5; k++)
                dk[k] = d[k];
#pragma unroll
            for (int k = 0; k < 4; k++)
                if ((leaf & (1u << k)) && dk[k + 1] > dk[k])
                    len += dk[k + 1] - dk[k];
            if (base == 0) {
#pragma unroll
                for (int k = 0; k < 5; k++)
                    d0[k] = dk[k];
                leaf0 = leaf;
            } else {
                sh.recKey[i] = len;
            }
        }
        if (base == 0) {
            len0 = len;
            incl0 = waveInclusiveScan(len);
            total += readLane(incl0, 63);
        } else {
            total += waveSum(len);
        }
    }
    unsigned long long start = 0;
    if (lane == 0)
        start = atomicAdd(&g.counters[3], (unsigned long long)total);
    start = (uint64_t(readLane(uint32_t(start >> 32), 0)) << 32) | readLane(uint32_t(start), 0);
    const bool fits = start + total <= g.capacity;
    if (lane == 0) {
        g.counts[tile] = total;
        g.offsets[tile] = uint32_t(start);
        if (!fits)
            atomicOr(&g.counters[1], 1ull);
        g.tileWork[tile] = nrec; // node visits
        g.tileWork[g.tileCount + tile] = deepest;
        if constexpr (Cfg::Level > 0) {
            // the smallest queued configuration this tile's records and stack fit, for its
            // next frame
            const uint32_t level = nrec <= TileListLarge::RecCap && deepest <= REG_STACK4_CAP    ? 1u
                                   : nrec <= TileListHuge::RecCap && deepest <= REG_STACK8_CAP   ? 2u
                                   : nrec <= TileListGiant::RecCap ? 3u
                                                                   : 4u;
            g.tileLevel[tile] = level;
        }
    }
    if (!fits)
        return true;
    {
        uint32_t at = uint32_t(start) + incl0 - len0;
        if (len0) {
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (leaf0 & (1u << k))
                    for (uint32_t t = d0[k]; t < d0[k + 1]; t++)
                        g.triOut[at++] = t;
        }
    }
    uint32_t run = uint32_t(start) + (nEmit ? readLane(incl0, 63) : 0u);
    for (uint32_t base = 64; base < nEmit; base += 64) {
        const uint32_t i = base + lane;
        const uint32_t len = i < nEmit ? sh.recKey[i] : 0u;
        const uint32_t incl = waveInclusiveScan(len);
        uint32_t at = run + incl - len;
        run += readLane(incl, 63);
        if (len) {
            const uint32_t rec = order(i);
            const uint32_t meta = sh.recMeta[rec];
            const uint32_t leaf = metaMask(meta) & metaLeaf(meta);
            const uint32_t* d = nodes[recNode(rec)].data;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t t0 = d[k], t1 = d[k + 1];
                if (leaf & (1u << k))
                    for (uint32_t t = t0; t < t1; t++)
                        g.triOut[at++] = t;
            }
        }
    }
    return true;
}

// every tile, standard configuration; a tile it cannot hold is queued for the large one
__device__ inline void tileListSmallOne(const TileListArgs& g, TileListLds<TileListSmall>& sh, uint32_t tile,
                                        uint32_t lane) {
    if (!tileListBody<TileListSmall>(g, sh, tile, lane) && lane == 0) {
        const unsigned long long slot = atomicAdd(&g.counters[g.nextCounter], 1ull);
        g.nextQueue[slot] = tile;
    }
}

// The LDS records (680 x 12 bytes, 8160 bytes: the emission order shares the metas' words)
// let a CU hold 20 of these waves. A record capacity below what a scene's densest tiles
// need costs more than the waves it adds: the tiles it sends to the large configuration
// run after this kernel, each a serial walk of one wave (the densest tile of the stereo
// room eye holds 652 records).
__global__ __launch_bounds__(64) void tileListKernel(TileListArgs g) {
    __shared__ TileListLds<TileListSmall> sh;
    tileListSmallOne(g, sh, blockIdx.x, threadIdx.x);
}

__global__ __launch_bounds__(64) void tileListBatchKernel(const TileListArgs* __restrict__ all,
                                                          const uint32_t* __restrict__ start, uint32_t n,
                                                          uint32_t swizzle) {
    __shared__ TileListLds<TileListSmall> sh;
    const uint32_t unit = xcdUnit(blockIdx.x, gridDim.x, swizzle);
    const uint32_t v = batchView(start, n, unit);
    tileListSmallOne(batchArgs(all, v), sh, unit - start[v], threadIdx.x);
}

// The tiles a smaller configuration queued, persistent workgroups: this workgroup takes
// the entries `first`, `first` + stride, ... of the queue; a tile this configuration
// cannot hold either is queued for the next.
template <class Cfg>
__device__ inline void tileListQueueRun(const TileListArgs& g, TileListLds<Cfg>& sh, uint32_t first, uint32_t stride,
                                        uint32_t queued, uint32_t lane) {
    for (uint32_t i = first; i < queued; i += stride) {
        const uint32_t tile = uniform(g.queue[i]);
        waveLdsSync(); // the previous tile's LDS reads are done
        if (!tileListBody<Cfg>(g, sh, tile, lane) && lane == 0) {
            const unsigned long long slot = atomicAdd(&g.counters[g.nextCounter], 1ull);
            g.nextQueue[slot] = tile;
        }
    }
}

template <class Cfg>
__global__ __launch_bounds__(64) void tileListQueueKernel(TileListArgs g) {
    __shared__ TileListLds<Cfg> sh;
    const uint32_t queued = uniform(uint32_t(g.counters[g.queueCounter]));
    tileListQueueRun<Cfg>(g, sh, blockIdx.x, gridDim.x, queued, threadIdx.x);
}

// every view's queue in turn, the entries of all of them dealt to the workgroups as one
// sequence
template <class Cfg>
__global__ __launch_bounds__(64) void tileListQueueBatchKernel(const TileListArgs* __restrict__ all, uint32_t n) {
    __shared__ TileListLds<Cfg> sh;
    uint32_t base = 0;
    for (uint32_t v = 0; v < n; v++) {
        const TileListArgs& g = all[v];
        const uint32_t queued = uniform(uint32_t(g.counters[g.queueCounter]));
        const uint32_t first = (blockIdx.x + gridDim.x - base % gridDim.x) % gridDim.x;
        tileListQueueRun<Cfg>(g, sh, first, gridDim.x, queued, threadIdx.x);
        base += queued;
    }
}

// ---------------------------------------------------------------------------------
// The reference's tile walk as written (traverseTiles, traverse_avx.h:412-527), one wave
// per tile, for the tiles no LDS configuration holds and those with a NaN or infinite key
// or an unsorted block stack. The walk is wave-uniform (one node per step, tested by every
// lane alike); the distance-sorted stack, up to TILE_STACK_SIZE entries, lives in LDS and
// the wave inserts into it together (LdsStack: the sequential insertion's position, one
// ballot per 64 entries, then a 64-wide move). The current node's entry stays on top
// while it is tested, as in the reference: a node with several passing internal
// children replaces its entry by them, each at its sorted position; a single one is
// visited directly. Count pass (emit = 0) and emit pass (emit = 1) walk alike, one after
// the other in the same wave: the count pass claims the list space and hands its start to
// the emit pass in `base`.
// ---------------------------------------------------------------------------------
struct TileWalkLds {
    float t[TILE_STACK_SIZE];
    uint32_t m[TILE_STACK_SIZE];
};

__device__ inline void tileWalkOne(const TileWalkArgs& a, TileWalkLds& sh, uint32_t tile, int emit,
                                   uint32_t lane, uint64_t& base) {
    const BVHNode* __restrict__ nodes = a.nodes;
    const uint64_t capacity = a.capacity;
    unsigned long long* __restrict__ counters = a.counters;
    uint32_t* __restrict__ counts = a.counts;
    uint32_t* __restrict__ triOut = a.triOut;
    {
        const uint32_t block = tile / TILES_PER_BLOCK;
        const uint32_t n0 = uniform(a.blockStackSize[block]);
        const FrustumPre& f = a.tilePre[tile];
        if (n0 == 0 || f.degenerate) {
            if (!emit && lane == 0) {
                counts[tile] = 0;
                a.offsets[tile] = 0u;
                a.tileWork[tile] = 0u;
                a.tileWork[a.tileCount + tile] = 0u;
            }
            return;
        }
        waveLdsSync(); // the previous tile's LDS reads are done
        LdsStack st;
        stackBindLds(st, sh.t, sh.m);
        bool irregular = false;
        {
            const StackEntry* bs = a.blockStacks + size_t(block) * BLOCK_STACK_SIZE;
            float t0 = 0.0f;
            uint32_t m0 = 0u;
            if (lane < n0) {
                t0 = bs[lane].tMin;
                m0 = bs[lane].nodeMask;
            }
            // what sends a tile here whatever its size: a NaN or infinite key, an unsorted
            // block stack (the walk's own keys are checked as it pushes them)
            const float above = __shfl_down(t0, 1);
            irregular = ballot((lane < n0 && !(fabsf(t0) <= FLT_MAX)) || (lane + 1u < n0 && t0 < above)) != 0ull;
            stackInit(st, t0, m0, n0, lane);
        }
        const uint64_t outBase = emit ? base : 0;
        uint32_t count = 0, visits = 0, maxSize = n0;
        bool listOverflow = false, stackOverflow = false;
        const int axis = f.distanceAxis;
        uint32_t topEntry = stackAt(st, n0 - 1u);
        uint32_t node = topEntry & STACK_NODE_MASK;
        uint32_t entryMask = topEntry >> 28;
        for (;;) {
            const BVHNode& n = nodes[node];
            visits++;
            uint32_t mask = testChildren(f, n);
            bool pop = true;
            if (mask) {
                uint32_t leaf = mask & n.leafMask;
                if (leaf) {
                    mask = testChildrenRefine(f, n, mask);
                    leaf = mask & n.leafMask;
                    for (int k = 0; k < 4; k++) {
                        if (!(leaf & (1u << k)))
                            continue;
                        const uint32_t t0 = n.data[k], t1 = n.data[k + 1];
                        if (emit) {
                            for (uint32_t t = t0 + lane; t < t1; t += 64) {
                                const uint64_t at = outBase + count + (t - t0);
                                if (at < capacity)
                                    triOut[at] = t;
                                else
                                    listOverflow = true;
                            }
                        }
                        count += t1 > t0 ? t1 - t0 : 0;
                    }
                }
                mask &= ~n.leafMask;
                if (mask && entryMask == 0) {
                    if ((mask & (mask - 1)) == 0) {
                        node = node + n.data[1 + __builtin_ctz(mask)];
                        continue;
                    }
                    const uint32_t nc = uint32_t(__builtin_popcount(mask));
                    if (st.size - 1u + nc > TILE_STACK_SIZE) {
                        stackOverflow = true;
                        break;
                    }
                    st.size--; // the children take the current entry's place
                    for (int k = 0; k < 4; k++)
                        if (mask & (1u << k)) {
                            irregular |= !(fabsf(n.bounds[axis][k]) <= FLT_MAX);
                            stackInsert(st, -n.bounds[axis][k], (node + n.data[1 + k]) & STACK_NODE_MASK, lane);
                        }
                    maxSize = st.size > maxSize ? st.size : maxSize;
                    // the closest entry may be an older one (possibly a block-level leaf entry)
                    topEntry = stackAt(st, st.size - 1u);
                    node = topEntry & STACK_NODE_MASK;
                    entryMask = topEntry >> 28;
                    pop = false;
                }
            }
            if (pop) {
                if (st.size == 1u)
                    break;
                st.size--;
                topEntry = stackAt(st, st.size - 1u);
                node = topEntry & STACK_NODE_MASK;
                entryMask = topEntry >> 28;
            }
        }
        const bool overflow = stackOverflow || ballot(listOverflow) != 0ull;
        unsigned long long start = 0;
        if (lane == 0) {
            if (!emit) {
                // the count pass claims the tile's list space; the emit pass fills it
                counts[tile] = count;
                start = atomicAdd(&counters[3], (unsigned long long)count);
                a.offsets[tile] = uint32_t(start);
                if (start + count > capacity)
                    atomicOr(&counters[1], 1ull);
                // the smallest queued configuration whose records and stack would hold this
                // walk (the walk's stack is the unpruned one, never shallower), unless the
                // tile can only be walked
                // (with a view stack limit below 64 every configuration has that limit)
                const bool full = a.stackLimit >= REG_STACK_CAP;
                const uint32_t capL = full ? REG_STACK4_CAP : a.stackLimit;
                const uint32_t capH = full ? REG_STACK8_CAP : a.stackLimit;
                const uint32_t capG = full ? REG_STACK16_CAP : a.stackLimit;
                const uint32_t capC = full ? REG_STACK16_CAP : a.stackLimit;
                a.tileLevel[tile] = irregular                                               ? 5u
                                    : visits <= TileListLarge::RecCap && maxSize <= capL    ? 1u
                                    : visits <= TileListHuge::RecCap && maxSize <= capH     ? 2u
                                    : visits <= TileListGiant::RecCap && maxSize <= capG    ? 3u
                                    : visits <= TileListColossal::RecCap && maxSize <= capC ? 4u
                                                                                            : 5u;
                a.tileWork[tile] = visits; // node visits
                a.tileWork[a.tileCount + tile] = maxSize;
            }
            if (overflow)
                atomicOr(&counters[1], 1ull);
        }
        if (!emit)
            base = uint64_t(readLane(uint32_t(start), 0)) | (uint64_t(readLane(uint32_t(start >> 32), 0)) << 32);
    }
}

// the count pass, then the emit pass, of one queued tile
__device__ inline void tileWalkTile(const TileWalkArgs& a, TileWalkLds& sh, uint32_t tile, uint32_t lane) {
    uint64_t base = 0;
    tileWalkOne(a, sh, tile, 0, lane, base);
    tileWalkOne(a, sh, tile, 1, lane, base);
}

__global__ __launch_bounds__(64) void tileWalkKernel(TileWalkArgs a) {
    __shared__ TileWalkLds sh;
    const uint32_t queued = uniform(uint32_t(a.counters[9]));
    for (uint32_t qi = blockIdx.x; qi < queued; qi += gridDim.x)
        tileWalkTile(a, sh, uniform(a.queue[qi]), threadIdx.x);
}

__global__ __launch_bounds__(64) void tileWalkBatchKernel(const TileWalkArgs* __restrict__ all, uint32_t n) {
    __shared__ TileWalkLds sh;
    uint32_t base = 0;
    for (uint32_t v = 0; v < n; v++) {
        const TileWalkArgs a = batchArgs(all, v);
        const uint32_t queued = uniform(uint32_t(a.counters[9]));
        for (uint32_t qi = (blockIdx.x + gridDim.x - base % gridDim.x) % gridDim.x; qi < queued; qi += gridDim.x)
            tileWalkTile(a, sh, uniform(a.queue[qi]), threadIdx.x);
        base += queued;
    }
}

FrustumSetupArgs frustumSetupArgs(const hvvr_gpu_view* v, const matrix4x4& m) {
    FrustumSetupArgs a;
    a.tileCam = v->tileFrustaCam.ptr;
    a.blockCam = v->blockFrustaCam.ptr;
    a.tileWorld = v->tileFrustaWorld.ptr;
    a.tilePre = v->tilePre.ptr;
    a.blockPre = v->blockPre.ptr;
    a.tileCount = v->tileCount;
    a.blockCount = v->blockCount;
    a.cameraToWorld = m;
    a.cull = v->cull;
    a.counters = reinterpret_cast<unsigned long long*>(v->counters.ptr);
    a.counterCount = uint32_t(v->counters.count);
    return a;
}

BlockTraverseArgs blockTraverseArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s) {
    BlockTraverseArgs a;
    a.nodes = s->nodes.ptr;
    a.blockPre = v->blockPre.ptr;
    a.stacks = v->blockStacks.ptr;
    a.stackSize = v->blockStackSize.ptr;
    return a;
}

TileWalkArgs tileWalkArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s) {
    TileWalkArgs a;
    a.nodes = s->nodes.ptr;
    a.tilePre = v->tilePre.ptr;
    a.blockStacks = v->blockStacks.ptr;
    a.blockStackSize = v->blockStackSize.ptr;
    a.queue = v->tileQueues.ptr + 4 * size_t(v->tileCount);
    a.counts = v->tileCounts.ptr;
    a.offsets = v->tileOffsets.ptr;
    a.triOut = v->triIndices.ptr;
    a.capacity = v->listCapacity;
    a.counters = reinterpret_cast<unsigned long long*>(v->counters.ptr);
    a.tileLevel = v->tileLevel.ptr;
    a.tileWork = v->tileWork.ptr;
    a.tileCount = v->tileCount;
    a.stackLimit = v->waveStackLimit;
    return a;
}

void launchFrustumSetup(hvvr_gpu_view* v, const matrix4x4& m, hipStream_t stream) {
    frustumSetupKernel<<<frustumSetupUnits(v), 128, 0, stream>>>(frustumSetupArgs(v, m));
}

void launchBlockTraverse(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream) {
    blockTraverseWaveKernel<<<v->blockCount, 64, 0, stream>>>(blockTraverseArgs(v, s));
}

void launchFrustumSetupBatch(const FrustumSetupArgs* all, const uint32_t* start, uint32_t viewCount, uint32_t units,
                             hipStream_t stream) {
    frustumSetupBatchKernel<<<units, 128, 0, stream>>>(all, start, viewCount);
}

void launchBlockTraverseBatch(const BlockTraverseArgs* all, const uint32_t* start, uint32_t viewCount,
                              uint32_t units, hipStream_t stream) {
    blockTraverseBatchKernel<<<units, 64, 0, stream>>>(all, start, viewCount);
}

// Rank of every tile among the tiles with a non-empty list (the reference's compacted
// tile index, traversal.cpp:281-292): one workgroup scans the list lengths in tile order.
enum : uint32_t { kRankThreads = 1024 };
__global__ __launch_bounds__(kRankThreads) void occupiedRankKernel(const uint32_t* __restrict__ counts, uint32_t n,
                                                                   uint32_t* __restrict__ rank) {
    __shared__ uint32_t waveTotals[kRankThreads / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += kRankThreads) {
        const uint32_t i = base + tid;
        const bool occupied = i < n && counts[i] != 0u;
        const uint64_t b = ballot(occupied);
        if (lane == 0)
            waveTotals[wave] = uint32_t(__builtin_popcountll(b));
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < kRankThreads / 64; w++) {
            before += w < wave ? waveTotals[w] : 0u;
            total += waveTotals[w];
        }
        if (i < n)
            rank[i] = carry + before + lanesBelow(b);
        carry += total;
        __syncthreads();
    }
}

void launchOccupiedRank(hvvr_gpu_view* v, hipStream_t stream) {
    occupiedRankKernel<<<1, kRankThreads, 0, stream>>>(v->tileCounts.ptr, v->tileCount, v->occRank.ptr);
}

// The frame's work statistics from the per-tile values (one workgroup): node visits and
// surviving list entries summed, the deepest stack as the maximum; a survivor is tested
// against the tile's 128 samples.
enum : uint32_t { kReduceThreads = 256 };
__global__ __launch_bounds__(kReduceThreads) void tileWorkReduceKernel(const uint32_t* __restrict__ work, uint32_t n,
                                                                       unsigned long long* __restrict__ counters) {
    __shared__ unsigned long long sums[3][kReduceThreads];
    __shared__ uint32_t deepest[kReduceThreads];
    const uint32_t tid = threadIdx.x;
    unsigned long long visits = 0, survivors = 0, tested = 0;
    uint32_t deep = 0;
    for (uint32_t i = tid; i < n; i += kReduceThreads) {
        visits += work[i];
        deep = max(deep, work[n + i]);
        survivors += work[2 * size_t(n) + i];
        tested += work[3 * size_t(n) + i];
    }
    sums[0][tid] = visits;
    sums[1][tid] = survivors;
    sums[2][tid] = tested;
    deepest[tid] = deep;
    __syncthreads();
    for (uint32_t half = kReduceThreads / 2; half > 0; half >>= 1) {
        if (tid < half) {
            sums[0][tid] += sums[0][tid + half];
            sums[1][tid] += sums[1][tid + half];
            sums[2][tid] += sums[2][tid + half];
            deepest[tid] = max(deepest[tid], deepest[tid + half]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        counters[0] = sums[1][0] * TILE_SIZE;
        counters[2] = deepest[0];
        counters[4] = sums[0][0];
        counters[10] = sums[2][0] * 64u; // (wave, survivor) tests x the wave's 64 samples
    }
}

void launchTileWorkReduce(hvvr_gpu_view* v, hipStream_t stream) {
    tileWorkReduceKernel<<<1, kReduceThreads, 0, stream>>>(v->tileWork.ptr, v->tileCount,
                                                          reinterpret_cast<unsigned long long*>(v->counters.ptr));
}

TileListArgs tileListArgs(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, int level) {
    uint32_t* queues = v->tileQueues.ptr; // [5][tileCount]: large, huge, giant, colossal, walk
    const size_t T = v->tileCount;
    TileListArgs g;
    g.nodes = s->nodes.ptr;
    g.tilePre = v->tilePre.ptr;
    g.blockStacks = v->blockStacks.ptr;
    g.blockStackSize = v->blockStackSize.ptr;
    g.counts = v->tileCounts.ptr;
    g.offsets = v->tileOffsets.ptr;
    g.triOut = v->triIndices.ptr;
    g.capacity = v->listCapacity;
    g.counters = reinterpret_cast<unsigned long long*>(v->counters.ptr);
    // level L takes queue L - 1 (counters[4 + L]) and fills queue L (counters[5 + L]); the
    // larger configurations' stack limit is the view's limit scaled past the standard one
    // (a test limit below 64 keeps every configuration short)
    g.queue = level > 0 ? queues + (level - 1) * T : nullptr;
    g.queueCounter = level > 0 ? uint32_t(4 + level) : 0u;
    g.nextQueue = queues + level * T;
    g.nextCounter = uint32_t(5 + level);
    const bool full = v->waveStackLimit >= REG_STACK_CAP;
    const uint32_t caps[5] = {v->waveStackLimit, REG_STACK4_CAP, REG_STACK8_CAP, REG_STACK16_CAP, REG_STACK16_CAP};
    g.stackLimit = level == 0 || !full ? v->waveStackLimit : caps[level];
    g.recScratch = v->recScratch.ptr;
    g.tileLevel = v->tileLevel.ptr;
    g.useHints = v->hintScene == s ? 1u : 0u;
    g.tileCount = v->tileCount;
    g.tileWork = v->tileWork.ptr;
    return g;
}

static uint32_t clampGroups(uint32_t want, uint32_t tiles) {
    return want < tiles ? want : (tiles ? tiles : 1u);
}

size_t tileListScratchWords() {
    const size_t large = size_t(kMaxLargeGroups) * recScratchWords<TileListLarge>();
    const size_t huge = size_t(kMaxHugeGroups) * recScratchWords<TileListHuge>();
    const size_t giant = size_t(kMaxGiantGroups) * recScratchWords<TileListGiant>();
    const size_t colossal = size_t(kMaxColossalGroups) * recScratchWords<TileListColossal>();
    return std::max(std::max(large, huge), std::max(giant, colossal));
}

void tileListGroupsFor(size_t words, uint32_t groups[4]) {
    const size_t per[4] = {recScratchWords<TileListLarge>(), recScratchWords<TileListHuge>(),
                           recScratchWords<TileListGiant>(), recScratchWords<TileListColossal>()};
    const uint32_t most[4] = {kMaxLargeGroups, kMaxHugeGroups, kMaxGiantGroups, kMaxColossalGroups};
    for (int k = 0; k < 4; k++)
        groups[k] = uint32_t(std::min<size_t>(most[k], words / per[k]));
}

// the arguments of configuration `level` with the record scratch and queues of `scratch`
static TileListArgs tileListArgsWith(const hvvr_gpu_view* v, const hvvr_gpu_scene* s, int level,
                                     const TileListScratch* scratch) {
    TileListArgs g = tileListArgs(v, s, level);
    if (scratch) {
        const size_t T = v->tileCount;
        g.recScratch = scratch->rec;
        g.queue = level > 0 ? scratch->queues + (level - 1) * T : nullptr;
        g.nextQueue = scratch->queues + level * T;
    }
    return g;
}

void launchTileLists(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream, const TileListScratch* scratch,
                     hipEvent_t* timing) {
    const LaunchShape shape;
    const uint32_t T = v->tileCount;
    const uint32_t large = scratch ? scratch->groups[0] : std::min(shape.largeGroups, uint32_t(kMaxLargeGroups));
    const uint32_t huge = scratch ? scratch->groups[1] : std::min(shape.hugeGroups, uint32_t(kMaxHugeGroups));
    const uint32_t giant = scratch ? scratch->groups[2] : std::min(shape.giantGroups, uint32_t(kMaxGiantGroups));
    const uint32_t colossal = scratch ? scratch->groups[3] : uint32_t(kMaxColossalGroups);
    if (timing)
        hipExtLaunchKernelGGL(tileListKernel, dim3(T), dim3(64), 0, stream, timing[0], timing[1], 0,
                              tileListArgsWith(v, s, 0, scratch));
    else
        tileListKernel<<<T, 64, 0, stream>>>(tileListArgsWith(v, s, 0, scratch));
    // 16 KB of LDS: ten workgroups per CU; 36 KB: four; 80 KB: two; 160 KB: one
    tileListQueueKernel<TileListLarge><<<clampGroups(large, T), 64, 0, stream>>>(tileListArgsWith(v, s, 1, scratch));
    tileListQueueKernel<TileListHuge><<<clampGroups(huge, T), 64, 0, stream>>>(tileListArgsWith(v, s, 2, scratch));
    tileListQueueKernel<TileListGiant><<<clampGroups(giant, T), 64, 0, stream>>>(tileListArgsWith(v, s, 3, scratch));
    tileListQueueKernel<TileListColossal>
        <<<clampGroups(colossal, T), 64, 0, stream>>>(tileListArgsWith(v, s, 4, scratch));
}

void launchTileListFallback(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream,
                            const TileListScratch* scratch) {
    const LaunchShape shape;
    const uint32_t T = v->tileCount;
    TileWalkArgs a = tileWalkArgs(v, s);
    if (scratch)
        a.queue = scratch->queues + 4 * size_t(T);
    tileWalkKernel<<<clampGroups(shape.walkGroups, T), 64, 0, stream>>>(a); // 32 KB of LDS each
}

void launchTileListsBatch(const TileListArgs* all, const uint32_t* start, uint32_t viewCount, uint32_t units,
                          const LaunchShape& shape, hipStream_t stream) {
    tileListBatchKernel<<<units, 64, 0, stream>>>(all, start, viewCount, shape.xcdSwizzle);
    tileListQueueBatchKernel<TileListLarge>
        <<<clampGroups(std::min(shape.largeGroups, uint32_t(kMaxLargeGroups)), units), 64, 0, stream>>>(
            all + viewCount, viewCount);
    tileListQueueBatchKernel<TileListHuge>
        <<<clampGroups(std::min(shape.hugeGroups, uint32_t(kMaxHugeGroups)), units), 64, 0, stream>>>(
            all + 2 * viewCount, viewCount);
    tileListQueueBatchKernel<TileListGiant>
        <<<clampGroups(std::min(shape.giantGroups, uint32_t(kMaxGiantGroups)), units), 64, 0, stream>>>(
            all + 3 * viewCount, viewCount);
    tileListQueueBatchKernel<TileListColossal>
        <<<clampGroups(uint32_t(kMaxColossalGroups), units), 64, 0, stream>>>(all + 4 * viewCount, viewCount);
}

void launchTileWalkBatch(const TileWalkArgs* all, uint32_t viewCount, uint32_t maxTiles, const LaunchShape& shape,
                         hipStream_t stream) {
    tileWalkBatchKernel<<<clampGroups(shape.walkGroups, maxTiles), 64, 0, stream>>>(all, viewCount);
}

} // namespace hvvr
