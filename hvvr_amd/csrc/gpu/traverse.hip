// Per-frame frustum setup and frustum-BVH traversal on the GPU. The reference runs
// both on the CPU with AVX2 (raycaster/render.cpp:282-365, raycaster/traversal.cpp:
// 99-330, raycaster/traverse_avx.h:293-527); here they are kernels whose output, the
// per-tile triangle lists, never leaves HBM.
//
//   frustumSetupKernel        one thread per tile or block frustum: camera->world
//                             transform (traversal.h:38-47) + precompute (traverse_avx.cpp)
//   blockTraverseWaveKernel   one wave64 per block: cooperative node tests, the <= 64-entry
//                             cull stack in LDS, then the distance sort of
//                             BlockFrame::sort as a parallel rank computation
//   tileTraverseWaveKernel    one wave64 per tile (wave_traverse.h). Run twice: count pass,
//                             then (after an exclusive scan of the counts) emit pass, so
//                             every tile's list lands contiguously in tile order with no
//                             host round trip and no atomics.
//   tileTraverseKernel        one thread per tile with the full 4096-entry stack in global
//                             memory; only re-runs tiles whose LDS stack overflowed.
//   scan kernels              3-phase exclusive scan of the per-tile counts.

#include "device_common.h"
#include "gpu_internal.h"
#include "wave_traverse.h"

namespace hvvr {

__global__ void frustumSetupKernel(const SimpleRayFrustum* __restrict__ tileCam, uint32_t tileCount,
                                   const SimpleRayFrustum* __restrict__ blockCam, uint32_t blockCount, matrix4x4 m,
                                   SimpleRayFrustum* __restrict__ tileWorld, FrustumPre* __restrict__ tilePre,
                                   FrustumPre* __restrict__ blockPre) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tileCount + blockCount)
        return;
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
    FrustumPre f;
    frustumPrecompute(o, d, f);
    if (isTile) {
        SimpleRayFrustum w;
        for (int v = 0; v < 4; v++) {
            w.origins[v] = o[v];
            w.directions[v] = d[v];
        }
        tileWorld[i] = w;
        tilePre[i] = f;
    } else {
        blockPre[i - tileCount] = f;
    }
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

// traverseBlocks (traverse_avx.h:316-410) + BlockFrame::sort (:293-314)
__global__ __launch_bounds__(64) void blockTraverseKernel(const BVHNode* __restrict__ nodes,
                                                          const FrustumPre* __restrict__ blockPre,
                                                          uint32_t blockCount, StackEntry* __restrict__ stacks,
                                                          uint32_t* __restrict__ stackSize) {
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= blockCount)
        return;
    const FrustumPre& f = blockPre[b];
    if (f.degenerate) {
        stackSize[b] = 0;
        return;
    }
    CullEntry cull[BLOCK_STACK_SIZE + 4];
    uint32_t top = 0;
    float tMin = 0.0f;
    uint32_t node = 0;
    int axis = f.distanceAxis, sizeAxis = f.distanceAxis & ~1;
    for (;;) {
        const BVHNode& n = nodes[node];
        uint32_t mask = testChildren(f, n);
        mask = testChildrenRefine(f, n, mask);
        if (mask) {
            uint32_t leaf = mask & n.leafMask;
            if (leaf) {
                mask &= ~leaf;
                uint32_t i = top;
                while (i > 0 && daz(cull[i - 1].bits) > 0.0f) {
                    cull[i] = cull[i - 1];
                    i--;
                }
                top++;
                cull[i].bits = leaf;
                cull[i].tMinTemp = tMin;
                cull[i].node = node;
            }
            for (int c = 0; c < 4; c++) {
                if (!(mask & (1u << c)))
                    continue;
                float dist = -n.bounds[axis][c];
                float size = ftz(n.bounds[sizeAxis][c] + n.bounds[sizeAxis + 1][c]);
                float key = daz(fbits(size));
                uint32_t i = top;
                while (i > 0 && daz(cull[i - 1].bits) > key) {
                    cull[i] = cull[i - 1];
                    i--;
                }
                cull[i].bits = fbits(size);
                cull[i].tMinTemp = dist;
                cull[i].node = node + n.data[1 + c];
                top++;
            }
        }
        if (top >= BLOCK_STACK_SIZE - 4 || top == 0 || cull[top - 1].bits <= 0xfu)
            break;
        top--;
        node = cull[top].node;
        tMin = cull[top].tMinTemp;
    }
    StackEntry* out = stacks + size_t(b) * BLOCK_STACK_SIZE;
    for (uint32_t t = 0; t < top; t++) {
        StackEntry e;
        e.tMin = cull[t].tMinTemp;
        uint32_t lm = cull[t].bits <= 0xfu ? cull[t].bits : 0u;
        e.nodeMask = (cull[t].node & STACK_NODE_MASK) | (lm << 28);
        uint32_t slot = t;
        while (slot > 0 && out[slot - 1].tMin < e.tMin) {
            out[slot] = out[slot - 1];
            slot--;
        }
        out[slot] = e;
    }
    stackSize[b] = top;
}

// traverseBlocks + BlockFrame::sort, one wave64 per block
__global__ __launch_bounds__(64) void blockTraverseWaveKernel(const BVHNode* __restrict__ nodes,
                                                              const FrustumPre* __restrict__ blockPre,
                                                              uint32_t blockCount, StackEntry* __restrict__ stacks,
                                                              uint32_t* __restrict__ stackSize) {
    __shared__ CullEntry cull[BLOCK_STACK_SIZE + 4];
    const uint32_t b = blockIdx.x;
    const uint32_t lane = threadIdx.x, c = lane & 3u, role = lane >> 2;
    WaveFrustum w;
    loadWaveFrustum(blockPre[b], role, w);
    if (w.degenerate) {
        if (lane == 0)
            stackSize[b] = 0;
        return;
    }
    const int axis = w.axis, sizeAxis = w.axis & ~1;
    uint32_t top = 0;
    float tMin = 0.0f;
    uint32_t node = 0;
    // insert e so that it lands below the run of top entries whose (DAZ) key exceeds `key`
    auto insert = [&](float key, const CullEntry& e) {
        int j = int(top) - 1 - int(lane);
        bool larger = j >= 0 && daz(cull[j >= 0 ? j : 0].bits) > key;
        uint64_t run = ~__ballot(larger);
        uint32_t n = run ? uint32_t(__builtin_ctzll(run)) : 64u;
        CullEntry moved;
        if (lane < n)
            moved = cull[j];
        waveLdsSync();
        if (lane < n)
            cull[j + 1] = moved;
        waveLdsSync();
        if (lane == 0)
            cull[top - n] = e;
        waveLdsSync();
        top++;
    };
    for (;;) {
        NodeLanes nl;
        loadNodeLanes(nodes, node, c, nl);
        const uint32_t leafMask = nodes[node].leafMask;
        uint32_t mask = waveTestChildren(w, nl, role);
        mask = waveTestRefine(w, nl, role, mask);
        if (mask) {
            uint32_t leaf = mask & leafMask;
            if (leaf) {
                mask &= ~leaf;
                CullEntry e;
                e.bits = leaf;
                e.tMinTemp = tMin;
                e.node = node;
                insert(0.0f, e);
            }
            for (int k = 0; k < 4; k++) {
                if (!(mask & (1u << k)))
                    continue;
                float size = ftz(nodes[node].bounds[sizeAxis][k] + nodes[node].bounds[sizeAxis + 1][k]);
                CullEntry e;
                e.bits = fbits(size);
                e.tMinTemp = -nodes[node].bounds[axis][k];
                e.node = node + nodes[node].data[1 + k];
                insert(daz(e.bits), e);
            }
        }
        if (top >= BLOCK_STACK_SIZE - 4 || top == 0 || cull[top - 1].bits <= 0xfu)
            break;
        top--;
        node = uniform(cull[top].node);
        tMin = cull[top].tMinTemp;
    }
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

// traverseTiles, one wave64 per tile (wave_traverse.h)
__global__ __launch_bounds__(64) void tileTraverseWaveKernel(const BVHNode* __restrict__ nodes,
                                                             const FrustumPre* __restrict__ tilePre,
                                                             const StackEntry* __restrict__ blockStacks,
                                                             const uint32_t* __restrict__ blockStackSize,
                                                             uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ offsets,
                                                             uint32_t* __restrict__ triOut, uint64_t capacity,
                                                             uint32_t* __restrict__ overflowTiles,
                                                             unsigned long long* __restrict__ counters,
                                                             uint32_t stackLimit, int emit) {
    __shared__ StackEntry stack[WAVE_STACK_SIZE];
    const uint32_t tile = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t block = tile / TILES_PER_BLOCK;
    const uint64_t outBase = emit ? offsets[tile] : 0;
    uint32_t count = 0;
    auto sink = [&](uint32_t t0, uint32_t t1) {
        uint32_t len = t1 - t0;
        if (emit) {
            for (uint32_t i = lane; i < len; i += 64) {
                uint64_t at = outBase + count + i;
                if (at < capacity)
                    triOut[at] = t0 + i;
            }
        }
        count += len;
    };
    uint32_t depth = 0;
    bool ok = waveTraverseTile(nodes, tilePre[tile], blockStacks + size_t(block) * BLOCK_STACK_SIZE,
                               blockStackSize[block], stack, stackLimit, lane, sink, &depth);
    if (lane == 0) {
        if (!ok) {
            overflowTiles[tile] = 1u; // re-run by the global-stack kernel
            count = 0;
        } else {
            if (emit && outBase + count > capacity)
                atomicOr(&counters[1], 1ull);
            atomicMax(&counters[2], (unsigned long long)depth);
        }
        if (!emit)
            counts[tile] = count;
    }
}

// traverseTiles (traverse_avx.h:412-527)
__global__ __launch_bounds__(64) void tileTraverseKernel(const BVHNode* __restrict__ nodes,
                                                         const FrustumPre* __restrict__ tilePre, uint32_t tileCount,
                                                         const StackEntry* __restrict__ blockStacks,
                                                         const uint32_t* __restrict__ blockStackSize,
                                                         StackEntry* __restrict__ scratch, uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ offsets,
                                                         uint32_t* __restrict__ triOut, uint64_t capacity,
                                                         const uint32_t* __restrict__ onlyTiles,
                                                         unsigned long long* __restrict__ counters, int emit) {
    uint32_t tile = blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= tileCount)
        return;
    if (onlyTiles && !onlyTiles[tile])
        return;
    uint32_t block = tile / TILES_PER_BLOCK;
    uint32_t n0 = blockStackSize[block];
    const FrustumPre& f = tilePre[tile];
    if (n0 == 0 || f.degenerate) {
        if (!emit)
            counts[tile] = 0;
        return;
    }
    auto S = [&](uint32_t slot) -> StackEntry& { return scratch[size_t(slot) * tileCount + tile]; };
    const StackEntry* bs = blockStacks + size_t(block) * BLOCK_STACK_SIZE;
    for (uint32_t i = 0; i < n0; i++)
        S(i) = bs[i];
    uint64_t outBase = emit ? offsets[tile] : 0;
    uint32_t count = 0;
    uint32_t top = n0 - 1;
    uint32_t maxTop = top;
    uint32_t node = S(top).nodeMask & STACK_NODE_MASK;
    uint32_t entryMask = S(top).nodeMask >> 28;
    bool overflow = false;
    int axis = f.distanceAxis;
    for (;;) {
        const BVHNode& n = nodes[node];
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
                    uint32_t t0 = n.data[k], t1 = n.data[k + 1];
                    if (emit) {
                        for (uint32_t t = t0; t < t1; t++) {
                            uint64_t at = outBase + count + (t - t0);
                            if (at < capacity)
                                triOut[at] = t;
                            else
                                overflow = true;
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
                bool first = true;
                for (int k = 0; k < 4; k++) {
                    if (!(mask & (1u << k)))
                        continue;
                    if (!first)
                        top++;
                    first = false;
                    if (top >= TILE_STACK_SIZE) {
                        overflow = true;
                        top = TILE_STACK_SIZE - 1;
                        break;
                    }
                    float dist = -n.bounds[axis][k];
                    uint32_t i = top;
                    while (i > 0 && S(i - 1).tMin < dist) {
                        S(i) = S(i - 1);
                        i--;
                    }
                    StackEntry e;
                    e.tMin = dist;
                    e.nodeMask = (node + n.data[1 + k]) & STACK_NODE_MASK;
                    S(i) = e;
                }
                if (overflow)
                    break;
                maxTop = top > maxTop ? top : maxTop;
                // the closest entry may be an older one (possibly a block-level leaf entry)
                StackEntry e = S(top);
                node = e.nodeMask & STACK_NODE_MASK;
                entryMask = e.nodeMask >> 28;
                pop = false;
            }
        }
        if (pop) {
            if (top == 0)
                break;
            top--;
            StackEntry e = S(top);
            node = e.nodeMask & STACK_NODE_MASK;
            entryMask = e.nodeMask >> 28;
        }
    }
    if (!emit)
        counts[tile] = count;
    if (overflow)
        atomicOr(&counters[1], 1ull);
    atomicMax(&counters[2], (unsigned long long)(maxTop + 1));
}

// ---- exclusive scan of tile counts: offsets[i] = sum(counts[0..i)), offsets[n] = total ----
enum { kScanThreads = 256, kScanPerThread = 4, kScanTile = kScanThreads * kScanPerThread };

__device__ inline uint32_t blockInclusiveScan(uint32_t v, uint32_t* lds) {
    uint32_t tid = threadIdx.x;
    lds[tid] = v;
    __syncthreads();
    for (uint32_t off = 1; off < kScanThreads; off <<= 1) {
        uint32_t add = tid >= off ? lds[tid - off] : 0u;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    return lds[tid];
}

__global__ __launch_bounds__(kScanThreads) void scanReduceKernel(const uint32_t* __restrict__ in, uint32_t n,
                                                                 uint32_t* __restrict__ partials) {
    __shared__ uint32_t lds[kScanThreads];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    uint32_t s = 0;
    for (int k = 0; k < kScanPerThread; k++)
        s += base + k < n ? in[base + k] : 0u;
    uint32_t inc = blockInclusiveScan(s, lds);
    if (threadIdx.x == kScanThreads - 1)
        partials[blockIdx.x] = inc;
}

__global__ __launch_bounds__(kScanThreads) void scanPartialsKernel(uint32_t* partials, uint32_t count,
                                                                   uint32_t* __restrict__ totalOut) {
    __shared__ uint32_t lds[kScanThreads];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < count; base += kScanThreads) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < count ? partials[i] : 0u;
        uint32_t inc = blockInclusiveScan(v, lds);
        if (i < count)
            partials[i] = carry + inc - v;
        uint32_t sum = lds[kScanThreads - 1];
        __syncthreads();
        carry += sum;
    }
    if (threadIdx.x == 0)
        *totalOut = carry;
}

__global__ __launch_bounds__(kScanThreads) void scanApplyKernel(const uint32_t* __restrict__ in, uint32_t n,
                                                                const uint32_t* __restrict__ partials,
                                                                uint32_t* __restrict__ out) {
    __shared__ uint32_t lds[kScanThreads];
    uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    uint32_t vals[kScanPerThread];
    uint32_t s = 0;
    for (int k = 0; k < kScanPerThread; k++) {
        vals[k] = base + k < n ? in[base + k] : 0u;
        s += vals[k];
    }
    uint32_t inc = blockInclusiveScan(s, lds);
    uint32_t run = partials[blockIdx.x] + inc - s;
    for (int k = 0; k < kScanPerThread; k++) {
        if (base + k < n)
            out[base + k] = run;
        run += vals[k];
    }
}

void launchFrustumSetup(hvvr_gpu_view* v, const matrix4x4& m, hipStream_t stream) {
    uint32_t n = v->tileCount + v->blockCount;
    frustumSetupKernel<<<(n + 127) / 128, 128, 0, stream>>>(v->tileFrustaCam.ptr, v->tileCount, v->blockFrustaCam.ptr,
                                                            v->blockCount, m, v->tileFrustaWorld.ptr, v->tilePre.ptr,
                                                            v->blockPre.ptr);
}

void launchBlockTraverse(hvvr_gpu_view* v, const hvvr_gpu_scene* s, hipStream_t stream) {
    blockTraverseWaveKernel<<<v->blockCount, 64, 0, stream>>>(s->nodes.ptr, v->blockPre.ptr, v->blockCount,
                                                              v->blockStacks.ptr, v->blockStackSize.ptr);
}

void launchTileTraverse(hvvr_gpu_view* v, const hvvr_gpu_scene* s, bool emit, hipStream_t stream) {
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(v->counters.ptr);
    tileTraverseWaveKernel<<<v->tileCount, 64, 0, stream>>>(
        s->nodes.ptr, v->tilePre.ptr, v->blockStacks.ptr, v->blockStackSize.ptr, v->tileCounts.ptr,
        v->tileOffsets.ptr, v->triIndices.ptr, v->listCapacity, v->overflowTiles.ptr, counters, v->waveStackLimit,
        emit ? 1 : 0);
    // tiles whose stack outgrew LDS: full-depth global-memory traversal, same results
    tileTraverseKernel<<<(v->tileCount + 63) / 64, 64, 0, stream>>>(
        s->nodes.ptr, v->tilePre.ptr, v->tileCount, v->blockStacks.ptr, v->blockStackSize.ptr,
        v->tileStackScratch.ptr, v->tileCounts.ptr, v->tileOffsets.ptr, v->triIndices.ptr, v->listCapacity,
        v->overflowTiles.ptr, counters, emit ? 1 : 0);
}

void launchScan(hvvr_gpu_view* v, hipStream_t stream) {
    uint32_t n = v->tileCount;
    uint32_t groups = (n + kScanTile - 1) / kScanTile;
    scanReduceKernel<<<groups, kScanThreads, 0, stream>>>(v->tileCounts.ptr, n, v->scanPartials.ptr);
    scanPartialsKernel<<<1, kScanThreads, 0, stream>>>(v->scanPartials.ptr, groups, v->tileOffsets.ptr + n);
    scanApplyKernel<<<groups, kScanThreads, 0, stream>>>(v->tileCounts.ptr, n, v->scanPartials.ptr,
                                                         v->tileOffsets.ptr);
}

} // namespace hvvr
