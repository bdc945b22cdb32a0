// CPU baseline of the render path. TEST INFRASTRUCTURE: only bench.py's cpu_baseline leg
// and tests/ load it (through oracle/pyoracle.py); the product never does.
//
// It is the reference's CPU stage as the reference runs it, beside a scalar restatement of
// its GPU kernels:
//   * per block, the frustum precompute (the oracle's or_frustum_setup, a restatement of
//     frustumTransform + the Frustum constructor, traversal.h:38-47, traverse_avx.cpp:22-211);
//   * the AVX2 traversal: Frustum::testBVHNodeChildren / ...Refine on __m256 lanes laid out
//     as the reference's (traverse_avx.h:104-217), traverseBlocks with its size-sorted cull
//     stack (:316-410), BlockFrame::sort (:293-314) and traverseTiles with its 4096-entry
//     distance-sorted stack and single-child shortcut (:412-527), one tile after another
//     from the block's sorted stack (traversal.cpp:99-211);
//   * then intersect and resolve of the block's samples through the oracle's scalar C
//     transcriptions of IntersectKernel and ResolveKernel (or_intersect_view, or_resolve_view).
// Work is split into tasks of 2 blocks (traversal.cpp:249-254) taken by worker threads
// (std::thread; the reference sizes its pool by hardware_concurrency, raycaster.cpp:33-37)
// that run the traversal with FTZ and DAZ set in MXCSR, as the reference's workers do
// (raycaster.cpp:20-28): a block-stack leaf entry, whose float key slot holds a 4-bit mask
// (a denormal), then compares as 0. Intersect and resolve run with IEEE denormals, as the
// GPU kernels they transcribe do.
//
// The traversal's output equals the scalar oracle's (or_build_tile_lists) list for list;
// tests/test_cpu_baseline.py checks that on the benchmark scenes.

#include "hvvr_oracle.h"

#include <immintrin.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

namespace {

// The reference's Frustum test data, swizzled for 8 lanes: lanes 0-3 take the four children's
// max bounds, lanes 4-7 their negated min bounds (the BVHNode rows xMax | xNegMin, ...).
struct alignas(32) AvxFrustum {
    __m256 plane[4][3]; // n-vertex selectors: plane normal components for max | negMin lanes
    __m128 planeDist[4];
    __m256 projX, projY, projZ;      // world-axis projections: lo | hi
    __m256 edge[2][3];               // 12 frustum-edge x world-axis axes, 8 per group
    __m256 edgeMin[2], edgeMax[2];   // the frustum's interval on each axis
    int axis;                        // distance-estimate row: xMax, xNegMin, yMax, ...
    int degenerate;
};

void avxFrustum(const o_frustum& f, AvxFrustum& a) {
    for (int p = 0; p < 4; p++) {
        for (int c = 0; c < 3; c++) {
            const float n = f.planeNeg[p][c], s = f.planePos[p][c];
            a.plane[p][c] = _mm256_setr_ps(n, n, n, n, s, s, s, s);
        }
        a.planeDist[p] = _mm_set1_ps(f.planeNeg[p][3]);
    }
    a.projX = _mm256_setr_ps(f.projLo[0], f.projLo[0], f.projLo[0], f.projLo[0], f.projHi[0], f.projHi[0],
                             f.projHi[0], f.projHi[0]);
    a.projY = _mm256_setr_ps(f.projLo[1], f.projLo[1], f.projLo[1], f.projLo[1], f.projHi[1], f.projHi[1],
                             f.projHi[1], f.projHi[1]);
    a.projZ = _mm256_setr_ps(f.projLo[2], f.projLo[2], f.projLo[2], f.projLo[2], f.projHi[2], f.projHi[2],
                             f.projHi[2], f.projHi[2]);
    alignas(32) float e[3][16] = {}, mn[16] = {}, mx[16] = {};
    for (int k = 0; k < 12; k++) {
        for (int c = 0; c < 3; c++)
            e[c][k] = f.edge[k][c];
        mn[k] = f.edgeMin[k];
        mx[k] = f.edgeMax[k];
    }
    for (int g = 0; g < 2; g++) {
        for (int c = 0; c < 3; c++)
            a.edge[g][c] = _mm256_load_ps(e[c] + 8 * g);
        a.edgeMin[g] = _mm256_load_ps(mn + 8 * g);
        a.edgeMax[g] = _mm256_load_ps(mx + 8 * g);
    }
    a.axis = f.axis;
    a.degenerate = f.degenerate;
}

// testBVHNodeChildren (traverse_avx.h:104-164)
inline uint32_t testChildren(const AvxFrustum& f, const o_node& n) {
    const __m256 X = _mm256_loadu_ps(&n.b[0][0]), Y = _mm256_loadu_ps(&n.b[2][0]), Z = _mm256_loadu_ps(&n.b[4][0]);
    __m128 pass = _mm_castsi128_ps(_mm_set1_epi32(-1));
    for (int p = 0; p < 4; p++) {
        const __m256 sel =
            _mm256_fmadd_ps(f.plane[p][0], X, _mm256_fmadd_ps(f.plane[p][1], Y, _mm256_mul_ps(f.plane[p][2], Z)));
        const __m128 d = _mm_add_ps(_mm256_castps256_ps128(sel), _mm256_extractf128_ps(sel, 1));
        pass = _mm_and_ps(pass, _mm_cmp_ps(d, f.planeDist[p], _CMP_LE_OQ));
    }
    uint32_t result = uint32_t(_mm_movemask_ps(pass));
    if (!result)
        return 0;
    const __m256 in = _mm256_and_ps(_mm256_and_ps(_mm256_cmp_ps(X, f.projX, _CMP_GE_OQ),
                                                  _mm256_cmp_ps(Y, f.projY, _CMP_GE_OQ)),
                                    _mm256_cmp_ps(Z, f.projZ, _CMP_GE_OQ));
    uint32_t m = uint32_t(_mm256_movemask_ps(in));
    m &= m >> 4;
    return result & m;
}

// testBVHNodeChildrenRefine (traverse_avx.h:166-217)
inline uint32_t testChildrenRefine(const AvxFrustum& f, const o_node& n, uint32_t prev) {
    uint32_t result = prev;
    const __m256 zero = _mm256_setzero_ps(), sign = _mm256_set1_ps(-0.0f);
    for (int g = 0; g < 2; g++) {
        const __m256 ex = f.edge[g][0], ey = f.edge[g][1], ez = f.edge[g][2];
        const __m256 px = _mm256_cmp_ps(ex, zero, _CMP_GT_OQ), py = _mm256_cmp_ps(ey, zero, _CMP_GT_OQ),
                     pz = _mm256_cmp_ps(ez, zero, _CMP_GT_OQ);
        const uint32_t lanes = g == 0 ? 0xffu : 0x0fu; // 12 axes: the second group's last 4 lanes are padding
        for (int c = 0; c < 4; c++) {
            const __m256 xM = _mm256_broadcast_ss(&n.b[0][c]), xN = _mm256_xor_ps(_mm256_broadcast_ss(&n.b[1][c]), sign);
            const __m256 yM = _mm256_broadcast_ss(&n.b[2][c]), yN = _mm256_xor_ps(_mm256_broadcast_ss(&n.b[3][c]), sign);
            const __m256 zM = _mm256_broadcast_ss(&n.b[4][c]), zN = _mm256_xor_ps(_mm256_broadcast_ss(&n.b[5][c]), sign);
            // p-vertex: the max bound where the axis component is positive, else the min bound;
            // the n-vertex the other way round
            const __m256 pX = _mm256_blendv_ps(xN, xM, px), pY = _mm256_blendv_ps(yN, yM, py),
                         pZ = _mm256_blendv_ps(zN, zM, pz);
            const __m256 nX = _mm256_blendv_ps(xM, xN, px), nY = _mm256_blendv_ps(yM, yN, py),
                         nZ = _mm256_blendv_ps(zM, zN, pz);
            const __m256 maxBox = _mm256_fmadd_ps(ex, pX, _mm256_fmadd_ps(ey, pY, _mm256_mul_ps(ez, pZ)));
            const __m256 minBox = _mm256_fmadd_ps(ex, nX, _mm256_fmadd_ps(ey, nY, _mm256_mul_ps(ez, nZ)));
            const __m256 sep = _mm256_or_ps(_mm256_cmp_ps(maxBox, f.edgeMin[g], _CMP_LT_OQ),
                                            _mm256_cmp_ps(f.edgeMax[g], minBox, _CMP_LT_OQ));
            if (uint32_t(_mm256_movemask_ps(sep)) & lanes)
                result &= ~(1u << c);
        }
    }
    return result;
}

// A stack entry: 16 bytes, the key in the first float so one SSE compare reads it.
struct alignas(16) Entry {
    union {
        float key; // cull stack: size estimate (internal) | tile stack: distance estimate
        uint32_t keyBits; // cull stack: the leaf mask of a leaf entry (a denormal under DAZ)
    };
    float tMin;    // cull stack: distance estimate
    uint32_t mask; // sorted / tile stack: 0 for an internal node, else the leaf children's mask
    uint32_t node;
};

enum : uint32_t { kBlockStack = 64, kTileStack = 4096, kNodeBits = 0x0fffffffu };

inline bool greaterSS(float a, float b) { return _mm_comigt_ss(_mm_set_ss(a), _mm_set_ss(b)); }
inline bool lessSS(float a, float b) { return _mm_comilt_ss(_mm_set_ss(a), _mm_set_ss(b)); }

// traverseBlocks + BlockFrame::sort: the block's distance-sorted stack (closest on top);
// returns its size. Index 0 of `cull`/`sorted` is the guard band.
uint32_t traverseBlock(const o_node* nodes, const AvxFrustum& f, Entry* cull, Entry* sorted) {
    if (f.degenerate)
        return 0;
    cull[0].key = -1.0f;
    Entry* cs = cull + 1;
    uint32_t top = 0;
    uint32_t node = 0;
    float tMin = 0.0f;
    const int sizeRow = f.axis & ~1;
    for (;;) {
        const o_node& n = nodes[node];
        uint32_t mask = testChildrenRefine(f, n, testChildren(f, n));
        if (mask) {
            const uint32_t leaf = mask & n.leafMask;
            if (leaf) {
                // leaf children go to the bottom of the stack, above earlier leaf entries
                mask &= ~leaf;
                uint32_t i = top;
                while (greaterSS(cs[int(i) - 1].key, 0.0f)) {
                    cs[i] = cs[i - 1];
                    i--;
                }
                top++;
                cs[i].keyBits = leaf;
                cs[i].tMin = tMin;
                cs[i].node = node;
            }
            if (mask) {
                float dist[4], size[4];
                for (int c = 0; c < 4; c++) {
                    dist[c] = -n.b[f.axis][c];
                    size[c] = n.b[sizeRow][c] + n.b[sizeRow + 1][c];
                }
                // internal children sorted by size, the largest on top
                uint32_t m = mask;
                bool first = true;
                while (m) {
                    if (!first)
                        top++;
                    first = false;
                    const uint32_t k = uint32_t(__builtin_ctz(m));
                    uint32_t i = top;
                    while (greaterSS(cs[int(i) - 1].key, size[k])) {
                        cs[i] = cs[i - 1];
                        i--;
                    }
                    cs[i].key = size[k];
                    cs[i].tMin = dist[k];
                    cs[i].node = node + n.data[1 + k];
                    m &= m - 1;
                }
                top++;
            }
        }
        // POP: stop near capacity, when empty, or when only leaf entries remain
        if (top >= kBlockStack - 4 || top == 0 || cs[top - 1].keyBits <= 0xfu)
            break;
        --top;
        node = cs[top].node;
        tMin = cs[top].tMin;
    }
    // BlockFrame::sort: re-insert by distance estimate, the closest on top
    sorted[0].key = std::numeric_limits<float>::infinity();
    Entry* ss = sorted + 1;
    for (uint32_t t = 0; t < top; t++) {
        Entry e;
        e.key = cs[t].tMin;
        e.tMin = 0.0f;
        e.mask = cs[t].keyBits > 0xfu ? 0u : cs[t].keyBits;
        e.node = cs[t].node;
        int slot = int(t);
        while (lessSS(ss[slot - 1].key, e.key)) {
            ss[slot] = ss[slot - 1];
            slot--;
        }
        ss[slot] = e;
    }
    return top;
}

// traverseTiles: the tile's list from a copy of the block's sorted stack; returns the
// count, or -1 when the stack or the output would overflow
int64_t traverseTile(const o_node* nodes, const Entry* blockSorted, uint32_t size, const AvxFrustum& f, Entry* frame,
                     std::vector<uint32_t>& out) {
    if (f.degenerate || size == 0)
        return 0;
    frame[0].key = std::numeric_limits<float>::infinity();
    Entry* st = frame + 1;
    memcpy(st, blockSorted + 1, sizeof(Entry) * size);
    int64_t count = 0;
    uint32_t top = size - 1;
    uint32_t node = st[top].node;
    for (;;) {
        const o_node& n = nodes[node];
        uint32_t mask = testChildren(f, n);
        bool popNext = true;
        if (mask) {
            uint32_t leaf = mask & n.leafMask;
            if (leaf) {
                mask = testChildrenRefine(f, n, mask);
                leaf = mask & n.leafMask;
                for (uint32_t e = leaf; e; e &= e - 1) {
                    const uint32_t k = uint32_t(__builtin_ctz(e));
                    for (uint32_t t = n.data[k]; t != n.data[k + 1]; t++) {
                        out.push_back(t);
                        count++;
                    }
                }
            }
            mask &= ~n.leafMask;
            // a leaf entry of the block stack does not re-spawn its internal siblings
            if (mask && st[top].mask == 0) {
                if ((mask & (mask - 1)) == 0) {
                    node = node + n.data[1 + __builtin_ctz(mask)]; // one hit child: test it directly
                    continue;
                }
                float dist[4];
                for (int c = 0; c < 4; c++)
                    dist[c] = -n.b[f.axis][c];
                bool first = true;
                for (uint32_t m = mask; m; m &= m - 1) {
                    if (!first)
                        top++;
                    first = false;
                    if (top >= kTileStack)
                        return -1;
                    const uint32_t k = uint32_t(__builtin_ctz(m));
                    int i = int(top);
                    while (lessSS(st[i - 1].key, dist[k])) {
                        st[i] = st[i - 1];
                        i--;
                    }
                    st[i].key = dist[k];
                    st[i].mask = 0;
                    st[i].node = node + n.data[1 + k];
                }
                node = st[top].node;
                popNext = false;
            }
        }
        if (popNext) {
            if (top == 0)
                break;
            top--;
            node = st[top].node;
        }
    }
    return count;
}

// FTZ and DAZ for the traversal, restored afterwards
struct FlushDenormals {
    unsigned saved;
    FlushDenormals() : saved(_mm_getcsr()) { _mm_setcsr(saved | 0x8040u); }
    ~FlushDenormals() { _mm_setcsr(saved); }
};

// Lists of one block's 64 tiles; counts[64]; returns the total or -1 on overflow.
int64_t blockLists(const o_node* nodes, const o_frustum& blockPre, const o_frustum* tilePre, Entry* cull,
                   Entry* sorted, Entry* tileFrame, uint32_t* counts, std::vector<uint32_t>& lists) {
    FlushDenormals ftz;
    AvxFrustum bf;
    avxFrustum(blockPre, bf);
    const uint32_t size = traverseBlock(nodes, bf, cull, sorted);
    int64_t total = 0;
    for (uint32_t t = 0; t < 64; t++) {
        if (!size) {
            counts[t] = 0;
            continue;
        }
        AvxFrustum tf;
        avxFrustum(tilePre[t], tf);
        const int64_t c = traverseTile(nodes, sorted, size, tf, tileFrame, lists);
        if (c < 0)
            return -1;
        counts[t] = uint32_t(c);
        total += c;
    }
    return total;
}

struct Scratch {
    std::vector<Entry> cull, sorted, tile;
    Scratch() : cull(kBlockStack + 8), sorted(kBlockStack + 8), tile(kTileStack + 8) {}
};

} // namespace

extern "C" {

/* or_build_tile_lists through the AVX2 traversal (same arguments and result). */
int64_t cb_build_tile_lists(const o_node* nodes, const o_frustum* blockPre, const o_frustum* tilePre,
                            uint32_t blockCount, uint32_t* counts, uint32_t* out, uint64_t cap) {
    Scratch s;
    std::vector<uint32_t> lists;
    uint64_t total = 0;
    for (uint32_t b = 0; b < blockCount; b++) {
        lists.clear();
        const int64_t n = blockLists(nodes, blockPre[b], tilePre + size_t(b) * 64, s.cull.data(), s.sorted.data(),
                                     s.tile.data(), counts + size_t(b) * 64, lists);
        if (n < 0)
            return -1;
        if (out) {
            if (total + lists.size() > cap)
                return -1;
            memcpy(out + total, lists.data(), lists.size() * 4);
        }
        total += lists.size();
    }
    return int64_t(total);
}

typedef struct {
    const o_beam* beams;            /* [blockCount * 8192] camera space */
    const o_sfrustum* tileFrusta;   /* [blockCount * 64] camera space */
    const o_sfrustum* blockFrusta;  /* [blockCount] */
    uint32_t blockCount, validSampleCount;
    float cameraToWorld[16];
} cb_view;

typedef struct {
    double seconds;             /* wall time of the timed run */
    uint64_t subsamples;        /* valid samples x A of the completed tasks */
    uint64_t blocks;            /* blocks rendered */
    double frustumSeconds, traverseSeconds, intersectSeconds, resolveSeconds; /* summed over threads */
    uint32_t threads;
    uint32_t hardwareConcurrency;
    int32_t status;             /* 0, or -1 when a tile's list overflowed its stack */
} cb_stats;

/* Renders the views' blocks in tasks of 2 (one pass over every block in a seeded order,
 * then again) on `threads` worker threads until `budgetSeconds` of wall time have passed;
 * tasks that finish inside the budget are counted. Scene arrays are the refit ones (world
 * space). */
int cb_run(const o_node* nodes, const o_tri* tris, const o_trishade* shade, const o_vertex* worldVerts,
           const o_material* materials, const o_lightenv* env, const cb_view* views, uint32_t viewCount,
           const ov2* offsets, uint32_t A, uint32_t threads, double budgetSeconds, uint32_t seed, cb_stats* stats) {
    typedef std::chrono::steady_clock clk;
    struct Task {
        uint32_t view, block, count;
    };
    std::vector<Task> tasks;
    for (uint32_t v = 0; v < viewCount; v++)
        for (uint32_t b = 0; b < views[v].blockCount; b += 2)
            tasks.push_back({v, b, std::min(2u, views[v].blockCount - b)});
    std::mt19937 rng(seed);
    std::shuffle(tasks.begin(), tasks.end(), rng);
    memset(stats, 0, sizeof(*stats));
    stats->hardwareConcurrency = std::thread::hardware_concurrency();
    if (threads == 0)
        threads = std::max(1u, stats->hardwareConcurrency);
    stats->threads = threads;
    std::atomic<uint64_t> next(0), subsamples(0), blocks(0);
    std::atomic<int> failed(0);
    std::vector<double> tF(threads), tT(threads), tI(threads), tR(threads);
    const auto start = clk::now();
    const auto deadline = start + std::chrono::duration_cast<clk::duration>(std::chrono::duration<double>(budgetSeconds));
    auto worker = [&](uint32_t w) {
        Scratch s;
        std::vector<o_sfrustum> world(64);
        std::vector<o_frustum> tilePre(64);
        o_frustum blockPre;
        std::vector<uint32_t> counts(64), lists, ids(size_t(8192) * A), colors(8192);
        std::vector<float> depth(size_t(8192) * A);
        auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
        while (!failed.load()) {
            const uint64_t i = next.fetch_add(1);
            const Task& task = tasks[i % tasks.size()];
            const cb_view& v = views[task.view];
            uint64_t done = 0;
            for (uint32_t b = task.block; b < task.block + task.count; b++) {
                auto t0 = clk::now();
                or_frustum_setup(v.tileFrusta + size_t(b) * 64, 64, v.cameraToWorld, world.data(), tilePre.data());
                or_frustum_setup(v.blockFrusta + b, 1, v.cameraToWorld, nullptr, &blockPre);
                auto t1 = clk::now();
                lists.clear();
                if (blockLists(nodes, blockPre, tilePre.data(), s.cull.data(), s.sorted.data(), s.tile.data(),
                               counts.data(), lists) < 0) {
                    failed.store(1);
                    return;
                }
                auto t2 = clk::now();
                const o_beam* beams = v.beams + size_t(b) * 8192;
                uint64_t surv = 0;
                or_intersect_view(beams, 8192, world.data(), 64, v.cameraToWorld, offsets, A, counts.data(),
                                  lists.data(), tris, ids.data(), depth.data(), &surv);
                auto t3 = clk::now();
                or_resolve_view(beams, 8192, v.cameraToWorld, offsets, A, ids.data(), tris, shade, worldVerts, materials,
                                env, 0, colors.data(), nullptr);
                auto t4 = clk::now();
                tF[w] += secs(t0, t1);
                tT[w] += secs(t1, t2);
                tI[w] += secs(t2, t3);
                tR[w] += secs(t3, t4);
                const uint64_t first = uint64_t(b) * 8192;
                done += first < v.validSampleCount ? std::min<uint64_t>(8192, v.validSampleCount - first) : 0;
            }
            if (clk::now() > deadline)
                return; // this task ended past the budget: not counted
            subsamples += done * A;
            blocks += task.count;
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t w = 0; w < threads; w++)
        pool.emplace_back(worker, w);
    for (auto& t : pool)
        t.join();
    stats->seconds = std::chrono::duration<double>(clk::now() - start).count();
    stats->seconds = std::min(stats->seconds, budgetSeconds);
    stats->subsamples = subsamples.load();
    stats->blocks = blocks.load();
    stats->frustumSeconds = std::accumulate(tF.begin(), tF.end(), 0.0);
    stats->traverseSeconds = std::accumulate(tT.begin(), tT.end(), 0.0);
    stats->intersectSeconds = std::accumulate(tI.begin(), tI.end(), 0.0);
    stats->resolveSeconds = std::accumulate(tR.begin(), tR.end(), 0.0);
    stats->status = failed.load() ? -1 : 0;
    return stats->status;
}

} // extern "C"
