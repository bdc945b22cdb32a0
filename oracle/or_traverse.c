/* Oracle: frustum setup and frustum-BVH traversal (test infrastructure only).
 *
 * Semantics follow the reference's default traversal build (TRAVERSAL_MODE ==
 * TRAVERSAL_AVX, raycaster/traversal.h:230-232), restated lane by lane in scalar C:
 *   - Frustum ctor            raycaster/traverse_avx.cpp:22-211
 *   - testBVHNodeChildren     raycaster/traverse_avx.h:104-164
 *   - ...Refine               raycaster/traverse_avx.h:166-217
 *   - traverseBlocks          raycaster/traverse_avx.h:316-410
 *   - BlockFrame::sort        raycaster/traverse_avx.h:293-314
 *   - traverseTiles           raycaster/traverse_avx.h:412-527 (incl. the single-child
 *                             fast path, :489-496, which the scalar traverse_ref lacks)
 *   - per-tile driver         raycaster/traversal.cpp:99-211
 * The reference runs with DAZ, under which a block-stack leaf entry (whose float
 * "tDelta" slot holds a 4-bit leaf mask, i.e. a denormal) compares as 0; that is
 * emulated explicitly with daz() below.
 */
#include "or_math.h"

#define ROOT 0u
#define NODE_BITS 0x0fffffffu
#define BLOCK_STACK 64
#define TILE_STACK 4096

/* frustumTransform (raycaster/traversal.h:38-47) + Frustum ctor */
static void frustum_pre(const ov3 o[4], const ov3 d[4], o_frustum* f) {
    memset(f, 0, sizeof(*f));
    int degenerate = isinf(o[0].x);
    float plane[4][4];
    ov3 avg = v3(0.0f, 0.0f, 0.0f);
    for (int p = 0; p < 4; p++) {
        if (degenerate) {
            plane[p][0] = plane[p][1] = plane[p][2] = 0.0f;
            plane[p][3] = -INFINITY;
        } else {
            ov3 n = cross3(d[p], d[(p + 1) % 4]);
            plane[p][0] = n.x;
            plane[p][1] = n.y;
            plane[p][2] = n.z;
            plane[p][3] = dot3(o[p], n);
        }
        avg = add3(avg, d[p]);
    }
    f->degenerate = degenerate;
    /* distance-estimate axis: index into node bounds (xMax, xNegMin, yMax, ...) */
    if (fabsf(avg.x) > fabsf(avg.y)) {
        if (fabsf(avg.x) > fabsf(avg.z)) f->axis = avg.x < 0 ? 0 : 1;
        else f->axis = avg.z < 0 ? 4 : 5;
    } else {
        if (fabsf(avg.y) > fabsf(avg.z)) f->axis = avg.y < 0 ? 2 : 3;
        else f->axis = avg.z < 0 ? 4 : 5;
    }
    for (int p = 0; p < 4; p++) {
        for (int a = 0; a < 3; a++) {
            float c = plane[p][a];
            f->planeNeg[p][a] = c < 0.0f ? c : 0.0f;
            f->planePos[p][a] = c > 0.0f ? -c : 0.0f;
        }
        f->planeNeg[p][3] = plane[p][3];
        f->planePos[p][3] = 0.0f;
    }
    /* projection onto the world axes, extended to FLT_MAX-scaled infinity */
    for (int a = 0; a < 3; a++) {
        float lo = 0.0f, hi = 0.0f;
        for (int v = 0; v < 4; v++) {
            float oa = comp3(o[v], a), da = comp3(d[v], a);
            float l = -oa + FLT_MAX * maxps(0.0f, -da);
            float h = oa + FLT_MAX * maxps(0.0f, da);
            lo = v == 0 ? l : maxps(lo, l);
            hi = v == 0 ? h : maxps(hi, h);
        }
        f->projLo[a] = -lo;
        f->projHi[a] = -hi;
    }
    /* frustum-edge x world-axis separating axes */
    const float eps = 0.000001f;
    for (int v = 0; v < 4; v++) {
        ov3 ax[3] = {v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 1)};
        for (int a = 0; a < 3; a++) {
            ov3 e = cross3(d[v], ax[a]);
            f->edge[v * 3 + a][0] = e.x;
            f->edge[v * 3 + a][1] = e.y;
            f->edge[v * 3 + a][2] = e.z;
        }
    }
    for (int e = 0; e < 12; e++) {
        float ex = f->edge[e][0], ey = f->edge[e][1], ez = f->edge[e][2];
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int v = 0; v < 4; v++) {
            float d0 = fmaf(ex, o[v].x, fmaf(ey, o[v].y, ez * o[v].z));
            mn = minps(mn, d0);
            mx = maxps(mx, d0);
            float d1 = fmaf(ex, d[v].x, fmaf(ey, d[v].y, ez * d[v].z));
            mn = minps(d1 + eps, 0.0f) * FLT_MAX + mn;
            mx = maxps(d1 - eps, 0.0f) * FLT_MAX + mx;
        }
        f->edgeMin[e] = mn;
        f->edgeMax[e] = mx;
    }
}

/* planeCullsFrustum (render.cpp:271-279): all 4 corner rays, at the origin and 10^4 along
 * the ray, on the positive side of the plane (normal p[0..2], dist p[3]) */
static int plane_culls(const float* p, const ov3 o[4], const ov3 d[4]) {
    ov3 n = v3(p[0], p[1], p[2]);
    int allOut = 1;
    for (int v = 0; v < 4; v++) {
        allOut = allOut && dot3(n, o[v]) > p[3];
        allOut = allOut && dot3(n, add3(o[v], scale3(d[v], 10000.0f))) > p[3];
    }
    return allOut;
}

void or_frustum_setup(const o_sfrustum* in, uint32_t n, const float m[16], o_sfrustum* worldOut, o_frustum* preOut) {
    or_frustum_setup_cull(in, n, m, NULL, worldOut, preOut);
}

void or_frustum_setup_cull(const o_sfrustum* in, uint32_t n, const float m[16], const float* planes,
                           o_sfrustum* worldOut, o_frustum* preOut) {
    for (uint32_t i = 0; i < n; i++) {
        ov3 o[4], d[4];
        /* an infinite first origin marks a culled (degenerate) frustum (render.cpp:331-338) */
        int culled = isinf(in[i].o[0].x);
        for (int v = 0; v < 4; v++) {
            float t[4];
            m4mul(m, in[i].o[v].x, in[i].o[v].y, in[i].o[v].z, 1.0f, t);
            o[v] = culled ? v3(INFINITY, INFINITY, INFINITY) : v3(t[0], t[1], t[2]);
            d[v] = culled ? v3(0.0f, 0.0f, 0.0f) : m3mul(m, in[i].d[v]);
        }
        /* blockTransformAndCullTask (render.cpp:311-340) */
        if (planes && !isinf(planes[3]) && !culled) {
            for (int k = 0; k < 4; k++)
                culled = culled || plane_culls(planes + 4 * k, o, d);
            if (culled)
                for (int v = 0; v < 4; v++) {
                    o[v] = v3(INFINITY, INFINITY, INFINITY);
                    d[v] = v3(0.0f, 0.0f, 0.0f);
                }
        }
        if (worldOut) {
            memcpy(worldOut[i].o, o, sizeof(o));
            memcpy(worldOut[i].d, d, sizeof(d));
        }
        if (preOut)
            frustum_pre(o, d, &preOut[i]);
    }
}

uint32_t or_test_children(const o_frustum* f, const o_node* n) {
    uint32_t result = 0xf;
    uint32_t planeMask = 0;
    for (int c = 0; c < 4; c++) {
        int pass = 1;
        for (int p = 0; p < 4; p++) {
            const float* ng = f->planeNeg[p];
            const float* ps = f->planePos[p];
            float dmax = fmaf(ng[0], n->b[0][c], fmaf(ng[1], n->b[2][c], ng[2] * n->b[4][c]));
            float dneg = fmaf(ps[0], n->b[1][c], fmaf(ps[1], n->b[3][c], ps[2] * n->b[5][c]));
            float dist = dmax + dneg;
            if (!(dist <= ng[3])) pass = 0;
        }
        if (pass) planeMask |= 1u << c;
    }
    result &= planeMask;
    if (!result) return 0;
    uint32_t projMask = 0;
    for (int c = 0; c < 4; c++) {
        int pass = 1;
        for (int a = 0; a < 3; a++)
            if (!(n->b[2 * a][c] >= f->projLo[a] && n->b[2 * a + 1][c] >= f->projHi[a])) pass = 0;
        if (pass) projMask |= 1u << c;
    }
    return result & projMask;
}

uint32_t or_test_children_refine(const o_frustum* f, const o_node* n, uint32_t prev) {
    uint32_t result = prev;
    for (int c = 0; c < 4; c++) {
        for (int e = 0; e < 12; e++) {
            float ex = f->edge[e][0], ey = f->edge[e][1], ez = f->edge[e][2];
            float px = ex > 0.0f ? n->b[0][c] : -n->b[1][c];
            float py = ey > 0.0f ? n->b[2][c] : -n->b[3][c];
            float pz = ez > 0.0f ? n->b[4][c] : -n->b[5][c];
            float nx = ex > 0.0f ? -n->b[1][c] : n->b[0][c];
            float ny = ey > 0.0f ? -n->b[3][c] : n->b[2][c];
            float nz = ez > 0.0f ? -n->b[5][c] : n->b[4][c];
            float maxBox = fmaf(ex, px, fmaf(ey, py, ez * pz));
            float minBox = fmaf(ex, nx, fmaf(ey, ny, ez * nz));
            if (maxBox < f->edgeMin[e] || f->edgeMax[e] < minBox) result &= ~(1u << c);
        }
    }
    return result;
}

/* DAZ view of a float stored as raw bits (denormals and zeros compare as 0) */
static inline float daz(uint32_t bits) {
    return (bits & 0x7f800000u) == 0 ? 0.0f : bitsf(bits);
}
/* FTZ of an arithmetic result */
static inline float ftz(float x) {
    return (fbits(x) & 0x7f800000u) == 0 ? bitsf(fbits(x) & 0x80000000u) : x;
}

typedef struct {
    uint32_t bits; /* leaf mask (<= 0xf) or float bits of the size estimate */
    float tMinTemp;
    uint32_t node;
} cull_entry;

uint32_t or_block_traverse(const o_node* nodes, const o_frustum* f, o_stack* sorted) {
    if (f->degenerate) return 0;
    cull_entry cull[BLOCK_STACK + 4];
    uint32_t top = 0;
    float tMin = 0.0f;
    uint32_t node = ROOT;
    int sizeAxis = f->axis & ~1;
    for (;;) {
        const o_node* n = &nodes[node];
        uint32_t mask = or_test_children(f, n);
        mask = or_test_children_refine(f, n, mask);
        if (mask) {
            uint32_t leaf = mask & n->leafMask;
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
            if (mask) {
                float dist[4], size[4];
                for (int c = 0; c < 4; c++) {
                    dist[c] = -n->b[f->axis][c];
                    size[c] = ftz(n->b[sizeAxis][c] + n->b[sizeAxis + 1][c]);
                }
                for (int c = 0; c < 4; c++) {
                    if (!(mask & (1u << c))) continue;
                    float key = daz(fbits(size[c]));
                    uint32_t i = top;
                    while (i > 0 && daz(cull[i - 1].bits) > key) {
                        cull[i] = cull[i - 1];
                        i--;
                    }
                    cull[i].bits = fbits(size[c]);
                    cull[i].tMinTemp = dist[c];
                    cull[i].node = node + n->data[1 + c];
                    top++;
                }
            }
        }
        /* POP */
        if (top >= BLOCK_STACK - 4 || top == 0 || cull[top - 1].bits <= 0xfu) break;
        top--;
        node = cull[top].node;
        tMin = cull[top].tMinTemp;
    }
    /* re-sort by distance, closest on top */
    for (uint32_t t = 0; t < top; t++) {
        o_stack e;
        e.tMin = cull[t].tMinTemp;
        uint32_t m = cull[t].bits <= 0xfu ? cull[t].bits : 0u;
        e.nodeMask = (cull[t].node & NODE_BITS) | (m << 28);
        uint32_t slot = t;
        while (slot > 0 && sorted[slot - 1].tMin < e.tMin) {
            sorted[slot] = sorted[slot - 1];
            slot--;
        }
        sorted[slot] = e;
    }
    return top;
}

int64_t or_tile_traverse(const o_node* nodes, const o_stack* blockStack, uint32_t stackSize, const o_frustum* f,
                         uint32_t* out, uint64_t cap) {
    if (f->degenerate || stackSize == 0) return 0;
    static __thread o_stack stack[TILE_STACK];
    memcpy(stack, blockStack, sizeof(o_stack) * stackSize);
    uint64_t count = 0;
    uint32_t top = stackSize - 1;
    uint32_t node = stack[top].nodeMask & NODE_BITS;
    for (;;) {
        const o_node* n = &nodes[node];
        uint32_t mask = or_test_children(f, n);
        int pop = 1;
        if (mask) {
            uint32_t leaf = mask & n->leafMask;
            if (leaf) {
                mask = or_test_children_refine(f, n, mask);
                leaf = mask & n->leafMask;
                for (int k = 0; k < 4; k++) {
                    if (!(leaf & (1u << k))) continue;
                    for (uint32_t t = n->data[k]; t < n->data[k + 1]; t++) {
                        if (out) {
                            if (count >= cap) return -1;
                            out[count] = t;
                        }
                        count++;
                    }
                }
            }
            mask &= ~n->leafMask;
            if (mask && (stack[top].nodeMask >> 28) == 0) {
                float dist[4];
                for (int c = 0; c < 4; c++) dist[c] = -n->b[f->axis][c];
                if ((mask & (mask - 1)) == 0) {
                    /* single hit child: test it next without touching the stack */
                    int k = __builtin_ctz(mask);
                    node = node + n->data[1 + k];
                    continue;
                }
                int first = 1;
                for (int k = 0; k < 4; k++) {
                    if (!(mask & (1u << k))) continue;
                    if (!first) top++;
                    first = 0;
                    if (top >= TILE_STACK) return -1;
                    uint32_t i = top;
                    while (i > 0 && stack[i - 1].tMin < dist[k]) {
                        stack[i] = stack[i - 1];
                        i--;
                    }
                    stack[i].tMin = dist[k];
                    stack[i].nodeMask = (node + n->data[1 + k]) & NODE_BITS;
                }
                node = stack[top].nodeMask & NODE_BITS;
                pop = 0;
            }
        }
        if (pop) {
            if (top == 0) break;
            top--;
            node = stack[top].nodeMask & NODE_BITS;
        }
    }
    return (int64_t)count;
}

int64_t or_build_tile_lists(const o_node* nodes, const o_frustum* blockPre, const o_frustum* tilePre,
                            uint32_t blockCount, uint32_t* counts, uint32_t* out, uint64_t cap) {
    uint64_t total = 0;
    o_stack sorted[BLOCK_STACK];
    for (uint32_t b = 0; b < blockCount; b++) {
        uint32_t n = or_block_traverse(nodes, &blockPre[b], sorted);
        for (uint32_t t = 0; t < 64; t++) {
            uint32_t tile = b * 64 + t;
            if (!n) {
                counts[tile] = 0;
                continue;
            }
            int64_t c = or_tile_traverse(nodes, sorted, n, &tilePre[tile], out ? out + total : NULL,
                                         out ? cap - total : 0);
            if (c < 0) return -1;
            counts[tile] = (uint32_t)c;
            total += (uint64_t)c;
        }
    }
    return (int64_t)total;
}
