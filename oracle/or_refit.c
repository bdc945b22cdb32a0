/* Oracle: scene refit from scratch (test infrastructure only).
 *
 *   SceneTransformVerticesKernel     raycaster/gpu_scene_state.cu:88-132
 *   ScenePrecomputeTrianglesKernel   raycaster/gpu_scene_state.cu:167-244
 *   SceneRefitPre/IntFixup kernels   raycaster/gpu_scene_state.cu:134-165, 246-270
 *   internal refit                   raycaster/gpu_scene_state.cu:273-375
 *
 * The reference accumulates leaf bounds with atomicMax on order-preserving integer
 * images of the floats (FloatFlip, sort.h:14-32) and internal bounds with max() over a
 * child's four slots; both are exact, so a from-scratch recomputation in any order gives
 * the same bits. max is taken in the FloatFlip order (total, -0 < +0) throughout.
 * Empty leaves keep the reset value -inf in all six bound arrays.
 */
#include "or_math.h"
#include <stdlib.h>

static void node_bounds_max(const o_node* n, float out[6]) {
    for (int k = 0; k < 6; k++)
        out[k] = maxTotal(n->b[k][0], maxTotal(n->b[k][1], maxTotal(n->b[k][2], n->b[k][3])));
}

int or_refit(const o_vertex* verts, uint32_t vertexCount, const uint32_t* vertexToTransform, const float* transforms,
             const o_trishade* shade, uint32_t triCount, const o_node* nodes, uint32_t nodeCount,
             o_vertex* worldVerts, o_tri* trisOut, o_node* nodesOut) {
    for (uint32_t v = 0; v < vertexCount; v++) {
        const float* M = transforms + 16 * (uint64_t)vertexToTransform[v];
        o_vertex o = verts[v];
        float fp[4];
        m4mul(M, o.pos.x, o.pos.y, o.pos.z, 1.0f, fp);
        float invW = 1.0f / fp[3];
        o.pos = v3(fp[0] * invW, fp[1] * invW, fp[2] * invW);
        ov3 n = v3(h2f(verts[v].normal[0]), h2f(verts[v].normal[1]), h2f(verts[v].normal[2]));
        n = norm3(m3mul(M, n));
        o.normal[0] = f2h(n.x);
        o.normal[1] = f2h(n.y);
        o.normal[2] = f2h(n.z);
        o.normal[3] = f2h(0.0f);
        worldVerts[v] = o;
    }
    for (uint32_t t = 0; t < triCount; t++) {
        ov3 a = worldVerts[shade[t].idx[0]].pos, b = worldVerts[shade[t].idx[1]].pos, c = worldVerts[shade[t].idx[2]].pos;
        trisOut[t].v0 = a;
        trisOut[t].e0 = sub3(b, a);
        trisOut[t].e1 = sub3(c, a);
    }
    for (uint32_t i = 0; i < nodeCount; i++) {
        nodesOut[i] = nodes[i];
        for (int k = 0; k < 6; k++)
            for (int c = 0; c < 4; c++) nodesOut[i].b[k][c] = -INFINITY;
    }
    /* leaves: bounds of their triangles */
    for (uint32_t i = 0; i < nodeCount; i++) {
        o_node* n = &nodesOut[i];
        for (int c = 0; c < 4; c++) {
            if (!((n->leafMask >> c) & 1u)) continue;
            for (uint32_t t = n->data[c]; t < n->data[c + 1] && t < triCount; t++) {
                ov3 p[3] = {worldVerts[shade[t].idx[0]].pos, worldVerts[shade[t].idx[1]].pos,
                            worldVerts[shade[t].idx[2]].pos};
                float mx[3], nm[3];
                for (int a = 0; a < 3; a++) {
                    float q0 = comp3(p[0], a), q1 = comp3(p[1], a), q2 = comp3(p[2], a);
                    mx[a] = maxTotal(q0, maxTotal(q1, q2));
                    nm[a] = maxTotal(-q0, maxTotal(-q1, -q2));
                }
                for (int a = 0; a < 3; a++) {
                    n->b[2 * a][c] = maxTotal(n->b[2 * a][c], mx[a]);
                    n->b[2 * a + 1][c] = maxTotal(n->b[2 * a + 1][c], nm[a]);
                }
            }
        }
    }
    /* internal: children complete before parents (children have larger indices) */
    for (uint32_t i = nodeCount; i-- > 0;) {
        o_node* n = &nodesOut[i];
        for (int c = 0; c < 4; c++) {
            if ((n->leafMask >> c) & 1u) continue;
            uint32_t child = i + n->data[1 + c];
            if (child <= i || child >= nodeCount) return -1;
            float b[6];
            node_bounds_max(&nodesOut[child], b);
            for (int k = 0; k < 6; k++) n->b[k][c] = b[k];
        }
    }
    return 0;
}
