// Scene update on the GPU: world-space vertices, precomputed triangles and the
// bottom-up BVH bound refit (raycaster/gpu_scene_state.cu:88-451).
//
// MI355X layout choices:
//  * leaf bounds are accumulated with 32-bit atomicMax on the order-preserving integer
//    image of each float (FloatFlip), exactly like the reference; the reset and the
//    un-flip each touch only the 96 bound bytes of a node with one thread per float;
//  * internal bounds are refit level by level (deepest parents first), one thread per
//    child->parent task, instead of the reference's persistent spin-wait kernel
//    (gpu_scene_state.cu:273-330), which relies on independent forward progress
//    between warps that AMD hardware does not guarantee.
// max() everywhere is the FloatFlip total order, so results are independent of the
// order in which atomics land and bit-identical to a from-scratch recomputation.

#include "device_common.h"
#include "gpu_internal.h"

namespace hvvr {

__global__ void transformVerticesKernel(const ShadingVertex* __restrict__ in, ShadingVertex* __restrict__ out,
                                        const uint32_t* __restrict__ vertexToTransform,
                                        const matrix4x4* __restrict__ transforms, uint32_t vertexCount) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= vertexCount)
        return;
    ShadingVertex v = in[i];
    matrix4x4 M = transforms[vertexToTransform[i]];
    vector4 p = M * vector4(v.pos, 1.0f);
    float invW = 1.0f / p.w;
    v.pos = vector3(p.x * invW, p.y * invW, p.z * invW);
    vector3 n(halfToFloat(v.normal[0]), halfToFloat(v.normal[1]), halfToFloat(v.normal[2]));
    n = normalize(M.upper3x3() * n);
    v.normal[0] = floatToHalfRNE(n.x);
    v.normal[1] = floatToHalfRNE(n.y);
    v.normal[2] = floatToHalfRNE(n.z);
    v.normal[3] = floatToHalfRNE(0.0f);
    out[i] = v;
}

// every bound float := FloatFlip(-inf); one thread per float, 24 per node
__global__ void resetBoundsKernel(BVHNode* __restrict__ nodes, uint32_t nodeCount) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodeCount * 24u)
        return;
    uint32_t node = i / 24u, f = i % 24u;
    uint32_t* b = reinterpret_cast<uint32_t*>(&nodes[node].bounds[0][0]);
    b[f] = floatFlip(0xff800000u);
}

__global__ void precomputeTrianglesKernel(const ShadingVertex* __restrict__ verts,
                                          const PrecomputedTriangleShade* __restrict__ shade,
                                          const uint32_t* __restrict__ triToNodeChild,
                                          PrecomputedTriangleIntersect* __restrict__ out, BVHNode* nodes,
                                          uint32_t triCount) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= triCount)
        return;
    PrecomputedTriangleShade s = shade[t];
    vector3 a = verts[s.indices[0]].pos, b = verts[s.indices[1]].pos, c = verts[s.indices[2]].pos;
    PrecomputedTriangleIntersect o;
    o.v0 = a;
    o.edge0 = b - a;
    o.edge1 = c - a;
    out[t] = o;
    uint32_t nc = triToNodeChild[t];
    uint32_t parent = nc >> 2, slot = nc & 3u;
    uint32_t* base = reinterpret_cast<uint32_t*>(&nodes[parent].bounds[0][0]);
    for (int axis = 0; axis < 3; axis++) {
        float q0 = a[axis], q1 = b[axis], q2 = c[axis];
        float mx = maxTotal(q0, maxTotal(q1, q2));
        float nm = maxTotal(-q0, maxTotal(-q1, -q2));
        atomicMax(base + (2 * axis) * 4 + slot, floatFlip(fbits(mx)));
        atomicMax(base + (2 * axis + 1) * 4 + slot, floatFlip(fbits(nm)));
    }
}

__global__ void unflipBoundsKernel(BVHNode* __restrict__ nodes, uint32_t nodeCount) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nodeCount * 24u)
        return;
    uint32_t node = i / 24u, f = i % 24u;
    uint32_t* b = reinterpret_cast<uint32_t*>(&nodes[node].bounds[0][0]);
    b[f] = floatUnflip(b[f]);
}

// one child->parent task per thread: parent slot := max over the child's 4 slots
__global__ void refitLevelKernel(const uint32_t* __restrict__ taskChild, const uint32_t* __restrict__ taskNodeChild,
                                 uint32_t start, uint32_t count, BVHNode* nodes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    uint32_t child = taskChild[start + i];
    uint32_t nc = taskNodeChild[start + i];
    uint32_t parent = nc >> 2, slot = nc & 3u;
    const float4* cb = reinterpret_cast<const float4*>(&nodes[child].bounds[0][0]);
    for (int k = 0; k < 6; k++) {
        float4 v = cb[k];
        nodes[parent].bounds[k][slot] = maxTotal(v.x, maxTotal(v.y, maxTotal(v.z, v.w)));
    }
}

void launchSceneRefit(hvvr_gpu_scene* s, hipStream_t stream) {
    const uint32_t bs = 256;
    if (s->vertexCount)
        transformVerticesKernel<<<(s->vertexCount + bs - 1) / bs, bs, 0, stream>>>(
            s->vertsModel.ptr, s->vertsWorld.ptr, s->vertexToTransform.ptr, s->transforms.ptr, s->vertexCount);
    uint32_t boundFloats = s->nodeCount * 24u;
    resetBoundsKernel<<<(boundFloats + bs - 1) / bs, bs, 0, stream>>>(s->nodes.ptr, s->nodeCount);
    if (s->triangleCount)
        precomputeTrianglesKernel<<<(s->triangleCount + bs - 1) / bs, bs, 0, stream>>>(
            s->vertsWorld.ptr, s->triShade.ptr, s->triToNodeChild.ptr, s->triIntersect.ptr, s->nodes.ptr,
            s->triangleCount);
    unflipBoundsKernel<<<(boundFloats + bs - 1) / bs, bs, 0, stream>>>(s->nodes.ptr, s->nodeCount);
    for (const RefitLevel& lv : s->levels)
        refitLevelKernel<<<(lv.count + bs - 1) / bs, bs, 0, stream>>>(s->taskChild.ptr, s->taskNodeChild.ptr, lv.start,
                                                                      lv.count, s->nodes.ptr);
}

} // namespace hvvr
