// Scene flattening: per-model meshes -> global vertex/triangle/node/material arrays,
// with "root fixup" nodes linking the per-model BVHs under one root
// (raycaster/scene_update.cpp:22-143).

#include "host_internal.h"

#include <algorithm>

namespace hvvr {

void buildScene(const std::vector<SceneModel>& models, FlatScene& out) {
    out.verts.clear();
    out.vertexToTransform.clear();
    out.triShade.clear();
    out.nodes.clear();
    out.materials.clear();
    out.transforms.clear();

    uint32_t modelCount = uint32_t(models.size());
    if (modelCount == 0) {
        // a root whose four children are empty leaves (scene_update.cpp:25-36)
        BVHNode root = {};
        root.leafMask = 0xf;
        out.nodes.push_back(root);
        return;
    }

    uint32_t rootNodeCount = (modelCount + 1) / 3;
    size_t vertexCount = 0, triCount = 0, nodeCount = rootNodeCount, materialCount = 0;
    for (const SceneModel& m : models) {
        vertexCount += m.mesh->verts.size();
        triCount += m.mesh->triShade.size();
        nodeCount += m.mesh->nodes.size();
        materialCount += m.mesh->materials.size();
    }
    out.verts.reserve(vertexCount);
    out.vertexToTransform.reserve(vertexCount);
    out.triShade.reserve(triCount);
    out.nodes.assign(nodeCount, BVHNode{});
    out.materials.reserve(materialCount);

    std::vector<uint32_t> modelRoot(modelCount);
    uint32_t vertexOffset = 0, triOffset = 0, materialOffset = 0, nodeOffset = rootNodeCount;
    for (uint32_t mi = 0; mi < modelCount; mi++) {
        const MeshData& mesh = *models[mi].mesh;
        out.transforms.push_back(models[mi].modelToWorld);
        for (const SimpleMaterial& mat : mesh.materials)
            out.materials.push_back(mat);
        for (const ShadingVertex& v : mesh.verts) {
            out.verts.push_back(v);
            out.vertexToTransform.push_back(mi);
        }
        for (PrecomputedTriangleShade t : mesh.triShade) {
            t.indices[0] += vertexOffset;
            t.indices[1] += vertexOffset;
            t.indices[2] += vertexOffset;
            t.material += materialOffset;
            out.triShade.push_back(t);
        }
        modelRoot[mi] = nodeOffset;
        for (size_t n = 0; n < mesh.nodes.size(); n++) {
            const TopologyNode& topo = mesh.nodes[n];
            BVHNode& node = out.nodes[nodeOffset + n];
            node.leafMask = topo.leafMask;
            node.data[0] = triOffset + topo.data[0];
            for (int k = 0; k < 4; k++)
                node.data[1 + k] = topo.isLeaf(k) ? triOffset + topo.data[1 + k] : topo.data[1 + k];
        }
        vertexOffset += uint32_t(mesh.verts.size());
        triOffset += uint32_t(mesh.triShade.size());
        materialOffset += uint32_t(mesh.materials.size());
        nodeOffset += uint32_t(mesh.nodes.size());
    }

    if (rootNodeCount > 0) {
        // child slot sequence: empty leaves, then fixup nodes 1..R-1, then model roots;
        // fixup node i owns slots [4i, 4i+4) (scene_update.cpp:109-142)
        std::vector<int64_t> slots;
        slots.reserve(rootNodeCount * 4);
        size_t emptySlots = size_t(rootNodeCount) * 4 - modelCount - (rootNodeCount - 1);
        for (size_t i = 0; i < emptySlots; i++)
            slots.push_back(-1);
        for (uint32_t i = 1; i < rootNodeCount; i++)
            slots.push_back(i);
        for (uint32_t r : modelRoot)
            slots.push_back(r);
        for (uint32_t i = 0; i < rootNodeCount; i++) {
            BVHNode& node = out.nodes[i];
            node.leafMask = 0;
            node.data[0] = 0;
            for (int k = 0; k < 4; k++) {
                int64_t target = slots[size_t(i) * 4 + k];
                if (target < 0) {
                    node.data[1 + k] = 0;
                    node.leafMask |= 1u << k;
                } else {
                    node.data[1 + k] = uint32_t(target - int64_t(i));
                }
            }
        }
    }
}

} // namespace hvvr
