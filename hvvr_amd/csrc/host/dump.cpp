// Scene and ray dumps for cross-checks with external raytracers (raycaster/render.cpp:28-201):
// the scene's world-space triangles as OFF, the meshes' BVH topology as BFF and a camera's
// subsample rays as RFF, each in the reference's text form or its binary form (the four
// letters of the format with a 'b', then the counts as int32 and the packed arrays; the
// file name gets a trailing 'b').
//
// One difference: the reference's binary RFF writer writes the first N rays of the unculled
// array after counting N rays without NaN directions; here the rays without NaN are
// written, as its text writer does.

#include "host_internal.h"

#include "../../../include/hvvr_gpu.h"

#include <stdio.h>
#include <string.h>

#include <memory>

namespace hvvr {

namespace {

struct FileCloser {
    void operator()(FILE* f) const {
        if (f)
            fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

File create(const std::string& name) {
    File f(fopen(name.c_str(), "wb"));
    if (!f)
        throw RaycasterError("cannot create " + name);
    return f;
}

void put(FILE* f, const void* p, size_t bytes, const std::string& name) {
    if (bytes && fwrite(p, 1, bytes, f) != bytes)
        throw RaycasterError("write failed: " + name);
}

} // namespace

void dumpRaysToRFF(const std::string& filename, const std::vector<SimpleRay>& rays, bool binary) {
    std::vector<SimpleRay> kept;
    for (const SimpleRay& r : rays)
        if (r.direction.x == r.direction.x)
            kept.push_back(r);
    if (kept.size() > 0x7fffffffu)
        throw RaycasterError("dumpRaysToRFF: too many rays");
    if (binary) {
        const std::string name = filename + "b";
        File f = create(name);
        const int32_t count = int32_t(kept.size());
        put(f.get(), "RFFb", 4, name);
        put(f.get(), &count, 4, name);
        put(f.get(), kept.data(), kept.size() * sizeof(SimpleRay), name);
        return;
    }
    File f = create(filename);
    fprintf(f.get(), "RFF\n%d\n", int(kept.size()));
    for (const SimpleRay& r : kept)
        fprintf(f.get(), "%.9g %.9g %.9g %.9g %.9g %.9g\n", r.origin.x, r.origin.y, r.origin.z, r.direction.x,
                r.direction.y, r.direction.z);
}

void dumpSceneToOFF(const std::string& filename, const std::vector<const Model*>& models, bool binary) {
    std::vector<vector3> vertices;
    std::vector<uint32_t> tris;
    for (const Model* model : models) {
        const MeshData& mesh = model->getMesh();
        const uint32_t base = uint32_t(vertices.size());
        for (const PrecomputedTriangleShade& t : mesh.triShade)
            for (int k = 0; k < 3; k++)
                tris.push_back(t.indices[k] + base);
        const matrix4x4 m(model->getTransform());
        for (const ShadingVertex& v : mesh.verts) {
            const vector4 p = m * vector4(v.pos, 1.0f);
            vertices.push_back(vector3(p.x, p.y, p.z));
        }
    }
    if (vertices.size() > 0x7fffffffu || tris.size() / 3 > 0x7fffffffu)
        throw RaycasterError("dumpSceneToOFF: scene too large");
    const int32_t counts[3] = {int32_t(vertices.size()), int32_t(tris.size() / 3), 0};
    if (binary) {
        const std::string name = filename + "b";
        File f = create(name);
        put(f.get(), "OFFb", 4, name);
        put(f.get(), counts, sizeof(counts), name);
        put(f.get(), vertices.data(), vertices.size() * sizeof(vector3), name);
        put(f.get(), tris.data(), tris.size() * 4, name);
        return;
    }
    File f = create(filename);
    fprintf(f.get(), "OFF\n%d %d %d\n", counts[0], counts[1], counts[2]);
    for (const vector3& v : vertices)
        fprintf(f.get(), "%f %f %f\n", v.x, v.y, v.z);
    for (size_t t = 0; t < tris.size(); t += 3)
        fprintf(f.get(), "3 %d %d %d\n", int(tris[t]), int(tris[t + 1]), int(tris[t + 2]));
}

void dumpBVH(const std::string& filename, const std::vector<const Model*>& models, bool binary) {
    // per node: the leaf mask and data[0..4] (first triangle / child offsets), 24 bytes
    std::vector<uint32_t> nodes;
    for (const Model* model : models)
        for (const TopologyNode& n : model->getMesh().nodes) {
            nodes.push_back(n.leafMask);
            for (int k = 0; k < 5; k++)
                nodes.push_back(n.data[k]);
        }
    const int32_t count = int32_t(nodes.size() / 6);
    if (binary) {
        const std::string name = filename + "b";
        File f = create(name);
        put(f.get(), "BFFb", 4, name);
        put(f.get(), &count, 4, name);
        put(f.get(), nodes.data(), nodes.size() * 4, name);
        return;
    }
    File f = create(filename);
    fprintf(f.get(), "BFF\n%d\n", count);
    for (size_t i = 0; i < nodes.size(); i += 6)
        fprintf(f.get(), "%u %u %u %u %u %u\n", nodes[i], nodes[i + 1], nodes[i + 2], nodes[i + 3], nodes[i + 4],
                nodes[i + 5]);
}

} // namespace hvvr
