// The reference's binary scene format (include/hvvr/model_import.h documents the layout):
// loadBin / saveBin (samples_shared/model_import_bin.cpp:69-168), load by extension and
// createObjects (model_import.cpp:15-63).

#include "hvvr/model_import.h"

#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <memory>

namespace model_import {

using namespace hvvr;

namespace {

enum : uint32_t { kBinVersion = 1 };

struct BinHeader {
    uint32_t version, textureCount, lightCount, meshCount;
};
struct BinTexture {
    uint32_t format, width, height, stride;
};
struct BinMesh {
    transform xf;
    uint32_t vertexCount, triangleCount, nodeCount, materialCount;
};
static_assert(sizeof(BinHeader) == 16 && sizeof(BinTexture) == 16 && sizeof(BinMesh) == 48, "bin layout");
static_assert(sizeof(transform) == 32 && sizeof(LightUnion) == 56 && sizeof(ShadingVertex) == 32 &&
                  sizeof(PrecomputedTriangleShade) == 16 && sizeof(TopologyNode) == 32 && sizeof(SimpleMaterial) == 64,
              "bin element layout");

struct FileCloser {
    void operator()(FILE* f) const {
        if (f)
            fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

bool failWith(std::string* error, const std::string& msg) {
    if (error)
        *error = msg;
    return false;
}

// Reads exactly `count` elements or fails; counts are checked against the bytes left in
// the file before anything is allocated for them.
struct Reader {
    FILE* f;
    uint64_t left;
    bool read(void* dst, uint64_t bytes) {
        if (bytes > left)
            return false;
        if (bytes && fread(dst, 1, bytes, f) != bytes)
            return false;
        left -= bytes;
        return true;
    }
    template <class T>
    bool readArray(std::vector<T>& v, uint32_t count) {
        if (uint64_t(count) * sizeof(T) > left)
            return false;
        v.resize(count);
        return read(v.data(), uint64_t(count) * sizeof(T));
    }
};

// The references inside one mesh: triangle vertex and material indices, leaf triangle
// ranges and child offsets of the node list (graphics_types.h:126-151: an internal
// child's data is the offset of its node from the parent, a leaf's the triangle range
// boundary).
bool meshConsistent(const MeshData& m, std::string& why) {
    for (const PrecomputedTriangleShade& t : m.triShade) {
        for (uint32_t k = 0; k < 3; k++)
            if (t.indices[k] >= m.verts.size()) {
                why = "triangle vertex index out of range";
                return false;
            }
        if (t.material >= m.materials.size()) {
            why = "triangle material index out of range";
            return false;
        }
    }
    if (m.nodes.empty() && !m.triShade.empty()) {
        why = "mesh has triangles but no BVH nodes";
        return false;
    }
    for (size_t n = 0; n < m.nodes.size(); n++) {
        const TopologyNode& node = m.nodes[n];
        for (uint32_t c = 0; c < 4; c++) {
            if (node.isLeaf(int(c))) {
                const uint32_t a = node.data[c], b = node.data[c + 1];
                if (a > b || b > m.triShade.size()) {
                    why = "leaf triangle range out of order or out of range";
                    return false;
                }
            } else {
                const uint64_t child = uint64_t(n) + node.data[c + 1];
                if (node.data[c + 1] == 0 || child >= m.nodes.size()) {
                    why = "child node offset out of range";
                    return false;
                }
            }
        }
    }
    return true;
}

} // namespace

bool load(const char* path, Model& model, std::string* error) {
    const char* ext = path ? strrchr(path, '.') : nullptr;
    if (!ext)
        return failWith(error, "no file extension");
    if (strcasecmp(ext + 1, "bin") == 0)
        return loadBin(path, model, error);
    return failWith(error, std::string("unsupported scene format '") + ext + "' (FBX import is not built)");
}

bool loadBin(const char* path, Model& model, std::string* error) {
    model = Model();
    File file(fopen(path, "rb"));
    if (!file)
        return failWith(error, std::string("cannot open ") + path);
    if (fseek(file.get(), 0, SEEK_END) != 0)
        return failWith(error, "cannot size the file");
    const long size = ftell(file.get());
    if (size < 0 || fseek(file.get(), 0, SEEK_SET) != 0)
        return failWith(error, "cannot size the file");
    Reader r{file.get(), uint64_t(size)};
    BinHeader h;
    if (!r.read(&h, sizeof(h)))
        return failWith(error, "truncated header");
    if (h.version != kBinVersion)
        return failWith(error, "bin file version " + std::to_string(h.version) + ", expected 1");
    for (uint32_t n = 0; n < h.textureCount; n++) {
        BinTexture bt;
        if (!r.read(&bt, sizeof(bt)))
            return failWith(error, "truncated texture header");
        Texture tex;
        tex.tex.format = TextureFormat(bt.format);
        tex.tex.width = bt.width;
        tex.tex.height = bt.height;
        tex.tex.strideElements = bt.stride;
        const size_t bytes = getTextureSize(bt.stride, bt.height, tex.tex.format);
        if (!bytes || bt.stride < bt.width)
            return failWith(error, "texture " + std::to_string(n) + ": unknown format or stride below width");
        if (bytes > r.left || !r.readArray(tex.bytes, uint32_t(bytes)))
            return failWith(error, "truncated texture data");
        tex.tex.data = tex.bytes.data();
        model.textures.push_back(std::move(tex));
    }
    if (!r.readArray(model.lights, h.lightCount))
        return failWith(error, "truncated lights");
    for (uint32_t n = 0; n < h.meshCount; n++) {
        BinMesh bm;
        if (!r.read(&bm, sizeof(bm)))
            return failWith(error, "truncated mesh header");
        Mesh mesh;
        mesh.transform = bm.xf;
        if (!r.readArray(mesh.data.verts, bm.vertexCount) || !r.readArray(mesh.data.triShade, bm.triangleCount) ||
            !r.readArray(mesh.data.nodes, bm.nodeCount) || !r.readArray(mesh.data.materials, bm.materialCount))
            return failWith(error, "truncated mesh " + std::to_string(n));
        std::string why;
        if (!meshConsistent(mesh.data, why))
            return failWith(error, "mesh " + std::to_string(n) + ": " + why);
        model.meshes.push_back(std::move(mesh));
    }
    if (r.left != 0)
        return failWith(error, std::to_string(r.left) + " bytes after the last mesh");
    return true;
}

bool saveBin(const char* path, const Model& model, std::string* error) {
    File file(fopen(path, "wb"));
    if (!file)
        return failWith(error, std::string("cannot create ") + path);
    bool ok = true;
    auto write = [&](const void* p, size_t bytes) {
        if (ok && bytes)
            ok = fwrite(p, 1, bytes, file.get()) == bytes;
    };
    BinHeader h = {kBinVersion, uint32_t(model.textures.size()), uint32_t(model.lights.size()),
                   uint32_t(model.meshes.size())};
    write(&h, sizeof(h));
    for (const Texture& t : model.textures) {
        BinTexture bt = {uint32_t(t.tex.format), t.tex.width, t.tex.height, t.tex.strideElements};
        const size_t bytes = getTextureSize(bt.stride, bt.height, t.tex.format);
        if (!bytes || !t.tex.data)
            return failWith(error, "texture with an unknown format or no texels");
        write(&bt, sizeof(bt));
        write(t.tex.data, bytes);
    }
    write(model.lights.data(), model.lights.size() * sizeof(LightUnion));
    for (const Mesh& m : model.meshes) {
        BinMesh bm = {m.transform, uint32_t(m.data.verts.size()), uint32_t(m.data.triShade.size()),
                      uint32_t(m.data.nodes.size()), uint32_t(m.data.materials.size())};
        write(&bm, sizeof(bm));
        write(m.data.verts.data(), m.data.verts.size() * sizeof(ShadingVertex));
        write(m.data.triShade.data(), m.data.triShade.size() * sizeof(PrecomputedTriangleShade));
        write(m.data.nodes.data(), m.data.nodes.size() * sizeof(TopologyNode));
        write(m.data.materials.data(), m.data.materials.size() * sizeof(SimpleMaterial));
    }
    if (!ok)
        return failWith(error, std::string("write failed: ") + path);
    return true;
}

bool createObjects(Raycaster& raycaster, Model& model) {
    for (const Texture& t : model.textures)
        if (!raycaster.createTexture(t.tex))
            return false;
    for (const LightUnion& l : model.lights)
        if (!raycaster.createLight(l))
            return false;
    for (const Mesh& m : model.meshes) {
        hvvr::Model* created = raycaster.createModel(m.data);
        if (!created)
            return false;
        created->setTransform(m.transform);
    }
    return true;
}

} // namespace model_import
