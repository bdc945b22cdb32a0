#pragma once
// Scene files for the hvvr::Raycaster C++ API: the reference's "model_import" layer
// (samples_shared/model_import.h:19-69, model_import.cpp:15-63, model_import_bin.cpp:14-170).
//
// The binary ".bin" format is the reference's, byte for byte (little endian, packed):
//   header   u32 version = 1, u32 textureCount, u32 lightCount, u32 meshCount      16 B
//   textures per texture: u32 format (TextureFormat), u32 width, u32 height,
//            u32 strideElements (16 B), then getTextureSize(stride, height, format)
//            bytes of texels
//   lights   LightUnion (56 B) each
//   meshes   per mesh: transform (quaternion x y z w, translation x y z, scale; 32 B),
//            u32 vertexCount, triangleCount, nodeCount, materialCount (48 B in all),
//            then ShadingVertex[vertexCount] (32 B), PrecomputedTriangleShade
//            [triangleCount] (16 B), TopologyNode[nodeCount] (32 B) and
//            SimpleMaterial[materialCount] (64 B)
// A mesh's vertices and triangles are stored after GenerateTopology (its node list is the
// BVH topology over them), so a loaded mesh goes to Raycaster::createModel as it is.
//
// Differences from the reference: loadBin checks every read and the mesh data's
// internal references (triangle vertex and material indices, leaf triangle ranges, child
// offsets) and fails on a truncated or inconsistent file instead of reading past it;
// failures describe themselves through `error`. FBX import is not part of this build
// (the reference compiles it out by default, MODEL_IMPORT_ENABLE_FBX 0).

#include "raycaster.h"

#include <string>
#include <vector>

namespace model_import {

struct Texture {
    hvvr::TextureData tex = {};   // tex.data points into `bytes`
    std::vector<uint8_t> bytes;
    Texture() = default;
    Texture(Texture&&) = default; // a moved vector keeps its buffer, so tex.data stays valid
    Texture& operator=(Texture&&) = default;
    Texture(const Texture&) = delete;
    Texture& operator=(const Texture&) = delete;
};

struct Mesh {
    hvvr::transform transform;
    hvvr::MeshData data;
};

struct Model {
    std::vector<Texture> textures;
    std::vector<hvvr::LightUnion> lights;
    std::vector<Mesh> meshes;
};

// By file extension (case-insensitive): ".bin" -> loadBin; anything else fails.
bool load(const char* path, Model& model, std::string* error = nullptr);
bool loadBin(const char* path, Model& model, std::string* error = nullptr);
bool saveBin(const char* path, const Model& model, std::string* error = nullptr);
// createTexture / createLight / createModel (+ setTransform) for everything in the model.
bool createObjects(hvvr::Raycaster& raycaster, Model& model);

} // namespace model_import
