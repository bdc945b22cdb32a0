// C interface over the host-side building blocks (include/hvvr_host.h).

#include "../../../include/hvvr_host.h"
#include "hvvr/model_import.h"
#include "host_internal.h"

#include <string.h>
#include <string>

using namespace hvvr;

namespace {
thread_local std::string gLastError;

int fail(const std::string& msg, int code = -1) {
    gLastError = msg;
    return code;
}

template <class T>
int copyOut(const std::vector<T>& v, void* dst, uint64_t bytes) {
    uint64_t need = uint64_t(v.size()) * sizeof(T);
    if (bytes < need)
        return fail("destination buffer too small", -2);
    if (need)
        memcpy(dst, v.data(), need);
    return 0;
}
} // namespace

struct hvvr_host_scene {
    FlatScene flat;
    uint32_t modelCount = 0;
};



extern "C" {

const char* hvvr_host_last_error(void) {
    return gLastError.c_str();
}

} // extern "C"

void hvvrSetHostError(const std::string& msg) {
    gLastError = msg;
}

extern "C" {

static int finishScene(std::vector<MeshData>& meshes, const std::vector<matrix4x4>& xforms,
                       const std::vector<LightUnion>& lights, hvvr_host_scene** out) {
    std::string err;
    for (MeshData& m : meshes)
        if (!GenerateTopology(m, &err))
            return fail(err);
    std::vector<SceneModel> models;
    for (size_t i = 0; i < meshes.size(); i++)
        models.push_back({&meshes[i], xforms[i]});
    hvvr_host_scene* s = new hvvr_host_scene();
    buildScene(models, s->flat);
    s->flat.lights = lights;
    s->modelCount = uint32_t(meshes.size());
    *out = s;
    return 0;
}

int hvvr_host_scene_create(int kind, uint64_t seed, uint32_t meshCount, hvvr_host_scene** out) {
    if (!out)
        return fail("null output");
    SceneGenResult gen;
    if (kind == HVVR_SCENE_HEIGHTFIELD_BOXES) {
        generateHeightfieldBoxesScene(seed, gen);
    } else if (kind == HVVR_SCENE_ROOM) {
        if (meshCount == 0)
            return fail("room scene needs meshCount > 0");
        generateRoomScene(seed, meshCount, gen);
    } else {
        return fail("unknown scene kind");
    }
    return finishScene(gen.meshes, gen.modelToWorld, gen.lights, out);
}

int hvvr_host_scene_save_bin(int kind, uint64_t seed, uint32_t meshCount, const char* path) {
    if (!path)
        return fail("null path");
    SceneGenResult gen;
    if (kind == HVVR_SCENE_HEIGHTFIELD_BOXES)
        generateHeightfieldBoxesScene(seed, gen);
    else if (kind == HVVR_SCENE_ROOM && meshCount > 0)
        generateRoomScene(seed, meshCount, gen);
    else
        return fail("unknown scene kind or no meshes");
    model_import::Model model;
    model.lights = gen.lights;
    std::string err;
    for (size_t i = 0; i < gen.meshes.size(); i++) {
        if (!GenerateTopology(gen.meshes[i], &err))
            return fail(err);
        model.meshes.push_back({gen.transforms[i], std::move(gen.meshes[i])});
    }
    if (!model_import::saveBin(path, model, &err))
        return fail(err);
    return 0;
}

int hvvr_host_scene_load_bin(const char* path, hvvr_host_scene** out) {
    if (!path || !out)
        return fail("null argument");
    model_import::Model model;
    std::string err;
    if (!model_import::loadBin(path, model, &err))
        return fail(std::string(path) + ": " + err);
    // the stored meshes carry their topology: flattened as they are (createModel +
    // setTransform of model_import.cpp:47-59)
    std::vector<SceneModel> models;
    for (const model_import::Mesh& m : model.meshes)
        models.push_back({&m.data, matrix4x4(m.transform)});
    hvvr_host_scene* s = new hvvr_host_scene();
    buildScene(models, s->flat);
    s->flat.lights = model.lights;
    s->modelCount = uint32_t(models.size());
    *out = s;
    return 0;
}

int hvvr_host_scene_create_from_meshes(uint32_t meshCount, const uint32_t* vertexStart, const void* verts,
                                       const uint32_t* triStart, const uint32_t* triIndices,
                                       const uint32_t* triMaterial, const float* transforms16,
                                       uint32_t materialCount, const void* materials, uint32_t lightCount,
                                       const void* lights, hvvr_host_scene** out) {
    if (!out || (meshCount && (!vertexStart || !verts || !triStart || !triIndices || !triMaterial)))
        return fail("null argument");
    std::vector<MeshData> meshes(meshCount);
    std::vector<matrix4x4> xforms(meshCount, matrix4x4::identity());
    const ShadingVertex* v = (const ShadingVertex*)verts;
    for (uint32_t m = 0; m < meshCount; m++) {
        MeshData& md = meshes[m];
        if (vertexStart[m + 1] < vertexStart[m] || triStart[m + 1] < triStart[m])
            return fail("mesh ranges must be non-decreasing");
        md.verts.assign(v + vertexStart[m], v + vertexStart[m + 1]);
        for (uint32_t t = triStart[m]; t < triStart[m + 1]; t++) {
            PrecomputedTriangleShade ts;
            for (int k = 0; k < 3; k++)
                ts.indices[k] = triIndices[3 * size_t(t) + k];
            ts.material = triMaterial[t];
            if (ts.material >= materialCount)
                return fail("triangle material index out of range");
            md.triShade.push_back(ts);
        }
        if (transforms16)
            memcpy(&xforms[m], transforms16 + 16 * size_t(m), 64);
    }
    int rc = finishScene(meshes, xforms,
                         std::vector<LightUnion>((const LightUnion*)lights, (const LightUnion*)lights + lightCount),
                         out);
    if (rc == 0) {
        // materials are shared scene-wide: indices were not offset per mesh
        (*out)->flat.materials.assign((const SimpleMaterial*)materials,
                                      (const SimpleMaterial*)materials + materialCount);
    }
    return rc;
}

void hvvr_host_scene_destroy(hvvr_host_scene* scene) {
    delete scene;
}

int hvvr_host_scene_counts(const hvvr_host_scene* s, uint64_t counts[7]) {
    if (!s || !counts)
        return fail("null argument");
    counts[0] = s->flat.verts.size();
    counts[1] = s->flat.triShade.size();
    counts[2] = s->flat.nodes.size();
    counts[3] = s->flat.materials.size();
    counts[4] = s->flat.transforms.size();
    counts[5] = s->flat.lights.size();
    counts[6] = s->modelCount;
    return 0;
}

int hvvr_host_scene_copy(const hvvr_host_scene* s, int array, void* dst, uint64_t bytes) {
    if (!s || !dst)
        return fail("null argument");
    switch (array) {
        case HVVR_SCENE_VERTS: return copyOut(s->flat.verts, dst, bytes);
        case HVVR_SCENE_VERTEX_TRANSFORM: return copyOut(s->flat.vertexToTransform, dst, bytes);
        case HVVR_SCENE_TRI_SHADE: return copyOut(s->flat.triShade, dst, bytes);
        case HVVR_SCENE_NODES: return copyOut(s->flat.nodes, dst, bytes);
        case HVVR_SCENE_MATERIALS: return copyOut(s->flat.materials, dst, bytes);
        case HVVR_SCENE_TRANSFORMS: return copyOut(s->flat.transforms, dst, bytes);
        case HVVR_SCENE_LIGHTS: return copyOut(s->flat.lights, dst, bytes);
        default: return fail("unknown scene array");
    }
}

int hvvr_host_view_preset_params(int preset, uint32_t eye, uint32_t width, uint32_t height, float params[8]) {
    if (!params)
        return fail("null output");
    CameraPreset cam;
    if (preset == HVVR_VIEW_PRESET_HEIGHTFIELD)
        cam = heightfieldCamera();
    else if (preset == HVVR_VIEW_PRESET_ROOM_EYE)
        cam = roomEyeCamera(eye, width ? width : 1080, height ? height : 1200);
    else
        return fail("unknown view preset");
    const float p[8] = {cam.viewportLower.x, cam.viewportLower.y, cam.viewportUpper.x, cam.viewportUpper.y,
                        cam.barrelK1,        cam.barrelK2,        cam.lens.radius,     cam.lens.focalDistance};
    memcpy(params, p, sizeof(p));
    return 0;
}

int hvvr_host_sample_to_camera(const float vp[4], float out[9]) {
    if (!vp || !out)
        return fail("null argument");
    const matrix3x3 m = sampleToCamera(vector2(vp[0], vp[1]), vector2(vp[2], vp[3]));
    memcpy(out, &m, sizeof(m));
    return 0;
}

int hvvr_host_view_create_preset(int preset, uint32_t eye, uint32_t width, uint32_t height, hvvr_host_view** out) {
    if (!out)
        return fail("null output");
    CameraPreset cam;
    if (preset == HVVR_VIEW_PRESET_HEIGHTFIELD) {
        cam = heightfieldCamera();
        if (width && height) {
            cam.width = width;
            cam.height = height;
        }
    } else if (preset == HVVR_VIEW_PRESET_ROOM_EYE) {
        cam = roomEyeCamera(eye, width ? width : 1080, height ? height : 1200);
    } else {
        return fail("unknown view preset");
    }
    std::vector<Sample> samples = getGridSamples(cam.width, cam.height);
    if (cam.barrelK1 != 0.0f || cam.barrelK2 != 0.0f)
        applyBarrelDistortion(samples, cam.barrelK1, cam.barrelK2);
    hvvr_host_view* v = new hvvr_host_view();
    v->sampleToCam = sampleToCamera(cam.viewportLower, cam.viewportUpper);
    buildSampleSetup(samples.data(), uint32_t(samples.size()), 1, cam.width, cam.height,
                     Sample2Dto3DMappingSettings(v->sampleToCam, cam.lens), v->setup);
    v->cameraToWorld = matrix4x4(cam.cameraToWorld);
    v->lens = cam.lens;
    *out = v;
    return 0;
}

int hvvr_host_view_create_from_samples(const void* samples, uint32_t count, uint32_t split, uint32_t width,
                                       uint32_t height, const float vp[4], float lensRadius, float focalDistance,
                                       hvvr_host_view** out) {
    return hvvr_host_view_create_spherical(samples, count, split, width, height, vp, 0.0f, 0.0f, lensRadius,
                                           focalDistance, out);
}

int hvvr_host_view_create_spherical(const void* samples, uint32_t count, uint32_t split, uint32_t width,
                                    uint32_t height, const float vp[4], float fovXDegrees, float fovYDegrees,
                                    float lensRadius, float focalDistance, hvvr_host_view** out) {
    if (!out || !samples || !vp || split < 1 || split > 3)
        return fail("invalid argument");
    if (!(fovXDegrees >= 0.0f && fovYDegrees >= 0.0f && fovXDegrees <= 360.0f && fovYDegrees <= 180.0f))
        return fail("spherical section: angles in [0, 360] x [0, 180] degrees");
    hvvr_host_view* v = new hvvr_host_view();
    ThinLens lens = {lensRadius, focalDistance};
    v->sampleToCam = sampleToCamera(vector2(vp[0], vp[1]), vector2(vp[2], vp[3]));
    // Camera::get2DSampleMappingSettings (camera.cpp:206-212): a nonzero horizontal angle
    // selects the spherical section
    Sample2Dto3DMappingSettings mapping =
        fovXDegrees > 0.0f ? Sample2Dto3DMappingSettings::sphericalSection(v->sampleToCam, lens, fovXDegrees, fovYDegrees)
                           : Sample2Dto3DMappingSettings(v->sampleToCam, lens);
    buildSampleSetup((const Sample*)samples, count, split, width, height, mapping, v->setup);
    v->lens = lens;
    *out = v;
    return 0;
}

void hvvr_host_view_destroy(hvvr_host_view* view) {
    delete view;
}

int hvvr_host_view_counts(const hvvr_host_view* v, uint64_t counts[7]) {
    if (!v || !counts)
        return fail("null argument");
    counts[0] = v->setup.validSampleCount;
    counts[1] = v->setup.sampleCount;
    counts[2] = v->setup.blockCount;
    counts[3] = v->setup.width;
    counts[4] = v->setup.height;
    counts[5] = v->setup.splitColorSamples;
    counts[6] = v->setup.tileFrusta.size();
    return 0;
}

int hvvr_host_view_copy(const hvvr_host_view* v, int array, void* dst, uint64_t bytes) {
    if (!v || !dst)
        return fail("null argument");
    switch (array) {
        case HVVR_VIEW_BEAMS: return copyOut(v->setup.beams, dst, bytes);
        case HVVR_VIEW_TILE_FRUSTA: return copyOut(v->setup.tileFrusta, dst, bytes);
        case HVVR_VIEW_BLOCK_FRUSTA: return copyOut(v->setup.blockFrusta, dst, bytes);
        case HVVR_VIEW_REMAP: return copyOut(v->setup.remap, dst, bytes);
        default: return fail("unknown view array");
    }
}

int hvvr_host_view_camera(const hvvr_host_view* v, float m[16], float lens[2]) {
    if (!v)
        return fail("null argument");
    if (m)
        memcpy(m, &v->cameraToWorld, 64);
    if (lens) {
        lens[0] = v->lens.radius;
        lens[1] = v->lens.focalDistance;
    }
    return 0;
}

int hvvr_host_view_sample_space(const hvvr_host_view* v, float s2c[9], float bounds[4]) {
    if (!v)
        return fail("null argument");
    if (s2c)
        memcpy(s2c, &v->sampleToCam, 36);
    if (bounds) {
        bounds[0] = v->setup.boundsLower.x;
        bounds[1] = v->setup.boundsLower.y;
        bounds[2] = v->setup.boundsUpper.x;
        bounds[3] = v->setup.boundsUpper.y;
    }
    return 0;
}

int hvvr_host_foveated_create(const float pattern[4], hvvr_host_foveated** out) {
    if (!out || !pattern)
        return fail("null argument");
    FoveatedSamplePattern p;
    p.degreeTrackingError = pattern[0];
    p.maxFOVDegrees = pattern[1];
    p.marSlope = pattern[2];
    p.fovealMARDegrees = pattern[3];
    hvvr_host_foveated* f = new hvvr_host_foveated();
    std::string err;
    if (!buildFoveatedSetup(p, f->setup, &err)) {
        delete f;
        return fail(err);
    }
    *out = f;
    return 0;
}

void hvvr_host_foveated_destroy(hvvr_host_foveated* f) {
    delete f;
}

int hvvr_host_foveated_counts(const hvvr_host_foveated* f, uint64_t counts[6]) {
    if (!f || !counts)
        return fail("null argument");
    counts[0] = f->setup.validSampleCount;
    counts[1] = f->setup.sampleCount;
    counts[2] = f->setup.blockCount;
    counts[3] = f->setup.samplesPerRing;
    counts[4] = f->setup.ringCount;
    counts[5] = f->setup.tileFrusta.size();
    return 0;
}

int hvvr_host_foveated_copy(const hvvr_host_foveated* f, int array, void* dst, uint64_t bytes) {
    if (!f || !dst)
        return fail("null argument");
    switch (array) {
        case HVVR_FOVEATED_BEAMS: return copyOut(f->setup.beams, dst, bytes);
        case HVVR_FOVEATED_TILE_FRUSTA: return copyOut(f->setup.tileFrusta, dst, bytes);
        case HVVR_FOVEATED_BLOCK_FRUSTA: return copyOut(f->setup.blockFrusta, dst, bytes);
        case HVVR_FOVEATED_POLAR_REMAP: return copyOut(f->setup.polarRemap, dst, bytes);
        default: return fail("unknown foveated array");
    }
}

int hvvr_host_foveated_params(const hvvr_host_foveated* f, float emap[6], float* maxEccentricity) {
    if (!f)
        return fail("null argument");
    if (emap) {
        const EccentricityMap& e = f->setup.emap;
        float v[6] = {e.m, e.w0, e.S, e.invLogA, e.invW0, e.invM};
        memcpy(emap, v, sizeof(v));
    }
    if (maxEccentricity)
        *maxEccentricity = f->setup.maxEccentricity;
    return 0;
}

int hvvr_host_foveated_frame(const float cameraToWorld[16], const float eyeDir[3], const float s2c[9],
                             const float bounds[4], float history[25], float out[57]) {
    if (!cameraToWorld || !eyeDir || !s2c || !bounds || !history || !out)
        return fail("null argument");
    matrix4x4 m;
    memcpy(&m, cameraToWorld, 64);
    matrix3x3 sampleToCam;
    memcpy(&sampleToCam, s2c, 36);
    FoveatedHistory h;
    memcpy(&h.worldToEyePrevious, history, 64);
    memcpy(&h.eyePreviousToSamplePrevious, history + 16, 36);
    FoveatedFrame fr;
    foveatedFrame(m, vector3(eyeDir[0], eyeDir[1], eyeDir[2]), sampleToCam, vector2(bounds[0], bounds[1]),
                  vector2(bounds[2], bounds[3]), h, fr);
    memcpy(history, &h.worldToEyePrevious, 64);
    memcpy(history + 16, &h.eyePreviousToSamplePrevious, 36);
    memcpy(out, &fr.eyeToWorld, 64);
    memcpy(out + 16, fr.cullPlanes, 64);
    memcpy(out + 32, &fr.sampleToEye, 36);
    memcpy(out + 41, &fr.eyeToSamplePrevious, 64);
    return 0;
}

int hvvr_host_grid_samples(uint32_t width, uint32_t height, float k1, float k2, void* dst) {
    if (!dst)
        return fail("null argument");
    std::vector<Sample> s = getGridSamples(width, height);
    if (k1 != 0.0f || k2 != 0.0f)
        applyBarrelDistortion(s, k1, k2);
    memcpy(dst, s.data(), s.size() * sizeof(Sample));
    return 0;
}

int hvvr_host_subsample_offsets(uint32_t msaaRate, float spinAngle, float* out) {
    if (!out || msaaRate < 1 || msaaRate > 32)
        return fail("msaaRate must be in [1, 32]");
    subsampleOffsets(msaaRate, spinAngle, (vector2*)out);
    return 0;
}

int hvvr_host_lens_table(uint32_t msaaRate, float* out) {
    if (!out || msaaRate < 1 || msaaRate > 32)
        return fail("msaaRate must be in [1, 32]");
    std::vector<vector2> t;
    lensPositionTable(msaaRate, t);
    memcpy(out, t.data(), t.size() * sizeof(vector2));
    return 0;
}

int hvvr_host_transform_matrix(const float tr[3], const float rot[4], float scale, float out[16]) {
    if (!tr || !rot || !out)
        return fail("null argument");
    matrix4x4 m(transform(vector3(tr[0], tr[1], tr[2]), quaternion(rot[0], rot[1], rot[2], rot[3]), scale));
    memcpy(out, &m, 64);
    return 0;
}

} // extern "C"
