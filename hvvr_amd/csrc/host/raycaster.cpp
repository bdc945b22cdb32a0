// hvvr::Raycaster / Camera / Model / Light over the GPU C ABI (include/hvvr_gpu.h).
//
// Object model and call sequence follow the reference's public API:
//   Raycaster ctor/dtor/reinit          raycaster/raycaster.cpp:30-76
//   createCamera/destroyCamera          raycaster/raycaster.cpp:78-91
//   createLight/createModel/...         raycaster/scene.cpp:43-80
//   render(elapsedTime)                 raycaster/render.cpp:204-218, :382-395
//   updateScene/uploadScene             raycaster/scene_update.cpp:145-174
//   Camera setters, setSamples, jitter  raycaster/camera.cpp:70-209
//   Model::setTransform                 raycaster/model.cpp:20-30
//
// Differences, all on the host side: a transform-only change re-uploads the model
// matrices and refits on the GPU instead of rebuilding the flattened scene; lights are
// re-uploaded whenever one is created or destroyed (the reference only uploads them with
// the geometry); the enabled cameras of one render() go to the GPU as one batched
// submission; a camera with a CPU render target gets its image copied back at the end of
// render() (the reference copies only to 3D-API targets, render.cpp:259-268).

#include <algorithm>
#include "host_internal.h"

#include "hvvr_gpu.h"

#include <random>
#include <string.h>

namespace hvvr {

namespace {

void check(int rc, const char* what) {
    if (rc != HVVR_GPU_OK)
        throw RaycasterError(std::string(what) + ": " + hvvr_gpu_last_error());
}

bool validMsaaRate(uint32_t a) {
    return a == 1 || a == 2 || a == 4 || a == 8 || a == 16 || a == 32;
}

template <class T>
void eraseOwned(std::vector<std::unique_ptr<T>>& v, T* p, const char* what) {
    for (auto it = v.begin(); it != v.end(); ++it) {
        if (it->get() == p) {
            v.erase(it);
            return;
        }
    }
    throw RaycasterError(std::string(what) + ": object does not belong to this raycaster");
}

} // namespace

// ---------------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------------
Model::Model(Raycaster& raycaster, MeshData&& meshData)
    : _raycaster(raycaster), _meshData(std::move(meshData)), _transform(transform::identity()) {}

void Model::setTransform(const transform& t) {
    _transform = t;
    _raycaster.transformsChanged();
}

// ---------------------------------------------------------------------------------
// Camera
// ---------------------------------------------------------------------------------
Camera::Camera(Raycaster& raycaster, const FloatRect& viewport, float apertureRadius)
    : _raycaster(raycaster), _viewport(viewport), _lens{apertureRadius, 1.0f} {
    // 16 frame jitters in [0,1)^2 from a default-seeded mt19937 (camera.cpp:74-79)
    std::uniform_real_distribution<float> uniform(0.0f, 1.0f);
    std::mt19937 generator;
    for (int i = 0; i < 16; ++i) {
        float x = uniform(generator);
        float y = uniform(generator);
        _frameJitters.push_back(vector2(x, y));
    }
}

Camera::~Camera() {
    if (_view)
        hvvr_gpu_view_destroy(_view);
}

void Camera::setViewport(const FloatRect& viewport) {
    _viewport = viewport;
    _viewDirty = true;
}

void Camera::setLens(const ThinLens& lens) {
    _lens = lens;
    _viewDirty = true;
}

void Camera::setAperture(float radius) {
    _lens.radius = radius;
    _viewDirty = true;
}

void Camera::setFocalDepth(float depth) {
    _lens.focalDistance = depth;
    _viewDirty = true;
}

void Camera::setRenderTarget(const ImageResourceDescriptor& target) {
    if (target.memoryType != ImageResourceDescriptor::MemoryType::CPU_MEMORY)
        throw RaycasterError("Camera::setRenderTarget: only CPU_MEMORY render targets are supported");
    if (target.stride < target.width)
        throw RaycasterError("Camera::setRenderTarget: stride (in pixels) is smaller than the width");
    bool resized = target.width != _renderTarget.width || target.height != _renderTarget.height;
    _renderTarget = target;
    if (resized)
        _viewDirty = true;
}

void Camera::setSamples(const Sample* samples, uint32_t count, uint32_t splitColorSamples) {
    if (!samples && count)
        throw RaycasterError("Camera::setSamples: null samples");
    if (splitColorSamples < 1 || splitColorSamples > 3)
        throw RaycasterError("Camera::setSamples: splitColorSamples must be 1, 2 or 3");
    _samples.assign(samples, samples + count);
    _split = splitColorSamples;
    _sampleDataGiven = false;
    _viewDirty = true;
    // the reference builds the sample hierarchy here (camera.cpp:135-156)
    rebuildView();
}

void Camera::setSphericalWarpSettings(float fovXDegrees, float fovYDegrees) {
    if (!(fovXDegrees >= 0.0f && fovYDegrees >= 0.0f && fovXDegrees <= 360.0f && fovYDegrees <= 180.0f))
        throw RaycasterError("Camera::setSphericalWarpSettings: angles in [0, 360] x [0, 180] degrees");
    _fovXDegrees = fovXDegrees;
    _fovYDegrees = fovYDegrees;
    if (!_sampleDataGiven)
        _viewDirty = true;
}

// camera.cpp:206-212
Sample2Dto3DMappingSettings Camera::get2DSampleMappingSettings() const {
    if (_fovXDegrees > 0.0f)
        return Sample2Dto3DMappingSettings::sphericalSection(getSampleToCamera(), _lens, _fovXDegrees, _fovYDegrees);
    return Sample2Dto3DMappingSettings(getSampleToCamera(), _lens);
}

void Camera::setSampleData(const SampleData& sampleData) {
    const SampleData& d = sampleData;
    if (d.blockCount * 8192u != d.sampleCount || d.beams.size() != d.sampleCount ||
        d.tileFrusta.size() != size_t(d.blockCount) * 64 || d.blockFrusta.size() != d.blockCount ||
        d.validSampleCount > d.sampleCount || d.splitColorSamples < 1 || d.splitColorSamples > 3 ||
        d.remap.size() != size_t(d.width) * d.height * d.splitColorSamples)
        throw RaycasterError("Camera::setSampleData: inconsistent sample data");
    _sampleData = sampleData;
    _sampleDataGiven = true;
    _viewDirty = true;
    rebuildView();
}

matrix3x3 Camera::getSampleToCamera() const {
    return sampleToCamera(_viewport.lower, _viewport.upper);
}

void Camera::rebuildView() {
    if (_view) {
        hvvr_gpu_view_destroy(_view);
        _view = nullptr;
    }
    _validSampleCount = 0;
    const bool haveSamples = _sampleDataGiven ? _sampleData.validSampleCount > 0 : !_samples.empty();
    if (!haveSamples || _renderTarget.width == 0 || _renderTarget.height == 0) {
        _viewDirty = true; // retried on the next render once samples and a target exist
        return;
    }
    if (_sampleDataGiven) {
        if (_sampleData.width != _renderTarget.width || _sampleData.height != _renderTarget.height)
            throw RaycasterError("Camera::setSampleData: the sample data maps a " + std::to_string(_sampleData.width) +
                                 "x" + std::to_string(_sampleData.height) + " render target, not " +
                                 std::to_string(_renderTarget.width) + "x" + std::to_string(_renderTarget.height));
    } else {
        for (const Sample& s : _samples)
            if (s.pixelLocation.x >= _renderTarget.width || s.pixelLocation.y >= _renderTarget.height)
                throw RaycasterError("Camera::setSamples: sample pixel location outside the render target");
    }

    if (_raycaster._spec.mode == RayCasterGPUMode::GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT) {
        rebuildFoveatedView();
        return;
    }

    if (!_sampleDataGiven)
        buildSampleSetup(_samples.data(), uint32_t(_samples.size()), _split, _renderTarget.width,
                         _renderTarget.height, get2DSampleMappingSettings(), _sampleData);
    const SampleData& setup = _sampleData;

    hvvr_gpu_view_desc d = {};
    d.msaaRate = _raycaster._spec.msaaRate;
    d.sampleCount = setup.sampleCount;
    d.validSampleCount = setup.validSampleCount;
    d.beams = setup.beams.data();
    d.tileFrusta = setup.tileFrusta.data();
    d.blockFrusta = setup.blockFrusta.data();
    d.remap = setup.remap.data();
    d.width = setup.width;
    d.height = setup.height;
    d.split = setup.splitColorSamples;
    d.lensRadius = _lens.radius;
    d.focalDistance = _lens.focalDistance;
    check(hvvr_gpu_view_create(_raycaster._ctx, &d, &_view), "Camera: view creation failed");
    _validSampleCount = setup.validSampleCount;
    _viewDirty = false;
}

// polarSpaceFoveatedSetup (foveated.cpp:86-113) + GPUCamera::registerPolarFoveatedSamples
// (gpu_camera.cu:194-230): the spec's polar pattern as an eye-space view whose image is
// the render target
void Camera::rebuildFoveatedView() {
    const auto& sp = _raycaster._spec.foveatedSamplePattern;
    FoveatedSamplePattern p;
    p.degreeTrackingError = sp.degreeTrackingError;
    p.maxFOVDegrees = sp.maxFOVDegrees;
    p.marSlope = sp.marSlope;
    p.fovealMARDegrees = sp.fovealMARDegrees;
    FoveatedSetup setup;
    std::string err;
    if (!buildFoveatedSetup(p, setup, &err))
        throw RaycasterError("Camera: " + err);
    hvvr_gpu_view_desc d = {};
    d.msaaRate = _raycaster._spec.msaaRate;
    d.sampleCount = setup.sampleCount;
    d.validSampleCount = setup.validSampleCount;
    d.beams = setup.beams.data();
    d.tileFrusta = setup.tileFrusta.data();
    d.blockFrusta = setup.blockFrusta.data();
    d.remap = nullptr;
    d.width = _renderTarget.width;
    d.height = _renderTarget.height;
    d.split = 1;
    d.lensRadius = _lens.radius;
    d.focalDistance = _lens.focalDistance;
    check(hvvr_gpu_view_create(_raycaster._ctx, &d, &_view), "Camera: foveated view creation failed");
    hvvr_gpu_foveated_desc f = {};
    f.polarRemap = (const uint32_t*)setup.polarRemap.data();
    f.samplesPerRing = setup.samplesPerRing;
    f.ringCount = setup.ringCount;
    const EccentricityMap& e = setup.emap;
    float emap[6] = {e.m, e.w0, e.S, e.invLogA, e.invW0, e.invM};
    memcpy(f.eccentricityMap, emap, sizeof(emap));
    f.maxEccentricity = setup.maxEccentricity;
    f.taaAlpha = 0.05f;
    f.taaStddevMultiplier = 1.0f;
    f.contrastEnhancement = 1;
    f.contrastFe = 1.0f;
    check(hvvr_gpu_view_set_foveated(_view, &f), "Camera: foveated setup failed");
    _validSampleCount = setup.validSampleCount;
    _worldToEyePrevious = matrix4x4(vector4(0.0f), vector4(0.0f), vector4(0.0f), vector4(0.0f));
    _eyePreviousToSamplePrevious = matrix3x3::identity();
    _viewDirty = false;
}

// the frame's eye transforms, screen cull planes and reprojection (render.cpp:454-488)
void Camera::foveatedFrame(hvvr_gpu_foveated_frame& out) {
    // sample bounds of the camera's own samples (camera.cpp:36-42)
    vector2 lo(1.0f, 1.0f), hi(0.0f, 0.0f);
    if (_sampleDataGiven) {
        lo = _sampleData.boundsLower;
        hi = _sampleData.boundsUpper;
    }
    for (const Sample& s : _samples) {
        if (_sampleDataGiven)
            break;
        lo = vector2(std::min(lo.x, s.position.x), std::min(lo.y, s.position.y));
        hi = vector2(std::max(hi.x, s.position.x), std::max(hi.y, s.position.y));
    }
    FoveatedHistory h;
    h.worldToEyePrevious = _worldToEyePrevious;
    h.eyePreviousToSamplePrevious = _eyePreviousToSamplePrevious;
    FoveatedFrame fr;
    hvvr::foveatedFrame(getCameraToWorld(), _eyeDir, getSampleToCamera(), lo, hi, h, fr);
    _worldToEyePrevious = h.worldToEyePrevious;
    _eyePreviousToSamplePrevious = h.eyePreviousToSamplePrevious;
    memcpy(out.eyeToWorld, &fr.eyeToWorld, sizeof(out.eyeToWorld));
    memcpy(out.cullPlanes, fr.cullPlanes, sizeof(out.cullPlanes));
    memcpy(out.sampleToEye, &fr.sampleToEye, sizeof(out.sampleToEye));
    memcpy(out.eyeToSamplePrevious, &fr.eyeToSamplePrevious, sizeof(out.eyeToSamplePrevious));
}

void Camera::extractImage() {
    if (!_view || !_renderTarget.data)
        return;
    check(hvvr_gpu_view_read_image(_view, (uint32_t*)_renderTarget.data, uint32_t(_renderTarget.stride)),
          "Camera::extractImage");
}

vector2 Camera::nextJitter() {
    // camera.cpp:205-209
    vector2 j = _frameJitters[_frameCount % _frameJitters.size()];
    ++_frameCount;
    return vector2(j.x * 0.5f + 0.5f, j.y * 0.5f + 0.5f);
}

// ---------------------------------------------------------------------------------
// Raycaster
// ---------------------------------------------------------------------------------
Raycaster::Raycaster(const RayCasterSpecification& spec) : _spec(spec) {
    if (_spec.outputTo3DApi)
        throw RaycasterError("Raycaster: 3D-API output (D3D11/GL interop) is not supported; use CPU render targets");
    if (!validMsaaRate(_spec.msaaRate))
        throw RaycasterError("Raycaster: msaaRate must be 1, 2, 4, 8, 16 or 32");
    check(hvvr_gpu_context_create(_spec.device, &_ctx), "Raycaster: GPU context creation failed");
}

Raycaster::~Raycaster() {
    _cameras.clear(); // views before the context
    if (_scene)
        hvvr_gpu_scene_destroy(_scene);
    if (_ctx)
        hvvr_gpu_context_destroy(_ctx);
}

void Raycaster::reinit(RayCasterGPUMode mode) {
    // raycaster.cpp:57-76: drops cameras and the GPU scene, keeps models and lights
    if (_ctx)
        hvvr_gpu_context_synchronize(_ctx);
    _cameras.clear();
    if (_scene) {
        hvvr_gpu_scene_destroy(_scene);
        _scene = nullptr;
    }
    if (_ctx) {
        hvvr_gpu_context_destroy(_ctx);
        _ctx = nullptr;
    }
    _spec.mode = mode;
    check(hvvr_gpu_context_create(_spec.device, &_ctx), "Raycaster::reinit: GPU context creation failed");
    _sceneDirty = true;
    _lightsDirty = true;
}

Camera* Raycaster::createCamera(const FloatRect& viewport, float apertureRadius) {
    _cameras.emplace_back(new Camera(*this, viewport, apertureRadius));
    return _cameras.back().get();
}

void Raycaster::destroyCamera(Camera* camera) {
    eraseOwned(_cameras, camera, "Raycaster::destroyCamera");
}

size_t getTextureSize(uint32_t strideElements, uint32_t height, TextureFormat format) {
    const size_t n = size_t(strideElements) * height;
    switch (format) {
        case TextureFormat::r8g8b8a8_unorm_srgb:
        case TextureFormat::r8g8b8a8_unorm:
        case TextureFormat::r11g11b10_float:
        case TextureFormat::r32_float:
            return n * 4;
        case TextureFormat::r16g16b16a16_unorm:
        case TextureFormat::r16g16b16a16_float:
            return n * 8;
        case TextureFormat::r32g32b32a32_float:
            return n * 16;
        default:
            return 0;
    }
}

Texture::Texture(const TextureData& data, uint32_t id) : _textureID(id), _data(data) {
    const size_t bytes = getTextureSize(data.strideElements, data.height, data.format);
    if (data.format == TextureFormat::none || data.format >= TextureFormat::count || data.strideElements < data.width)
        throw RaycasterError("Raycaster::createTexture: invalid texture description");
    if (bytes && !data.data)
        throw RaycasterError("Raycaster::createTexture: null texels");
    _texels.assign(data.data, data.data + bytes);
    _data.data = _texels.data();
}

Texture* Raycaster::createTexture(const TextureData& textureData) {
    if (_nextTextureID >= SimpleMaterial::badTextureIndex)
        throw RaycasterError("Raycaster::createTexture: texture IDs exhausted (65535 is reserved for 'none')");
    _textures.emplace_back(new Texture(textureData, _nextTextureID++));
    _texturesDirty = true;
    return _textures.back().get();
}

void Raycaster::destroyTexture(Texture* texture) {
    eraseOwned(_textures, texture, "Raycaster::destroyTexture");
    _texturesDirty = true;
}

void Raycaster::uploadTextures() {
    // indexed by ID: destroyed IDs below the highest live one hold a black 1x1 stand-in
    uint32_t count = 0;
    for (const auto& t : _textures)
        count = std::max(count, t->getID() + 1);
    static const uint8_t black[4] = {0, 0, 0, 0};
    std::vector<hvvr_gpu_texture_desc> descs(count, hvvr_gpu_texture_desc{uint32_t(TextureFormat::r8g8b8a8_unorm), 1,
                                                                          1, 1, black});
    for (const auto& t : _textures) {
        const TextureData& d = t->getData();
        descs[t->getID()] = {uint32_t(d.format), d.width, d.height, d.strideElements, d.data};
    }
    check(hvvr_gpu_scene_set_textures(_scene, count, descs.data()), "Raycaster: texture upload failed");
    _texturesDirty = false;
}

Light* Raycaster::createLight(const LightUnion& lightUnion) {
    _lights.emplace_back(new Light(lightUnion));
    _lightsDirty = true;
    return _lights.back().get();
}

void Raycaster::destroyLight(Light* light) {
    eraseOwned(_lights, light, "Raycaster::destroyLight");
    _lightsDirty = true;
}

Model* Raycaster::createModel(MeshData&& meshData) {
    if (meshData.nodes.empty())
        throw RaycasterError("Raycaster::createModel: mesh has no BVH (call GenerateTopology first)");
    _models.emplace_back(new Model(*this, std::move(meshData)));
    _sceneDirty = true;
    return _models.back().get();
}

Model* Raycaster::createModel(const MeshData& meshData) {
    return createModel(MeshData(meshData));
}

void Raycaster::destroyModel(Model* model) {
    eraseOwned(_models, model, "Raycaster::destroyModel");
    _sceneDirty = true;
}

void Raycaster::setSceneDirty() {
    _sceneDirty = true;
}

void Raycaster::updateScene() {
    if (_sceneDirty) {
        if (_scene) {
            hvvr_gpu_scene_destroy(_scene);
            _scene = nullptr;
        }
        if (!_models.empty()) {
            std::vector<SceneModel> models;
            models.reserve(_models.size());
            for (const auto& m : _models)
                models.push_back({&m->getMesh(), matrix4x4(m->getTransform())});
            FlatScene flat;
            buildScene(models, flat);
            if (flat.triShade.empty())
                throw RaycasterError("Raycaster: the scene's models hold no triangles");
            hvvr_gpu_scene_desc d = {};
            d.vertices = flat.verts.data();
            d.vertexToTransform = flat.vertexToTransform.data();
            d.vertexCount = uint32_t(flat.verts.size());
            d.triangles = flat.triShade.data();
            d.triangleCount = uint32_t(flat.triShade.size());
            d.nodes = flat.nodes.data();
            d.nodeCount = uint32_t(flat.nodes.size());
            d.materials = flat.materials.data();
            d.materialCount = uint32_t(flat.materials.size());
            d.transforms = (const float*)flat.transforms.data();
            d.transformCount = uint32_t(flat.transforms.size());
            check(hvvr_gpu_scene_create(_ctx, &d, &_scene), "Raycaster: scene upload failed");
        }
        _sceneDirty = false;
        _transformsDirty = false;
        _lightsDirty = true;
        _texturesDirty = true;
    }
    if (!_scene)
        return;
    if (_transformsDirty) {
        std::vector<matrix4x4> xf;
        xf.reserve(_models.size());
        for (const auto& m : _models)
            xf.push_back(matrix4x4(m->getTransform()));
        check(hvvr_gpu_scene_set_transforms(_scene, (const float*)xf.data(), uint32_t(xf.size())),
              "Raycaster: transform upload failed");
        _transformsDirty = false;
    }
    if (_lightsDirty) {
        std::vector<LightUnion> lights;
        lights.reserve(_lights.size());
        for (const auto& l : _lights)
            lights.push_back(l->getLightUnion());
        check(hvvr_gpu_scene_set_lights(_scene, lights.data(), uint32_t(lights.size()), nullptr),
              "Raycaster: light upload failed");
        _lightsDirty = false;
    }
    if (_texturesDirty)
        uploadTextures();
    check(hvvr_gpu_scene_update(_scene), "Raycaster: scene update failed");
}

void Raycaster::dumpSceneAndRays(Camera* camera, const std::string& prefix, bool binary) {
    if (!camera || !camera->_view)
        throw RaycasterError("Raycaster::dumpSceneAndRays: the camera has no view yet (render it first)");
    if (_spec.mode == RayCasterGPUMode::GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT)
        throw RaycasterError("Raycaster::dumpSceneAndRays: the foveated mode has no image-order rays");
    std::vector<const Model*> models;
    for (const auto& m : _models)
        models.push_back(m.get());
    dumpBVH(prefix + "bvh_dump.bff", models, binary);
    dumpSceneToOFF(prefix + "scene_dump.off", models, binary);
    const uint64_t n = uint64_t(camera->_validSampleCount) * _spec.msaaRate;
    std::vector<SimpleRay> rays(n);
    const matrix4x4 m = camera->getCameraToWorld();
    check(hvvr_gpu_view_dump_rays(camera->_view, (const float*)&m, 1, rays.data(), n),
          "Raycaster::dumpSceneAndRays: ray dump failed");
    dumpRaysToRFF(prefix + "ray_dump.rff", rays, binary);
}

void Raycaster::render(double elapsedTime) {
    (void)elapsedTime;
    updateScene();
    if (!_scene)
        return; // no scene geometry is loaded (render.cpp:205-207)

    std::vector<Camera*> active;
    for (auto& c : _cameras) {
        if (!c->getEnabled())
            continue;
        if (c->_viewDirty)
            c->rebuildView();
        if (c->_view)
            active.push_back(c.get());
    }
    if (active.empty())
        return;

    if (_spec.mode == RayCasterGPUMode::GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT) {
        // renderFoveatedPolarSpaceCudaReconstruct (render.cpp:454-488), camera by camera
        std::vector<hvvr_gpu_foveated_frame> frames(active.size());
        std::vector<vector2> jitter(active.size());
        for (size_t i = 0; i < active.size(); i++) {
            active[i]->foveatedFrame(frames[i]);
            jitter[i] = active[i]->nextJitter();
            check(hvvr_gpu_view_render_foveated(active[i]->_view, _scene, &frames[i], &jitter[i].x),
                  "Raycaster::render");
        }
        for (size_t i = 0; i < active.size(); i++) {
            if (growLists(active[i]->_view)) {
                check(hvvr_gpu_view_foveated_rewind(active[i]->_view), "Raycaster::render");
                check(hvvr_gpu_view_render_foveated(active[i]->_view, _scene, &frames[i], &jitter[i].x),
                      "Raycaster::render");
            }
            active[i]->extractImage();
        }
        return;
    }

    std::vector<hvvr_gpu_view*> views;
    std::vector<matrix4x4> cameraToWorld;
    std::vector<vector2> jitter;
    for (Camera* c : active) {
        views.push_back(c->_view);
        cameraToWorld.push_back(c->getCameraToWorld());
        jitter.push_back(c->nextJitter());
    }
    check(hvvr_gpu_render_views(_scene, uint32_t(views.size()), views.data(), (const float*)cameraToWorld.data(),
                                (const float*)jitter.data()),
          "Raycaster::render");
    for (size_t i = 0; i < active.size(); i++) {
        if (growLists(views[i]))
            check(hvvr_gpu_view_render(views[i], _scene, (const float*)&cameraToWorld[i], &jitter[i].x),
                  "Raycaster::render");
        active[i]->extractImage();
    }
}

// After a frame: true when its triangle lists overflowed the view's storage, which has
// then grown to what the frame needed (plus a quarter), for the caller to render again.
// (The reference stops with a message instead, traverse_avx.h:463.)
bool Raycaster::growLists(hvvr_gpu_view* view) {
    uint64_t need = 0, cap = 0;
    int overflowed = 0;
    check(hvvr_gpu_view_list_demand(view, &need, &cap, &overflowed), "Raycaster::render");
    if (!overflowed || need <= cap)
        return false;
    check(hvvr_gpu_view_reserve_lists(view, std::min<uint64_t>(need + need / 4, 0xffffffffull)),
          "Raycaster::render");
    return true;
}

} // namespace hvvr
