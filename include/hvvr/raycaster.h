#pragma once
// hvvr::Raycaster C++ API for the MI355X render path: the same object model and entry
// points as the reference's public API (raycaster/raycaster.h:38-115, camera.h:51-146,
// model.h:26-57, light.h:19-77), implemented on the GPU C ABI of hvvr_gpu.h.
//
// Differences a caller sees:
//  * RayCasterSpecification::msaaRate selects the MSAA rate at run time (the reference
//    compiles it in as COLOR_MODE_MSAA_RATE = 16, raycaster_common.h:27);
//  * errors throw hvvr::RaycasterError instead of printing and exiting;
//  * render targets are CPU memory (ImageResourceDescriptor::MemoryType::CPU_MEMORY);
//    D3D11/GL interop and textures are outside this build;
//  * depth of field follows the camera's lens (setLens / setAperture / setFocalDepth);
//  * in the foveated mode (GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT) each camera traces
//    the spec's polar sample pattern around its eye direction (setEyeDir, camera space)
//    and reconstructs its render target; the camera's own samples set the screen
//    rectangle tiles are culled to, as in the reference.

#include "types.h"

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct hvvr_gpu_context;
struct hvvr_gpu_scene;
struct hvvr_gpu_view;
struct hvvr_gpu_foveated_frame;

namespace hvvr {

class RaycasterError : public std::runtime_error {
public:
    explicit RaycasterError(const std::string& what) : std::runtime_error(what) {}
};

// graphics_types.h:25-31
struct FloatRect {
    vector2 lower, upper;
    FloatRect() = default;
    FloatRect(const vector2& l, const vector2& u) : lower(l), upper(u) {}
};

// graphics_types.h:37-81 (CPU memory only)
struct ImageResourceDescriptor {
    enum class MemoryType { CPU_MEMORY, CUDA_MEMORY, DX_TEXTURE, OPENGL_TEXTURE };
    MemoryType memoryType = MemoryType::CPU_MEMORY;
    void* data = nullptr;
    uint32_t width = 0;
    uint32_t height = 0;
    size_t stride = 0; // in pixels
};

// model.h:26-31
struct MeshData {
    std::vector<ShadingVertex> verts;
    std::vector<PrecomputedTriangleShade> triShade;
    std::vector<TopologyNode> nodes;
    std::vector<SimpleMaterial> materials;
};

// samples.h:17-29
struct Sample {
    vector2ui pixelLocation;
    vector2 position;
    struct Extents {
        vector2 minorAxis;
        float majorAxisLength;
    } extents;
};

// raycaster_spec.h:22-50
enum class RaycasterOutputFormat { COLOR_RGBA8 };
enum class RayCasterGPUMode { GPU_INTERSECT_AND_RECONSTRUCT_DEFERRED_MSAA_RESOLVE, GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT };
struct RayCasterSpecification {
    RayCasterGPUMode mode = RayCasterGPUMode::GPU_INTERSECT_AND_RECONSTRUCT_DEFERRED_MSAA_RESOLVE;
    size_t threadCount = 0;
    struct FoveatedSamplePattern {
        float degreeTrackingError = 0.5f;
        float maxFOVDegrees = 110.0f;
        float marSlope = 0.022f;
        float fovealMARDegrees = 1.0f / 60.0f;
    } foveatedSamplePattern;
    RaycasterOutputFormat outputFormat = RaycasterOutputFormat::COLOR_RGBA8;
    bool outputTo3DApi = false;
    uint32_t msaaRate = 16; // 1, 2, 4, 8, 16 or 32
    int device = 0;
};

// How sample-space positions in [0,1]^2 become camera-space rays
// (sample_hierarchy.h:73-95): the perspective transform of the camera's viewport, or a
// spherical section of fovXDegrees x fovYDegrees (yaw and pitch linear in the sample
// position, sample_hierarchy.cpp:27-35).
struct Sample2Dto3DMappingSettings {
    enum class MappingType { Perspective, SphericalSection };
    MappingType type = MappingType::Perspective;
    matrix3x3 sampleToCamera = matrix3x3::identity();
    ThinLens thinLens = {0.0f, 1.0f};
    float fovXDegrees = 0.0f;
    float fovYDegrees = 0.0f;
    Sample2Dto3DMappingSettings() {}
    Sample2Dto3DMappingSettings(const matrix3x3& s2c, ThinLens lens) : sampleToCamera(s2c), thinLens(lens) {}
    static Sample2Dto3DMappingSettings sphericalSection(const matrix3x3& s2c, ThinLens lens, float fovXDegrees,
                                                       float fovYDegrees) {
        Sample2Dto3DMappingSettings r(s2c, lens);
        r.type = MappingType::SphericalSection;
        r.fovXDegrees = fovXDegrees;
        r.fovYDegrees = fovYDegrees;
        return r;
    }
};

// A camera's sample hierarchy (camera.h:28-47): samples clustered into blocks of 8192 and
// tiles of 128 (padded to whole blocks), camera-space beams and tile / block frusta, and
// the image-location -> sample map. Built from raw samples by the constructor, or taken
// from another camera (Camera::getSampleData / setSampleData).
struct SampleData {
    Sample2Dto3DMappingSettings settings2DTo3D;
    uint32_t validSampleCount = 0;
    uint32_t sampleCount = 0; // padded to a multiple of 8192
    uint32_t blockCount = 0;
    uint32_t splitColorSamples = 1;
    uint32_t width = 0, height = 0;             // the render target the remap addresses
    std::vector<DirectionalBeam> beams;         // [sampleCount]
    std::vector<SimpleRayFrustum> tileFrusta;   // [blockCount * 64]
    std::vector<SimpleRayFrustum> blockFrusta;  // [blockCount]
    std::vector<int32_t> remap;                 // imageLocationToSampleIndex: [width * height * split], -1 = none
    vector2 boundsLower, boundsUpper;           // sampleBounds
    SampleData() = default;
    // camera.cpp:21-64
    SampleData(const Sample* rawSamples, uint32_t rawSampleCount, uint32_t splitColorSamples,
               const Sample2Dto3DMappingSettings& settings2DTo3D, uint32_t rtWidth, uint32_t rtHeight);
};

// SAH BVH4 topology for a mesh, reordering triangles and vertices (shared/bvh.h:15).
bool GenerateTopology(MeshData& mesh);
// Grid samples of a width x height render target (samples.cpp:21-42).
std::vector<Sample> getGridSamples(size_t width, size_t height);

// texture.h:16-80: texel formats, a texture's description and its byte size
enum class TextureFormat : uint32_t {
    none = 0,
    r8g8b8a8_unorm_srgb,
    r8g8b8a8_unorm,
    r16g16b16a16_unorm,
    r32g32b32a32_float,
    r16g16b16a16_float,
    r11g11b10_float,
    r32_float,
    count
};
struct TextureData {
    const uint8_t* data;
    TextureFormat format;
    uint32_t width;
    uint32_t height;
    uint32_t strideElements;
};
// bytes of strideElements x height texels; 0 for `none` or an unknown format
size_t getTextureSize(uint32_t strideElements, uint32_t height, TextureFormat format);

class Raycaster;

// A texture the raycaster holds (texture.h:36-54); getID() is the index materials use in
// their texture fields (IDs count up from 0 in creation order, as the reference's atlas
// does). The next updateScene uploads every live texture; shading reads them with wrap
// addressing, trilinear mip filtering and up to 8 anisotropic probes. A material naming a
// destroyed texture reads a black 1x1 stand-in; one naming an ID never created shades
// with its constant colour.
class Texture {
public:
    Texture(const TextureData& data, uint32_t id);
    uint32_t getID() const { return _textureID; }
    const TextureData& getData() const { return _data; }

protected:
    uint32_t _textureID;
    TextureData _data;         // data points into _texels
    std::vector<uint8_t> _texels;
};

class Light {
public:
    explicit Light(const LightUnion& u) : _lightUnion(u) {}
    const LightUnion& getLightUnion() const { return _lightUnion; }

protected:
    LightUnion _lightUnion;
};

class Model {
public:
    Model(Raycaster& raycaster, MeshData&& meshData);
    Model(const Model&) = delete;
    Model& operator=(const Model&) = delete;
    const MeshData& getMesh() const { return _meshData; }
    void setTransform(const transform& t);
    const transform& getTransform() const { return _transform; }

protected:
    Raycaster& _raycaster;
    MeshData _meshData;
    transform _transform;
};

class Camera {
    friend class Raycaster;

public:
    Camera(Raycaster& raycaster, const FloatRect& viewport, float apertureRadius);
    ~Camera();
    Camera(const Camera&) = delete;
    Camera& operator=(const Camera&) = delete;

    void setEnabled(bool enabled) { _enabled = enabled; }
    bool getEnabled() const { return _enabled; }
    void setEyeDir(const vector3& eyeDir) { _eyeDir = eyeDir; }
    const vector3& getEyeDir() const { return _eyeDir; }
    void setViewport(const FloatRect& viewport);
    const FloatRect& getViewport() const { return _viewport; }
    void setLens(const ThinLens& lens);
    const ThinLens& getLens() const { return _lens; }
    void setAperture(float radius);
    float getAperture() const { return _lens.radius; }
    void setFocalDepth(float depth);
    void setRenderTarget(const ImageResourceDescriptor& target);
    void setSamples(const Sample* samples, uint32_t count, uint32_t splitColorSamples);
    // Nonzero angles: rays from a spherical section of fovXDegrees x fovYDegrees instead of
    // the viewport's perspective transform (camera.cpp:206-224); applies from the next
    // setSamples (or render, if samples are set).
    void setSphericalWarpSettings(float fovXDegrees, float fovYDegrees);
    // The sample hierarchy in use (camera.cpp:135-160). setSampleData replaces it as a
    // whole (its remap addresses the render target's size); setSamples rebuilds it.
    void setSampleData(const SampleData& sampleData);
    const SampleData& getSampleData() const { return _sampleData; }
    Sample2Dto3DMappingSettings get2DSampleMappingSettings() const;
    uint32_t getSampleCount() const { return _validSampleCount; }
    matrix3x3 getSampleToCamera() const;
    void setCameraToWorld(const transform& cameraToWorld) { _cameraToWorld = cameraToWorld; }
    matrix4x4 getCameraToWorld() const { return matrix4x4(_cameraToWorld); }
    const vector3& getTranslation() const { return _cameraToWorld.translation; }
    // Copies the last rendered image into the CPU render target (render() does this for
    // every enabled camera that has one).
    void extractImage();

protected:
    void rebuildView();
    void rebuildFoveatedView();
    void foveatedFrame(hvvr_gpu_foveated_frame& out);
    vector2 nextJitter();

    Raycaster& _raycaster;
    FloatRect _viewport;
    ThinLens _lens = {0.0f, 1.0f};
    bool _enabled = true;
    vector3 _eyeDir = vector3(0.0f, 0.0f, -1.0f);
    ImageResourceDescriptor _renderTarget;
    transform _cameraToWorld = transform::identity();
    std::vector<Sample> _samples;
    SampleData _sampleData;
    bool _sampleDataGiven = false; // setSampleData: render from _sampleData as given
    float _fovXDegrees = 0.0f, _fovYDegrees = 0.0f;
    uint32_t _split = 1;
    uint32_t _validSampleCount = 0;
    bool _viewDirty = true;
    hvvr_gpu_view* _view = nullptr;
    std::vector<vector2> _frameJitters;
    uint32_t _frameCount = 0;
    // foveated mode: previous-frame eye transforms (camera.h:121-122)
    matrix4x4 _worldToEyePrevious = matrix4x4(vector4(0.0f), vector4(0.0f), vector4(0.0f), vector4(0.0f));
    matrix3x3 _eyePreviousToSamplePrevious = matrix3x3::identity();
};

class Raycaster {
    friend class Camera;
    friend class Model;

public:
    explicit Raycaster(const RayCasterSpecification& spec);
    ~Raycaster();
    Raycaster(const Raycaster&) = delete;
    Raycaster& operator=(const Raycaster&) = delete;

    Camera* createCamera(const FloatRect& viewport, float apertureRadius);
    void destroyCamera(Camera* camera);
    // raycaster.h:51-52: the texels are copied
    Texture* createTexture(const TextureData& textureData);
    void destroyTexture(Texture* texture);
    Light* createLight(const LightUnion& lightUnion);
    void destroyLight(Light* light);
    Model* createModel(MeshData&& meshData);
    Model* createModel(const MeshData& meshData);
    void destroyModel(Model* model);
    void setSceneDirty();
    void render(double elapsedTime);
    // dumpSceneAndRays (render.cpp:175-201, compiled in there by DUMP_SCENE_AND_RAYS): the
    // models' BVH topology (<prefix>bvh_dump.bff), world-space triangles
    // (<prefix>scene_dump.off) and the camera's subsample rays in image order
    // (<prefix>ray_dump.rff) for camera-to-world as last set; binary forms add a 'b' to each
    // name. The camera must have been rendered (its sample setup is on the GPU).
    void dumpSceneAndRays(Camera* camera, const std::string& prefix = "", bool binary = true);
    void reinit(RayCasterGPUMode mode);

    // GPU handles, for callers that mix this API with the C ABI (hvvr_gpu.h)
    hvvr_gpu_context* gpuContext() const { return _ctx; }
    hvvr_gpu_scene* gpuScene() const { return _scene; }

protected:
    void updateScene();
    void uploadTextures();
    bool growLists(hvvr_gpu_view* view);
    void transformsChanged() { _transformsDirty = true; }

    RayCasterSpecification _spec;
    std::vector<std::unique_ptr<Camera>> _cameras;
    std::vector<std::unique_ptr<Texture>> _textures;
    uint32_t _nextTextureID = 0;
    std::vector<std::unique_ptr<Light>> _lights;
    std::vector<std::unique_ptr<Model>> _models;
    bool _sceneDirty = false;
    bool _transformsDirty = false;
    bool _lightsDirty = true;
    bool _texturesDirty = false;
    hvvr_gpu_context* _ctx = nullptr;
    hvvr_gpu_scene* _scene = nullptr;
};

} // namespace hvvr
