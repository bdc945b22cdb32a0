// Renders a scene through the hvvr::Raycaster C++ API (include/hvvr/raycaster.h), the
// way an application of the reference drives its raycaster (projects/hvvr_samples/
// modelviewer/modelviewer.cpp: create models, lights and a camera, set samples and a
// render target, call render() every frame).
//
//   render_scene <scene> <out.rgba> [frames] [foveated | spherical=FX,FY | share-samples]
//                [size=WxH] [msaa=N] [camera=TX,TY,TZ,QX,QY,QZ,QW] [dump=PREFIX[,text]]
//
// dump=PREFIX writes Raycaster::dumpSceneAndRays(camera, PREFIX) after the first frame
// (binary OFF / BFF / RFF files, or the text forms with ",text").
// <scene> is either a file in the reference's binary scene format (model_import.h; loaded
// with model_import::load and createObjects, as modelviewer.cpp:71-77 does), rendered at
// size (default 640x480) and msaa (default 4) from `camera` (translation and rotation
// quaternion, default identity) through the viewport (-W/H, -1)..(W/H, 1) that
// modelviewer.cpp:160 sets, or an "HVSC" test scene file (below), which carries its own
// camera, size and MSAA rate.
//
// "foveated" renders in GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT mode with the feb2017
// sample pattern (raycaster_spec.h:38-47), gaze straight ahead. "spherical=FX,FY" sets
// Camera::setSphericalWarpSettings(FX, FY) before the samples (rays from a spherical
// section of FX x FY degrees, as modelviewer.cpp:168 does for its wide view).
// "share-samples" adds a second camera that takes the first one's sample hierarchy
// through getSampleData / setSampleData (camera.h:96-97); its image follows the first's.
//
// scene.bin (little endian, packed):
//   char[4] "HVSC", u32 version = 1
//   u32 modelCount; per model:
//     u32 vertexCount, triangleCount, materialCount
//     f32 translation[3], rotation[4] (quaternion x y z w), scale
//     ShadingVertex[vertexCount] (32 B), u32 indices[3*triangleCount] (model-local),
//     u32 material[triangleCount] (model-local), SimpleMaterial[materialCount] (64 B)
//   u32 lightCount, LightUnion[lightCount] (56 B)
//   u32 width, height, msaaRate
//   f32 viewport[4] (lower.x lower.y upper.x upper.y)
//   f32 cameraTranslation[3], cameraRotation[4], cameraScale
//   u32 moveModel; if 1: f32 translation[3], rotation[4], scale for model 0, applied
//     after the first frame (a transform-only update)
//
// out.rgba: one width*height RGBA8 image after the first render, and a second one after
// the model-0 move when moveModel is 1. Prints one JSON line with host-inclusive frame
// times (render() through the image copy into the CPU render target).

#include "hvvr/model_import.h"
#include "hvvr/raycaster.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

using namespace hvvr;

namespace {

struct Reader {
    FILE* f;
    template <class T>
    T get() {
        T v;
        read(&v, sizeof(T));
        return v;
    }
    void read(void* dst, size_t bytes) {
        if (bytes && fread(dst, 1, bytes, f) != bytes)
            throw std::runtime_error("scene file truncated");
    }
    transform getTransform() {
        float t[8];
        read(t, sizeof(t));
        return transform(vector3(t[0], t[1], t[2]), quaternion(t[3], t[4], t[5], t[6]), t[7]);
    }
};

} // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr,
                "usage: %s scene out.rgba [frames] [foveated | spherical=FX,FY | share-samples] [size=WxH] "
                "[msaa=N] [camera=TX,TY,TZ,QX,QY,QZ,QW]\n",
                argv[0]);
        return 2;
    }
    int frames = 1;
    std::string mode;
    uint32_t optW = 640, optH = 480, optMsaa = 4;
    float cam[7] = {0, 0, 0, 0, 0, 0, 1};
    std::string dumpPrefix;
    bool dumpBinary = true;
    for (int i = 3; i < argc; i++) {
        const std::string a = argv[i];
        if (!a.empty() && a.find_first_not_of("0123456789") == std::string::npos) {
            frames = std::max(1, atoi(a.c_str()));
        } else if (a.rfind("size=", 0) == 0) {
            if (sscanf(a.c_str() + 5, "%ux%u", &optW, &optH) != 2 || !optW || !optH) {
                fprintf(stderr, "size=WxH: two positive integers\n");
                return 2;
            }
        } else if (a.rfind("dump=", 0) == 0) {
            dumpPrefix = a.substr(5);
            const size_t comma = dumpPrefix.rfind(",text");
            if (comma != std::string::npos && comma + 5 == dumpPrefix.size()) {
                dumpPrefix.resize(comma);
                dumpBinary = false;
            }
        } else if (a.rfind("msaa=", 0) == 0) {
            optMsaa = uint32_t(atoi(a.c_str() + 5));
        } else if (a.rfind("camera=", 0) == 0) {
            if (sscanf(a.c_str() + 7, "%f,%f,%f,%f,%f,%f,%f", cam, cam + 1, cam + 2, cam + 3, cam + 4, cam + 5,
                       cam + 6) != 7) {
                fprintf(stderr, "camera=TX,TY,TZ,QX,QY,QZ,QW: seven numbers\n");
                return 2;
            }
        } else {
            mode = a;
        }
    }
    try {
        FILE* f = fopen(argv[1], "rb");
        if (!f)
            throw std::runtime_error(std::string("cannot open ") + argv[1]);
        Reader r{f};
        char magic[4];
        r.read(magic, 4);
        const bool testScene = memcmp(magic, "HVSC", 4) == 0;
        if (testScene && r.get<uint32_t>() != 1)
            throw std::runtime_error("not a version-1 HVSC scene file");

        struct ModelIn {
            MeshData mesh;
            transform xf;
        };
        std::vector<ModelIn> models(testScene ? r.get<uint32_t>() : 0u);
        for (ModelIn& m : models) {
            uint32_t nv = r.get<uint32_t>(), nt = r.get<uint32_t>(), nm = r.get<uint32_t>();
            m.xf = r.getTransform();
            m.mesh.verts.resize(nv);
            r.read(m.mesh.verts.data(), nv * sizeof(ShadingVertex));
            std::vector<uint32_t> idx(3 * size_t(nt)), mat(nt);
            r.read(idx.data(), idx.size() * 4);
            r.read(mat.data(), mat.size() * 4);
            m.mesh.triShade.resize(nt);
            for (uint32_t t = 0; t < nt; t++) {
                for (int k = 0; k < 3; k++)
                    m.mesh.triShade[t].indices[k] = idx[3 * size_t(t) + k];
                m.mesh.triShade[t].material = mat[t];
            }
            m.mesh.materials.resize(nm);
            r.read(m.mesh.materials.data(), nm * sizeof(SimpleMaterial));
        }
        std::vector<LightUnion> lights;
        uint32_t width = optW, height = optH, msaa = optMsaa;
        float vp[4] = {-float(optW) / float(optH), -1.0f, float(optW) / float(optH), 1.0f};
        transform cameraToWorld(vector3(cam[0], cam[1], cam[2]), quaternion(cam[3], cam[4], cam[5], cam[6]), 1.0f);
        uint32_t moveModel = 0;
        transform moved = transform::identity();
        if (testScene) {
            lights.resize(r.get<uint32_t>());
            r.read(lights.data(), lights.size() * sizeof(LightUnion));
            width = r.get<uint32_t>();
            height = r.get<uint32_t>();
            msaa = r.get<uint32_t>();
            r.read(vp, sizeof(vp));
            cameraToWorld = r.getTransform();
            moveModel = r.get<uint32_t>();
            if (moveModel)
                moved = r.getTransform();
        }
        fclose(f);
        model_import::Model binModel;
        if (!testScene) {
            std::string err;
            if (!model_import::load(argv[1], binModel, &err))
                throw std::runtime_error(std::string(argv[1]) + ": " + err);
        }

        float fovX = 0.0f, fovY = 0.0f;
        if (mode.rfind("spherical=", 0) == 0 && sscanf(mode.c_str() + 10, "%f,%f", &fovX, &fovY) != 2)
            throw std::runtime_error("spherical=FX,FY: two angles in degrees");
        RayCasterSpecification spec;
        spec.msaaRate = msaa;
        if (mode == "foveated") {
            spec.mode = RayCasterGPUMode::GPU_FOVEATED_POLAR_SPACE_CUDA_RECONSTRUCT;
            spec.foveatedSamplePattern.degreeTrackingError = 7.0f;
            spec.foveatedSamplePattern.fovealMARDegrees = 1.0f / 20.0f;
            spec.foveatedSamplePattern.marSlope = 0.015f;
            spec.foveatedSamplePattern.maxFOVDegrees = 90.0f;
        }
        Raycaster raycaster(spec);
        std::vector<Model*> handles;
        for (ModelIn& m : models) {
            GenerateTopology(m.mesh);
            Model* model = raycaster.createModel(std::move(m.mesh));
            model->setTransform(m.xf);
            handles.push_back(model);
        }
        for (const LightUnion& l : lights)
            raycaster.createLight(l);
        if (!testScene && !model_import::createObjects(raycaster, binModel))
            throw std::runtime_error("createObjects failed");

        Camera* camera = raycaster.createCamera(FloatRect(vector2(vp[0], vp[1]), vector2(vp[2], vp[3])), 0.0f);
        std::vector<uint32_t> image(size_t(width) * height, 0);
        ImageResourceDescriptor target;
        target.memoryType = ImageResourceDescriptor::MemoryType::CPU_MEMORY;
        target.data = image.data();
        target.width = width;
        target.height = height;
        target.stride = width;
        camera->setRenderTarget(target);
        if (fovX > 0.0f)
            camera->setSphericalWarpSettings(fovX, fovY);
        std::vector<Sample> samples = getGridSamples(width, height);
        camera->setSamples(samples.data(), uint32_t(samples.size()), 1);
        camera->setCameraToWorld(cameraToWorld);
        std::vector<uint32_t> image2;
        if (mode == "share-samples") {
            Camera* second = raycaster.createCamera(camera->getViewport(), 0.0f);
            image2.assign(image.size(), 0);
            ImageResourceDescriptor t2 = target;
            t2.data = image2.data();
            second->setRenderTarget(t2);
            second->setSampleData(camera->getSampleData());
            second->setCameraToWorld(cameraToWorld);
        }

        FILE* out = fopen(argv[2], "wb");
        if (!out)
            throw std::runtime_error(std::string("cannot open ") + argv[2]);
        std::vector<double> ms;
        for (int i = 0; i < frames; i++) {
            auto t0 = std::chrono::steady_clock::now();
            raycaster.render(0.0);
            auto t1 = std::chrono::steady_clock::now();
            ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
            if (i == 0 && !dumpPrefix.empty())
                raycaster.dumpSceneAndRays(camera, dumpPrefix, dumpBinary);
        }
        fwrite(image.data(), 4, image.size(), out);
        if (!image2.empty())
            fwrite(image2.data(), 4, image2.size(), out);
        if (moveModel && !handles.empty()) {
            handles[0]->setTransform(moved);
            raycaster.render(0.0);
            fwrite(image.data(), 4, image.size(), out);
        }
        fclose(out);
        double first = ms[0], rest = 0.0;
        for (size_t i = 1; i < ms.size(); i++)
            rest += ms[i];
        printf("{\"frames\": %d, \"first_ms\": %.3f, \"mean_ms_after_first\": %.3f, \"samples\": %u}\n", frames,
               first, ms.size() > 1 ? rest / double(ms.size() - 1) : first, camera->getSampleCount());
    } catch (const RaycasterError& e) {
        fprintf(stderr, "RaycasterError: %s\n", e.what());
        return 3;
    } catch (const std::exception& e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
