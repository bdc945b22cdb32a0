// Renders a scene through the hvvr::Raycaster C++ API (include/hvvr/raycaster.h), the
// way an application of the reference drives its raycaster (projects/hvvr_samples/
// modelviewer/modelviewer.cpp: create models, lights and a camera, set samples and a
// render target, call render() every frame).
//
//   render_scene <scene.bin> <out.rgba> [frames] [foveated | spherical=FX,FY | share-samples]
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

#include "hvvr/raycaster.h"

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
        fprintf(stderr, "usage: %s scene.bin out.rgba [frames] [foveated | spherical=FX,FY | share-samples]\n",
                argv[0]);
        return 2;
    }
    int frames = argc > 3 ? atoi(argv[3]) : 1;
    if (frames < 1)
        frames = 1;
    try {
        FILE* f = fopen(argv[1], "rb");
        if (!f)
            throw std::runtime_error(std::string("cannot open ") + argv[1]);
        Reader r{f};
        char magic[4];
        r.read(magic, 4);
        if (memcmp(magic, "HVSC", 4) != 0 || r.get<uint32_t>() != 1)
            throw std::runtime_error("not a version-1 HVSC scene file");

        struct ModelIn {
            MeshData mesh;
            transform xf;
        };
        std::vector<ModelIn> models(r.get<uint32_t>());
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
        std::vector<LightUnion> lights(r.get<uint32_t>());
        r.read(lights.data(), lights.size() * sizeof(LightUnion));
        uint32_t width = r.get<uint32_t>(), height = r.get<uint32_t>(), msaa = r.get<uint32_t>();
        float vp[4];
        r.read(vp, sizeof(vp));
        transform cameraToWorld = r.getTransform();
        uint32_t moveModel = r.get<uint32_t>();
        transform moved = transform::identity();
        if (moveModel)
            moved = r.getTransform();
        fclose(f);

        const std::string mode = argc > 4 ? argv[4] : "";
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
