// Seeded synthetic scenes and camera presets for the benchmark configurations.
// Meshes are closed, wound counter-clockwise seen from outside (the intersection test
// rejects back faces), and every vertex is referenced (GenerateTopology requires it).

#include "host_internal.h"

#include <math.h>

namespace hvvr {

namespace {

ShadingVertex makeVertex(vector3 p, vector3 n, vector2 uv) {
    ShadingVertex v;
    v.pos = p;
    n = normalize(n);
    v.normal[0] = floatToHalfRNE(n.x);
    v.normal[1] = floatToHalfRNE(n.y);
    v.normal[2] = floatToHalfRNE(n.z);
    v.normal[3] = floatToHalfRNE(0.0f);
    v.uv = uv;
    v.pad0 = 0;
    return v;
}

void addTri(MeshData& m, uint32_t a, uint32_t b, uint32_t c, uint32_t material) {
    PrecomputedTriangleShade t;
    t.indices[0] = a;
    t.indices[1] = b;
    t.indices[2] = c;
    t.material = material;
    m.triShade.push_back(t);
}

// grid face: origin + s*u + t*v for s,t in [0,1], normal = cross(u, v) direction
void addGridFace(MeshData& m, vector3 origin, vector3 u, vector3 v, uint32_t nu, uint32_t nv, uint32_t material) {
    vector3 n = cross(u, v);
    uint32_t base = uint32_t(m.verts.size());
    for (uint32_t j = 0; j <= nv; j++)
        for (uint32_t i = 0; i <= nu; i++) {
            float s = float(i) / float(nu), t = float(j) / float(nv);
            m.verts.push_back(makeVertex(origin + u * s + v * t, n, vector2(s, t)));
        }
    for (uint32_t j = 0; j < nv; j++)
        for (uint32_t i = 0; i < nu; i++) {
            uint32_t a = base + j * (nu + 1) + i;
            uint32_t b = a + 1;
            uint32_t c = a + (nu + 1);
            uint32_t d = c + 1;
            addTri(m, a, b, d, material);
            addTri(m, a, d, c, material);
        }
}

// axis-aligned box [lo, hi] with faces subdivided (nx, ny, nz) along each axis
void addBox(MeshData& m, vector3 lo, vector3 hi, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t material) {
    vector3 e = hi - lo;
    vector3 X(e.x, 0, 0), Y(0, e.y, 0), Z(0, 0, e.z);
    // +x face: normal +x = cross(Y, Z)
    addGridFace(m, vector3(hi.x, lo.y, lo.z), Y, Z, ny, nz, material);
    // -x face: cross(Z, Y) = -x
    addGridFace(m, vector3(lo.x, lo.y, lo.z), Z, Y, nz, ny, material);
    // +y face: cross(Z, X) = +y
    addGridFace(m, vector3(lo.x, hi.y, lo.z), Z, X, nz, nx, material);
    // -y face: cross(X, Z) = -y
    addGridFace(m, vector3(lo.x, lo.y, lo.z), X, Z, nx, nz, material);
    // +z face: cross(X, Y) = +z
    addGridFace(m, vector3(lo.x, lo.y, hi.z), X, Y, nx, ny, material);
    // -z face: cross(Y, X) = -z
    addGridFace(m, vector3(lo.x, lo.y, lo.z), Y, X, ny, nx, material);
}

// cylinder of radius 0.5, height 1 around +y, centred at the origin: 2*seg side + 2*seg cap triangles
void addCylinder(MeshData& m, uint32_t seg, uint32_t material) {
    uint32_t base = uint32_t(m.verts.size());
    for (uint32_t ring = 0; ring < 2; ring++) {
        float y = ring ? 0.5f : -0.5f;
        for (uint32_t i = 0; i < seg; i++) {
            float a = 2.0f * kPi * float(i) / float(seg);
            vector3 p(0.5f * cosf(a), y, -0.5f * sinf(a));
            m.verts.push_back(makeVertex(p, vector3(p.x, 0.0f, p.z), vector2(float(i) / seg, float(ring))));
        }
    }
    for (uint32_t i = 0; i < seg; i++) {
        uint32_t i1 = (i + 1) % seg;
        uint32_t b0 = base + i, b1 = base + i1, t0 = base + seg + i, t1 = base + seg + i1;
        // counter-clockwise around +y seen from outside
        addTri(m, b0, b1, t1, material);
        addTri(m, b0, t1, t0, material);
    }
    for (uint32_t cap = 0; cap < 2; cap++) {
        float y = cap ? 0.5f : -0.5f;
        vector3 n(0.0f, cap ? 1.0f : -1.0f, 0.0f);
        uint32_t center = uint32_t(m.verts.size());
        m.verts.push_back(makeVertex(vector3(0.0f, y, 0.0f), n, vector2(0.5f, 0.5f)));
        uint32_t ringBase = uint32_t(m.verts.size());
        for (uint32_t i = 0; i < seg; i++) {
            float a = 2.0f * kPi * float(i) / float(seg);
            m.verts.push_back(makeVertex(vector3(0.5f * cosf(a), y, -0.5f * sinf(a)), n, vector2(0.0f, 0.0f)));
        }
        for (uint32_t i = 0; i < seg; i++) {
            uint32_t a = ringBase + i, b = ringBase + (i + 1) % seg;
            if (cap)
                addTri(m, center, a, b, material);
            else
                addTri(m, center, b, a, material);
        }
    }
}

SimpleMaterial ggxMaterial(float roughness, vector3 albedo) {
    SimpleMaterial mat = {};
    mat.emissive = vector4(0.0f, 0.0f, 0.0f, 0.0f);
    mat.diffuse = vector4(albedo, 1.0f);
    mat.specular = vector4(0.04f, 0.04f, 0.04f, 1.0f);
    mat.glossiness = 1.0f - roughness;
    mat.opacity = 1.0f;
    mat.textureIDsAndShadingModel =
        SimpleMaterial::buildShadingCode(ShadingModelPhong, SimpleMaterial::badTextureIndex,
                                         SimpleMaterial::badTextureIndex, SimpleMaterial::badTextureIndex);
    return mat;
}

quaternion randomRotation(Pcg32& rng) {
    // uniform random unit quaternion (Shoemake)
    double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
    double a = sqrt(1.0 - u1), b = sqrt(u1);
    double t2 = 6.283185307179586 * u2, t3 = 6.283185307179586 * u3;
    return quaternion(float(a * sin(t2)), float(a * cos(t2)), float(b * sin(t3)), float(b * cos(t3)));
}

LightUnion directionalLight(vector3 dir, vector3 power) {
    LightUnion l;
    l.type = LightTypeDirectional;
    l.directional.Direction = normalize(dir);
    l.directional.Power = power;
    return l;
}

LightUnion pointLight(vector3 pos, vector3 color, float falloffStart, float falloffEnd) {
    LightUnion l;
    l.type = LightTypePoint;
    l.point.Position = pos;
    l.point.FalloffEnd = falloffEnd;
    l.point.Color = color;
    l.point.FalloffScale = 1.0f / (falloffEnd - falloffStart);
    return l;
}

} // namespace

void generateHeightfieldBoxesScene(uint64_t seed, SceneGenResult& out) {
    out = SceneGenResult();
    Pcg32 rng(seed, 1);
    const uint32_t cells = 64;
    const float extent = 16.0f;
    // heightfield: one model, per-triangle materials
    {
        MeshData m;
        std::vector<float> h((cells + 1) * (cells + 1));
        for (float& v : h)
            v = rng.uniformf(0.0f, 1.0f);
        float step = extent / float(cells);
        auto H = [&](int x, int z) {
            x = x < 0 ? 0 : (x > int(cells) ? int(cells) : x);
            z = z < 0 ? 0 : (z > int(cells) ? int(cells) : z);
            return h[size_t(z) * (cells + 1) + x];
        };
        for (uint32_t z = 0; z <= cells; z++)
            for (uint32_t x = 0; x <= cells; x++) {
                vector3 n(H(int(x) - 1, int(z)) - H(int(x) + 1, int(z)), 2.0f * step,
                          H(int(x), int(z) - 1) - H(int(x), int(z) + 1));
                m.verts.push_back(makeVertex(vector3(x * step, H(int(x), int(z)), z * step), n,
                                             vector2(float(x) / cells, float(z) / cells)));
            }
        uint32_t material = 0;
        for (uint32_t z = 0; z < cells; z++)
            for (uint32_t x = 0; x < cells; x++) {
                uint32_t a = z * (cells + 1) + x, b = a + 1, c = a + (cells + 1), d = c + 1;
                // counter-clockwise seen from +y
                addTri(m, a, c, d, material++);
                addTri(m, a, d, b, material++);
            }
        for (uint32_t i = 0; i < material; i++) {
            float roughness = rng.uniformf(0.1f, 0.9f);
            vector3 albedo(rng.uniformf(0, 1), rng.uniformf(0, 1), rng.uniformf(0, 1));
            m.materials.push_back(ggxMaterial(roughness, albedo));
        }
        out.meshes.push_back(std::move(m));
        out.modelToWorld.push_back(matrix4x4::identity());
        out.transforms.push_back(transform::identity());
    }
    // 256 axis-aligned boxes, one model each, 12 triangles with per-triangle materials
    for (uint32_t b = 0; b < 256; b++) {
        MeshData m;
        vector3 c(rng.uniformf(0.0f, extent), 0.0f, rng.uniformf(0.0f, extent));
        vector3 e;
        e.x = float(0.1 * pow(20.0, rng.uniform()));
        e.y = float(0.1 * pow(20.0, rng.uniform()));
        e.z = float(0.1 * pow(20.0, rng.uniform()));
        c.y = rng.uniformf(0.0f, 1.0f);
        addBox(m, c - e * 0.5f, c + e * 0.5f, 1, 1, 1, 0);
        for (uint32_t t = 0; t < m.triShade.size(); t++) {
            m.triShade[t].material = t;
            float roughness = rng.uniformf(0.1f, 0.9f);
            vector3 albedo(rng.uniformf(0, 1), rng.uniformf(0, 1), rng.uniformf(0, 1));
            m.materials.push_back(ggxMaterial(roughness, albedo));
        }
        out.meshes.push_back(std::move(m));
        out.modelToWorld.push_back(matrix4x4::identity());
        out.transforms.push_back(transform::identity());
    }
    out.lights.push_back(directionalLight(vector3(0.4f, 0.8f, 0.45f), vector3(2.5f, 2.4f, 2.2f)));
}

void generateRoomScene(uint64_t seed, uint32_t meshCount, SceneGenResult& out) {
    out = SceneGenResult();
    Pcg32 rng(seed, 2);
    for (uint32_t i = 0; i < meshCount; i++) {
        MeshData m;
        bool cylinder = (rng.next() & 1u) != 0;
        if (cylinder)
            addCylinder(m, 64, 0);
        else
            addBox(m, vector3(-0.5f), vector3(0.5f), 4, 4, 6, 0);
        float roughness = rng.uniformf(0.1f, 0.9f);
        vector3 albedo(rng.uniformf(0, 1), rng.uniformf(0, 1), rng.uniformf(0, 1));
        m.materials.push_back(ggxMaterial(roughness, albedo));
        float scale = float(exp(0.7 * rng.normal()));
        quaternion rot = randomRotation(rng);
        vector3 pos(rng.uniformf(0.0f, 30.0f), rng.uniformf(0.0f, 15.0f), rng.uniformf(0.0f, 30.0f));
        out.transforms.push_back(transform(pos, rot, scale));
        out.modelToWorld.push_back(matrix4x4(out.transforms.back()));
        out.meshes.push_back(std::move(m));
    }
    out.lights.push_back(directionalLight(vector3(-0.3f, 0.9f, 0.3f), vector3(1.2f, 1.2f, 1.1f)));
    for (int i = 0; i < 4; i++) {
        vector3 p(rng.uniformf(3.0f, 27.0f), rng.uniformf(3.0f, 12.0f), rng.uniformf(3.0f, 27.0f));
        vector3 c(rng.uniformf(2.0f, 6.0f), rng.uniformf(2.0f, 6.0f), rng.uniformf(2.0f, 6.0f));
        out.lights.push_back(pointLight(p, c, 2.0f, 18.0f));
    }
}

CameraPreset heightfieldCamera() {
    CameraPreset c;
    c.width = 640;
    c.height = 480;
    float t = tanf(30.0f * kPi / 180.0f);
    c.viewportLower = vector2(-t * 4.0f / 3.0f, -t);
    c.viewportUpper = vector2(t * 4.0f / 3.0f, t);
    c.barrelK1 = 0.0f;
    c.barrelK2 = 0.0f;
    float pitch = -40.0f * kPi / 180.0f;
    c.cameraToWorld = transform(vector3(8.0f, 8.0f, 19.0f), quaternion::fromAxisAngle(vector3(1, 0, 0), pitch), 1.0f);
    c.lens = {0.0f, 1.0f};
    return c;
}

CameraPreset roomEyeCamera(uint32_t eye, uint32_t width, uint32_t height) {
    CameraPreset c;
    c.width = width;
    c.height = height;
    c.barrelK1 = 0.22f;
    c.barrelK2 = 0.24f;
    // 100 degree horizontal field at the centre row after distortion (magnification 1.46
    // at the horizontal edge)
    float sx = tanf(50.0f * kPi / 180.0f) / 1.46f;
    float sy = sx * float(height) / float(width);
    c.viewportLower = vector2(-sx, -sy);
    c.viewportUpper = vector2(sx, sy);
    float ipd = 0.064f;
    float x = 15.0f + (eye ? 0.5f : -0.5f) * ipd;
    c.cameraToWorld = transform(vector3(x, 7.5f, 29.0f), quaternion::fromAxisAngle(vector3(0, 1, 0), 0.0f), 1.0f);
    c.lens = {0.0f, 1.0f};
    return c;
}

} // namespace hvvr
