// Sample generation and the camera-space sample hierarchy (per camera configuration,
// not per frame):
//   * grid samples (raycaster/samples.cpp:21-42) and a radial barrel distortion of the
//     sample positions (the reference takes distorted samples through
//     Camera::setSamples, camera.cpp:135-138);
//   * clustering into blocks of 8192 / tiles of 128 (sample_clustering.h:17-49);
//   * padded 2-D tile/block bounds (sample_hierarchy.cpp:154-190), 3-D camera-space
//     frusta (get3DFrustumFrom2D, :46-123) and per-sample DirectionalBeams
//     (generateFrom2D, :129-152), through the viewport's perspective transform or a
//     spherical section (sphericalUVToDirection, :27-35);
//   * the image-location -> sample remap (camera.cpp:57-63).

#include "host_internal.h"
#include "cluster.h"

#include <algorithm>
#include <limits>
#include <random>

namespace hvvr {

std::vector<Sample> getGridSamples(size_t width, size_t height) {
    float invW = 1.0f / float(width);
    float invH = 1.0f / float(height);
    std::vector<Sample> s(width * height);
    for (uint32_t y = 0; y < height; y++) {
        for (uint32_t x = 0; x < width; x++) {
            Sample& o = s[size_t(y) * width + x];
            o.pixelLocation = {x, y};
            o.position = vector2((float(x) + 0.5f) * invW, (float(y) + 0.5f) * invH);
            if (width > height) {
                o.extents.minorAxis = vector2(invW, 0.0f);
                o.extents.majorAxisLength = invH;
            } else {
                o.extents.minorAxis = vector2(0.0f, invH);
                o.extents.majorAxisLength = invW;
            }
        }
    }
    return s;
}

void applyBarrelDistortion(std::vector<Sample>& samples, float k1, float k2) {
    // isotropic frame: unit = half the sample-space width in pixels
    if (samples.empty())
        return;
    uint32_t w = 0, h = 0;
    for (const Sample& s : samples) {
        w = std::max(w, s.pixelLocation.x + 1);
        h = std::max(h, s.pixelLocation.y + 1);
    }
    double aspect = double(h) / double(w);
    for (Sample& s : samples) {
        double ux = 2.0 * double(s.position.x) - 1.0;
        double uy = (2.0 * double(s.position.y) - 1.0) * aspect;
        double r2 = ux * ux + uy * uy;
        double f = 1.0 + k1 * r2 + k2 * r2 * r2;
        double g = 1.0 + 3.0 * k1 * r2 + 5.0 * k2 * r2 * r2; // radial magnification d(r f)/dr
        s.position.x = float(0.5 + 0.5 * ux * f);
        s.position.y = float(0.5 + 0.5 * (uy * f) / aspect);
        s.extents.minorAxis = vector2(float(s.extents.minorAxis.x * g), float(s.extents.minorAxis.y * g));
        s.extents.majorAxisLength = float(s.extents.majorAxisLength * g);
    }
}

matrix3x3 sampleToCamera(vector2 lower, vector2 upper) {
    return matrix3x3(vector3(upper.x - lower.x, 0.0f, 0.0f), vector3(0.0f, lower.y - upper.y, 0.0f),
                     vector3(lower.x, upper.y, -1.0f));
}

namespace {

struct SortedSample {
    Sample s;
    uint32_t channel;
};

struct Unpacked {
    vector2 center, majorAxis, minorAxis;
};
// samples.cpp:144-159 (ellipse axes scaled by sqrt(2)/2)
Unpacked unpackSample(const Sample& s) {
    const float extentModifier = 0.70710678118f;
    Unpacked u;
    u.center = s.position;
    u.minorAxis = vector2(s.extents.minorAxis.x * extentModifier, s.extents.minorAxis.y * extentModifier);
    float rescale = s.extents.majorAxisLength * extentModifier / length(s.extents.minorAxis);
    u.majorAxis = vector2(-u.minorAxis.y * rescale, u.minorAxis.x * rescale);
    return u;
}

// 2-D bounds stored as (xMin, -xMax, yMin, -yMax) (sample_hierarchy.h:19-58)
struct Bounds2D {
    float d[4];
    void setEmpty() {
        for (float& f : d)
            f = std::numeric_limits<float>::infinity();
    }
    void merge(float x, float y) {
        d[0] = fminv(d[0], x);
        d[1] = fminv(d[1], -x);
        d[2] = fminv(d[2], y);
        d[3] = fminv(d[3], -y);
    }
    void merge(const Bounds2D& o) {
        for (int i = 0; i < 4; i++)
            d[i] = fminv(d[i], o.d[i]);
    }
    float xMin() const { return d[0]; }
    float xMax() const { return -d[1]; }
    float yMin() const { return d[2]; }
    float yMax() const { return -d[3]; }
};

// A sample-space position as a direction on a spherical section: yaw and pitch linear in
// the position, 0 at the centre (sample_hierarchy.cpp:27-35).
vector3 sphericalUVToDirection(vector2 uv, float fovXDegrees, float fovYDegrees) {
    const float radiansPerDegree = 3.14159265358979f / 180.0f;
    float yaw = (uv.x - 0.5f) * (fovXDegrees * radiansPerDegree);
    float pitch = -(uv.y - 0.5f) * (fovYDegrees * radiansPerDegree);
    return normalize(vector3(sinf(yaw) * cosf(pitch), sinf(pitch), -cosf(yaw) * cosf(pitch)));
}

// sample_hierarchy.cpp:46-123
SimpleRayFrustum frustumFrom2D(const Bounds2D& b, const Sample2Dto3DMappingSettings& m) {
    const matrix3x3& s2c = m.sampleToCamera;
    const ThinLens lens = m.thinLens;
    const float farPlane = -100.0f;
    vector3 nearPoints[4] = {vector3(-lens.radius, -lens.radius, 0.0f), vector3(+lens.radius, -lens.radius, 0.0f),
                             vector3(+lens.radius, +lens.radius, 0.0f), vector3(-lens.radius, +lens.radius, 0.0f)};
    vector3 rayDirections[4] = {s2c * vector3(b.xMin(), b.yMax(), 1.0f), s2c * vector3(b.xMax(), b.yMax(), 1.0f),
                                s2c * vector3(b.xMax(), b.yMin(), 1.0f), s2c * vector3(b.xMin(), b.yMin(), 1.0f)};
    if (m.type == Sample2Dto3DMappingSettings::MappingType::SphericalSection) {
        // the corner rays of the padded sample-space rectangle on the sphere, from the
        // lens square (the reference fits no far plane here)
        const vector2 uv[4] = {vector2(b.xMin(), b.yMax()), vector2(b.xMax(), b.yMax()), vector2(b.xMax(), b.yMin()),
                               vector2(b.xMin(), b.yMin())};
        SimpleRayFrustum f;
        for (int i = 0; i < 4; i++) {
            f.origins[i] = nearPoints[i];
            f.directions[i] = sphericalUVToDirection(uv[i], m.fovXDegrees, m.fovYDegrees);
        }
        return f;
    }
    vector3 focus[4];
    for (int i = 0; i < 4; i++)
        focus[i] = rayDirections[i] * (lens.focalDistance / -rayDirections[i].z);
    float farXMax = -std::numeric_limits<float>::infinity();
    float farXMin = std::numeric_limits<float>::infinity();
    float farYMax = farXMax, farYMin = farXMin;
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 4; j++) {
            vector3 rayPos = nearPoints[i];
            vector3 rayDir = normalize(focus[j] - rayPos);
            vector3 p = rayPos + rayDir * (farPlane / rayDir.z);
            farXMax = fmaxv(farXMax, p.x);
            farXMin = fminv(farXMin, p.x);
            farYMax = fmaxv(farYMax, p.y);
            farYMin = fminv(farYMin, p.y);
        }
    }
    vector3 farPoints[4] = {vector3(farXMin, farYMin, farPlane), vector3(farXMax, farYMin, farPlane),
                            vector3(farXMax, farYMax, farPlane), vector3(farXMin, farYMax, farPlane)};
    SimpleRayFrustum f;
    for (int i = 0; i < 4; i++) {
        f.origins[i] = nearPoints[i];
        f.directions[i] = normalize(farPoints[i] - nearPoints[i]);
    }
    return f;
}

} // namespace

void buildSampleSetup(const Sample* raw, uint32_t count, uint32_t split, uint32_t width, uint32_t height,
                      const Sample2Dto3DMappingSettings& mapping, SampleSetup& out) {
    out = SampleSetup();
    out.settings2DTo3D = mapping;
    const matrix3x3& s2c = mapping.sampleToCamera;
    const bool spherical = mapping.type == Sample2Dto3DMappingSettings::MappingType::SphericalSection;
    out.splitColorSamples = split;
    out.width = width;
    out.height = height;
    out.validSampleCount = count;
    if (count == 0)
        return;
    std::vector<SortedSample> sorted(count);
    for (uint32_t n = 0; n < count; n++)
        sorted[n] = {raw[n], n % split};
    uint32_t blockCount = (count + BLOCK_SIZE - 1) / BLOCK_SIZE;
    naiveXYCluster(sorted, blockCount, [](const SortedSample& x) { return x.s.position; });

    out.boundsLower = vector2(1.0f, 1.0f);
    out.boundsUpper = vector2(0.0f, 0.0f);
    for (const SortedSample& s : sorted) {
        out.boundsLower.x = std::min(out.boundsLower.x, s.s.position.x);
        out.boundsLower.y = std::min(out.boundsLower.y, s.s.position.y);
        out.boundsUpper.x = std::max(out.boundsUpper.x, s.s.position.x);
        out.boundsUpper.y = std::max(out.boundsUpper.y, s.s.position.y);
    }

    out.blockCount = blockCount;
    out.sampleCount = blockCount * BLOCK_SIZE;
    out.beams.resize(out.sampleCount);
    out.tileFrusta.resize(size_t(blockCount) * TILES_PER_BLOCK);
    out.blockFrusta.resize(blockCount);

    uint32_t maxIndex = count - 1;
    uint32_t sampleIndex = 0;
    for (uint32_t b = 0; b < blockCount; b++) {
        Bounds2D block;
        block.setEmpty();
        for (uint32_t t = 0; t < TILES_PER_BLOCK; t++) {
            Bounds2D tile;
            tile.setEmpty();
            for (uint32_t i = 0; i < TILE_SIZE; i++) {
                const Sample& s = sorted[sampleIndex].s;
                float major = s.extents.majorAxisLength;
                tile.merge(s.position.x + major, s.position.y + major);
                tile.merge(s.position.x - major, s.position.y - major);
                // padding samples repeat the last valid sample
                uint32_t slot = b * BLOCK_SIZE + t * TILE_SIZE + i;
                Unpacked u = unpackSample(s);
                DirectionalBeam& beam = out.beams[slot];
                if (spherical) {
                    const float fx = mapping.fovXDegrees, fy = mapping.fovYDegrees;
                    beam.centerRay = sphericalUVToDirection(u.center, fx, fy);
                    beam.du = sphericalUVToDirection(u.center + u.majorAxis, fx, fy) - beam.centerRay;
                    beam.dv = sphericalUVToDirection(u.center + u.minorAxis, fx, fy) - beam.centerRay;
                } else {
                    beam.centerRay = s2c * vector3(u.center, 1.0f);
                    beam.du = s2c * vector3(u.majorAxis, 0.0f);
                    beam.dv = s2c * vector3(u.minorAxis, 0.0f);
                }
                sampleIndex = std::min(sampleIndex + 1, maxIndex);
            }
            // the infinite cull rect leaves the bounds unchanged
            block.merge(tile);
            out.tileFrusta[size_t(b) * TILES_PER_BLOCK + t] = frustumFrom2D(tile, mapping);
        }
        out.blockFrusta[b] = frustumFrom2D(block, mapping);
    }

    out.remap.assign(size_t(width) * height * split, -1);
    for (uint32_t i = 0; i < count; i++) {
        vector2ui loc = sorted[i].s.pixelLocation;
        size_t idx = (size_t(loc.y) * width + loc.x) * split + sorted[i].channel;
        if (idx < out.remap.size())
            out.remap[idx] = int32_t(i);
    }
}

SampleData::SampleData(const Sample* rawSamples, uint32_t rawSampleCount, uint32_t splitColorSamples,
                       const Sample2Dto3DMappingSettings& settings2DTo3D, uint32_t rtWidth, uint32_t rtHeight) {
    buildSampleSetup(rawSamples, rawSampleCount, splitColorSamples, rtWidth, rtHeight, settings2DTo3D, *this);
}

} // namespace hvvr
