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
#include "../common/sample_setup.h"

#include <string.h>

#include <algorithm>
#include <limits>
#include <random>

namespace hvvr {

static_assert(sizeof(Sample) == sizeof(RawSample), "Sample and RawSample share a layout");

std::vector<Sample> getGridSamples(size_t width, size_t height) {
    std::vector<Sample> s(width * height);
    for (uint32_t y = 0; y < height; y++)
        for (uint32_t x = 0; x < width; x++) {
            RawSample r = gridSample(x, y, uint32_t(width), uint32_t(height));
            memcpy(&s[size_t(y) * width + x], &r, sizeof(r));
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
    for (Sample& s : samples) {
        RawSample r;
        memcpy(&r, &s, sizeof(r));
        barrelDistort(r, w, h, k1, k2);
        memcpy(&s, &r, sizeof(r));
    }
}

matrix3x3 sampleToCamera(vector2 lower, vector2 upper) {
    return matrix3x3(vector3(upper.x - lower.x, 0.0f, 0.0f), vector3(0.0f, lower.y - upper.y, 0.0f),
                     vector3(lower.x, upper.y, -1.0f));
}

SampleMapping sampleMapping(const Sample2Dto3DMappingSettings& m) {
    SampleMapping o;
    o.sampleToCamera = m.sampleToCamera;
    o.lens = m.thinLens;
    o.fovXDegrees = m.fovXDegrees;
    o.fovYDegrees = m.fovYDegrees;
    o.spherical = m.type == Sample2Dto3DMappingSettings::MappingType::SphericalSection ? 1 : 0;
    return o;
}

namespace {

struct SortedSample {
    Sample s;
    uint32_t channel;
};

RawSample toRaw(const Sample& s) {
    RawSample r;
    memcpy(&r, &s, sizeof(r));
    return r;
}

} // namespace

void buildSampleSetup(const Sample* raw, uint32_t count, uint32_t split, uint32_t width, uint32_t height,
                      const Sample2Dto3DMappingSettings& mapping, SampleSetup& out) {
    out = SampleSetup();
    out.settings2DTo3D = mapping;
    const SampleMapping m = sampleMapping(mapping);
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
                // padding samples repeat the last valid sample
                const RawSample s = toRaw(sorted[sampleIndex].s);
                tile.mergeSample(s);
                out.beams[b * BLOCK_SIZE + t * TILE_SIZE + i] = sampleBeam(s, m);
                sampleIndex = std::min(sampleIndex + 1, maxIndex);
            }
            // the infinite cull rect leaves the bounds unchanged
            block.merge(tile);
            out.tileFrusta[size_t(b) * TILES_PER_BLOCK + t] = frustumFrom2D(tile, m);
        }
        out.blockFrusta[b] = frustumFrom2D(block, m);
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
