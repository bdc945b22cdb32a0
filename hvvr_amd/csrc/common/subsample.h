#pragma once
// Spiral subsample pattern shared by the host library (tests, camera setup) and the
// GPU library (per-view offset tables): tapLocation / getSubsampleUnitOffset
// (raycaster/gpu_samples.h:23-61) with startOffset 0.5, evaluated in double and
// rounded once so that every consumer of a table sees the same bits.

#include "hvvr/types.h"

#include <math.h>
#include <random>
#include <vector>

namespace hvvr {

// Mara et al. 2016 spiral-turn table (raycaster/gpu_samples.h:23-34); a published
// table of constants that has to be identical.
static const int kSpiralTurns[100] = {
    1,  1,  1,  2,  3,  2,  5,  2,  3,  2,  3,  3,  5,  5,  3,  4,  7,  5,  5,  7,
    9,  8,  5,  5,  7,  7,  7,  8,  5,  8,  11, 12, 7,  10, 13, 8,  11, 8,  7,  14,
    11, 11, 13, 12, 13, 19, 17, 13, 11, 18, 19, 11, 11, 14, 17, 21, 15, 16, 17, 18,
    13, 17, 11, 17, 19, 18, 25, 18, 19, 19, 29, 21, 19, 27, 31, 29, 21, 18, 17, 29,
    31, 31, 23, 18, 25, 26, 25, 23, 19, 34, 19, 27, 21, 25, 39, 29, 17, 21, 27, 29};

inline void subsampleOffsets(uint32_t msaaRate, float spinAngle, vector2* out) {
    // tapLocation / getSubsampleUnitOffset (gpu_samples.h:35-61), startOffset 0.5
    float turns = float(kSpiralTurns[msaaRate < 100 ? msaaRate : 99]) * 6.28f;
    for (uint32_t i = 0; i < msaaRate; i++) {
        float alpha = (float(i) + 0.5f) * (1.0f / float(msaaRate));
        double angle = double(alpha) * double(turns) + double(spinAngle);
        double radius = sqrt(double(alpha));
        out[i] = vector2(float(cos(angle) * radius), float(sin(angle) * radius));
    }
}

// Depth-of-field lens-position table (GPUCamera::initLookupTables, raycaster/gpu_camera.cu:59-88).
inline void lensPositionTable(uint32_t msaaRate, std::vector<vector2>& out) {
    // 4 rows x 128 samples x A subsamples, entry [row * 128 * A + subsample * 128 + sample]:
    // a random rotation per sample from a default-seeded mt19937, subsample (7s+7) mod A
    // of the rotated spiral (gpu_camera.cu:65-85). The uniform draw is mt19937's 32-bit
    // output * 2^-32 (this library's mapping; the reference's
    // uniform_real_distribution<float> bits are implementation-defined).
    const uint32_t tiles = 4;
    out.assign(size_t(tiles) * TILE_SIZE * msaaRate, vector2());
    std::mt19937 gen;
    std::vector<vector2> spiral(msaaRate);
    for (uint32_t tile = 0; tile < tiles; tile++) {
        for (uint32_t sample = 0; sample < TILE_SIZE; sample++) {
            float rotation = float(double(gen()) * (1.0 / 4294967296.0)) * 6.283185307180f;
            subsampleOffsets(msaaRate, rotation, spiral.data());
            for (uint32_t sub = 0; sub < msaaRate; sub++)
                out[size_t(tile) * TILE_SIZE * msaaRate + size_t(sub) * TILE_SIZE + sample] =
                    spiral[(sub * 7 + 7) % msaaRate];
        }
    }
}

} // namespace hvvr
