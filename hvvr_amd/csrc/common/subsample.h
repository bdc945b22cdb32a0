#pragma once
// Spiral subsample pattern shared by the host library (tests, camera setup) and the
// GPU library (per-view offset tables): tapLocation / getSubsampleUnitOffset
// (raycaster/gpu_samples.h:23-61) with startOffset 0.5, evaluated in double and
// rounded once so that every consumer of a table sees the same bits.

#include "hvvr/types.h"

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

} // namespace hvvr
