#pragma once
// Polar foveated sampling: the eccentricity map and the polar-texture coordinate
// mapping, shared by the host sample setup and the reconstruction kernels.
//
//   EccentricityMap                        raycaster/foveated.h:29-112
//   FoveatedSamplePattern                  raycaster/raycaster_spec.h:29-47
//   EccentricityToTexCoordMapping + helpers raycaster/gpu_foveated.cu:75-130

#include "hvvr/types.h"

#include <math.h>

namespace hvvr {

// the reference's float constants (shared/constants_math.h:15-17)
constexpr float kPi = 3.141592653589f;
constexpr float kRadiansPerDegree = kPi / 180.0f;

struct FoveatedSamplePattern {
    float degreeTrackingError = 0.5f;
    float maxFOVDegrees = 110.0f;
    float marSlope = 0.022f;
    float fovealMARDegrees = 1.0f / 60.0f;
    // RayCasterSpecification::feb2017FoveatedDemoSettings
    static FoveatedSamplePattern feb2017() {
        FoveatedSamplePattern p;
        p.degreeTrackingError = 7.0f;
        p.fovealMARDegrees = 1.0f / 20.0f;
        p.marSlope = 0.015f;
        p.maxFOVDegrees = 90.0f;
        return p;
    }
};

// Ring index <-> eccentricity (radians). Sample spacing is w0 out to eccentricity S, then
// grows linearly with eccentricity beyond it (slope m): the continuous form of
// E(n+1) = E(n) + w0 + m * max(E(n) - S, 0).
struct EccentricityMap {
    float m = 0.0f, w0 = 0.0f, S = 0.0f;
    float invLogA = 0.0f; // 1 / log(1 + m)
    float invW0 = 0.0f, invM = 0.0f;

    HVVR_HD float apply(float i) const {
        float e0 = i * w0;
        float x = (e0 - S) * invW0;
        float gx = (w0 * (powf(1.0f + m, x) - 1.0f)) * invM;
        float e1 = gx + S;
        return (e0 < S) ? e0 : e1;
    }
    HVVR_HD float applyInverse(float E) const {
        float i0 = E * invW0;
        float sDivW = S * invW0;
        float x = logf((m * (i0 - sDivW)) + 1.0f) * invLogA;
        float i1 = x + sDivW;
        return (E < S) ? i0 : i1;
    }
};

// marSlope, foveal MAR and tracking error in degrees; the reciprocals in double
inline EccentricityMap makeEccentricityMap(float marSlope, float fovealMARDegrees, float trackingErrorDegrees) {
    EccentricityMap e;
    e.m = marSlope;
    e.w0 = fovealMARDegrees * kRadiansPerDegree;
    e.S = trackingErrorDegrees * kRadiansPerDegree;
    e.invLogA = float(1.0 / log(1.0 + double(e.m)));
    e.invW0 = float(1.0 / double(e.w0));
    e.invM = float(1.0 / double(e.m));
    return e;
}

// polar texture: x = azimuth column (samplesPerRing wide, wraps), y = ring (ringCount high)
struct EccentricityToTexCoord {
    EccentricityMap emap;
    float texMapSize;       // ring count
    float invTexMapSize;
    float invMaxEccentricity;
};

HVVR_HD float eccentricityToTexCoord(float eccentricity, const EccentricityToTexCoord& t) {
    return (t.emap.applyInverse(eccentricity) + 0.5f) * t.invTexMapSize;
}

// unit direction (eye space, looking down -z) at azimuth theta and eccentricity e
HVVR_HD vector3 angularEyeCoordToDirection(float theta, float e) {
    float z = -cosf(e);
    float xyLength = sqrtf(1.0f - z * z);
    return vector3(cosf(-theta) * xyLength, sinf(-theta) * xyLength, z);
}

HVVR_HD void eyeSpaceDirectionToAngularEyeCoord(const vector3& dir, float& theta, float& eccentricity) {
    eccentricity = acosf(-dir.z);
    theta = -atan2f(dir.y, dir.x);
}

HVVR_HD vector2 angularEyeCoordToPolarTextureCoord(float theta, float eccentricity, const EccentricityToTexCoord& t) {
    return vector2((theta + kPi) / (2.0f * kPi), eccentricityToTexCoord(eccentricity, t));
}

} // namespace hvvr
