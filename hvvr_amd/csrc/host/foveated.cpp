// Polar foveated path, host side:
//   * eye-space polar samples (raycaster/samples.cpp:44-142);
//   * their order: clustered by (azimuth, eccentricity) into blocks and tiles, with the
//     sample -> polar texel map (foveated.cpp:19-84);
//   * eye-space tile and block frusta fitted to the beams (frusta.cu:26-121);
//   * the per-frame eye transforms, screen cull planes and the previous-frame
//     reprojection (render.cpp:397-416, 454-488).

#include "cluster.h"
#include "host_internal.h"

#include <algorithm>
#include <limits>
#include <thread>

namespace hvvr {

namespace {

// I + K sin + K^2 (1 - cos), K the cross-product matrix of the axis (a rotation by
// `radians` about a unit axis; vector_math.h:1083-1088)
matrix3x3 axisAngleMatrix(const vector3& a, float radians) {
    float c = cosf(radians), s = sinf(radians);
    matrix3x3 k(vector3(0.0f, a.z, -a.y), vector3(-a.z, 0.0f, a.x), vector3(a.y, -a.x, 0.0f));
    matrix3x3 k2 = k * k;
    float t = 1.0f - c;
    matrix3x3 r;
    r.m0 = vector3(1.0f, 0.0f, 0.0f) + k.m0 * s + k2.m0 * t;
    r.m1 = vector3(0.0f, 1.0f, 0.0f) + k.m1 * s + k2.m1 * t;
    r.m2 = vector3(0.0f, 0.0f, 1.0f) + k.m2 * s + k2.m2 * t;
    return r;
}

float at(const matrix4x4& m, int row, int col) {
    const vector4& c = col == 0 ? m.m0 : col == 1 ? m.m1 : col == 2 ? m.m2 : m.m3;
    return row == 0 ? c.x : row == 1 ? c.y : row == 2 ? c.z : c.w;
}

// general inverse by cofactors (in double, rounded once)
matrix4x4 invert4(const matrix4x4& m) {
    double a[4][4], inv[4][4];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            a[r][c] = at(m, r, c);
    auto minor3 = [&](int skipR, int skipC) {
        double b[3][3];
        for (int r = 0, rr = 0; r < 4; r++) {
            if (r == skipR)
                continue;
            for (int c = 0, cc = 0; c < 4; c++) {
                if (c == skipC)
                    continue;
                b[rr][cc++] = a[r][c];
            }
            rr++;
        }
        return b[0][0] * (b[1][1] * b[2][2] - b[1][2] * b[2][1]) - b[0][1] * (b[1][0] * b[2][2] - b[1][2] * b[2][0]) +
               b[0][2] * (b[1][0] * b[2][1] - b[1][1] * b[2][0]);
    };
    double det = 0.0;
    for (int c = 0; c < 4; c++)
        det += ((c & 1) ? -1.0 : 1.0) * a[0][c] * minor3(0, c);
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++)
            inv[c][r] = (((r + c) & 1) ? -1.0 : 1.0) * minor3(r, c) / det;
    matrix4x4 out;
    vector4* cols[4] = {&out.m0, &out.m1, &out.m2, &out.m3};
    for (int c = 0; c < 4; c++)
        *cols[c] = vector4(float(inv[0][c]), float(inv[1][c]), float(inv[2][c]), float(inv[3][c]));
    return out;
}

vector3 xyz(const vector4& v) {
    return vector3(v.x, v.y, v.z);
}

// Plane transform (graphics_types.h:195-200): a point of the plane goes through M, the
// normal through the inverse transpose
vector4 transformPlane(const matrix4x4& m, const vector3& normal, float dist) {
    vector3 o = xyz(m * vector4(normal * dist, 1.0f));
    matrix4x4 inv = invert4(m);
    // (inverse(m))^T * (normal, 0): component j is column j of the inverse dotted with it
    vector3 n(dot(xyz(inv.m0), normal), dot(xyz(inv.m1), normal), dot(xyz(inv.m2), normal));
    return vector4(n, dot(o, n));
}

SimpleRayFrustum degenerateFrustum() {
    SimpleRayFrustum f;
    for (int i = 0; i < 4; i++) {
        f.origins[i] = vector3(INFINITY, INFINITY, INFINITY);
        f.directions[i] = vector3(0.0f, 0.0f, 0.0f);
    }
    return f;
}

// One frustum around beams [0, n): over `orientations` rotations about the mean
// direction, the one with the smallest uv bounding box of the beam footprints
// (each beam's centre +- its larger differential, times `slop`) on the z = 1 plane.
SimpleRayFrustum fitFrustum(const DirectionalBeam* s, size_t n, float slop, int orientations) {
    if (n == 0)
        return degenerateFrustum();
    vector3 dominant(0.0f);
    for (size_t i = 0; i < n; i++)
        dominant = dominant + s[i].centerRay;
    dominant = normalize(dominant);
    matrix3x3 rot = matrix3x3::rotationFromZAxis(dominant);
    float bestArea = INFINITY;
    matrix3x3 bestRot = rot;
    vector2 bestMin(INFINITY), bestMax(-INFINITY);
    const float range = (kPi / 2.0f) * 0.8f;
    for (int o = 0; o < orientations; o++) {
        matrix3x3 cur = axisAngleMatrix(vector3(0.0f, 0.0f, 1.0f), (range * o / float(orientations)) - (range / 2.0f)) * rot;
        matrix3x3 inv = invert(cur);
        vector2 lo(INFINITY), hi(-INFINITY);
        for (size_t i = 0; i < n; i++) {
            vector3 v = inv * s[i].centerRay;
            vector2 uv(v.x / v.z, v.y / v.z);
            v = inv * (s[i].du + s[i].centerRay);
            float r = length(uv - vector2(v.x / v.z, v.y / v.z));
            v = inv * (s[i].dv + s[i].centerRay);
            r = std::max(r, length(uv - vector2(v.x / v.z, v.y / v.z)));
            r *= slop;
            lo = vector2(std::min(lo.x, uv.x - r), std::min(lo.y, uv.y - r));
            hi = vector2(std::max(hi.x, uv.x + r), std::max(hi.y, uv.y + r));
        }
        float area = (hi.x - lo.x) * (hi.y - lo.y);
        if (o == 0 || area < bestArea) {
            bestArea = area;
            bestRot = cur;
            bestMin = lo;
            bestMax = hi;
        }
    }
    auto toDir = [&](float u, float v) { return bestRot * normalize(vector3(u, v, 1.0f)); };
    SimpleRayFrustum f;
    for (int i = 0; i < 4; i++)
        f.origins[i] = vector3(0.0f, 0.0f, 0.0f);
    f.directions[0] = toDir(bestMin.x, bestMax.y);
    f.directions[1] = toDir(bestMax.x, bestMax.y);
    f.directions[2] = toDir(bestMax.x, bestMin.y);
    f.directions[3] = toDir(bestMin.x, bestMin.y);
    return f;
}

template <class Fn>
void parallelFor(size_t n, Fn fn) {
    unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64 || threads == 1) {
        for (size_t i = 0; i < n; i++)
            fn(i);
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; t++)
        pool.emplace_back([&, t] {
            for (size_t i = t; i < n; i += threads)
                fn(i);
        });
    for (auto& th : pool)
        th.join();
}

} // namespace

bool polarFoveatedSamples(const FoveatedSamplePattern& p, EccentricityMap& emap, uint32_t& samplesPerRing,
                          uint32_t& ringCount, std::vector<DirectionalBeam>& out, std::string* error) {
    emap = makeEccentricityMap(p.marSlope, p.fovealMARDegrees, p.degreeTrackingError);
    if (!(p.marSlope > 0.0f) || !(p.fovealMARDegrees > 0.0f) || !(p.degreeTrackingError >= 0.0f) ||
        !(p.maxFOVDegrees > 0.0f) || !(p.maxFOVDegrees < 180.0f)) {
        if (error)
            *error = "foveated pattern needs marSlope > 0, fovealMAR > 0, tracking error >= 0, 0 < maxFOV < 180";
        return false;
    }
    // ring and per-ring counts first (samples.cpp:59-77)
    const float maxE = p.maxFOVDegrees * kRadiansPerDegree;
    const size_t sampleLimit = size_t(1) << 26;
    size_t spr = 0, rings = 0;
    float E = 0.0f;
    while (E <= maxE) {
        float w = emap.apply(float(rings) + 1.0f) - E;
        float around = 2.0f * kPi * sinf(E);
        size_t onRing = (size_t)std::ceil(around / w);
        spr = std::max(spr, onRing);
        ++rings;
        E = emap.apply(float(rings));
        if (rings * std::max<size_t>(spr, 1) > sampleLimit) {
            if (error)
                *error = "foveated pattern needs more than 2^26 samples";
            return false;
        }
    }
    if (spr == 0)
        spr = 1;
    samplesPerRing = uint32_t(spr);
    ringCount = uint32_t(rings);
    out.assign(spr * rings, DirectionalBeam());
    // concentric rings; du = offset to the next ring on the tangent plane, dv = the
    // azimuthal spacing along the ring, perpendicular to it
    E = 0.0f;
    for (size_t r = 0; r < rings; r++) {
        float eNext = emap.apply(float(r) + 1.0f);
        float around = 2.0f * kPi * sinf(E);
        vector3 base = normalize(quaternion::fromAxisAngle(vector3(0, 1, 0), E) * vector3(0, 0, -1));
        vector3 zenithBase = normalize(quaternion::fromAxisAngle(vector3(0, 1, 0), eNext) * vector3(0, 0, -1));
        for (size_t i = 0; i < spr; i++) {
            float rotation = (float(i) + 0.5f) / float(spr) * 2.0f * kPi;
            quaternion q = quaternion::fromAxisAngle(vector3(0, 0, -1), rotation);
            vector3 dir = normalize(q * base);
            vector3 zenithDir = normalize(q * zenithBase);
            vector3 onTangent = zenithDir * (1.0f / dot(zenithDir, dir));
            vector3 zenithDiff = onTangent - dir;
            vector3 azimuthUnit = cross(dir, normalize(zenithDiff));
            DirectionalBeam& b = out[spr * r + i];
            b.centerRay = dir;
            b.du = zenithDiff;
            b.dv = azimuthUnit * around / float(spr);
        }
        E = eNext;
    }
    return true;
}

bool buildFoveatedSetup(const FoveatedSamplePattern& p, FoveatedSetup& out, std::string* error) {
    out = FoveatedSetup();
    out.pattern = p;
    std::vector<DirectionalBeam> polar;
    if (!polarFoveatedSamples(p, out.emap, out.samplesPerRing, out.ringCount, polar, error))
        return false;
    out.maxEccentricity = p.maxFOVDegrees * kRadiansPerDegree;
    const uint32_t count = uint32_t(polar.size());
    out.validSampleCount = count;
    out.blockCount = (count + BLOCK_SIZE - 1) / BLOCK_SIZE;
    out.sampleCount = out.blockCount * BLOCK_SIZE;

    // cluster by (azimuth, eccentricity), eccentricity as y (foveated.cpp:19-46). The
    // gaze-direction ring has no azimuth; the reference's normalize(0) makes it NaN
    // there, which leaves its sort order undefined, so the azimuth is taken from the
    // unnormalised (x, y): the same angle everywhere else, 0 or +-pi on that ring.
    struct Point {
        vector2 position;
        uint32_t index;
    };
    std::vector<Point> pts(count);
    for (uint32_t i = 0; i < count; i++) {
        const vector3& v = polar[i].centerRay;
        pts[i].position = vector2(atan2f(v.y, v.x), acosf(-v.z));
        pts[i].index = i;
    }
    naiveXYCluster(pts, out.blockCount, [](const Point& x) { return x.position; });

    out.beams.assign(out.sampleCount, DirectionalBeam());
    out.polarRemap.assign(count, vector2ui());
    for (uint32_t k = 0; k < count; k++) {
        uint32_t i = pts[k].index;
        out.beams[k] = polar[i];
        out.polarRemap[k] = vector2ui{i % out.samplesPerRing, i / out.samplesPerRing};
    }
    // padding repeats the last sample (its results are never remapped)
    for (uint32_t k = count; k < out.sampleCount; k++)
        out.beams[k] = out.beams[count - 1];

    // frusta over the valid beams of each tile / block; a tile with none is degenerate
    const uint32_t tileCount = out.blockCount * TILES_PER_BLOCK;
    out.tileFrusta.assign(tileCount, SimpleRayFrustum());
    out.blockFrusta.assign(out.blockCount, SimpleRayFrustum());
    parallelFor(size_t(tileCount) + out.blockCount, [&](size_t j) {
        if (j < tileCount) {
            size_t begin = j * TILE_SIZE, end = std::min<size_t>(begin + TILE_SIZE, count);
            out.tileFrusta[j] = fitFrustum(out.beams.data() + begin, end > begin ? end - begin : 0, 2.0f, 63);
        } else {
            size_t b = j - tileCount;
            size_t begin = b * BLOCK_SIZE, end = std::min<size_t>(begin + BLOCK_SIZE, count);
            out.blockFrusta[b] = fitFrustum(out.beams.data() + begin, end > begin ? end - begin : 0, 2.0f, 63);
        }
    });
    return true;
}

void foveatedFrame(const matrix4x4& cameraToWorld, const vector3& eyeDir, const matrix3x3& sampleToCam,
                   vector2 boundsLower, vector2 boundsUpper, FoveatedHistory& history, FoveatedFrame& out) {
    matrix3x3 eyeToCamera = matrix3x3::rotationFromZAxis(-eyeDir);
    out.eyeToWorld = cameraToWorld * matrix4x4(eyeToCamera, vector3(0.0f));
    matrix3x3 cameraToEye = invert(eyeToCamera);
    // cullRectToCullPlanes: the sample-bounds rectangle as 4 eye-space planes through the
    // eye, pushed out by 1 cm, in world space
    vector2 corners[4] = {vector2(boundsUpper.x, boundsUpper.y), vector2(boundsLower.x, boundsUpper.y),
                          vector2(boundsLower.x, boundsLower.y), vector2(boundsUpper.x, boundsLower.y)};
    for (int i = 0; i < 4; i++) {
        vector3 d0 = sampleToCam * vector3(corners[i].x, corners[i].y, 1.0f);
        vector3 d1 = sampleToCam * vector3(corners[(i + 1) % 4].x, corners[(i + 1) % 4].y, 1.0f);
        vector3 n = cameraToEye * normalize(cross(d1, d0));
        out.cullPlanes[i] = transformPlane(out.eyeToWorld, n, -0.01f);
    }
    out.sampleToEye = cameraToEye * sampleToCam;
    matrix4x4 eyeToEyePrevious = history.worldToEyePrevious * out.eyeToWorld;
    out.eyeToSamplePrevious = matrix4x4(history.eyePreviousToSamplePrevious, vector3(0.0f)) * eyeToEyePrevious;
    history.worldToEyePrevious = invert4(out.eyeToWorld);
    history.eyePreviousToSamplePrevious = invert(out.sampleToEye);
}

} // namespace hvvr
