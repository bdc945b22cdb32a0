/* Scalar vector helpers for the oracle (test infrastructure only). */
#ifndef OR_MATH_H
#define OR_MATH_H
#include "hvvr_oracle.h"
#include <float.h>
#include <math.h>
#include <string.h>

static inline ov3 v3(float x, float y, float z) { ov3 r = {x, y, z}; return r; }
static inline ov3 add3(ov3 a, ov3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline ov3 sub3(ov3 a, ov3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline ov3 mul3(ov3 a, ov3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline ov3 scale3(ov3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline ov3 neg3(ov3 a) { return v3(-a.x, -a.y, -a.z); }
static inline float dot3(ov3 a, ov3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline ov3 cross3(ov3 a, ov3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline ov3 norm3(ov3 a) {
    float l = sqrtf(dot3(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}
static inline float comp3(ov3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* column-major 4x4 times (v, w), left-to-right sums (vector_math.h:1273-1278) */
static inline void m4mul(const float m[16], float x, float y, float z, float w, float out[4]) {
    for (int r = 0; r < 4; r++)
        out[r] = m[0 * 4 + r] * x + m[1 * 4 + r] * y + m[2 * 4 + r] * z + m[3 * 4 + r] * w;
}
/* upper 3x3 times v (vector_math.h:1129-1132) */
static inline ov3 m3mul(const float m[16], ov3 v) {
    return v3(m[0] * v.x + m[4] * v.y + m[8] * v.z, m[1] * v.x + m[5] * v.y + m[9] * v.z,
              m[2] * v.x + m[6] * v.y + m[10] * v.z);
}

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* _mm256_min_ps / _mm256_max_ps lane semantics: second operand unless the first wins */
static inline float minps(float a, float b) { return a < b ? a : b; }
static inline float maxps(float a, float b) { return a > b ? a : b; }

/* radix-sortable float order (sort.h:14-32): total order with -0 < +0 */
static inline uint32_t flipf(float f) {
    uint32_t u = fbits(f);
    uint32_t mask = (uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u;
    return u ^ mask;
}
static inline float maxTotal(float a, float b) { return flipf(a) >= flipf(b) ? a : b; }

/* IEEE binary16 round-to-nearest-even */
static inline uint16_t f2h(float f) {
    uint32_t x = fbits(f);
    uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) {
        if (ax < 0x33000000u) return (uint16_t)sign;
        uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
        int shift = 126 - (int)(ax >> 23);
        uint32_t h = mant >> shift, rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = (ax >> 13) - (112u << 10), rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}
static inline float h2f(uint16_t h) {
    /* exact conversion via double arithmetic */
    int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    double v;
    if (e == 0) v = ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = ldexp((double)(m | 0x400), e - 25);
    return (float)((h & 0x8000) ? -v : v);
}

#endif
