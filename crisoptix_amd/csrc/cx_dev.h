// Device arithmetic shared by every kernel: vectors with glm's evaluation order, the PCG32
// and TEA generators, and the project-written sine/cosine.
//
// Every expression is written so that, compiled with -ffp-contract=off (no FMA formation)
// and IEEE division/sqrt, each operation rounds once.  That makes the per-pixel results a
// deterministic function of the inputs, identical on every build that honours IEEE-754.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CX_HD __host__ __device__ __forceinline__

namespace cx {

struct V3 {
    float x, y, z;
};

CX_HD V3 v3(float s) { return V3{s, s, s}; }
CX_HD V3 v3(float a, float b, float c) { return V3{a, b, c}; }
CX_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CX_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CX_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
CX_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
CX_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
CX_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
CX_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
CX_HD float comp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// glm::dot: (a.x*b.x + a.y*b.y) + a.z*b.z
CX_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CX_HD float length(V3 v) { return sqrtf(dot(v, v)); }
// glm::distance(p0, p1) = length(p1 - p0)
CX_HD float distance(V3 p0, V3 p1) { return length(p1 - p0); }
// glm::normalize: v * (1 / sqrt(dot(v, v)))
CX_HD V3 normalize(V3 v) { return v * (1.0f / sqrtf(dot(v, v))); }
// glm::reflect(I, N) = I - N * dot(N, I) * 2
CX_HD V3 reflect(V3 i, V3 n) { return i - n * dot(n, i) * 2.0f; }
// glm::max(x, y) = (x < y) ? y : x ; glm::min(x, y) = (y < x) ? y : x
CX_HD float gmax(float x, float y) { return (x < y) ? y : x; }
CX_HD float gmin(float x, float y) { return (y < x) ? y : x; }
CX_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
CX_HD bool bad(float v) { return isnan(v) || isinf(v); }
CX_HD bool any_bad(V3 v) { return bad(v.x) || bad(v.y) || bad(v.z); }
CX_HD bool is_zero(V3 v) { return v.x == 0.0f && v.y == 0.0f && v.z == 0.0f; }
// Rec.709 luminance (include/graph/luminance.h:6-9)
CX_HD float luminance(V3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// Sine/cosine for BSDF sample angles: octant reduction by a three-part pi/4 (Cody-Waite) and
// the single-precision minimax polynomials of Cephes sinf.c (S. L. Moshier).
CX_HD void sincos_cw(float x, float* s, float* c) {
    float sign_s = 1.0f;
    if (x < 0.0f) {
        x = -x;
        sign_s = -1.0f;
    }
    int j = (int)(x * 1.27323954473516f);
    float y = (float)j;
    if (j & 1) {
        j += 1;
        y += 1.0f;
    }
    j &= 7;
    const float z = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    const float zz = z * z;
    const float sp = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
    const float cp =
        ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz - 0.5f * zz +
        1.0f;
    float rs, rc;
    if (j == 0) {
        rs = sp;
        rc = cp;
    } else if (j == 2) {
        rs = cp;
        rc = -sp;
    } else if (j == 4) {
        rs = -sp;
        rc = -cp;
    } else {
        rs = -cp;
        rc = sp;
    }
    *s = sign_s * rs;
    *c = rc;
}

CX_HD float pow5(float x) {
    const float x2 = x * x;
    return x2 * x2 * x;
}

// ---- random numbers (include/util/random.h) ----

// SampleTEA64 (random.h:24-38): 4 rounds of TEA keyed by the reference's constants.
CX_HD uint64_t tea64(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
#pragma unroll
    for (int n = 0; n < 4; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return (uint64_t)v0 | ((uint64_t)v1 << 32);
}

// Every PCG32 of the path uses the default stream (random.h:47), so the increment is a
// constant and only the 64-bit state is stored.
static constexpr uint64_t PCG_MULT = 0x5851f42d4c957f2dULL;
static constexpr uint64_t PCG_INC = (0xda3e39cb94b95bdbULL << 1) | 1ULL;
static constexpr uint64_t PCG_DEFAULT_STATE = 0x853c49e6748fea9bULL;

struct Rng {
    uint64_t s;
    CX_HD uint32_t u32() {
        const uint64_t old = s;
        s = old * PCG_MULT + PCG_INC;
        const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        const uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    // low word from the first draw
    CX_HD uint64_t u64() {
        const uint64_t lo = u32();
        const uint64_t hi = u32();
        return lo | (hi << 32);
    }
    CX_HD float f() {
        const uint32_t bits = (u32() >> 9) | 0x3f800000u;
        return __builtin_bit_cast(float, bits) - 1.0f;
    }
};

// pcg32_srandom with the default stream (random.h:52-59)
CX_HD Rng rng_seed(uint64_t initstate) {
    Rng r;
    r.s = 0;
    r.u32();
    r.s += initstate;
    r.u32();
    return r;
}

}  // namespace cx
