// ORACLE -- test infrastructure only.  Nothing in the product may include, link or call
// this file; it exists so tests/, the smoke check and the CPU baseline of bench.py can
// check the HIP kernels against a plain host restatement of the reference's shader logic.
//
// Float vector arithmetic restated from glm with the exact evaluation order glm uses
// (left-to-right component sums, normalize = v * (1 / sqrt(dot(v, v))), ...), so that a
// build with -ffp-contract=off gives one IEEE result per expression.
#pragma once

#include <cmath>
#include <cstdint>

namespace orc {

struct V2 { float x, y; };

struct V3 {
    float x, y, z;
    V3() : x(0.0f), y(0.0f), z(0.0f) {}
    V3(float s) : x(s), y(s), z(s) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(const V3& a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(const V3& a, const V3& b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline V3 operator*(const V3& a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, const V3& a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(const V3& a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3 operator/(const V3& a, const V3& b) { return V3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline V3& operator+=(V3& a, const V3& b) { a = a + b; return a; }
inline V3& operator*=(V3& a, const V3& b) { a = a * b; return a; }
inline V3& operator*=(V3& a, float s) { a = a * s; return a; }
inline V3& operator/=(V3& a, float s) { a = a / s; return a; }
inline bool operator==(const V3& a, const V3& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// glm::dot for vec3: (a.x*b.x + a.y*b.y) + a.z*b.z
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(const V3& v) { return std::sqrt(dot(v, v)); }
// glm::distance(p0, p1) = length(p1 - p0)
inline float distance(const V3& p0, const V3& p1) { return length(p1 - p0); }
// glm::normalize(v) = v * inversesqrt(dot(v, v)), inversesqrt(x) = 1 / sqrt(x)
inline V3 normalize(const V3& v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline V3 cross(const V3& a, const V3& b) {
    return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
// glm::reflect(I, N) = I - N * dot(N, I) * 2
inline V3 reflect(const V3& I, const V3& N) { return I - N * dot(N, I) * 2.0f; }
// glm::max(x, y) = (x < y) ? y : x ; glm::min(x, y) = (y < x) ? y : x
inline float gmax(float x, float y) { return (x < y) ? y : x; }
inline float gmin(float x, float y) { return (y < x) ? y : x; }
inline V3 gmax(const V3& a, float s) { return V3(gmax(a.x, s), gmax(a.y, s), gmax(a.z, s)); }
inline V3 gmax(float s, const V3& a) { return V3(gmax(s, a.x), gmax(s, a.y), gmax(s, a.z)); }
inline float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
// glm::mix(x, y, a) = x * (1 - a) + y * a
inline V3 mix(const V3& x, const V3& y, float a) { return x * (1.0f - a) + y * a; }
inline bool any_nan_inf(const V3& v) {
    return std::isnan(v.x) || std::isnan(v.y) || std::isnan(v.z) ||
           std::isinf(v.x) || std::isinf(v.y) || std::isinf(v.z);
}
inline bool bad(float v) { return std::isnan(v) || std::isinf(v); }

// GetLuminance (include/graph/luminance.h:6-9)
inline float luminance(const V3& c) { return dot(c, V3(0.2126f, 0.7152f, 0.0722f)); }

// 3x3 matrix stored as three columns, glm::mat3 * vec3 evaluation order:
// row r = (m[0][r]*v.x + m[1][r]*v.y) + m[2][r]*v.z
struct M3 { V3 c0, c1, c2; };
inline V3 mul(const M3& m, const V3& v) {
    return V3(m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z,
              m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
              m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z);
}

// Sine and cosine for the angles the BSDF samples, phi in [0, 2*pi).  Same method on
// the CPU and the GPU: octant reduction with the three-part pi/4 of Cody and Waite,
// then the single-precision minimax polynomials of the Cephes library (S. L. Moshier,
// "Methods and Programs for Mathematical Functions", sinf.c).  No FMA, so every step
// is one IEEE operation.
inline void sincos_cw(float x, float* s, float* c) {
    float sign_s = 1.0f;
    if (x < 0.0f) { x = -x; sign_s = -1.0f; }
    int j = (int)(x * 1.27323954473516f);
    float y = (float)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    const float z = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    const float zz = z * z;
    const float sp = ((-1.9515295891e-4f * zz + 8.3321608736e-3f) * zz - 1.6666654611e-1f) * zz * z + z;
    const float cp = ((2.443315711809948e-5f * zz - 1.388731625493765e-3f) * zz + 4.166664568298827e-2f) * zz * zz
                     - 0.5f * zz + 1.0f;
    float rs, rc;
    switch (j) {
        case 0: rs = sp; rc = cp; break;
        case 2: rs = cp; rc = -sp; break;
        case 4: rs = -sp; rc = -cp; break;
        default: rs = -cp; rc = sp; break;
    }
    *s = sign_s * rs;
    *c = rc;
}

// glm::pow(max(0, 1 - VdotH), 5) of FresnelSchlick (shader/brdf.cu:29-37) as products.
inline float pow5(float x) { const float x2 = x * x; return x2 * x2 * x; }

}  // namespace orc
