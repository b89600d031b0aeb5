// Oracle — TEST INFRASTRUCTURE ONLY. Never linked into or called by the product path.
//
// Scalar / 3x3 arithmetic of the reference's WGSL prelude, restated with the exact
// operation order of the WGSL source text (compile with -ffp-contract=off, no fast-math):
//   matrixOps.wgsl:8-27   mat3x3Inverse        matrixOps.wgsl:33-45  polar rotation
//   matrixOps.wgsl:47-53  outerProduct         mpmOps.wgsl:14-53     cell / B-spline helpers
//   randomNumberGenerators.wgsl:3-38 hash1, hash3
// The WGSL built-ins that the spec leaves implementation-defined are given one fixed
// definition here (and the same one in the HIP kernels):
//   determinant(M)  = glm-style cofactor expansion along the first row of the column-major
//                     matrix: m00(m11 m22 - m21 m12) - m10(m01 m22 - m21 m02) + m20(m01 m12 - m11 m02)
//   exp, log, pow   = orc_expf / orc_logf / orc_powf below (Cody–Waite reduction + Taylor /
//                     atanh series), accurate to a few ulp, bit-reproducible on CPU and GPU.
//   f32 -> i32      = truncation toward zero, saturating, NaN -> 0 (WGSL conversion rules).
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>
#include <climits>

namespace orc {

// ---------------------------------------------------------------- integer hashes
// randomNumberGenerators.wgsl:3-9
static inline uint32_t hash1(uint32_t n) {
    uint32_t h32 = n + 374761393u;
    h32 = 668265263u * ((h32 << 17) | (h32 >> (32 - 17)));
    h32 = 2246822519u * (h32 ^ (h32 >> 15));
    h32 = 3266489917u * (h32 ^ (h32 >> 13));
    return h32 ^ (h32 >> 16);
}
// randomNumberGenerators.wgsl:24-38
static inline uint32_t hash3(uint32_t px, uint32_t py, uint32_t pz) {
    const uint32_t p2 = 2246822519u, p3 = 3266489917u, p4 = 668265263u, p5 = 374761393u;
    uint32_t h32 = pz + p5 + px * p3;
    h32 = p4 * ((h32 << 17) | (h32 >> (32 - 17)));
    h32 += py * p3;
    h32 = p4 * ((h32 << 17) | (h32 >> (32 - 17)));
    h32 = p2 * (h32 ^ (h32 >> 15));
    h32 = p3 * (h32 ^ (h32 >> 13));
    return h32 ^ (h32 >> 16);
}
// randomNumberGenerators.wgsl:40-56
static inline uint32_t hash4(uint32_t px, uint32_t py, uint32_t pz, uint32_t pw) {
    const uint32_t p2 = 2246822519u, p3 = 3266489917u, p4 = 668265263u, p5 = 374761393u;
    uint32_t h32 = pw + p5 + px * p3;
    h32 = p4 * ((h32 << 17) | (h32 >> (32 - 17)));
    h32 += py * p3;
    h32 = p4 * ((h32 << 17) | (h32 >> (32 - 17)));
    h32 += pz * p3;
    h32 = p4 * ((h32 << 17) | (h32 >> (32 - 17)));
    h32 = p2 * (h32 ^ (h32 >> 15));
    h32 = p3 * (h32 ^ (h32 >> 13));
    return h32 ^ (h32 >> 16);
}

// WGSL i32 addition wraps (two's complement); C++ signed overflow is undefined, so stencil cell
// offsets (c + o, where c can be the saturated i32 of an infinite position) add through u32.
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// ---------------------------------------------------------------- conversions
static inline float bits_f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t f_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// WGSL i32(f32): truncate toward zero, saturate, NaN -> 0.
static inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static inline int32_t f2i(double x) {   // FP64 mode of the cell-number conversion
    if (x != x) return 0;
    if (x >= 2147483648.0) return INT32_MAX;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

// ---------------------------------------------------------------- exp / log / pow (FP32)
// exp: Cody–Waite reduction x = n ln2 + r, |r| <= ln2/2, e^r by its degree-7 Taylor
// polynomial in Horner form, result scaled by 2^n through the exponent bits.
static inline float ldexp_n(float p, int n) {
    if (n > 127) { p = p * bits_f(0x7f000000u); n -= 127; if (n > 127) n = 127; }
    if (n < -126) { p = p * bits_f(0x0c800000u); n += 102; if (n < -126) { p = p * bits_f(0x0c800000u); n += 102; } if (n < -126) n = -126; }
    return p * bits_f((uint32_t)(n + 127) << 23);
}
static inline float orc_expf(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return INFINITY;
    if (x < -103.97208404541015625f) return 0.0f;
    const float LOG2E = 1.44269502162933349609375f;
    const float RND = 12582912.0f;                    // 1.5 * 2^23: round-to-nearest-even trick
    float t = x * LOG2E;
    float n = (t + RND) - RND;
    const float LN2_HI = 0.693359375f;                // 355/512, exact products n*LN2_HI
    const float LN2_LO = -2.12194440e-4f;
    float r = (x - n * LN2_HI) - n * LN2_LO;
    float p = 1.98412698e-4f;                         // 1/7!
    p = p * r + 1.38888889e-3f;                       // 1/6!
    p = p * r + 8.33333377e-3f;                       // 1/5!
    p = p * r + 4.16666679e-2f;                       // 1/4!
    p = p * r + 1.66666672e-1f;                       // 1/3!
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return ldexp_n(p, (int)n);
}
// log: x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1),
// atanh series through s^9.
static inline float orc_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    uint32_t b = f_bits(x);
    if ((b >> 23) == 0) { x = x * 8388608.0f; b = f_bits(x); e = -23; }  // subnormal
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = bits_f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float z = s * s;
    float q = 0.111111112f;                           // 1/9
    q = q * z + 0.142857149f;                         // 1/7
    q = q * z + 0.200000003f;                         // 1/5
    q = q * z + 0.333333343f;                         // 1/3
    q = q * z + 1.0f;
    float lm = (2.0f * s) * q;
    float fe = (float)e;
    return fe * 0.693359375f + (lm + fe * -2.12194440e-4f);
}
static inline float orc_powf(float x, float y) {
    if (x != x || y != y) return NAN;
    if (x == 1.0f) return 1.0f;
    if (x == 0.0f) return (y < 0.0f) ? INFINITY : ((y == 0.0f) ? 1.0f : 0.0f);
    if (x < 0.0f) return NAN;                         // only |J| reaches pow on the path
    return orc_expf(y * orc_logf(x));
}

// R-generic front ends: FP32 uses the functions above, FP64 uses libm.
static inline float  r_exp(float x)  { return orc_expf(x); }
static inline double r_exp(double x) { return std::exp(x); }
static inline float  r_pow(float x, float y)  { return orc_powf(x, y); }
static inline double r_pow(double x, double y) { return std::pow(x, y); }
static inline float  r_sqrt(float x)  { return std::sqrt(x); }
static inline double r_sqrt(double x) { return std::sqrt(x); }

// ---------------------------------------------------------------- vec3 / mat3 (column-major)
template <class R> struct V3 { R x, y, z; };
template <class R> struct M3 { V3<R> c[3]; };   // c[i] = column i; M[i][j] = column i, row j

template <class R> static inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> static inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> static inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> static inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> static inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> static inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> static inline V3<R> cmul(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> static inline V3<R> cdiv(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> static inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> static inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> static inline R length(V3<R> a) { return r_sqrt(dot(a, a)); }
template <class R> static inline V3<R> normalize(V3<R> a) { R l = length(a); return a / l; }
template <class R> static inline V3<R> vmin(V3<R> a, V3<R> b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
template <class R> static inline V3<R> vmax(V3<R> a, V3<R> b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }

template <class R> static inline M3<R> ident() { M3<R> m; m.c[0] = {1, 0, 0}; m.c[1] = {0, 1, 0}; m.c[2] = {0, 0, 1}; return m; }
template <class R> static inline M3<R> zero3() { M3<R> m; m.c[0] = {0, 0, 0}; m.c[1] = {0, 0, 0}; m.c[2] = {0, 0, 0}; return m; }
template <class R> static inline M3<R> madd(const M3<R>& a, const M3<R>& b) { M3<R> m; for (int i = 0; i < 3; ++i) m.c[i] = a.c[i] + b.c[i]; return m; }
template <class R> static inline M3<R> msub(const M3<R>& a, const M3<R>& b) { M3<R> m; for (int i = 0; i < 3; ++i) m.c[i] = a.c[i] - b.c[i]; return m; }
template <class R> static inline M3<R> mscale(R s, const M3<R>& a) { M3<R> m; for (int i = 0; i < 3; ++i) m.c[i] = s * a.c[i]; return m; }
template <class R> static inline M3<R> mscale_r(const M3<R>& a, R s) { M3<R> m; for (int i = 0; i < 3; ++i) m.c[i] = a.c[i] * s; return m; }
template <class R> static inline M3<R> transpose(const M3<R>& a) {
    M3<R> m;
    m.c[0] = {a.c[0].x, a.c[1].x, a.c[2].x};
    m.c[1] = {a.c[0].y, a.c[1].y, a.c[2].y};
    m.c[2] = {a.c[0].z, a.c[1].z, a.c[2].z};
    return m;
}
// M * v = (c0 v.x + c1 v.y) + c2 v.z
template <class R> static inline V3<R> mvmul(const M3<R>& a, V3<R> v) {
    return (a.c[0] * v.x + a.c[1] * v.y) + a.c[2] * v.z;
}
template <class R> static inline M3<R> mmul(const M3<R>& a, const M3<R>& b) {
    M3<R> m; for (int j = 0; j < 3; ++j) m.c[j] = mvmul(a, b.c[j]); return m;
}
template <class R> static inline R det3(const M3<R>& m) {
    const R m00 = m.c[0].x, m01 = m.c[0].y, m02 = m.c[0].z;
    const R m10 = m.c[1].x, m11 = m.c[1].y, m12 = m.c[1].z;
    const R m20 = m.c[2].x, m21 = m.c[2].y, m22 = m.c[2].z;
    return (m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02)) + m20 * (m01 * m12 - m11 * m02);
}
// matrixOps.wgsl:8-27
template <class R> static inline M3<R> inverse3(const M3<R>& m) {
    R det = det3(m);
    if (det == (R)0) return ident<R>();
    R inv = (R)1 / det;
    M3<R> r;
    r.c[0] = {m.c[1].y * m.c[2].z - m.c[1].z * m.c[2].y,
              m.c[0].z * m.c[2].y - m.c[0].y * m.c[2].z,
              m.c[0].y * m.c[1].z - m.c[0].z * m.c[1].y};
    r.c[1] = {m.c[1].z * m.c[2].x - m.c[1].x * m.c[2].z,
              m.c[0].x * m.c[2].z - m.c[0].z * m.c[2].x,
              m.c[0].z * m.c[1].x - m.c[0].x * m.c[1].z};
    r.c[2] = {m.c[1].x * m.c[2].y - m.c[1].y * m.c[2].x,
              m.c[0].y * m.c[2].x - m.c[0].x * m.c[2].y,
              m.c[0].x * m.c[1].y - m.c[0].y * m.c[1].x};
    return mscale_r(r, inv);
}
// matrixOps.wgsl:33-45 — four Newton steps R <- (R + R^-T)/2
template <class R> static inline M3<R> polar_rotation(const M3<R>& f) {
    M3<R> r = f;
    for (int i = 0; i < 4; ++i) {
        M3<R> rit = transpose(inverse3(r));
        r = mscale((R)0.5, madd(r, rit));
    }
    return r;
}
// matrixOps.wgsl:47-53
template <class R> static inline M3<R> outer(V3<R> u, V3<R> v) {
    M3<R> m; m.c[0] = u * v.x; m.c[1] = u * v.y; m.c[2] = u * v.z; return m;
}

}  // namespace orc
