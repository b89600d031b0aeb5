// Device arithmetic for the MPM step on gfx950.
//
// Every expression keeps the operation order of the reference WGSL prelude, and the library
// is compiled with -ffp-contract=off, so each product/sum is one IEEE FP32 rounding exactly as
// the source text reads (HIP's default `/` and sqrtf are correctly rounded). That is what
// makes the fixed-point grid sums, and every particle field downstream of them, reproducible
// bit for bit on the host.
//   shaderPrelude/matrixOps.wgsl:8-53     inverse (adjugate / det), 4-step Newton polar, outer
//   shaderPrelude/mpmOps.wgsl:14-53       cell number, fractional position, B-spline weights
//   shaderPrelude/randomNumberGenerators.wgsl:3-56  hash1 / hash3 / hash4
// WGSL built-ins with implementation-defined results get one fixed definition:
//   determinant   m00(m11 m22 - m21 m12) - m10(m01 m22 - m21 m02) + m20(m01 m12 - m11 m02)
//   exp/log/pow   Cody–Waite reduction + degree-7 Taylor (exp), atanh series (log),
//                 pow(x,y) = exp(y log x); a few ulp, identical on host and device
//   i32(f32)      truncation toward zero, saturating, NaN -> 0
//   normalize(v)  v / sqrt(dot(v,v)),  length(v) = sqrt(dot(v,v))
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPM_DEV __device__ __forceinline__

namespace mpm {

// ------------------------------------------------------------------ hashes
MPM_DEV uint32_t rotl17(uint32_t h) { return (h << 17) | (h >> 15); }
MPM_DEV uint32_t hash1(uint32_t n) {
    uint32_t h = n + 374761393u;
    h = 668265263u * rotl17(h);
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
MPM_DEV uint32_t hash3(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t h = z + 374761393u + x * 3266489917u;
    h = 668265263u * rotl17(h);
    h += y * 3266489917u;
    h = 668265263u * rotl17(h);
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}
MPM_DEV uint32_t hash4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    uint32_t h = w + 374761393u + x * 3266489917u;
    h = 668265263u * rotl17(h);
    h += y * 3266489917u;
    h = 668265263u * rotl17(h);
    h += z * 3266489917u;
    h = 668265263u * rotl17(h);
    h = 2246822519u * (h ^ (h >> 15));
    h = 3266489917u * (h ^ (h >> 13));
    return h ^ (h >> 16);
}

// ------------------------------------------------------------------ scalar helpers
MPM_DEV float as_f(uint32_t u) { return __uint_as_float(u); }
MPM_DEV uint32_t as_u(float f) { return __float_as_uint(f); }

// i32(f32) with WGSL semantics: v_cvt_i32_f32 truncates toward zero, saturates out-of-range
// values and converts NaN to 0 in hardware (a plain C cast leaves those cases undefined).
MPM_DEV int32_t f2i(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Fixed-point transfer term i32((x / d) * S) for a divisor d fixed for the whole launch, with
// the quotient formed from rd = RN(1/d) as q0 = RN(xc rd), e = xc - q0 d (exact by FMA),
// q = RN(q0 + e rd), where xc is x clamped to +-2^100 (NaN passes through). For |x| in
// [2^-80, 2^80] q is the IEEE quotient; outside it both forms truncate to 0 or saturate. The
// host uses this form for a (d, S) pair only after a device check over all 2^32 float inputs
// found it equal to i32((x / d) * S) everywhere (k_verify_fixed_div); otherwise it divides.
MPM_DEV int32_t fixed_div(float x, float d, float rd, float S, bool fast) {
    if (!fast) return f2i((x / d) * S);
    const float ax = __uint_as_float(__float_as_uint(x) & 0x7fffffffu);
    const float xc = ax > 0x1p100f ? __uint_as_float((__float_as_uint(x) & 0x80000000u) | 0x71800000u) : x;
    const float q0 = xc * rd;
    const float e = __builtin_fmaf(-q0, d, xc);
    return f2i(__builtin_fmaf(e, rd, q0) * S);
}
// i32(x / d) for |x| <= 2^100 by the same FMA-corrected quotient without the clamp. The host
// uses it only for a divisor that a device check over every float with |x| <= 2^100 found equal
// to i32(x / d) (k_verify_fixed_div, mode 1).
MPM_DEV int32_t quot_i32(float x, float d, float rd) {
    const float q0 = x * rd;
    const float e = __builtin_fmaf(-q0, d, x);
    return f2i(__builtin_fmaf(e, rd, q0));
}
// x / d for |x| <= 2^100 by the same FMA-corrected quotient; used only for a divisor that a device
// check found bit-identical to the IEEE quotient for every float with |x| <= 2^100 (mode 2).
MPM_DEV float qdiv(float x, float d, float rd) {
    const float q0 = x * rd;
    const float e = __builtin_fmaf(-q0, d, x);
    return __builtin_fmaf(e, rd, q0);
}

MPM_DEV float scale_pow2(float p, int n) {
    if (n > 127) { p = p * as_f(0x7f000000u); n -= 127; if (n > 127) n = 127; }
    if (n < -126) {
        p = p * as_f(0x0c800000u); n += 102;
        if (n < -126) { p = p * as_f(0x0c800000u); n += 102; }
        if (n < -126) n = -126;
    }
    return p * as_f((uint32_t)(n + 127) << 23);
}
MPM_DEV float fexp(float x) {
    if (x != x) return x;
    if (x > 88.72283935546875f) return __builtin_huge_valf();
    if (x < -103.97208404541015625f) return 0.0f;
    float t = x * 1.44269502162933349609375f;
    float n = (t + 12582912.0f) - 12582912.0f;
    float r = (x - n * 0.693359375f) - n * -2.12194440e-4f;
    float p = 1.98412698e-4f;
    p = p * r + 1.38888889e-3f;
    p = p * r + 8.33333377e-3f;
    p = p * r + 4.16666679e-2f;
    p = p * r + 1.66666672e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return scale_pow2(p, (int)n);
}
MPM_DEV float flog(float x) {
    if (x != x) return x;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return -__builtin_huge_valf();
    if (x == __builtin_huge_valf()) return x;
    int e = 0;
    uint32_t b = as_u(x);
    if ((b >> 23) == 0) { x = x * 8388608.0f; b = as_u(x); e = -23; }
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = as_f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float z = s * s;
    float q = 0.111111112f;
    q = q * z + 0.142857149f;
    q = q * z + 0.200000003f;
    q = q * z + 0.333333343f;
    q = q * z + 1.0f;
    float lm = (2.0f * s) * q;
    float fe = (float)e;
    return fe * 0.693359375f + (lm + fe * -2.12194440e-4f);
}
MPM_DEV float fpow(float x, float y) {
    if (x != x || y != y) return __builtin_nanf("");
    if (x == 1.0f) return 1.0f;
    if (x == 0.0f) return (y < 0.0f) ? __builtin_huge_valf() : ((y == 0.0f) ? 1.0f : 0.0f);
    if (x < 0.0f) return __builtin_nanf("");
    return fexp(y * flog(x));
}
MPM_DEV float fabs_(float x) { return as_f(as_u(x) & 0x7fffffffu); }

// ------------------------------------------------------------------ vec3 / mat3 (column-major)
struct v3 { float x, y, z; };
struct m3 { v3 c[3]; };

MPM_DEV v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
MPM_DEV v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
MPM_DEV v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
MPM_DEV v3 operator*(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
MPM_DEV v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
MPM_DEV v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
MPM_DEV v3 cmul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
MPM_DEV v3 cdiv(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
MPM_DEV float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
MPM_DEV v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
MPM_DEV float length(v3 a) { return __builtin_sqrtf(dot(a, a)); }
MPM_DEV v3 normalize(v3 a) { float l = length(a); return a / l; }
MPM_DEV float mabs_max(v3 a) { return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z)); }
MPM_DEV v3 vmin(v3 a, v3 b) { return mk(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
MPM_DEV v3 vmax(v3 a, v3 b) { return mk(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }

MPM_DEV m3 ident() { m3 m; m.c[0] = mk(1, 0, 0); m.c[1] = mk(0, 1, 0); m.c[2] = mk(0, 0, 1); return m; }
MPM_DEV m3 zero3() { m3 m; m.c[0] = mk(0, 0, 0); m.c[1] = mk(0, 0, 0); m.c[2] = mk(0, 0, 0); return m; }
MPM_DEV m3 madd(const m3& a, const m3& b) { m3 m; m.c[0] = a.c[0] + b.c[0]; m.c[1] = a.c[1] + b.c[1]; m.c[2] = a.c[2] + b.c[2]; return m; }
MPM_DEV m3 msub(const m3& a, const m3& b) { m3 m; m.c[0] = a.c[0] - b.c[0]; m.c[1] = a.c[1] - b.c[1]; m.c[2] = a.c[2] - b.c[2]; return m; }
MPM_DEV m3 mscale(float s, const m3& a) { m3 m; m.c[0] = s * a.c[0]; m.c[1] = s * a.c[1]; m.c[2] = s * a.c[2]; return m; }
MPM_DEV m3 mscale_r(const m3& a, float s) { m3 m; m.c[0] = a.c[0] * s; m.c[1] = a.c[1] * s; m.c[2] = a.c[2] * s; return m; }
MPM_DEV m3 transpose(const m3& a) {
    m3 m;
    m.c[0] = mk(a.c[0].x, a.c[1].x, a.c[2].x);
    m.c[1] = mk(a.c[0].y, a.c[1].y, a.c[2].y);
    m.c[2] = mk(a.c[0].z, a.c[1].z, a.c[2].z);
    return m;
}
MPM_DEV v3 mvmul(const m3& a, v3 v) { return (a.c[0] * v.x + a.c[1] * v.y) + a.c[2] * v.z; }
MPM_DEV m3 mmul(const m3& a, const m3& b) { m3 m; m.c[0] = mvmul(a, b.c[0]); m.c[1] = mvmul(a, b.c[1]); m.c[2] = mvmul(a, b.c[2]); return m; }
MPM_DEV float det3(const m3& m) {
    return (m.c[0].x * (m.c[1].y * m.c[2].z - m.c[2].y * m.c[1].z)
          - m.c[1].x * (m.c[0].y * m.c[2].z - m.c[2].y * m.c[0].z))
          + m.c[2].x * (m.c[0].y * m.c[1].z - m.c[1].y * m.c[0].z);
}
MPM_DEV m3 adj_scaled(const m3& m, float det) {
    const float inv = 1.0f / det;
    m3 r;
    r.c[0] = mk(m.c[1].y * m.c[2].z - m.c[1].z * m.c[2].y,
                m.c[0].z * m.c[2].y - m.c[0].y * m.c[2].z,
                m.c[0].y * m.c[1].z - m.c[0].z * m.c[1].y);
    r.c[1] = mk(m.c[1].z * m.c[2].x - m.c[1].x * m.c[2].z,
                m.c[0].x * m.c[2].z - m.c[0].z * m.c[2].x,
                m.c[0].z * m.c[1].x - m.c[0].x * m.c[1].z);
    r.c[2] = mk(m.c[1].x * m.c[2].y - m.c[1].y * m.c[2].x,
                m.c[0].y * m.c[2].x - m.c[0].x * m.c[2].y,
                m.c[0].x * m.c[1].y - m.c[0].y * m.c[1].x);
    return mscale_r(r, inv);
}
// A singular matrix inverts to the identity (matrixOps.wgsl:8-12). The test is made for the whole
// wave first: when no lane has det = 0 (the usual case) no lane carries the identity's selects.
MPM_DEV m3 inverse3(const m3& m) {
    const float det = det3(m);
    if (__builtin_expect(__ballot(det == 0.0f) == 0ull, 1)) return adj_scaled(m, det);
    if (det == 0.0f) return ident();
    return adj_scaled(m, det);
}
MPM_DEV m3 polar_rotation(const m3& f) {
    m3 r = f;
#pragma unroll
    for (int i = 0; i < 4; ++i) r = mscale(0.5f, madd(r, transpose(inverse3(r))));
    return r;
}
MPM_DEV m3 outer(v3 u, v3 v) { m3 m; m.c[0] = u * v.x; m.c[1] = u * v.y; m.c[2] = u * v.z; return m; }

}  // namespace mpm
