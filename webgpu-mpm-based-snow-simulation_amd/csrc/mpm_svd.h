// 3x3 singular value decomposition for the SVD plasticity option (the snow model's singular-value
// clamp of Stomakhin et al. 2013, "A material point method for snow simulation", §7: F_E = U S V^T,
// S clamped to [1 - theta_c, 1 + theta_s], F_P absorbing the rest). No reference counterpart: the
// reference decides plasticity from diag(R^T F_E) after a 4-step Newton polar decomposition
// (plasticityOps.wgsl:7-48), which stays the default.
//
// Method: the cyclic Jacobi eigenvalue method on the symmetric A^T A (Jacobi 1846; the rotation of
// Golub & Van Loan, Matrix Computations, Alg. 8.5.1 sym.schur2: tau = (a_qq - a_pp) / (2 a_pq),
// t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c), the approach of
// McAdams et al. 2011 ("Computing the singular value decomposition of 3x3 matrices with minimal
// branching and elementary floating point operations") with exact rather than approximate Givens
// rotations: SVD_SWEEPS sweeps over the pivots (0,1), (0,2), (1,2) accumulate V; then B = A V,
// whose columns are sigma_j u_j, so sigma_j = |b_j| and no QR step is needed. The clamped matrix
// is sum_j (clamp(sigma_j) / sigma_j) b_j v_j^T.
//
// Templated on the scalar type: FP32 for the default path and FP64 for the fallback on a
// near-degenerate F (sigma_min < SVD_DEGENERATE sigma_max, or a non-finite sigma), where the
// eigenvalues of A^T A lose sigma_min's relative accuracy in FP32. Every expression is one IEEE
// rounding per operation in the written order (the library is built with -ffp-contract=off and
// correctly rounded division and square root), so the CPU oracle reproduces it bit for bit.
// Matrices are column-major arrays: a[3 * column + row].
#pragma once
#include "mpm_math.h"

namespace mpm {

enum : int { SVD_SWEEPS = 6 };
#define MPM_SVD_DEGENERATE 1e-2

template <class T> MPM_DEV T svd_sqrt(T x);
template <> MPM_DEV float svd_sqrt<float>(float x) { return __builtin_sqrtf(x); }
template <> MPM_DEV double svd_sqrt<double>(double x) { return __builtin_sqrt(x); }
template <class T> MPM_DEV T svd_abs(T x) { return x < T(0) ? -x : x; }
template <class T> MPM_DEV T dot3(const T* a, const T* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// One Jacobi rotation of the symmetric s (full 3x3, s[3 * i + j]) zeroing s_pq; V <- V J.
template <class T> MPM_DEV void jacobi_rotate(T* s, T* v, int p, int q) {
    const T apq = s[3 * p + q];
    if (apq == T(0)) return;
    const int r = 3 - p - q;
    const T tau = (s[3 * q + q] - s[3 * p + p]) / (T(2) * apq);
    const T t = (tau >= T(0) ? T(1) : T(-1)) / (svd_abs(tau) + svd_sqrt(T(1) + tau * tau));
    const T c = T(1) / svd_sqrt(T(1) + t * t);
    const T sn = t * c;
    const T spp = s[3 * p + p] - t * apq;
    const T sqq = s[3 * q + q] + t * apq;
    const T srp = c * s[3 * r + p] - sn * s[3 * r + q];
    const T srq = sn * s[3 * r + p] + c * s[3 * r + q];
    s[3 * p + p] = spp; s[3 * q + q] = sqq;
    s[3 * p + q] = T(0); s[3 * q + p] = T(0);
    s[3 * r + p] = srp; s[3 * p + r] = srp;
    s[3 * r + q] = srq; s[3 * q + r] = srq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T vp = v[3 * p + k], vq = v[3 * q + k];
        v[3 * p + k] = c * vp - sn * vq;
        v[3 * q + k] = sn * vp + c * vq;
    }
}

// v (columns v_j), b = A v (columns b_j), sigma_j = |b_j|.
template <class T> MPM_DEV void svd3(const T* a, T* v, T* b, T* sig) {
    T s[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) s[3 * i + j] = dot3(a + 3 * i, a + 3 * j);
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = (k % 4 == 0) ? T(1) : T(0);
    for (int sweep = 0; sweep < SVD_SWEEPS; ++sweep) {
        jacobi_rotate(s, v, 0, 1);
        jacobi_rotate(s, v, 0, 2);
        jacobi_rotate(s, v, 1, 2);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int r = 0; r < 3; ++r) b[3 * j + r] = (a[r] * v[3 * j] + a[3 + r] * v[3 * j + 1]) + a[6 + r] * v[3 * j + 2];
        sig[j] = svd_sqrt(dot3(b + 3 * j, b + 3 * j));
    }
}

// Degenerate: a singular value that is zero, not finite (or beyond 3e38), or below
// MPM_SVD_DEGENERATE of the largest.
template <class T> MPM_DEV bool svd_degenerate(const T* sig) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
        if (!(sig[j] > T(0)) || !(sig[j] < (T)3.0e38)) return true;
    const T mn = sig[0] < sig[1] ? (sig[0] < sig[2] ? sig[0] : sig[2]) : (sig[1] < sig[2] ? sig[1] : sig[2]);
    const T mx = sig[0] > sig[1] ? (sig[0] > sig[2] ? sig[0] : sig[2]) : (sig[1] > sig[2] ? sig[1] : sig[2]);
    return mn < T(MPM_SVD_DEGENERATE) * mx;
}

template <class T> MPM_DEV bool svd_in_band(const T* sig, T lo, T hi) {
    return lo <= sig[0] && lo <= sig[1] && lo <= sig[2] && sig[0] <= hi && sig[1] <= hi && sig[2] <= hi;
}

// out = sum_j (clamp(sigma_j) / sigma_j) b_j v_j^T
template <class T> MPM_DEV void svd_clamp(const T* v, const T* b, const T* sig, T lo, T hi, T* out) {
    T f[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        T t = sig[j] < lo ? lo : sig[j];
        t = hi < t ? hi : t;
        f[j] = t / sig[j];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            out[3 * k + r] = (((f[0] * v[k]) * b[r] + (f[1] * v[3 + k]) * b[3 + r])) + (f[2] * v[6 + k]) * b[6 + r];
}

// det / inverse (adjugate over det, identity for det == 0) / product, column-major, in the order
// of matrixOps.wgsl:8-27 as the FP32 helpers of mpm_math.h evaluate them.
template <class T> MPM_DEV T det3a(const T* m) {
    return (m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2])) + m[6] * (m[1] * m[5] - m[4] * m[2]);
}
template <class T> MPM_DEV void inverse3a(const T* m, T* o) {
    const T det = det3a(m);
    if (det == T(0)) {
#pragma unroll
        for (int k = 0; k < 9; ++k) o[k] = (k % 4 == 0) ? T(1) : T(0);
        return;
    }
    const T inv = T(1) / det;
    T r[9];
    r[0] = m[4] * m[8] - m[5] * m[7]; r[1] = m[2] * m[7] - m[1] * m[8]; r[2] = m[1] * m[5] - m[2] * m[4];
    r[3] = m[5] * m[6] - m[3] * m[8]; r[4] = m[0] * m[8] - m[2] * m[6]; r[5] = m[2] * m[3] - m[0] * m[5];
    r[6] = m[3] * m[7] - m[4] * m[6]; r[7] = m[1] * m[6] - m[0] * m[7]; r[8] = m[0] * m[4] - m[1] * m[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = r[k] * inv;
}
template <class T> MPM_DEV void mmul3a(const T* a, const T* b, T* o) {
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int r = 0; r < 3; ++r) o[3 * j + r] = (a[r] * b[3 * j] + a[3 + r] * b[3 * j + 1]) + a[6 + r] * b[3 * j + 2];
}

// The SVD clamp of F_e (column-major fe) with F_p (fp): false when every singular value is in
// [lo, hi] (nothing changes), else the new F_e and F_p = F_e_new^-1 (F_e F_p). With fp64_fallback a
// near-degenerate F_e is decomposed and updated in FP64 and rounded once at the end; *used_fp64
// tells whether it was.
MPM_DEV bool plasticity_svd(const float* fe, const float* fp, float lo, float hi, bool fp64_fallback,
                            float* fe_new, float* fp_new, bool* used_fp64) {
    float v[9], b[9], sig[3];
    svd3<float>(fe, v, b, sig);
    *used_fp64 = fp64_fallback && svd_degenerate<float>(sig);
    if (!*used_fp64) {
        if (svd_in_band<float>(sig, lo, hi)) return false;
        float inv[9], F[9];
        svd_clamp<float>(v, b, sig, lo, hi, fe_new);
        inverse3a<float>(fe_new, inv);
        mmul3a<float>(fe, fp, F);
        mmul3a<float>(inv, F, fp_new);
        return true;
    }
    double a[9], vd[9], bd[9], sd[3], nfe[9], inv[9], F[9], p[9], nfp[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) { a[k] = (double)fe[k]; p[k] = (double)fp[k]; }
    svd3<double>(a, vd, bd, sd);
    if (svd_in_band<double>(sd, (double)lo, (double)hi)) return false;
    svd_clamp<double>(vd, bd, sd, (double)lo, (double)hi, nfe);
    inverse3a<double>(nfe, inv);
    mmul3a<double>(a, p, F);
    mmul3a<double>(inv, F, nfp);
#pragma unroll
    for (int k = 0; k < 9; ++k) { fe_new[k] = (float)nfe[k]; fp_new[k] = (float)nfp[k]; }
    return true;
}

}  // namespace mpm
