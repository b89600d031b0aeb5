// Oracle — TEST INFRASTRUCTURE ONLY. The 3x3 SVD of the SVD plasticity option, written from the
// published method (not from the library's source):
//   * eigen-decomposition of the symmetric A^T A by cyclic Jacobi rotations (Jacobi 1846; the
//     rotation of Golub & Van Loan, Matrix Computations, Algorithm 8.5.1 sym.schur2:
//     tau = (s_qq - s_pp) / (2 s_pq), t = sign(tau) / (|tau| + sqrt(1 + tau^2)),
//     c = 1 / sqrt(1 + t^2), s = t c; diagonal s_pp - t s_pq, s_qq + t s_pq; V <- V J), six sweeps
//     over the pivots (0,1), (0,2), (1,2), a rotation skipped when s_pq is exactly 0
//     (McAdams et al. 2011 use the same Jacobi-on-A^T A structure);
//   * B = A V, whose columns are sigma_j u_j: sigma_j = |b_j|;
//   * Stomakhin et al. 2013 §7: clamp sigma_j to [1 - theta_c, 1 + theta_s]; the new F_E is
//     sum_j (clamp(sigma_j) / sigma_j) b_j v_j^T and F_P = F_E_new^-1 F_E F_P (the reference's
//     update of F_P, plasticityOps.wgsl:36-47).
// FP32 by default; with the FP64 fallback a near-degenerate F_E (sigma_min < 1e-2 sigma_max, or a
// non-finite sigma) is decomposed and updated in double and rounded once. Each operation is one
// IEEE rounding in the written order (-ffp-contract=off), matching the device bit for bit.
#pragma once
#include <cmath>
#include "orc_math.h"

namespace orc {

template <class R> struct Svd3 { M3<R> V, B; R sig[3]; };

template <class R> static inline R mat_el(const M3<R>& m, int row, int col) { return (&m.c[col].x)[row]; }
template <class R> static inline R& mat_el(M3<R>& m, int row, int col) { return (&m.c[col].x)[row]; }

// S = A^T A; Jacobi sweeps on S accumulating V; B = A V; sigma_j = |b_j|.
template <class R> static inline Svd3<R> svd3(const M3<R>& A) {
    R S[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S[i][j] = dot(A.c[i], A.c[j]);
    Svd3<R> out;
    out.V = ident<R>();
    const int piv[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int sweep = 0; sweep < 6; ++sweep) {
        for (const auto& pq : piv) {
            const int p = pq[0], q = pq[1], r = 3 - p - q;
            const R apq = S[p][q];
            if (apq == (R)0) continue;
            const R tau = (S[q][q] - S[p][p]) / ((R)2 * apq);
            const R sgn = tau >= (R)0 ? (R)1 : (R)-1;
            const R t = sgn / (std::fabs(tau) + r_sqrt((R)1 + tau * tau));
            const R c = (R)1 / r_sqrt((R)1 + t * t);
            const R s = t * c;
            const R spp = S[p][p] - t * apq, sqq = S[q][q] + t * apq;
            const R srp = c * S[r][p] - s * S[r][q];
            const R srq = s * S[r][p] + c * S[r][q];
            S[p][p] = spp; S[q][q] = sqq; S[p][q] = S[q][p] = (R)0;
            S[r][p] = S[p][r] = srp; S[r][q] = S[q][r] = srq;
            const V3<R> vp = out.V.c[p], vq = out.V.c[q];
            out.V.c[p] = c * vp - s * vq;
            out.V.c[q] = s * vp + c * vq;
        }
    }
    for (int j = 0; j < 3; ++j) {
        out.B.c[j] = mvmul(A, out.V.c[j]);
        out.sig[j] = r_sqrt(dot(out.B.c[j], out.B.c[j]));
    }
    return out;
}

// Degenerate: a singular value that is zero, not finite (or beyond 3e38), or below 1e-2 of the largest.
template <class R> static inline bool svd_degenerate(const R sig[3]) {
    for (int j = 0; j < 3; ++j)
        if (!(sig[j] > (R)0) || !(sig[j] < (R)3.0e38)) return true;
    const R mn = std::min(std::min(sig[0], sig[1]), sig[2]);
    const R mx = std::max(std::max(sig[0], sig[1]), sig[2]);
    return mn < (R)1e-2 * mx;
}

// sum_j (clamp(sigma_j) / sigma_j) b_j v_j^T
template <class R> static inline M3<R> svd_clamped(const Svd3<R>& d, R lo, R hi) {
    R f[3];
    for (int j = 0; j < 3; ++j) {
        R t = d.sig[j] < lo ? lo : d.sig[j];
        t = hi < t ? hi : t;
        f[j] = t / d.sig[j];
    }
    M3<R> out;
    for (int k = 0; k < 3; ++k)
        for (int row = 0; row < 3; ++row)
            mat_el(out, row, k) = ((f[0] * mat_el(d.V, k, 0)) * mat_el(d.B, row, 0) + (f[1] * mat_el(d.V, k, 1)) * mat_el(d.B, row, 1)) +
                                  (f[2] * mat_el(d.V, k, 2)) * mat_el(d.B, row, 2);
    return out;
}

// The option's plasticity update of one particle: false when every singular value is in [lo, hi].
// *fp64 reports whether the FP64 path was taken (fallback on and F_E near-degenerate).
static inline bool plasticity_svd(M3<float>& Fe, M3<float>& Fp, float lo, float hi, bool fallback, bool* fp64) {
    const Svd3<float> d = svd3(Fe);
    *fp64 = fallback && svd_degenerate(d.sig);
    if (!*fp64) {
        if (lo <= d.sig[0] && lo <= d.sig[1] && lo <= d.sig[2] && d.sig[0] <= hi && d.sig[1] <= hi && d.sig[2] <= hi) return false;
        const M3<float> nFe = svd_clamped(d, lo, hi);
        const M3<float> nFp = mmul(inverse3(nFe), mmul(Fe, Fp));
        Fe = nFe; Fp = nFp;
        return true;
    }
    M3<double> A, P;
    for (int c = 0; c < 3; ++c) {
        A.c[c] = {(double)Fe.c[c].x, (double)Fe.c[c].y, (double)Fe.c[c].z};
        P.c[c] = {(double)Fp.c[c].x, (double)Fp.c[c].y, (double)Fp.c[c].z};
    }
    const Svd3<double> d64 = svd3(A);
    const double l = lo, h = hi;
    if (l <= d64.sig[0] && l <= d64.sig[1] && l <= d64.sig[2] && d64.sig[0] <= h && d64.sig[1] <= h && d64.sig[2] <= h) return false;
    const M3<double> nFe = svd_clamped(d64, l, h);
    const M3<double> nFp = mmul(inverse3(nFe), mmul(A, P));
    for (int c = 0; c < 3; ++c) {
        Fe.c[c] = {(float)nFe.c[c].x, (float)nFe.c[c].y, (float)nFe.c[c].z};
        Fp.c[c] = {(float)nFp.c[c].x, (float)nFp.c[c].y, (float)nFp.c[c].z};
    }
    return true;
}

}  // namespace orc
