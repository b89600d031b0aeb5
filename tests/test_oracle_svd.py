"""Pins the oracle's 3x3 SVD of the SVD plasticity option (oracle/orc_svd.h) against numpy's FP64
SVD on adversarial matrices: well-conditioned, repeated singular values, reflections, exactly
diagonal, rotations, near-singular and singular; and the plasticity update built on it against
U clamp(S) V^T in FP64 (Stomakhin et al. 2013 §7) and against the invariant F_E F_P = F_E' F_P'.
The FP64 fallback must be taken where FP32 Jacobi on A^T A cannot resolve the smallest singular
value, and must then be accurate."""
import numpy as np
import pytest

LO, HI = float(np.float32(1 - 2.5e-2)), float(np.float32(1 + 7.5e-3))


def rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def f32(m):
    return np.asarray(m, np.float32).astype(np.float64)


def np_clamp(F, lo=LO, hi=HI):
    u, s, vt = np.linalg.svd(F)
    return u @ np.diag(np.clip(s, lo, hi)) @ vt, s


def check_decomposition(orc, F, fp64, rtol):
    sig, V, B = orc.svd3(F, fp64=fp64)
    ref = np.linalg.svd(F, compute_uv=False)
    assert np.allclose(np.sort(sig)[::-1], ref, rtol=rtol, atol=rtol * ref[0]), (sig, ref)
    assert np.allclose(V.T @ V, np.eye(3), atol=4e-7 if not fp64 else 1e-7)
    assert np.allclose(B @ V.T, F, atol=2e-6 * np.abs(F).max())          # A = B V^T
    G = B.T @ B                                                          # columns orthogonal
    assert np.abs(G - np.diag(np.diag(G))).max() <= 2e-6 * ref[0] ** 2


def test_svd_well_conditioned_and_reflections(orc):
    rng = np.random.default_rng(80)
    for i in range(300):
        F = f32(rot(rng) @ np.diag(rng.uniform(0.5, 2.0, 3)) @ rot(rng).T)
        if i % 3 == 0:
            F[:, 0] *= -1                                                # a reflection (det < 0)
        check_decomposition(orc, F, False, 2e-6)


def test_svd_repeated_and_diagonal(orc):
    rng = np.random.default_rng(81)
    for i in range(200):
        s = rng.uniform(0.8, 1.2)
        t = s if i % 2 else rng.uniform(0.8, 1.2)
        F = f32(rot(rng) @ np.diag([s, s, t]) @ rot(rng).T)
        check_decomposition(orc, F, False, 2e-6)
    for d in ([1.1, 0.9, 1.0], [0.97, 0.97, 0.97], [2.0, -0.5, 1.0]):
        sig, V, B = orc.svd3(np.diag(d))
        assert np.array_equal(np.sort(sig), np.sort(np.abs(np.float32(d)).astype(np.float64)))  # no rotation needed


@pytest.mark.parametrize("kappa", [1e2, 1e3, 1e5])
def test_svd_near_singular_needs_fp64(orc, kappa):
    """sigma_min = sigma_max / kappa: the FP32 path loses sigma_min's relative accuracy (A^T A squares
    kappa); the FP64 path keeps it to 1e-6."""
    rng = np.random.default_rng(82)
    worst32, worst64 = 0.0, 0.0
    for _ in range(50):
        F = f32(rot(rng) @ np.diag([1.0, rng.uniform(0.5, 1.0), 1.0 / kappa]) @ rot(rng).T)
        ref = np.linalg.svd(F, compute_uv=False)
        s32 = np.sort(orc.svd3(F)[0])[::-1]
        s64 = np.sort(orc.svd3(F, fp64=True)[0])[::-1]
        worst32 = max(worst32, abs(s32[2] - ref[2]) / ref[2])
        worst64 = max(worst64, abs(s64[2] - ref[2]) / ref[2])
    assert worst64 < 1e-6
    if kappa >= 1e3:
        assert worst32 > 10 * worst64


def test_plasticity_update_against_fp64(orc):
    """F_E' = U clamp(S) V^T, F_P' = F_E'^-1 F_E F_P; in-band F_E unchanged; near-singular F_E through
    the fallback."""
    rng = np.random.default_rng(83)
    n_changed = 0
    for i in range(400):
        if i % 4 == 0:
            s = rng.uniform(LO + 1e-3, HI - 1e-3, 3)                         # in band
        elif i % 4 == 1:
            s = rng.uniform(0.9, 1.05, 3)                                    # straddles both edges
        elif i % 4 == 2:
            s = np.array([rng.uniform(0.9, 1.05), rng.uniform(0.9, 1.05), rng.uniform(1e-4, 5e-3)])   # near-singular
        else:
            s = rng.uniform(0.95, 1.0, 3) * np.array([1, 1, -1])             # reflection
        Fe = f32(rot(rng) @ np.diag(s) @ rot(rng).T)
        Fp = f32(np.eye(3) + 0.05 * rng.standard_normal((3, 3)))
        fallback = i % 2 == 0
        changed, Fe2, Fp2, took64 = orc.plasticity_svd(Fe, Fp, LO, HI, fallback=fallback)
        want, sv = np_clamp(Fe)
        degenerate = sv.min() < 1e-2 * sv.max()
        assert took64 == (fallback and degenerate)
        inband = bool(np.all((sv >= LO) & (sv <= HI)))
        if inband:
            assert not changed and np.array_equal(Fe2, Fe) and np.array_equal(Fp2, Fp)
            continue
        n_changed += changed
        tol = 2e-6 if (not degenerate or took64) else 5e-2
        assert np.abs(Fe2 - want).max() < tol, (i, np.abs(Fe2 - want).max())
        total = Fe @ Fp
        assert np.abs(Fe2 @ Fp2 - total).max() < (1e-5 if (not degenerate or took64) else 5e-2) * np.abs(total).max()
    assert n_changed > 200


def test_svd_and_polar_agree_on_diagonal_f(orc):
    """For a diagonal F_E both decide plasticity on the same values: the clamp of the reference's
    diag(R^T F_E) check and the SVD clamp give the same clamped diagonal (to rounding)."""
    from helpers import uniforms_blob
    import __graft_entry__
    m = __graft_entry__.load_package()
    rng = np.random.default_rng(84)
    n = 200
    d = rng.uniform(0.9, 1.05, (n, 3)).astype(np.float32)
    p = m.make_particles(rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32))
    for i in range(n):
        p["Fe"][i, :, :3] = np.diag(d[i])
    uni = uniforms_blob(np.float32(1e-9), (0, 0, 0), (1, 1, 1), (32, 32, 32))
    outs = []
    for mode in (0, 1):
        o = orc.OracleSim(p, uni, gravity=(0.0, 0.0, 0.0), youngs_modulus=0.0, plasticity_mode=mode)
        o.step(0, 1)
        outs.append(o.particles()["Fe"][:, :, :3].astype(np.float64))
    for i in range(n):
        assert np.allclose(np.diag(outs[0][i]), np.clip(d[i], LO, HI), atol=2e-6)
        assert np.allclose(outs[0][i], outs[1][i], atol=3e-6)
