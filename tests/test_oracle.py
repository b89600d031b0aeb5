"""Pins the CPU oracle (FP32-faithful restatement of the reference substep) against facts that do
not come from the oracle's own formulas: an external hash library, libm/numpy in FP64, closed
forms of the B-spline and the Newton polar iteration, brute force, and conservation laws."""
import ctypes
import struct

import numpy as np
import pytest

from helpers import uniforms_blob


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ulp_dist(a, b):
    """Distance in float32 ulps between float32 arrays a and b (finite values)."""
    ai = a.astype(np.float32).view(np.int32).astype(np.int64)
    bi = b.astype(np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


# ---------------------------------------------------------------- hashes (randomNumberGenerators.wgsl:3-56)
def test_hashes_are_xxh32_of_le_words(orc):
    """hash1/hash3/hash4 equal xxHash32 (the xxhash C library) of the little-endian words with the
    seed shifted by the input length: h = seed + P5 + len is the reference's start value."""
    xxhash = pytest.importorskip("xxhash")
    L = orc.lib()
    rng = np.random.default_rng(0)
    vals = rng.integers(0, 2**32, size=(400, 4), dtype=np.uint64)
    for x, y, z, w in vals.tolist():
        assert L.orc_hash1(x) == xxhash.xxh32_intdigest(b"\x00" * 4, seed=(x - 4) % 2**32)
        assert L.orc_hash3(x, y, z) == xxhash.xxh32_intdigest(struct.pack("<II", x, y), seed=(z - 8) % 2**32)
        assert L.orc_hash4(x, y, z, w) == xxhash.xxh32_intdigest(struct.pack("<III", x, y, z), seed=(w - 12) % 2**32)


# ---------------------------------------------------------------- exp / log / pow / i32
def test_exp_within_2ulp_of_fp64(orc):
    L = orc.lib()
    x = np.concatenate([np.linspace(-87, 88, 20001), np.random.default_rng(1).uniform(-5, 5, 5000)]).astype(np.float32)
    got = orc.f32_vec(L.orc_exp, x)
    want = np.exp(x.astype(np.float64)).astype(np.float32)
    assert ulp_dist(got, want).max() <= 2
    assert L.orc_exp(0.0) == 1.0 and L.orc_exp(-200.0) == 0.0 and np.isinf(L.orc_exp(100.0))
    assert np.isnan(L.orc_exp(float("nan")))


def test_log_pow_within_4ulp_of_fp64(orc):
    L = orc.lib()
    rng = np.random.default_rng(2)
    x = np.concatenate([np.geomspace(1e-30, 1e30, 8001), rng.uniform(0.5, 2.0, 4000), [1e-40, 1.0]]).astype(np.float32)
    lg = orc.f32_vec(L.orc_log, x)
    want = np.log(x.astype(np.float64))
    nz = np.abs(want) > 1e-3
    assert ulp_dist(lg[nz], want[nz].astype(np.float32)).max() <= 4
    assert np.all(np.abs(lg[~nz] - want[~nz]) <= 4 * np.spacing(np.float32(1e-3)))
    j = rng.uniform(0.2, 5.0, 4000).astype(np.float32)
    pw = orc.f32_vec(lambda a: L.orc_pow(a, np.float32(-0.3333333)), j)
    wantp = (j.astype(np.float64) ** np.float64(np.float32(-0.3333333))).astype(np.float32)
    assert ulp_dist(pw, wantp).max() <= 4
    assert L.orc_pow(0.0, -0.5) == float("inf") and L.orc_pow(1.0, -7.0) == 1.0


def test_f32_to_i32_is_truncating_and_saturating(orc):
    L = orc.lib()
    cases = {0.0: 0, -0.0: 0, 1.9: 1, -1.9: -1, 2.5e9: 2**31 - 1, -2.5e9: -2**31, float("inf"): 2**31 - 1,
             float("-inf"): -2**31, float("nan"): 0, 2147483520.0: 2147483520, -2147483648.0: -2**31}
    for x, want in cases.items():
        assert L.orc_f2i(x) == want, x


# ---------------------------------------------------------------- B-spline (mpmOps.wgsl:34-53)
def test_bspline_closed_forms(orc):
    L = orc.lib()
    f = np.random.default_rng(3).uniform(0.0, 1.0, (2000, 3)).astype(np.float32)
    for row in f:
        w = np.zeros(9, np.float32)
        d = np.zeros(9, np.float32)
        L.orc_bspline(_p(np.ascontiguousarray(row)), _p(w), _p(d))
        w = w.reshape(3, 3).astype(np.float64)
        d = d.reshape(3, 3).astype(np.float64)
        ff = row.astype(np.float64)
        # textbook quadratic B-spline N(r) at node distances r = f + 1/2 - i (nodes at cell centres)
        for i in range(3):
            r = np.abs(ff + 0.5 - i)
            n = np.where(r < 0.5, 0.75 - r * r, np.where(r < 1.5, 0.5 * (1.5 - r) ** 2, 0.0))
            assert np.allclose(w[i], n, atol=2e-7)
        assert np.allclose(w.sum(0), 1.0, atol=3e-7)                        # partition of unity
        pos = np.array([-1.0, 0.0, 1.0])[:, None] + 0.5 - ff[None, :]
        assert np.allclose((w * pos).sum(0), 0.0, atol=5e-7)                # first moment
        assert np.allclose((w * pos * pos).sum(0), 0.25, atol=5e-7)         # second moment (APIC D = 4 B)
        assert np.allclose(d.sum(0), 0.0, atol=2e-7)                        # derivative sums to 0


# ---------------------------------------------------------------- 3x3 algebra (matrixOps.wgsl:8-45)
def _mat(L, fn, m):
    m = np.ascontiguousarray(m.T.reshape(9), np.float32)   # row-major numpy -> column-major
    o = np.zeros(9, np.float32)
    getattr(L, fn)(_p(m), _p(o))
    return o.reshape(3, 3).T.astype(np.float64)


def test_det_and_inverse_against_fp64(orc):
    L = orc.lib()
    rng = np.random.default_rng(4)
    for _ in range(500):
        m = (np.eye(3) + 0.4 * rng.standard_normal((3, 3))).astype(np.float32)
        mc = np.ascontiguousarray(m.T.reshape(9))
        det = L.orc_det3(_p(mc))
        assert abs(det - np.linalg.det(m.astype(np.float64))) <= 1e-5 * max(1.0, np.abs(m).max() ** 3)
        inv = _mat(L, "orc_inverse3", m)
        want = np.linalg.inv(m.astype(np.float64))
        assert np.allclose(inv, want, rtol=2e-5, atol=2e-5 * np.abs(want).max())
    singular = np.array([[1, 2, 3], [2, 4, 6], [0, 1, 1]], np.float32)
    assert np.array_equal(_mat(L, "orc_inverse3", singular), np.eye(3))     # det == 0 -> identity


def test_polar_newton_closed_forms(orc):
    L = orc.lib()
    rng = np.random.default_rng(5)
    for _ in range(100):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] *= -1
        r = _mat(L, "orc_polar3", q.astype(np.float32))
        assert np.abs(r - q).max() < 1e-6                                   # rotations are fixed points
    for _ in range(100):
        s = rng.uniform(0.3, 3.0, 3)
        r = _mat(L, "orc_polar3", np.diag(s).astype(np.float32))
        t = s.copy()
        for _k in range(4):                                                 # s <- (s + 1/s) / 2, four times
            t = 0.5 * (t + 1.0 / t)
        assert np.allclose(np.diag(r), t, rtol=3e-6)
        assert np.abs(r - np.diag(np.diag(r))).max() == 0.0


# ---------------------------------------------------------------- collider primitives (colliderPrelude.wgsl:37-102)
def test_closest_point_triangle_brute_force(orc):
    L = orc.lib()
    rng = np.random.default_rng(6)
    uv = np.array([(u, v) for u in np.linspace(0, 1, 121) for v in np.linspace(0, 1, 121) if u + v <= 1.0])
    for _ in range(150):
        a, b, c = rng.standard_normal((3, 3)).astype(np.float32)
        p = (rng.standard_normal(3) * 1.5).astype(np.float32)
        out = np.zeros(3, np.float32)
        L.orc_closest_point_triangle(_p(p), _p(a), _p(b), _p(c), _p(out))
        samples = a[None] + uv[:, :1] * (b - a)[None] + uv[:, 1:] * (c - a)[None]
        dmin = np.sqrt(((samples - p) ** 2).sum(1)).min()
        dgot = np.sqrt(((out.astype(np.float64) - p) ** 2).sum())
        assert dgot <= dmin + 1e-5
        # the returned point lies on the triangle (barycentric coordinates in [0,1])
        m = np.stack([b - a, c - a], 1).astype(np.float64)
        coef, *_ = np.linalg.lstsq(m, out.astype(np.float64) - a, rcond=None)
        assert coef.min() >= -1e-4 and coef.sum() <= 1 + 1e-4


def test_ray_triangle_against_linear_solve(orc):
    L = orc.lib()
    rng = np.random.default_rng(7)
    hits = 0
    for _ in range(2000):
        a, b, c = rng.standard_normal((3, 3)).astype(np.float32)
        o = (rng.standard_normal(3) * 2).astype(np.float32)
        if _ % 2:   # aimed at a random point of the triangle, through it at t = 0.8
            u0, v0 = rng.uniform(0, 1, 2)
            if u0 + v0 > 1:
                u0, v0 = 1 - u0, 1 - v0
            d = ((a + u0 * (b - a) + v0 * (c - a) - o) / 0.8).astype(np.float32)
        else:
            d = (rng.standard_normal(3) * 2).astype(np.float32)
        t = L.orc_intersect_ray_triangle(_p(o), _p(d), _p(a), _p(b), _p(c))
        m = np.stack([-d, b - a, c - a], 1).astype(np.float64)
        if abs(np.linalg.det(m)) < 1e-3:
            continue
        tt, u, v = np.linalg.solve(m, (o - a).astype(np.float64))
        inside = u > 1e-4 and v > 1e-4 and u + v < 1 - 1e-4 and tt > 1e-4
        outside = u < -1e-4 or v < -1e-4 or u + v > 1 + 1e-4 or tt < -1e-4
        if inside:
            hits += 1
            assert abs(t - tt) < 1e-4 * max(1.0, tt)
        elif outside:
            assert t == -1.0
    assert hits > 500


# ---------------------------------------------------------------- one-substep laws of the restated step
def test_pbmpm_first_transfer_conserves_mass_and_momentum(orc):
    from helpers import particles_in_spheres
    p = particles_in_spheres(3000, np.random.default_rng(8), vx=2.0)
    dt = np.float32(1.0 / 600)
    p["dx"][:, :3] = p["vel"] * dt                                       # dx = v dt, D = 0 (rigid translation)
    uni = uniforms_blob(dt, (0, 0, 0), (1, 1, 1), (128, 128, 128))
    o = orc.OracleSim(p, uni, pbmpm_iterations=1, pbmpm_force_divisor=1.0, gravity=(0.0, 0.0, 0.0),
                      hardening_pbmpm=1e9)                                # relax -> 1: D becomes the solve correction
    o.run_stage(1, 1)
    xyz, vals = o.dump_grid()
    S = 65536.0
    mass = vals[:, 0].sum() / S
    # each of the 27 per-node contributions truncates by < 1 fixed-point unit
    assert abs(mass - p["mass"].sum()) <= 27 * p.shape[0] / S
    # F_e = I, D = 0: the solve leaves D = 0 (rotation I, J = 1), so momentum is m v exactly up to truncation
    mom = vals[:, 1:].sum(0) / S
    want = (p["mass"][:, None] * p["vel"]).sum(0)
    assert np.all(np.abs(mom - want) <= 27 * p.shape[0] / S * 3)


def test_pbmpm_gravity_scatter_sums_to_m_g_dt_over_k(orc):
    from helpers import particles_in_spheres
    p = particles_in_spheres(2000, np.random.default_rng(9), vx=0.0)
    dt = np.float32(1.0 / 600)
    uni = uniforms_blob(dt, (0, 0, 0), (1, 1, 1), (128, 128, 128))
    k = 4
    o = orc.OracleSim(p, uni, pbmpm_iterations=k, pbmpm_force_divisor=float(k), gravity=(0.0, -9.8, 0.0))
    o.run_stage(1, 1)
    xyz, vals = o.dump_grid()
    S = 65536.0
    my = vals[:, 2].sum() / S
    want = p["mass"].sum() * np.float32(-9.8) / k * dt
    assert abs(my - want) <= 27 * p.shape[0] / S                          # truncation toward zero per term


def test_pbmpm_g2p_reproduces_uniform_velocity(orc):
    """All particles translating with one velocity: G2P returns that velocity (to fixed-point
    resolution) and B ~ 0 because sum_i w_i (x_i - x_p) = 0."""
    from helpers import particles_in_spheres
    p = particles_in_spheres(4000, np.random.default_rng(10), vx=0.0)
    p["vel"][:, 0] = 1.5
    dt = np.float32(1.0 / 600)
    p["dx"][:, :3] = p["vel"] * dt
    uni = uniforms_blob(dt, (0, 0, 0), (1, 1, 1), (128, 128, 128))
    o = orc.OracleSim(p, uni, pbmpm_iterations=1, pbmpm_force_divisor=1.0, gravity=(0.0, 0.0, 0.0))
    o.step(1, 1)
    q = o.particles()
    assert np.allclose(q["vel"][:, 0], 1.5, atol=2e-3)
    assert np.abs(q["vel"][:, 1:]).max() < 2e-3
    assert np.abs(q["D"][:, :, :3]).max() < 2e-3


def test_explicit_grid_update_free_fall(orc):
    """gridUpdate.cs.wgsl:40-52: a resting cloud with F = I gets v = g dt on every node with mass."""
    from helpers import particles_in_spheres
    p = particles_in_spheres(3000, np.random.default_rng(11), vx=0.0)
    dt = np.float32(1e-4)
    uni = uniforms_blob(dt, (0, 0, 0), (1, 1, 1), (64, 64, 64))
    o = orc.OracleSim(p, uni, gravity=(0.0, -9.8, 0.0))
    o.run_stage(0, 2)
    xyz, vals = o.dump_grid()
    m = vals[:, 0]
    sel = m > 0
    # node momentum after the update is g dt m, truncated to the 2^-16 fixed-point grid
    want = np.float64(np.float32(-9.8) * dt) * m[sel]
    assert np.abs(vals[sel, 2] - want).max() < 1.0 + 1e-6 * np.abs(want).max()
    assert np.abs(vals[sel, 1]).max() <= 1 and np.abs(vals[sel, 3]).max() <= 1


def test_plasticity_closed_forms(orc):
    """plasticityOps.wgsl:7-48 through one explicit substep of isolated particles at rest with D = 0:
    F_e in the [1-theta_c, 1+theta_s] band is untouched; a diagonal F_e is clamped per axis and
    F_e F_p is preserved."""
    from helpers import particles_in_spheres
    p = particles_in_spheres(3, np.random.default_rng(12), vx=0.0)
    p["pos"] = [[0.2, 0.2, 0.2], [0.5, 0.5, 0.5], [0.8, 0.8, 0.8]]
    p["Fe"][0, :, :3] = np.diag([1.003, 0.99, 1.0])
    p["Fe"][1, :, :3] = np.diag([1.2, 0.9, 1.001])
    p["Fe"][2, :, :3] = np.diag([0.5, 1.0, 1.05])
    fe0 = p["Fe"][:, :, :3].astype(np.float64).copy()
    uni = uniforms_blob(np.float32(1e-9), (0, 0, 0), (1, 1, 1), (32, 32, 32))
    o = orc.OracleSim(p, uni, gravity=(0.0, 0.0, 0.0), youngs_modulus=0.0)
    o.step(0, 1)
    q = o.particles()
    fe = q["Fe"][:, :, :3].astype(np.float64)
    fp = q["Fp"][:, :, :3].astype(np.float64)
    assert np.allclose(fe[0], fe0[0], atol=1e-7)
    lo, hi = np.float32(1 - 2.5e-2), np.float32(1 + 7.5e-3)
    assert np.allclose(np.diag(fe[1]), np.clip(np.diag(fe0[1]), lo, hi), atol=2e-6)
    assert np.allclose(np.diag(fe[2]), np.clip(np.diag(fe0[2]), lo, hi), atol=2e-6)
    for i in range(3):   # column-major storage: F = Fe Fp as matrices (transposed consistently)
        assert np.allclose(fp[i] @ fe[i], fe0[i], atol=5e-6)


def test_fp32_restatement_tracks_fp64_twin(orc):
    """The FP32-faithful oracle against the FP64 twin of the same algorithm (same fixed-point grid)
    after 1 and 10 PBMPM substeps: positions within 1e-5 of the domain, F within 1e-4."""
    from helpers import particles_in_spheres
    p = particles_in_spheres(6000, np.random.default_rng(13), vx=2.0)
    uni = uniforms_blob(np.float32(1.0 / 600), (0, 0, 0), (1, 1, 1), (128, 128, 128))
    kw = dict(pbmpm_iterations=4, pbmpm_force_divisor=4.0, gravity=(0.0, -9.8, 0.0))
    a = orc.OracleSim(p, uni, mode=0, **kw)
    b = orc.OracleSim(p, uni, mode=1, **kw)
    for n, (tol_x, tol_f) in ((1, (1e-6, 1e-5)), (9, (1e-5, 1e-4))):
        a.step(1, n)
        b.step(1, n)
        x, y = a.state_f64(), b.state_f64()
        assert np.abs(x[:, 0:3] - y[:, 0:3]).max() < tol_x
        assert np.abs(x[:, 7:16] - y[:, 7:16]).max() < tol_f
