"""Pins the oracle's PBMPM constraint solve, explicit stress, wall response, collider response and
G2P velocity gradient against facts that do not come from the oracle's own formulas: closed forms
of the 4-step Newton polar iteration on diagonal, rotated and symmetric matrices; the textbook
corotated stress evaluated in FP64 with numpy's SVD polar factor; the B-spline's reproduction of
affine fields; and the geometry of a ray crossing a plane. Each case is chosen so that a plausible
slip (swapped blend weights, a sign of an exponent, a transposed product, a dropped factor) fails.

Reference: pbmpm.wgsl:16-39 (solve), stressTensorOps.wgsl:9-22,39-56 (PK1, hardening),
integrateParticles.wgsl:33-64 (walls), colliderPrelude.wgsl:293-356 (collider response),
gridToParticle.cs.wgsl:28-80 (explicit G2P), :82-130 (PBMPM G2P)."""
import numpy as np
import pytest

from helpers import uniforms_blob

EXPO = float(np.float32(-0.3333333))   # pbmpm.wgsl:23 volume exponent
BLEND = 0.95                            # pbmpm.wgsl:26


def newton4(s):
    """The polar iteration R <- (R + R^-T)/2 on a diagonal entry, four times (matrixOps.wgsl:33-45)."""
    r = np.asarray(s, np.float64)
    for _ in range(4):
        r = 0.5 * (r + 1.0 / r)
    return r


def rotation(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def relax(jp, hardening=20.0):
    return 0.5 + 0.5 * (1.0 - np.exp(-hardening / jp))


def f32(m):
    return np.asarray(m, np.float32).astype(np.float64)


# ---------------------------------------------------------------- PBMPM solve (pbmpm.wgsl:16-39)
@pytest.mark.parametrize("s", [0.8, 0.93, 1.0, 1.07, 1.3])
def test_pbmpm_solve_scaled_identity(orc, s):
    """F_e = s I, D = 0, J_p = 1: trial = s I, R = r4(s) I, |det|^e = s^(3e), so
    corr = (r4(s) (0.95 s^(3e) + 0.05) / s - 1) I and relax = 1 - exp(-20)/2."""
    s = float(np.float32(s))
    got = orc.solve_constraints(s * np.eye(3), np.eye(3), np.zeros((3, 3)))
    c = newton4(s) * (BLEND * s ** (3 * EXPO) + (1 - BLEND)) / s - 1.0
    assert np.allclose(got, relax(1.0) * c * np.eye(3), atol=3e-6), (got, c)


def test_pbmpm_solve_rotated_anisotropic_with_relaxation(orc):
    """F_e = Q diag(s), D = 0: the Newton iterates are Q diag(r_n(s)), so
    corr = Q diag(r4(s) (0.95 v + 0.05) / s) Q^T - I with v = (s1 s2 s3)^e; the relax factor
    0.5 + 0.5 (1 - exp(-h / J_p)) scales it, J_p = det F_p."""
    rng = np.random.default_rng(20)
    for jp, h in ((1.0, 20.0), (5.0, 20.0), (20.0, 20.0), (60.0, 20.0), (3.0, 7.0)):
        for _ in range(20):
            q = rotation(rng)
            s = rng.uniform(0.8, 1.25, 3)
            fe = f32(q @ np.diag(s))
            # the closed form needs F_e exactly Q diag(s): take s from the float32 matrix's SVD
            u, sv, vt = np.linalg.svd(fe)
            fp = np.diag([jp, 1.0, 1.0])
            got = orc.solve_constraints(fe, fp, np.zeros((3, 3)), hardening_pbmpm=h)
            v = np.prod(sv) ** EXPO
            # fe = u diag(sv) vt: Newton iterates u diag(r_n) vt; corr = u diag(r4 c / sv) u^T - I
            corr = u @ np.diag(newton4(sv) * (BLEND * v + (1 - BLEND)) / sv) @ u.T - np.eye(3)
            want = relax(jp, h) * corr
            assert np.allclose(got, want, atol=4e-6), (jp, h, got, want)


def test_pbmpm_solve_symmetric_displacement(orc):
    """F_e = I and I + D = Q diag(t) Q^T (symmetric positive definite): the Newton iterates stay
    Q diag(r_n(t)) Q^T, so corr = Q diag(r4(t) (0.95 (t1 t2 t3)^e + 0.05)) Q^T - I and
    D' = D + relax (corr - D)."""
    rng = np.random.default_rng(21)
    for _ in range(40):
        q = rotation(rng)
        t = rng.uniform(0.85, 1.2, 3)
        d = f32(q @ np.diag(t) @ q.T - np.eye(3))
        w, vecs = np.linalg.eigh(np.eye(3) + d)
        jp = rng.uniform(0.5, 40.0)
        got = orc.solve_constraints(np.eye(3), np.diag([1.0, jp, 1.0]), d)
        corr = vecs @ np.diag(newton4(w) * (BLEND * np.prod(w) ** EXPO + (1 - BLEND))) @ vecs.T - np.eye(3)
        want = d + relax(jp) * (corr - d)
        assert np.allclose(got, want, atol=4e-6), (got, want)


def test_pbmpm_solve_general_against_svd_polar(orc):
    """Random F_e and D near the identity (the 4 Newton steps have converged to FP32 there):
    trial = (I + D) F_e, R = U V^T of its SVD, corr = (0.95 R |det trial|^e + 0.05 R) F_e^-1 - I.
    The order of the products matters here: (I + D) F_e vs F_e (I + D) and corr's F_e^-1 on the
    right both differ by far more than the tolerance."""
    rng = np.random.default_rng(22)
    worst = 0.0
    for _ in range(200):
        fe = f32(np.eye(3) + 0.12 * rng.standard_normal((3, 3)))
        d = f32(0.06 * rng.standard_normal((3, 3)))
        jp = rng.uniform(0.6, 30.0)
        got = orc.solve_constraints(fe, np.diag([1.0, 1.0, jp]), d)
        trial = (np.eye(3) + d) @ fe
        u, sv, vt = np.linalg.svd(trial)
        if sv.min() < 0.7 or sv.max() > 1.35 or np.linalg.det(trial) <= 0:
            continue
        r = u @ vt
        corr = (BLEND * r * abs(np.linalg.det(trial)) ** EXPO + (1 - BLEND) * r) @ np.linalg.inv(fe) - np.eye(3)
        want = d + relax(jp) * (corr - d)
        worst = max(worst, np.abs(got - want).max())
        # the product order matters by far more than the tolerance
        trial_swapped = fe @ (np.eye(3) + d)
        assert np.abs(trial_swapped - trial).max() > 1e-3
    assert worst < 2e-5, worst


# ---------------------------------------------------------------- explicit stress (stressTensorOps.wgsl)
def lame(E=1.4e5, nu=0.2):
    return np.float32(E / (2 * (1 + nu))), np.float32(E * nu / ((1 + nu) * (1 - 2 * nu)))


def pk1_fp64(F, Jp, xi, E=1.4e5, nu=0.2):
    """Stomakhin et al. 2013 snow stress: P = 2 mu (F - R) + lambda (J - 1) J F^-T with
    mu, lambda = mu0, lambda0 * exp(xi (1 - J_p)); R the polar factor of F (numpy SVD)."""
    mu0, la0 = lame(E, nu)
    e = np.exp(xi * (1.0 - Jp))
    u, s, vt = np.linalg.svd(F)
    R = u @ vt
    J = np.linalg.det(F)
    return 2 * mu0 * e * (F - R) + la0 * e * (J - 1) * J * np.linalg.inv(F).T


def test_explicit_pk1_diagonal_closed_form(orc):
    """F = diag(s), F_p = I: P = diag(2 mu (s - r4(s)) + lambda (J - 1) J / s)."""
    mu, la = lame()
    for s in ([1.1, 1.0, 1.0], [0.9, 1.05, 1.2], [0.7, 0.8, 1.4]):
        s = np.asarray(np.float32(s), np.float64)
        got = orc.pk1_explicit(np.diag(s), np.eye(3))
        J = np.prod(s)
        want = np.diag(2 * mu * (s - newton4(s)) + la * (J - 1) * J / s)
        assert np.allclose(got, want, rtol=1e-5, atol=1e-5 * np.abs(want).max())


@pytest.mark.parametrize("xi", [5.0, 10.0])
def test_explicit_pk1_hardened_against_fp64(orc, xi):
    rng = np.random.default_rng(23)
    for _ in range(200):
        q = rotation(rng)
        u2 = rotation(rng)
        F = f32(q @ np.diag(rng.uniform(0.85, 1.2, 3)) @ u2.T)
        jp = float(np.float32(rng.uniform(0.85, 1.15)))
        Fp = np.diag([jp, 1.0, 1.0])
        got = orc.pk1_explicit(F, Fp, hardening_explicit=xi)
        want = pk1_fp64(F, jp, xi)
        assert np.abs(got - want).max() <= 3e-5 * np.abs(want).max() + 0.05, (got, want)
        # P is not symmetric for a general F: a transposed P fails the P2G moment test below
    assert np.abs(want - want.T).max() > 10.0


def test_explicit_p2g_moment_is_minus_V_dt_PT(orc):
    """One particle at rest: every node takes -V P^T grad(w_i) dt (particleToGrid.cs.wgsl:61-75),
    and sum_i grad(w_i) (x_i - x_c)^T = I for the quadratic B-spline, so the first moment of the
    deposited momentum is -V dt P^T; the momenta sum to 0 and the masses to m."""
    from helpers import particles_in_spheres
    rng = np.random.default_rng(24)
    S = float(2 ** 20)
    dt = np.float32(1e-3)
    h = 1.0 / 64
    for trial in range(6):
        p = particles_in_spheres(1, rng, vx=0.0)
        p["pos"][0] = rng.uniform(0.3, 0.7, 3)
        p["vel"][0] = 0.0
        F = f32(rotation(rng) @ np.diag(rng.uniform(0.85, 1.2, 3)) @ rotation(rng).T)
        jp = float(np.float32(rng.uniform(0.85, 1.15)))
        p["Fe"][0, :, :3] = F.T            # column-major storage
        p["Fp"][0, :, :3] = np.diag([1.0, jp, 1.0])
        uni = uniforms_blob(dt, (0, 0, 0), (1, 1, 1), (64, 64, 64), S=S)
        o = orc.OracleSim(p, uni, gravity=(0.0, 0.0, 0.0), hardening_explicit=10.0)
        o.run_stage(0, 1)
        xyz, vals = o.dump_grid()
        c0 = np.floor(p["pos"][0].astype(np.float64) / h)
        off = (xyz - c0) * h
        mom = vals[:, 1:] / S
        mass = vals[:, 0] / S
        moment = mom.T @ off                       # sum_i mom_i (x_i - x_c)^T
        V = 1.0 / 400.0
        want = -V * float(dt) * pk1_fp64(F, jp, 10.0).T
        assert np.abs(mom.sum(0)).max() <= 27 * 3 / S
        assert abs(mass.sum() - 1.0) <= 27 / S
        assert np.abs(moment - want).max() <= 3e-4 * np.abs(want).max() + 27 * 3 * h / S, (trial, moment, want)


# ---------------------------------------------------------------- walls (integrateParticles.wgsl:33-64)
def test_walls_clamp_and_halve(orc):
    """A particle whose x + dx leaves [gridMin, gridMax) on an axis ends on that wall with dx and v
    of that axis multiplied by -1/2; the other axes and particles inside are untouched.
    F_e = (I + D) F_e with D on the left."""
    import __graft_entry__
    m = __graft_entry__.load_package()
    rng = np.random.default_rng(25)
    n = 7
    pos = np.full((n, 3), 0.5, np.float32)
    dxs = np.zeros((n, 3), np.float32)
    vel = np.zeros((n, 3), np.float32)
    lo, hi = np.array([0.0, -0.5, 0.25], np.float32), np.array([1.0, 0.5, 2.25], np.float32)
    cases = []
    for a in range(3):
        for side in (0, 1):
            i = 2 * a + side
            pos[i] = (lo + hi) / 2
            wall = lo[a] if side == 0 else hi[a]
            pos[i, a] = wall + (0.01 if side == 0 else -0.01)
            dxs[i] = rng.uniform(-0.003, 0.003, 3)
            dxs[i, a] = -0.05 if side == 0 else 0.05
            vel[i] = rng.uniform(-2, 2, 3)
            cases.append((i, a, wall))
    pos[6] = (lo + hi) / 2
    dxs[6] = [0.01, -0.02, 0.03]
    vel[6] = [1.0, 2.0, 3.0]
    p = m.make_particles(pos, vel=vel)
    p["dx"][:, :3] = dxs
    q = rotation(rng)
    d = f32(1e-3 * rng.standard_normal((3, 3)))
    p["Fe"][6, :, :3] = f32(q).T
    p["D"][6, :, :3] = d.T
    uni = uniforms_blob(np.float32(1e-3), lo, hi, (16, 16, 16))
    o = orc.OracleSim(p, uni)
    o.integrate_only()
    r = o.particles()
    for i, a, wall in cases:
        want_pos = (pos[i] + dxs[i]).astype(np.float64)
        want_pos[a] = wall
        assert np.allclose(r["pos"][i], want_pos, atol=1e-7), (i, r["pos"][i], want_pos)
        want_dx = dxs[i].astype(np.float64).copy()
        want_dx[a] *= -0.5
        want_v = vel[i].astype(np.float64).copy()
        want_v[a] *= -0.5
        assert np.array_equal(r["dx"][i, :3], want_dx.astype(np.float32))
        assert np.array_equal(r["vel"][i], want_v.astype(np.float32))
    assert np.allclose(r["pos"][6], pos[6] + dxs[6], atol=1e-7)
    assert np.array_equal(r["vel"][6], vel[6]) and np.array_equal(r["dx"][6, :3], dxs[6])
    fe = r["Fe"][6, :, :3].T.astype(np.float64)
    assert np.allclose(fe, (np.eye(3) + d) @ f32(q), atol=2e-7)
    assert np.abs(f32(q) @ (np.eye(3) + d) - (np.eye(3) + d) @ f32(q)).max() > 1e-5


# ---------------------------------------------------------------- collider response (colliderPrelude.wgsl:293-356)
def one_triangle_collider(y=0.5):
    """One large triangle in the plane y = `y` over the domain, as packed collider words
    (indices, then xyz vertices) and one 32-byte leaf node; face normal (0, -1, 0)."""
    verts = np.array([[-1.0, y, -1.0], [3.0, y, -1.0], [-1.0, y, 3.0]], np.float32)
    words = np.concatenate([np.array([0, 1, 2], np.uint32), verts.ravel().view(np.uint32)])
    mn, mx = verts.min(0), verts.max(0)
    node = np.concatenate([mn.view(np.uint32), [0], mx.view(np.uint32), [np.uint32(0x80000001)]]).astype(np.uint32)
    return words, node, mn, mx


def run_collider(orc, pos, dx, coll_vel=(0.0, 0.0, 0.0), dt=1e-3, T=None):
    import __graft_entry__
    m = __graft_entry__.load_package()
    words, node, mn, mx = one_triangle_collider()
    p = m.make_particles(np.asarray(pos, np.float32))
    p["dx"][:, :3] = dx
    p["vel"] = np.asarray(dx, np.float32) / np.float32(dt)
    uni = uniforms_blob(np.float32(dt), (0, 0, 0), (1, 1, 1), (32, 32, 32), collider=(mn, mx, 3, T))
    uni[320:332] = np.asarray(coll_vel, np.float32).view(np.uint8)
    o = orc.OracleSim(p, uni)
    o.set_collider(words, node)
    o.integrate_only()
    return p, o.particles()


def test_collider_ccd_snap_and_velocity(orc):
    """A particle crossing the plane y = 0.5 in one step: hit = prev + t dx with t from the plane,
    hit normal = face normal oriented along the ray, position = hit + 0.05 n. Static collider:
    v_rel . n > 0, so v = v_rel (+ 0). Collider moving against the ray (v_c . n < dx/dt . n):
    v = 0.9 v_t + 0.2 v_c / dt."""
    dt = 1e-3
    prev = np.array([[0.4, 0.52, 0.6], [0.3, 0.47, 0.7], [0.6, 0.515, 0.2]], np.float32)
    dx = np.array([[0.01, -0.04, 0.005], [-0.002, 0.05, 0.001], [0.003, -0.03, -0.004]], np.float32)
    for cv in ((0.0, 0.0, 0.0), (0.5, 0.0, -0.25)):
        p, r = run_collider(orc, prev, dx, coll_vel=cv, dt=dt)
        for i in range(2):
            t = (0.5 - float(prev[i, 1])) / float(dx[i, 1])
            hit = prev[i].astype(np.float64) + t * dx[i].astype(np.float64)
            n = np.array([0.0, np.sign(dx[i, 1]), 0.0])
            assert np.allclose(r["pos"][i], hit + 0.05 * n, atol=2e-6), (i, r["pos"][i], hit + 0.05 * n)
            vrel = dx[i].astype(np.float64) / dt - np.asarray(cv)
            vn = vrel @ n
            assert vn > 0
            want = vrel + np.asarray(cv) * 0.2 / dt
            assert np.allclose(r["vel"][i], want, rtol=1e-5, atol=1e-4), (i, r["vel"][i], want)
            assert np.allclose(r["dx"][i, :3], want * dt, rtol=1e-5, atol=1e-7)
    # the collider moving with the ray faster than the particle: v_rel . n < 0 -> friction branch
    cv = np.array([0.2, -80.0, 0.1])
    p, r = run_collider(orc, prev[2:3], dx[2:3], coll_vel=cv, dt=dt)
    t = (0.5 - float(prev[2, 1])) / float(dx[2, 1])
    hit = prev[2].astype(np.float64) + t * dx[2].astype(np.float64)
    n = np.array([0.0, -1.0, 0.0])
    vrel = dx[2].astype(np.float64) / dt - cv
    vn = vrel @ n
    assert vn < 0
    vt = vrel - vn * n
    want = vt * (1 - 0.1) + cv * 0.2 / dt
    assert np.allclose(r["pos"][0], hit + 0.05 * n, atol=2e-6)
    assert np.allclose(r["vel"][0], want, rtol=1e-5, atol=1e-3), (r["vel"][0], want)


def test_collider_proximity_push_out(orc):
    """No crossing and |dx| <= 1e-4 (no CCD): a particle within 0.05 of the plane is moved to
    0.05 from its closest point on the side it is on; approaching it loses its normal velocity
    (friction 0), leaving it keeps v_rel; one farther than 0.05 is untouched."""
    dt = 1e-3
    prev = np.array([[0.4, 0.53, 0.6], [0.3, 0.48, 0.7], [0.6, 0.52, 0.2], [0.5, 0.58, 0.5]], np.float32)
    dx = np.array([[2e-5, -5e-5, 1e-5], [-3e-5, -4e-5, 2e-5], [1e-5, 5e-5, -2e-5], [1e-5, -5e-5, 1e-5]], np.float32)
    p, r = run_collider(orc, prev, dx, dt=dt)
    for i in range(3):
        cur = prev[i].astype(np.float64) + dx[i]
        side = np.sign(cur[1] - 0.5)
        n = np.array([0.0, side, 0.0])
        cp = np.array([cur[0], 0.5, cur[2]])
        assert np.allclose(r["pos"][i], cp + 0.05 * n, atol=2e-6), (i, r["pos"][i])
        vrel = dx[i].astype(np.float64) / dt
        vn = vrel @ n
        want = vrel - vn * n if vn < 0 else vrel
        assert np.allclose(r["vel"][i], want, atol=1e-5), (i, r["vel"][i], want)
    assert np.allclose(r["pos"][3], prev[3] + dx[3], atol=1e-7)
    assert np.array_equal(r["vel"][3], p["vel"][3])


def test_collider_transform_moves_the_plane(orc):
    """With colliderTransformMat translating by (0, 0.1, 0) the local plane y = 0.5 sits at y = 0.6
    in world space: the same closed form applies about 0.6."""
    T = np.eye(4)
    T[1, 3] = 0.1
    dt = 1e-3
    prev = np.array([[0.4, 0.62, 0.6]], np.float32)
    dx = np.array([[0.01, -0.04, 0.005]], np.float32)
    p, r = run_collider(orc, prev, dx, dt=dt, T=T)
    t = (0.6 - float(prev[0, 1])) / float(dx[0, 1])
    hit = prev[0].astype(np.float64) + t * dx[0].astype(np.float64)
    assert np.allclose(r["pos"][0], hit + 0.05 * np.array([0, -1.0, 0]), atol=3e-6)


# ---------------------------------------------------------------- G2P gradients (gridToParticle.cs.wgsl)
@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("grad_scale", [1.0, 2.0])
def test_g2p_reproduces_affine_grid_field(orc, method, grad_scale):
    """Grid velocity v_i = a + M c_i (c_i the node's integer cell coordinates): the quadratic B-spline
    reproduces it, so v = a + M (x_p / h - 1/2), dx = v dt; the explicit gradient in cell units
    sum_i v_i (dw_i/df)^T = M, and the PBMPM D = 4 sum_i w_i v_i (c_i - x_p)^T = M (second moment
    1/4). Both give D = M dt grad_scale. M is not symmetric, so a transposed product fails."""
    from helpers import particles_in_spheres
    rng = np.random.default_rng(26 + method)
    p = particles_in_spheres(400, rng, vx=0.0)
    h = 1.0 / 32
    dt = np.float32(1e-3)
    a = np.array([0.3, -0.2, 0.1])
    M = np.array([[0.010, 0.004, -0.006], [-0.003, 0.008, 0.002], [0.005, -0.007, -0.012]])
    uni = uniforms_blob(dt, (0, 0, 0), (1, 1, 1), (32, 32, 32))
    o = orc.OracleSim(p, uni, grad_scale=grad_scale)
    o.g2p_affine(method, a, M)
    r = o.particles()
    xp = p["pos"].astype(np.float64) / h - 0.5
    v = a[None] + xp @ M.T
    assert np.allclose(r["vel"], v, atol=2e-6), np.abs(r["vel"] - v).max()
    assert np.allclose(r["dx"][:, :3], v * float(dt), atol=2e-9)
    D = r["D"][:, :, :3].transpose(0, 2, 1).astype(np.float64) / (float(dt) * grad_scale)
    err = np.abs(D - M[None]).max()
    assert err < 2e-5, err
    assert np.abs(M - M.T).max() > 1e-3
