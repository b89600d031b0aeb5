"""The plasticity shortcut of the transfer passes (band_clear in csrc/mpm_kernels.h): a particle whose
F_e has every singular value in [lo + 5e-4, hi - 5e-4] by Gershgorin bounds on F_e^T F_e skips the
reference's polar-decomposition check (plasticityOps.wgsl:7-48), which must then leave F_e and F_p
unchanged. Checked here against the oracle, which always runs the full check: F_e with singular
values on both sides of the shortcut's edges and of the band's own edges, arbitrary rotations and
reflections. (The device restates the same test; tests/test_gpu_edge.py compares it bit for bit.)"""
import numpy as np
import pytest

from helpers import uniforms_blob

MARGIN = 5e-4


def _band_clear(F, lo, hi):
    """Restatement of band_clear on float32 column-major matrices F[n, col, row]."""
    F = F.astype(np.float32)
    blo, bhi = float(lo) + MARGIN, float(hi) - MARGIN
    lo2, hi2 = np.float32(blo * blo), np.float32(bhi * bhi)
    if float(lo2) < blo * blo:
        lo2 = np.nextafter(lo2, np.float32(2))
    if float(hi2) > bhi * bhi:
        hi2 = np.nextafter(hi2, np.float32(0))

    def dot(a, b):
        return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]

    c = [F[:, i, :] for i in range(3)]
    c00, c11, c22 = dot(c[0], c[0]), dot(c[1], c[1]), dot(c[2], c[2])
    c01, c02, c12 = np.abs(dot(c[0], c[1])), np.abs(dot(c[0], c[2])), np.abs(dot(c[1], c[2]))
    r0, r1, r2 = c01 + c02, c01 + c12, c02 + c12
    return ((c00 + r0 <= hi2) & (c11 + r1 <= hi2) & (c22 + r2 <= hi2) &
            (c00 - r0 >= lo2) & (c11 - r1 >= lo2) & (c22 - r2 >= lo2))


def _haar(rng, n):
    q, r = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    return q * np.sign(np.diagonal(r, axis1=1, axis2=2))[:, None, :]


# the default band, and a wide one near the limits of where the host enables the shortcut
@pytest.mark.parametrize("theta_c,theta_s", [(2.5e-2, 7.5e-3), (0.195, 0.245)])
def test_shortcut_never_hides_a_clamp(mpm, orc, theta_c, theta_s):
    rng = np.random.default_rng(41)
    lo, hi = np.float32(1.0 - theta_c), np.float32(1.0 + theta_s)
    n = 200000
    # singular values around the shortcut's edges (lo + m, hi - m) and the band's edges (lo, hi)
    centres = np.array([float(lo), float(lo) + MARGIN, float(hi) - MARGIN, float(hi), 1.0])
    sv = centres[rng.integers(0, centres.size, (n, 3))] + rng.uniform(-3e-4, 3e-4, (n, 3))
    sv[: n // 4] = rng.uniform(float(lo) + MARGIN, float(hi) - MARGIN, (n // 4, 3))   # deep inside
    u, v = _haar(rng, n), _haar(rng, n)
    u[::3, :, 0] *= -1.0                                                             # reflections
    v[1::3] = np.eye(3)      # R^T F = diag(s): the check compares the singular values themselves
    fe = np.einsum("nij,nj,nkj->nik", u, sv, v).astype(np.float32)   # row-major F
    p = mpm.make_particles(rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32))
    p["Fe"][:, :, :3] = np.swapaxes(fe, 1, 2)
    uni = uniforms_blob(np.float32(1e-4), (0, 0, 0), (1, 1, 1), (32, 32, 32))
    # no stress and no gravity: the grid velocity is 0, so D = 0 and only plasticity can change F_e
    o = orc.OracleSim(p, uni, gravity=(0.0, 0.0, 0.0), youngs_modulus=0.0, theta_c=theta_c, theta_s=theta_s)
    o.step(0, 1)
    out = o.particles()
    clamped = np.any(out["Fe"].reshape(n, -1) != p["Fe"].reshape(n, -1), axis=1)
    clear = _band_clear(p["Fe"][:, :, :3], lo, hi)
    assert not np.any(clear & clamped), np.flatnonzero(clear & clamped)[:10]
    assert clear.mean() > 0.1 and clamped.mean() > 0.1 and (~clear & ~clamped).mean() > 0.05
    # the shortcut is decided on the singular values with a margin: near-edge values on the safe
    # side of it take the full check, which leaves them unchanged too
    deep = np.arange(n) < n // 4
    assert clear[deep].mean() > 0.5
