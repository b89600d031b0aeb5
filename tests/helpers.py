"""Shared helpers for the parity tests: building matched product / oracle simulations and
comparing their outputs keyed by cell coordinate (block numbering is not deterministic)."""
import numpy as np

# creation parameters shared by both sides (they share the same uniform bytes as well)
SHARED_KEYS = ("pbmpm_iterations", "pbmpm_force_divisor", "gravity", "youngs_modulus", "poissons_ratio",
               "hardening_explicit", "hardening_pbmpm", "pbmpm_blend", "pbmpm_volume_exponent", "theta_c",
               "theta_s", "density", "grad_scale", "collider_margin", "friction_ccd", "friction_proximity",
               "collider_velocity_factor", "wall_mode", "wall_cells", "plasticity_mode")


def make_pair(mpm, orc, sc, fp64=False, with_collider=True, **sim_over):
    """Product MpmSimulation + OracleSim on the same particles, uniforms and collider."""
    p = sc["particles"]
    kw = dict(sc["sim"])
    kw.update(sim_over)
    sim = mpm.MpmSimulation(p.shape[0], **kw)
    if with_collider and sc.get("collider") is not None:
        v, idx = sc["collider"]
        sim.set_collider(v, idx)
    sim.upload(p)
    okw = {k: v for k, v in kw.items() if k in SHARED_KEYS}
    o = orc.OracleSim(p, sim.uniform_bytes(), fp64=fp64, hash_map_size=200_003, max_blocks=70_000, **okw)
    if with_collider and sc.get("collider") is not None:
        o.set_collider(sim.colliderManager.colliderData, sim.colliderManager.bvhNodes)
    return sim, o


def keyed(xyz, vals):
    order = np.lexsort((xyz[:, 0], xyz[:, 1], xyz[:, 2]))
    return xyz[order], vals[order]


def assert_grid_equal(sim, o, what=""):
    gx, gv = keyed(*sim.dump_grid())
    ox, ov = keyed(*o.dump_grid())
    assert gx.shape == ox.shape, f"{what}: {gx.shape[0]} vs {ox.shape[0]} allocated cells"
    assert np.array_equal(gx, ox), f"{what}: allocated cell sets differ"
    ov_i = ov.astype(np.int64).astype(np.int32)
    bad = np.nonzero((gv != ov_i).any(axis=1))[0]
    assert bad.size == 0, (f"{what}: {bad.size} cells differ, first at {gx[bad[0]]}: "
                           f"gpu {gv[bad[0]]} oracle {ov_i[bad[0]]}")
    return gv


def assert_particles_equal(mpm, got, want, what=""):
    a = mpm.state_columns(got).astype(np.float32)
    b = mpm.state_columns(want).astype(np.float32)
    ab, bb = a.view(np.uint32), b.view(np.uint32)
    bad = np.argwhere(ab != bb)
    assert bad.size == 0, (f"{what}: {bad.shape[0]} particle words differ; first particle {bad[0][0]} "
                           f"column {bad[0][1]}: gpu {a[bad[0][0], bad[0][1]]!r} oracle {b[bad[0][0], bad[0][1]]!r}")


def uniforms_blob(dt, grid_min, grid_max, res, S=65536.0, collider=None, interaction=None):
    """The 37,376-byte uniform buffer as GpuUniformsBufferManager.ts:54-182 lays it out
    (oracle-only CPU tests build it without the product library)."""
    b = np.zeros(37376, np.uint8)

    def f(off, vals):
        a = np.asarray(vals, np.float32).ravel()
        b[off:off + a.nbytes] = a.view(np.uint8)

    def u(off, vals):
        a = np.asarray(vals, np.uint32).ravel()
        b[off:off + a.nbytes] = a.view(np.uint8)

    f(0, [dt]); f(4, [S]); f(16, grid_min); f(32, grid_max); u(208, res)
    cd = [(float(grid_max[a]) - float(grid_min[a])) / res[a] for a in range(3)]
    f(352, cd)
    eye = np.eye(4, dtype=np.float32).T.ravel()
    f(256, eye); f(432, eye)
    if collider is not None:
        mn, mx, n_idx, T = collider
        f(224, mn); f(240, mx); u(348, [n_idx]); u(364, [1])
        if T is not None:
            f(256, np.asarray(T, np.float32).T.ravel()); f(432, np.linalg.inv(np.asarray(T, np.float64)).T.ravel())
    if interaction is not None:
        pos, strength, radius, mode = interaction
        f(496, pos); f(508, [strength]); f(512, [radius]); u(516, [1]); u(520, [mode])
    return b


def particles_in_spheres(n, rng, vx=2.0):
    """n particles (unit mass, F = I) uniformly in the two r = 0.15 spheres at (0.3,0.5,0.5) and
    (0.7,0.5,0.5) of [0,1]^3, moving +vx / -vx along x."""
    import __graft_entry__
    m = __graft_entry__.load_package()
    pts = []
    for c in ((0.3, 0.5, 0.5), (0.7, 0.5, 0.5)):
        q = np.empty((0, 3))
        k = n // 2 if c[0] < 0.5 else n - n // 2
        while q.shape[0] < k:
            r = rng.uniform(-0.15, 0.15, (4 * k + 8, 3))
            q = np.concatenate([q, r[(r * r).sum(1) < 0.15 ** 2]])
        pts.append(q[:k] + np.asarray(c))
    vel = np.zeros((n, 3), np.float32)
    vel[: n // 2, 0] = vx
    vel[n // 2:, 0] = -vx
    return m.make_particles(np.concatenate(pts).astype(np.float32), vel=vel)


def particles_from_state(mpm, s):
    """192-byte records from 37 FP32 state words per particle (pos vel mass F_e F_p dx D,
    matrices column-major), the layout of state_columns."""
    s = np.asarray(s, np.float32)
    n = s.shape[0]
    p = mpm.make_particles(s[:, 0:3], vel=s[:, 3:6], mass=s[:, 6])
    p["Fe"][:, :, :3] = s[:, 7:16].reshape(n, 3, 3)
    p["Fp"][:, :, :3] = s[:, 16:25].reshape(n, 3, 3)
    p["dx"][:, :3] = s[:, 25:28]
    p["D"][:, :, :3] = s[:, 28:37].reshape(n, 3, 3)
    return p


def load_golden(name):
    """A whole-scene fixture of tests/golden/ as a dict (meta decoded from JSON)."""
    import json
    import pathlib
    z = np.load(pathlib.Path(__file__).resolve().parent / "golden" / f"{name}.npz")
    d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(bytes(d["meta"]).decode())
    return d


def golden_digests():
    import json
    import pathlib
    return json.loads((pathlib.Path(__file__).resolve().parent / "golden" / "digests.json").read_text())


GOLDEN_WHOLE = ("explicit_cube_4096", "pbmpm_spheres_4096")
