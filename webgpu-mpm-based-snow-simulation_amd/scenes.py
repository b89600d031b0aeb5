"""Seeded synthetic scenes for the workloads the step is measured and tested on.

  explicit_small  65,536 particles, 64^3 grid over [0,1]^3, 0.25-side cube at (0.5,0.6,0.5),
                  v = (0,-1,0), 16 per cell with uniform jitter, dt = 1e-4, sticky 3-cell walls.
  pbmpm_two_spheres  1,048,576 particles in two r = 0.15 spheres at (0.3,0.5,0.5) and
                  (0.7,0.5,0.5) moving +/-2 m/s along x, 128^3 grid, k = 4, dt = 1/600,
                  12-triangle box collider of side 0.2 at the domain centre.
  explicit_slab_terrain  4,194,304 particles in a slab 0.6 x 0.1 x 0.6 at y = 0.7 over a
                  64 x 64-vertex heightfield (7,938 triangles), 256^3 grid, dt = 5e-5, Young's modulus
                  log-uniform in [5e4, 2e5] per particle.
  ensemble_scenes  64 independent scenes for the batched step: N log-uniform in [8,192, 524,288],
                  grid by count (32^3 / 64^3 / 128^3), half explicit, half PBMPM with k in 2..6,
                  sphere or cube emitters with velocity U[-3,3]^3.
Gravity is -9.8 along +y (the scenes are y-up). Masses use one unit per particle so the
2^16 fixed-point grid sums keep their resolution (the reference's random masses in [0,1]
would truncate to zero at these particle spacings).
"""
import numpy as np

from .particles import make_particles


def box_mesh(center, side):
    """12-triangle axis-aligned cube (the geometry of the reference's static/box.glb, scaled)."""
    c = np.asarray(center, np.float64)
    h = 0.5 * side
    v = np.array([[x, y, z] for z in (-h, h) for y in (-h, h) for x in (-h, h)], np.float64) + c
    # vertex index = x + 2y + 4z
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, cc, d in quads:
        tris += [(a, b, cc), (a, cc, d)]
    return v.astype(np.float32), np.asarray(tris, np.uint32).ravel()


def heightfield_mesh(nx, nz, extent, hmin, hmax, rng):
    """Seeded low-poly terrain: nx x nz vertices over [0,extent]^2, heights U[hmin,hmax]."""
    xs = np.linspace(0.0, extent, nx)
    zs = np.linspace(0.0, extent, nz)
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    Y = rng.uniform(hmin, hmax, size=X.shape)
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1).astype(np.float32)
    tris = []
    for j in range(nz - 1):
        for i in range(nx - 1):
            a = j * nx + i
            b, c, d = a + 1, a + nx, a + nx + 1
            tris += [(a, c, b), (b, c, d)]
    return v, np.asarray(tris, np.uint32).ravel()


def jittered_cells(lo_cell, hi_cell, h, ppc, rng):
    """ppc particles per cell, uniform jitter inside each cell of the integer box [lo, hi)."""
    cells = np.stack(np.meshgrid(*[np.arange(lo_cell[a], hi_cell[a]) for a in range(3)], indexing="ij"), -1).reshape(-1, 3)
    base = np.repeat(cells, ppc, axis=0).astype(np.float64)
    return ((base + rng.random(base.shape)) * h).astype(np.float32)


def sphere_points(center, radius, n, rng):
    out = np.empty((0, 3))
    while out.shape[0] < n:
        q = rng.uniform(-radius, radius, size=(2 * n, 3))
        q = q[(q * q).sum(1) < radius * radius]
        out = np.concatenate([out, q])
    return (out[:n] + np.asarray(center)).astype(np.float32)


def scene(name, n_particles=None, seed=None):
    """Returns dict(particles, sim kwargs, collider (vertices, indices) or None, method, substeps)."""
    if name == "explicit_small":
        rng = np.random.default_rng(0 if seed is None else seed)
        res, h = 64, 1.0 / 64
        lo = np.array([0.5 - 0.125, 0.6 - 0.125, 0.5 - 0.125])
        lo_c = np.round(lo / h).astype(int)
        pos = jittered_cells(lo_c, lo_c + 16, h, 16, rng)
        if n_particles is not None:
            pos = pos[rng.permutation(pos.shape[0])[:n_particles]]
        p = make_particles(pos, vel=(0.0, -1.0, 0.0))
        sim = dict(grid_res=(res, res, res), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1e-4,
                   gravity=(0.0, -9.8, 0.0), wall_mode=1, wall_cells=3, hardening_explicit=10.0)
        return dict(particles=p, sim=sim, collider=None, method=0, substeps=100)
    if name == "pbmpm_two_spheres":
        rng = np.random.default_rng(1 if seed is None else seed)
        n = 1_048_576 if n_particles is None else n_particles
        a = sphere_points((0.3, 0.5, 0.5), 0.15, n // 2, rng)
        b = sphere_points((0.7, 0.5, 0.5), 0.15, n - n // 2, rng)
        vel = np.zeros((n, 3), np.float32)
        vel[: n // 2, 0] = 2.0
        vel[n // 2:, 0] = -2.0
        p = make_particles(np.concatenate([a, b]), vel=vel)
        dt = 1.0 / 600
        # PBMPM carries motion in the displacement (P2G deposits m (dx + D off) / dt,
        # pbmpm.wgsl:41-76), so the initial velocity enters as dx = v dt
        p["dx"][:, :3] = vel * np.float32(dt)
        sim = dict(grid_res=(128, 128, 128), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=dt,
                   gravity=(0.0, -9.8, 0.0), pbmpm_iterations=4, pbmpm_force_divisor=4.0)
        return dict(particles=p, sim=sim, collider=box_mesh((0.5, 0.5, 0.5), 0.2), method=1, substeps=600)
    if name == "explicit_slab_terrain":
        rng = np.random.default_rng(2 if seed is None else seed)
        res, h = 256, 1.0 / 256
        n = 4_194_304 if n_particles is None else n_particles
        lo = np.array([0.2, 0.65, 0.2])
        pos = (lo + rng.random((n, 3)) * np.array([0.6, 0.1, 0.6])).astype(np.float32)
        p = make_particles(pos)
        youngs = np.exp(rng.uniform(np.log(5e4), np.log(2e5), n)).astype(np.float32)
        sim = dict(grid_res=(res, res, res), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=5e-5,
                   gravity=(0.0, -9.8, 0.0))
        return dict(particles=p, sim=sim, collider=heightfield_mesh(64, 64, 1.0, 0.05, 0.25, rng), method=0, substeps=2000,
                    youngs=youngs)
    raise KeyError(name)


def ensemble_scenes(n_scenes=64, seed=3, n_min=8192, n_max=524288):
    """Scenes of the batched workload (seed of scene s: 1000 seed + s). Each entry:
    dict(particles, sim = MpmEnsemble scene kwargs including method and n_particles)."""
    out = []
    for s in range(n_scenes):
        rng = np.random.default_rng(1000 * seed + s)
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        res = 32 if n < 32768 else (64 if n < 131072 else 128)
        explicit = s % 2 == 0
        # emitter of about 12 particles per cell
        vol = n * (1.0 / res) ** 3 / 12.0
        centre = rng.uniform(0.35, 0.65, 3)
        if rng.random() < 0.5:
            r = (3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0)
            pos = sphere_points(centre, r, n, rng)
        else:
            side = vol ** (1.0 / 3.0)
            pos = (centre - 0.5 * side + side * rng.random((n, 3))).astype(np.float32)
        v = rng.uniform(-3.0, 3.0, 3).astype(np.float32)
        p = make_particles(pos, vel=v)
        sim = dict(n_particles=n, grid_res=(res, res, res), grid_min=(0, 0, 0), grid_max=(1, 1, 1),
                   gravity=(0.0, -9.8, 0.0))
        if explicit:
            sim.update(method=0, dt=1e-4, wall_mode=1, wall_cells=3, hardening_explicit=10.0)
        else:
            k = int(rng.integers(2, 7))
            dt = 1.0 / 600
            p["dx"][:, :3] = v * np.float32(dt)
            sim.update(method=1, dt=dt, pbmpm_iterations=k, pbmpm_force_divisor=float(k))
        out.append(dict(particles=p, sim=sim))
    return out
