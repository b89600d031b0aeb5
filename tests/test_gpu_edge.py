"""GPU parity at full size and on the edge cases the reference handles silently."""
import numpy as np
import pytest

from helpers import make_pair, assert_grid_equal, assert_particles_equal

pytestmark = pytest.mark.gpu


def test_full_pbmpm_scene_one_substep_bitexact(mpm, orc):
    """The 1,048,576-particle two-sphere scene (k = 4, box collider): one substep, every field, and
    two more, binned from the keys the LAST and collider passes made."""
    sc = mpm.scenes.scene("pbmpm_two_spheres")
    sim, o = make_pair(mpm, orc, sc)
    sim.step(1, 1)
    o.step(1, 1)
    assert_particles_equal(mpm, sim.download(), o.particles(), "1M PBMPM substep")
    sim.step(1, 2)
    o.step(1, 2)
    assert_particles_equal(mpm, sim.download(), o.particles(), "1M PBMPM substeps 2-3")
    st = sim.stats()
    assert st["particles_unbinned"] == 0 and st["blocks_dropped"] == 0 and st["probe_failures"] == 0
    sim.close()


def test_full_pbmpm_scene_is_deterministic(mpm):
    """Block numbering and particle order inside a block come from atomics, yet two contexts
    produce identical particle records after 30 substeps (integer grid sums are order-free)."""
    sc = mpm.scenes.scene("pbmpm_two_spheres")
    outs = []
    for _ in range(2):
        with mpm.MpmSimulation(sc["particles"].shape[0], **sc["sim"]) as sim:
            sim.set_collider(*sc["collider"])
            sim.upload(sc["particles"])
            sim.step(1, 30)
            outs.append(sim.download())
    assert np.array_equal(outs[0].view(np.uint8), outs[1].view(np.uint8))
    assert np.isfinite(mpm.state_columns(outs[0])).all()


def test_explicit_small_scene_substeps_bitexact(mpm, orc):
    """The 65,536-particle explicit scene with sticky walls, 12 substeps."""
    sc = mpm.scenes.scene("explicit_small")
    sim, o = make_pair(mpm, orc, sc)
    sim.step(0, 12)
    o.step(0, 12)
    assert_particles_equal(mpm, sim.download(), o.particles(), "explicit 65k x12")
    sim.close()


def _scene(mpm, pos, vel, **sim):
    p = mpm.make_particles(np.asarray(pos, np.float32), vel=np.asarray(vel, np.float32))
    return dict(particles=p, sim=sim, collider=None)


def _random_deformations(rng, n, lo, hi):
    """n FP32 3x3 matrices U diag(s) V^T with s uniform in [lo, hi] (column-major record layout)."""
    u, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    v, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    s = rng.uniform(lo, hi, (n, 3))
    m = np.einsum("nij,nj,nkj->nik", u, s, v).astype(np.float32)
    return np.swapaxes(m, 1, 2)


@pytest.mark.parametrize("collider", [False, True])
def test_plasticity_clamps_bitexact(mpm, orc, collider):
    """Particles whose F_e lies outside the plastic band [1 - theta_c, 1 + theta_s] and whose F_p
    is not the identity, some of them outside the grid: every substep clamps, so F_p (kept per
    particle id) and det F_p (which the relax factor and the hardening read) change while the
    particles are re-sorted. Both methods, with and without the collider (whose launch then
    integrates the unbinned particles)."""
    rng = np.random.default_rng(31)
    n = 4000
    pos = rng.uniform(0.1, 0.9, (n, 3)).astype(np.float32)
    pos[:40, 1] = -0.1                     # unbinned: integrated outside the transfer passes
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    for method in (1, 0):
        p = mpm.make_particles(pos, vel=vel)
        p["Fe"][:, :, :3] = _random_deformations(rng, n, 0.9, 1.1)
        p["Fp"][:, :, :3] = _random_deformations(rng, n, 0.95, 1.05)
        sc = dict(particles=p, sim=dict(grid_res=(32, 32, 32), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1.0 / 600,
                                        gravity=(0.0, -9.8, 0.0), pbmpm_iterations=3, pbmpm_force_divisor=3.0),
                  collider=mpm.scenes.box_mesh((0.5, 0.5, 0.5), 0.2) if collider else None)
        if method == 0:
            sc["sim"]["dt"] = 1e-4
        sim, o = make_pair(mpm, orc, sc)
        fp0 = p["Fp"].copy()
        for s in range(3):
            sim.step(method, 1)
            o.step(method, 1)
            out = sim.download()
            assert_particles_equal(mpm, out, o.particles(), f"method {method} collider {collider} substep {s}")
            if s == 0:      # the walls clamp them into the grid for the later substeps
                assert sim.stats()["particles_unbinned"] == 40
        changed = np.any(out["Fp"].reshape(n, -1) != fp0.reshape(n, -1), axis=1)
        assert changed.mean() > 0.5, changed.mean()
        sim.close()


def test_plastic_band_edges_bitexact(mpm, orc):
    """F_e with singular values on both sides of the band's edges and of the inner edges where the
    passes skip the polar-decomposition check (lo + 5e-4, hi - 5e-4), arbitrary rotations and
    reflections, some axis-aligned: both methods, bit for bit against the oracle, which always
    runs the full check."""
    rng = np.random.default_rng(33)
    n = 8000
    lo, hi = 1.0 - 2.5e-2, 1.0 + 7.5e-3
    centres = np.array([lo, lo + 5e-4, hi - 5e-4, hi, 1.0])
    sv = centres[rng.integers(0, centres.size, (n, 3))] + rng.uniform(-3e-4, 3e-4, (n, 3))
    q1, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q2, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q2[1::3] = np.eye(3)
    fe = np.einsum("nij,nj,nkj->nik", q1, sv, q2).astype(np.float32)
    pos = rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32)
    for method in (1, 0):
        p = mpm.make_particles(pos, vel=rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32))
        p["Fe"][:, :, :3] = np.swapaxes(fe, 1, 2)
        sc = dict(particles=p, sim=dict(grid_res=(32, 32, 32), grid_min=(0, 0, 0), grid_max=(1, 1, 1),
                                        dt=1.0 / 600 if method == 1 else 1e-4, gravity=(0.0, -9.8, 0.0),
                                        pbmpm_iterations=3, pbmpm_force_divisor=3.0), collider=None)
        sim, o = make_pair(mpm, orc, sc)
        for s in range(2):
            sim.step(method, 1)
            o.step(method, 1)
            assert_particles_equal(mpm, sim.download(), o.particles(), f"method {method} substep {s}")
        sim.close()


def test_particles_on_and_outside_the_walls(mpm, orc):
    """Particles at exactly gridMax (cell == res, finding 13), outside the grid and a
    non-multiple-of-4 resolution: binned / unbinned sets, grids and states match the oracle."""
    rng = np.random.default_rng(20)
    n = 3000
    pos = rng.uniform(0.02, 0.98, (n, 3)).astype(np.float32)
    pos[:50, 0] = 1.0                       # exactly on the max wall
    pos[50:80, 1] = -0.01                   # below the min wall
    pos[80:100, 2] = 1.05                   # above the max wall
    vel = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    for res in (30, 32):
        sc = _scene(mpm, pos, vel, grid_res=(res, res, res), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1.0 / 600,
                    gravity=(0.0, -9.8, 0.0), pbmpm_iterations=3, pbmpm_force_divisor=3.0)
        for method in (1, 0):
            sim, o = make_pair(mpm, orc, sc)
            sim.debug_stage(method, 1)
            o.run_stage(method, 1)
            assert_grid_equal(sim, o, f"res {res} method {method} P2G")
            sim.close()
            sim, o = make_pair(mpm, orc, sc)
            for s in range(3):
                sim.step(method, 1)
                o.step(method, 1)
                assert_particles_equal(mpm, sim.download(), o.particles(), f"res {res} method {method} substep {s}")
            assert sim.stats()["particles_unbinned"] == o.stats()["particles_unbinned"]
            sim.close()


@pytest.mark.parametrize("table", [0, 4093])
def test_block_capacity_drops_are_counted(mpm, table):
    """More stencil blocks than max_blocks: the overflow is dropped and counted (mapAffectedBlocks
    .wgsl:27-31) and the step still completes with finite state (direct block addressing and the
    open-addressing table)."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=20000, seed=21)
    with mpm.MpmSimulation(20000, **dict(sc["sim"], max_blocks=200, hash_map_size=table)) as sim:
        sim.upload(sc["particles"])
        sim.step(1, 2)
        st = sim.stats()
        assert st["blocks_allocated"] == 200 and st["blocks_dropped"] > 0
        assert np.isfinite(mpm.state_columns(sim.download())).all()


def test_pointer_interaction_bitexact(mpm, orc):
    """updateGrid's repel / attract force (pbmpm.wgsl:104-134) with the pointer inside the cloud."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=6000, seed=22)
    for mode in (0, 1):
        sim, o = make_pair(mpm, orc, sc, with_collider=False)
        um = sim.uniformsManager
        um.writeInteractionPos([38.0, 64.0, 64.0])        # grid-cell units
        um.writeInteractionStrength(30.0)
        um.writeInteractionRadius(12.0)
        um.writeIsInteracting(True)
        um.writeInteractionMode(mode)
        o.set_uniforms(sim.uniform_bytes())
        for s in range(3):
            sim.step(1, 1)
            o.step(1, 1)
            assert_particles_equal(mpm, sim.download(), o.particles(), f"interaction mode {mode} substep {s}")
        sim.close()


def test_moved_and_scaled_collider_bitexact(mpm, orc):
    """Collider transform and velocity uniforms (updateColliderTransformMat / updateColliderVel,
    GpuSnowPipelineRunner.svelte.ts:383-390): unit box scaled 0.25 and translated."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=8000, seed=23)
    v, idx = mpm.scenes.box_mesh((0.0, 0.0, 0.0), 1.0)
    T = np.eye(4)
    T[:3, :3] *= 0.25
    T[:3, 3] = (0.5, 0.48, 0.5)
    sim, o = make_pair(mpm, orc, dict(sc, collider=None))
    cm = sim.set_collider(v, idx, transform=T)
    sim.updateColliderVel([0.0, 0.5, 0.0])
    o.set_uniforms(sim.uniform_bytes())
    o.set_collider(cm.colliderData, cm.bvhNodes)
    for s in range(4):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"moved collider substep {s}")
    sim.close()


def _chain_bvh(n_tris, lo, hi):
    """Right-leaning chain: internal node i has the leaf of triangle i on its left and internal
    node i + 1 (or the last leaf) on its right; every box is the whole mesh's."""
    m = n_tris - 1
    nodes = np.zeros((m + n_tris, 8), np.uint32)
    nodes.view(np.float32)[:, 0:3] = lo
    nodes.view(np.float32)[:, 4:7] = hi
    for i in range(m):
        nodes[i, 3] = m + i
        nodes[i, 7] = i + 1 if i + 1 < m else m + n_tris - 1
    for j in range(n_tris):
        nodes[m + j, 3] = j
        nodes[m + j, 7] = 0x80000000 | 1
    return nodes


def test_bvh_deeper_than_the_traversal_stack_bitexact(mpm, orc):
    """colliderPrelude.wgsl:215-291 pushes children only while fewer than 31 entries are pending:
    in a chain with more than 31 right turns the deeper triangles are never visited. The device
    walk (threaded records) must skip exactly those; a flat leaf with every triangle gives other
    particles, so the limit is exercised."""
    rng = np.random.default_rng(41)
    v, idx = mpm.scenes.heightfield_mesh(12, 12, 1.0, 0.3, 0.45, rng)
    n = 6000
    pos = np.stack([rng.uniform(0.05, 0.95, n), rng.uniform(0.47, 0.55, n), rng.uniform(0.05, 0.95, n)], 1).astype(np.float32)
    p = mpm.make_particles(pos, vel=(0.0, -2.0, 0.0))
    sc = dict(particles=p, sim=dict(grid_res=(64, 64, 64), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1e-4,
                                    gravity=(0.0, -9.8, 0.0)), collider=None)
    sim, o = make_pair(mpm, orc, sc)
    cm = sim.set_collider(v, idx)
    n_tris = cm.numIndices // 3
    chain = _chain_bvh(n_tris, np.min(v, 0), np.max(v, 0))
    cm.bvhNodes, cm.numBvhNodes = chain, chain.shape[0]
    cm.upload(sim)
    o.set_uniforms(sim.uniform_bytes())
    o.set_collider(cm.colliderData, chain)
    flat = np.zeros((1, 8), np.uint32)
    flat.view(np.float32)[0, 0:3] = np.min(v, 0)
    flat.view(np.float32)[0, 4:7] = np.max(v, 0)
    flat[0, 7] = n_tris | 0x80000000
    of = orc.OracleSim(p, sim.uniform_bytes(), hash_map_size=200_003, max_blocks=70_000, gravity=(0.0, -9.8, 0.0))
    of.set_collider(cm.colliderData, flat)
    for s in range(3):
        sim.step(0, 1)
        o.step(0, 1)
        of.step(0, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"chain BVH substep {s}")
    differ = (mpm.state_columns(o.particles()) != mpm.state_columns(of.particles())).any(1).sum()
    assert differ > 50, differ
    sim.close()


def test_bvh_that_is_not_a_tree_is_rejected(mpm):
    """A node reached twice (shared child or a cycle) would make the traversal revisit or loop:
    mpm_set_collider refuses it and leaves the collider it had in place."""
    v, idx = mpm.scenes.box_mesh((0.5, 0.5, 0.5), 0.2)
    with mpm.MpmSimulation(1000, grid_res=(32, 32, 32), grid_min=(0, 0, 0), grid_max=(1, 1, 1)) as sim:
        cm = sim.set_collider(v, idx)
        lo, hi = np.min(v, 0), np.max(v, 0)
        for shape in ("shared", "cycle"):
            nodes = np.zeros((3, 8), np.uint32)
            nodes.view(np.float32)[:, 0:3] = lo
            nodes.view(np.float32)[:, 4:7] = hi
            nodes[0, 3], nodes[0, 7] = 1, (1 if shape == "shared" else 2)
            nodes[1, 3], nodes[1, 7] = 0, 0x80000000 | 12
            if shape == "shared":
                nodes[2, 3], nodes[2, 7] = 0, 0x80000000 | 12
            else:
                nodes[2, 3], nodes[2, 7] = 1, 0          # internal node whose right child is the root
            rc = sim.lib.mpm_set_collider(sim.ctx, mpm.managers._ptr(cm.colliderData), cm.colliderData.size, cm.numIndices,
                                          mpm.managers._ptr(nodes), nodes.shape[0])
            assert rc == mpm._lib.MPM_E_INVALID, (shape, rc)
            assert "tree" in sim.lib.mpm_last_error(sim.ctx).decode()
