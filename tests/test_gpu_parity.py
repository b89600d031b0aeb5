"""GPU parity: the HIP path through the C ABI against the FP32-faithful oracle.

Bar: bit-exact for the i32 grid accumulators, the allocated block set, the per-block
particle histogram, and (because every kernel restates the reference arithmetic term by
term with the same exp/log/pow) for every FP32 particle field after full substeps.
"""
import ctypes

import numpy as np
import pytest

from helpers import make_pair, assert_grid_equal, assert_particles_equal, keyed

pytestmark = pytest.mark.gpu


def _sel(mpm, kind, a, b=None, out_w=1):
    lib = mpm.load_library()
    a = np.ascontiguousarray(a, np.float32)
    n = a.size // (9 if kind in (6, 7, 8) else 1)
    out = np.zeros(n * out_w, np.float32)
    bb = None if b is None else np.ascontiguousarray(b, np.float32)
    rc = lib.mpm_selftest(kind, a.ctypes.data_as(ctypes.c_void_p),
                          None if bb is None else bb.ctypes.data_as(ctypes.c_void_p),
                          out.ctypes.data_as(ctypes.c_void_p), n)
    assert rc == 0
    return out


def test_device_arithmetic_bitwise(mpm, orc):
    rng = np.random.default_rng(5)
    L = orc.lib()
    x = np.concatenate([rng.uniform(-100, 90, 4000), rng.uniform(-1, 1, 2000), [0.0, -0.0, 88.7, -103.9, 1e-40, 3e38]]).astype(np.float32)
    g = _sel(mpm, 0, x)
    want = orc.f32_vec(L.orc_exp, x)
    assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
    xp = np.abs(x) + np.float32(1e-30)
    assert np.array_equal(_sel(mpm, 1, xp).view(np.uint32), orc.f32_vec(L.orc_log, xp).view(np.uint32))
    y = rng.uniform(-2, 2, x.size).astype(np.float32)
    assert np.array_equal(_sel(mpm, 2, xp, y).view(np.uint32), orc.f32_vec(L.orc_pow, xp, y).view(np.uint32))
    assert np.array_equal(_sel(mpm, 3, xp), np.sqrt(xp))
    d = rng.uniform(0.1, 10, x.size).astype(np.float32)
    assert np.array_equal(_sel(mpm, 4, x, d), x / d)
    f = rng.uniform(-3e9, 3e9, 3000).astype(np.float32)
    f = np.concatenate([f, [np.nan, -0.5, 0.5, 2.5e9, -2.5e9]]).astype(np.float32)
    gi = _sel(mpm, 9, f).view(np.int32)
    wi = np.array([L.orc_f2i(float(v)) for v in f], np.int32)
    assert np.array_equal(gi, wi)


def test_device_hash3(mpm, orc):
    L = orc.lib()
    rng = np.random.default_rng(6)
    a = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 1000, dtype=np.uint64).astype(np.uint32)
    got = _sel(mpm, 5, a.view(np.float32), b.view(np.float32)).view(np.uint32)
    want = np.array([L.orc_hash3(int(a[i]), int(b[i]), i) for i in range(a.size)], np.uint32)
    assert np.array_equal(got, want)


def test_device_matrix_ops(mpm, orc):
    L = orc.lib()
    rng = np.random.default_rng(7)
    m = (np.eye(3)[None] + 0.3 * rng.standard_normal((500, 3, 3))).astype(np.float32).reshape(500, 9)
    det_g = _sel(mpm, 6, m)
    inv_g = _sel(mpm, 7, m, out_w=9).reshape(500, 9)
    pol_g = _sel(mpm, 8, m, out_w=9).reshape(500, 9)
    for i in range(500):
        mi = np.ascontiguousarray(m[i])
        o = np.zeros(9, np.float32)
        assert det_g[i].view(np.uint32) == np.float32(L.orc_det3(mi.ctypes.data_as(ctypes.c_void_p))).view(np.uint32)
        L.orc_inverse3(mi.ctypes.data_as(ctypes.c_void_p), o.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(inv_g[i].view(np.uint32), o.view(np.uint32))
        L.orc_polar3(mi.ctypes.data_as(ctypes.c_void_p), o.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(pol_g[i].view(np.uint32), o.view(np.uint32))


def _small_pbmpm(mpm, n=4096, seed=11):
    return mpm.scenes.scene("pbmpm_two_spheres", n_particles=n, seed=seed)


# hash_map_size 0: blocks addressed directly by key; nonzero: the reference's open-addressing table
@pytest.mark.parametrize("k,table", [(1, 0), (3, 0), (4, 0), (3, 65537)])
def test_pbmpm_stage_grids_bitexact(mpm, orc, k, table):
    sc = _small_pbmpm(mpm)
    for stage in range(0, k + 1):
        sim, o = make_pair(mpm, orc, sc, pbmpm_iterations=k, pbmpm_force_divisor=float(k), hash_map_size=table)
        sim.debug_stage(1, stage)
        o.run_stage(1, stage)
        gb = sim.dump_blocks()
        ob, ocnt = o.active_blocks()
        assert sorted(map(tuple, gb.tolist())) == sorted(map(tuple, ob.tolist())), f"stage {stage}: allocated block sets differ"
        if stage == 0:
            pb, pc = sim.dump_particle_blocks()
            got = {tuple(b): int(c) for b, c in zip(pb.tolist(), pc.tolist())}
            want = {tuple(b): int(c) for b, c in zip(ob.tolist(), ocnt.tolist()) if c}
            assert got == want, "per-block particle histogram differs"
        else:
            assert_grid_equal(sim, o, f"k={k} stage {stage}")
        sim.close()


@pytest.mark.parametrize("table", [0, 65537])
def test_pbmpm_substeps_bitexact(mpm, orc, table):
    sc = _small_pbmpm(mpm, n=8192, seed=12)
    sim, o = make_pair(mpm, orc, sc, hash_map_size=table)
    for s in range(6):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"PBMPM substep {s + 1}")
    sim.close()


@pytest.mark.parametrize("chunk", [256, 2048])
def test_pbmpm_chunk_sizes_bitexact(mpm, orc, chunk):
    """Blocks of ~1,000 particles (16 per cell) in work chunks of 256 (one round each) and of 2,048
    (four or five rounds per chunk: FIRST's cell-interleaved row order cycles through all three of
    its count buffers). The grid sums are integers, so the results match the oracle for any chunking."""
    rng = np.random.default_rng(17)
    n = 8192
    pos = (np.array([0.375, 0.5, 0.375]) + rng.uniform(0.0, 0.25, (n, 3))).astype(np.float32)
    vel = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    sc = dict(particles=mpm.make_particles(pos, vel=vel),
              sim=dict(grid_res=(32, 32, 32), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1.0 / 600,
                       gravity=(0.0, -9.8, 0.0), pbmpm_iterations=4, pbmpm_force_divisor=4.0),
              collider=None)
    sim, o = make_pair(mpm, orc, sc, chunk_particles=chunk)
    for s in range(3):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"chunk {chunk} substep {s + 1}")
    st = sim.stats()
    if chunk == 2048:     # one chunk per particle block, the eight dense ones of ~1,000 particles
        assert st["work_chunks"] == st["particle_blocks"]
    else:
        assert st["work_chunks"] >= 32
    sim.close()


def test_explicit_stages_and_substeps_bitexact(mpm, orc):
    sc = mpm.scenes.scene("explicit_small", n_particles=8192, seed=3)
    for stage in (1, 2):
        sim, o = make_pair(mpm, orc, sc)
        sim.debug_stage(0, stage)
        o.run_stage(0, stage)
        assert_grid_equal(sim, o, f"explicit stage {stage}")
        sim.close()
    sim, o = make_pair(mpm, orc, sc)
    for s in range(5):
        sim.step(0, 1)
        o.step(0, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"explicit substep {s + 1}")
    sim.close()


@pytest.mark.parametrize("case", ["scale_not_pow2", "huge_momentum_lanes"])
def test_explicit_p2g_general_form_bitexact(mpm, orc, case):
    """Explicit P2G folds the fixed-point scale into momentum, mass and dt only for S = 2^k and
    particles with |m|, |m v|, |V P F^T| <= 2^40; everything else takes the general form, per lane.
    Deformed particles (F_e off the identity, so the stress term is live) against the oracle: a
    scale of 10^5 (every particle general), and a few lanes of momentum 2^41 among folded ones."""
    rng = np.random.default_rng(15)
    sc = mpm.scenes.scene("explicit_small", n_particles=6000, seed=15)
    p = sc["particles"].copy()
    n = p.shape[0]
    u, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    sv = rng.uniform(0.97, 1.01, (n, 3))
    p["Fe"][:, :, :3] = np.swapaxes(np.einsum("nij,nj,nkj->nik", u, sv, u).astype(np.float32), 1, 2)
    if case == "huge_momentum_lanes":
        p["vel"][::97, 0] = np.float32(2.0 ** 41)
    sc = dict(sc, particles=p)

    def pair():
        sim, o = make_pair(mpm, orc, sc)
        if case == "scale_not_pow2":
            sim.uniformsManager.writeFixedPointScale(1e5)
            o.set_uniforms(sim.uniform_bytes())
        return sim, o

    sim, o = pair()
    sim.debug_stage(0, 1)
    o.run_stage(0, 1)
    assert_grid_equal(sim, o, f"{case} P2G")
    sim.close()
    sim, o = pair()
    for s in range(3):
        sim.step(0, 1)
        o.step(0, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"{case} substep {s + 1}")
    sim.close()


def test_graph_replay_matches_eager(mpm, orc):
    sc = _small_pbmpm(mpm, n=6000, seed=13)
    a, _ = make_pair(mpm, orc, sc, use_graph=1)
    b, _ = make_pair(mpm, orc, sc, use_graph=0)
    for _ in range(3):
        a.step(1, 4)
        b.step(1, 4)
        x, y = a.download(), b.download()
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    a.close(); b.close()


def test_multi_substep_graph_matches_eager_and_oracle(mpm, orc):
    """Runs of 8 substeps replay one graph holding all of them (both buffer parities captured in
    sequence); remainders take the one-substep graph. With the graphs prepared up front, 19 substeps
    in calls of 11 and 8 equal 19 eager substeps and the oracle; without, a call of 40 captures the
    run graph itself and equals 40 eager substeps. Both methods."""
    for method, sc in ((1, _small_pbmpm(mpm, n=6000, seed=21)), (0, mpm.scenes.scene("explicit_small", n_particles=6000, seed=22))):
        a, o = make_pair(mpm, orc, sc, use_graph=1)
        b, _ = make_pair(mpm, orc, sc, use_graph=0)
        c, _ = make_pair(mpm, orc, sc, use_graph=1)
        a.prepare_graphs(method)
        a.step(method, 11)
        a.step(method, 8)
        b.step(method, 19)
        o.step(method, 19)
        x, y = a.download(), b.download()
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"method {method}: graph runs differ from eager"
        assert_particles_equal(mpm, x, o.particles(), f"method {method}, 19 substeps through the graphs")
        assert a.stats()["substeps"] == 19
        c.step(method, 40)
        b.step(method, 21)
        assert np.array_equal(c.download().view(np.uint8), b.download().view(np.uint8)), f"method {method}: 40-substep call differs"
        assert c.stats()["substeps"] == 40
        a.close()
        b.close()
        c.close()


def test_keys_carried_between_substeps_follow_the_state(mpm, orc):
    """A substep leaves the next one's block keys (made where positions become final). They are
    used only for the state and grid they describe: after the grid is moved between calls
    (writeGridMinCoords / writeGridMaxCoords, 2.5 cells in x and 1.25 in y, so keys change and some
    particles leave the grid), after a fresh upload, across a switch between the explicit and PBMPM
    methods and in the 8-substep run graphs, every substep still equals the oracle given the same
    uniforms and particles."""
    sc = _small_pbmpm(mpm, n=6000, seed=31)
    sim, o = make_pair(mpm, orc, sc, use_graph=1)
    sim.step(1, 3)
    o.step(1, 3)
    assert_particles_equal(mpm, sim.download(), o.particles(), "before the grid moves")
    um = sim.uniformsManager
    mn = np.frombuffer(um.read(16, 12), np.float32).copy()
    mx = np.frombuffer(um.read(32, 12), np.float32).copy()
    res = np.frombuffer(um.read(208, 12), np.uint32)
    cell = (mx - mn) / res.astype(np.float32)
    shift = np.array([2.5, 1.25, 0.0], np.float32) * cell
    um.writeGridMinCoords(list(mn + shift))
    um.writeGridMaxCoords(list(mx + shift))
    o.set_uniforms(sim.uniform_bytes())
    for s in range(3):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"moved grid, substep {s}")
    # a fresh upload of the same particles: the next substep bins them afresh
    p0 = o.particles()
    sim.upload(p0)
    sim.step(1, 2)
    o.step(1, 2)
    assert_particles_equal(mpm, sim.download(), o.particles(), "after a re-upload")
    # explicit substeps bin from keys made by PBMPM's LAST and the other way round
    sim.step(0, 2)
    o.step(0, 2)
    sim.step(1, 2)
    o.step(1, 2)
    assert_particles_equal(mpm, sim.download(), o.particles(), "after switching methods")
    # the run graphs (8 substeps in one replay, prepared for the steady state): a call that starts
    # right after the grid moves back runs its first substep eagerly, then replays
    sim.prepare_graphs(1)
    um.writeGridMinCoords(list(mn))
    um.writeGridMaxCoords(list(mx))
    o.set_uniforms(sim.uniform_bytes())
    sim.step(1, 11)
    o.step(1, 11)
    assert_particles_equal(mpm, sim.download(), o.particles(), "run graphs after the grid moved back")
    sim.close()


@pytest.mark.parametrize("method", [1, 0])
def test_non_cubic_partial_block_grid_substeps_bitexact(mpm, orc, method):
    """A 30 x 22 x 41 grid over a non-cubic box (blocks of 4 cells, the last block along each axis
    partly outside the grid, cells of three different sizes) with the box collider: the keys a
    substep makes for the next use the grid's own block counts per axis, for both methods."""
    rng = np.random.default_rng(61)
    n = 7000
    pos = (np.array([0.2, 0.15, 0.25]) + rng.uniform(0.0, 1.0, (n, 3)) * np.array([0.6, 0.5, 0.5])).astype(np.float32)
    vel = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    p = mpm.make_particles(pos, vel=vel)
    sim_kw = dict(grid_res=(30, 22, 41), grid_min=(0.0, 0.0, 0.0), grid_max=(1.0, 0.75, 1.1), gravity=(0.0, -9.8, 0.0))
    if method == 1:
        p["dx"][:, :3] = vel * np.float32(1 / 600)
        sim_kw.update(dt=1.0 / 600, pbmpm_iterations=3, pbmpm_force_divisor=3.0)
    else:
        sim_kw.update(dt=1e-4)
    sc = dict(particles=p, sim=sim_kw, collider=mpm.scenes.box_mesh((0.5, 0.4, 0.55), 0.25))
    sim, o = make_pair(mpm, orc, sc, use_graph=1)
    for s in range(5):
        sim.step(method, 1)
        o.step(method, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"method {method}, non-cubic grid, substep {s}")
    sim.close()


def test_explicit_with_collider_substeps_bitexact(mpm, orc):
    """Explicit substeps against a box collider that the falling cube overlaps: the G2P pass makes
    the next substep's keys of the particles it finalises and the collider pass those of the queued
    ones, over several substeps, in graphs and eagerly."""
    sc = mpm.scenes.scene("explicit_small", n_particles=12000, seed=41)
    sc = dict(sc, collider=mpm.scenes.box_mesh((0.5, 0.46, 0.5), 0.2))
    for use_graph in (1, 0):
        sim, o = make_pair(mpm, orc, sc, use_graph=use_graph)
        for s in range(6):
            sim.step(0, 1)
            o.step(0, 1)
            assert_particles_equal(mpm, sim.download(), o.particles(), f"explicit + collider substep {s} (graph {use_graph})")
        assert sim.stats()["collider_queued"] > 0
        sim.close()


def test_divisor_fast_path_verified(mpm):
    """P2G's reciprocal+FMA quotient is enabled only for (dt, scale) pairs whose exhaustive device
    check (all 2^32 float numerators) matched i32((x / dt) * scale); the scene time steps pass."""
    sim = mpm.MpmSimulation(64, grid_res=(128, 128, 128), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1.0 / 600)
    lib = sim.lib
    fast = {}
    for d in (1.0 / 600, 1.0 / 384, 1.0 / 192, 1e-4, 5e-5, 1.0 / 128):
        f = ctypes.c_int(-1)
        assert lib.mpm_divisor_fast(sim.ctx, ctypes.c_float(d), ctypes.c_float(65536.0), ctypes.byref(f)) == 0
        fast[d] = f.value
        print("divisor", d, "fast" if f.value else "ieee")
    assert fast[1.0 / 600] == 1 and fast[1e-4] == 1
    sim.close()


@pytest.mark.parametrize("case", ["scale_not_pow2", "huge_dx_lanes"])
def test_pbmpm_p2g_general_form_bitexact(mpm, orc, case):
    """P2G folds the fixed-point scale into the mass only for S = 2^k and particles with
    |m|, |dx|, |D| <= 2^36; everything else takes the general form, per lane. Both cases are
    compared against the oracle: a scale of 10^5 (every particle general), and a few particles
    with |dx| = 2^40 scattered through the waves of an otherwise folded pass."""
    sc = _small_pbmpm(mpm, n=6000, seed=14)
    if case == "huge_dx_lanes":
        p = sc["particles"].copy()
        p["dx"][::97, 0] = np.float32(2.0 ** 40)
        sc = dict(sc, particles=p)
    for stage in (1, 2):
        sim, o = make_pair(mpm, orc, sc, pbmpm_iterations=3, pbmpm_force_divisor=3.0)
        if case == "scale_not_pow2":
            sim.uniformsManager.writeFixedPointScale(1e5)
            o.set_uniforms(sim.uniform_bytes())
        sim.debug_stage(1, stage)
        o.run_stage(1, stage)
        assert_grid_equal(sim, o, f"{case} stage {stage}")
        sim.close()
    sim, o = make_pair(mpm, orc, sc, pbmpm_iterations=3, pbmpm_force_divisor=3.0)
    if case == "scale_not_pow2":
        sim.uniformsManager.writeFixedPointScale(1e5)
        o.set_uniforms(sim.uniform_bytes())
    for s in range(3):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"{case} substep {s + 1}")
    sim.close()
