"""GPU checks of the batched-scene step (no reference counterpart): every scene of a batch must
come out bit-identical to the same scene stepped alone in its own context (the single-scene path
is itself bit-exact against the oracle), in the persistent one-launch schedule and the
launch-per-phase schedule, for several chunk sizes, with graphs and without, and across the
autotuner (which must leave the state as it found it)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_batch(mpm):
    """Six scenes: explicit and PBMPM, grids 32^3 / 48^3 / 64^3, k = 2..5, uneven sizes, one
    scene starting partly below its grid (unbinned in the first substep, then clamped onto the
    wall), sticky walls on one."""
    rng = np.random.default_rng(51)
    out = []
    specs = [(0, 32, 3000, None), (1, 64, 9000, 3), (1, 32, 2500, 2), (0, 48, 7000, None), (1, 48, 5000, 5), (1, 32, 4000, 4)]
    for i, (method, res, n, k) in enumerate(specs):
        pos = rng.uniform(0.25, 0.75, (n, 3)).astype(np.float32)
        if i == 5:
            pos[:300, 1] = rng.uniform(-0.2, -0.01, 300)         # below the grid: unbinned
        vel = rng.uniform(-2, 2, 3).astype(np.float32)
        p = mpm.make_particles(pos, vel=vel)
        sim = dict(n_particles=n, grid_res=(res, res, res), grid_min=(0, 0, 0), grid_max=(1, 1, 1), gravity=(0.0, -9.8, 0.0))
        if method == 0:
            sim.update(method=0, dt=1e-4, wall_mode=1 if i == 0 else 0, wall_cells=3)
        else:
            p["dx"][:, :3] = vel * np.float32(1 / 600)
            sim.update(method=1, dt=1 / 600, pbmpm_iterations=k, pbmpm_force_divisor=float(k))
        out.append(dict(particles=p, sim=sim))
    return out


def _alone(mpm, sc, substeps):
    kw = dict(sc["sim"])
    method = kw.pop("method")
    n = kw.pop("n_particles")
    with mpm.MpmSimulation(n, **kw) as sim:
        sim.upload(sc["particles"])
        sim.step(method, substeps)
        return sim.download()


def _same(a, b):
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.fixture(scope="module")
def batch(mpm):
    scenes = _small_batch(mpm)
    ref = [_alone(mpm, sc, 4) for sc in scenes]
    return scenes, ref


@pytest.mark.parametrize("persistent,chunk,wg", [(True, 512, 2), (False, 512, 2), (True, 256, 3), (True, 1024, 1), (False, 1024, 4), (False, 2048, 3), (False, 256, 2)])
def test_batch_equals_scenes_alone(mpm, batch, persistent, chunk, wg):
    scenes, ref = batch
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
        ens.set_tuning(chunk, wg, persistent)
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        ens.step(4)
        ens.sync()
        for i in range(len(scenes)):
            got = ens.download(i)
            assert _same(got, ref[i]), f"scene {i} differs (persistent={persistent}, chunk={chunk})"
        st = ens.stats()
        assert st["n_particles"] == sum(sc["sim"]["n_particles"] for sc in scenes)
        assert st["blocks_dropped"] == 0 and st["probe_failures"] == 0


def test_batch_with_block_table_equals_scenes_alone(mpm, batch):
    """hash_map_size given: the batch bins through the open-addressing block table instead of
    direct key addressing; the results are the same bits."""
    scenes, ref = batch
    with mpm.MpmEnsemble([dict(sc["sim"], hash_map_size=400009) for sc in scenes]) as ens:
        ens.set_tuning(512, 2, False)
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        ens.step(4)
        for i in range(len(scenes)):
            assert _same(ens.download(i), ref[i]), f"scene {i}"
        assert ens.stats()["probe_failures"] == 0


def test_batch_eager_equals_graph(mpm, batch):
    scenes, ref = batch
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes], use_graph=0) as ens:
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        ens.step(4)
        for i in range(len(scenes)):
            assert _same(ens.download(i), ref[i]), f"scene {i}"


def test_autotune_restores_state(mpm, batch):
    scenes, ref = batch
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        best, cands = ens.autotune(trial_substeps=2)
        assert len(cands) == 9 and all(c["ms_per_substep"] > 0 and c["persistent"] == 0 for c in cands)
        assert best["ms_per_substep"] == min(c["ms_per_substep"] for c in cands)
        assert ens.tuning()["chunk_particles"] == best["chunk_particles"]
        for i, sc in enumerate(scenes):
            assert _same(ens.download(i), sc["particles"]), f"scene {i} changed by the autotuner"
        ens.step(4)
        for i in range(len(scenes)):
            assert _same(ens.download(i), ref[i]), f"scene {i} after autotune"


def test_reupload_one_scene_midway(mpm, batch):
    """Particles of one scene can be replaced after stepping (rows found by particle id)."""
    scenes, ref = batch
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        ens.step(3)
        ens.upload(2, scenes[2]["particles"])
        ens.step(1)
        assert _same(ens.download(0), _alone(mpm, scenes[0], 4))
        assert _same(ens.download(2), _alone(mpm, scenes[2], 1))


def test_ensemble_workload_runs(mpm):
    """The 64-scene batched workload (about 7M particles): a few substeps, finite state."""
    scenes = mpm.scenes.ensemble_scenes()
    with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
        for i, sc in enumerate(scenes):
            ens.upload(i, sc["particles"])
        ens.step(3)
        ens.sync()
        for i in (0, 1, 63):
            assert np.isfinite(mpm.state_columns(ens.download(i))).all()
        assert ens.stats()["blocks_dropped"] == 0


def test_batch_scenes_match_oracle_directly(mpm, orc, batch):
    """Three scenes of the batch (explicit with sticky walls, PBMPM k = 3 and k = 5, one of them
    partly below its grid) against the oracle run on that scene alone, bit for bit, in both
    schedules: not only transitively through the single-scene path."""
    from helpers import uniforms_blob, SHARED_KEYS, assert_particles_equal
    scenes, _ = batch
    for persistent in (False, True):
        with mpm.MpmEnsemble([sc["sim"] for sc in scenes]) as ens:
            ens.set_tuning(512, 2, persistent)
            for i, sc in enumerate(scenes):
                ens.upload(i, sc["particles"])
            ens.step(4)
            ens.sync()
            for i in (0, 1, 4, 5):
                kw = scenes[i]["sim"]
                uni = uniforms_blob(np.float32(kw["dt"]), kw["grid_min"], kw["grid_max"], kw["grid_res"])
                o = orc.OracleSim(scenes[i]["particles"], uni, hash_map_size=200003, max_blocks=70000,
                                  **{k: v for k, v in kw.items() if k in SHARED_KEYS})
                o.step(kw["method"], 4)
                assert_particles_equal(mpm, ens.download(i), o.particles(), f"batch scene {i} (persistent={persistent})")
