"""GPU checks of the large explicit workload: per-particle Young's modulus bit-exact against the
oracle, uniform per-particle moduli identical to the scene-wide constant, and the full 4,194,304-
particle slab over the 7,938-triangle terrain (256^3 grid) against the oracle for three substeps."""
import numpy as np
import pytest

from helpers import make_pair, assert_particles_equal

pytestmark = pytest.mark.gpu


def test_per_particle_youngs_bitexact(mpm, orc):
    sc = mpm.scenes.scene("explicit_small", n_particles=20000, seed=71)
    E = np.exp(np.random.default_rng(72).uniform(np.log(5e4), np.log(2e5), 20000)).astype(np.float32)
    sim, o = make_pair(mpm, orc, sc)
    sim.set_particle_youngs(E)
    o.set_particle_youngs(E)
    for s in range(4):
        sim.step(0, 1)
        o.step(0, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"per-particle E substep {s}")
    sim.close()


def test_uniform_youngs_equals_the_constant(mpm):
    sc = mpm.scenes.scene("explicit_small", n_particles=20000, seed=73)
    outs = []
    for per_particle in (False, True):
        with mpm.MpmSimulation(20000, **sc["sim"]) as sim:
            sim.upload(sc["particles"])
            if per_particle:
                sim.set_particle_youngs(np.full(20000, 1.4e5, np.float32))
            sim.step(0, 5)
            outs.append(sim.download())
    assert np.array_equal(outs[0].view(np.uint8), outs[1].view(np.uint8))


def test_terrain_full_size_substep_bitexact(mpm, orc):
    sc = mpm.scenes.scene("explicit_slab_terrain")
    n = sc["particles"].shape[0]
    sim, o = make_pair(mpm, orc, sc, max_blocks=300_000)
    sim.set_particle_youngs(sc["youngs"])
    o.set_particle_youngs(sc["youngs"])
    sim.step(0, 1)
    o.step(0, 1)
    assert_particles_equal(mpm, sim.download(), o.particles(), "4M terrain substep")
    # two more: binned from the keys the G2P pass made
    sim.step(0, 2)
    o.step(0, 2)
    assert_particles_equal(mpm, sim.download(), o.particles(), "4M terrain substeps 2-3")
    st = sim.stats()
    assert st["n_particles"] == n and st["blocks_dropped"] == 0 and st["probe_failures"] == 0
    sim.step(0, 50)
    assert np.isfinite(mpm.state_columns(sim.download())).all()
    sim.close()
