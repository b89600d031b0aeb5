"""GPU parity of the render-side splats (mcDensity.cs.wgsl calculateDensity,
calculateGridMass.wgsl) against the oracle: every cell bit-exact, on uploaded particles and on
the state after PBMPM substeps (sorted order), for cubic and non-cubic grids."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_splats_bitexact(mpm, orc):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=200000, seed=61)
    p = sc["particles"].copy()
    rng = np.random.default_rng(62)
    p["mass"] = rng.uniform(0.0, 1.0, p.shape[0]).astype(np.float32) ** 2
    p["pos"][:50] = 0.0                     # on the min wall: the mass-grid quirk
    with mpm.MpmSimulation(p.shape[0], **sc["sim"]) as sim:
        sim.set_collider(*sc["collider"])
        sim.upload(p)
        for step in range(2):
            state = sim.download()
            for res in ((64, 64, 64), (37, 41, 29)):
                got = sim.splat_density(res)
                want = orc.splat(0, state, (0, 0, 0), (1, 1, 1), 65536.0, res)
                assert np.array_equal(got, want), f"density {res} step {step}"
            got = sim.splat_mass()
            want = orc.splat(1, state, (0, 0, 0), (1, 1, 1), 65536.0, (128, 128, 128))
            assert np.array_equal(got, want), f"mass step {step}"
            sim.step(1, 3)
