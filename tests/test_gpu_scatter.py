"""GPU parity of the particle scatter init (particleScatter.cs.wgsl) against the oracle: every
field of every record bit-exact, for a box (every first candidate accepted) and an octahedron
(rejection sampling), then one PBMPM substep from the scattered state."""
import numpy as np
import pytest

from helpers import assert_particles_equal
from test_scatter_oracle import _box_soup, _octahedron_soup, _xyzw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", ["box", "octahedron"])
def test_scatter_bitexact(mpm, orc, shape):
    soup = _box_soup() if shape == "box" else _octahedron_soup()
    n = 40000
    with mpm.MpmSimulation(n, grid_res=(64, 64, 64), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1 / 600) as sim:
        mm = sim.set_mesh(soup)
        sim.scatterParticlesInMeshVolume()
        got = sim.download()
        want = orc.scatter(_xyzw(soup), np.float32(mm.minCoords), np.float32(mm.maxCoords), n)
        assert np.array_equal(got.view(np.float32).reshape(n, 48).view(np.uint32), want.view(np.uint32)), shape


def test_scatter_then_step_matches_oracle(mpm, orc):
    soup = _octahedron_soup()
    n = 20000
    with mpm.MpmSimulation(n, grid_res=(64, 64, 64), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1 / 600,
                           gravity=(0.0, -9.8, 0.0), pbmpm_iterations=3, pbmpm_force_divisor=3.0) as sim:
        sim.set_mesh(soup)
        sim.scatterParticlesInMeshVolume()
        start = sim.download()
        o = orc.OracleSim(start, sim.uniform_bytes(), hash_map_size=200_003, max_blocks=70_000, gravity=(0.0, -9.8, 0.0),
                          pbmpm_iterations=3, pbmpm_force_divisor=3.0)
        sim.step(1, 2)
        o.step(1, 2)
        assert_particles_equal(mpm, sim.download(), o.particles(), "scatter + 2 PBMPM substeps")
