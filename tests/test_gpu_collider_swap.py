"""A collider replaced after substeps have been replayed from a captured hipGraph: the indexed box is
swapped for the same box as a triangle soup (same triangles, BVH and index count, so every
substep parameter is unchanged, but the collider words are reallocated), and then for a moved
box of the same size. The substeps after each swap must match the oracle bit for bit; a replay
of the graph captured before the swap would read the freed collider buffers."""
import numpy as np
import pytest

from helpers import make_pair, assert_particles_equal

pytestmark = pytest.mark.gpu


def test_collider_swap_after_graph_replay_bitexact(mpm, orc):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=6000, seed=21)
    sim, o = make_pair(mpm, orc, sc)
    assert sim.config.use_graph
    for s in range(2):
        sim.step(1, 1)
        o.step(1, 1)
    assert_particles_equal(mpm, sim.download(), o.particles(), "before the swap")
    v, idx = sc["collider"]
    soup = v[idx].reshape(-1, 3)
    for what, (vv, ii) in (("triangle soup", (soup, np.arange(idx.size, dtype=np.uint32))),
                           ("moved box", mpm.scenes.box_mesh((0.5, 0.45, 0.52), 0.2))):
        print(f"collider swap: {what}", flush=True)
        cm = sim.set_collider(vv, ii)
        o.set_uniforms(sim.uniform_bytes())
        o.set_collider(cm.colliderData, cm.bvhNodes)
        for s in range(2):
            sim.step(1, 1)
            o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"after the swap to the {what}")
    sim.close()
