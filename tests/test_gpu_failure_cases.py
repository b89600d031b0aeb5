"""Conditions the reference handles silently (mapAffectedBlocks.wgsl:11,27-31,52;
countParticlesPerBlock.wgsl:16-20) on the GPU path: particles with non-finite positions, and an
open-addressing block table forced to exhaust its probes."""
import numpy as np
import pytest

from helpers import make_pair, assert_grid_equal

pytestmark = pytest.mark.gpu


def _nan_equal(a, b):
    """Bit-equal words, or both NaN (a NaN's payload is not defined by the arithmetic)."""
    ab, bb = a.view(np.uint32), b.view(np.uint32)
    return (ab == bb) | (np.isnan(a) & np.isnan(b))


@pytest.mark.parametrize("method", [0, 1])
def test_nonfinite_particles_match_oracle(mpm, orc, method):
    """Particles with a NaN or infinite position or displacement: i32(NaN) = 0 bins them at cell 0
    and every grid contribution they make truncates to 0, so the grid and every finite particle
    equal the oracle bit for bit; the non-finite particles stay non-finite in the same words.
    The count is reported in mpm_stats.nonfinite_particles."""
    name = "pbmpm_two_spheres" if method else "explicit_small"
    sc = mpm.scenes.scene(name, n_particles=5000, seed=41)
    p = sc["particles"].copy()
    p["pos"][10] = [np.nan, 0.5, 0.5]
    p["pos"][20] = [0.5, np.inf, 0.5]
    p["pos"][30] = [-np.inf, np.nan, 0.2]
    p["dx"][40, :3] = [np.nan, 0.0, 0.0]
    p["vel"][50] = [np.inf, 0.0, 0.0]
    sc = dict(sc, particles=p)
    sim, o = make_pair(mpm, orc, sc)
    print(f"non-finite particles, method {method}", flush=True)
    pos = p["pos"]
    for s in range(3):
        # counted at binning: the positions the substep starts from (an infinite coordinate can be
        # clamped to a wall by integrate and be finite afterwards)
        expect = int((~np.isfinite(pos)).any(axis=1).sum())
        sim.step(method, 1)
        o.step(method, 1)
        got = sim.download()
        a = mpm.state_columns(got).astype(np.float32)
        b = mpm.state_columns(o.particles()).astype(np.float32)
        ok = _nan_equal(a, b)
        assert ok.all(), f"substep {s + 1}: {int((~ok).sum())} words differ, first {np.argwhere(~ok)[0]}"
        assert sim.stats()["nonfinite_particles"] == expect, (s, expect)
        pos = got["pos"]
    finite = np.isfinite(a).all(axis=1)
    assert finite.sum() >= p.shape[0] - 10
    sim.close()
    sim, o = make_pair(mpm, orc, sc)
    sim.debug_stage(method, 1)
    o.run_stage(method, 1)
    assert_grid_equal(sim, o, "grid with non-finite particles")
    assert sim.stats()["nonfinite_particles"] == 3
    sim.close()


def test_probe_exhaustion_is_counted_and_contained(mpm, orc):
    """A block table of 31 slots for a scene touching ~95 blocks: inserts exhaust their probes
    (max_probes 4). The failures are counted, no more than 31 blocks are allocated, every allocated
    block is one the full table allocates too, particles whose own block is missing are counted
    unbinned (mpm_stats), and the state stays finite over substeps."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=4000, seed=42)
    sc = dict(sc, sim=dict(sc["sim"], grid_res=(32, 32, 32)))
    sim, o = make_pair(mpm, orc, sc, hash_map_size=31, max_probes=4)
    print("probe exhaustion: 31-slot table", flush=True)
    full, _ = make_pair(mpm, orc, sc, hash_map_size=65537)
    full.debug_stage(1, 0)
    full_blocks = set(map(tuple, full.dump_blocks().tolist()))
    full.close()
    assert len(full_blocks) > 62
    sim.debug_stage(1, 0)
    st = sim.stats()
    blocks = set(map(tuple, sim.dump_blocks().tolist()))
    assert st["probe_failures"] > 0
    assert 0 < len(blocks) <= 31 and blocks <= full_blocks
    assert st["blocks_allocated"] == len(blocks)
    assert st["particles_binned"] + st["particles_unbinned"] == 4000 and st["particles_unbinned"] > 0
    # particle blocks (runs) whose own block is allocated hold exactly the binned particles
    pb, cnt = sim.dump_particle_blocks()
    held = sum(int(c) for b, c in zip(map(tuple, pb.tolist()), cnt) if b in blocks)
    assert held == st["particles_binned"]
    sim.step(1, 3)
    out = mpm.state_columns(sim.download())
    assert np.isfinite(out).all()
    sim.close()
