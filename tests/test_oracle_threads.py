"""The oracle's threaded passes (OpenMP, atomic i32 grid sums, CAS block table, atomic binning
cursors; the CPU baseline of the bench runs them) give the serial oracle's values bit for bit:
every particle word after substeps of both methods with the box collider, and every cell of the
cell-keyed grid (block numbering and the order within a block differ, no value does)."""
import numpy as np
import pytest

from helpers import uniforms_blob, keyed, SHARED_KEYS


def _pair(mpm, orc, name, n, seed, threads):
    sc = mpm.scenes.scene(name, n_particles=n, seed=seed)
    kw = sc["sim"]
    coll = None
    words = nodes = None
    if sc["collider"] is not None:
        cm = mpm.GpuColliderBufferManager(*sc["collider"])
        words, nodes = cm.colliderData, cm.bvhNodes
        coll = (cm.minCoords, cm.maxCoords, cm.numIndices, None)
    uni = uniforms_blob(np.float32(kw["dt"]), kw["grid_min"], kw["grid_max"], kw["grid_res"], collider=coll)
    okw = {k: v for k, v in kw.items() if k in SHARED_KEYS}
    out = []
    for t in (1, threads):
        o = orc.OracleSim(sc["particles"], uni, hash_map_size=200003, max_blocks=70000, **okw)
        if words is not None:
            o.set_collider(words, nodes)
        assert o.set_threads(t) == t
        out.append(o)
    return sc["method"], out


@pytest.mark.parametrize("name,n", [("pbmpm_two_spheres", 20000), ("explicit_small", 20000)])
def test_threaded_oracle_equals_serial(mpm, orc, name, n):
    method, (a, b) = _pair(mpm, orc, name, n, 71, 8)
    a.step(method, 3)
    b.step(method, 3)
    pa = mpm.state_columns(a.particles()).astype(np.float32)
    pb = mpm.state_columns(b.particles()).astype(np.float32)
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))
    ka, va = keyed(*a.dump_grid())
    kb, vb = keyed(*b.dump_grid())
    assert np.array_equal(ka, kb) and np.array_equal(va, vb)
    assert a.stats() == b.stats()


def test_threads_only_in_fp32_mode(mpm, orc):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=100, seed=1)
    uni = uniforms_blob(np.float32(1 / 600), (0, 0, 0), (1, 1, 1), (32, 32, 32))
    assert orc.OracleSim(sc["particles"], uni, mode=1).set_threads(8) == 1
