"""mpm_splat_noalloc (the allocation-free variant of mpm_splat) through the C ABI against mpm_splat:
both splat kinds, two calls in a row with the second on a smaller grid and a smaller particle set
(a stale accumulation grid would show), a grid beyond the array's maximum rejected, and a fresh
context whose only splats are allocation-free never gets the context grid that mpm_splat grows."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MPM_E_RANGE = -4


def _sim(mpm, n, seed, res=(64, 64, 64)):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=n, seed=seed)
    sim = mpm.MpmSimulation(n, **dict(sc["sim"], grid_res=res))
    sim.set_collider(*sc["collider"])
    sim.upload(sc["particles"])
    sim.step(1, 2)
    return sim


def _noalloc(sim, kind, res):
    r = np.ascontiguousarray(res if res is not None else [0, 0, 0], np.uint32)
    cells = int(np.prod(r)) if kind == 0 else int(np.prod(sim.uniformsManager.read(208, 12).view(np.uint32)))
    out = np.full(cells, 0xDEADBEEF, np.uint32)
    rc = sim.lib.mpm_splat_noalloc(sim.ctx, kind, r.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), out.size)
    return rc, out


def test_splat_noalloc_matches_splat_on_two_calls(mpm):
    lib = mpm.load_library()
    big = _sim(mpm, 60000, 61)
    small = _sim(mpm, 7000, 62, res=(32, 32, 32))
    try:
        # no splat has run on `small`: its context grid must stay unallocated throughout
        assert lib.mpm_splat_grid(small.ctx) is None
        for sim, res in ((big, (48, 40, 56)), (small, (20, 24, 16))):
            print(f"density splat {res}, {sim.nParticles} particles", flush=True)
            rc, got = _noalloc(sim, 0, res)
            assert rc == 0, lib.mpm_last_error(sim.ctx)
            want = sim.splat_density(res) if sim is big else None
            if want is None:    # the reference values for `small` from a second context
                with _sim(mpm, 7000, 62, res=(32, 32, 32)) as twin:
                    want = twin.splat_density(res)
            assert np.array_equal(got, want.ravel()), "density splat differs"
            assert got.sum() > 0
        for sim in (big, small):
            print(f"mass splat, {sim.nParticles} particles", flush=True)
            rc, got = _noalloc(sim, 1, None)
            assert rc == 0, lib.mpm_last_error(sim.ctx)
            if sim is big:
                want = sim.splat_mass()
            else:
                with _sim(mpm, 7000, 62, res=(32, 32, 32)) as twin:
                    want = twin.splat_mass()
            assert np.array_equal(got, want.ravel()), "mass splat differs"
        assert lib.mpm_splat_grid(small.ctx) is None
        # beyond the array: 160 x 128 x 128 cells > 2^21
        rc, _ = _noalloc(small, 0, (160, 128, 128))
        assert rc == MPM_E_RANGE and b"2^21" in lib.mpm_last_error(small.ctx)
        # and a call after the rejected one still matches
        rc, got = _noalloc(big, 0, (48, 40, 56))
        assert rc == 0 and np.array_equal(got, big.splat_density((48, 40, 56)).ravel())
    finally:
        big.close()
        small.close()
