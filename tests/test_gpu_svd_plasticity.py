"""The SVD plasticity option (plasticity_mode 1: FP32 SVD; 2: with the FP64 fallback for
near-degenerate F_e) through the C ABI against the oracle, bit for bit: F_e with singular values
around the band edges, near-singular and reflected, in both methods; and full PBMPM substeps of the
two-sphere scene with the box collider. The FP64 count in mpm_stats equals the oracle's."""
import numpy as np
import pytest

from helpers import make_pair, assert_particles_equal

pytestmark = pytest.mark.gpu


def _adversarial_fe(rng, n):
    lo, hi = 1.0 - 2.5e-2, 1.0 + 7.5e-3
    centres = np.array([lo, lo + 5e-4, hi - 5e-4, hi, 1.0, 0.9, 1.05])
    sv = centres[rng.integers(0, centres.size, (n, 3))] + rng.uniform(-3e-4, 3e-4, (n, 3))
    sv[::7, 2] = rng.uniform(1e-4, 8e-3, sv[::7].shape[0])            # near-singular: FP64 in mode 2
    q1, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q2, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    q2[1::5] = np.eye(3)
    return np.einsum("nij,nj,nkj->nik", q1, sv, q2).astype(np.float32)


@pytest.mark.parametrize("mode", [1, 2])
def test_svd_plasticity_adversarial_f_bitexact(mpm, orc, mode):
    rng = np.random.default_rng(90 + mode)
    n = 8000
    fe = _adversarial_fe(rng, n)
    pos = rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32)
    for method in (1, 0):
        print(f"svd plasticity mode {mode}, method {method}", flush=True)
        p = mpm.make_particles(pos, vel=rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32))
        p["Fe"][:, :, :3] = np.swapaxes(fe, 1, 2)
        sc = dict(particles=p, sim=dict(grid_res=(32, 32, 32), grid_min=(0, 0, 0), grid_max=(1, 1, 1),
                                        dt=1.0 / 600 if method == 1 else 1e-4, gravity=(0.0, -9.8, 0.0),
                                        pbmpm_iterations=3, pbmpm_force_divisor=3.0,
                                        youngs_modulus=1.4e5 if method == 1 else 0.0), collider=None)
        sim, o = make_pair(mpm, orc, sc, plasticity_mode=mode)
        for s in range(2):
            before = o.stats()["svd_fp64"]
            sim.step(method, 1)
            o.step(method, 1)
            assert_particles_equal(mpm, sim.download(), o.particles(), f"mode {mode} method {method} substep {s}")
            assert sim.stats()["svd_fp64_particles"] == o.stats()["svd_fp64"] - before
        if mode == 2:
            assert o.stats()["svd_fp64"] > 0
        sim.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_svd_plasticity_pbmpm_scene_bitexact(mpm, orc, mode):
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=8192, seed=93)
    sim, o = make_pair(mpm, orc, sc, plasticity_mode=mode)
    print(f"svd plasticity mode {mode}, two spheres", flush=True)
    for s in range(5):
        sim.step(1, 1)
        o.step(1, 1)
        assert_particles_equal(mpm, sim.download(), o.particles(), f"mode {mode} substep {s}")
    sim.close()
