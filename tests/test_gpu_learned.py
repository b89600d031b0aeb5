"""GPU checks of the learned PBMPM warm start (no reference counterpart): the matrix-core forward
and backward against the numpy oracle (oracle/mlp.py, bfloat16 emulated), the warm-start
inference applied to particle records, the per-particle fallback, and on-device training."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def _net(seed, out_scale=1.0):
    p = mlp.init_params(seed, out_scale=out_scale)
    norm = np.zeros((4, 16), np.float32)
    norm[0, :16] = np.linspace(-0.5, 0.5, 16)
    norm[1, :16] = np.linspace(0.5, 2.0, 16)
    norm[0, 13:16] = 0.0                     # dprev: centred, ~1e-5 scale
    norm[1, 13:16] = 2e-5
    norm[2, :12] = np.linspace(-1e-3, 1e-3, 12)
    norm[3, :12] = np.linspace(1e-3, 4e-3, 12)
    norm[3, 12:] = 1.0
    return p, norm.ravel()


def test_mlp_gradient_matches_oracle(mpm):
    """Gradient of the MSE loss over 1,024 samples: every layer's weights and biases, the
    standardised outputs and the loss, against the bfloat16-emulating oracle. Tolerances: outputs
    2% of their largest magnitude, gradient 2% in norm and 5% of the largest entry per entry
    (single bfloat16 rounding flips of intermediate activations are 2^-8 relative)."""
    rng = np.random.default_rng(11)
    n = 1024
    p, norm = _net(3)
    nm = norm.reshape(4, 16)
    feats = (nm[0] + nm[1] * rng.standard_normal((n, 16))).astype(np.float32)
    targs = (nm[2] + nm[3] * rng.standard_normal((n, 16))).astype(np.float32)
    targs[:, 12:] = 0
    g, y, loss = mpm.mlp_grad(p, norm, feats, targs)
    go, yo, lo = mlp.forward_backward(p, norm, feats, targs, emulate=True)
    assert np.abs(y - yo).max() <= 0.02 * np.abs(yo).max()
    assert abs(loss - lo) <= 0.01 * lo
    for name, a, b in (("W1", mlp.W1, mlp.B1), ("b1", mlp.B1, mlp.W2), ("W2", mlp.W2, mlp.B2),
                       ("b2", mlp.B2, mlp.W3), ("W3", mlp.W3, mlp.B3), ("b3", mlp.B3, mlp.NPARAM)):
        ga, gb = g[a:b], go[a:b]
        assert np.linalg.norm(ga - gb) <= 0.02 * np.linalg.norm(gb) + 1e-12, name
        assert np.abs(ga - gb).max() <= 0.05 * np.abs(gb).max() + 1e-12, name
    assert np.all(g[mlp.W3:mlp.B3].reshape(64, 16)[:, mlp.NOUT:] == 0)


def _records(mpm, n, rng):
    """Particles in a 64^3 unit grid with random near-rotation F_e, F_p, D, velocity, mass."""
    pos = rng.uniform(0.2, 0.8, (n, 3)).astype(np.float32)
    p = mpm.make_particles(pos, vel=rng.standard_normal((n, 3)).astype(np.float32),
                           mass=rng.uniform(0.5, 1.5, n).astype(np.float32))
    fe = np.empty((n, 3, 3))
    for i in range(n):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        q *= np.sign(np.linalg.det(q))
        fe[i] = q @ np.diag(rng.uniform(0.9, 1.1, 3))
    fp = np.eye(3) + 0.05 * rng.standard_normal((n, 3, 3))
    d = 2e-3 * rng.standard_normal((n, 3, 3))
    p["Fe"][:, :, :3] = np.swapaxes(fe, 1, 2)           # column-major storage
    p["Fp"][:, :, :3] = np.swapaxes(fp, 1, 2)
    p["D"][:, :, :3] = np.swapaxes(d, 1, 2)
    return p, fe, fp, d


def _special_fe(rng):
    """F_e whose singular values the features must get right where a 3x3 eigen-solver is weakest:
    the identity and pure rotations (all three equal; F^T F diagonal, every Jacobi pivot 0), diagonal
    stretches (pivots 0, values distinct), and rotated stretches with two equal singular values."""
    out = [np.eye(3), np.diag([1.1, 0.95, 1.0]), np.diag([0.9, 1.0, 1.1])]
    for sig in ([1.0, 1.0, 1.0], [1.05, 1.05, 0.92], [1.08, 0.97, 0.97], [0.93, 1.02, 0.93]):
        for _ in range(6):
            u, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            v, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            u *= np.sign(np.linalg.det(u))
            v *= np.sign(np.linalg.det(v))
            out.append(u @ np.diag(sig) @ v.T)
    return np.array(out)


@pytest.mark.parametrize("sig_sd", [None, 0.03])
def test_warmstart_inference_matches_oracle(mpm, sig_sd):
    """One warm start over 4,000 particles: D and dx plus the destandardised predictions per
    particle, against the oracle network on features from numpy's SVD; a particle whose dprev is
    0 (no iteration reached it) keeps plain PBMPM and counts as a fallback. The first particles
    carry the special F_e of _special_fe. With sig_sd the singular-value features are standardised
    as a trained set standardises them (mean 1, standard deviation sig_sd): an error of 1e-3 in a
    singular value then moves its network input by 1e-3 / sig_sd."""
    rng = np.random.default_rng(12)
    n = 4000
    p, fe, fp, d = _records(mpm, n, rng)
    sp = _special_fe(rng)
    fe[:len(sp)] = sp
    p["Fe"][:len(sp), :, :3] = np.swapaxes(sp, 1, 2)
    p["dx"][:, :3] = 1e-3 * rng.standard_normal((n, 3))
    dprev = (2e-5 * rng.standard_normal((n, 3))).astype(np.float32)
    dprev[::50] = 0.0
    params, norm = _net(4, out_scale=0.5)
    if sig_sd is not None:
        nm = norm.reshape(4, 16)
        nm[0, :3] = 1.0
        nm[1, :3] = sig_sd
        norm = nm.ravel()
    with mpm.MpmSimulation(n, grid_res=(64, 64, 64), grid_min=(0, 0, 0), grid_max=(1, 1, 1), dt=1 / 600,
                           pbmpm_iterations=2, pbmpm_force_divisor=2.0) as sim:
        sim.upload(p)
        sim.set_learn_params(params, norm)
        sim.learn_enable(True, limit=1e6)
        sim.learn_dprev(dprev)
        assert np.array_equal(sim.learn_dprev(), dprev)
        sim.debug_stage(1, -2)                           # MPM_STAGE_WARM_START
        out = sim.download()
        assert sim.stats()["mlp_fallbacks"] == n // 50
    feats = np.stack([mlp.features(fe[i], fp[i], d[i], dprev[i]) for i in range(n)]).astype(np.float32)
    _, y, _ = mlp.forward_backward(params, norm, feats, np.zeros((n, 16), np.float32), emulate=True)
    nm = norm.reshape(4, 16)
    pred = y[:, :12] * nm[3, :12] + nm[2, :12]
    used = np.any(dprev != 0, axis=1)
    want_d = d + np.where(used[:, None, None], pred[:, :9].reshape(n, 3, 3).transpose(0, 2, 1), 0.0)  # column-major outputs
    want_dx = p["dx"][:, :3].astype(np.float64) + np.where(used[:, None], pred[:, 9:12], 0.0)
    got = mpm.matrix(out["D"])
    err = np.abs(got - want_d).reshape(n, 9) / nm[3, :9]
    assert np.quantile(err, 0.99) < 0.02 and err.max() < 0.1
    errx = np.abs(out["dx"][:, :3] - want_dx) / nm[3, 9:12]
    assert np.quantile(errx, 0.99) < 0.02 and errx.max() < 0.1
    k = len(sp)
    assert err[:k].max() < 0.1 and errx[:k].max() < 0.1
    assert np.array_equal(out["dx"][~used, :3], p["dx"][~used, :3]) and np.array_equal(out["D"][~used], p["D"][~used])
    # everything but D and dx is untouched
    assert np.array_equal(out["pos"], p["pos"]) and np.array_equal(out["Fe"], p["Fe"]) and np.array_equal(out["vel"], p["vel"])


def test_fallback_keeps_plain_pbmpm_bitexact(mpm):
    """With a limit no prediction meets, every particle falls back: the substeps equal plain PBMPM
    bit for bit and the fallback count is the particle count."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=20000, seed=31)
    params, norm = _net(5)
    outs = []
    for learned in (False, True):
        with mpm.MpmSimulation(20000, **sc["sim"]) as sim:
            sim.set_collider(*sc["collider"])
            sim.upload(sc["particles"])
            if learned:
                sim.set_learn_params(params, norm)
                sim.learn_enable(True, limit=1e-30)
            sim.step(1, 5)
            outs.append(sim.download())
            if learned:
                assert sim.stats()["mlp_fallbacks"] == 20000
    assert np.array_equal(outs[0].view(np.uint8), outs[1].view(np.uint8))


def test_collect_and_train_reduces_loss(mpm):
    """Pairs logged from k = 16 substeps, then Adam on the matrix cores: the standardised loss
    (about 1 per output for an untrained network) falls."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=32768, seed=32)
    with mpm.MpmSimulation(32768, **sc["sim"]) as sim:
        sim.set_collider(*sc["collider"])
        sim.upload(sc["particles"])
        sim.step(1, 20)
        npairs = sim.learn_collect(40, k_ref=16, every=2)
        assert npairs == 19 * 32768          # substeps 2, 4, .., 38 (substep 0 only produces dprev)
        first = sim.learn_train(50, batch=4096, lr=1e-3, seed=1)
        later = sim.learn_train(1500, batch=4096, lr=1e-3, seed=2)
        assert np.isfinite(first) and np.isfinite(later)
        assert later < 0.85 * first and later < 0.8
        params, norm = sim.learn_params()
        assert np.all(np.isfinite(params)) and np.all(norm.reshape(4, 16)[[1, 3]] > 0)
        sim.learn_enable(True)
        sim.step(1, 10)
        assert np.isfinite(mpm.state_columns(sim.download())).all()


def test_residual_tracking(mpm):
    """Per-iteration residuals of plain PBMPM fall with the iteration index on the two-sphere scene."""
    sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=32768, seed=33)
    with mpm.MpmSimulation(32768, **sc["sim"]) as sim:
        sim.set_collider(*sc["collider"])
        sim.upload(sc["particles"])
        sim.set_iterations(8)
        sim.set_residual_tracking(True)
        sim.step(1, 10)
        mean, peak, ns = sim.residuals()
        assert ns == 10
        assert mean[0] == 0 and np.all(mean[1:8] > 0) and np.all(mean[8:] == 0)
        assert mean[7] < mean[1] and np.all(peak[1:8] >= mean[1:8])


def test_trained_warm_start_gain_over_plain(mpm):
    """What the trained network gains. PBMPM's iterations relax the constraints towards a fixed
    point, so accuracy is measured against the most-iterated run (k = 64): the warm start at
    k = 16, trained on another sampling of the scene (trajectory at 16 iterations, targets from
    64-iteration substeps from the same states), must end closer to the 64-iteration run than
    plain PBMPM at 48 iterations does (3x fewer iterations; about 62 measured). Two-sphere scene,
    262,144 particles, 200 substeps; rms position error over the particles."""
    n, k, k_ref, sub = 262144, 16, 64, 200

    def make(seed):
        sc = mpm.scenes.scene("pbmpm_two_spheres", n_particles=n, seed=seed)
        sim = mpm.MpmSimulation(n, **sc["sim"])
        sim.set_collider(*sc["collider"])
        sim.upload(sc["particles"])
        return sim

    with make(41) as tr:
        assert tr.learn_collect(sub, k_ref=k_ref, every=10, k_run=k) == 19 * n
        tr.learn_train(4000, batch=65536, lr=1e-3, seed=3)
        tr.learn_train(2000, batch=65536, lr=2e-4, seed=4)
        net = tr.learn_params()

    def final_pos(kk, learned):
        with make(42) as sim:
            sim.set_iterations(kk)
            if learned:
                sim.set_learn_params(*net)
                sim.learn_enable(True)
            sim.step(1, sub)
            p = sim.download()["pos"][:, :3].astype(np.float64)
        assert np.isfinite(p).all()
        return p

    ref = final_pos(k_ref, False)

    def err(p):
        return float(np.sqrt((np.linalg.norm(p - ref, axis=1) ** 2).mean()))

    e_learn, e48, e16 = err(final_pos(k, True)), err(final_pos(48, False)), err(final_pos(k, False))
    print(f"rms error vs 64 iterations: learned k=16 {e_learn:.4f}, plain k=48 {e48:.4f}, plain k=16 {e16:.4f}")
    assert e_learn < e48 and e_learn < 0.25 * e16
