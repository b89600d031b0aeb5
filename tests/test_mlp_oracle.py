"""CPU checks that pin the warm-start network oracle (oracle/mlp.py) before it judges the device:
closed values of the bfloat16 rounding and of GELU, the analytic gradient against central finite
differences of the float64 loss, and the bfloat16 emulation staying within rounding distance of
the float64 network."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0);
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 (odd) and 1 + 2^-6 (even).
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 3.0e38, 1e-40], np.float32)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 226 / 128 * 2.0 ** 127, 0.0], np.float32)  # 3e38 = 225.7/128 * 2^127
    got = mlp.bf16(x)
    assert got[:6].tolist() == want[:6].tolist()
    assert got[6] == np.float32(2.0 ** -133)                      # 1e-40 = 1.09 bfloat16 subnormal ulps (2^-133)
    assert np.all(mlp.bf16(got) == got)                           # idempotent


def test_gelu_closed_values():
    assert mlp.gelu(0.0) == 0.0
    assert mlp.gelu_grad(0.0) == 0.5
    assert abs(mlp.gelu(20.0) - 20.0) < 1e-12 and abs(mlp.gelu(-20.0)) < 1e-12
    assert abs(mlp.gelu_grad(20.0) - 1.0) < 1e-9 and abs(mlp.gelu_grad(-20.0)) < 1e-9
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp.gelu_grad(z))) < 1e-8
    # the tanh form stays within 1e-3 of the exact x * Phi(x)
    from math import erf, sqrt
    exact = np.array([v * 0.5 * (1 + erf(v / sqrt(2))) for v in z])
    assert np.max(np.abs(exact - mlp.gelu(z))) < 1e-3


def _problem(n=48, seed=3):
    rng = np.random.default_rng(seed)
    feats = rng.standard_normal((n, 16)).astype(np.float32) * 2 + 0.5
    targs = rng.standard_normal((n, 16)).astype(np.float32) * 1e-3
    targs[:, 12:] = 0
    norm = np.zeros((4, 16), np.float32)
    norm[0] = 0.5
    norm[1] = 2.0
    norm[2] = 0.0
    norm[3] = 1e-3
    p = mlp.init_params(seed, out_scale=1.0)
    return p, norm.ravel(), feats, targs


def test_analytic_gradient_matches_finite_differences():
    p, norm, f, t = _problem()
    g, _, loss = mlp.forward_backward(p, norm, f, t, emulate=False)
    assert abs(loss / f.shape[0] - mlp.loss64(p, norm, f, t)) < 1e-9 * max(1.0, loss)
    rng = np.random.default_rng(7)
    live = [i for i in range(mlp.NPARAM)
            if not (mlp.W3 <= i < mlp.B3 and (i - mlp.W3) % 16 >= mlp.NOUT)
            and not (mlp.B3 + mlp.NOUT <= i)]
    idx = rng.choice(live, 80, replace=False)
    idx = np.concatenate([idx, [mlp.B1 + 5, mlp.B2 + 60, mlp.B3 + 2, mlp.W3 + 16 * 63 + 8]])
    p64 = p.astype(np.float64)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p64[i]))
        a, b = p64.copy(), p64.copy()
        a[i] += h
        b[i] -= h
        fd = (mlp.loss64(a, norm, f, t) - mlp.loss64(b, norm, f, t)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, np.abs(g).max()), (i, fd, g[i])
    # padding parameters get exactly zero gradient
    assert np.all(g[mlp.W3:mlp.B3].reshape(64, 16)[:, mlp.NOUT:] == 0)
    assert np.all(g[mlp.B3 + mlp.NOUT:] == 0)


def test_bf16_emulation_is_close_to_float64():
    p, norm, f, t = _problem(n=256, seed=5)
    g32, y32, l32 = mlp.forward_backward(p, norm, f, t, emulate=True)
    g64, y64, l64 = mlp.forward_backward(p, norm, f, t, emulate=False)
    assert np.abs(y32 - y64).max() < 0.05 * np.abs(y64).max()
    assert np.linalg.norm(g32 - g64) < 0.03 * np.linalg.norm(g64)
    assert abs(l32 - l64) < 0.02 * l64


def test_features_of_a_rotated_stretch():
    """Singular values of Fe = R diag(s) U^T largest first; det Fp; D column-major; dprev."""
    th = 0.3
    r = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    u = np.array([[1, 0, 0], [0, np.cos(0.7), -np.sin(0.7)], [0, np.sin(0.7), np.cos(0.7)]])
    fe = r @ np.diag([1.1, 0.9, 1.05]) @ u.T
    fp = np.diag([1.0, 2.0, 0.5]) * 1.01
    d = np.arange(9.0).reshape(3, 3) * 1e-3
    f = mlp.features(fe, fp, d, [3e-5, -4e-5, 1e-6])
    assert np.allclose(f[:3], [1.1, 1.05, 0.9], atol=1e-12)
    assert abs(f[3] - np.linalg.det(fp)) < 1e-12
    assert np.allclose(f[4:13], d.T.ravel())
    assert np.array_equal(f[13:16], [3e-5, -4e-5, 1e-6])
