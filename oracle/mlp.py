"""Oracle of the learned PBMPM warm start — TEST INFRASTRUCTURE ONLY (tests/, bench.py).

The warm-start network has no counterpart in the reference (it is an addition on top of the
PBMPM solver of src/lib/gpu/mpm/pbmpm.wgsl), so this module restates the network from its
definition, in numpy:

  x  = standardise(features)            features = [singular values of Fe, largest first (3), det Fp,
                                                    D (9, column-major), dprev (3)]
  z1 = x  W1 + b1,  h1 = gelu(z1)       W1 [16][64]
  z2 = h1 W2 + b2,  h2 = gelu(z2)       W2 [64][64]
  y  = h2 W3 + b3                        W3 [64][16]  (outputs 12..15 are padding)
  loss = sum over samples and outputs 0..11 of (y - standardise(target))^2
  targets = [D change (9, column-major), dx change (3)]

gelu is the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).

Two evaluations:
  - emulate=True mirrors the device arithmetic: every matrix-core operand (x, h1, h2, W1..W3 and
    the back-propagated dY, dZ2, dZ1) is rounded to bfloat16 (round to nearest even) and the
    products accumulate in float32; biases, GELU and the bias gradients stay in float32.
  - emulate=False is the same network in float64 with no rounding.
Pinned by: central finite differences of the float64 loss (analytic gradient), closed values of
gelu / gelu' and of the bfloat16 rounding (tests/test_mlp_oracle.py).
"""
import numpy as np

NIN, NOUT, H = 16, 12, 64
W1, B1 = 0, 16 * 64
W2 = B1 + 64
B2 = W2 + 64 * 64
W3 = B2 + 64
B3 = W3 + 64 * 16
NPARAM = B3 + 16


def bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params):
    p = np.asarray(params)
    return (p[W1:B1].reshape(16, 64), p[B1:W2], p[W2:B2].reshape(64, 64), p[B2:W3],
            p[W3:B3].reshape(64, 16), p[B3:NPARAM])


def pack(w1, b1, w2, b2, w3, b3):
    return np.concatenate([np.ravel(a) for a in (w1, b1, w2, b2, w3, b3)]).astype(np.float32)


def gelu(z):
    t = np.tanh(0.7978845608 * (z + 0.044715 * z * z * z))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    t = np.tanh(0.7978845608 * (z + 0.044715 * z * z * z))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * 0.7978845608 * (1.0 + 0.134145 * z * z)


def init_params(seed=0, out_scale=0.1):
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((16, 64)) / np.sqrt(NIN)
    w1[NIN:] = 0.0
    w2 = rng.standard_normal((64, 64)) / 8.0
    w3 = rng.standard_normal((64, 16)) / 8.0 * out_scale
    w3[:, NOUT:] = 0.0
    b1 = 0.1 * rng.standard_normal(64)
    b2 = 0.1 * rng.standard_normal(64)
    b3 = np.zeros(16)
    b3[:NOUT] = 0.05 * rng.standard_normal(NOUT)
    return pack(w1, b1, w2, b2, w3, b3)


def standardise(feats, targs, norm):
    nm = np.asarray(norm, np.float32).reshape(4, 16)
    x = (np.asarray(feats, np.float32) - nm[0]) / nm[1]
    x[:, NIN:] = 0.0
    t = (np.asarray(targs, np.float32) - nm[2]) / nm[3]
    return x.astype(np.float32), t.astype(np.float32)


def forward_backward(params, norm, feats, targs, emulate=True):
    """Returns (gradient of the mean-over-samples loss [NPARAM], y [n][16], sum of squared errors)."""
    x, t = standardise(feats, targs, norm)
    n = x.shape[0]
    if emulate:
        ft, rnd = np.float32, bf16
    else:
        ft, rnd = np.float64, (lambda a: np.asarray(a, np.float64))
        x, t = x.astype(np.float64), t.astype(np.float64)
    w1, b1, w2, b2, w3, b3 = (np.asarray(a, ft) for a in unpack(params))
    w1r, w2r, w3r = rnd(w1).astype(ft), rnd(w2).astype(ft), rnd(w3).astype(ft)
    xr = rnd(x).astype(ft)
    z1 = (xr @ w1r).astype(ft) + b1
    h1 = rnd(gelu(z1)).astype(ft)
    z2 = (h1 @ w2r).astype(ft) + b2
    h2 = rnd(gelu(z2)).astype(ft)
    y = (h2 @ w3r).astype(ft) + b3
    e = (y - t)
    e[:, NOUT:] = 0.0
    loss = float(np.sum(e.astype(np.float64) ** 2))
    dy = (2.0 * e * ft(1.0 / n)).astype(ft)
    dyr = rnd(dy).astype(ft)
    g_w3 = h2.T @ dyr
    g_b3 = dy.sum(0)
    dz2 = (dyr @ w3r.T).astype(ft) * gelu_grad(z2).astype(ft)
    dz2r = rnd(dz2).astype(ft)
    g_w2 = h1.T @ dz2r
    g_b2 = dz2.sum(0)
    dz1 = (dz2r @ w2r.T).astype(ft) * gelu_grad(z1).astype(ft)
    dz1r = rnd(dz1).astype(ft)
    g_w1 = xr.T @ dz1r
    g_b1 = dz1.sum(0)
    grad = np.concatenate([g_w1.ravel(), g_b1, g_w2.ravel(), g_b2, g_w3.ravel(), g_b3])
    return grad, y, loss


def loss64(params, norm, feats, targs):
    """Mean-over-samples loss of the float64 network (finite-difference target)."""
    x, t = standardise(feats, targs, norm)
    w1, b1, w2, b2, w3, b3 = (np.asarray(a, np.float64) for a in unpack(params))
    y = gelu(gelu(x.astype(np.float64) @ w1 + b1) @ w2 + b2) @ w3 + b3
    e = (y - t)[:, :NOUT]
    return float(np.sum(e * e) / x.shape[0])


def features(fe, fp, d, dprev):
    """Network input of one particle (3x3 matrices as numpy [row][col]; stored column-major)."""
    s = np.linalg.svd(np.asarray(fe, np.float64), compute_uv=False)      # largest first
    f = np.zeros(16)
    f[0:3] = s
    f[3] = np.linalg.det(np.asarray(fp, np.float64))
    f[4:13] = np.asarray(d, np.float64).T.ravel()     # column-major
    f[13:16] = dprev
    return f
