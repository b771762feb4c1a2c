"""CPU restatement of the learned-order scorer (test infrastructure only).

The product never imports this module: tests use it as the checker of the
device kernels (ltl-learning-on-gpus_amd/csrc/scorer.hip).  The scorer has no
counterpart in the reference (its candidate order is fixed by the loops of
code/ltli6463.cu:1209-1440); this file restates the network the device
computes, so that the device forward, gradient and Adam step can be checked.

Network: x (24 features, zero-padded to 32) -> 128 (ReLU) -> 128 (ReLU) -> 1.
Parameter vector layout: W1[128][32], b1[128], W2[128][128], b2[128],
w3[128], b3.  Loss: mean squared error mean_i (y_i - t_i)^2.

With bf16=True the arithmetic rounds to BF16 (nearest even) exactly where
the device feeds its matrix cores: the input rows, W1, the first hidden
activation and W2 in the forward pass; dL/dz2 and dL/dz1 in the backward
pass; everything else is float32.  With bf16=False the same formulas run
in float64 without rounding; central finite differences pin that form.
"""
import numpy as np

P_W1 = 0
P_B1 = P_W1 + 128 * 32
P_W2 = P_B1 + 128
P_B2 = P_W2 + 128 * 128
P_W3 = P_B2 + 128
P_B3 = P_W3 + 128
P_N = P_B3 + 1
FEATURES = 24


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(np.shape(a))


def unpack(p):
    return (p[P_W1:P_B1].reshape(128, 32), p[P_B1:P_W2], p[P_W2:P_B2].reshape(128, 128),
            p[P_B2:P_W3], p[P_W3:P_B3], p[P_B3])


def _prep(p, x, bf):
    dt = np.float32 if bf else np.float64
    p = np.asarray(p, dtype=dt)
    x = np.asarray(x, dtype=dt).reshape(-1, FEATURES)
    X = np.zeros((x.shape[0], 32), dtype=dt)
    X[:, :FEATURES] = x
    r = bf16 if bf else (lambda a: a)
    return p, X, r


def forward(p, x, bf16_mode=True, _cache=False):
    """Scores of feature rows x (n, 24)."""
    p, X, r = _prep(p, x, bf16_mode)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xr = r(X)
    z1 = Xr @ r(W1).T + b1
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(W2).T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    if _cache:
        return y, (Xr, z1, h1, z2, h2, W2, w3, r)
    return y


def grad(p, x, t, bf16_mode=True):
    """(gradient over the parameter vector, loss) of mean_i (y_i - t_i)^2."""
    y, (Xr, z1, h1, z2, h2, W2, w3, r) = forward(p, x, bf16_mode, _cache=True)
    t = np.asarray(t, dtype=y.dtype).reshape(-1)
    n = t.shape[0]
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    g = np.zeros(P_N, dtype=y.dtype)
    g[P_W3:P_B3] = h2.T @ dy
    g[P_B3] = dy.sum()
    dz2 = (dy[:, None] * w3[None, :]) * (z2 > 0)
    g[P_B2:P_W3] = dz2.sum(0)
    dz2r = r(dz2)
    g[P_W2:P_B2] = (dz2r.T @ h1).reshape(-1)
    dh1 = dz2r @ r(W2)
    dz1 = dh1 * (z1 > 0)
    g[P_B1:P_W2] = dz1.sum(0)
    g[P_W1:P_B1] = (r(dz1).T @ Xr).reshape(-1)
    return g, loss


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma and Ba, 2015); returns p, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_like(seed):
    """Random parameters in the device initialiser's ranges (for tests)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P_N, dtype=np.float32)
    W1 = rng.uniform(-1, 1, (128, 32)) * np.sqrt(6 / 24)
    W1[:, FEATURES:] = 0
    p[P_W1:P_B1] = W1.reshape(-1)
    p[P_B1:P_W2] = rng.uniform(-0.1, 0.1, 128)
    p[P_W2:P_B2] = rng.uniform(-1, 1, 128 * 128) * np.sqrt(6 / 128)
    p[P_B2:P_W3] = rng.uniform(-0.1, 0.1, 128)
    p[P_W3:P_B3] = rng.uniform(-1, 1, 128) * np.sqrt(3 / 128)
    p[P_B3] = 0.5
    return p
