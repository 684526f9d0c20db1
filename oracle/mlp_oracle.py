"""CPU oracle of the FP32 cost-model MLP -- TEST INFRASTRUCTURE ONLY.

float64 numpy restatement of the 6 -> 64 -> 64 -> 1 network (ReLU, ReLU,
linear), its mean-squared-error gradients and one bias-corrected Adam step,
with the flat parameter layout of include/primesieve_gpu.h:
W1[64][6], b1[64], W2[64][64], b2[64], W3[64], b3.  Only tests/ and the
benchmark's checks use it, as the checker of the HIP kernels.

There is no reference counterpart (the reference has no learned component);
the oracle is pinned by central finite differences of its own loss and by
a hand-evaluated Adam step (tests/test_mlp_oracle.py).
"""
import numpy as np

NIN, H = 6, 64
SIZES = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)   # 4673


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def init_params(seed=0):
    """He-style initialisation."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / NIN), (H, NIN)), "b1": np.zeros(H),
         "W2": rng.normal(0, np.sqrt(2.0 / H), (H, H)), "b2": np.zeros(H),
         "W3": rng.normal(0, np.sqrt(1.0 / H), H), "b3": np.zeros(1)}
    return pack(d)


def forward(p, X):
    d = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ d["W1"].T + d["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ d["W2"].T + d["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ d["W3"] + d["b3"][0]
    return y, (X, a1, h1, a2, h2)


def loss(p, X, T):
    y, _ = forward(p, X)
    return float(np.mean((y - np.asarray(T, dtype=np.float64)) ** 2))


def gradients(p, X, T):
    """Gradients of the mean squared error, and the loss."""
    d = unpack(p)
    y, (X, a1, h1, a2, h2) = forward(p, X)
    T = np.asarray(T, dtype=np.float64)
    n = X.shape[0]
    dy = 2.0 * (y - T) / n
    g = {"W3": h2.T @ dy, "b3": np.array([dy.sum()])}
    g2 = np.outer(dy, d["W3"]) * (a2 > 0)
    g["W2"] = g2.T @ h1
    g["b2"] = g2.sum(0)
    g1 = (g2 @ d["W2"]) * (a1 > 0)
    g["W1"] = g1.T @ X
    g["b1"] = g1.sum(0)
    return pack(g), float(np.mean((y - T) ** 2))


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (step number t >= 1)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** t)
    vhat = v / (1 - b2 ** t)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v
