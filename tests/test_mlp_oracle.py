"""Pins the MLP oracle (oracle/mlp_oracle.py) on the CPU: its gradients
against central finite differences of its own loss, and its Adam step
against a hand-evaluated update."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as mo   # noqa: E402


def test_param_count():
    assert mo.NPARAMS == 4673


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(4)
    p = mo.init_params(1) + rng.normal(0, 0.05, mo.NPARAMS)
    X = rng.normal(0, 1, (37, 6))
    T = rng.normal(0, 1, 37)
    g, l = mo.gradients(p, X, T)
    assert abs(l - mo.loss(p, X, T)) < 1e-12
    idx = rng.choice(mo.NPARAMS, 120, replace=False)
    h = 1e-6
    for i in list(idx) + [0, 383, 384, 447, 448, 4543, 4544, 4607, 4608, 4671, 4672]:
        e = np.zeros(mo.NPARAMS)
        e[i] = h
        fd = (mo.loss(p + e, X, T) - mo.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_adam_single_step_by_hand():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.1, -0.3, 0.0])
    m0 = np.zeros(3)
    v0 = np.zeros(3)
    p1, m1, v1 = mo.adam(p, g, m0, v0, 1, lr=0.01)
    # step 1 from zero moments: mhat = g, vhat = g^2 -> update = lr * g / (|g| + eps)
    want = p - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, want, rtol=0, atol=1e-12)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)


def test_pack_unpack_layout():
    p = np.arange(mo.NPARAMS, dtype=np.float64)
    d = mo.unpack(p)
    assert d["W1"][1, 0] == 6 and d["b1"][0] == 384 and d["W2"][0, 1] == 449
    assert d["b2"][0] == 4544 and d["W3"][0] == 4608 and d["b3"][0] == 4672
    assert np.array_equal(mo.pack(d), p)
