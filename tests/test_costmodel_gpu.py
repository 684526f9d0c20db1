"""GPU checks of the cost model: the HIP forward, backward and Adam kernels
against the float64 oracle (oracle/mlp_oracle.py), training that lowers the
loss, and -- on the committed timing dataset measured on an MI355X -- the
regret of the trained tuner against grid search and its gain over the fixed
default geometry."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo   # noqa: E402

DATA = os.path.join(ROOT, "primesieve-gpu_amd", "data", "costmodel_timings.npz")

# FP32 kernels vs float64 oracle: relative tolerance on outputs / gradients
RTOL = 2e-4
ATOL = 2e-5


@pytest.fixture(scope="module")
def cm(gpu):
    import costmodel
    return costmodel


def _setup(seed, n=300):
    rng = np.random.default_rng(seed)
    p = (mo.init_params(seed) + rng.normal(0, 0.02, mo.NPARAMS)).astype(np.float32)
    X = rng.normal(0, 1, (n, 6)).astype(np.float32)
    T = rng.normal(0, 1, n).astype(np.float32)
    return p, X, T


def test_forward_matches_oracle(cm):
    p, X, _ = _setup(1)
    net = cm.MLP()
    net.set_params(p)
    y = net.forward_raw(X)
    want, _ = mo.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.allclose(y, want, rtol=RTOL, atol=ATOL), np.abs(y - want).max()


def test_gradients_match_oracle(cm):
    p, X, T = _setup(2)
    net = cm.MLP()
    net.set_params(p)
    g, loss = net.gradients_raw(X, T)
    want, wl = mo.gradients(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - wl) <= 1e-4 * abs(wl) + 1e-6
    scale = np.abs(want).max()
    assert np.allclose(g, want, rtol=RTOL, atol=ATOL * scale), np.abs(g - want).max()


def test_adam_matches_oracle(cm):
    p, X, T = _setup(3)
    net = cm.MLP()
    net.set_params(p)
    rng = np.random.default_rng(5)
    pp = p.astype(np.float64)
    m = np.zeros(mo.NPARAMS)
    v = np.zeros(mo.NPARAMS)
    for t in range(1, 4):
        g = rng.normal(0, 1e-2, mo.NPARAMS).astype(np.float32)
        net.adam_step(g, 1e-3)
        pp, m, v = mo.adam(pp, g.astype(np.float64), m, v, t, 1e-3)
    got = net.get_params()
    assert np.allclose(got, pp, rtol=1e-5, atol=1e-6), np.abs(got - pp).max()


def test_training_lowers_loss(cm):
    rng = np.random.default_rng(6)
    X = rng.normal(0, 1, (4096, 6))
    y = np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 3] + 0.1 * X[:, 4]
    net = cm.MLP(seed=1)
    losses = net.fit(X, y, epochs=30, batch=128, lr=3e-3, seed=2)
    assert losses[-1] < 0.1 * losses[0], losses


@pytest.mark.skipif(not os.path.exists(DATA), reason="timing dataset not collected")
def test_tuner_regret_on_held_out_queries(cm):
    """Train on the training queries' timings, choose a geometry for each of
    the held-out queries, and compare with grid search (the measured best) and
    with the fixed default."""
    d = np.load(DATA, allow_pickle=False)
    qtr, ttr, qte, tte = d["q_train"], d["t_train"], d["q_test"], d["t_test"]
    grid = [tuple(g) for g in d["grid"]]
    qtr = [tuple(int(v) for v in q) for q in qtr]
    qte = [tuple(int(v) for v in q) for q in qte]
    net = cm.MLP(seed=0)
    net.fit(cm.design_matrix(qtr, grid), np.log2(ttr.reshape(-1)), epochs=60, batch=256, lr=2e-3, seed=0)
    chosen, _ = cm.choose(net, qte, grid)
    r = cm.regret(chosen, tte, grid.index(cm.DEFAULT))
    print("regret", r, flush=True)
    assert r["total_regret"] < r["default_total_regret"]
    assert r["gain_over_default"] > 1.0


def test_runtime_grid_search_on_a_batch(cm, gpu, golden):
    """The fallback tuner times a seeded sample of a batch under each tuning
    of a (segment bytes, waves, bucket threshold) grid and returns the
    fastest; a batch run under the chosen tuning keeps the reference's
    counts."""
    q = np.load(os.path.join(ROOT, "tests", "golden", "config3_queries.npy"), allow_pickle=False)
    exp = golden("config3_expected.json")
    idx = sorted(int(i) for i in exp)[:96]
    qs = [(int(q[i, 0]), int(q[i, 1]), int(q[i, 2])) for i in idx]
    grid = [(131072, 16, 7), (131072, 16, 4), (65536, 8, 12), (163840, 16, 0)]
    best, times = cm.grid_search_batch(qs, grid=grid, sample=48, seed=3, reps=1)
    print("grid search", best, times, flush=True)
    assert best in grid and set(times) == set(grid) and all(t > 0 for t in times.values())
    cm.apply_tuning(best)
    try:
        t = gpu.get_tuning()
        # segment bytes are rounded down to what fits in LDS beside the segment's tail
        assert t["work_group_size"] == 64 * best[1] and best[0] - 4096 < t["sieve_bytes"] <= best[0]
        assert gpu.count_batch(qs) == [exp[str(i)] for i in idx]
    finally:
        gpu.set_tuning()


def test_plan_batch_confirms_the_model_by_timing(cm, gpu, golden):
    """The model's batch tuning is kept only when timing a sample of the
    batch confirms it against the default geometry; the packing order is a
    permutation and the packed batch keeps the reference's counts."""
    q = np.load(os.path.join(ROOT, "tests", "golden", "config3_queries.npy"), allow_pickle=False)
    exp = golden("config3_expected.json")
    idx = sorted(int(i) for i in exp)[:64]
    qs = [(int(q[i, 0]), int(q[i, 1]), int(q[i, 2])) for i in idx]
    net = cm.MLP(seed=4)
    proposal, _ = cm.choose_for_batch(net, [(a, b - a + 1, k) for a, b, k in qs])
    (sb, waves), order, times = cm.plan_batch(net, qs, sample=32, seed=1, reps=1)
    print("proposal", proposal, "chosen", (sb, waves), times, flush=True)
    assert sorted(order.tolist()) == list(range(len(qs)))
    if proposal == cm.DEFAULT:
        assert (sb, waves) == cm.DEFAULT and times == {}
    else:
        pick = proposal + (cm.DEFAULT3[2],)
        assert set(times) == {cm.DEFAULT3, pick}
        assert (sb, waves) == min(times, key=times.get)[:2]
    gpu.set_tuning(sieve_bytes=sb, work_group_size=64 * waves)
    try:
        got = gpu.count_batch([qs[i] for i in order])
        assert [got[j] for j in np.argsort(order)] == [exp[str(i)] for i in idx]
    finally:
        gpu.set_tuning()
