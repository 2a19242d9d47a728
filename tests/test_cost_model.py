"""Task cost model: the CPU checker pinned by finite differences (CPU), and
the GPU kernels (forward, back-propagation, Adam, captured training) checked
against it, plus what the trained model gains over the flop-count scheduler
on measured task times (GPU)."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(MO.NPARAM)
    p[MO.OW1:MO.OB1] = rng.normal(0, np.sqrt(2 / 7), MO.OB1 - MO.OW1)
    p[MO.OB1:MO.OW2] = rng.normal(0, 0.1, MO.H)
    p[MO.OW2:MO.OB2] = rng.normal(0, np.sqrt(2 / 64), MO.H * MO.H)
    p[MO.OB2:MO.OW3] = rng.normal(0, 0.1, MO.H)
    p[MO.OW3:MO.OB3] = rng.normal(0, 0.125, MO.H)
    p[MO.OB3] = 0.3
    return p


def batch(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.1, 1.0, (n, 7))
    x[:, 6] = rng.integers(0, 2, n)
    t = np.sin(3 * x[:, 0]) * x[:, 3] + 0.5 * x[:, 6] - x[:, 1] * x[:, 4]
    return x, t


def test_oracle_gradient_matches_finite_differences():
    p = rand_params(1)
    x, t = batch(2, 40)
    _, g = MO.loss_grad(p, x, t)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(MO.NPARAM, 60, replace=False),
                          [MO.OB1, MO.OB2, MO.OW3, MO.OB3, MO.OW2 + 5]])
    h = 1e-6
    for k in idx:
        e = np.zeros(MO.NPARAM)
        e[k] = h
        fd = (MO.loss(p + e, x, t) - MO.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(g[k])), (k, fd, g[k])


def test_oracle_planted_gradient_mistake_is_caught():
    """A transposed W2 in the backward pass (a plausible mistake) fails the
    finite-difference check."""
    p = rand_params(4)
    x, t = batch(5, 30)
    W1, b1, W2, b2, w3, b3 = MO.unpack(p)
    y, a1, a2 = MO.forward(p, x)
    dy = 2 * (y - t) / len(t)
    d2 = np.outer(dy, w3) * (a2 > 0)
    d1_bad = (d2 @ W2) * (a1 > 0)  # should be W2.T
    g_bad = (x.T @ d1_bad).ravel()
    k = 3
    e = np.zeros(MO.NPARAM)
    e[k] = 1e-6
    fd = (MO.loss(p + e, x, t) - MO.loss(p - e, x, t)) / 2e-6
    assert abs(fd - g_bad[k]) > 1e-6 * (1 + abs(fd))


def test_oracle_adam_matches_published_update():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    q, m1, m2 = MO.adam_step(p, g, np.zeros(2), np.zeros(2), 1, 0.1)
    # first step: m^ = g, v^ = g^2 -> p - lr * sign(g) (up to eps)
    assert np.allclose(q, p - 0.1 * np.sign(g), atol=1e-7)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_gpu_loss_and_gradient_match_oracle(xt):
    m = xt.CostModel(0, seed=11)
    p = rand_params(6).astype(np.float32)
    m.set_params(p)
    x, t = batch(7, 300)
    loss, g = m.loss_grad(x.astype(np.float32), t.astype(np.float32))
    rl, rg = MO.loss_grad(p.astype(np.float64), x.astype(np.float32).astype(np.float64),
                          t.astype(np.float32).astype(np.float64))
    print(f"[gpu] mlp loss {loss!r} vs {rl!r}; max |dg| {np.max(np.abs(g - rg)):.3e} "
          f"of max |g| {np.max(np.abs(rg)):.3e}", flush=True)
    assert abs(loss - rl) <= 1e-5 * abs(rl)
    assert np.max(np.abs(g - rg)) <= 1e-5 * np.max(np.abs(rg)) + 1e-7
    m.close()


@pytest.mark.gpu
def test_gpu_adam_step_and_determinism(xt):
    p = rand_params(8).astype(np.float32)
    x, t = batch(9, 200)
    x32, t32 = x.astype(np.float32), t.astype(np.float32)
    m = xt.CostModel(0, seed=1)
    m.set_params(p)
    hist = m.train(x32, t32, 1, 1e-3)
    q = m.get_params()
    rl, rg = MO.loss_grad(p.astype(np.float64), x32.astype(np.float64), t32.astype(np.float64))
    rq, _, _ = MO.adam_step(p.astype(np.float64), rg, np.zeros(MO.NPARAM), np.zeros(MO.NPARAM),
                            1, 1e-3)
    assert abs(hist[0] - rl) <= 1e-5 * rl
    # first Adam step moves each parameter by ~lr * sign(g) (where |g| >> eps)
    big = np.abs(rg) > 1e-4
    assert np.max(np.abs(q[big] - rq[big])) <= 2e-6
    # two trainings from the same start give the same bits
    a = xt.CostModel(0, seed=5)
    b = xt.CostModel(0, seed=5)
    ha = a.train(x32, t32, 200, 3e-3)
    hb = b.train(x32, t32, 200, 3e-3)
    assert np.array_equal(a.get_params(), b.get_params()) and np.array_equal(ha, hb)
    assert ha[-1] < 0.2 * ha[0], (ha[0], ha[-1])
    for mm in (m, a, b):
        mm.close()


@pytest.mark.gpu
def test_learned_costs_beat_flop_count(xt):
    """On measured per-task times of an uneven-tile workload, the model
    trained (on the GPU) on half of the tasks predicts the other half's
    times better than the flop count at the best single rate.  Its LPT
    packing of eight queues is reported beside the flop count's (simulated on
    the measured times) but not asserted: with box-level work stealing in the
    persistent kernel the queue order does not change the run time on one GPU
    (measured on the whole O=48, V=384 workload: driver order, flop count and
    model within 0.3%), so the gain the model shows is in its predictions."""
    tiling = xt.Tiling.uneven(24, 160, (6, 20), (16, 48), seed=3)
    ctx = xt.CCSDTriples(0)
    ctx.set_tiling(tiling)
    ctx.synth_upload(20241202)
    ctx.prepare(xt.PREC_MIXED)
    n = ctx.num_tasks()
    _, flops, _ = ctx.flops()
    r = ctx.run(precision=xt.PREC_MIXED, record_task_time=1)
    ms = np.asarray(r["task_ms"], dtype=np.float64)
    assert np.all(ms > 0)
    feats = ctx.task_features(xt.PREC_MIXED)
    tr, te = np.arange(0, n, 2), np.arange(1, n, 2)
    y = np.log(ms) - np.log(flops)
    mu, sd = y[tr].mean(), y[tr].std()
    model = xt.CostModel(0, seed=2)
    hist = model.train(feats[tr], (y[tr] - mu) / sd, 1500, 3e-3)
    pred = mu + sd * model.predict(feats[te]).astype(np.float64)
    rmse_model = np.sqrt(np.mean((pred - y[te]) ** 2))
    rmse_flops = np.sqrt(np.mean((mu - y[te]) ** 2))

    def makespan(cost):
        order = np.argsort(-cost, kind="stable")
        load = np.zeros(8)
        for k in order:
            load[np.argmin(load)] += ms[te][k]
        return load.max()

    ms_model = makespan(flops[te] * np.exp(pred))
    ms_flops = makespan(flops[te])
    print(f"[gpu] {n} tasks, loss {hist[0]:.3f} -> {hist[-1]:.4f}; held-out log-time RMSE "
          f"model {rmse_model:.4f} vs flops {rmse_flops:.4f}; LPT makespan model {ms_model:.2f} "
          f"vs flops {ms_flops:.2f} ms (lower bound {ms[te].sum() / 8:.2f})", flush=True)
    assert rmse_model < 0.5 * rmse_flops
    # the library path: learn_schedule installs costs and the learned schedule
    # gives the analytic schedule's bits
    st = ctx.learn_schedule(model, xt.PREC_MIXED, n_sample=min(128, n), steps=800)
    print(f"[gpu] learn_schedule: {st}", flush=True)
    assert st["fit_rmse_model"] < st["fit_rmse_flops"]
    a = ctx.run(precision=xt.PREC_MIXED, schedule=xt.SCHED_ANALYTIC)
    b = ctx.run(precision=xt.PREC_MIXED, schedule=xt.SCHED_LEARNED)
    assert a["e4"] == b["e4"] and a["e5"] == b["e5"]
    assert np.array_equal(a["task_partials"], b["task_partials"])
    model.close()
    ctx.close()
