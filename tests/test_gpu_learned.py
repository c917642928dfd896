"""The frontier predictor's HIP kernels against the float64 oracle, and the
learned schedule end to end (reports unchanged, predictions measured)."""
import numpy as np
import pytest

import oracle
from oracle import mlp_ref
from ngap_amd import learned, synth
from ngap_amd.mlp import MLP

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def mlp():
    m = MLP(0, seed=3)
    yield m
    m.close()


def test_forward_and_gradient_match_oracle(mlp):
    rng = np.random.default_rng(0)
    p = mlp.params.astype(np.float64)
    for n in (1, 17, 300):
        x = rng.normal(size=(n, 264)).astype(np.float32)
        t = rng.normal(size=n).astype(np.float32)
        say("mlp forward/gradient", n)
        y = mlp.predict(x)
        y_ref, _ = mlp_ref.forward(p, x.astype(np.float64))
        assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4)   # FP32 vs float64
        g, loss = mlp.gradient(x, t)
        g_ref, loss_ref = mlp_ref.gradient(p, x.astype(np.float64), t.astype(np.float64))
        assert np.isclose(loss, loss_ref, rtol=1e-4)
        scale = np.abs(g_ref).max()
        assert np.max(np.abs(g - g_ref)) <= 1e-4 * scale + 1e-6


def test_pair_inference_equals_dense(mlp):
    rng = np.random.default_rng(1)
    hist = rng.random((40, 256)).astype(np.float32)
    hist /= hist.sum(1, keepdims=True)
    stats = rng.random((7, 8)).astype(np.float32)
    pairs = np.stack([rng.integers(0, 7, 500), rng.integers(0, 40, 500)], 1)
    say("mlp pairs")
    y1 = mlp.predict_pairs(hist, stats, pairs)
    y2 = mlp.predict(learned.features(hist, stats, pairs))
    assert np.array_equal(y1, y2)


def test_training_fits_a_known_function():
    rng = np.random.default_rng(2)
    x = rng.random((20000, 264)).astype(np.float32)
    t = (3 * x[:, 0] - 2 * x[:, 260] + x[:, 100] * x[:, 101]).astype(np.float32)
    m = MLP(0, seed=5)
    say("mlp training")
    l0 = float(np.mean((m.predict(x) - t) ** 2))
    loss = m.train(x, t, epochs=20, batch=512, lr=2e-3)
    l1 = float(np.mean((m.predict(x) - t) ** 2))
    assert l1 < 0.05 * l0 and loss < 0.05 * l0
    m.close()


def test_learned_schedule_end_to_end(engine):
    rng = np.random.default_rng(9)
    groups = synth.batch_groups(rng, 48, lo=50, hi=8000)
    chunk = 1 << 14
    train = synth.drifting_input(np.random.default_rng(10), 1 << 20, chunk, 1e-3, groups, 16)
    test = synth.drifting_input(np.random.default_rng(11), 1 << 19, chunk, 1e-3, groups, 16)
    stats = learned.group_stats(groups)
    engine.load_groups(groups, 2)
    say("traces")
    engine.set_input(train, chunk)
    pairs, y = learned.collect_traces(engine, train, chunk, len(groups))
    x = learned.features(learned.window_hist(train, chunk), stats, pairs)
    m = MLP(0, seed=1)
    say("train", x.shape)
    m.train(x, y, epochs=8, batch=2048, lr=2e-3)
    engine.set_input(test, chunk)
    tp, ty = learned.collect_traces(engine, test, chunk, len(groups))
    pred = m.predict_pairs(learned.window_hist(test, chunk), stats, tp)
    mae = float(np.mean(np.abs(pred - ty)))
    base = float(np.mean(np.abs(y.mean() - ty)))   # a predictor without features
    say("mae", mae, "constant predictor", base)
    assert mae < base
    S = len(test) // chunk
    cost = np.zeros((len(groups), S * learned.windows_per_chunk(chunk)))
    cost[tp[:, 0], tp[:, 1]] = learned.predicted_cost(pred)
    waves = engine.batched_waves()
    assert waves > 0
    tasks, tcost = learned.schedule_from_costs(cost, len(groups), S, chunk, waves=waves)
    # the predicted peak sizes the worklist spill area below the static
    # default of 65,536 items per wave (the memory the network saves)
    assert learned.spill_capacity(pred, tcost) < 65536
    want, _ = oracle.run(groups, test, chunk)
    say("learned schedule run")
    st = engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2,
                    unique=True, tasks=tasks, spill_capacity=learned.spill_capacity(pred, tcost),
                    fallback=True)
    assert np.array_equal(engine.reports(), want)
    assert st["num_tasks"] == len(tasks)
    ctasks, ccost = learned.cross_schedule_from_costs(cost, S, chunk, waves=waves)
    say("learned all-group schedule run")
    st = engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2,
                    unique=True, tasks=ctasks, spill_capacity=learned.spill_capacity(pred, ccost),
                    fallback=True)
    assert np.array_equal(engine.reports(), want)
    assert st["num_tasks"] == len(ctasks)
    m.close()
