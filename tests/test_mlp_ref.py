"""The MLP oracle (oracle/mlp_ref.py) pinned by central finite differences,
and the learned-schedule host logic."""
import numpy as np

from ngap_amd import learned
from ngap_amd.groups import make_group
from oracle import mlp_ref


def _params(rng):
    p = []
    for shape, fan in (((264, 64), 264), ((64,), 1), ((64, 64), 64), ((64,), 1), ((64,), 64), ((1,), 1)):
        p.append(rng.uniform(-1, 1, shape).ravel() * np.sqrt(6.0 / fan) * (0.1 if fan == 1 else 1))
    return np.concatenate(p)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    x = rng.normal(size=(12, 264))
    t = rng.normal(size=12)
    g, _ = mlp_ref.gradient(p, x, t)
    idx = np.concatenate([rng.choice(264 * 64, 40, replace=False),
                          264 * 64 + np.arange(64), rng.integers(264 * 64 + 64, p.size, 60),
                          [p.size - 1]])
    eps = 1e-6
    for i in idx:
        a, b = p.copy(), p.copy()
        a[i] += eps
        b[i] -= eps
        fd = (mlp_ref.loss(a, x, t) - mlp_ref.loss(b, x, t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_window_histograms_and_tiling():
    data = np.arange(3 * 1000, dtype=np.int64).astype(np.uint8)
    h = learned.window_hist(data, 1000, 256)
    assert h.shape == (3 * 4, 256)
    assert np.allclose(h.sum(1), 1.0)
    t = learned.fixed_tiling(2, 3, 1000, 256)
    assert t.shape == (2 * 3 * 4, 4)
    # each (group, chunk) tiles [0, 1000)
    for g in range(2):
        for c in range(3):
            rows = t[(t[:, 0] == g) & (t[:, 1] == c)]
            assert rows[0, 2] == 0 and rows[-1, 3] == 1000
            assert np.all(rows[1:, 2] == rows[:-1, 3])
    p = learned.pairs_of(t, 1000, 256)
    assert p.max(0).tolist() == [1, 11]


def test_schedule_merges_cheap_windows_and_tiles_exactly():
    rng = np.random.default_rng(1)
    G, S, L, W = 3, 4, 4096 * 8, 4096
    cost = rng.exponential(1.0, (G, S * 8))
    cost[0] *= 100  # one heavy group
    tasks, tc = learned.schedule_from_costs(cost, G, S, L, waves=4, window=W)
    assert np.all(np.diff(tc) <= 1e-12)  # longest first
    for g in range(G):
        for c in range(S):
            rows = tasks[(tasks[:, 0] == g) & (tasks[:, 1] == c)]
            rows = rows[np.argsort(rows[:, 2])]
            assert rows[0, 2] == 0 and rows[-1, 3] == L
            assert np.all(rows[1:, 2] == rows[:-1, 3])
    heavy = tasks[tasks[:, 0] == 0]
    light = tasks[tasks[:, 0] != 0]
    assert (heavy[:, 3] - heavy[:, 2]).mean() < (light[:, 3] - light[:, 2]).mean()


def test_cross_schedule_splits_heavy_windows_and_tiles_exactly():
    rng = np.random.default_rng(5)
    G, S, L, W = 5, 3, 4096 * 8, 4096
    cost = rng.exponential(1.0, (G, S * 8))
    cost[:, 5] *= 200  # one heavy window in chunk 0
    tasks, tc = learned.cross_schedule_from_costs(cost, S, L, waves=4, window=W)
    assert np.all(tasks[:, 0] == learned.ALL_GROUPS)
    assert np.all(np.diff(tc) <= 1e-9) and np.isclose(tc.sum(), cost.sum())
    for c in range(S):
        rows = tasks[tasks[:, 1] == c]
        rows = rows[np.argsort(rows[:, 2])]
        assert rows[0, 2] == 0 and rows[-1, 3] == L
        assert np.all(rows[1:, 2] == rows[:-1, 3])
    heavy = tasks[(tasks[:, 1] == 0) & (tasks[:, 2] >= 5 * W) & (tasks[:, 3] <= 6 * W)]
    assert len(heavy) > 1 and np.all(heavy[:, 3] - heavy[:, 2] >= 256)


def test_group_stats():
    g = make_group(3, [(0, 0), (0, 1), (1, 2)], [list(range(256)), [1], [2, 3]],
                   all_input=[0], reporting=[2])
    s = learned.group_stats([g])[0]
    assert np.isclose(s[0], np.log1p(3)) and np.isclose(s[1], np.log1p(3))
    assert np.isclose(s[2], np.log1p(1)) and s[3] == 0 and np.isclose(s[4], np.log1p(1))
    assert np.isclose(s[5], 1.0) and np.isclose(s[6], 1 / 3) and np.isclose(s[7], np.log1p(1))


def test_cross_schedule_in_position_order():
    rng = np.random.default_rng(6)
    G, S, L, W = 4, 3, 4096 * 8, 4096
    cost = rng.exponential(1.0, (G, S * 8))
    by_cost, c1 = learned.cross_schedule_from_costs(cost, S, L, waves=4, window=W)
    by_pos, c2 = learned.cross_schedule_from_costs(cost, S, L, waves=4, window=W,
                                                   order="position")
    key = lambda t: sorted(map(tuple, t.tolist()))
    assert key(by_cost) == key(by_pos) and np.isclose(c1.sum(), c2.sum())
    assert np.all(np.diff(by_pos[:, 1].astype(np.int64)) >= 0)
    same_chunk = by_pos[1:, 1] == by_pos[:-1, 1]
    assert np.all(by_pos[1:, 2][same_chunk] == by_pos[:-1, 3][same_chunk])
