"""Runtime features on the GPU: graph replays, host work items, overflow
fallback, the request server (alone == in the mix == oracle)."""
import numpy as np
import pytest

import oracle
from ngap_amd import learned, NgapError, synth
from ngap_amd.groups import make_group
from ngap_amd.serve import Server, run_poisson

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _workload(seed=21, n=12):
    rng = np.random.default_rng(seed)
    groups = [synth.regex_group(rng, int(s), dotstar=(i % 4 == 0))
              for i, s in enumerate(np.exp(rng.uniform(np.log(200), np.log(5000), n)))]
    data = synth.zipf_bytes(rng, 1 << 17, 0.9)
    return groups, data


def test_graph_replays_match_eager(engine):
    groups, data = _workload()
    want, _ = oracle.run(groups, data, 1 << 13)
    engine.load_groups(groups, 2)
    engine.set_input(data, 1 << 13)
    eager = dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2,
                 unique=True, window=2048)
    say("eager")
    t_eager = min(engine.run(**eager)["kernel_ms"] for _ in range(3))
    assert np.array_equal(engine.reports(), want)
    st0 = engine.run(**eager)
    opts = engine.options(use_graph=True, **eager)
    for rep in range(4):
        say("graph replay", rep)
        st = engine.run(opts)
        t_graph = st["kernel_ms"]
        assert np.array_equal(engine.reports(), want)
        # the counter reset is part of the graph: every replay counts from zero
        assert st["items"] == st0["items"] > 0 and st["reports_raw"] == st0["reports_raw"]
        # each replay is timed itself: not a stale event from an earlier launch
        assert 0 < t_graph < 4 * t_eager + 0.5, (t_graph, t_eager)
        # eager work between replays (a slower engine: a stale time would show)
        engine.run(algorithm="nonblockinggroups", schedule="plain")
        assert np.array_equal(engine.reports(), want)


def test_host_work_items(engine):
    groups, data = _workload(seed=22)
    L = 1 << 13
    S = len(data) // L
    want, _ = oracle.run(groups, data, L)
    engine.load_groups(groups, 3)
    engine.set_input(data, L)
    rng = np.random.default_rng(0)
    tasks = []
    for c in range(S):
        for g in range(len(groups)):
            cuts = sorted(set([0, L] + list(rng.integers(1, L, rng.integers(0, 5)))))
            tasks += [(g, c, a, b) for a, b in zip(cuts, cuts[1:])]
    order = rng.permutation(len(tasks))
    tasks = [tasks[i] for i in order]
    say("random tilings, random order")
    st = engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=3,
                    tasks=tasks)
    assert st["num_tasks"] == len(tasks)
    assert np.array_equal(engine.reports(), want)
    assert engine.tasks().shape == (len(tasks), 4)
    with pytest.raises(NgapError):   # traces only when asked for
        engine.task_items()
    say("random tilings with traces")
    st = engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=3,
                    tasks=tasks, task_traces=True)
    assert np.array_equal(engine.reports(), want)
    items = engine.task_items()
    assert items.shape == (len(tasks),) and int(items.sum()) == st["items"]
    # the same options again: the run returns a schedule id that the next
    # run passes back (no comparison of the work items), with equal reports
    say("schedule id reuse")
    opts = engine.options(algorithm="nonblockingallgroups", schedule="batched",
                          precompute_depth=3, tasks=tasks)
    first = engine.run(opts)["tasks_id"]
    assert first != 0 and opts.tasks_id == first
    for _ in range(2):
        assert engine.run(opts)["tasks_id"] == first
        assert np.array_equal(engine.reports(), want)
    # new input retires the id: the next run compares and gets a new one
    engine.set_input(data, L)
    assert engine.run(opts)["tasks_id"] != first
    assert np.array_equal(engine.reports(), want)
    # a gap or an overlap is rejected before anything runs
    say("bad tilings")
    with pytest.raises(NgapError):
        engine.run(algorithm="nonblockingallgroups", schedule="batched", tasks=tasks[1:])
    with pytest.raises(NgapError):
        engine.run(algorithm="nonblockingallgroups", schedule="batched", tasks=tasks + tasks[:1])


def test_all_group_work_items(engine):
    """Chunks tiled by all-group work items (roots of every group from the
    cross-group index) beside chunks tiled per group, in one launch."""
    groups, data = _workload(seed=23)
    L = 1 << 13
    S = len(data) // L
    want, _ = oracle.run(groups, data, L)
    engine.load_groups(groups, 3)
    engine.set_input(data, L)
    rng = np.random.default_rng(1)
    ALL = -1
    tasks = []
    for c in range(S):
        if c % 2 == 0:
            cuts = sorted(set([0, L] + list(rng.integers(1, L, rng.integers(0, 6)))))
            tasks += [(ALL, c, a, b) for a, b in zip(cuts, cuts[1:])]
        else:
            tasks += [(g, c, 0, L) for g in range(len(groups))]
    tasks = [tasks[i] for i in rng.permutation(len(tasks))]
    for d in (2, 3):
        say("all-group work items, depth", d)
        st = engine.run(algorithm="nonblockingallgroups", schedule="batched",
                        precompute_depth=d, unique=True, tasks=tasks)
        assert st["num_tasks"] == len(tasks)
        assert np.array_equal(engine.reports(), want)
    done = engine.tasks()
    assert int(np.sum(done[:, 0] == 0xFFFFFFFF)) == sum(1 for t in tasks if t[0] == ALL)
    say("all-group work items rejected")
    with pytest.raises(NgapError):   # depth 1 has no cross-group index
        engine.run(algorithm="nonblockinggroups", schedule="batched", tasks=tasks)
    with pytest.raises(NgapError):   # a chunk mixing both kinds
        engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2,
                   tasks=tasks + [(0, 0, 0, L)])


def test_schedule_cache_follows_input_and_groups(engine):
    groups, data = _workload(seed=25)
    L = 1 << 13
    engine.load_groups(groups, 2)
    for k, d in enumerate([data, data[::-1].copy(), data]):
        want, _ = oracle.run(groups, d, L)
        engine.set_input(d, L)
        for x in (0, 1):
            say("schedule cache, input", k, "cross", x)
            engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2,
                       cross_groups=x)
            assert np.array_equal(engine.reports(), want)
    g2 = groups[:5]
    want, _ = oracle.run(g2, data, L)
    engine.load_groups(g2, 2)
    engine.set_input(data, L)
    say("schedule cache, fewer groups")
    engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2)
    assert np.array_equal(engine.reports(), want)


def test_ring_capacity_changes_nothing_but_speed(engine):
    groups, data = _workload(seed=26)
    L = 1 << 13
    want, _ = oracle.run(groups, data, L)
    engine.load_groups(groups, 3)
    engine.set_input(data, L)
    for ring in (64, 128, 2048):
        for x in (0, 1):
            say("ring capacity", ring, "cross", x)
            engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=3,
                       ring_capacity=ring, cross_groups=x, spill_capacity=1 << 16)
            assert np.array_equal(engine.reports(), want)
    # the plain and blocking launches keep their own ring whatever the option says
    engine.run(algorithm="blockinggroups", schedule="batched", ring_capacity=2048)
    assert np.array_equal(engine.reports(), want)
    with pytest.raises(NgapError):
        engine.run(algorithm="nonblockingallgroups", schedule="batched", ring_capacity=96)


def test_root_pass_changes_nothing_but_speed(engine):
    groups, data = _workload(seed=27)
    L = 1 << 13
    want, _ = oracle.run(groups, data, L)
    engine.load_groups(groups, 3)
    engine.set_input(data, L)
    for depth in (1, 2, 3):
        for x in (0, 1):
            for cap in (0, 1):
                say("root pass depth", depth, "cross", x, "capacity", cap)
                engine.run(algorithm="nonblockingallgroups", schedule="batched",
                           precompute_depth=depth, cross_groups=x, root_pass=True,
                           root_capacity=cap, unique=True, window=1500)
                assert np.array_equal(engine.reports(), want)
    # graph replays of a root-pass run
    opts = engine.options(algorithm="nonblockingallgroups", schedule="batched",
                          precompute_depth=3, root_pass=True, use_graph=True)
    for rep in range(3):
        say("root pass graph replay", rep)
        engine.run(opts)
        assert np.array_equal(engine.reports(), want)
        engine.run(algorithm="nonblockinggroups", schedule="plain")
    # root areas far too small (16 all-input states match every byte): the
    # work items are re-run without a root pass even with the fallback off
    lg, ldata, _ = synth.literal64(input_len=1 << 16)
    lwant, _ = oracle.run(lg, ldata, 1 << 14)
    engine.load_groups(lg, 1)
    engine.set_input(ldata, 1 << 14)
    say("root areas overflow")
    st = engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=1,
                    root_pass=True, root_capacity=1, fallback=False)
    assert st["fallback_tasks"] > 0
    assert np.array_equal(engine.reports(), lwant)


def test_autotuner_picks_knobs_not_results(engine):
    groups, data = _workload(seed=24)
    L = 1 << 13
    want, _ = oracle.run(groups, data, L)
    engine.load_groups(groups, 2)
    engine.set_input(data, L)
    say("autotune")
    best, ms = engine.autotune(max_runs=10, algorithm="nonblockingallgroups",
                               precompute_depth=2, unique=True)
    assert ms > 0 and best.schedule == 1 and best.fallback == 1
    assert best.window in (0, 256, 512, 1024, 2048, 4096, 8192)
    engine.run(best)
    assert np.array_equal(engine.reports(), want)


def test_overflow_fallback_restores_exact_reports(engine):
    n = 1 + 4 + 16 + 64
    edges = [(0, 1 + i) for i in range(4)] + [(1 + i, 5 + 4 * i + j) for i in range(4) for j in range(4)]
    edges += [(5 + i, 21 + 4 * i + j) for i in range(16) for j in range(4)]
    g = make_group(n, edges, [[ord("a")]] * n, all_input=[0], reporting=list(range(21, n)))
    data = (b"a" * 3000 + b"b") * 4
    want, _ = oracle.run([g], data, len(data) // 4, 4)
    engine.load_groups([g], 1)
    engine.set_input(data, len(data) // 4, 4)
    say("fallback")
    st = engine.run(algorithm="nonblockinggroups", schedule="batched", spill_capacity=64,
                    window=1024, fallback=True)
    assert st["fallback_tasks"] > 0
    assert np.array_equal(engine.reports(), want)
    # the caller's work items (as a learned schedule passes them) with
    # worklists sized too small: the same fallback, the same reports
    say("fallback, caller work items")
    tasks = learned.fixed_tiling(1, 4, len(data) // 4, window=512)
    st = engine.run(algorithm="nonblockinggroups", schedule="batched", spill_capacity=64,
                    ring_capacity=64, tasks=tasks, fallback=True)
    assert st["fallback_tasks"] > 0 and st["num_tasks"] == len(tasks)
    assert np.array_equal(engine.reports(), want)


def test_server_alone_equals_mix_equals_oracle(engine):
    groups = [synth.regex_group(np.random.default_rng(40 + i), 2000 + 1500 * i,
                                dotstar=(i == 2)) for i in range(5)]
    engine.load_groups(groups, 2)
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(24):
        L = int(np.exp(rng.uniform(np.log(1024), np.log(1 << 17))))
        reqs.append((int(rng.integers(0, len(groups))), synth.zipf_bytes(rng, L, 0.8)))
    # oracle: each request as one chunk of its own group (group index 0 here,
    # fixed up to the real group id)
    expect = []
    for g, d in reqs:
        r, _ = oracle.run([groups[g]], d, len(d), 1)
        expect.append(r | np.uint64(g << 22 << 32))
    srv = Server(engine, max_inflight=3, algorithm="nonblockingallgroups", precompute_depth=2,
                 unique=True)
    say("server: requests alone")
    alone = []
    for g, d in reqs:
        rid = srv.submit(g, d)
        alone.append(srv.drain([rid])[rid][0])
    say("server: requests in the mix")
    ids = [srv.submit(g, d) for g, d in reqs]
    mix = srv.drain(ids)
    for i, (g, d) in enumerate(reqs):
        assert np.array_equal(alone[i], expect[i]), i
        assert np.array_equal(mix[ids[i]][0], expect[i]), i
    say("server: poisson replay")
    trace = [(0.002 * i, g, d) for i, (g, d) in enumerate(reqs)]
    res = run_poisson(srv, trace)
    assert res["requests"] == len(reqs) and res["p99_ms"] >= res["p50_ms"] > 0
    for i in range(len(reqs)):
        assert np.array_equal(res["reports"][i], expect[i])
    srv.close()


def test_server_budget_binds_and_results_stay_exact(engine):
    """A 64 MiB HBM budget: launches never hold more than the budget, queued
    requests start as earlier launches free their memory, a request larger
    than the budget alone fails, and every other result equals the oracle."""
    groups = [synth.regex_group(np.random.default_rng(60 + i), 1500 + 1000 * i,
                                dotstar=(i == 1)) for i in range(4)]
    engine.load_groups(groups, 3)
    rng = np.random.default_rng(61)
    reqs = []
    for i in range(32):
        L = int(np.exp(rng.uniform(np.log(8 << 10), np.log(1 << 20))))
        reqs.append((int(rng.integers(0, len(groups))), synth.zipf_bytes(rng, L, 1.1)))
    expect = []
    for g, d in reqs:
        r, _ = oracle.run([groups[g]], d, len(d), 1)
        expect.append(r | np.uint64(g << 22 << 32))
    budget = 64 << 20
    srv = Server(engine, budget_bytes=budget, max_inflight=4, algorithm="nonblockingallgroups",
                 precompute_depth=3, unique=True)
    say("server budget: oversized request")
    big = srv.submit(0, synth.zipf_bytes(rng, 40 << 20, 1.1))
    say("server budget: 32 requests at once")
    ids = [srv.submit(g, d) for g, d in reqs]
    got = srv.drain(ids)
    st = srv.stats()
    say("server budget stats", st)
    for _ in range(100):
        state, rep, _ = srv.take(big)
        if state == 3:
            break
        srv.poll(block=True)
    assert state == 3, "a request larger than the budget fails"
    assert st["rejected"] == 1
    assert 0 < st["peak_held_bytes"] <= budget
    assert st["peak_inflight_bytes"] <= budget
    # about 17 MiB queued at once do not fit one launch under this budget
    assert st["launches"] >= 3
    assert st["redone"] == 0
    for i in range(len(reqs)):
        assert np.array_equal(got[ids[i]][0], expect[i]), i
    srv.close()


def test_server_rejects_report_off(engine):
    groups = [synth.regex_group(np.random.default_rng(70), 500)]
    engine.load_groups(groups, 2)
    with pytest.raises(NgapError):
        Server(engine, report_off=True, algorithm="nonblockingallgroups", precompute_depth=2)
