"""Full-size workloads against the oracle.  Every full run below is the
engine's own run over the whole input (no truncated copies): its reports are
compared with the oracle's on every chunk (literal64, regex20k, the learned
workload) or on a spread of whole chunks (batch1024).  Inputs whose matches are
too rare to exercise the report path (regex20k and batch1024 over plain
Zipf(1.1) bytes hold 0 and 1 reports) are also run with pattern instances
planted at 1e-4 per byte, so each check holds hundreds of reports."""
import numpy as np
import pytest

import oracle
from ngap_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def in_chunks(reports, chunk, picks):
    """The reports whose offset lies in one of the chunks `picks`."""
    off = (reports & np.uint64(0xFFFFFFFF)).astype(np.int64)
    return reports[np.isin(off // chunk, np.asarray(picks))]


def chunk_mask(num_groups, num_seg, picks):
    m = np.zeros(num_groups * num_seg, np.uint8)
    for c in picks:
        m[c * num_groups:(c + 1) * num_groups] = 1
    return m


def run_configs(engine, configs, want, label, picks=None, chunk=None):
    for cfg in configs:
        cfg = dict(cfg)
        cfg.setdefault("algorithm", "nonblockingallgroups")
        say(label, cfg)
        engine.run(unique=True, **cfg)
        got = engine.reports()
        if picks is not None:
            got = in_chunks(got, chunk, picks)
        assert np.array_equal(got, want), (label, cfg, got.size, want.size)


def test_literal64_full_size_bit_exact(engine):
    groups, data, chunk = synth.literal64()
    want, _ = oracle.run(groups, data, chunk, 1)
    engine.load_groups(groups, 3)
    engine.set_input(data, chunk, 1)
    assert want.size > 20  # 4^-L per position for L = 8..16: tens of matches
    run_configs(engine, [dict(schedule=s, precompute_depth=d) for s in ("plain", "batched")
                         for d in (1, 2, 3)], want, "literal64 full")


REGEX20K_CONFIGS = [dict(schedule="plain", precompute_depth=2),
                    dict(schedule="batched", precompute_depth=3, use_graph=True),
                    dict(schedule="batched", precompute_depth=2, window=4096),
                    dict(schedule="batched", precompute_depth=3, window=1000),
                    dict(schedule="batched", precompute_depth=1, window=65536),
                    dict(schedule="batched", precompute_depth=2, window=4096,
                         algorithm="nonblockingpcgroups")]


def test_regex20k_full_size_every_chunk(engine):
    """20,000 patterns in one group (474k states), 16 MiB in 64 chunks of
    256 KiB: the whole run against the oracle on every chunk, for the
    workload's Zipf(1.1) input and for the same input with planted pattern
    instances."""
    groups, data, chunk = synth.regex20k()
    engine.load_groups(groups, 3)
    want, _ = oracle.run(groups, data, chunk, memo_depth=3)
    engine.set_input(data, chunk)
    run_configs(engine, REGEX20K_CONFIGS, want, "regex20k full")
    planted = synth.plant(np.random.default_rng(3101), data.copy(), groups, 1e-4)
    want, _ = oracle.run(groups, planted, chunk, memo_depth=3)
    say("regex20k planted: oracle reports", want.size)
    assert want.size > 500
    engine.set_input(planted, chunk)
    run_configs(engine, REGEX20K_CONFIGS, want, "regex20k planted full")


def test_batch1024_full_size_spread_chunks(engine):
    """1,024 automata of 50..50,000 states (7.19M), 8 MiB in 128 chunks: the
    whole run, compared with the oracle on 32 whole chunks spread over the
    input (first and last included), for the workload's input (one report in
    8 MiB) and with planted instances (other 32 chunks)."""
    groups, data, chunk = synth.batch1024()
    S = len(data) // chunk
    engine.load_groups(groups, 3)
    configs = [dict(schedule="batched", precompute_depth=3, use_graph=True),
               dict(schedule="batched", precompute_depth=2, cross_groups=0),
               dict(schedule="plain", precompute_depth=3)]
    for label, d, picks in (("batch1024", data, list(range(0, S, 4)) + [S - 1]),
                            ("batch1024 planted",
                             synth.plant(np.random.default_rng(3102), data.copy(), groups, 1e-4),
                             list(range(2, S, 4)))):
        want, _ = oracle.run(groups, d, chunk, S, unit_mask=chunk_mask(len(groups), S, picks),
                             memo_depth=3)
        say(label, "oracle reports on", len(picks), "chunks:", want.size)
        if "planted" in label:
            assert want.size > 100
        engine.set_input(d, chunk)
        run_configs(engine, configs, want, label, picks, chunk)


def test_learned_workload_full_size_every_chunk(engine):
    """The headline workload: 1,024 automata (7.34M states), 8 MiB of
    drifting-Zipf input with planted pattern instances, 128 chunks of 64 KiB.
    The whole run against the oracle on every chunk (about 1,100 reports),
    for the plain engine and the batched engine's schedules."""
    groups, data, chunk = synth.learned()
    engine.load_groups(groups, 3)
    want, _ = oracle.run(groups, data, chunk, memo_depth=3)
    say("learned: oracle reports", want.size)
    assert want.size > 1000
    engine.set_input(data, chunk)
    run_configs(engine, [dict(schedule="plain", precompute_depth=2),
                         dict(schedule="batched", precompute_depth=3, use_graph=True),
                         dict(schedule="batched", precompute_depth=3, window=1024),
                         dict(schedule="batched", precompute_depth=2, cross_groups=0),
                         dict(schedule="batched", precompute_depth=3, root_pass=True)],
                want, "learned full")


def test_serving_workload_full_size(engine):
    """The serving workload's 64 resident automata (1k-200k states, 2.5M in
    all) at depth 3 and the first 64 of its requests (1 KiB-4 MiB, the largest
    3.8 MiB), with pattern instances planted at 1e-3 per byte, queued at once:
    every request's reports equal the oracle's on that request alone."""
    from ngap_amd.serve import Server
    groups = synth.serving_groups()
    say("serving workload: load", sum(g.num_states for g in groups), "states")
    engine.load_groups(groups, 3)
    rng = np.random.default_rng(2005)
    reqs = []
    for _, g, d in synth.serving_requests(seed=2004, n=64):
        d = d.copy()
        synth.plant(rng, d, [groups[g]], 1e-3)
        reqs.append((g, d))
    assert max(len(d) for _, d in reqs) > (3 << 20)
    srv = Server(engine, max_inflight=4, algorithm="nonblockingallgroups", precompute_depth=3,
                 unique=True)
    say("serving workload: 64 requests at once")
    ids = [srv.submit(g, d) for g, d in reqs]
    got = srv.drain(ids)
    srv.close()
    total = 0
    for i, (g, d) in enumerate(reqs):
        want, _ = oracle.run([groups[g]], d, len(d), 1)
        want = want | np.uint64(g << 22 << 32)
        total += want.size
        assert np.array_equal(got[ids[i]][0], want), (i, g, len(d))
    assert total > 1000
