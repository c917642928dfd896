"""Full-size workloads.  literal64 (64 ACGT literals, 1 MiB, one chunk) is
checked bit for bit against the oracle; regex20k (16 MiB, 64 chunks) through
size-independent properties: every schedule, window and algorithm gives the
same report set, and a truncated copy of sampled chunks matches the oracle."""
import sys

import numpy as np
import pytest

import oracle
from ngap_amd import Engine, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_literal64_full_size_bit_exact(engine):
    groups, data, chunk = synth.literal64()
    want, _ = oracle.run(groups, data, chunk, 1)
    engine.load_groups(groups, 3)
    engine.set_input(data, chunk, 1)
    assert want.size > 20  # 4^-L per position for L = 8..16: tens of matches
    for sched in ("plain", "batched"):
        for d in (1, 2, 3):
            say("literal64 full", sched, d)
            engine.run(algorithm="nonblockingallgroups", schedule=sched, precompute_depth=d,
                       unique=True)
            assert np.array_equal(engine.reports(), want)


def test_regex20k_full_size_invariants(engine):
    groups, data, chunk = synth.regex20k()
    engine.load_groups(groups, 3)
    engine.set_input(data, chunk)
    base = None
    for cfg in [dict(schedule="plain", precompute_depth=2),
                dict(schedule="batched", precompute_depth=2, window=4096),
                dict(schedule="batched", precompute_depth=3, window=1000),
                dict(schedule="batched", precompute_depth=1, window=65536),
                dict(schedule="batched", precompute_depth=2, window=4096, algorithm="nonblockingpcgroups")]:
        cfg = dict(cfg)
        cfg.setdefault("algorithm", "nonblockingallgroups")
        say("regex20k full", cfg)
        st = engine.run(unique=True, **cfg)
        got = engine.reports()
        if base is None:
            base, items = got, st["items"]
        assert np.array_equal(got, base)
    # truncated sampled chunks against the oracle
    picks = [0, 31, 63]
    L = 8192
    sample = np.concatenate([data[c * chunk: c * chunk + L] for c in picks])
    want, _ = oracle.run(groups, sample, L, len(picks))
    engine.set_input(sample, L)
    say("regex20k sampled chunks vs oracle")
    engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2)
    assert np.array_equal(engine.reports(), want)


def test_learned_workload_full_size_schedules_agree(engine):
    """1,024 automata (log-uniform 50..50,000 states, a fifth with a leading
    dot-star) over 8 MiB of drifting-Zipf input with planted pattern instances,
    64 KiB chunks: the plain engine, the batched engine with all-group work
    items at depth 3 and with per-group work items at depth 2 give the same
    report set (about a thousand reports); truncated copies of sampled chunks
    match the oracle."""
    groups, data, chunk = synth.learned()
    engine.load_groups(groups, 3)
    engine.set_input(data, chunk)
    base = None
    for cfg in [dict(schedule="plain", precompute_depth=2),
                dict(schedule="batched", precompute_depth=3, use_graph=True),
                dict(schedule="batched", precompute_depth=2, cross_groups=0)]:
        say("learned full", cfg)
        engine.run(algorithm="nonblockingallgroups", unique=True, **cfg)
        got = engine.reports()
        if base is None:
            base = got
            assert got.size > 500
        assert np.array_equal(got, base)
    picks = [0, 64, 127]
    L = 8192
    sample = np.concatenate([data[c * chunk: c * chunk + L] for c in picks])
    want, _ = oracle.run(groups, sample, L, len(picks))
    engine.set_input(sample, L)
    say("learned sampled chunks vs oracle", want.size)
    engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=3)
    assert np.array_equal(engine.reports(), want)
