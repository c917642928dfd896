"""HIP engine vs the oracle: bit-exact sorted unique report sets.

Every case prints its name before touching the GPU.  Sizes keep the oracle
to seconds; the full-size workloads are checked through properties in
test_gpu_fullsize.py.
"""
import os
import random
import sys

import numpy as np
import pytest

import oracle
from ngap_amd import Engine, Frontend, NgapError, synth
from ngap_amd.groups import make_group
from tests.conftest import GOLDEN
from tests import nfa_helpers as H

pytestmark = pytest.mark.gpu

ALGOS = ["blockinggroups", "NAPgroups", "nonblockinggroups", "nonblockingr1groups", "nonblockingr2groups",
         "nonblockingpcgroups", "nonblockingallgroups"]


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def check_all(engine, groups, data, chunk, num_seg=None, configs=None, is_dup=False, depth=3):
    want, _ = oracle.run(groups, data, chunk, num_seg, is_dup=is_dup)
    engine.load_groups(groups, max_precompute_depth=depth)
    engine.set_input(data, chunk, num_seg)
    configs = configs or [dict(algorithm=a) for a in ALGOS]
    for cfg in configs:
        cfg = dict(cfg)
        cfg.setdefault("is_dup", is_dup)
        say(cfg)
        engine.run(**cfg)
        got = engine.reports()
        assert got.shape == want.shape and np.array_equal(got, want), \
            (cfg, got.size, want.size)
    return want


def test_apple(engine):
    g = Frontend.load(os.path.join(GOLDEN, "apple.anml"), 1).groups()
    data = open(os.path.join(GOLDEN, "apple_input.txt"), "rb").read()
    cfgs = []
    for a in ALGOS:
        for sched in ("plain", "batched"):
            for d in (0, 1, 2, 3):
                cfgs.append(dict(algorithm=a, schedule=sched, precompute_depth=d, unique=d % 2))
    want = check_all(engine, g, data, 81, 1, cfgs)
    assert [hex(x) for x in want] == ["0x400000005", "0x40000002f", "0x400000040", "0x40000004b"]


def test_literal64_small(engine):
    groups, data, _ = synth.literal64(input_len=1 << 16)
    cfgs = [dict(algorithm=a, schedule=s, precompute_depth=d)
            for a in ("nonblockinggroups", "nonblockingpcgroups", "nonblockingallgroups")
            for s in ("plain", "batched") for d in (1, 2, 3)]
    check_all(engine, groups, data, 1 << 16, 1, cfgs)


@pytest.mark.parametrize("seed", range(3))
def test_regex_patterns_multi_chunk(engine, seed):
    rng = np.random.default_rng(seed)
    g = synth.regex_group(rng, 20000, n_patterns=800)
    data = synth.zipf_bytes(rng, 1 << 18, 1.1)
    cfgs = [dict(algorithm=a, schedule=s, precompute_depth=d, unique=u, window=w)
            for a, s, d, u, w in [("nonblockingallgroups", "plain", 2, 1, 0),
                                  ("nonblockingallgroups", "batched", 2, 1, 4096),
                                  ("nonblockingallgroups", "batched", 3, 0, 1000),
                                  ("nonblockingpcgroups", "plain", 3, 1, 0),
                                  ("nonblockingr1groups", "batched", 1, 0, 8192),
                                  ("NAPgroups", "plain", 0, 0, 0)]]
    # the reference's scheduling knobs never change the report set
    cfgs += [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=3,
                  data_buffer_fetch_size=25600, add_aas_start=256, add_aas_interval=256,
                  active_threshold=t) for s, t in (("plain", 16), ("batched", 0))]
    # one group through the cross-group index (forced)
    cfgs += [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=d,
                  unique=1, cross_groups=2, window=w) for d, w in ((2, 0), (3, 777))]
    check_all(engine, [g], data, 1 << 14, None, cfgs)


def test_many_groups_with_dot_star(engine):
    rng = np.random.default_rng(11)
    groups = [synth.regex_group(rng, int(s), dotstar=(i % 5 == 0))
              for i, s in enumerate(np.exp(rng.uniform(np.log(50), np.log(3000), 40)))]
    data = synth.zipf_bytes(rng, 1 << 17, 1.1)
    cfgs = [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=d, unique=1)
            for s in ("plain", "batched") for d in (1, 2, 3)]
    # batched with one group per work item, and cross-group work items with
    # forced windows
    cfgs += [dict(algorithm=a, schedule="batched", precompute_depth=d, unique=u,
                  cross_groups=x, window=w)
             for a, d, u, x, w in [("nonblockingallgroups", 2, 1, 0, 0),
                                   ("nonblockingallgroups", 3, 0, 0, 2048),
                                   ("nonblockingallgroups", 2, 0, 1, 256),
                                   ("nonblockingpcgroups", 3, 1, 1, 1000),
                                   ("nonblockingallgroups", 3, 1, 2, 8192)]]
    check_all(engine, groups, data, 1 << 13, None, cfgs)


@pytest.mark.parametrize("seed", range(3))
def test_class_repeat_chains(engine, seed):
    r = random.Random(200 + seed)
    alpha = list(b"abcd")
    specs = [H.random_spec(r, r.randint(2, 6), alpha) for _ in range(30)]
    g, _ = H.chains_group(specs)
    data = bytes(r.choice(alpha) for _ in range(5000))
    cfgs = [dict(algorithm=a, schedule=s, precompute_depth=d, unique=u)
            for a in ("nonblockinggroups", "nonblockingallgroups")
            for s in ("plain", "batched") for d in (1, 3) for u in (0, 1)]
    cfgs += [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=d,
                  unique=u, cross_groups=2, window=w)
             for d, u, w in ((2, 0, 0), (3, 1, 0), (3, 0, 64))]
    check_all(engine, [g], data, 250, None, cfgs)


@pytest.mark.parametrize("is_dup", [False, True])
def test_start_of_data_chunks(engine, is_dup):
    r = random.Random(31)
    alpha = list(b"ab")
    specs = [H.random_spec(r, 3, alpha) for _ in range(6)]
    g, _ = H.chains_group(specs, start="start-of-data")
    data = bytes(r.choice(alpha) for _ in range(640))
    cfgs = [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=d)
            for s in ("plain", "batched") for d in (1, 2)]
    cfgs += [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=d,
                  cross_groups=2, window=w) for d, w in ((2, 16), (3, 0))]
    check_all(engine, [g, g], data, 64, None, cfgs, is_dup=is_dup)


@pytest.mark.parametrize("seed", range(3))
def test_anml_with_degree_splits(engine, tmp_path, seed):
    path = tmp_path / f"rand{seed}.anml"
    from ngap_amd.groups import write_anml
    write_anml(path, H.random_anml_states(seed))
    groups = Frontend.load(str(path), 3).groups()
    r = random.Random(seed)
    data = bytes(r.choice(b"abcxyzAB019\n\t -[]\\") for _ in range(3000))
    cfgs = [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=d, unique=1)
            for s in ("plain", "batched") for d in (1, 2, 3)]
    check_all(engine, groups, data, 300, None, cfgs)


def test_wide_fanout_direct_api(engine):
    # states with more than four successors (continuation records)
    n = 40
    edges = [(0, j) for j in range(1, 12)] + [(j, 12 + (j % 5)) for j in range(1, 12)]
    edges += [(12 + k, 20 + k) for k in range(5)] + [(20, 20)]
    classes = [[ord("a")]] + [[ord("b"), ord("c")]] * 11 + [[ord("c"), ord("d")]] * 8 + \
        [list(range(256))] * (n - 20)
    g = make_group(n, edges, classes, all_input=[0], reporting=list(range(12, 25)))
    data = b"abcdcdab" * 300
    cfgs = [dict(algorithm=a, schedule=s, precompute_depth=d, unique=u)
            for a in ("nonblockinggroups", "nonblockingallgroups")
            for s in ("plain", "batched") for d in (1, 2, 3) for u in (0, 1)]
    cfgs += [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=d,
                  unique=u, cross_groups=2) for d in (2, 3) for u in (0, 1)]
    check_all(engine, [g], data, 600, None, cfgs)


def test_chunk_of_one_byte_and_no_reports(engine):
    g = make_group(2, [(0, 1)], [[ord("a")], [ord("b")]], all_input=[0], reporting=[1])
    check_all(engine, [g], b"ab" * 64, 1, None,
              [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=d)
               for s in ("plain", "batched") for d in (1, 3)] +
              [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=d,
                    cross_groups=2) for d in (2, 3)])
    check_all(engine, [g, g], b"ab" * 64, 2, None,
              [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=d,
                    cross_groups=2) for d in (2, 3)])
    check_all(engine, [g], b"zz" * 64, 16, None,
              [dict(algorithm="nonblockingallgroups", schedule="plain")])


def test_overflow_is_reported_not_wrong(engine):
    # every byte reports: 4096 reports against a capacity of 10
    g = make_group(1, [], [[ord("a")]], all_input=[0], reporting=[0])
    engine.load_groups([g], 2)
    engine.set_input(b"a" * 4096, 4096, 1)
    say("tiny result capacity")
    with pytest.raises(NgapError) as ei:
        engine.run(algorithm="nonblockingallgroups", precompute_depth=2, result_capacity=10)
    assert ei.value.code == 4
    engine.run(algorithm="nonblockingallgroups", precompute_depth=2)
    assert engine.reports().size == 4096


def test_report_off_counts_without_writing(engine):
    # --report-off: the reports are counted, not written; with one literal
    # all-input state (no duplicates possible) the count is exact
    g = make_group(1, [], [[ord("a")]], all_input=[0], reporting=[0])
    engine.load_groups([g], 2)
    engine.set_input(b"ab" * 2048, 4096, 1)
    for sched in ("plain", "batched"):
        say("report-off", sched)
        st = engine.run(algorithm="nonblockingallgroups", schedule=sched, precompute_depth=2,
                        report_off=True, result_capacity=10)
        assert st["reports_raw"] == 2048
        with pytest.raises(NgapError):   # nothing to read back after --report-off
            engine.reports()
        engine.run(algorithm="nonblockingallgroups", schedule=sched, precompute_depth=2)
        assert engine.reports().size == 2048


def test_spill_overflow_is_reported(engine):
    # a 4-ary tree of 'a' states three levels deep below an all-input root:
    # every position creates 4 + 16 + 64 items, far beyond a 512-item ring when
    # the spill capacity is forced to 64 items per wave.
    n = 1 + 4 + 16 + 64
    edges = [(0, 1 + i) for i in range(4)] + [(1 + i, 5 + 4 * i + j) for i in range(4) for j in range(4)]
    edges += [(5 + i, 21 + 4 * i + j) for i in range(16) for j in range(4)]
    g = make_group(n, edges, [[ord("a")]] * n, all_input=[0], reporting=list(range(21, n)))
    data = b"a" * 20000
    want, _ = oracle.run([g], data, len(data), 1)
    engine.load_groups([g], 1)
    engine.set_input(data, len(data), 1)
    for sched in ("plain", "batched"):
        say("spill overflow", sched)
        with pytest.raises(NgapError) as ei:
            engine.run(algorithm="nonblockinggroups", schedule=sched, spill_capacity=64,
                       window=20000)
        assert ei.value.code == 4 and ei.value.stats["overflow_units"] >= 1
        engine.run(algorithm="nonblockinggroups", schedule=sched, spill_capacity=1 << 16,
                   window=20000)
        assert np.array_equal(engine.reports(), want)


def test_reports_bit_exact_after_reload(engine):
    # reloading groups / input between runs keeps results exact
    rng = np.random.default_rng(9)
    for k in range(3):
        g = synth.regex_group(rng, 3000 * (k + 1))
        data = synth.zipf_bytes(rng, 1 << 15, 1.1)
        check_all(engine, [g], data, 1 << 13, None,
                  [dict(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=2)])


@pytest.mark.parametrize("is_dup", [False, True])
def test_empty_sod_only_and_ragged_input(engine, is_dup):
    # a group without states, a group with start-of-data states only, an
    # all-input group and a regex group; the input's tail beyond the last
    # whole chunk is dropped by both the engine and the oracle
    r = random.Random(77)
    alpha = list(b"abc")
    sod_g, _ = H.chains_group([H.random_spec(r, 3, alpha) for _ in range(5)], start="start-of-data")
    aas_g, _ = H.chains_group([H.random_spec(r, 4, alpha) for _ in range(8)])
    empty = make_group(0, [], [])
    regex = synth.regex_group(np.random.default_rng(5), 400)
    data = bytes(r.choice(alpha) for _ in range(1000)) + b"xyz"
    cfgs = [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=d, cross_groups=x)
            for s, d, x in (("plain", 1, 1), ("plain", 3, 1), ("batched", 1, 1),
                            ("batched", 2, 0), ("batched", 2, 1), ("batched", 3, 1))]
    cfgs.append(dict(algorithm="blockinggroups", schedule="plain"))
    check_all(engine, [empty, sod_g, aas_g, regex, empty], data, 100, None, cfgs, is_dup=is_dup)


def test_a_thousand_and_twenty_four_groups(engine):
    # the group-id budget of the report encoding (10 bits) used in full
    r = random.Random(1024)
    alpha = list(b"acgt")
    groups = [H.chains_group([H.random_spec(r, r.randint(1, 4), alpha)
                              for _ in range(r.randint(1, 3))])[0] for _ in range(1024)]
    data = bytes(r.choice(alpha) for _ in range(1 << 13))
    cfgs = [dict(algorithm="nonblockingallgroups", schedule=s, precompute_depth=d, unique=u)
            for s, d, u in (("plain", 2, 1), ("batched", 2, 0), ("batched", 3, 1))]
    want = check_all(engine, groups, data, 1 << 11, None, cfgs)
    assert int(np.max(want >> np.uint64(54))) == 1023 or want.size == 0


def _top_id_group(n, pattern):
    """A group of n states whose only edges form literal chains of `pattern`
    placed at the top of the id range (the rest are isolated states): its
    reports carry state ids up to n - 1."""
    from ngap_amd.groups import Group
    L = len(pattern)
    first = n - L
    row = np.zeros(n + 1, np.int32)
    row[first + 1:] = np.minimum(np.arange(1, n - first + 1), L - 1)  # edges i -> i + 1 in the chain
    col = np.arange(first + 1, n, dtype=np.int32)
    sets = np.zeros(n * 8, np.uint32)
    attrs = np.zeros(n, np.uint8)
    for i, b in enumerate(pattern):
        v = first + i
        sets[v * 8 + b // 32] |= np.uint32(1 << (b % 32))
    attrs[first] |= 1          # all-input start
    attrs[n - 1] |= 2          # reports
    return Group(row=row, col=col, sets=sets, attrs=attrs,
                 aas=np.array([first], np.int32), sod=np.zeros(0, np.int32))


def test_state_ids_at_the_22_bit_limit(engine):
    """Group 1 has 2^22 states with its reporting chain at the top ids: the
    report encoding (state | group << 22) carries state 4,194,303 of group 1
    exactly; a group of 2^22 + 1 states is rejected before anything runs."""
    n = 1 << 22
    g0 = _top_id_group(64, b"abc")
    g1 = _top_id_group(n, b"abcd")
    rng = np.random.default_rng(5)
    data = rng.choice(np.frombuffer(b"abcd", np.uint8), 1 << 14)
    want, _ = oracle.run([g0, g1], data, 1 << 12)
    grp = want >> np.uint64(54)
    st = (want >> np.uint64(32)) & np.uint64((1 << 22) - 1)
    assert want.size > 10 and int(st.max()) == n - 1 and set(grp.tolist()) == {0, 1}
    engine.load_groups([g0, g1], 3)
    engine.set_input(data, 1 << 12)
    for sched, d in (("plain", 1), ("batched", 2), ("batched", 3)):
        say("22-bit ids", sched, d)
        engine.run(algorithm="nonblockingallgroups", schedule=sched, precompute_depth=d,
                   unique=True)
        assert np.array_equal(engine.reports(), want)
    say("22-bit ids: one state too many")
    with pytest.raises(NgapError):
        engine.load_groups([g0, _top_id_group(n + 1, b"abcd")], 1)
