"""The oracle (oracle/oracle.cpp) pinned against known answers that do not
come from its own formula: the reference's apple example (README.md:156-183),
substring search, Python ``re`` on every window, chunk/start-of-data rules.
"""
import os
import random

import numpy as np
import pytest

import oracle
from ngap_amd import Frontend
from ngap_amd.groups import make_group
from tests.conftest import GOLDEN
from tests import nfa_helpers as H


def test_apple_known_answer():
    g = Frontend.load(os.path.join(GOLDEN, "apple.anml"), 1).groups()
    data = open(os.path.join(GOLDEN, "apple_input.txt"), "rb").read()
    rep, st = oracle.run(g, data, 81, 1)
    assert [hex(x) for x in rep] == ["0x400000005", "0x40000002f", "0x400000040", "0x40000004b"]
    assert st["reports"] == 4


@pytest.mark.parametrize("seed", range(4))
def test_literal_patterns_are_substring_search(seed):
    r = random.Random(seed)
    pats = [bytes(r.choice(b"ACGT") for _ in range(r.randint(3, 7))) for _ in range(12)]
    specs = [[(H.LIT, [c]) for c in p] for p in pats]
    g, owner = H.chains_group(specs)
    data = bytes(r.choice(b"ACGT") for _ in range(3000))
    L, S = 500, 6
    rep, _ = oracle.run([g], data, L, S)
    got = H.reports_to_pattern_offsets(rep, owner)
    want = set()
    for pi, p in enumerate(pats):
        for c in range(S):
            chunk = data[c * L:(c + 1) * L]
            i = chunk.find(p)
            while i >= 0:
                want.add((pi, c * L + i + len(p) - 1))
                i = chunk.find(p, i + 1)
    assert got == want
    # exact state ids: the last state of each literal chain reports
    last = {pi: max(s for s, o in owner.items() if o == pi) for pi in range(len(pats))}
    ids = ((rep >> np.uint64(32)) & np.uint64((1 << 22) - 1)).astype(int)
    assert all(int(s) == last[owner[int(s)]] for s in ids)


@pytest.mark.parametrize("seed", range(6))
def test_class_and_repeat_patterns_match_python_re(seed):
    r = random.Random(100 + seed)
    alpha = list(b"abcdef")
    specs = [H.random_spec(r, r.randint(2, 5), alpha) for _ in range(6)]
    g, owner = H.chains_group(specs)
    data = bytes(r.choice(alpha) for _ in range(240))
    L, S = 80, 3
    rep, _ = oracle.run([g], data, L, S)
    assert H.reports_to_pattern_offsets(rep, owner) == H.brute_force(specs, data, L, S)


@pytest.mark.parametrize("is_dup", [False, True])
def test_start_of_data_only_in_first_chunk_unless_dup(is_dup):
    r = random.Random(7)
    alpha = list(b"ab")
    specs = [H.random_spec(r, 3, alpha) for _ in range(4)]
    g, owner = H.chains_group(specs, start="start-of-data")
    data = bytes(r.choice(alpha) for _ in range(64))
    rep, _ = oracle.run([g], data, 16, 4, is_dup=is_dup)
    want = H.brute_force(specs, data, 16, 4, anchored=True, is_dup=is_dup)
    assert H.reports_to_pattern_offsets(rep, owner) == want
    assert want  # the case is not vacuous


def test_group_bits_and_order():
    # two groups: the same literal in each; ids carry group << 22 and the
    # output is sorted by (offset, id) (code/src/ngap/ngap.cu:19-30).
    g0 = make_group(2, [(0, 1)], [[ord("x")], [ord("y")]], all_input=[0], reporting=[1])
    g1 = make_group(1, [], [[ord("y")]], all_input=[0], reporting=[0])
    rep, _ = oracle.run([g0, g1], b"xyxy", 4, 1)
    assert [hex(x) for x in rep] == ["0x100000001", "0x40000000000001", "0x100000003",
                                      "0x40000000000003"]


def test_self_loop_dot_star_set_semantics():
    # all-input '*' that loops on itself, feeding 'a' -> 'b' (report): the
    # reports are every 'ab' end, exactly once.
    star = list(range(256))
    g = make_group(3, [(0, 0), (0, 1), (1, 2)], [star, [ord("a")], [ord("b")]],
                   all_input=[0], reporting=[2])
    data = b"zabab" * 20
    rep, _ = oracle.run([g], data, len(data), 1)
    offs = [int(x & 0xFFFFFFFF) for x in rep]
    assert offs == [i + 1 for i in range(len(data) - 1) if data[i:i + 2] == b"ab"]


def test_empty_and_reportless():
    g = make_group(1, [], [[ord("q")]], all_input=[0])
    rep, st = oracle.run([g], b"qqqq", 2, 2)
    assert rep.size == 0 and st["matched_aas"] == 4


def test_work_counts_on_a_closed_form():
    # one all-input literal 'a' -> literal 'b' (report) over "ab"*n, one chunk:
    # every 'a' matches (n), every 'b' position tests the successor (n) and
    # matches; counts follow directly.
    n = 50
    g = make_group(2, [(0, 1)], [[ord("a")], [ord("b")]], all_input=[0], reporting=[1])
    _, st = oracle.run([g], b"ab" * n, 2 * n, 1)
    assert st["matched_aas"] == n
    assert st["tested_non_aas"] == n
    assert st["reports"] == n
    assert st["matched_advance"] == 2 * n - 1  # all but the final 'b' advance
    assert st["successors"] == n  # only the 'a' items have a successor


@pytest.mark.parametrize("k", [1, 2, 3])
def test_memo_work_counts_on_a_closed_form(k):
    # all-input 'a' -> 'b' -> 'c' (report) over "abc"*n: every chain is a-b-c.
    # depth k: states at chain depth k-1 come from memo lists; items at chain
    # depth >= k-1 that have a successor are expanded.
    n = 40
    g = make_group(3, [(0, 1), (1, 2)], [[ord("a")], [ord("b")], [ord("c")]],
                   all_input=[0], reporting=[2])
    _, st = oracle.run([g], b"abc" * n, 3 * n, 1, memo_depth=k)
    assert st["memo_list_items"] == n              # one chain depth per k (0, 1, 2)
    assert st["expanded_items"] == {1: 2 * n, 2: n, 3: 0}[k]
    assert st["expanded_tests"] == st["expanded_items"]
    assert st["reports"] == n
    assert st["root_positions"] == n               # every 'a'
    if k >= 2:                                     # (counted for memo depth >= 2)
        assert st["level2_lists"] == n             # every 'b' (one step below a root)
    assert st["deepest_lists"] == n                # 'a', 'b' or 'c' for k = 1, 2, 3


def test_memo_depth_restarts_at_all_input_successors():
    # all-input '*' with a self-loop feeding 'a' -> 'b' (report): the loop edge
    # re-enters an all-input state, so chains never grow through it.
    star = list(range(256))
    g = make_group(3, [(0, 0), (0, 1), (1, 2)], [star, [ord("a")], [ord("b")]],
                   all_input=[0], reporting=[2])
    data = b"ab" * 30
    _, st = oracle.run([g], data, len(data), 1, memo_depth=2)
    assert st["root_positions"] == len(data)
    # 'a' is enabled one position after a '*' match: every 'a' but the first
    assert st["memo_list_items"] == 29            # each such 'a' at depth 1
    assert st["expanded_items"] == 29             # each expands to 'b'
    assert st["reports"] == 29
    assert st["level2_lists"] == 29               # positions of those 'a's


def _with_sod_on_all_input(g):
    """The group with its first all-input state also listed as start-of-data
    (the frontier then holds it with the start-of-data depth at a chunk's
    first position)."""
    from ngap_amd.groups import Group
    sod = np.unique(np.concatenate([g.sod, g.aas[:1]])).astype(np.int32)
    return Group(g.row, g.col, g.sets, g.attrs, g.aas, sod)


@pytest.mark.parametrize("seed", range(3))
def test_implicit_all_input_equals_reference_loop(seed):
    """The fast mode (all-input states visited through per-byte lists after a
    chunk's first position) gives the reports and every work count of the
    reference's loop as written, on automata with dot-star loops, start-of-data
    states (one of them also all-input), chunks with and without duplication
    and memo depths 1-3."""
    from ngap_amd import synth
    rng = np.random.default_rng(900 + seed)
    groups = [synth.regex_group(rng, int(s), dotstar=(i % 3 == 0))
              for i, s in enumerate([120, 900, 2500, 400])]
    groups.append(_with_sod_on_all_input(H.chains_group(
        [[(H.LIT, [65]), (H.CLS, [66, 67]), (H.LIT, [68])], [(H.LIT, [66]), (H.LIT, [67])]])[0]))
    data = synth.drifting_input(rng, 1 << 14, 2048, 5e-3, groups[:4])
    data[::97] = np.frombuffer(b"ABCD", np.uint8)[np.arange(data[::97].size) % 4]
    for K in (1, 2, 3):
        for dup in (False, True):
            a, sa = oracle.run(groups, data, 2048, is_dup=dup, memo_depth=K, implicit_aas=False)
            b, sb = oracle.run(groups, data, 2048, is_dup=dup, memo_depth=K, implicit_aas=True)
            assert a.size > 10
            assert np.array_equal(a, b), (K, dup)
            assert sa == sb, (K, dup)
