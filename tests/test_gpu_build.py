"""The device image builder (HIP kernels) against the host builder: every
image array byte for byte, on random automata and on the literal64 and
regex20k workloads, at memo depths 1-3; reports from the device-built image
stay oracle-exact; malformed groups fail the same way on both."""
import random

import numpy as np
import pytest

import oracle
from ngap_amd import Engine, Frontend, NgapError, synth
from ngap_amd.groups import Group, make_group, write_anml
from tests import nfa_helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def compare_builds(engine, groups, depth, label):
    say(label, "host build, depth", depth)
    engine.load_groups(groups, depth, build="host")
    info_h = engine.image_info()
    host = {n: engine.image_array(n) for n in Engine.IMAGE_ARRAYS}
    say(label, "device build, depth", depth)
    engine.load_groups(groups, depth, build="device")
    info_d = engine.image_info()
    assert info_d["built_on_device"] == 1 and info_h["built_on_device"] == 0
    assert info_d["bytes"] == info_h["bytes"]
    for n in Engine.IMAGE_ARRAYS:
        dev = engine.image_array(n)
        assert dev.size == host[n].size, (label, depth, n, dev.size, host[n].size)
        if not np.array_equal(dev, host[n]):
            bad = np.nonzero(dev != host[n])[0]
            raise AssertionError(f"{label} depth {depth}: {n} differs at {bad.size} bytes, "
                                 f"first at byte {bad[0]}")
    return info_h, info_d


def _random_groups(seed):
    rng = np.random.default_rng(seed)
    r = random.Random(seed)
    groups = [synth.regex_group(rng, int(s), dotstar=(i % 3 == 0))
              for i, s in enumerate([60, 700, 3000, 150, 9000])]
    # chains with start-of-data starts, an empty group, and a state with a
    # dozen successors (continuation records), unsorted and repeated edges
    groups.append(H.chains_group([H.random_spec(r, 4, list(b"abc")) for _ in range(6)],
                                 start="start-of-data")[0])
    groups.append(make_group(0, [], []))
    n = 30
    edges = [(0, j) for j in range(1, 14)] + [(j, 14 + (j % 5)) for j in range(1, 14)]
    edges += [(14 + k, 20 + k) for k in range(5)] + [(20, 20), (21, 0)]
    classes = [[ord("a")]] + [[ord("b"), ord("c")]] * 13 + [[ord("c")]] * 6 + \
        [list(range(256))] * (n - 20)
    wide = make_group(n, edges, classes, all_input=[0, 25], reporting=list(range(14, 26)),
                      start_of_data=[3, 4])
    col = wide.col.copy()
    for v in range(n):  # reverse each row's order and repeat its first edge
        a, b = wide.row[v], wide.row[v + 1]
        col[a:b] = col[a:b][::-1]
    row = wide.row.copy()
    col2, row2 = [], [0]
    for v in range(n):
        seg = list(col[row[v]:row[v + 1]])
        if seg:
            seg.append(seg[0])
        col2 += seg
        row2.append(len(col2))
    groups.append(Group(np.array(row2, np.int32), np.array(col2, np.int32), wide.sets,
                        wide.attrs, wide.aas, wide.sod))
    return groups


@pytest.mark.parametrize("seed", [0, 1])
def test_device_build_equals_host_build_random(engine, seed):
    groups = _random_groups(seed)
    for depth in (1, 2, 3):
        compare_builds(engine, groups, depth, f"random {seed}")
    # reports from the device-built image (depth 3) against the oracle
    rng = np.random.default_rng(100 + seed)
    data = synth.drifting_input(rng, 1 << 15, 2048, 3e-3, groups[:5])
    data[::37] = np.frombuffer(b"abcd", np.uint8)[np.arange(data[::37].size) % 4]
    want, _ = oracle.run(groups, data, 2048)
    engine.set_input(data, 2048)
    for sched, d, x in (("plain", 3, 1), ("batched", 3, 1), ("batched", 2, 0)):
        say("device-built image", sched, d, x)
        engine.run(algorithm="nonblockingallgroups", schedule=sched, precompute_depth=d,
                   cross_groups=x, unique=True)
        assert np.array_equal(engine.reports(), want)


def test_device_build_equals_host_build_anml(engine, tmp_path):
    """Automata from ANML through the front-end (degree splits, self-loops)."""
    groups = []
    for seed in range(3):
        path = tmp_path / f"rand{seed}.anml"
        write_anml(path, H.random_anml_states(seed, n=120))
        groups += Frontend.load(str(path), 3).groups()
    compare_builds(engine, groups, 3, "anml")


def test_device_build_equals_host_build_workloads(engine):
    for name in ("literal64", "regex20k"):
        groups, data, chunk = getattr(synth, name)()
        info_h, info_d = compare_builds(engine, groups, 3, name)
        say(name, "host load ms", round(info_h["build_ms"], 1), "device load ms",
            round(info_d["build_ms"], 1))


def test_device_build_rejects_malformed_groups(engine):
    good = make_group(3, [(0, 1), (1, 2)], [[97], [98], [99]], all_input=[0], reporting=[2])
    bad_col = Group(good.row, np.array([1, 7], np.int32), good.sets, good.attrs, good.aas,
                    good.sod)
    bad_row = Group(np.array([0, 2, 1, 2], np.int32), good.col, good.sets, good.attrs,
                    good.aas, good.sod)
    bad_aas = Group(good.row, good.col, good.sets, good.attrs, np.array([5], np.int32),
                    good.sod)
    for label, g in (("column", bad_col), ("row", bad_row), ("all-input", bad_aas)):
        msgs = []
        for build in ("host", "device"):
            say("malformed", label, build)
            with pytest.raises(NgapError) as ei:
                engine.load_groups([good, g], 3, build=build)
            msgs.append(str(ei.value))
        assert msgs[0] == msgs[1], msgs
    engine.load_groups([good], 3)  # the engine stays usable
