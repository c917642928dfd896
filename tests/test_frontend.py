"""Front-end (ANML -> groups): the state-id namespace must equal the
reference's.  Pinned by (a) dumps the reference's own front-end produced
(committed fixtures, tests/golden/make_frontend_golden.py), (b) the same
dumper live when oracle/_ref is built, and (c) explicit VASim symbol sets."""
import glob
import hashlib
import json
import os
import re

import numpy as np
import pytest

import oracle
from ngap_amd import Frontend, parse_symbol_set, synth
from ngap_amd.groups import write_anml
from tests.conftest import GOLDEN
from tests import nfa_helpers as H

FIX = os.path.join(GOLDEN, "frontend")


def bits(*bytes_):
    out = np.zeros(256, bool)
    for b in bytes_:
        out[b] = True
    return out


def rng_bits(a, b):
    return bits(*range(ord(a) if isinstance(a, str) else a, (ord(b) if isinstance(b, str) else b) + 1))


SYMBOL_SETS = [
    ("*", np.ones(256, bool)),
    (".", ~bits(10)),
    ("a", bits(ord("a"))),
    ("[abc]", bits(*b"abc")),
    ("[a-c]", rng_bits("a", "c")),
    ("[^a]", ~bits(ord("a"))),
    ("\\x41", bits(0x41)),
    ("[\\x00-\\x1f]", rng_bits(0, 0x1F)),
    ("\\d", rng_bits("0", "9")),
    ("[\\w]", rng_bits("0", "9") | rng_bits("A", "Z") | rng_bits("a", "z") | bits(ord("_"))),
    ("\\s", bits(9, 10, 11, 12, 13, 32)),
    ("[a\\-z]", bits(*b"a-z")),      # escaped '-' is literal
    ("[-a]", bits(*b"-a")),          # '-' right after '[' is literal
    ("[\\n\\t]", bits(10, 9)),
    ("[\\]\\[]", bits(*b"[]")),
    ("[\\\\]", bits(ord("\\"))),
    ("[A-Fa-f0-9]", rng_bits("A", "F") | rng_bits("a", "f") | rng_bits("0", "9")),
    ("[^\\x0a]", ~bits(10)),
    ("[\\x30-\\x39A]", rng_bits("0", "9") | bits(ord("A"))),
    ("{", bits(ord("{"))),
]


@pytest.mark.parametrize("text,want", SYMBOL_SETS, ids=[s for s, _ in SYMBOL_SETS])
def test_symbol_sets(text, want):
    assert np.array_equal(parse_symbol_set(text), want)


def test_symbol_set_errors():
    from ngap_amd import NgapError
    with pytest.raises(NgapError):
        parse_symbol_set("{abc}")   # the reference exits: curly braces
    with pytest.raises(NgapError):
        parse_symbol_set("[abc")    # unbalanced brackets


def _golden_cases():
    out = []
    for dump in sorted(glob.glob(os.path.join(FIX, "*.dump"))):
        m = re.match(r"(.*)_g(\d+)_m(-?\d+)\.dump$", os.path.basename(dump))
        anml = os.path.join(FIX, m.group(1) + ".anml")
        if not os.path.exists(anml):
            anml = os.path.join(GOLDEN, m.group(1) + ".anml")
        out.append((anml, int(m.group(2)), int(m.group(3)), dump))
    return out


@pytest.mark.parametrize("anml,gnum,maxsz,dump", _golden_cases(),
                         ids=[os.path.basename(c[3]) for c in _golden_cases()])
def test_matches_reference_dumps(anml, gnum, maxsz, dump):
    ours = Frontend.load(anml, gnum, maxsz).dump()
    assert ours == open(dump).read()


have_ref = os.path.exists(oracle.REF_DUMPER)


@pytest.mark.skipif(not have_ref, reason="reference front-end dumper not built")
@pytest.mark.parametrize("seed", range(20, 44))
def test_random_anml_against_live_reference(tmp_path, seed):
    path = str(tmp_path / "r.anml")
    write_anml(path, H.random_anml_states(seed, n=40 + seed * 3, n_cc_hint=2 + seed % 6))
    for gnum in (1, 3, 1 + seed % 11):
        ref = oracle.ref_dump(path, gnum)
        if ref.startswith("groups 0"):
            with pytest.raises(Exception):
                Frontend.load(path, gnum)
            continue
        assert Frontend.load(path, gnum).dump() == ref


def test_group_count_lowered_to_component_count():
    fe = Frontend.load(os.path.join(GOLDEN, "apple.anml"), 10)
    assert fe.num_groups == 1  # one component: --group-num 10 becomes 1


def test_unsupported_elements_and_suffix(tmp_path):
    p = tmp_path / "x.anml"
    p.write_text('<anml><automata-network id="a"><counter id="c"/></automata-network></anml>')
    with pytest.raises(Exception):
        Frontend.load(str(p), 1)
    q = tmp_path / "x.xml"
    q.write_text("<anml/>")
    with pytest.raises(Exception):
        Frontend.load(str(q), 1)


def test_entities_and_forward_references(tmp_path):
    text = ('<?xml version="1.0"?>\n<!-- c -->\n<anml version="1.0"><automata-network id="n">'
            '<state-transition-element id="a" symbol-set="[&lt;&amp;]" start="all-input">'
            '<activate-on-match element="b"/></state-transition-element>'
            "<state-transition-element id='b' symbol-set='&#x41;'><report-on-match/>"
            "</state-transition-element></automata-network></anml>")
    g = Frontend.from_text(text, 1).groups()[0]
    assert g.num_states == 2 and list(g.col) == [1] and list(g.aas) == [0]
    assert np.array_equal(parse_symbol_set("[<&]"),
                          np.array([(g.sets[b // 32] >> (b % 32)) & 1 for b in range(256)], bool))
    assert g.attrs[1] & 2


LARGE = json.load(open(os.path.join(FIX, "large.json")))


@pytest.mark.parametrize("name", sorted(LARGE))
def test_large_automata_match_reference_digest(tmp_path, name):
    """Benchmark-scale automata (regex20k's 474k states, 64 batch1024
    automata, a random automaton with 30 % wide states that the out-degree
    limit splits) through the front-end: the dump equals, by SHA-256 and
    per-group counts, the reference front-end's dump of the same ANML."""
    from tests.frontend_large import write_case
    case = LARGE[name]
    path = str(tmp_path / f"{name}.anml")
    groups = write_case(name, path)
    ours = Frontend.load(path, case["group_num"], case["max_nfa_size"]).dump()
    assert re.findall(r"^group \d+ .*$", ours, re.M) == case["groups"]
    assert hashlib.sha256(ours.encode()).hexdigest() == case["sha256"]
    if name == "regex20k":
        # the workload's own arrays are what the front-end makes of its ANML
        # (the id pipeline is the identity for it)
        assert hashlib.sha256(synth.dump_groups(groups).encode()).hexdigest() == case["sha256"]
