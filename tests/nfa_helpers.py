"""Small automata and independent answers for the tests.

Independent answers come from the mathematics of the patterns, not from the
oracle's frontier formula: literal chains are substring search, class chains
and bounded repeats are Python ``re`` on every (start, end) window.
"""
import random
import re

import numpy as np

from ngap_amd.groups import make_group

LIT, CLS, REP = 0, 1, 2


def random_spec(r, n_elems, alphabet=range(256), p=(0.6, 0.3, 0.1), klo=2, khi=6):
    alphabet = list(alphabet)
    spec = []
    for _ in range(n_elems):
        u = r.random()
        if u < p[0]:
            spec.append((LIT, [r.choice(alphabet)]))
        else:
            k = r.randint(min(klo, len(alphabet)), min(khi, len(alphabet)))
            spec.append((CLS if u < p[0] + p[1] else REP, sorted(r.sample(alphabet, k))))
    return spec


def spec_regex(spec):
    parts = []
    for kind, b in spec:
        cls = b"".join(b"\\x%02x" % x for x in b)
        if kind == LIT:
            parts.append(b"\\x%02x" % b[0])
        elif kind == CLS:
            parts.append(b"[" + cls + b"]")
        else:
            parts.append(b"[" + cls + b"]{1,4}")
    return re.compile(b"".join(parts), re.DOTALL)


def chains_group(specs, start="all-input"):
    """One group with one chain component per spec.  Returns (group,
    pattern_of_state) where pattern_of_state maps reporting states to specs."""
    n = 0
    classes, edges, reporting, starts, owner = [], [], [], [], {}
    for pi, spec in enumerate(specs):
        prev_exits = None
        for ei, (kind, b) in enumerate(spec):
            ids = list(range(n, n + (4 if kind == REP else 1)))
            n += len(ids)
            for _ in ids:
                classes.append(b)
            if kind == REP:
                edges += list(zip(ids, ids[1:]))
            if ei == 0:
                starts.append(ids[0])
            if prev_exits is not None:
                edges += [(x, ids[0]) for x in prev_exits]
            prev_exits = ids
        reporting += prev_exits
        for x in prev_exits:
            owner[x] = pi
    if start == "all-input":
        g = make_group(n, edges, classes, all_input=starts, reporting=reporting)
    else:
        g = make_group(n, edges, classes, start_of_data=starts, reporting=reporting)
    return g, owner


def brute_force(specs, data, chunk_len, num_seg, anchored=False, is_dup=False):
    """Set of (pattern, offset): a match of spec ending at offset inside one
    chunk (starting anywhere in it, or at the chunk start when anchored and
    the chunk holds start-of-data states)."""
    out = set()
    rx = [spec_regex(s) for s in specs]
    for c in range(num_seg):
        cs, ce = c * chunk_len, (c + 1) * chunk_len
        if anchored and not (c == 0 or is_dup):
            continue
        for end in range(cs, ce):
            for pi, r in enumerate(rx):
                starts = [cs] if anchored else range(max(cs, end - 200), end + 1)
                for s in starts:
                    if r.fullmatch(data[s:end + 1]):
                        out.add((pi, end))
                        break
    return out


def reports_to_pattern_offsets(reports, owner, group_index=0):
    r = np.asarray(reports, np.uint64)
    off = (r & np.uint64(0xFFFFFFFF)).astype(np.int64)
    hi = (r >> np.uint64(32)).astype(np.int64)
    assert np.all((hi >> 22) == group_index)
    st = hi & ((1 << 22) - 1)
    return {(owner[int(s)], int(o)) for s, o in zip(st, off)}


def random_anml_states(seed, n=60, n_cc_hint=4, self_loop_p=0.15, wide_p=0.15):
    """Random ANML states (dicts for write_anml): mostly forward edges, some
    states with 5..8 successors (so the out-degree limit splits them), a few
    back edges and self-loops, several components, assorted symbol-set
    spellings."""
    r = random.Random(seed)
    spellings = ["*", ".", "[a-c]", "[^a]", "\\x41", "[\\x00-\\x1f]", "\\d", "[\\w]", "\\s",
                 "[a\\-z]", "[-a]", "[xyz]", "a", "[\\x30-\\x39A]", "[^\\x0a]", "[\\n\\t]",
                 "[A-Fa-f0-9]", "[\\]\\[]", "b", "[\\\\]"]
    comp = [r.randrange(n_cc_hint) for _ in range(n)]
    states = []
    for i in range(n):
        later = [j for j in range(i + 1, n) if comp[j] == comp[i]]
        deg = r.randint(5, 8) if r.random() < wide_p else r.randint(0, 2)
        nxt = r.sample(later, min(deg, len(later)))
        earlier = [j for j in range(i) if comp[j] == comp[i]]
        if earlier and r.random() < 0.1:
            nxt.append(r.choice(earlier))
        if r.random() < self_loop_p:
            nxt.insert(r.randrange(len(nxt) + 1), i)
        st = r.choice(["", "", "", "all-input", "start-of-data"])
        states.append(dict(id=f"n{i}_{r.randrange(1000)}", set=r.choice(spellings), start=st,
                           report=r.random() < 0.3, next=[]))
        states[-1]["_next"] = nxt
    for s in states:
        s["next"] = [states[j]["id"] for j in s.pop("_next")]
    return states
