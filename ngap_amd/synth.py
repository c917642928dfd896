"""Seeded synthetic automata and inputs for the benchmark workloads.

Workloads (sizes as stated for each):
  literal64  64 literal patterns over {A,C,G,T}, lengths 8..16, each a chain
             whose first state is all-input and whose last state reports;
             1 MiB of iid uniform ACGT, one chunk.
  regex20k   20,000 patterns of 6..40 positions: a literal byte (p=0.6), a
             class of 2..32 bytes (p=0.3) or a class repeated {1,4} (p=0.1).
             Patterns are grouped into connected components whose sizes follow
             a truncated power law (alpha 1.5, 5..5,000 states); inside a
             component the patterns share prefixes through a trie whose nodes
             have at most four successors, so no state exceeds the out-degree
             limit of 4 the reference front-end enforces.  The root of each
             component is all-input.  16 MiB Zipf(1.1) bytes, 256 KiB chunks.
  batch1024  1,024 automata, log-uniform 50..50,000 states, one group each;
             20 % start with an all-input '*' state that loops on itself and
             feeds the pattern tries (leading dot-star).  8 MiB Zipf(1.1),
             64 KiB chunks.
  learned    batch1024 automata; per-chunk Zipf exponent
             s = 1.2 - 0.4 cos(2 pi c / 128) and planted pattern instances at
             1e-4 per byte.
  serving    64 resident automata, log-uniform 1k..200k states (regex20k
             generator); requests of log-uniform 1 KiB..4 MiB Zipf(1.1).

Zipf(s) over 256 symbols: rank k (1..256) has probability ~ k^-s and is byte
k-1.  States of a component are numbered contiguously in breadth-first order
and components follow each other, which is the numbering the reference
front-end assigns to the same automaton written as ANML (no splits happen,
so the id pipeline is the identity).
"""
import random

import numpy as np

from .groups import Group

LIT, CLS, REP = 0, 1, 2


def zipf_probs(s, n=256):
    p = np.arange(1, n + 1, dtype=np.float64) ** (-float(s))
    return p / p.sum()


def zipf_bytes(rng, n, s=1.1):
    return rng.choice(256, size=int(n), p=zipf_probs(s)).astype(np.uint8)


def acgt_bytes(rng, n):
    return np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(n))]


def _class_pool(rng, n=4096, lo=2, hi=32):
    """n random classes of k distinct bytes, k uniform in [lo, hi]."""
    perm = np.argsort(rng.random((n, 256)), axis=1)[:, :hi]
    ks = rng.integers(lo, hi + 1, n)
    return [sorted(perm[i, :ks[i]].tolist()) for i in range(n)]


class _Builder:
    """Accumulates states (class, flags) and edges for one group."""

    def __init__(self):
        self.classes = []   # arrays of bytes
        self.aas = []
        self.report = []
        self.edges = []
        self.patterns = []  # per pattern: list of (kind, bytes)

    def state(self, bytes_, all_input=False):
        i = len(self.classes)
        self.classes.append(bytes_)
        if all_input:
            self.aas.append(i)
        return i

    def build(self):
        n = len(self.classes)
        lens = np.fromiter((len(b) for b in self.classes), np.int64, n)
        flat = np.concatenate([np.asarray(b, np.int64) for b in self.classes]) if n else \
            np.zeros(0, np.int64)
        owner = np.repeat(np.arange(n, dtype=np.int64), lens)
        sets = np.zeros(n * 8, np.uint32)
        np.bitwise_or.at(sets, owner * 8 + flat // 32,
                         np.left_shift(np.uint32(1), (flat % 32).astype(np.uint32)))
        e = np.array(self.edges, np.int64).reshape(-1, 2)
        key = np.unique(e[:, 0] * (1 << 32) + e[:, 1])
        src, dst = key >> 32, key & 0xFFFFFFFF
        row = np.zeros(n + 1, np.int64)
        np.add.at(row, src + 1, 1)
        row = np.cumsum(row).astype(np.int32)
        col = dst.astype(np.int32)
        attrs = np.zeros(n, np.uint8)
        aas = np.array(sorted(self.aas), np.int32)
        attrs[aas] |= 1
        if self.report:
            attrs[np.array(self.report, np.int64)] |= 2
        return Group(row, col, sets, attrs, aas, np.zeros(0, np.int32))


def _gen_trie(r, B, n_patterns, root_all_input, pool, root=None, p_new=0.2,
              len_lo=6, len_hi=40, root_literal=True):
    """Adds one component: a trie of patterns.  Returns the root state id.
    r is a random.Random (the structure generator); classes come from pool."""
    # element: [kind, bytes, children, end]
    def rand_elem():
        u = r.random()
        if u < 0.6:
            return [LIT, [r.randrange(256)], [], False]
        if u < 0.9:
            return [CLS, pool[r.randrange(len(pool))], [], False]
        return [REP, pool[r.randrange(len(pool))], [], False]

    if root is None:
        first = [LIT, [r.randrange(256)], [], False] if root_literal else rand_elem()
    else:
        first = root
    elems = [first]
    for _ in range(n_patterns):
        L = r.randint(len_lo, len_hi)
        cur = 0
        spec = [(elems[0][0], elems[0][1])]
        for _d in range(1, L):
            e = elems[cur]
            maxc = 3 if e[0] == REP else 4
            kids = e[2]
            if kids and (len(kids) >= maxc or r.random() > p_new):
                cur = kids[r.randrange(len(kids))]
            else:
                elems.append(rand_elem())
                kids.append(len(elems) - 1)
                cur = len(elems) - 1
            spec.append((elems[cur][0], elems[cur][1]))
        elems[cur][3] = True
        B.patterns.append(spec)
    # breadth-first state numbering
    order = [0]
    for i in order:
        order.extend(elems[i][2])
    entry, exits = {}, {}
    for i in order:
        kind, b, _, _ = elems[i]
        if kind == REP:
            ids = [B.state(b) for _ in range(4)]
            for a, c in zip(ids, ids[1:]):
                B.edges.append((a, c))
            entry[i], exits[i] = ids[0], ids
        else:
            s = B.state(b, all_input=(i == 0 and root_all_input))
            entry[i], exits[i] = s, [s]
    for i in order:
        for c in elems[i][2]:
            for x in exits[i]:
                B.edges.append((x, entry[c]))
        if elems[i][3]:
            B.report.extend(exits[i])
    return entry[0]


def _powerlaw_sizes(rng, total, alpha=1.5, lo=5, hi=5000):
    out = []
    a1 = 1.0 - alpha
    while sum(out) < total:
        u = rng.random()
        x = (lo ** a1 + u * (hi ** a1 - lo ** a1)) ** (1.0 / a1)
        out.append(int(x))
    return out


def regex_group(rng, target_states, n_patterns=None, dotstar=False):
    """One group of trie components (regex20k generator).  With ``dotstar`` the
    group is one component: an all-input '*' state with a self-loop feeding
    up to three tries whose roots are not all-input."""
    B = _Builder()
    r = random.Random(int(rng.integers(0, 2**63 - 1)))
    pool = _class_pool(rng, 4096 if target_states > 20000 else 512)
    per_pattern = 23.5  # states per pattern after prefix sharing (measured)
    if n_patterns is None:
        n_patterns = max(1, int(round(target_states / per_pattern)))
    if dotstar:
        star = B.state(np.arange(256), all_input=True)
        B.edges.append((star, star))
        k = min(3, n_patterns)
        shares = np.full(k, n_patterns // k)
        shares[: n_patterns % k] += 1
        for s in shares:
            root = _gen_trie(r, B, int(s), False, pool)
            B.edges.append((star, root))
    else:
        sizes = np.array(_powerlaw_sizes(rng, target_states), np.float64)
        if len(sizes) > n_patterns:
            sizes = sizes[:n_patterns]
        # exactly n_patterns, at least one per component (largest remainder)
        share = (n_patterns - len(sizes)) * sizes / sizes.sum()
        counts = 1 + np.floor(share).astype(np.int64)
        rest = n_patterns - int(counts.sum())
        if rest > 0:
            counts[np.argsort(-(share - np.floor(share)))[:rest]] += 1
        for c in counts:
            _gen_trie(r, B, int(c), True, pool)
    g = B.build()
    g.patterns = B.patterns
    return g


def literal64(seed=1000, n_patterns=64, input_len=1 << 20):
    rng = np.random.default_rng(seed)
    B = _Builder()
    acgt = np.frombuffer(b"ACGT", np.uint8)
    for _ in range(n_patterns):
        L = int(rng.integers(8, 17))
        syms = acgt[rng.integers(0, 4, L)]
        prev = None
        for j, s in enumerate(syms):
            v = B.state(np.array([int(s)]), all_input=(j == 0))
            if prev is not None:
                B.edges.append((prev, v))
            prev = v
        B.report.append(prev)
        B.patterns.append(bytes(syms))
    g = B.build()
    g.patterns = B.patterns
    data = acgt_bytes(np.random.default_rng(seed + 1000), input_len)
    return [g], data, input_len


def regex20k(seed=1001, n_patterns=20000, target_states=400000, input_len=16 << 20,
             chunk=256 << 10):
    rng = np.random.default_rng(seed)
    g = regex_group(rng, target_states, n_patterns=n_patterns)
    data = zipf_bytes(np.random.default_rng(seed + 1000), input_len, 1.1)
    return [g], data, chunk


def batch_groups(rng, n=1024, lo=50, hi=50000, dotstar_frac=0.2):
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
    dots = rng.random(n) < dotstar_frac
    return [regex_group(rng, int(s), dotstar=bool(d)) for s, d in zip(sizes, dots)]


def batch1024(seed=1002, n=1024, input_len=8 << 20, chunk=64 << 10):
    rng = np.random.default_rng(seed)
    groups = batch_groups(rng, n)
    data = zipf_bytes(np.random.default_rng(seed + 1000), input_len, 1.1)
    return groups, data, chunk


def sample_instance(rng, spec):
    """A byte string matching one pattern spec [(kind, bytes), ...]."""
    out = []
    for kind, b in spec:
        reps = int(rng.integers(1, 5)) if kind == REP else 1
        b = list(b)
        for _ in range(reps):
            out.append(int(b[int(rng.integers(0, len(b)))]))
    return bytes(out)


def drifting_input(rng, input_len, chunk, planted_rate=1e-4, groups=None, period=128):
    n_chunks = input_len // chunk
    parts = []
    for c in range(n_chunks):
        s = 1.2 - 0.4 * np.cos(2 * np.pi * c / period)
        parts.append(zipf_bytes(rng, chunk, s))
    data = np.concatenate(parts) if parts else np.zeros(0, np.uint8)
    if groups:
        plant(rng, data, groups, planted_rate)
    return data


def plant(rng, data, groups, rate):
    """Writes instances of the groups' patterns into data (in place) at
    `rate` per byte, each at a uniform position."""
    specs = [(gi, p) for gi, g in enumerate(groups) for p in getattr(g, "patterns", [])]
    n = data.size
    for _ in range(int(rate * n)):
        gi, spec = specs[int(rng.integers(0, len(specs)))]
        inst = sample_instance(rng, spec)
        pos = int(rng.integers(0, max(1, n - len(inst))))
        data[pos:pos + len(inst)] = np.frombuffer(inst, np.uint8)[: n - pos]
    return data


def learned(seed=1003, n=1024, input_len=8 << 20, chunk=64 << 10):
    rng = np.random.default_rng(seed)
    groups = batch_groups(rng, n)
    data = drifting_input(np.random.default_rng(seed + 1000), input_len, chunk, 1e-4, groups)
    return groups, data, chunk


def serving_groups(seed=1004, n=64, lo=1000, hi=200000):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)
    return [regex_group(rng, int(s)) for s in sizes]


def serving_requests(seed=2004, n=256, n_groups=64, lo=1 << 10, hi=4 << 20, rate=None):
    """(arrival_s, group, bytes) triples; Poisson arrivals at `rate` per second
    (all at time 0 when rate is None)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for _ in range(n):
        if rate:
            t += float(rng.exponential(1.0 / rate))
        L = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        out.append((t, int(rng.integers(0, n_groups)), zipf_bytes(rng, L, 1.1)))
    return out


def write_groups_anml(path, groups):
    """Writes every group's states into one ANML document, group after group
    in state order, with ids g<group>_<state> and VASim hex-range symbol
    sets (the front-end reads the same automaton back)."""
    from .groups import symbol_set_string, _xml_attr
    cache = {}
    with open(path, "w") as f:
        f.write('<anml version="1.0">\n<automata-network id="synthetic">\n')
        for gi, g in enumerate(groups):
            sets = np.ascontiguousarray(g.sets).reshape(g.num_states, 8)
            sod = set(g.sod.tolist())
            lines = []
            for v in range(g.num_states):
                key = sets[v].tobytes()
                ss = cache.get(key)
                if ss is None:
                    bits = np.unpackbits(sets[v].view(np.uint8), bitorder="little").astype(bool)
                    ss = cache[key] = _xml_attr(symbol_set_string(bits))
                start = ' start="all-input"' if g.attrs[v] & 1 else \
                    (' start="start-of-data"' if v in sod else "")
                lines.append(f'<state-transition-element id="g{gi}_{v}" symbol-set="{ss}"{start}>')
                for u in g.col[g.row[v]:g.row[v + 1]]:
                    lines.append(f'<activate-on-match element="g{gi}_{u}"/>')
                if g.attrs[v] & 2:
                    lines.append('<report-on-match reportcode="1"/>')
                lines.append("</state-transition-element>")
            f.write("\n".join(lines))
            f.write("\n")
        f.write("</automata-network>\n</anml>\n")


def dump_groups(groups):
    """The groups in the front-end dump format (ngap_frontend_dump), with the
    state names write_groups_anml gives them: equal to the front-end's dump of
    that ANML exactly when loading it gives these groups back unchanged."""
    out = [f"groups {len(groups)}"]
    for gi, g in enumerate(groups):
        n = g.num_states
        out.append(f"group {gi} states {n} edges {g.num_edges} aas {g.aas.size} "
                   f"sod {g.sod.size}")
        sets = np.ascontiguousarray(g.sets).reshape(n, 8)
        for v in range(n):
            words = "".join(f"{int(w):08x}" for w in sets[v][::-1])
            succ = "".join(f" {int(u)}" for u in g.col[g.row[v]:g.row[v + 1]])
            out.append(f"{v} g{gi}_{v} a{int(g.attrs[v]) & 3} s{words} ->{succ}")
    return "\n".join(out) + "\n"


def to_anml_states(group, prefix="s"):
    """ANML state dicts for write_anml (ids s0, s1, ... in state order)."""
    n = group.num_states
    sets = group.sets.reshape(n, 8)
    out = []
    for v in range(n):
        bits = np.array([(sets[v, b // 32] >> (b % 32)) & 1 for b in range(256)], bool)
        st = "all-input" if group.attrs[v] & 1 else ("start-of-data" if v in set(group.sod.tolist()) else "")
        nxt = [f"{prefix}{u}" for u in group.col[group.row[v]:group.row[v + 1]]]
        out.append(dict(id=f"{prefix}{v}", set=bits, start=st, report=bool(group.attrs[v] & 2),
                        next=nxt))
    return out
