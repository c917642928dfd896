"""NFA groups on the host: the arrays ``Graph``/``Csr`` hold per group in the
reference (code/include/commons/graph.h:101-248), as numpy arrays.

A :class:`Group` is what the engine and the oracle consume.  Groups come from
the ANML front-end (:func:`load_anml`, reference state-id namespace) or from
the synthetic generators in :mod:`ngap_amd.synth`.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib


@dataclass
class Group:
    row: np.ndarray          # int32 [n+1], edges sorted by (src, dst)
    col: np.ndarray          # int32 [E]
    sets: np.ndarray         # uint32 [n*8], bit b in word b/32
    attrs: np.ndarray        # uint8 [n], bit0 all-input, bit1 report
    aas: np.ndarray          # int32 all-input states, ascending
    sod: np.ndarray          # int32 start-of-data states, ascending
    names: Optional[List[str]] = field(default=None, repr=False)

    @property
    def num_states(self):
        return int(self.attrs.shape[0])

    @property
    def num_edges(self):
        return int(self.col.shape[0])

    def desc(self):
        """ctypes descriptor; the arrays must stay alive while it is used."""
        for a, dt in ((self.row, np.int32), (self.col, np.int32), (self.sets, np.uint32),
                      (self.attrs, np.uint8), (self.aas, np.int32), (self.sod, np.int32)):
            assert a.dtype == dt and a.flags.c_contiguous, (a.dtype, dt)
        d = _lib.GroupDesc()
        d.num_states = self.num_states
        d.row_offsets = self.row.ctypes.data
        d.col_indices = self.col.ctypes.data if self.col.size else None
        d.symbol_sets = self.sets.ctypes.data
        d.node_attrs = self.attrs.ctypes.data
        d.num_always_active = int(self.aas.size)
        d.always_active = self.aas.ctypes.data if self.aas.size else None
        d.num_start = int(self.sod.size)
        d.start_states = self.sod.ctypes.data if self.sod.size else None
        return d

    def reports_states(self):
        return np.nonzero(self.attrs & 2)[0]


def make_group(n, edges, classes, all_input=(), start_of_data=(), reporting=(), names=None):
    """Builds a Group from a state count, an edge list and per-state byte sets.

    ``classes`` is a list of n iterables of byte values (or 256-bool arrays).
    """
    edges = sorted(set((int(a), int(b)) for a, b in edges))
    row = np.zeros(n + 1, np.int32)
    col = np.array([b for _, b in edges], np.int32)
    for a, _ in edges:
        row[a + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    sets = np.zeros(n * 8, np.uint32)
    for v, cl in enumerate(classes):
        cl = np.asarray(cl)
        bytes_ = np.nonzero(cl)[0] if cl.dtype == bool else cl.astype(np.int64)
        for b in bytes_:
            sets[v * 8 + b // 32] |= np.uint32(1 << (b % 32))
    attrs = np.zeros(n, np.uint8)
    aas = np.array(sorted(set(all_input)), np.int32)
    sod = np.array(sorted(set(start_of_data) - set(all_input)), np.int32)
    attrs[aas] |= 1
    for r in reporting:
        attrs[r] |= 2
    return Group(row, col, sets, attrs, aas, sod, names)


def _group_from_desc(d, names):
    n = d.num_states
    def arr(ptr, count, ct, dt):
        if count == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).astype(dt).copy()
    row = arr(d.row_offsets, n + 1, C.c_int32, np.int32)
    col = arr(d.col_indices, int(row[-1]), C.c_int32, np.int32)
    sets = arr(d.symbol_sets, n * 8, C.c_uint32, np.uint32)
    attrs = arr(d.node_attrs, n, C.c_uint8, np.uint8)
    aas = arr(d.always_active, d.num_always_active, C.c_int32, np.int32)
    sod = arr(d.start_states, d.num_start, C.c_int32, np.int32)
    return Group(row, col, sets, attrs, aas, sod, names)


class Frontend:
    """ANML -> groups with the reference's state ids (C++ front-end)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def load(cls, path, group_num=10, max_nfa_size=-1, quit_degree=True, remove_degree=False):
        h = C.c_void_p()
        _lib.check(_lib.lib().ngap_frontend_load(str(path).encode(), group_num, max_nfa_size,
                                                 int(quit_degree), int(remove_degree), C.byref(h)))
        return cls(h)

    @classmethod
    def from_text(cls, text, group_num=10, max_nfa_size=-1):
        h = C.c_void_p()
        _lib.check(_lib.lib().ngap_frontend_load_text(text.encode(), group_num, max_nfa_size,
                                                      C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.lib().ngap_frontend_free(self._h)
            self._h = None

    @property
    def num_groups(self):
        return _lib.lib().ngap_frontend_num_groups(self._h)

    def groups(self, with_names=True):
        L = _lib.lib()
        out = []
        for g in range(self.num_groups):
            d = _lib.GroupDesc()
            _lib.check(L.ngap_frontend_group(self._h, g, C.byref(d)))
            names = None
            if with_names:
                names = [L.ngap_frontend_state_name(self._h, g, s).decode()
                         for s in range(d.num_states)]
            out.append(_group_from_desc(d, names))
        return out

    def dump(self):
        return _lib.lib().ngap_frontend_dump(self._h).decode()


def load_anml(path, group_num=10, max_nfa_size=-1):
    return Frontend.load(path, group_num, max_nfa_size).groups()


def parse_symbol_set(s):
    """VASim symbol-set string -> 256-bool numpy array (C++ parser)."""
    w = (C.c_uint32 * 8)()
    _lib.check(_lib.lib().ngap_parse_symbol_set(s.encode("latin-1"), C.byref(w)))
    out = np.zeros(256, bool)
    for i in range(256):
        out[i] = (w[i // 32] >> (i % 32)) & 1
    return out


def _xml_attr(s):
    return (s.replace("&", "&amp;").replace('"', "&quot;").replace("<", "&lt;")
            .replace(">", "&gt;"))


def symbol_set_string(bits):
    """A VASim symbol-set string matching exactly the given 256-bool set."""
    bits = np.asarray(bits, bool)
    if bits.all():
        return "*"
    on = np.nonzero(bits)[0]
    parts = []
    i = 0
    while i < len(on):
        j = i
        while j + 1 < len(on) and on[j + 1] == on[j] + 1:
            j += 1
        a, b = on[i], on[j]
        parts.append(f"\\x{a:02x}" if a == b else f"\\x{a:02x}-\\x{b:02x}")
        i = j + 1
    return "[" + "".join(parts) + "]"


def write_anml(path, states, network_id="synthetic"):
    """Writes ANML.  ``states``: list of dicts with keys id, set (256 bools or a
    symbol-set string), start ('', 'all-input', 'start-of-data'), report
    (bool), next (list of ids)."""
    with open(path, "w") as f:
        f.write('<anml version="1.0">\n')
        f.write(f'<automata-network id="{network_id}">\n')
        for s in states:
            ss = s["set"] if isinstance(s["set"], str) else symbol_set_string(s["set"])
            st = f' start="{s["start"]}"' if s.get("start") else ""
            f.write(f'<state-transition-element id="{_xml_attr(s["id"])}" '
                    f'symbol-set="{_xml_attr(ss)}"{st}>\n')
            for t in s.get("next", ()):
                f.write(f'<activate-on-match element="{_xml_attr(t)}"/>\n')
            if s.get("report"):
                f.write('<report-on-match reportcode="1"/>\n')
            f.write("</state-transition-element>\n")
        f.write("</automata-network>\n</anml>\n")
