"""ORACLE — test infrastructure only (see oracle.cpp).  Imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg, never by the product.

``run`` restates ngap::automataGroupsReference (code/src/ngap/ngap.cu:2171-
2268) with set semantics and returns the sorted unique reports in the
reference encoding plus the algorithmic work counts the roofline uses.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_DUMPER = os.path.join(_HERE, "_ref", "ref_frontend_dump")


class OracleGroup(C.Structure):
    _fields_ = [("num_states", C.c_int32), ("row", C.c_void_p), ("col", C.c_void_p),
                ("sets", C.c_void_p), ("attrs", C.c_void_p), ("num_aas", C.c_int32),
                ("aas", C.c_void_p), ("num_sod", C.c_int32), ("sod", C.c_void_p)]


class OracleStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "tested_non_aas", "matched_aas", "matched_advance", "successors", "reports",
        "enabled_items", "matched_items", "memo_list_items", "expanded_items",
        "expanded_tests", "root_positions", "level2_lists", "deepest_lists")]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_run_ex.restype = C.c_int
        L.oracle_run_ex.argtypes = [C.c_int32, C.POINTER(OracleGroup), C.c_void_p, C.c_uint32,
                                    C.c_uint32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                    C.POINTER(OracleStats)]
        _lib = L
    return _lib


def run(groups, data, chunk_len, num_seg=None, is_dup=False, threads=0, unit_mask=None,
        memo_depth=1, implicit_aas=True):
    """Reports (uint64 array, reference encoding, sorted unique) and stats.

    implicit_aas=False runs the reference's loop as written (every all-input
    state in the frontier, tested at every position: the CPU baseline);
    True leaves them out of the frontier after a chunk's first position and
    visits those whose class holds the byte, with the same reports and counts
    (checked by tests/test_oracle.py)."""
    buf = np.frombuffer(bytes(data), np.uint8) if isinstance(data, (bytes, bytearray)) \
        else np.ascontiguousarray(data, np.uint8)
    if num_seg is None:
        num_seg = buf.size // chunk_len
    assert chunk_len * num_seg <= buf.size
    og = (OracleGroup * len(groups))()
    for i, g in enumerate(groups):
        og[i] = OracleGroup(g.num_states, g.row.ctypes.data, g.col.ctypes.data if g.col.size else None,
                            g.sets.ctypes.data, g.attrs.ctypes.data, int(g.aas.size),
                            g.aas.ctypes.data if g.aas.size else None, int(g.sod.size),
                            g.sod.ctypes.data if g.sod.size else None)
    mask = None
    if unit_mask is not None:
        mask = np.ascontiguousarray(unit_mask, np.uint8)
        assert mask.size == len(groups) * num_seg
    n = C.c_uint64()
    st = OracleStats()
    L = lib()
    L.oracle_run_ex(len(groups), og, buf.ctypes.data, chunk_len, num_seg, int(is_dup),
                    mask.ctypes.data if mask is not None else None, int(threads),
                    int(memo_depth), int(bool(implicit_aas)), None, 0, C.byref(n), C.byref(st))
    out = np.zeros(n.value, np.uint64)
    L.oracle_run_ex(len(groups), og, buf.ctypes.data, chunk_len, num_seg, int(is_dup),
                    mask.ctypes.data if mask is not None else None, int(threads),
                    int(memo_depth), int(bool(implicit_aas)), out.ctypes.data, n.value,
                    C.byref(n), C.byref(st))
    stats = {k: getattr(st, k) for k, _ in OracleStats._fields_}
    return out, stats


def algorithmic_bytes_reference_model(stats, unit_bytes):
    """The reference algorithm's bytes (every all-input state tested at every
    position), summed over units: L + 4*T_ne + 4*M_aas + 8*M + 4*S + 8*R, where
    unit_bytes = L summed over (group, chunk) units."""
    return (unit_bytes + 4 * stats["tested_non_aas"] + 4 * stats["matched_aas"]
            + 8 * stats["matched_advance"] + 4 * stats["successors"] + 8 * stats["reports"])


def algorithmic_bytes(stats, unit_bytes):
    """Bytes a prefix-memoized engine (depth k, the stats' memo_depth) must move
    at minimum, summed over units: 1 B of input per (group, position); an 8 B
    memo index entry per position where an all-input state matches; 4 B per
    memo list entry materialised; per expanded item one 16 B successor record
    and a 4 B class word per non-all-input successor; 8 B per report."""
    return (unit_bytes + 8 * stats["root_positions"] + 4 * stats["memo_list_items"]
            + 16 * stats["expanded_items"] + 4 * stats["expanded_tests"]
            + 8 * stats["reports"])


def algorithmic_bytes_cross(stats, positions, memo_depth):
    """Bytes of the all-group engine (memo depth >= 2), which reads the
    cross-group index of the deepest level once per position instead of
    scanning every group: 1 B of input and an 8 B index slot per position.
    Depth 2: a 16 B index entry per non-empty list and 4 B per list entry;
    depth 3: 8 B per list entry, a (state, group) pair read straight from the
    3-gram slot's run.  Then the same expanded items and reports as
    algorithmic_bytes."""
    if memo_depth == 2:
        roots = 16 * stats["deepest_lists"] + 4 * stats["memo_list_items"]
    else:
        roots = 8 * stats["memo_list_items"]
    return (9 * positions + roots + 16 * stats["expanded_items"] + 4 * stats["expanded_tests"]
            + 8 * stats["reports"])


def ref_dump(anml_path, group_num, max_nfa_size=-1):
    """Group dump from the reference's own front-end (needs oracle/_ref)."""
    args = [REF_DUMPER, str(anml_path), str(group_num)]
    if max_nfa_size != -1:
        args.append(str(max_nfa_size))
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    m = re.search(r"^groups ", out, re.M)
    return out[m.start():]
