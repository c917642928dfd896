"""ngap_run_noalloc (device input and report buffers from the caller, the
run's scratch memory in the library's fixed device array) against ngap_run's
reports on the same input, on consecutive calls with shrinking inputs."""
import ctypes as C

import numpy as np
import pytest

import oracle
from ngap_amd import _lib, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


class DeviceBuffer:
    """Device memory from the HIP runtime the library uses (the test's own
    allocation: the entry point under test allocates nothing)."""

    def __init__(self, nbytes):
        self.hip = C.CDLL("libamdhip64.so")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(max(1, nbytes))) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, C.c_void_p(arr.ctypes.data), C.c_size_t(arr.nbytes),
                                  C.c_int(1)) == 0

    def download(self, n, dtype):
        out = np.zeros(n, dtype)
        if n:
            assert self.hip.hipMemcpy(C.c_void_p(out.ctypes.data), self.p, C.c_size_t(out.nbytes),
                                      C.c_int(2)) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def run_noalloc(engine, opts, data, chunk, cap=1 << 20):
    d_in = DeviceBuffer(len(data))
    d_rep = DeviceBuffer(cap * 8)
    try:
        d_in.upload(np.asarray(data, np.uint8))
        st = _lib.RunStats()
        code = _lib.lib().ngap_run_noalloc(engine._h, C.byref(opts), d_in.p, chunk,
                                           len(data) // chunk, d_rep.p, cap, C.byref(st))
        assert code == _lib.NGAP_OK, _lib.lib().ngap_last_error()
        raw = d_rep.download(st.reports_raw, np.uint64)
    finally:
        d_in.free()
        d_rep.free()
    u = np.unique(raw)  # device encoding (offset << 32) | id, sorted
    return (u << np.uint64(32)) | (u >> np.uint64(32)), st


def test_noalloc_matches_run_on_consecutive_calls(engine):
    rng = np.random.default_rng(77)
    groups = [synth.regex_group(rng, int(s), dotstar=(i % 4 == 0))
              for i, s in enumerate([300, 4000, 900, 12000, 2500, 600])]
    engine.load_groups(groups, 3)
    big = synth.drifting_input(rng, 1 << 18, 1 << 14, 2e-3, groups)
    calls = [("all-group depth 3", big, 1 << 14, dict(precompute_depth=3)),
             ("all-group depth 3, smaller", big[: 1 << 16], 1 << 13, dict(precompute_depth=3)),
             ("per-group depth 2, smaller", big[: 1 << 15], 1 << 12,
              dict(precompute_depth=2, cross_groups=0)),
             ("all-group depth 3, window 512", big[: 1 << 17], 1 << 14,
              dict(precompute_depth=3, window=512))]
    for label, data, chunk, kw in calls:
        opts = engine.options(algorithm="nonblockingallgroups", schedule="batched", unique=True,
                              **kw)
        engine.set_input(data, chunk)
        engine.run(opts)
        want = engine.reports()
        oracle_want, _ = oracle.run(groups, data, chunk, memo_depth=kw["precompute_depth"])
        assert np.array_equal(want, oracle_want)
        say("noalloc", label, len(data), chunk, "reports", want.size)
        got, st = run_noalloc(engine, opts, data, chunk)
        assert want.size > 20
        assert np.array_equal(got, want), label
        assert st.num_tasks > 0 and st.overflow_units == 0


def test_noalloc_rejects_oversized_schedules(engine):
    rng = np.random.default_rng(78)
    groups = [synth.regex_group(rng, 500) for _ in range(8)]
    engine.load_groups(groups, 2)
    data = synth.zipf_bytes(rng, 1 << 18, 1.1)
    # 8 groups x 2^18 one-position work items: beyond the array's flags
    opts = engine.options(algorithm="nonblockingallgroups", schedule="batched",
                          precompute_depth=2, cross_groups=0, window=1)
    d_in = DeviceBuffer(data.size)
    d_rep = DeviceBuffer(8 << 16)
    d_in.upload(data)
    st = _lib.RunStats()
    code = _lib.lib().ngap_run_noalloc(engine._h, C.byref(opts), d_in.p, 1 << 18, 1, d_rep.p,
                                       1 << 16, C.byref(st))
    d_in.free()
    d_rep.free()
    assert code == _lib.NGAP_ERR_INVALID
    assert b"work items" in _lib.lib().ngap_last_error()
