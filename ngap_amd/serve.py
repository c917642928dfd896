"""Serving loop over the engine's request server (C ABI ngap_server_*).

Requests (group, bytes) arrive over time; the server keeps the automata
resident, batches whatever is waiting into launches as slots and HBM budget
free up (continuous batching), and reports each request's latency from
submission to completion.  :func:`run_poisson` replays an arrival trace in
real time and returns p50/p99 latency and aggregate input MB/s.
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check

QUEUED, INFLIGHT, DONE, FAILED = 0, 1, 2, 3


class Server:
    def __init__(self, engine, budget_bytes=0, max_inflight=2, **run_opts):
        run_opts.setdefault("schedule", "batched")
        self._opts = engine.options(**run_opts)
        self._h = C.c_void_p()
        check(_lib.lib().ngap_server_create(engine._h, int(budget_bytes), int(max_inflight),
                                            C.byref(self._opts), C.byref(self._h)))
        self._engine = engine  # keep the engine alive

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().ngap_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, group, data):
        buf = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8)
                                   if isinstance(data, (bytes, bytearray)) else data, np.uint8)
        rid = C.c_uint64()
        check(_lib.lib().ngap_server_submit(self._h, int(group), buf.ctypes.data, buf.size,
                                            C.byref(rid)))
        return rid.value

    def poll(self, block=False):
        check(_lib.lib().ngap_server_poll(self._h, int(block)))

    def take(self, rid):
        """(state, reports or None, latency_ms).  Reports are sorted, unique,
        reference encoding with offsets relative to the request."""
        st, n, lat = C.c_int32(), C.c_uint64(), C.c_double()
        check(_lib.lib().ngap_server_take(self._h, rid, C.byref(st), None, 0, C.byref(n),
                                          C.byref(lat)))
        if st.value != DONE:
            return st.value, None, lat.value
        out = np.zeros(n.value, np.uint64)
        check(_lib.lib().ngap_server_take(self._h, rid, C.byref(st), out.ctypes.data, n.value,
                                          C.byref(n), C.byref(lat)))
        return st.value, out, lat.value

    def pending(self):
        q, f = C.c_uint64(), C.c_uint64()
        check(_lib.lib().ngap_server_pending(self._h, C.byref(q), C.byref(f)))
        return q.value, f.value

    def stats(self):
        """Memory and admission figures (ngap_server_stats) as a dict."""
        info = _lib.ServerInfo()
        check(_lib.lib().ngap_server_stats(self._h, C.byref(info)))
        return {f: getattr(info, f) for f, _ in info._fields_}

    def drain(self, ids):
        """Polls until every id is done; returns {id: (reports, latency_ms)}."""
        out = {}
        left = list(ids)
        while left:
            self.poll(block=True)
            still = []
            for rid in left:
                st, rep, lat = self.take(rid)
                if st == DONE:
                    out[rid] = (rep, lat)
                elif st == FAILED:
                    raise RuntimeError(f"request {rid} failed")
                else:
                    still.append(rid)
            left = still
        return out


def run_poisson(server, requests):
    """Replays (arrival_s, group, bytes) requests in real time.  Returns a
    dict with p50/p99 latency (ms), aggregate input MB/s over the makespan,
    and each request's reports in submission order."""
    t0 = time.perf_counter()
    ids = []
    total = 0
    for arrival, group, data in requests:
        while time.perf_counter() - t0 < arrival:
            server.poll(block=False)
        ids.append(server.submit(group, data))
        total += len(data)
        server.poll(block=False)
    res = server.drain(ids)
    span = time.perf_counter() - t0
    lat = np.array([res[i][1] for i in ids])
    return {
        "requests": len(ids),
        "input_bytes": int(total),
        "makespan_s": span,
        "mbps": total / span / 1e6,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "reports": [res[i][0] for i in ids],
    }
