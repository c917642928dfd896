"""Python mirror of the engine's C ABI (include/ngap_amd.h).

Mirrors the reference launcher's flow (code/src/ngap/ngap.cu:1477-2077):
groups are loaded once (device image + memo tables), the concatenated input
is uploaded once, and each :meth:`Engine.run` is one timed launch whose
reports come back sorted by (offset, id) and de-duplicated, in the reference
encoding ``((state | group << 22) << 32) | offset``.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import ALGORITHMS, SCHED_BATCHED, SCHED_PLAIN, NgapError, check


class Engine:
    def __init__(self, device=0):
        self._h = C.c_void_p()
        check(_lib.lib().ngap_create(int(device), C.byref(self._h)))
        self.num_groups = 0
        self.max_depth = 0

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().ngap_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load_groups(self, groups, max_precompute_depth=2, build=None):
        """Builds the automata image: ``build`` "device" (HIP kernels),
        "host" (host threads) or None (ngap_load_groups' choice); both give
        the same arrays."""
        descs = (_lib.GroupDesc * len(groups))(*[g.desc() for g in groups])
        if build is None:
            check(_lib.lib().ngap_load_groups(self._h, descs, len(groups),
                                              int(max_precompute_depth)))
        else:
            where = {"device": _lib.BUILD_DEVICE, "host": _lib.BUILD_HOST}[build]
            check(_lib.lib().ngap_load_groups_ex(self._h, descs, len(groups),
                                                 int(max_precompute_depth), where))
        self.num_groups = len(groups)
        self.max_depth = max(1, int(max_precompute_depth))

    IMAGE_ARRAYS = ("meta", "succ4", "cls", "sod", "beg1", "beg2", "beg3", "nrep1", "nrep2",
                    "nrep3", "list1", "list2", "list3", "sel2", "sel3", "mask1", "xr1", "xe1",
                    "xmask1", "xr2", "xe2", "xsod", "xs3", "xp3", "x2rep")

    def image_info(self):
        info = _lib.ImageInfo()
        check(_lib.lib().ngap_image_info(self._h, C.byref(info)))
        return {f: getattr(info, f) for f, _ in info._fields_}

    def image_array(self, name):
        """A device image array as raw bytes (numpy uint8)."""
        n = C.c_uint64()
        check(_lib.lib().ngap_image_array(self._h, name.encode(), None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint8)
        if n.value:
            check(_lib.lib().ngap_image_array(self._h, name.encode(), out.ctypes.data, n.value,
                                              C.byref(n)))
        return out

    def set_input(self, data, chunk_len, num_seg=None):
        buf = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8)
                                   if isinstance(data, (bytes, bytearray)) else data,
                                   dtype=np.uint8)
        if num_seg is None:
            num_seg = buf.size // chunk_len
        need = int(chunk_len) * int(num_seg)
        if need > buf.size:
            raise ValueError("input shorter than chunk_len * num_seg")
        self._input = buf[:need]
        check(_lib.lib().ngap_set_input(self._h, self._input.ctypes.data, int(chunk_len),
                                        int(num_seg)))
        self.chunk_len, self.num_seg = int(chunk_len), int(num_seg)

    @staticmethod
    def options(algorithm="nonblockingallgroups", schedule="plain", precompute_depth=0,
                unique=False, is_dup=False, active_threshold=20, report_off=False,
                result_capacity=54619400, window=0, spill_capacity=0, persistent_blocks=0,
                use_graph=False, data_buffer_fetch_size=128, add_aas_start=0,
                add_aas_interval=1, tasks=None, fallback=False, cross_groups=1,
                ring_capacity=0, task_traces=False, root_pass=None, root_capacity=0):
        o = _lib.RunOpts()
        _lib.lib().ngap_default_opts(C.byref(o))
        if algorithm not in ALGORITHMS:
            raise ValueError(f"not supported algoritm {algorithm}")
        o.algorithm = ALGORITHMS[algorithm]
        o.schedule = SCHED_BATCHED if schedule == "batched" else SCHED_PLAIN
        o.precompute_depth = int(precompute_depth)
        o.unique = int(bool(unique))
        o.is_dup = int(bool(is_dup))
        o.active_threshold = int(active_threshold)
        o.report_off = int(bool(report_off))
        o.result_capacity = int(result_capacity)
        o.window = int(window)
        o.spill_capacity = int(spill_capacity)
        o.persistent_blocks = int(persistent_blocks)
        o.use_graph = int(bool(use_graph))
        o.data_buffer_fetch_size = int(data_buffer_fetch_size)
        o.add_aas_start = int(add_aas_start)
        o.add_aas_interval = int(add_aas_interval)
        o.fallback = int(bool(fallback))
        o.cross_groups = int(cross_groups)
        o.ring_capacity = int(ring_capacity)
        o.task_traces = int(bool(task_traces))
        if root_pass is not None:  # None keeps the library default
            o.root_pass = int(bool(root_pass))
        o.root_capacity = int(root_capacity)
        if tasks is not None:
            # group -1 (or ALL_GROUPS) = a work item for every group
            arr = np.ascontiguousarray(np.asarray(tasks, np.int64).astype(np.uint32)
                                       .reshape(-1, 4))
            o.tasks = arr.ctypes.data
            o.num_tasks = arr.shape[0]
            o._keep = arr  # keep the array alive with the options
        return o

    def batched_waves(self, ring_capacity=0):
        """Waves of the batched engine's persistent grid."""
        b, w = C.c_uint32(), C.c_uint32()
        check(_lib.lib().ngap_batched_grid(self._h, int(ring_capacity), C.byref(b), C.byref(w)))
        return b.value * w.value

    def autotune(self, max_runs=16, **kw):
        """Batched-engine knobs (window, LDS ring, persistent blocks, spill
        capacity, root pass, active threshold) timed on the loaded input;
        returns (options, best kernel ms)."""
        base = self.options(**kw)
        best = _lib.RunOpts()
        ms = C.c_double()
        check(_lib.lib().ngap_autotune(self._h, C.byref(base), int(max_runs), C.byref(best),
                                       C.byref(ms)))
        return best, ms.value

    def run(self, opts=None, **kw):
        """One launch.  Returns a dict of stats; raises NgapError on failure
        (overflow included: ``err.code == NGAP_ERR_OVERFLOW``)."""
        if opts is None:
            opts = self.options(**kw)
        st = _lib.RunStats()
        code = _lib.lib().ngap_run(self._h, C.byref(opts), C.byref(st))
        if code == _lib.NGAP_OK and opts.tasks:
            # the options own a private copy of their work items, so later runs
            # with them may skip the engine's comparison of the work items
            opts.tasks_id = st.tasks_id
        stats = dict(kernel_ms=st.kernel_ms, reports_raw=st.reports_raw,
                     overflow_units=st.overflow_units, items=st.items,
                     num_units=st.num_units, num_tasks=st.num_tasks,
                     fallback_tasks=st.fallback_tasks, fallback_ms=st.fallback_ms,
                     tasks_id=st.tasks_id)
        if code != _lib.NGAP_OK:
            err = NgapError(code, _lib.lib().ngap_last_error().decode(errors="replace"))
            err.stats = stats
            raise err
        return stats

    def reports(self):
        """Sorted unique reports of the last run (uint64, reference encoding)."""
        n = C.c_uint64()
        check(_lib.lib().ngap_get_reports(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint64)
        if n.value:
            check(_lib.lib().ngap_get_reports(self._h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def tasks(self):
        """Work items of the last batched run: uint32 [n, 4] (group, chunk,
        begin, end), begin/end relative to the chunk start, in run order."""
        n = C.c_uint64()
        check(_lib.lib().ngap_get_tasks(self._h, None, 0, C.byref(n)))
        out = np.zeros((n.value, 4), np.uint32)
        if n.value:
            check(_lib.lib().ngap_get_tasks(self._h, out.ctypes.data, n.value, C.byref(n)))
        return out

    def task_items(self):
        n = C.c_uint64()
        check(_lib.lib().ngap_get_task_items(self._h, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint64)
        if n.value:
            check(_lib.lib().ngap_get_task_items(self._h, out.ctypes.data, n.value, C.byref(n)))
        return out


def decode_reports(reports):
    """(offset, group, state) arrays from reference-encoded reports."""
    r = np.asarray(reports, np.uint64)
    hi = (r >> np.uint64(32)).astype(np.uint32)
    return (r & np.uint64(0xFFFFFFFF)).astype(np.uint32), hi >> 22, hi & ((1 << 22) - 1)
