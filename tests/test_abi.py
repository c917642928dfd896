"""The C-ABI library loads and exports every symbol include/ngap_amd.h
declares; the front-end entry points work without a GPU."""
import ctypes as C
import os
import re

import pytest

from ngap_amd import _lib
from tests.conftest import ROOT, have_gpu


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ngap_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ngap_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python mirror binds each of them
    assert set(syms) <= set(_lib.SIGNATURES), set(syms) - set(_lib.SIGNATURES)


def test_default_options_follow_reference_defaults():
    o = _lib.RunOpts()
    _lib.lib().ngap_default_opts(C.byref(o))
    # ngap_option.h:52-68, common_func.h:50-58
    assert o.data_buffer_fetch_size == 128 and o.add_aas_start == 0
    assert o.add_aas_interval == 1 and o.active_threshold == 20
    assert o.unique == 0 and o.precompute_depth == 0
    assert o.result_capacity == 54619400


@pytest.mark.skipif(have_gpu(), reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    h = C.c_void_p()
    code = _lib.lib().ngap_create(0, C.byref(h))
    assert code == _lib.NGAP_ERR_DEVICE
    assert b"device" in _lib.lib().ngap_last_error()


def test_null_arguments_are_errors():
    L = _lib.lib()
    assert L.ngap_run(None, None, None) == _lib.NGAP_ERR_INVALID
    assert L.ngap_frontend_load(None, 1, -1, 1, 0, None) == _lib.NGAP_ERR_INVALID
    assert L.ngap_frontend_free(None) == _lib.NGAP_OK


def test_every_entry_point_rejects_null_handles():
    """Each status-returning entry point, called with NULL handles and
    pointers (and zero sizes), returns NGAP_ERR_INVALID with a message before
    it touches a device; the destroy/free calls accept NULL like free()."""
    text = open(os.path.join(ROOT, "include", "ngap_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\bint\s+(ngap_[a-z_0-9]+)\s*\(", text)))
    assert len(names) >= 30
    raw = C.CDLL(_lib.LIB_PATH)   # no argtypes: every argument is passed as 0 / NULL
    raw.ngap_last_error.restype = C.c_char_p
    for n in names:
        f = getattr(raw, n)
        f.restype = C.c_int
        rc = f(*([None] * 10))
        if n.endswith(("_destroy", "_free")):
            assert rc == _lib.NGAP_OK, n
        else:
            assert rc == _lib.NGAP_ERR_INVALID, (n, rc)
            assert raw.ngap_last_error(), n


def test_ctypes_structs_match_the_header_layout(tmp_path):
    """Size and every field offset of the ctypes mirrors equal what a C
    compiler lays out from include/ngap_amd.h."""
    pairs = [("ngap_task", _lib.Task), ("ngap_group_desc", _lib.GroupDesc),
             ("ngap_run_opts", _lib.RunOpts), ("ngap_run_stats", _lib.RunStats),
             ("ngap_server_info", _lib.ServerInfo), ("ngap_image_summary", _lib.ImageInfo)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ngap_amd.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out.splitlines()}
    for cname, py in pairs:
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


_OOM_SCRIPT = r"""
import ctypes as C, os, resource, sys
L = C.CDLL(sys.argv[1])
L.ngap_last_error.restype = C.c_char_p
n = 200000
text = ('<anml version="1.0"><automata-network id="x">' + "".join(
    f'<state-transition-element id="s{i}" symbol-set="[ab]" start="all-input">'
    f'<activate-on-match element="s{(i + 1) % n}"/></state-transition-element>'
    for i in range(n)) + '</automata-network></anml>').encode()
vm = int(open("/proc/self/statm").read().split()[0]) * os.sysconf("SC_PAGE_SIZE")
resource.setrlimit(resource.RLIMIT_AS, (vm + (64 << 20), resource.getrlimit(resource.RLIMIT_AS)[1]))
h = C.c_void_p()
rc = L.ngap_frontend_load_text(text, 1, -1, C.byref(h))
print(rc, L.ngap_last_error().decode())
"""


def test_host_memory_exhaustion_is_a_status():
    """A C++ allocation failure inside an entry point (the front-end parsing a
    200,000-state network under a 64 MiB address-space margin) comes back as
    NGAP_ERR_HOST with a message, not as an exception through the C ABI."""
    import subprocess
    import sys
    out = subprocess.run([sys.executable, "-c", _OOM_SCRIPT, _lib.LIB_PATH],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    rc, msg = out.stdout.strip().split(" ", 1)
    assert int(rc) == _lib.NGAP_ERR_HOST and "memory" in msg, out.stdout
