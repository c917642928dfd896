"""bench.py's contract on the host: one GPU only, a JSON line that says so."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_rocprof_timeline_union(tmp_path):
    """The timeline summary counts overlapping kernels once, splits copies by
    direction and cuts the run to its last milliseconds."""
    import sqlite3
    from ngap_amd import rocprof_summary as R
    db = str(tmp_path / "run_results.db")
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, start int, end int)")
    con.execute("create table memory_copies (name text, start int, end int, size int)")
    con.executemany("insert into kernels values (?, ?, ?)",
                    [("a", 0, 100), ("k_x", 1000, 2000), ("k_x", 1500, 2500), ("b", 3000, 3100)])
    con.executemany("insert into memory_copies values (?, ?, ?, ?)",
                    [("H2D", 900, 1100, 4096), ("H2D", 2400, 2600, 4096), ("D2H", 3050, 3200, 64)])
    con.commit()
    con.close()
    t = R.timeline(db, "k_x")
    assert t["span"] == 3200 - 1000
    assert t["kernel_busy"] == 1500 + 100          # [1000, 2500) and [3000, 3100)
    assert t["copy_busy"] == 200 + 150             # the copy from 900 starts before the cut
    assert t["any_busy"] == 1600 + 200             # [1000, 2600) plus [3000, 3200)
    assert t["copies"]["H2D"][:2] == (1, 4096) and t["copies"]["D2H"][:2] == (1, 64)
    t = R.timeline(db, tail_ns=300)
    assert t["span"] == 300 and t["kernel_busy"] == 100 and t["dispatches"] == 1
