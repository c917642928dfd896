"""Host check of the device image's index tables (memo rank tables, the
cross-group root index, start-of-data pairs) against brute-force memo lists:
tests/layout_check.cpp compiled with g++ against ngap_amd/csrc/layout.cpp (host
code only, no GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_image_tables_match_brute_force(tmp_path):
    exe = tmp_path / "layout_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-o", str(exe),
                    os.path.join(ROOT, "tests", "layout_check.cpp"),
                    os.path.join(ROOT, "ngap_amd", "csrc", "layout.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
