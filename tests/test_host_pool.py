"""The host thread pool behind the pinned staging copies (server submit and
ngap_set_input), and the guard that turns host exceptions into NGAP_ERR_HOST
at the C ABI, compiled on the CPU with g++."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_pool_parts_and_copies(tmp_path):
    exe = tmp_path / "host_pool_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-o", str(exe),
                    os.path.join(ROOT, "tests", "host_pool_check.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("PASS"), out.stdout + out.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_pool_under_thread_sanitizer(tmp_path):
    exe = tmp_path / "host_pool_tsan"
    build = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
                            "-o", str(exe), os.path.join(ROOT, "tests", "host_pool_check.cpp")],
                           capture_output=True, text=True)
    if build.returncode != 0:
        pytest.skip("no ThreadSanitizer runtime: " + build.stderr[-200:])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    if "unexpected memory mapping" in out.stderr:
        pytest.skip("ThreadSanitizer cannot run on this kernel's memory layout")
    assert out.returncode == 0 and out.stdout.startswith("PASS"), out.stdout + out.stderr[-4000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_c_guard_turns_exceptions_into_status(tmp_path):
    exe = tmp_path / "c_guard_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", str(exe),
                    os.path.join(ROOT, "tests", "c_guard_check.cpp"),
                    os.path.join(ROOT, "ngap_amd", "csrc", "errors.cpp")], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.startswith("PASS"), out.stdout + out.stderr
