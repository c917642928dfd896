"""The host-side front-end, image builder and the CLI's --validation check
under AddressSanitizer and UBSan (tests/asan_host_check.cpp, compiled with g++
on the CPU): every ANML fixture through the pipeline and the builder at depths
1-3, the CLI's check against the oracle on seeded inputs, and malformed groups
that must be rejected."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_frontend_and_builder_under_asan(tmp_path):
    exe = tmp_path / "asan_host_check"
    src = os.path.join(ROOT, "ngap_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-pthread", "-o", str(exe),
                    os.path.join(ROOT, "tests", "asan_host_check.cpp"),
                    os.path.join(src, "frontend.cpp"), os.path.join(src, "layout.cpp"),
                    os.path.join(src, "cli_validate.cpp"), "-fopenmp"],
                   check=True)
    anml = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.anml")) +
                  glob.glob(os.path.join(ROOT, "tests", "golden", "frontend", "*.anml")))
    assert anml
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(exe)] + anml, capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and out.stdout.strip().endswith("PASS"), \
        (out.stdout + out.stderr)[-4000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_mutated_anml_under_asan(tmp_path):
    """Seeded mutations of every ANML fixture and of symbol sets: errors, never
    crashes or undefined behaviour (tests/fuzz_anml_check.cpp)."""
    exe = tmp_path / "fuzz_anml_check"
    src = os.path.join(ROOT, "ngap_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-pthread", "-o", str(exe),
                    os.path.join(ROOT, "tests", "fuzz_anml_check.cpp"),
                    os.path.join(src, "frontend.cpp"), os.path.join(src, "layout.cpp")],
                   check=True)
    anml = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.anml")) +
                  glob.glob(os.path.join(ROOT, "tests", "golden", "frontend", "*.anml")))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(exe), "600"] + anml, capture_output=True, text=True, timeout=600,
                         env=env)
    assert out.returncode == 0 and out.stdout.startswith("PASS"), (out.stdout + out.stderr)[-4000:]
    parsed = int(out.stdout.split()[2])
    assert parsed > 50   # some mutants still parse and go through grouping and the builder
