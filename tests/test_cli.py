"""The `ngap` command-line drop-in: flags, stdout contract, validation."""
import os
import re
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT, have_gpu

NGAP = os.path.join(ROOT, "ngap_amd", "ngap")
APPLE = ["-a", os.path.join(GOLDEN, "apple.anml"), "-i", os.path.join(GOLDEN, "apple_input.txt"),
         "--app-name=apple", "--input-start-pos=0", "--input-len=81",
         "--split-entire-inputstream-to-chunk-size=81", "--group-num=1",
         "--duplicate-input-stream=1", "--unique=false", "--unique-frequency=10",
         "--use-soa=false", "--result-capacity=54619400", "--use-uvm=false",
         "--data-buffer-fetch-size=25600000", "--add-aan-start=256", "--add-aas-interval=32",
         "--active-threshold=10", "--precompute-cutoff=-1", "--precompute-depth=3",
         "--compress-prec-table=true", "--report-off=false", "--validation=true"]


def run(args, timeout=120):
    return subprocess.run([NGAP] + args, capture_output=True, text=True, timeout=timeout)


def test_rejects_unknown_flag_and_missing_inputs():
    r = run(["--no-such-flag=1"])
    assert r.returncode == 1 and "Error in command line" in r.stderr
    r = run(["--algorithm=NAPgroups"])
    assert r.returncode == 1


@pytest.mark.skipif(have_gpu(), reason="checks the front-end half without a device")
def test_front_end_half_without_device():
    r = run(APPLE + ["--algorithm=nonblockingallgroups"])
    assert "total_state_number = 5" in r.stdout
    assert "always_active_state_number = 1" in r.stdout
    assert "grouped_nfas.size = 1" in r.stdout
    assert "no HIP device" in r.stderr


def test_unknown_algorithm_message():
    r = run(APPLE + ["--algorithm=graph"])
    if "no HIP device" in r.stderr:
        pytest.skip("no device")
    assert "not supported algoritm graph" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("algo", ["blockinggroups", "NAPgroups", "nonblockinggroups", "nonblockingr1groups",
                                  "nonblockingr2groups", "nonblockingpcgroups",
                                  "nonblockingallgroups"])
@pytest.mark.parametrize("schedule", ["plain", "batched"])
def test_apple_readme_run(algo, schedule):
    print("[case] cli", algo, schedule, flush=True)
    r = run(APPLE + [f"--algorithm={algo}", f"--schedule={schedule}"])
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "Validation PASS!" in out
    assert "Result(4): \n0x400000005, 0x40000002f, 0x400000040, 0x40000004b, \n" in out
    assert "Reference result(4):" in out
    assert re.search(r"ngap elapsed time: \S+ seconds, throughput = \S+ MB/s", out)
    assert out.rstrip().endswith("FINISHED!")


@pytest.mark.gpu
def test_tuning_and_duplicated_input(tmp_path):
    print("[case] cli tuning + duplicate input", flush=True)
    r = run(APPLE[:4] + ["--algorithm=nonblockingallgroups", "--input-len=81",
                         "--split-entire-inputstream-to-chunk-size=81", "--group-num=1",
                         "--duplicate-input-stream=3", "--tuning=true"])
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Validation PASS!") == 2  # plain + autotuned batched
    assert "[Tuning] Tuning completed" in r.stdout
    # three copies of the input, three chunks, start-of-data would be per chunk
    assert "Unique results number: 12" in r.stdout
