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


@pytest.mark.gpu
def test_write_reports_without_validation(tmp_path):
    print("[case] cli --write-reports", flush=True)
    out = tmp_path / "reports.txt"
    args = [a for a in APPLE if not a.startswith("--validation")]
    r = run(args + ["--algorithm=nonblockingallgroups", "--validation=false",
                    "--write-reports=true", f"--report-filename={out}"])
    assert r.returncode == 0, r.stderr
    assert "Reports written: 4" in r.stdout
    assert out.read_text().split() == ["0x400000005", "0x40000002f", "0x400000040",
                                       "0x40000004b"]
    assert "accepted and have no effect" in r.stderr  # the reference's AAS knobs


@pytest.mark.gpu
def test_quick_validation_counts(tmp_path):
    print("[case] cli --quick-validation", flush=True)
    r = run(APPLE + ["--algorithm=nonblockingallgroups", "--quick-validation=4"])
    assert r.returncode == 0, r.stderr
    assert "Quick Validation PASS!" in r.stdout
    assert re.search(r"throughput = \S+ MB/s", r.stdout)
    r = run(APPLE + ["--algorithm=nonblockingallgroups", "--quick-validation=5"])
    assert "Quick Validation FAILED!" in r.stdout and "WRONGTHRPUT(" in r.stdout


def _bursty_anml(path):
    """An all-input state whose successors fan out 4-way over three levels of
    all-byte states (84 states): every position enables 84 items, so a
    one-item worklist spill area overflows."""
    from ngap_amd.groups import write_anml
    states, level, nid = [], ["r"], 0
    states.append(dict(id="r", set="*", start="all-input", report=False, next=[]))
    by_id = {"r": states[0]}
    for depth in range(3):
        nxt = []
        for parent in level:
            for _ in range(4):
                sid = f"s{nid}"
                nid += 1
                st = dict(id=sid, set="*", start="", report=depth == 2, next=[])
                states.append(st)
                by_id[sid] = st
                by_id[parent]["next"].append(sid)
                nxt.append(sid)
        level = nxt
    write_anml(path, states)


@pytest.mark.gpu
def test_try_adaptive_aas_retries_after_overflow(tmp_path):
    print("[case] cli --try-adaptive-aas", flush=True)
    anml = tmp_path / "bursty.anml"
    _bursty_anml(anml)
    data = tmp_path / "in.bin"
    data.write_bytes(bytes(range(256)) * 16)
    args = ["-a", str(anml), "-i", str(data), "--algorithm=nonblockingallgroups",
            "--group-num=1", "--split-entire-inputstream-to-chunk-size=4096",
            "--precompute-depth=1", "--spill-capacity=1", "--validation=true"]
    r = run(args)
    assert "WRONGTHRPUT(" in r.stdout and "Try adaptive aas" not in r.stdout
    r = run(args + ["--try-adaptive-aas=true"])
    assert r.returncode == 0, r.stderr
    assert "worklist spill capacity exceeded" in r.stdout
    assert "Try adaptive aas" in r.stdout
    assert r.stdout.count("Validation PASS!") == 1
    assert r.stdout.rstrip().endswith("FINISHED!")


@pytest.mark.gpu
def test_regex20k_anml_through_the_cli(tmp_path):
    """The regex20k automaton (474k states) written as ANML and scanned by the
    CLI with validation: the front-end, the device image and the engine on a
    4 MiB slice of its input with planted pattern instances."""
    import numpy as np
    from ngap_amd import synth
    print("[case] cli regex20k anml", flush=True)
    groups, data, chunk = synth.regex20k()
    anml = tmp_path / "regex20k.anml"
    synth.write_groups_anml(str(anml), groups)
    inp = tmp_path / "regex20k.bin"
    sl = synth.plant(np.random.default_rng(3103), data[: 4 << 20].copy(), groups, 1e-4)
    inp.write_bytes(sl.tobytes())
    r = run(["-a", str(anml), "-i", str(inp), "--algorithm=nonblockingallgroups",
             "--group-num=1", f"--split-entire-inputstream-to-chunk-size={chunk}",
             "--precompute-depth=3", "--unique=true", "--validation=true"], timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Validation PASS!" in r.stdout, r.stdout[-2000:]
    n = int(re.search(r"Unique results number: (\d+)", r.stdout).group(1))
    assert n > 100
    assert "total_state_number = 473702" in r.stdout
