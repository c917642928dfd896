"""Benchmark: input MB/s of the non-blocking automata engine on one MI355X.

Default workload (regex20k): 20,000 random regex patterns (~474k states, one
group, all components rooted at all-input states), 16 MiB of Zipf(1.1) input
in 256 KiB chunks.  A step is one launch over the whole input (plain schedule:
one workgroup per (group, chunk) unit; batched: persistent waves over 4 KiB
work items), inputs resident in HBM, reports kept on the device.

Printed JSON (one line): value = input MB/s of the batched engine (wall clock
over the timed steps, each step = counter reset + matching launch + overflow
check), plus the plain engine's MB/s, a bit-exactness check (batched == plain
in full, both == oracle on sampled chunks), the roofline of the dominant
kernel and the CPU baseline (the oracle timed on host cores).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "input MB/s, bit-exact reports: batched+learned vs plain ngAP; p50/p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def hip_sync():
    hip = C.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def workload(name):
    from ngap_amd import synth
    if name == "regex20k":
        groups, data, chunk = synth.regex20k()
    elif name == "literal64":
        groups, data, chunk = synth.literal64()
    elif name == "batch1024":
        groups, data, chunk = synth.batch1024()
    elif name == "learned":
        groups, data, chunk = synth.learned()
    else:
        raise SystemExit(f"unknown workload {name}")
    return groups, data, chunk


def sample_stats(groups, data, chunk, depth, n_units=4, sample_len=16384, threads=0):
    """Oracle work counts on the first sample_len bytes of n_units chunks,
    scaled to bytes per input byte (chunk-start effects are negligible)."""
    import oracle
    S = len(data) // chunk
    picks = np.linspace(0, S - 1, min(n_units, S)).astype(int)
    sample = np.concatenate([data[c * chunk: c * chunk + min(sample_len, chunk)] for c in picks])
    L = min(sample_len, chunk)
    t = time.time()
    _, st = oracle.run(groups, sample, L, len(picks), threads=threads, memo_depth=depth)
    dt = time.time() - t
    unit_bytes = sample.size * len(groups)
    per_byte = oracle.algorithmic_bytes(st, unit_bytes) / sample.size
    per_byte_ref = oracle.algorithmic_bytes_reference_model(st, unit_bytes) / sample.size
    per_byte_cross = oracle.algorithmic_bytes_cross(st, sample.size, depth) / sample.size
    return per_byte, per_byte_ref, per_byte_cross, st, sample.size, dt, len(picks)


def cpu_baseline(groups, data, chunk, budget_s=15.0):
    """The oracle (a port of the reference's CPU validator) timed on host
    cores over a bounded sample: one unit per core, each a slice of a chunk."""
    import oracle
    cores = min(16, os.cpu_count() or 1)
    # calibrate on 2 KiB of one unit
    L0 = min(2048, chunk)
    t = time.time()
    oracle.run(groups, data[:L0], L0, 1, threads=1)
    rate = L0 / max(1e-6, time.time() - t)  # bytes/s per core
    L = int(min(chunk, max(256, rate * budget_s)))
    S = len(data) // chunk
    units = min(cores, S)
    sample = np.concatenate([data[c * chunk: c * chunk + L] for c in range(units)])
    t = time.time()
    oracle.run(groups, sample, L, units, threads=units)
    dt = time.time() - t
    return dict(value=round(sample.size / dt / 1e6, 6), unit="MB/s", cores=units, kind="port",
                sample=f"{units} units x {L} bytes (first bytes of chunks 0..{units - 1}), "
                       f"{dt:.1f} s, oracle with unique=true semantics")


def copy_calibration(nbytes, reps):
    """Device-to-device copies of a known size: under rocprofv3 --pmc the
    FETCH_SIZE / WRITE_SIZE of this copy kernel give the counters' scale."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    hip_sync()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def learned_comparison(depth, steps, workload_name):
    """batch1024 / learned workloads: plain (one workgroup per unit) vs the
    batched engine with its static schedule vs the batched engine with the
    MLP-shaped schedule; MLP trained on traces of an input with a different
    seed.  Returns a dict (MB/s, speedups, MAE, overflow rate, bit-exactness)."""
    from ngap_amd import Engine, learned, synth
    from ngap_amd.mlp import MLP
    t0 = time.time()
    rng = np.random.default_rng(1002 if workload_name == "batch1024" else 1003)
    groups = synth.batch_groups(rng, 1024)
    chunk = 64 << 10
    n = 8 << 20
    if workload_name == "batch1024":
        test = synth.zipf_bytes(np.random.default_rng(2002), n, 1.1)
        train = synth.zipf_bytes(np.random.default_rng(3002), n, 1.1)
    else:
        test = synth.drifting_input(np.random.default_rng(2003), n, chunk, 1e-4, groups)
        train = synth.drifting_input(np.random.default_rng(3003), n, chunk, 1e-4, groups)
    log(f"{workload_name}: {sum(g.num_states for g in groups)} states, generated in "
        f"{time.time() - t0:.1f} s")
    eng = Engine(0)
    eng.load_groups(groups, max(depth, 1))
    stats = learned.group_stats(groups)
    # traces of the training input -> at most 10^6 samples
    eng.set_input(train, chunk)
    t1 = time.time()
    pairs, y = learned.collect_traces(eng, train, chunk, len(groups), depth)
    x = learned.features(learned.window_hist(train, chunk), stats, pairs)
    if x.shape[0] > 1000000:
        keep = np.random.default_rng(0).choice(x.shape[0], 1000000, replace=False)
        x, y = x[keep], y[keep]
    m = MLP(0, seed=7)
    loss = m.train(x, y, epochs=4, batch=4096, lr=2e-3)
    log(f"trained on {x.shape[0]} traces in {time.time() - t1:.1f} s, loss {loss:.4f}")
    # test input
    eng.set_input(test, chunk)
    total = len(test)
    common = dict(algorithm="nonblockingallgroups", precompute_depth=depth, unique=True)
    out = {"workload": workload_name, "groups": len(groups),
           "states": int(sum(g.num_states for g in groups)), "input_bytes": total,
           "chunk": chunk, "train_traces": int(x.shape[0]), "train_loss": loss}
    res = {}
    wall, kern, st = timed(eng, eng.options(schedule="plain", **common), steps, 1)
    res["plain"] = (wall, eng.reports())
    wall, kern, st = timed(eng, eng.options(schedule="batched", **common), steps, 1)
    res["static"] = (wall, eng.reports())
    t2 = time.time()
    S = total // chunk
    K = learned.windows_per_chunk(chunk)
    all_pairs = np.stack(np.meshgrid(np.arange(len(groups)), np.arange(S * K), indexing="ij"),
                         -1).reshape(-1, 2).astype(np.uint32)
    hist = learned.window_hist(test, chunk)
    pred = m.predict_pairs(hist, stats, all_pairs)
    cost = learned.predicted_cost(pred).reshape(len(groups), S * K)
    blocks_waves = 256 * 4 * 4
    if depth >= 2:   # all-group work items (cross-group root index)
        tasks, tcost = learned.cross_schedule_from_costs(cost, S, chunk, blocks_waves)
    else:
        tasks, tcost = learned.schedule_from_costs(cost, len(groups), S, chunk, blocks_waves)
    spill = learned.spill_capacity(pred, tcost)
    sched_s = time.time() - t2
    opts = eng.options(schedule="batched", tasks=tasks, spill_capacity=spill, fallback=True,
                       use_graph=True, **common)
    wall, kern, st = timed(eng, opts, steps, 1)
    res["learned"] = (wall, eng.reports())
    out["learned_fallback_tasks"] = int(st["fallback_tasks"])
    out["learned_overflow_rate"] = st["fallback_tasks"] / max(1, st["num_tasks"])
    out["learned_tasks"] = int(st["num_tasks"])
    out["learned_spill_capacity"] = spill
    out["schedule_host_s"] = sched_s
    # prediction quality on the test input
    tp, ty = learned.collect_traces(eng, test, chunk, len(groups), depth)
    tpred = m.predict_pairs(hist, stats, tp)
    out["mae_log_items"] = float(np.mean(np.abs(tpred - ty)))
    out["mae_constant_predictor"] = float(np.mean(np.abs(np.mean(y) - ty)))
    for k, (wall, _) in res.items():
        out[f"{k}_mbps"] = total / (wall * 1e-3) / 1e6
    out["speedup_learned_vs_plain"] = res["plain"][0] / res["learned"][0]
    out["speedup_static_vs_plain"] = res["plain"][0] / res["static"][0]
    out["bit_exact"] = bool(np.array_equal(res["plain"][1], res["static"][1]) and
                            np.array_equal(res["plain"][1], res["learned"][1]))
    out["reports"] = int(res["plain"][1].size)
    log(f"{workload_name}: {out}")
    m.close()
    eng.close()
    return out


def serving_bench(rates, n_requests=128):
    """Serving workload: 64 resident automata (1k..200k states), Poisson
    requests of 1 KiB..4 MiB Zipf(1.1) bytes; p50/p99 latency and aggregate
    MB/s per arrival rate; every request's reports checked against the same
    request processed alone."""
    from ngap_amd import Engine, synth
    from ngap_amd.serve import Server, run_poisson
    t0 = time.time()
    groups = synth.serving_groups()
    eng = Engine(0)
    eng.load_groups(groups, 2)
    log(f"serving: {sum(g.num_states for g in groups)} states loaded in {time.time() - t0:.1f} s")
    srv = Server(eng, budget_bytes=200 << 30, max_inflight=4, algorithm="nonblockingallgroups",
                 precompute_depth=2, unique=True)
    base = synth.serving_requests(seed=2004, n=n_requests)
    alone = []
    for _, g, d in base:
        rid = srv.submit(g, d)
        alone.append(srv.drain([rid])[rid][0])
    out = {"groups": len(groups), "requests": n_requests,
           "input_bytes": int(sum(len(d) for _, _, d in base)), "rates": []}
    for rate in rates:
        rng = np.random.default_rng(int(rate))
        t = np.cumsum(rng.exponential(1.0 / rate, len(base)))
        trace = [(float(t[i]), g, d) for i, (_, g, d) in enumerate(base)]
        res = run_poisson(srv, trace)
        exact = all(np.array_equal(res["reports"][i], alone[i]) for i in range(len(base)))
        row = {k: res[k] for k in ("p50_ms", "p99_ms", "mean_ms", "mbps", "makespan_s")}
        row.update(rate_per_s=rate, bit_exact_vs_alone=exact)
        out["rates"].append(row)
        log(f"serving rate {rate}/s: {row}")
    srv.close()
    eng.close()
    return out


def timed(eng, opts, steps, warmup):
    for _ in range(warmup):
        eng.run(opts)
    hip_sync()
    ks = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = eng.run(opts)
        ks.append(st["kernel_ms"])
    hip_sync()
    wall = (time.perf_counter() - t0) / steps * 1e3
    return wall, float(np.mean(ks)), st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="regex20k")
    ap.add_argument("--depth", type=int, default=2)
    ap.add_argument("--window", type=int, default=0,
                    help="batched work-item positions, 0 = the engine's choice")
    ap.add_argument("--cross-groups", type=int, default=1, choices=[0, 1, 2],
                    help="batched static schedule: all-group work items (1 = when several groups)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--only", choices=["plain", "batched"], default=None)
    ap.add_argument("--learned", default="learned", choices=["none", "batch1024", "learned"],
                    help="also run the plain / static / learned comparison on this workload")
    ap.add_argument("--serving", default="100,1000,10000",
                    help="arrival rates (per second) for the serving run; 'none' skips it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    ap.add_argument("--oracle-check", action="store_true",
                    help="also compare reports with the oracle on three full chunks (slow)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration << 20, max(1, a.steps))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the (group, chunk) units would shard across GPUs with no exchange)"}))
        sys.exit(2)

    from ngap_amd import Engine
    t = time.time()
    groups, data, chunk = workload(a.workload)
    total = (len(data) // chunk) * chunk
    log(f"workload {a.workload}: {len(groups)} groups, {sum(g.num_states for g in groups)} states, "
        f"{total} bytes, chunk {chunk} ({time.time() - t:.1f} s)")
    eng = Engine(0)
    t = time.time()
    eng.load_groups(groups, max_precompute_depth=max(a.depth, 1))
    eng.set_input(data, chunk)
    log(f"loaded ({time.time() - t:.1f} s)")

    common = dict(algorithm="nonblockingallgroups", precompute_depth=a.depth, unique=True)
    cross = a.depth >= 2 and (a.cross_groups == 2 or (a.cross_groups == 1 and len(groups) > 1))
    res = {}
    for sched in (["plain", "batched"] if a.only is None else [a.only]):
        opts = eng.options(schedule=sched, window=a.window, cross_groups=a.cross_groups, **common)
        wall, kern, st = timed(eng, opts, a.steps, a.warmup)
        res[sched] = dict(wall_ms=wall, kernel_ms=kern, stats=st, reports=eng.reports())
        log(f"{sched}: {wall:.3f} ms/step wall, {kern:.3f} ms kernel, "
            f"{total / wall / 1e3:.1f} MB/s, items {st['items']}, reports {res[sched]['reports'].size}")

    main_s = "batched" if "batched" in res else "plain"
    bit_exact = None
    if not a.no_check and len(res) == 2:
        bit_exact = bool(np.array_equal(res["plain"]["reports"], res["batched"]["reports"]))
    oracle_check = None
    per_byte = None
    if not a.no_check:
        per_byte, per_byte_ref, per_byte_cross, st, nbytes, dt, nunits = sample_stats(
            groups, data, chunk, max(1, a.depth))
        if main_s == "batched" and cross:
            per_byte = per_byte_cross
        log(f"oracle sample: {nbytes} bytes over {nunits} chunks in {dt:.1f} s, "
            f"{per_byte:.2f} algorithmic bytes per input byte at memo depth {a.depth} "
            f"({per_byte_ref:.1f} for the reference algorithm); {st}")
        # reports on sampled chunks: engine == oracle restricted to those chunks
        import oracle
        S = len(data) // chunk
        mask = np.zeros(len(groups) * S, np.uint8)
        picks = [0, S // 2, S - 1]
        for c in picks:
            mask[c * len(groups):(c + 1) * len(groups)] = 1
        if a.oracle_check or a.workload == "literal64":
            want, _ = oracle.run(groups, data, chunk, S, unit_mask=mask)
            got = res[main_s]["reports"]
            off = (got & np.uint64(0xFFFFFFFF)).astype(np.int64)
            keep = np.isin(off // chunk, picks)
            oracle_check = bool(np.array_equal(got[keep], want))
            log(f"oracle check on chunks {picks}: {oracle_check}")

    wall = res[main_s]["wall_ms"]
    kern = res[main_s]["kernel_ms"]
    value = total / (wall * 1e-3) / 1e6
    roofline = None
    if per_byte is not None:
        achieved = per_byte * total / (kern * 1e-3) / 1e9
        roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 5), traffic=None,
                        kernel="k_batched" if main_s == "batched" else "k_plain",
                        work_items="all-group" if main_s == "batched" and cross else "per-group",
                        algorithmic_bytes_per_input_byte=round(per_byte, 3),
                        reference_model_bytes_per_input_byte=round(per_byte_ref, 2))
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "MB/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded generators, ngap_amd/synth.py)",
        "config": {"workload": a.workload, "groups": len(groups),
                   "states": int(sum(g.num_states for g in groups)), "input_bytes": int(total),
                   "chunk": int(chunk), "algorithm": "nonblockingallgroups",
                   "precompute_depth": a.depth, "window": a.window,
                   "all_group_work_items": bool(cross)},
        "kernel_ms": round(kern, 4), "kernel_mbps": round(total / (kern * 1e-3) / 1e6, 3),
        "plain_mbps": round(total / (res["plain"]["wall_ms"] * 1e-3) / 1e6, 3) if "plain" in res else None,
        "batched_mbps": round(total / (res["batched"]["wall_ms"] * 1e-3) / 1e6, 3) if "batched" in res else None,
        "speedup_batched_vs_plain": round(res["plain"]["wall_ms"] / res["batched"]["wall_ms"], 3)
        if len(res) == 2 else None,
        "reports": int(res[main_s]["reports"].size),
        "bit_exact_batched_vs_plain": bit_exact, "bit_exact_vs_oracle_sampled": oracle_check,
        "roofline": roofline,
    }
    eng.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(groups, data, chunk)
        log(f"cpu baseline: {out['cpu_baseline']}")
    if a.learned != "none":
        out["learned"] = learned_comparison(max(1, a.depth), a.steps, a.learned)
    if a.serving != "none":
        out["serving"] = serving_bench([float(r) for r in a.serving.split(",")])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
