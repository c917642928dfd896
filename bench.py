"""Benchmark: input MB/s of the non-blocking automata engine on one MI355X,
batched + learned schedule against the plain engine, bit-exact reports.

Default workload (learned): 1,024 automata, log-uniform 50..50,000 states (20 %
with a leading dot-star), 8 MiB in 64 KiB chunks whose byte distribution
drifts between Zipf s = 0.8 and 1.6, with pattern instances planted at 1e-4
per byte.  Other workloads: batch1024 (same automata, Zipf(1.1) input),
regex20k (20,000 patterns in one group, 16 MiB Zipf(1.1), 256 KiB chunks),
literal64.

A step is one pass of the engine over the whole input with inputs resident in
HBM: counter reset + matching launch + overflow check (+ fallback re-runs).
Engines timed on the same input: plain (one workgroup per (group, chunk)
unit), batched with the static schedule (hipGraph replay), batched with the
autotuned knobs, and batched with the learned schedule (the frontier-size MLP
trained on traces of a second input from the same generator, different seed).

Printed JSON (one line): value = input MB/s of the batched + learned engine
(wall clock over the timed steps), the other engines' MB/s and speedups, the
report-set checks, prediction MAE and overflow-fallback rate, the roofline of
the dominant kernel, the CPU baseline (the oracle on host cores over a bounded
sample), and the serving latencies.
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
    """(groups, test input, training input, chunk): the training input comes
    from the same generator with another seed (traces for the MLP)."""
    from ngap_amd import synth
    if name == "regex20k":
        groups, data, chunk = synth.regex20k()
        train = synth.zipf_bytes(np.random.default_rng(3001), len(data), 1.1)
    elif name == "literal64":
        groups, data, chunk = synth.literal64()
        train = synth.acgt_bytes(np.random.default_rng(3000), len(data))
    elif name == "batch1024":
        groups, data, chunk = synth.batch1024()
        train = synth.zipf_bytes(np.random.default_rng(3002), len(data), 1.1)
    elif name == "learned":
        groups, data, chunk = synth.learned()
        train = synth.drifting_input(np.random.default_rng(3003), len(data), chunk, 1e-4, groups)
    else:
        raise SystemExit(f"unknown workload {name}")
    return groups, data, train, chunk


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


def learned_schedule(eng, groups, train, test, chunk, depth):
    """Trains the frontier-size MLP on traces of `train` (at most 10^6
    (group, 4 KiB window) samples) and shapes the batched schedule of `test`
    from its predictions.  Leaves `test` loaded.  Returns (options, info)."""
    from ngap_amd import learned
    from ngap_amd.mlp import MLP
    stats = learned.group_stats(groups)
    eng.set_input(train, chunk)
    t0 = time.time()
    pairs, y = learned.collect_traces(eng, train, chunk, len(groups), depth)
    x = learned.features(learned.window_hist(train, chunk), stats, pairs)
    if x.shape[0] > 1000000:
        keep = np.random.default_rng(0).choice(x.shape[0], 1000000, replace=False)
        x, y = x[keep], y[keep]
    m = MLP(0, seed=7)
    epochs, batch = learned.train_settings(x.shape[0])
    loss = m.train(x, y, epochs=epochs, batch=batch, lr=2e-3)
    t_train = time.time() - t0
    eng.set_input(test, chunk)
    t1 = time.time()
    S = len(test) // chunk
    K = learned.windows_per_chunk(chunk)
    all_pairs = np.stack(np.meshgrid(np.arange(len(groups)), np.arange(S * K), indexing="ij"),
                         -1).reshape(-1, 2).astype(np.uint32)
    hist = learned.window_hist(test, chunk)
    pred = m.predict_pairs(hist, stats, all_pairs)
    waves = eng.batched_waves()
    cost = learned.predicted_cost(pred).reshape(len(groups), S * K)
    if depth >= 2 and len(groups) > 1:   # all-group work items
        tasks, tcost = learned.cross_schedule_from_costs(cost, S, chunk, waves)
    else:
        tasks, tcost = learned.schedule_from_costs(cost, len(groups), S, chunk, waves)
    spill = learned.spill_capacity(pred, tcost)
    t_sched = time.time() - t1
    # prediction quality on the test input's own traces
    tp, ty = learned.collect_traces(eng, test, chunk, len(groups), depth)
    tpred = m.predict_pairs(hist, stats, tp)
    info = {"train_traces": int(x.shape[0]), "train_loss": round(float(loss), 5),
            "train_s": round(t_train, 2), "schedule_host_s": round(t_sched, 3),
            "tasks": int(len(tasks)), "spill_capacity": int(spill), "waves": int(waves),
            "mae_log_items": round(float(np.mean(np.abs(tpred - ty))), 4),
            "mae_constant_predictor": round(float(np.mean(np.abs(np.mean(y) - ty))), 4)}
    m.close()
    opts = eng.options(schedule="batched", tasks=tasks, spill_capacity=spill, fallback=True,
                       use_graph=True, algorithm="nonblockingallgroups",
                       precompute_depth=depth, unique=True)
    return opts, info


def compare(name, depth, steps, warmup, learned_on=True, tune=True):
    """All engines on one workload; returns (result dict, per-engine runs,
    groups, test input, chunk)."""
    from ngap_amd import Engine
    t = time.time()
    groups, test, train, chunk = workload(name)
    total = (len(test) // chunk) * chunk
    log(f"workload {name}: {len(groups)} groups, {sum(g.num_states for g in groups)} states, "
        f"{total} bytes, chunk {chunk} ({time.time() - t:.1f} s)")
    eng = Engine(0)
    eng.load_groups(groups, max_precompute_depth=max(depth, 1))
    common = dict(algorithm="nonblockingallgroups", precompute_depth=depth, unique=True)
    runs = {}

    def run(label, opts):
        wall, kern, st = timed(eng, opts, steps, warmup)
        runs[label] = dict(wall_ms=wall, kernel_ms=kern, stats=st, reports=eng.reports())
        log(f"{name} {label}: {wall:.3f} ms/step wall, {kern:.3f} ms kernel, "
            f"{total / wall / 1e3:.1f} MB/s, items {st['items']}, tasks {st['num_tasks']}, "
            f"fallback {st['fallback_tasks']}, reports {runs[label]['reports'].size}")

    info = None
    if learned_on:
        lopts, info = learned_schedule(eng, groups, train, test, chunk, max(depth, 1))
        log(f"{name} learned schedule: {info}")
    else:
        eng.set_input(test, chunk)
    run("plain", eng.options(schedule="plain", **common))
    run("static", eng.options(schedule="batched", use_graph=True, **common))
    if tune:
        topts, tms = eng.autotune(max_runs=12, **common)
        topts.use_graph = 1
        run("tuned", topts)
        tuned_knobs = dict(window=int(topts.window), persistent_blocks=int(topts.persistent_blocks),
                           spill_capacity=int(topts.spill_capacity))
    if learned_on:
        run("learned", lopts)
    eng.close()
    ref = runs["plain"]["reports"]
    res = {"workload": name, "groups": len(groups),
           "states": int(sum(g.num_states for g in groups)), "input_bytes": int(total),
           "chunk": int(chunk), "reports": int(ref.size),
           "bit_exact_vs_plain": all(bool(np.array_equal(r["reports"], ref))
                                     for r in runs.values())}
    for k, r in runs.items():
        res[f"{k}_mbps"] = round(total / (r["wall_ms"] * 1e-3) / 1e6, 3)
        res[f"{k}_kernel_ms"] = round(r["kernel_ms"], 4)
        if k != "plain":
            res[f"speedup_{k}_vs_plain"] = round(runs["plain"]["wall_ms"] / r["wall_ms"], 3)
    if tune:
        res["tuned_knobs"] = tuned_knobs
    if info:
        st = runs["learned"]["stats"]
        res["learned"] = dict(info, fallback_tasks=int(st["fallback_tasks"]),
                              overflow_rate=st["fallback_tasks"] / max(1, st["num_tasks"]))
    return res, runs, groups, test, chunk


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
    # one unmeasured pass with everything queued, so every in-flight slot has
    # grown its device buffers before the sweep (steady-state serving)
    run_poisson(srv, [(0.0, g, d) for _, g, d in base])
    out = {"groups": len(groups), "requests": n_requests,
           "input_bytes": int(sum(len(d) for _, _, d in base)), "rates": []}
    for rate in rates:
        if np.isfinite(rate):
            rng = np.random.default_rng(int(rate))
            t = np.cumsum(rng.exponential(1.0 / rate, len(base)))
        else:   # every request queued at once: the server's capacity
            t = np.zeros(len(base))
        trace = [(float(t[i]), g, d) for i, (_, g, d) in enumerate(base)]
        res = run_poisson(srv, trace)
        exact = all(np.array_equal(res["reports"][i], alone[i]) for i in range(len(base)))
        row = {k: res[k] for k in ("p50_ms", "p99_ms", "mean_ms", "mbps", "makespan_s")}
        row.update(rate_per_s=rate if np.isfinite(rate) else "all at once",
                   bit_exact_vs_alone=exact)
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
    ap.add_argument("--workload", default="learned",
                    choices=["learned", "batch1024", "regex20k", "literal64"])
    ap.add_argument("--depth", type=int, default=3,
                    help="--precompute-depth of both engines (the reference's best setting is 3)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned schedule")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner")
    ap.add_argument("--also", default="batch1024",
                    help="comma-separated further workloads compared the same way, or 'none'")
    ap.add_argument("--serving", default="100,1000,10000,inf",
                    help="arrival rates (per second, inf = all at once) for the serving run; "
                         "'none' skips it")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle work counts")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration << 20, max(1, a.steps))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the (group, chunk) units would shard across GPUs with no exchange)"}))
        sys.exit(2)

    res, runs, groups, data, chunk = compare(a.workload, a.depth, a.steps, a.warmup,
                                             learned_on=not a.no_learned, tune=not a.no_tune)
    main_s = "learned" if "learned" in runs else "static"
    total = res["input_bytes"]
    wall = runs[main_s]["wall_ms"]
    kern = runs[main_s]["kernel_ms"]
    roofline = None
    if not a.no_check:
        depth = max(1, a.depth)
        per_byte, per_byte_ref, per_byte_cross, st, nbytes, dt, nunits = sample_stats(
            groups, data, chunk, depth)
        cross = depth >= 2 and len(groups) > 1
        model = per_byte_cross if cross else per_byte
        log(f"oracle sample: {nbytes} bytes over {nunits} chunks in {dt:.1f} s, "
            f"{model:.2f} algorithmic bytes per input byte ({per_byte:.1f} group by group, "
            f"{per_byte_ref:.1f} for the reference algorithm); {st}")
        achieved = model * total / (kern * 1e-3) / 1e9
        # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
        # on this workload and engine (profiles/), scaled by the copy calibration
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
        if os.path.exists(tpath):
            t = json.load(open(tpath)).get(a.workload)
            if t and t.get("engine") == main_s and t.get("precompute_depth") == a.depth:
                traffic = t["bytes_per_launch"]
        roofline = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                        frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                        traffic_gbs=round(traffic / (kern * 1e-3) / 1e9, 1) if traffic else None,
                        kernel="k_batched_stream",
                        work_items="all-group" if cross else "per-group",
                        algorithmic_bytes_per_input_byte=round(model, 3),
                        per_group_model_bytes_per_input_byte=round(per_byte, 3),
                        reference_model_bytes_per_input_byte=round(per_byte_ref, 2))
    out = {
        "metric": METRIC, "value": res[f"{main_s}_mbps"], "unit": "MB/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded generators, ngap_amd/synth.py)",
        "config": {"workload": a.workload, "groups": res["groups"], "states": res["states"],
                   "input_bytes": total, "chunk": res["chunk"],
                   "algorithm": "nonblockingallgroups", "precompute_depth": a.depth,
                   "engine": f"batched+{main_s}"},
        "kernel_ms": round(kern, 4),
        "plain_mbps": res["plain_mbps"],
        "speedup_vs_plain": round(runs["plain"]["wall_ms"] / wall, 3),
        "bit_exact_vs_plain": res["bit_exact_vs_plain"],
        "engines": res,
        "roofline": roofline,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(groups, data, chunk)
        log(f"cpu baseline: {out['cpu_baseline']}")
    del runs, groups, data
    if a.also != "none":
        out["also"] = [compare(w, a.depth, a.steps, a.warmup, learned_on=not a.no_learned,
                               tune=not a.no_tune)[0]
                       for w in a.also.split(",") if w and w != a.workload]
    if a.serving != "none":
        out["serving"] = serving_bench([float(r) for r in a.serving.split(",")])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
