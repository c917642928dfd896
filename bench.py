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
unit), batched with the static schedule, batched one group per work item,
batched with the autotuned knobs, and batched with the learned schedule (the
frontier-size MLP trained on traces of a second input from the same
generator, different seed), with and without its per-group shape.

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


def cpu_info():
    """Host CPU model, online CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = nproc
    return model, nproc, usable


def host_threads(units):
    """Oracle threads: min(units, usable CPUs, 16).  The reference gives every
    (group, chunk) unit a thread (ngap.cu:2190); 16 is the host CPU share of
    one GPU on the measuring machine."""
    _, _, usable = cpu_info()
    return max(1, min(units, usable, 16))


def spread_chunks(S, n):
    return sorted(set(np.linspace(0, S - 1, min(n, S)).round().astype(int).tolist()))


def sample_stats(groups, data, chunk, depth, max_chunks=32):
    """Oracle work counts over whole chunks (the whole input, or max_chunks
    chunks spread over it, at least 1 MiB for the benchmark workloads), with
    the all-input states implicit (same counts as the reference loop, much
    faster).  Returns the byte models per input byte and the sample."""
    import oracle
    S = len(data) // chunk
    G = len(groups)
    picks = spread_chunks(S, max_chunks)
    mask = np.zeros(G * S, np.uint8)
    for c in picks:
        mask[c * G:(c + 1) * G] = 1
    threads = host_threads(G * len(picks))
    t = time.time()
    rep, st = oracle.run(groups, data, chunk, S, unit_mask=mask, threads=threads,
                         memo_depth=depth, implicit_aas=True)
    dt = time.time() - t
    positions = len(picks) * chunk
    unit_bytes = positions * G
    return dict(
        per_group=oracle.algorithmic_bytes(st, unit_bytes) / positions,
        reference=oracle.algorithmic_bytes_reference_model(st, unit_bytes) / positions,
        cross=oracle.algorithmic_bytes_cross(st, positions, depth) / positions,
        # engine-independent: the input once and each distinct report once
        floor=(positions + 8 * st["reports"]) / positions,
        stats=st, positions=positions, chunks=len(picks), seconds=dt, threads=threads,
        reports=int(rep.size))


def cpu_baseline(groups, data, chunk, depth, budget_s=25.0):
    """The oracle running the reference's CPU validator loop as written
    (every all-input state tested at every position, ngap.cu:2171-2268) on
    host threads over a bounded sample of whole-chunk prefixes: one unit per
    (group, chunk), threads = min(units, usable CPUs, 16); about `budget_s`
    seconds, and at least 1 MiB of input where that takes under 40 s."""
    import oracle
    model, nproc, usable = cpu_info()
    S = len(data) // chunk
    G = len(groups)
    # calibrate on 2 KiB of one chunk, one thread
    L0 = min(2048, chunk)
    t = time.time()
    oracle.run(groups, data[:L0], L0, 1, threads=1, implicit_aas=False, memo_depth=depth)
    rate = L0 / max(1e-6, time.time() - t)  # input bytes/s of one thread over all groups
    picks = spread_chunks(S, 16)
    threads = host_threads(G * len(picks))
    L = max(256, rate * threads * budget_s / len(picks))
    mib = (1 << 20) / len(picks)
    if L < mib and mib * len(picks) / (rate * threads) < 40.0:
        L = mib
    L = int(min(chunk, L))
    sample = np.concatenate([data[c * chunk: c * chunk + L] for c in picks])
    units = G * len(picks)
    threads = host_threads(units)
    t = time.time()
    oracle.run(groups, sample, L, len(picks), threads=threads, implicit_aas=False,
               memo_depth=depth)
    dt = time.time() - t
    return dict(value=round(sample.size / dt / 1e6, 6), unit="MB/s", cores=threads,
                kind="port", cpu_model=model, nproc=nproc, usable_cpus=usable,
                threads_rule="min((group, chunk) units, usable CPUs, 16)",
                sample=f"first {L} bytes of {len(picks)} chunks spread over the input "
                       f"({sample.size} bytes x {G} groups = {units} units), {dt:.1f} s; "
                       f"the reference's CPU validator loop with unique=true semantics")


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


# Batched engines run with --active-threshold 0, one of the reference's tuning
# values (ngap.cu:1999) and the one the autotuner picks on the 1,024-automaton
# workloads; the plain engine keeps the reference default.
BATCHED_THRESHOLD = 0
# A run is one kernel launch: a direct launch costs about 8 us less per run
# than replaying it from a captured graph (0.264 vs 0.272 ms wall per run on
# the learned workload), so the bench launches directly.
USE_GRAPH = False


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
    # per-group work items, and all-group ones where the cross-group index
    # exists (depth >= 2, one group included): the headline schedule
    ptasks, pcost = learned.schedule_from_costs(cost, len(groups), S, chunk, waves)
    if depth >= 2:
        # 1.5 work items per wave (more cost more than their balance saves,
        # fewer leave waves idle), claimed in input order (neighbouring
        # windows share their byte distribution, so the index lines they touch)
        tasks, tcost = learned.cross_schedule_from_costs(cost, S, chunk, waves, per_wave=1.5,
                                                         order="position")
    else:
        tasks, tcost = ptasks, pcost
    spill = learned.spill_capacity(pred, tcost)
    pspill = learned.spill_capacity(pred, pcost)
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
                       use_graph=USE_GRAPH, algorithm="nonblockingallgroups",
                       precompute_depth=depth, unique=True,
                       active_threshold=BATCHED_THRESHOLD)
    popts = None
    if depth >= 2:
        popts = eng.options(schedule="batched", tasks=ptasks, spill_capacity=pspill,
                            fallback=True, use_graph=USE_GRAPH, algorithm="nonblockingallgroups",
                            precompute_depth=depth, unique=True,
                            active_threshold=BATCHED_THRESHOLD)
        info["pergroup_tasks"] = int(len(ptasks))
    return opts, popts, info


def roofline_of(groups, data, chunk, depth, kern_ms, engine_label, workload, traffic_key):
    """Roofline object of the dominant kernel (k_batched_stream) for one
    workload: algorithmic bytes from the oracle's exact counts over whole
    chunks divided by the kernel's HIP-event time; traffic from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary when one exists for this
    workload and engine."""
    ss = sample_stats(groups, data, chunk, depth)
    cross = depth >= 2
    model = ss["cross"] if cross else ss["per_group"]
    total = (len(data) // chunk) * chunk
    log(f"{workload} oracle sample: {ss['chunks']} whole chunks ({ss['positions']} bytes) in "
        f"{ss['seconds']:.1f} s on {ss['threads']} threads: {model:.2f} algorithmic bytes per "
        f"input byte ({ss['per_group']:.1f} group by group, {ss['reference']:.1f} for the "
        f"reference algorithm, floor {ss['floor']:.4f}); {ss['stats']}")
    achieved = model * total / (kern_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath)).get(traffic_key)
        if t and t.get("engine") == engine_label and t.get("precompute_depth") == depth:
            traffic = t["bytes_per_launch"]
    return dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                traffic_gbs=round(traffic / (kern_ms * 1e-3) / 1e9, 1) if traffic else None,
                kernel="k_batched_stream", kernel_ms=round(kern_ms, 4),
                work_items="all-group" if cross else "per-group",
                algorithmic_bytes_per_input_byte=round(model, 3),
                floor_bytes_per_input_byte=round(ss["floor"], 5),
                floor_frac=round(ss["floor"] * total / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 6),
                per_group_model_bytes_per_input_byte=round(ss["per_group"], 3),
                reference_model_bytes_per_input_byte=round(ss["reference"], 2),
                sample=f"{ss['chunks']} whole chunks, {ss['positions']} bytes")


def end_to_end(eng, data, chunk, opts, steps):
    """Input upload + matching + report download, sort and unique per step
    (wall clock), with the time of each part."""
    t_in = t_run = t_out = 0.0
    n = 0
    for _ in range(steps):
        t0 = time.perf_counter()
        eng.set_input(data, chunk)
        t1 = time.perf_counter()
        eng.run(opts)
        t2 = time.perf_counter()
        n = eng.reports().size
        t3 = time.perf_counter()
        t_in += t1 - t0
        t_run += t2 - t1
        t_out += t3 - t2
    total = (len(data) // chunk) * chunk
    wall = (t_in + t_run + t_out) / steps
    return dict(mbps=round(total / wall / 1e6, 3), ms_per_step=round(wall * 1e3, 4),
                upload_ms=round(t_in / steps * 1e3, 4), run_ms=round(t_run / steps * 1e3, 4),
                reports_ms=round(t_out / steps * 1e3, 4), reports=int(n))


def compare(name, depth, steps, warmup, learned_on=True, tune=True, repeats=5, check=True,
            cpu=True, only=None):
    """All engines on one workload, each timed `repeats` times (interleaved)
    over `steps` steps; returns (result dict, per-engine runs, groups, input,
    chunk).  With `only`, that engine alone and no end-to-end leg: its
    launches are then the last ones of the matching kernel, what a profiled
    run summarises."""
    from ngap_amd import Engine
    t = time.time()
    groups, test, train, chunk = workload(name)
    total = (len(test) // chunk) * chunk
    log(f"workload {name}: {len(groups)} groups, {sum(g.num_states for g in groups)} states, "
        f"{total} bytes, chunk {chunk} ({time.time() - t:.1f} s)")
    eng = Engine(0)
    # the image built by host threads (then uploaded), then by the device
    # builder, which the runs use
    load = {}
    for where in ("host", "device"):
        eng.load_groups(groups, max_precompute_depth=max(depth, 1), build=where)
        iinfo = eng.image_info()
        load[where] = round(iinfo["build_ms"], 1)
    log(f"{name}: image of {iinfo['bytes'] / 1e6:.1f} MB built in {load['host']} ms on the host, "
        f"{load['device']} ms on the device")
    common = dict(algorithm="nonblockingallgroups", precompute_depth=depth, unique=True)
    info = None
    engines = {}
    lpopts = None
    if learned_on:
        lopts, lpopts, info = learned_schedule(eng, groups, train, test, chunk, max(depth, 1))
        log(f"{name} learned schedule: {info}")
    else:
        eng.set_input(test, chunk)
    engines["plain"] = eng.options(schedule="plain", **common)
    engines["static"] = eng.options(schedule="batched", use_graph=USE_GRAPH,
                                    active_threshold=BATCHED_THRESHOLD, **common)
    if depth >= 2:
        # batched, one group per work item (no cross-group index)
        engines["pergroup"] = eng.options(schedule="batched", use_graph=USE_GRAPH, cross_groups=0,
                                          active_threshold=BATCHED_THRESHOLD, **common)
    tuned_knobs = None
    if tune:
        topts, tms = eng.autotune(max_runs=24, **common)
        topts.use_graph = int(USE_GRAPH)
        engines["tuned"] = topts
        tuned_knobs = dict(window=int(topts.window), persistent_blocks=int(topts.persistent_blocks),
                           spill_capacity=int(topts.spill_capacity),
                           ring_capacity=int(topts.ring_capacity), root_pass=int(topts.root_pass),
                           active_threshold=int(topts.active_threshold))
    if learned_on:
        engines["learned"] = lopts
        if lpopts is not None:
            # the learned schedule of per-group work items, next to "pergroup"
            engines["learned_pergroup"] = lpopts
    if only:
        engines = {only: engines[only]}
    runs = {k: dict(walls=[], kerns=[]) for k in engines}
    for rep in range(repeats):
        for label, opts in engines.items():
            if label == "plain" and rep > 0:
                continue  # slow and steady: one repetition
            wall, kern, st = timed(eng, opts, steps, warmup if rep == 0 else 1)
            r = runs[label]
            r["walls"].append(wall)
            r["kerns"].append(kern)
            r["stats"] = st
            r["reports"] = eng.reports()
    for label, r in runs.items():
        r["wall_ms"] = float(np.median(r["walls"]))
        r["kernel_ms"] = float(np.median(r["kerns"]))
        st = r["stats"]
        log(f"{name} {label}: {r['wall_ms']:.3f} ms/step wall (runs {['%.4f' % w for w in r['walls']]}), "
            f"{r['kernel_ms']:.4f} ms kernel, {total / r['wall_ms'] / 1e3:.1f} MB/s, "
            f"items {st['items']}, tasks {st['num_tasks']}, fallback {st['fallback_tasks']}, "
            f"reports {r['reports'].size}")
    main_s = "learned" if learned_on else "static"
    if only:
        eng.close()
        return dict(workload=name, engine=only, kernel_ms=runs[only]["kernel_ms"],
                    wall_ms=runs[only]["wall_ms"], steps=steps, warmup=warmup,
                    reports=int(runs[only]["reports"].size)), runs, groups, test, chunk
    e2e = end_to_end(eng, test, chunk, engines[main_s], max(3, steps // 2))
    log(f"{name} end to end ({main_s}): {e2e}")
    eng.close()
    ref = runs["plain"]["reports"]
    res = {"workload": name, "groups": len(groups),
           "states": int(sum(g.num_states for g in groups)), "input_bytes": int(total),
           "chunk": int(chunk), "reports": int(ref.size), "load_groups_ms": load,
           "image_bytes": int(iinfo["bytes"]),
           "bit_exact_vs_plain": all(bool(np.array_equal(r["reports"], ref))
                                     for r in runs.values()),
           "repeats": repeats}
    for k, r in runs.items():
        res[f"{k}_mbps"] = round(total / (r["wall_ms"] * 1e-3) / 1e6, 3)
        res[f"{k}_kernel_ms"] = round(r["kernel_ms"], 4)
        if len(r["walls"]) > 1:
            res[f"{k}_mbps_runs"] = [round(total / (w * 1e-3) / 1e6, 1) for w in r["walls"]]
        if k != "plain":
            res[f"speedup_{k}_vs_plain"] = round(runs["plain"]["wall_ms"] / r["wall_ms"], 3)
    if "learned" in runs and "static" in runs:
        res["learned_vs_static"] = round(runs["static"]["wall_ms"] / runs["learned"]["wall_ms"], 4)
    if "learned_pergroup" in runs and "pergroup" in runs:
        res["learned_pergroup_vs_pergroup"] = round(
            runs["pergroup"]["wall_ms"] / runs["learned_pergroup"]["wall_ms"], 4)
    res["end_to_end"] = dict(e2e, engine=main_s,
                             what="set_input (H2D) + run + reports (D2H, sort, unique)")
    if tuned_knobs:
        res["tuned_knobs"] = tuned_knobs
    if info:
        st = runs["learned"]["stats"]
        res["learned"] = dict(info, fallback_tasks=int(st["fallback_tasks"]),
                              overflow_rate=st["fallback_tasks"] / max(1, st["num_tasks"]))
    if check:
        res["roofline"] = roofline_of(groups, test, chunk, max(1, depth),
                                      runs[main_s]["kernel_ms"], main_s, name, name)
    if cpu:
        res["cpu_baseline"] = cpu_baseline(groups, test, chunk, max(1, depth))
        log(f"{name} cpu baseline: {res['cpu_baseline']}")
    return res, runs, groups, test, chunk


ALGORITHMS = ["blockinggroups", "NAPgroups", "nonblockinggroups", "nonblockingr1groups",
              "nonblockingr2groups", "nonblockingpcgroups", "nonblockingallgroups"]


def breakdown(name, depth, steps, warmup):
    """Every --algorithm on one workload, with the reference's decomposition
    (plain: one workgroup per (group, chunk) unit) and with the batched engine
    (blockinggroups: plain only): the reference evaluation's breakdown of its
    optimisations (README.md:141-150) re-measured on this engine.  Every
    report set is compared with the first one."""
    from ngap_amd import Engine, NgapError
    groups, test, _, chunk = workload(name)
    total = (len(test) // chunk) * chunk
    eng = Engine(0)
    eng.load_groups(groups, max(depth, 1))
    eng.set_input(test, chunk)
    rows, ref = [], None
    for algo in ALGORITHMS:
        for sched in ("plain", "batched"):
            if algo == "blockinggroups" and sched == "batched":
                continue
            opts = eng.options(algorithm=algo, schedule=sched, precompute_depth=depth,
                               unique=True, use_graph=USE_GRAPH and sched == "batched",
                               fallback=True)
            try:
                wall, kern, st = timed(eng, opts, steps, warmup)
            except NgapError as e:  # e.g. a frontier beyond the blocking engine's capacity
                rows.append(dict(algorithm=algo, schedule=sched, error=str(e)))
                log(f"breakdown {name}: {rows[-1]}")
                continue
            rep = eng.reports()
            if ref is None:
                ref = rep
            rows.append(dict(algorithm=algo, schedule=sched,
                             mbps=round(total / (wall * 1e-3) / 1e6, 1),
                             kernel_ms=round(kern, 4), items=int(st["items"]),
                             bit_exact=bool(np.array_equal(rep, ref))))
            log(f"breakdown {name}: {rows[-1]}")
    eng.close()
    return dict(workload=name, input_bytes=int(total), groups=len(groups),
                precompute_depth=depth, steps=steps, rows=rows)


def serving_bench(rates, n_requests=1024, depth=3):
    """Serving workload: 64 resident automata (1k..200k states), Poisson
    requests of 1 KiB..4 MiB Zipf(1.1) bytes; p50/p99 latency and aggregate
    MB/s per arrival rate; every request's reports checked against the same
    request processed alone."""
    from ngap_amd import Engine, synth
    from ngap_amd.serve import Server, run_poisson
    t0 = time.time()
    groups = synth.serving_groups()
    t1 = time.time()
    eng = Engine(0)
    eng.load_groups(groups, depth)
    log(f"serving: {sum(g.num_states for g in groups)} states generated in {t1 - t0:.1f} s, "
        f"loaded at depth {depth} in {time.time() - t1:.1f} s")
    srv = Server(eng, budget_bytes=200 << 30, max_inflight=4, algorithm="nonblockingallgroups",
                 precompute_depth=depth, unique=True)
    base = synth.serving_requests(seed=2004, n=n_requests)
    alone = []
    for _, g, d in base:
        rid = srv.submit(g, d)
        alone.append(srv.drain([rid])[rid][0])
    # one unmeasured pass with everything queued, so every in-flight slot has
    # grown its device buffers before the sweep (steady-state serving)
    run_poisson(srv, [(0.0, g, d) for _, g, d in base])
    out = {"groups": len(groups), "requests": n_requests, "precompute_depth": depth,
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
    out["server"] = srv.stats()
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
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed repetitions of --steps steps per engine (median reported)")
    ap.add_argument("--requests", type=int, default=1024, help="serving requests per sweep")
    ap.add_argument("--also", default="batch1024,regex20k,literal64",
                    help="comma-separated further workloads compared the same way, or 'none'")
    ap.add_argument("--serving", default="100,1000,10000,inf",
                    help="arrival rates (per second, inf = all at once) for the serving run; "
                         "'none' skips it")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving sweep (prints its JSON object)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the oracle work counts")
    ap.add_argument("--engine", default=None,
                    choices=["plain", "static", "pergroup", "tuned", "learned", "learned_pergroup"],
                    help="time this engine alone on --workload (no other engine, workload or "
                         "end-to-end leg) and print its timing: the command to profile")
    ap.add_argument("--breakdown", default=None, metavar="WORKLOAD",
                    choices=["learned", "batch1024", "regex20k", "literal64"],
                    help="only time every --algorithm on WORKLOAD, plain and batched "
                         "(prints its JSON object)")
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

    if a.breakdown:
        print(json.dumps({"breakdown": breakdown(a.breakdown, a.depth, a.steps, a.warmup)}),
              flush=True)
        return
    if a.serving_only:
        print(json.dumps({"serving": serving_bench([float(r) for r in a.serving.split(",")],
                                                   n_requests=a.requests)}), flush=True)
        return
    if a.engine:
        res = compare(a.workload, a.depth, a.steps, a.warmup,
                      learned_on=a.engine.startswith("learned"), tune=a.engine == "tuned",
                      repeats=1, check=False, cpu=False, only=a.engine)[0]
        print(json.dumps({"profile": res}), flush=True)
        return
    res, runs, groups, data, chunk = compare(a.workload, a.depth, a.steps, a.warmup,
                                             learned_on=not a.no_learned, tune=not a.no_tune,
                                             repeats=a.repeats, check=not a.no_check,
                                             cpu=not a.no_cpu_baseline)
    main_s = "learned" if "learned" in runs else "static"
    total = res["input_bytes"]
    wall = runs[main_s]["wall_ms"]
    kern = runs[main_s]["kernel_ms"]
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
        "end_to_end": res["end_to_end"],
        "engines": res,
        "roofline": res.get("roofline"),
    }
    if "cpu_baseline" in res:
        out["cpu_baseline"] = res["cpu_baseline"]
    del runs, groups, data
    if a.also != "none":
        out["also"] = [compare(w, a.depth, a.steps, a.warmup, learned_on=not a.no_learned,
                               tune=not a.no_tune, repeats=a.repeats, check=not a.no_check,
                               cpu=not a.no_cpu_baseline)[0]
                       for w in a.also.split(",") if w and w != a.workload]
    if a.serving != "none":
        out["serving"] = serving_bench([float(r) for r in a.serving.split(",")],
                                       n_requests=a.requests)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
