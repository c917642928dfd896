"""Learned schedule for the batched engine.

Traces: the batched engine run over a fixed tiling of 4 KiB windows reports,
per (group, window) work item, the number of (state, offset) items it
expanded.  Features: the window's normalised 256-bin byte histogram and 8
statistics of the group (log states, log edges, log all-input states, log
start-of-data states, log reporting states, mean out-degree, fraction of
states whose class has >= 128 bytes, log self-loops).  Target: log(1 + items).

The trained predictor (ngap_amd.mlp, HIP kernels) is used only to shape the
schedule: consecutive windows of a (group, chunk) unit are merged into one
work item until its predicted cost reaches a target, work items run
longest-predicted-first, and the per-wave worklist spill capacity is sized
from the predictions.  Matching is untouched; a work item whose worklist
overflows is re-run by the engine with a larger worklist (fallback), and the
report set is the same under every schedule.
"""
import numpy as np

WINDOW = 4096
SCAN_COST = 0.02  # cost of scanning one position, in units of one expanded item


def windows_per_chunk(chunk, window=WINDOW):
    return (chunk + window - 1) // window


def window_hist(data, chunk, window=WINDOW):
    """[num_windows, 256] normalised byte histograms; window index
    w = c * windows_per_chunk + k covers chunk c, bytes [k*window, (k+1)*window)."""
    data = np.asarray(data, np.uint8)
    S = len(data) // chunk
    K = windows_per_chunk(chunk, window)
    out = np.zeros((S * K, 256), np.float32)
    for c in range(S):
        base = c * chunk
        for k in range(K):
            a, b = base + k * window, base + min((k + 1) * window, chunk)
            h = np.bincount(data[a:b], minlength=256).astype(np.float32)
            out[c * K + k] = h / max(1, b - a)
    return out


def group_stats(groups):
    """[G, 8] float32 automaton statistics."""
    out = np.zeros((len(groups), 8), np.float32)
    for i, g in enumerate(groups):
        n = max(1, g.num_states)
        deg = np.diff(g.row)
        src = np.repeat(np.arange(g.num_states), deg)
        selfloops = int(np.sum(src == g.col))
        width = np.unpackbits(np.ascontiguousarray(g.sets).view(np.uint8)
                              .reshape(g.num_states, 32), axis=1).sum(1)
        out[i] = [np.log1p(g.num_states), np.log1p(g.num_edges), np.log1p(g.aas.size),
                  np.log1p(g.sod.size), np.log1p(int(np.sum(g.attrs & 2) // 2)),
                  g.num_edges / n, float(np.mean(width >= 128)) if g.num_states else 0.0,
                  np.log1p(selfloops)]
    return out


def fixed_tiling(num_groups, num_seg, chunk, window=WINDOW):
    """Every (group, chunk, window) as one work item: uint32 [n, 4]."""
    K = windows_per_chunk(chunk, window)
    c, g, k = np.meshgrid(np.arange(num_seg), np.arange(num_groups), np.arange(K), indexing="ij")
    begin = k * window
    end = np.minimum(begin + window, chunk)
    return np.stack([g.ravel(), c.ravel(), begin.ravel(), end.ravel()], 1).astype(np.uint32)


def pairs_of(tasks, chunk, window=WINDOW):
    """(group, global window index) of fixed-tiling work items."""
    K = windows_per_chunk(chunk, window)
    t = np.asarray(tasks, np.int64)
    return np.stack([t[:, 0], t[:, 1] * K + t[:, 2] // window], 1).astype(np.uint32)


def collect_traces(engine, data, chunk, num_groups, depth=2, window=WINDOW):
    """Runs the fixed tiling once (inputs and groups already loaded) and
    returns (pairs [n, 2], log1p(items) [n])."""
    S = len(data) // chunk
    tasks = fixed_tiling(num_groups, S, chunk, window)
    engine.run(algorithm="nonblockingallgroups", schedule="batched", precompute_depth=depth,
               unique=True, tasks=tasks, fallback=True, task_traces=True)
    items = engine.task_items().astype(np.float64)
    done = engine.tasks()
    return pairs_of(done, chunk, window), np.log1p(items).astype(np.float32)


def train_settings(n, steps=2000):
    """(epochs, batch) giving about `steps` Adam steps (at least 4 epochs)
    for n training samples."""
    batch = int(min(4096, max(256, n // 16)))
    per_epoch = max(1, n // batch)
    return int(min(300, max(4, -(-steps // per_epoch)))), batch


def features(hist, stats, pairs):
    pairs = np.asarray(pairs, np.int64)
    return np.concatenate([hist[pairs[:, 1]], stats[pairs[:, 0]]], 1).astype(np.float32)


ALL_GROUPS = 0xFFFFFFFF


def cross_schedule_from_costs(cost, num_seg, chunk, waves, window=WINDOW, per_wave=2,
                              min_piece=256, order="cost"):
    """All-group work items (every group's roots in one window, precompute
    depth >= 2) shaped by per-(group, window) costs: a window's cost is the
    sum over groups, then windows are cut or merged as in
    schedule_from_costs.  Returns work items and their costs, ordered by
    decreasing cost, or with order="position" by chunk and window: work items
    claimed together then cover neighbouring input, whose byte distribution
    (and so the index lines it touches) is alike."""
    K = windows_per_chunk(chunk, window)
    wcost = np.asarray(cost, np.float64).reshape(-1, num_seg * K).sum(0)
    target = max(wcost.sum() / max(1, waves * per_wave), 1e-9)
    tasks, tcost = [], []
    for c in range(num_seg):
        _tile(tasks, tcost, ALL_GROUPS, c, wcost[c * K:(c + 1) * K], chunk, window, target,
              min_piece)
    tasks, tcost = np.asarray(tasks, np.uint32), np.asarray(tcost)
    if order == "position":
        idx = np.lexsort((tasks[:, 2], tasks[:, 1]))
    else:
        idx = np.argsort(-tcost, kind="stable")
    return tasks[idx], tcost[idx]


def schedule_from_costs(cost, num_groups, num_seg, chunk, waves, window=WINDOW,
                        per_wave=2, min_piece=256):
    """Per-group work items shaped by per-(group, window) costs: inside each
    (group, chunk), a window above the target total / (waves * per_wave) is cut
    into equal pieces (at least min_piece bytes) and consecutive windows below
    it are merged.  Returns work items ordered by decreasing cost and their
    costs.  cost: [num_groups, num_seg * K] indexed by (group, global window)."""
    K = windows_per_chunk(chunk, window)
    cost = np.asarray(cost, np.float64).reshape(num_groups, num_seg * K)
    target = max(cost.sum() / max(1, waves * per_wave), 1e-9)
    tasks, tcost = [], []
    for g in range(num_groups):
        for c in range(num_seg):
            _tile(tasks, tcost, g, c, cost[g, c * K:(c + 1) * K], chunk, window, target,
                  min_piece)
    order = np.argsort(-np.asarray(tcost), kind="stable")
    return np.asarray(tasks, np.uint32)[order], np.asarray(tcost)[order]


def _tile(tasks, tcost, g, c, row, chunk, window, target, min_piece):
    """Work items of one (group or all groups, chunk) from its window costs."""
    a, acc = 0, 0.0
    K = len(row)
    for k in range(K):
        lo, hi = k * window, min((k + 1) * window, chunk)
        w = row[k]
        if w >= target:
            if acc > 0:
                tasks.append((g, c, a, lo))
                tcost.append(acc)
            n = int(min(np.ceil(w / target), max(1, (hi - lo) // min_piece)))
            cuts = [lo + (hi - lo) * i // n for i in range(n + 1)]
            for x, y in zip(cuts, cuts[1:]):
                tasks.append((g, c, x, y))
                tcost.append(w / n)
            a, acc = hi, 0.0
            continue
        acc += w
        if acc >= target or k == K - 1:
            tasks.append((g, c, a, hi))
            tcost.append(acc)
            a, acc = hi, 0.0


def predicted_cost(pred_log_items, window=WINDOW):
    return np.expm1(np.maximum(pred_log_items, 0.0)) + SCAN_COST * window


def spill_capacity(pred_log_items, tasks_cost, floor=2048):
    """Worklist spill items per wave from the predictions: a quarter of the
    largest predicted work item, rounded up to a power of two."""
    peak = float(np.max(tasks_cost)) if len(tasks_cost) else 0.0
    cap = max(floor, int(peak / 4))
    return int(min(1 << 20, 1 << int(np.ceil(np.log2(cap)))))
