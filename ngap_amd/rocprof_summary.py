"""Markdown summaries of rocprofv3 result databases (the sqlite files that
`rocprofv3 --kernel-trace` / `--pmc` write with `-o NAME`).

    python -m ngap_amd.rocprof_summary kernels RUN_results.db [--last KERNEL N]
    python -m ngap_amd.rocprof_summary pmc RUN_results.db [--last KERNEL N]
    python -m ngap_amd.rocprof_summary timeline RUN_results.db [--from KERNEL]

`kernels` prints one row per kernel name (dispatches, average / min / max /
total duration in ns); `pmc` one row per (kernel, counter) with the average
counter value per dispatch, summed over the rows a dispatch has (one per XCC
or shader engine for some counters; FETCH_SIZE and WRITE_SIZE are in KiB).  With
--last, the last N dispatches of the kernel whose name contains KERNEL (in
dispatch order) are summarised on their own: the timed steps of a benchmark
that launches the same kernel for other purposes first.  `timeline` (a run
with --kernel-trace --memory-copy-trace) gives, from the first dispatch of the
kernel named by --from (or over the last --tail-ms milliseconds) to the end
of the run, the span, the time at least one
kernel ran (the union of their intervals), the time a copy ran, and the copies
by direction (count, bytes, busy time).
"""
import argparse
import sqlite3


def _rows(db, sql):
    con = sqlite3.connect(db)
    try:
        return list(con.execute(sql))
    finally:
        con.close()


def kernels(db):
    return _rows(db, "select name, count(*), avg(end - start), min(end - start), "
                     "max(end - start), sum(end - start) from kernels group by name "
                     "order by sum(end - start) desc")


def kernel_durations(db, pattern):
    return [r[0] for r in _rows(db, "select end - start from kernels where name like "
                                    f"'%{pattern}%' order by start")]


# A counter may come in several rows per dispatch (one per XCC / shader
# engine); rows are summed per dispatch first.
_PER_DISPATCH = ("select name, dispatch_id, counter_name, sum(counter_value) as v, "
                 "avg(duration) as d, min(start) as s from pmc_events "
                 "group by dispatch_id, counter_name")


def pmc(db):
    return _rows(db, f"select name, counter_name, count(*), avg(v), avg(d) from ({_PER_DISPATCH}) "
                     "group by name, counter_name")


def pmc_values(db, pattern):
    return _rows(db, f"select counter_name, v, d from ({_PER_DISPATCH}) where name "
                     f"like '%{pattern}%' order by s, dispatch_id")


def _union(iv):
    """Total length of the union of [start, end) intervals."""
    tot, cur_s, cur_e = 0, None, None
    for s, e in sorted(iv):
        if cur_e is None or s > cur_e:
            if cur_e is not None:
                tot += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    if cur_e is not None:
        tot += cur_e - cur_s
    return tot


def timeline(db, first=None, tail_ns=None):
    ks = _rows(db, "select name, start, end from kernels order by start")
    try:
        cs = _rows(db, "select name, start, end, size from memory_copies order by start")
    except sqlite3.OperationalError:
        cs = []
    t0 = min([k[1] for k in ks] + [c[1] for c in cs]) if ks or cs else 0
    if first:
        hit = [k[1] for k in ks if first in k[0]]
        t0 = hit[0] if hit else t0
    if tail_ns:
        t0 = max([k[2] for k in ks] + [c[2] for c in cs]) - tail_ns
    ks = [k for k in ks if k[1] >= t0]
    cs = [c for c in cs if c[1] >= t0]
    t1 = max([k[2] for k in ks] + [c[2] for c in cs]) if ks or cs else t0
    by = {}
    for name, s, e, size in cs:
        d = by.setdefault(name, [0, 0, []])
        d[0] += 1
        d[1] += size or 0
        d[2].append((s, e))
    return dict(span=t1 - t0, kernel_busy=_union([(k[1], k[2]) for k in ks]),
                copy_busy=_union([(c[1], c[2]) for c in cs]),
                any_busy=_union([(k[1], k[2]) for k in ks] + [(c[1], c[2]) for c in cs]),
                dispatches=len(ks),
                copies={n: (v[0], v[1], _union(v[2])) for n, v in by.items()})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["kernels", "pmc", "timeline"])
    ap.add_argument("db")
    ap.add_argument("--last", nargs=2, metavar=("KERNEL", "N"))
    ap.add_argument("--from", dest="first", metavar="KERNEL")
    ap.add_argument("--tail-ms", type=float, default=None)
    a = ap.parse_args()
    if a.mode == "timeline":
        t = timeline(a.db, a.first, int(a.tail_ms * 1e6) if a.tail_ms else None)
        span = max(1, t["span"])
        where = (f"the last {a.tail_ms} ms of the run" if a.tail_ms else
                 f"the first `{a.first or 'dispatch'}` to the end of the run")
        print(f"span {t['span'] / 1e6:.3f} ms, {where}; {t['dispatches']} dispatches")
        print(f"kernels running {t['kernel_busy'] / 1e6:.3f} ms ({100 * t['kernel_busy'] / span:.1f} %), "
              f"copies running {t['copy_busy'] / 1e6:.3f} ms ({100 * t['copy_busy'] / span:.1f} %), "
              f"either {t['any_busy'] / 1e6:.3f} ms ({100 * t['any_busy'] / span:.1f} %)\n")
        print("| copy | count | bytes | busy ms | GB/s while busy |")
        print("|---|---|---|---|---|")
        for n, (cnt, b, busy) in sorted(t["copies"].items()):
            print(f"| {n} | {cnt} | {b} | {busy / 1e6:.3f} | {b / max(1, busy):.2f} |")
    elif a.mode == "kernels":
        print("| kernel | dispatches | avg ns | min ns | max ns | total ns |")
        print("|---|---|---|---|---|---|")
        for r in kernels(a.db):
            print(f"| `{r[0][:80]}` | {r[1]} | {r[2]:.0f} | {r[3]} | {r[4]} | {r[5]} |")
        if a.last:
            d = kernel_durations(a.db, a.last[0])[-int(a.last[1]):]
            if d:
                print(f"\nlast {len(d)} dispatches of `{a.last[0]}`: avg {sum(d) / len(d):.0f} ns "
                      f"({', '.join(str(x) for x in d)})")
    else:
        print("| kernel | counter | dispatches | avg value | avg ns |")
        print("|---|---|---|---|---|")
        for r in pmc(a.db):
            print(f"| `{r[0][:80]}` | {r[1]} | {r[2]} | {r[3]:.3f} | {r[4]:.0f} |")
        if a.last:
            rows = pmc_values(a.db, a.last[0])
            by = {}
            for name, val, dur in rows:
                by.setdefault(name, []).append((val, dur))
            for name, vals in by.items():
                v = vals[-int(a.last[1]):]
                print(f"\nlast {len(v)} dispatches of `{a.last[0]}`, {name}: avg "
                      f"{sum(x for x, _ in v) / len(v):.3f} (durations avg "
                      f"{sum(y for _, y in v) / len(v):.0f} ns)")


if __name__ == "__main__":
    main()
