"""Markdown summaries of rocprofv3 result databases (the sqlite files that
`rocprofv3 --kernel-trace` / `--pmc` write with `-o NAME`).

    python -m ngap_amd.rocprof_summary kernels RUN_results.db [--last KERNEL N]
    python -m ngap_amd.rocprof_summary pmc RUN_results.db [--last KERNEL N]

`kernels` prints one row per kernel name (dispatches, average / min / max /
total duration in ns); `pmc` one row per (kernel, counter) with the average
counter value per dispatch, summed over the rows a dispatch has (one per XCC
or shader engine for some counters; FETCH_SIZE and WRITE_SIZE are in KiB).  With
--last, the last N dispatches of the kernel whose name contains KERNEL (in
dispatch order) are summarised on their own: the timed steps of a benchmark
that launches the same kernel for other purposes first.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["kernels", "pmc"])
    ap.add_argument("db")
    ap.add_argument("--last", nargs=2, metavar=("KERNEL", "N"))
    a = ap.parse_args()
    if a.mode == "kernels":
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
