"""Summaries of a rocprofv3 trace database (the default rocpd output of
``rocprofv3 --kernel-trace``): per-kernel launch statistics, and the busy time against the
span of the last dispatches (the idle time between kernels, e.g. graph-replay launch gaps).

    python -m computer_graphics_lab_00_amd.rocpd_summary RESULTS.db [--last N --steps K]
        [--range FIRST COUNT | --tail COUNT]

--range restricts the statistics to launches FIRST .. FIRST + COUNT - 1 of each kernel (in
time order), e.g. the window of steps a benchmark timed; --tail to the last COUNT launches of
each kernel.
"""
from __future__ import annotations

import argparse
import re
import sqlite3


def _short(name: str) -> str:
    """The kernel's name with its template arguments (demangled or mangled symbol)."""
    m = re.search(r"(k_[a-z_0-9]+)(<[^>()]*>)?", name)
    if not m:
        return name[:60]
    key = m.group(1) + (m.group(2).replace(" ", "") if m.group(2) else "")
    t = re.search(r"k_[a-z_0-9]+I((?:Lb\dE)+)E", name)   # mangled bool arguments: I Lb1E Lb0E ... E
    if t and not m.group(2):
        key += "<" + ",".join(x for x in re.findall(r"Lb(\d)E", t.group(1))) + ">"
    return key


def kernel_stats(db: str, first: int = 0, count: int | None = None,
                 tail: int | None = None) -> list[tuple[str, int, float, float, float, float]]:
    """(kernel, launches, avg us, min us, max us, total ms), largest total first; with `count`,
    over launches first .. first + count - 1 of each kernel in time order; with `tail`, over
    the last `tail` launches of each kernel."""
    rows = sorted(sqlite3.connect(db).execute("select start, end, name from kernels").fetchall())
    agg: dict[str, list[float]] = {}
    for s, e, name in rows:
        agg.setdefault(_short(name), []).append((e - s) / 1000.0)
    if count is not None:
        agg = {k: v[first:first + count] for k, v in agg.items() if v[first:first + count]}
    if tail is not None:
        agg = {k: v[-tail:] for k, v in agg.items()} if tail > 0 else {}
    out = [(k, len(v), sum(v) / len(v), min(v), max(v), sum(v) / 1000.0) for k, v in agg.items()]
    return sorted(out, key=lambda r: -r[5])


def busy_vs_span(db: str, last: int) -> dict:
    """Busy time, span and gaps of the last `last` dispatches of the library's kernels."""
    rows = sorted(sqlite3.connect(db).execute("select start, end, name from kernels").fetchall())
    rows = [r for r in rows if "k_" in r[2]][-last:]
    busy = sum(e - s for s, e, _ in rows)
    span = rows[-1][1] - rows[0][0]
    gaps = [rows[i + 1][0] - rows[i][1] for i in range(len(rows) - 1)]
    return {"dispatches": len(rows), "busy_ns": busy, "span_ns": span, "idle_fraction": (span - busy) / span,
            "mean_gap_ns": sum(gaps) / max(1, len(gaps)), "max_gap_ns": max(gaps) if gaps else 0}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m computer_graphics_lab_00_amd.rocpd_summary")
    ap.add_argument("db")
    ap.add_argument("--last", type=int, default=0, help="dispatches at the end of the trace to check for gaps")
    ap.add_argument("--steps", type=int, default=1, help="steps those dispatches cover")
    window = ap.add_mutually_exclusive_group()
    window.add_argument("--range", type=int, nargs=2, metavar=("FIRST", "COUNT"), default=None,
                        help="statistics over launches FIRST .. FIRST + COUNT - 1 of each kernel")
    window.add_argument("--tail", type=int, default=None, metavar="COUNT",
                        help="statistics over the last COUNT launches of each kernel")
    a = ap.parse_args(argv)
    stats = kernel_stats(a.db, *(a.range or (0, None)), tail=a.tail)
    total = sum(r[5] for r in stats)
    print(f"{'kernel':40s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'total_ms':>9s} {'pct':>6s}")
    for k, n, avg, mn, mx, tot in stats:
        print(f"{k:40s} {n:6d} {avg:9.2f} {mn:9.2f} {mx:9.2f} {tot:9.3f} {100 * tot / total:6.1f}")
    if a.last:
        g = busy_vs_span(a.db, a.last)
        print(f"last {g['dispatches']} dispatches ({g['dispatches'] / a.steps:.0f} per step): busy "
              f"{g['busy_ns'] / 1e6 / a.steps:.3f} ms/step, span {g['span_ns'] / 1e6 / a.steps:.3f} ms/step, idle "
              f"{100 * g['idle_fraction']:.2f} %, gap between kernels mean {g['mean_gap_ns'] / 1e3:.2f} us, "
              f"max {g['max_gap_ns'] / 1e3:.1f} us")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
