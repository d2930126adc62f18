"""Summaries of rocprofv3 SQLite databases (text for the profiles/ directory).

    python -m range_libc_amd.rocpd_summary DB            per-kernel durations
    python -m range_libc_amd.rocpd_summary --pmc DB ...  per-kernel counter means
"""
import sqlite3
import sys


def _short(name, n=70):
    return name if len(name) <= n else name[: n - 3] + "..."


def durations(path):
    c = sqlite3.connect(path)
    rows = c.execute(
        "select name, count(*), avg(end-start), min(end-start), max(end-start), sum(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    print(f"{'kernel':70s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s}")
    for name, n, avg, mn, mx, _ in rows:
        print(f"{_short(name):70s} {n:6d} {avg / 1e3:9.2f} {mn / 1e3:9.2f} {mx / 1e3:9.2f}")


def pmc(paths, match=None):
    """Mean per dispatch of every collected counter, per kernel."""
    acc = {}
    for path in paths:
        c = sqlite3.connect(path)
        for kname, counter, value, dur in c.execute(
                "select kernel_name, counter_name, value, duration from counters_collection"):
            if match and match not in kname:
                continue
            d = acc.setdefault(kname, {})
            s = d.setdefault(counter, [0.0, 0])
            s[0] += value
            s[1] += 1
            t = d.setdefault("_duration_ns", [0.0, 0])
            if counter == next(iter(d)) or t[1] < s[1]:
                t[0] += dur
                t[1] += 1
    for kname, d in acc.items():
        print(_short(kname, 100))
        for counter, (tot, n) in sorted(d.items()):
            print(f"    {counter:32s} {tot / n:16.1f}   (n={n})")
    return acc


if __name__ == "__main__":
    if sys.argv[1] == "--pmc":
        args = sys.argv[2:]
        match = None
        if args and args[0].startswith("--match="):
            match = args[0].split("=", 1)[1]
            args = args[1:]
        pmc(args, match)
    else:
        durations(sys.argv[1])
