"""Print a per-kernel duration table from a rocprofv3 SQLite database.

    python -m range_libc_amd.rocpd_summary gpurun_out/<dir>/run_results.db
"""
import sqlite3
import sys


def main(path):
    c = sqlite3.connect(path)
    rows = c.execute(
        "select name, count(*), avg(end-start), min(end-start), max(end-start), sum(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    print(f"{'kernel':70s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s}")
    for name, n, avg, mn, mx, _ in rows:
        short = name if len(name) <= 70 else name[:67] + "..."
        print(f"{short:70s} {n:6d} {avg / 1e3:9.2f} {mn / 1e3:9.2f} {mx / 1e3:9.2f}")


if __name__ == "__main__":
    main(sys.argv[1])
