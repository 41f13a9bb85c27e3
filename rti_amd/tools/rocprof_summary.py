"""Summarise rocprofv3 results (rocpd SQLite) into short text.

  python -m rti_amd.tools.rocprof_summary stats <results.db>
      per-kernel calls, total, average, min and max duration in ms
      (the --stats view, from the per-dispatch kernel records);
  python -m rti_amd.tools.rocprof_summary dispatches <results.db> <kernel-substring>
      duration of every dispatch of the matching kernels, in launch order;
  python -m rti_amd.tools.rocprof_summary pmc <results.db> <counter> [kernel-substring]
      per-dispatch counter values (summed over the counter's instances).
"""
import sqlite3
import sys


def stats(db):
    c = sqlite3.connect(db)
    rows = c.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                     "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = ["kernel | calls | total_ms | average_ms | min_ms | max_ms | percent"]
    for name, calls, tot, avg, lo, hi in rows:
        out.append(f"{name[:110]} | {calls} | {tot / 1e6:.3f} | {avg / 1e6:.3f} | {lo / 1e6:.3f} | "
                   f"{hi / 1e6:.3f} | {100.0 * tot / total:.2f}")
    return "\n".join(out)


def dispatches(db, kernel):
    c = sqlite3.connect(db)
    rows = c.execute("select name, duration from kernels where instr(name, ?) > 0 order by start",
                     (kernel,)).fetchall()
    return "\n".join(f"{i} | {name[:110]} | {d / 1e6:.3f} ms" for i, (name, d) in enumerate(rows))


def pmc(db, counter, kernel=None):
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, dispatch_id, counter_name, sum(value), max(duration) from counters_collection "
                     "where counter_name = ? group by dispatch_id order by dispatch_id", (counter,)).fetchall()
    out = []
    for name, disp, cname, val, dur in rows:
        if kernel and kernel not in name:
            continue
        out.append(f"{name[:110]} | dispatch {disp} | {cname} | {val}")
    return "\n".join(out)


if __name__ == "__main__":
    if sys.argv[1] == "stats":
        print(stats(sys.argv[2]))
    elif sys.argv[1] == "dispatches":
        print(dispatches(sys.argv[2], sys.argv[3]))
    else:
        print(pmc(sys.argv[2], sys.argv[3], sys.argv[4] if len(sys.argv) > 4 else None))
