"""Summaries of rocprofv3 result databases (rocpd SQLite) as markdown.

    python -m timstochrigidbodiessim_amd.rocpd_summary kernels RUN_results.db [--top N] [--min-us U]
    python -m timstochrigidbodiessim_amd.rocpd_summary traffic FETCH_results.db WRITE_results.db \
        --kernel NAME_SUBSTRING --copy-bytes BYTES [--algorithmic-bytes BYTES]

`kernels` prints per-kernel launch statistics from a --kernel-trace run.
`traffic` reads two --pmc runs (FETCH_SIZE, then WRITE_SIZE, in KiB per
dispatch) of the same program, which must also launch copy16_kernel (the
streaming copy of fcm_profile_copy) reading and writing --copy-bytes per
launch; each counter is scaled by copy-bytes / (what the counter reads for the
copy), and the HBM bytes per dispatch of the named kernel are printed, with
the machine-readable line `traffic_bytes_per_launch: N`."""
import argparse
import collections
import sqlite3
import sys


def _kernel_rows(db):
    c = sqlite3.connect(db)
    return c.execute("select name, start, end from kernels").fetchall()


def kernels(db, top, min_us):
    d = collections.defaultdict(list)
    for n, s, e in _kernel_rows(db):
        d[n.split("(")[0]].append((e - s) / 1e3)
    tot = sum(sum(v) for v in d.values())
    print("| kernel | calls | total ms | avg us | min us | max us | share | working calls | working avg us |")
    print("|---|---|---|---|---|---|---|---|---|")
    for n, v in sorted(d.items(), key=lambda kv: -sum(kv[1]))[:top]:
        w = [x for x in v if x >= min_us]
        wavg = f"{sum(w) / len(w):.1f}" if w else "-"
        print(f"| {n} | {len(v)} | {sum(v) / 1e3:.2f} | {sum(v) / len(v):.1f} | {min(v):.1f} | {max(v):.1f} | "
              f"{100 * sum(v) / tot:.1f}% | {len(w)} | {wavg} |")
    print(f"\nTotal kernel time: {tot / 1e3:.1f} ms over {sum(len(v) for v in d.values())} dispatches.  "
          f"'working' = launches of at least {min_us:g} us (the others returned at once: a gated GMRES step "
          f"queued after convergence).")


def _counter_means(db, counter):
    c = sqlite3.connect(db)
    rows = c.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection").fetchall()
    per = collections.defaultdict(lambda: collections.defaultdict(float))  # kernel -> dispatch -> sum
    for kn, cn, v, did in rows:
        if cn == counter:
            per[kn.split("(")[0]][did] += v
    return {k: sum(v.values()) / len(v) for k, v in per.items()}


def _find(means, sub):
    hits = [k for k in means if sub in k]
    if len(hits) != 1:
        raise SystemExit(f"kernel substring {sub!r} matches {hits}")
    return hits[0]


def traffic(fetch_db, write_db, kernel, copy_bytes, alg):
    f = _counter_means(fetch_db, "FETCH_SIZE")
    w = _counter_means(write_db, "WRITE_SIZE")
    kf, kw = _find(f, kernel), _find(w, kernel)
    cf, cw = _find(f, "copy16_kernel"), _find(w, "copy16_kernel")
    sf = copy_bytes / (f[cf] * 1024.0)  # bytes per counted KiB, from the copy
    sw = copy_bytes / (w[cw] * 1024.0)
    fetch = f[kf] * 1024.0 * sf
    write = w[kw] * 1024.0 * sw
    print(f"| quantity | value |\n|---|---|")
    print(f"| copy16_kernel FETCH_SIZE (KiB/dispatch) | {f[cf]:.0f} (copy reads {copy_bytes / 1024:.0f} KiB: scale {sf:.4f}) |")
    print(f"| copy16_kernel WRITE_SIZE (KiB/dispatch) | {w[cw]:.0f} (copy writes {copy_bytes / 1024:.0f} KiB: scale {sw:.4f}) |")
    print(f"| {kf} FETCH_SIZE (KiB/dispatch) | {f[kf]:.0f} |")
    print(f"| {kw} WRITE_SIZE (KiB/dispatch) | {w[kw]:.0f} |")
    print(f"| HBM bytes per launch, scaled (fetch + write) | {fetch / 1e6:.1f} MB + {write / 1e6:.1f} MB = "
          f"{(fetch + write) / 1e6:.1f} MB |")
    if alg:
        print(f"| algorithmic bytes per launch | {alg / 1e6:.1f} MB (traffic / algorithmic = {(fetch + write) / alg:.3f}) |")
    print(f"\ntraffic_bytes_per_launch: {int(round(fetch + write))}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--top", type=int, default=40)
    k.add_argument("--min-us", type=float, default=10.0)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_db")
    t.add_argument("write_db")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--algorithmic-bytes", type=float, default=None)
    a = ap.parse_args(argv)
    if a.cmd == "kernels":
        kernels(a.db, a.top, a.min_us)
    else:
        traffic(a.fetch_db, a.write_db, a.kernel, a.copy_bytes, a.algorithmic_bytes)


if __name__ == "__main__":
    sys.exit(main())
