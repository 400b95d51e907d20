"""Summaries of rocprofv3 result databases (SQLite, rocpd schema).

    python -m blackhorizon_amd.profiling stats <results.db> [--top N]
    python -m blackhorizon_amd.profiling dispatches <results.db> [--match trace]
    python -m blackhorizon_amd.profiling pmc <results.db> [--match trace]
    python -m blackhorizon_amd.profiling traffic <fetch.db> --write-db <write.db> --match trace

``traffic`` turns FETCH_SIZE and WRITE_SIZE (KiB per dispatch, from two
separate ``--pmc`` runs of the same command) into HBM bytes per dispatch of
the matching kernels, each counter scaled by what it reads for the layout
copy ``aos_to_soa_kernel`` in the same run: that kernel reads and writes
exactly 64 bytes per thread of its grid (n a multiple of 256).

``stats`` prints per-kernel dispatch counts and durations in ms (the
``--kernel-trace --stats`` view); ``dispatches`` every dispatch of the
matching kernels in launch order, with its grid and duration.  ``pmc`` prints, per dispatch of the
matching kernels, the collected counters and, for FP64 VALU counters, the
counter-based FLOP count 64 x (ADD + MUL + TRANS + 2 FMA) (wave-level
instruction counts times the wavefront width: inactive lanes included).
"""

from __future__ import annotations

import argparse
import collections
import sqlite3
import sys


def kernel_stats(db: str, top: int = 12) -> list[tuple]:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage "
                     "from top_kernels order by total_duration desc").fetchall()
    return rows[:top]


def dispatches(db: str, match: str = "") -> list[tuple]:
    """(dispatch id, kernel, grid, workgroup, duration ms) per dispatch of
    the matching kernels, in launch order."""
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, name, grid_x, workgroup_x, duration from kernels "
                     "order by start").fetchall()
    return [(d, n, g, w, dur * 1e-6) for d, n, g, w, dur in rows if match in n]


def pmc_by_dispatch(db: str, match: str = "") -> list[dict]:
    c = sqlite3.connect(db)
    rows = c.execute("select dispatch_id, kernel_name, counter_name, value, grid_size, "
                     "workgroup_size, vgpr_count, accum_vgpr_count, sgpr_count, "
                     "lds_block_size, scratch_size from counters_collection").fetchall()
    out: dict = collections.OrderedDict()
    for disp, name, cn, val, grid, wg, vg, ag, sg, lds, scr in rows:
        if match and match not in name:
            continue
        d = out.setdefault(disp, dict(dispatch=disp, kernel=name, grid=grid, workgroup=wg,
                                      vgpr=vg, agpr=ag, sgpr=sg, lds=lds, scratch=scr,
                                      counters={}))
        d["counters"][cn] = d["counters"].get(cn, 0.0) + float(val)
    for d in out.values():
        k = d["counters"]
        if "SQ_INSTS_VALU_FMA_F64" in k:
            d["fp64_flop_counter"] = 64.0 * (k.get("SQ_INSTS_VALU_ADD_F64", 0.0)
                                             + k.get("SQ_INSTS_VALU_MUL_F64", 0.0)
                                             + k.get("SQ_INSTS_VALU_TRANS_F64", 0.0)
                                             + 2.0 * k["SQ_INSTS_VALU_FMA_F64"])
    return list(out.values())


CALIB_KERNEL = "aos_to_soa_kernel"
CALIB_BYTES_PER_THREAD = 64


def _calibration(rows: list[dict], counter: str) -> float:
    """Known bytes / counted bytes of the layout copies in one run."""
    ratios = [d["grid"] * CALIB_BYTES_PER_THREAD / (d["counters"][counter] * 1024.0)
              for d in rows if CALIB_KERNEL in d["kernel"] and d["counters"].get(counter)]
    if not ratios:
        raise SystemExit(f"no {CALIB_KERNEL} dispatch with {counter} in this run")
    return sum(ratios) / len(ratios)


def traffic(fetch_db: str, write_db: str, match: str) -> dict:
    """Calibrated HBM read/write bytes per dispatch of the matching kernels
    (paired by order of appearance in the two runs)."""
    f_rows = pmc_by_dispatch(fetch_db)
    w_rows = pmc_by_dispatch(write_db)
    kf = _calibration(f_rows, "FETCH_SIZE")
    kw = _calibration(w_rows, "WRITE_SIZE")
    fm = [d for d in f_rows if match in d["kernel"]]
    wm = [d for d in w_rows if match in d["kernel"]]
    if len(fm) != len(wm) or not fm:
        raise SystemExit(f"{len(fm)} FETCH vs {len(wm)} WRITE dispatches match {match!r}")
    per = []
    for a, b in zip(fm, wm):
        rd = a["counters"]["FETCH_SIZE"] * 1024.0 * kf
        wr = b["counters"]["WRITE_SIZE"] * 1024.0 * kw
        per.append(dict(kernel=a["kernel"], grid=a["grid"], read_bytes=rd, write_bytes=wr))
    mean = sum(p["read_bytes"] + p["write_bytes"] for p in per) / len(per)
    return dict(fetch_factor=kf, write_factor=kw, dispatches=per, mean_bytes_per_dispatch=mean)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("what", choices=["stats", "dispatches", "pmc", "traffic"])
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--match", default="")
    ap.add_argument("--write-db", help="traffic: the WRITE_SIZE run's database")
    a = ap.parse_args(argv)
    if a.what == "traffic":
        t = traffic(a.db, a.write_db, a.match)
        print(f"calibration ({CALIB_KERNEL}, {CALIB_BYTES_PER_THREAD} B/thread each way): "
              f"FETCH_SIZE x {t['fetch_factor']:.4f}, WRITE_SIZE x {t['write_factor']:.4f}")
        for d in t["dispatches"]:
            print(f"{d['kernel'][:70]} grid={d['grid']} read={d['read_bytes'] / 1e6:.3f} MB "
                  f"write={d['write_bytes'] / 1e6:.3f} MB")
        print(f"mean HBM bytes per dispatch: {t['mean_bytes_per_dispatch']:.6g}")
        return 0
    if a.what == "dispatches":
        print(f"{'dispatch':>8} {'kernel':<80} {'grid':>9} {'wg':>4} {'ms':>10}")
        for disp, name, grid, wg, ms in dispatches(a.db, a.match):
            short = name if len(name) <= 80 else name[:77] + "..."
            print(f"{disp:>8} {short:<80} {grid:>9} {wg:>4} {ms:>10.3f}")
        return 0
    if a.what == "stats":
        print(f"{'kernel':<90} {'calls':>6} {'total_ms':>10} {'avg_ms':>9} {'pct':>6}")
        for name, calls, tot, avg, pct in kernel_stats(a.db, a.top):
            short = name if len(name) <= 90 else name[:87] + "..."
            print(f"{short:<90} {calls:>6} {tot / 1e3:>10.3f} {avg / 1e3:>9.4f} {pct:>6.2f}")
        return 0
    for d in pmc_by_dispatch(a.db, a.match):
        name = d["kernel"] if len(d["kernel"]) <= 70 else d["kernel"][:67] + "..."
        counters = " ".join(f"{k}={v:.6g}" for k, v in sorted(d["counters"].items()))
        extra = f" fp64_flop_counter={d['fp64_flop_counter']:.6g}" if "fp64_flop_counter" in d else ""
        print(f"dispatch {d['dispatch']} {name} grid={d['grid']} wg={d['workgroup']} "
              f"vgpr={d['vgpr']} sgpr={d['sgpr']} lds={d['lds']} scratch={d['scratch']}: "
              f"{counters}{extra}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
