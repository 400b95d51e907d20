"""Summaries of rocprofv3 result databases (SQLite, rocpd schema).

    python -m blackhorizon_amd.profiling stats <results.db> [--top N]
    python -m blackhorizon_amd.profiling pmc <results.db> [--match trace]

``stats`` prints per-kernel dispatch counts and durations in ms (the
``--kernel-trace --stats`` view).  ``pmc`` prints, per dispatch of the
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


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("what", choices=["stats", "pmc"])
    ap.add_argument("db")
    ap.add_argument("--top", type=int, default=12)
    ap.add_argument("--match", default="")
    a = ap.parse_args(argv)
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
