"""Text summaries of rocprofv3 result databases (the SQLite `*.db` a run writes
under its `-d` folder), for the kernel statistics kept with the benchmarks.

    python -m parode_amd.rocprof_summary trace DIR [--passes P] [--top K] [--dispatches K]
    python -m parode_amd.rocprof_summary pmc DIR [--top K]

trace: per kernel (and LDS size) over a `--kernel-trace` run: time per pass
(total / P), launches per pass, average duration, work-groups, work-group
size, LDS bytes, VGPRs, AGPRs and scratch bytes; optionally the last K
dispatches in launch order.

pmc: per kernel over a `--pmc` run that collected SQ_INSTS_VALU_MFMA_MOPS_F32,
SQ_INSTS_VALU_MFMA_MOPS_F64 and SQ_INSTS_MFMA: the matrix-core rate of each
kernel, MOPS x 512 flop / dispatch time (the MOPS counters count matrix
operations in units of 512 flop), and the fraction of MI355X's dense peak
(FP32 157.3, FP64 78.6 TFLOP/s).  For other counters: their per-dispatch
averages.
"""
import argparse
import collections
import glob
import sqlite3
import sys

PEAK_TF = {"F32": 157.3, "F64": 78.6}


def _db(folder):
    hits = sorted(glob.glob(folder + "/**/*.db", recursive=True))
    if not hits:
        sys.exit(f"no rocprofv3 database under {folder}")
    return sqlite3.connect(hits[0])


def _short(name):
    for pre in ("parode::k::g::(anonymous namespace)::", "parode::k::"):
        name = name.replace(pre, "")
    return name.split("(")[0]


def trace(args):
    c = _db(args.dir)
    rows = c.execute(
        "select name, count(*), sum(end-start)/1e6, avg(end-start)/1e3,"
        " max(grid_x)/max(workgroup_x), max(workgroup_x), max(lds_size), max(vgpr_count),"
        " max(accum_vgpr_count), max(scratch_size) from kernels group by name, lds_size"
        " order by 3 desc").fetchall()
    total = sum(r[2] for r in rows)
    p = float(args.passes)
    print(f"kernel time {total:.3f} ms in total, {total / p:.3f} ms per pass ({p:g} passes)")
    print(f"{'ms/pass':>8s} {'launches':>8s} {'avg us':>9s} {'wgs':>8s} {'wg':>5s} {'lds':>6s}"
          f" {'vgpr':>5s} {'agpr':>5s} {'scr':>4s} kernel")
    for r in rows[:args.top]:
        print(f"{r[2] / p:8.3f} {r[1] / p:8.1f} {r[3]:9.1f} {r[4]:8d} {r[5]:5d} {r[6]:6d}"
              f" {r[7]:5d} {r[8]:5d} {r[9]:4d} {_short(r[0])[:80]}")
    if args.dispatches:
        d = c.execute("select name, (end-start)/1e3, grid_x/workgroup_x, workgroup_x, lds_size"
                      " from kernels order by start desc limit ?",
                      (args.dispatches,)).fetchall()[::-1]
        print(f"\nlast {args.dispatches} dispatches: us, work-groups, wg size, lds, kernel")
        for nm, us, g, w, lds in d:
            print(f"{us:9.1f} {g:8d} {w:5d} {lds:6d} {_short(nm)[:80]}")


def pmc(args):
    c = _db(args.dir)
    rows = c.execute(
        "select dispatch_id, kernel_name, counter_name, sum(value), max(duration)"
        " from counters_collection group by dispatch_id, counter_name").fetchall()
    per = collections.defaultdict(dict)
    names, dur = {}, {}
    for disp, name, ctr, val, du in rows:
        per[disp][ctr] = val
        names[disp], dur[disp] = name, du
    agg = collections.defaultdict(collections.Counter)
    for disp, ctrs in per.items():
        a = agg[_short(names[disp])]
        a["n"] += 1
        a["dur"] += dur[disp]
        for k, v in ctrs.items():
            a[k] += v
    if "SQ_INSTS_VALU_MFMA_MOPS_F32" not in {k for a in agg.values() for k in a}:
        # other counters: their per-dispatch averages
        ctrs = sorted({k for a in agg.values() for k in a} - {"n", "dur"})
        print(f"{'kernel':50s} {'disp':>5s} {'ms':>8s} " + " ".join(f"{k[-20:]:>20s}" for k in ctrs))
        for nm, a in sorted(agg.items(), key=lambda kv: -kv[1]["dur"])[:args.top]:
            print(f"{nm[:50]:50s} {a['n']:5d} {a['dur'] / 1e6:8.3f} "
                  + " ".join(f"{a[k] / a['n']:20.4g}" for k in ctrs))
        return
    print(f"{'kernel':50s} {'disp':>5s} {'ms':>8s} {'F32 TF/s':>9s} {'of peak':>8s}"
          f" {'F64 TF/s':>9s} {'of peak':>8s} {'MFMA insts':>11s}")
    for nm, a in sorted(agg.items(), key=lambda kv: -kv[1]["dur"])[:args.top]:
        s = a["dur"] * 1e-9
        tf = {t: (a[f"SQ_INSTS_VALU_MFMA_MOPS_{t}"] * 512 / s / 1e12 if s else 0.0)
              for t in ("F32", "F64")}
        print(f"{nm[:50]:50s} {a['n']:5d} {a['dur'] / 1e6:8.3f} {tf['F32']:9.2f}"
              f" {tf['F32'] / PEAK_TF['F32']:8.3f} {tf['F64']:9.2f}"
              f" {tf['F64'] / PEAK_TF['F64']:8.3f} {a['SQ_INSTS_MFMA']:11.3g}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="mode", required=True)
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--passes", type=float, default=1.0)
    t.add_argument("--top", type=int, default=40)
    t.add_argument("--dispatches", type=int, default=0)
    q = sub.add_parser("pmc")
    q.add_argument("dir")
    q.add_argument("--top", type=int, default=40)
    args = ap.parse_args(argv)
    (trace if args.mode == "trace" else pmc)(args)


if __name__ == "__main__":
    main()
