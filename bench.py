"""Benchmark: one MI355X, the parallel-in-time LTI solve, state-steps/s.

Default workload (the configuration the metric is quoted on): Adams-Bashforth-
Moulton 4 LTI system, n=64 states (companion scan size 256), m=8 inputs, p=8
outputs, h=5e-4, N=2^20 output points, inputs sin(w t + phi) + N(0, 0.1^2)
noise sampled in FP64 on the h/2 grid, mixed precision (FP64 step map and
powers, FP64 top scan levels, FP32 lower levels).  A step = one whole solve of
that system (step map, powers, start-up, leaves, both sweeps, outputs) with
matrices and samples already resident in HBM.

Prints one JSON line.  Everything after the timed region (accuracy against the
FP64 oracle, the CPU baseline) is reported beside the throughput.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3e12, "fp64": 78.6e12, "hbm": 8.0e12}


def level_flops(na, n_scan):
    """Algorithmic flops of the level sweeps: 2 na^2 per pair, pairs of
    level l = ceil(N / 2^(l+1)), both sweeps."""
    L = max(0, (n_scan - 1).bit_length())
    pairs = sum((n_scan + (2 << l) - 1) // (2 << l) for l in range(L))
    return 2 * 2.0 * na * na * pairs, pairs


def cpu_baseline_leg(s, U64, y_gpu, target_s=12.0, full_reference=True):
    """The reference's CPU path (ABM4 PECE / classic RK4 with float state,
    TestsSources/TestParODEStaticLibCInterface.cpp:158-202, :336-461) timed on
    one host core over a bounded prefix, and the FP64 oracle as the accuracy
    check of the GPU result."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    A32, B32, C32, D32, x032 = s.A, s.B, s.C, s.D, s.x0
    U32 = U64.astype(np.float32)
    cpu = orc.cpu_abm_pece if s.method == 1 else orc.cpu_rk4_odeint
    probe = min(s.steps, 1 << 13)
    t0 = time.perf_counter()
    cpu(A32, B32, C32, D32, x032, U32[:2 * probe + 1], probe, s.h)
    dt = time.perf_counter() - t0
    sample = int(min(s.steps, max(probe, probe * target_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    cpu(A32, B32, C32, D32, x032, U32[:2 * sample + 1], sample, s.h)
    dt = time.perf_counter() - t0
    out = {"value": sample / dt, "unit": "state-steps/s", "cores": 1, "kind": "port",
           "sample": f"first {sample} of {s.steps} points of the same system, "
                     f"{'ABM4 PECE' if s.method == 1 else 'classic RK4'} with float state, "
                     f"1 thread, {dt:.1f} s"}
    acc = {}
    if full_reference:
        t0 = time.perf_counter()
        y64 = orc.simulate_f64(s.method, A32.astype(np.float64), B32.astype(np.float64),
                               C32.astype(np.float64), D32.astype(np.float64),
                               x032.astype(np.float64), U64, s.steps, s.h)
        acc = {"max_rel_err_vs_fp64": orc.max_rel_err(y_gpu, y64),
               "fp64_oracle_s": round(time.perf_counter() - t0, 2)}
    return out, acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--points", type=int, default=0, help="override N (output points)")
    ap.add_argument("--precision", default="mixed", choices=["plain", "mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)

    import torch
    from parode_amd import api, synth

    prec = {"plain": api.PLAIN32, "mixed": api.MIXED, "fp64": api.FP64}[args.precision]
    s = synth.config(args.config, steps=args.points or None)
    U64 = synth.samples_f64(s)
    samples = U64 if prec != api.PLAIN32 else U64.astype(np.float32)
    dev = torch.device("cuda:0")
    tA, tB, tC, tD, tx = (torch.from_numpy(np.ascontiguousarray(v)).to(dev)
                          for v in (s.A, s.B, s.C, s.D, s.x0))
    tU = torch.from_numpy(samples).to(dev)
    tY = torch.zeros((s.steps, s.p), dtype=torch.float32, device=dev)
    plan = api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, prec, samples.dtype == np.float64)
    stream = torch.cuda.Stream(device=dev)
    sp = stream.cuda_stream
    ptrs = (tA.data_ptr(), tB.data_ptr(), tC.data_ptr(), tD.data_ptr(), tx.data_ptr(),
            tU.data_ptr(), tY.data_ptr())
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        plan.execute(*ptrs, sp)
    torch.cuda.synchronize()

    # timed region: K solves back to back on one stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        plan.execute(*ptrs, sp)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    ms = ms_total / args.steps

    # per-phase GPU time (hipEvents on the same stream) and per-solve latency
    phases, lat = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        phases.append(plan.execute_timed(*ptrs, sp))
        lat.append((time.perf_counter() - t0) * 1e3)
    y = tY.cpu().numpy()
    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}

    na = s.n_scan
    n_scan = s.steps if s.method == 0 else s.steps - 3
    flops, pairs = level_flops(na, n_scan)
    sweep_ms = ph["up_sweep"] + ph["down_sweep"]
    achieved = flops / (sweep_ms * 1e-3)
    peak = PEAK["fp64"] if prec == api.FP64 else PEAK["fp32"]
    value = s.steps / (ms * 1e-3)

    line = {
        "metric": "state-steps/s at <=1e-6 rel err vs FP64, served vs plain scan; p50/p99 latency",
        "value": value,
        "unit": "state-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"plain": "fp32", "mixed": "fp32+fp64", "fp64": "fp64"}[args.precision],
        "data": "synthetic (seeded system and FP64 input samples, resident in HBM)",
        "config": {"workload": f"{'ABM4' if s.method else 'RK4'} LTI n={s.n} (scan {na}) m={s.m} "
                               f"p={s.p} h={s.h} N={s.steps} ({s.tag})",
                   "precision": args.precision, "points": s.steps, "scan_size": na},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "kernel": "level sweeps (k_level)",
                     "flops_per_solve": flops, "pairs_per_sweep": pairs, "sweep_ms": sweep_ms},
        "phases_ms": ph,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "plan_bytes": plan.footprint,
    }
    if not args.no_cpu or not args.no_check:
        cb, acc = cpu_baseline_leg(s, U64, y, full_reference=not args.no_check)
        if not args.no_cpu:
            line["cpu_baseline"] = cb
        line.update(acc)
    plan.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
