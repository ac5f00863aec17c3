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

# HBM bytes of the largest launch of the scan at the default workload, the
# second chunk pass (k_abm_chunks_fac, 2^20 elements in chunks of 128), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE counters, FETCH scaled by 2.0 (what the
# counter reads for a plain copy of known size): (134471 * 2.0 + 262364) KiB.
TRAFFIC_CHUNK_PASS2 = (134471 * 2.0 + 262364) * 1024


def level_flops(na, n_scan, n_leaf=None, chunk_log=0, factored=False):
    """Algorithmic flops of the scan (both sweeps): 2 na^2 per pair, pairs of
    level l = ceil(N / 2^(l+1)).  With n_leaf (ABM, n = na / 4: the companion
    map M has rows below 3n that are shifts by n, M^2 rows below 2n that are
    shifts by 2n, and the leaves are nonzero only in their last n rows), the
    two lowest levels need fewer products per pair:
      level 0: up 2 n n (M's last n rows times the leaf's last n rows),
               down 2 n na (only the last n rows, which the outputs read)
      level 1: up 2 (2n)(2n), down 2 (2n) na.
    With chunk_log = lc (chunk recurrences), levels below lc are replaced by
    two passes per element (ABM: 2 n na, the n x 4n dense block of M per step,
    or 2 n (2n) with the FP64 map factored as x3 + A w1 + A^2 w2; RK4: 2 na^2),
    and only levels >= lc run as tree levels.
    Returns (flops, tree pairs)."""
    L = max(0, (n_scan - 1).bit_length())
    per = [(n_scan + (2 << l) - 1) // (2 << l) for l in range(L)]
    if chunk_log:
        # levels below chunk_log (ABM: the n x 4n block per step; RK4: the
        # dense map), two passes over every element
        # (an FP64 ABM map factored through A and A^2: 2 n x n products)
        per_step = (2.0 * n_leaf * (2 * n_leaf) if factored else 2.0 * n_leaf * na) \
            if n_leaf else 2.0 * na * na
        flops = 2 * per_step * n_scan
        tree = per[chunk_log:]
        return flops + 2 * 2.0 * na * na * sum(tree), sum(tree)
    pairs = sum(per)
    flops = 2 * 2.0 * na * na * pairs
    if n_leaf and L:
        n = n_leaf
        flops -= per[0] * (2 * 2.0 * na * na - 2.0 * n * n - 2.0 * n * na)
        if L > 1:
            flops -= per[1] * (2 * 2.0 * na * na - 2.0 * (2 * n) * (2 * n) - 2.0 * (2 * n) * na)
    return flops, pairs


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


def _torch_samples(torch, s, dev, dtype):
    """u(t0 + k h/2), k = 0 .. 2N, of the system's sinusoid family, on the device."""
    k = torch.arange(2 * s.steps + 1, dtype=torch.float64, device=dev)
    t = (s.t0 + k * (s.h / 2.0))[:, None]
    c = torch.from_numpy(np.ascontiguousarray(s.ucoef)).to(dev)
    g = c[:, 4][None, :] * t + c[:, 5][None, :]
    U = c[:, 0] * torch.sin(g) + c[:, 1] * torch.cos(g) + c[:, 2] + c[:, 3] * t
    return U.to(dtype).contiguous()


def batch_legs(args, torch, api, synth):
    """Batch of independent systems (512 by default: RK4/ABM alternating, n
    log-uniform on [4, 256], N log-uniform on [2^10, 2^18]).

    plain  : FP32 plain precision, one system per solve, the solves queued back
             to back on one stream, inputs and outputs resident in HBM.
    served : the continuous-batching server, mixed precision, all requests
             submitted at once; end to end (host matrices in, host outputs
             out) from the first submission to the last completion.
    stream : the server under Poisson arrivals from the same distribution;
             p50/p99 latency from submission to outputs on the host."""
    dev = torch.device("cuda:0")
    systems = synth.batch_config(args.batch)
    total_steps = float(sum(s.steps for s in systems))
    weighted = float(sum(s.steps * s.n for s in systems))
    out = {"systems": len(systems), "state_steps": total_steps}

    # ---- plain
    plans, bufs = [], []
    for s in systems:
        pl = api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, api.PLAIN32, False)
        tens = [torch.from_numpy(np.ascontiguousarray(v)).to(dev)
                for v in (s.A, s.B, s.C, s.D, s.x0)]
        tens.append(_torch_samples(torch, s, dev, torch.float32))
        tens.append(torch.empty((s.steps, s.p), dtype=torch.float32, device=dev))
        plans.append(pl)
        bufs.append(tens)
    stream = torch.cuda.Stream(device=dev)
    sp = stream.cuda_stream

    def run_plain():
        for pl, t in zip(plans, bufs):
            pl.execute(*(x.data_ptr() for x in t), sp)

    run_plain()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, args.batch_reps)
    ev0.record(stream)
    for _ in range(reps):
        run_plain()
    ev1.record(stream)
    torch.cuda.synchronize()
    plain_ms = ev0.elapsed_time(ev1) / reps
    y_plain = [t[-1].cpu().numpy() for t in bufs[:8]]
    for pl in plans:
        pl.close()
    del plans, bufs
    torch.cuda.empty_cache()
    out["plain"] = {"ms": plain_ms, "state_steps_per_s": total_steps / (plain_ms * 1e-3),
                    "n_weighted_per_s": weighted / (plain_ms * 1e-3), "precision": "plain fp32"}

    # ---- served (all at once)
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=args.lanes, device_bytes=int(need * 1.1) + (1 << 30),
                     pinned_bytes=2 << 30, precision=api.MIXED)

    order = list(range(len(systems)))      # arrival order = generation order

    def serve_all():
        t0 = time.perf_counter()
        tickets = {}
        for i in order:
            s = systems[i]
            tickets[i] = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                                    coef=s.ucoef)
        lat = []
        for i in order:
            st, l, q = srv.wait(tickets[i][0])
            assert st == 0, st
            lat.append(l)
        return time.perf_counter() - t0, lat, [tickets[i][1] for i in range(len(systems))]

    serve_all()  # warm-up (library and arena first touch)
    wall, lat, ys = min((serve_all() for _ in range(reps)), key=lambda r: r[0])
    stats = srv.stats()
    out["served"] = {"ms": wall * 1e3, "state_steps_per_s": total_steps / wall,
                     "n_weighted_per_s": weighted / wall, "precision": "mixed",
                     "lanes": args.lanes, "latency_ms": {"p50": float(np.percentile(lat, 50)),
                                                         "p99": float(np.percentile(lat, 99))},
                     "peak_running": stats["peak_running"],
                     "arena_peak_bytes": stats["arena_peak"]}
    out["served_vs_plain"] = out["served"]["state_steps_per_s"] / out["plain"]["state_steps_per_s"]

    # ---- Poisson stream
    if args.stream_seconds > 0:
        import threading
        rng = np.random.default_rng(7)
        arrivals, t = [], 0.0
        while True:
            t += rng.exponential(1.0 / args.rate)
            if t > args.stream_seconds:
                break
            arrivals.append((t, int(rng.integers(len(systems)))))
        done, lock = [], threading.Lock()
        pending = []

        def waiter():
            while True:
                with lock:
                    item = pending.pop(0) if pending else None
                if item is None:
                    time.sleep(0.0005)
                    continue
                if item == "stop":
                    return
                tk, idx = item
                st, l, q = srv.wait(tk)
                done.append((st, l, q, systems[idx].steps))

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for ta, idx in arrivals:
            dt = start + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            s = systems[idx]
            tk, _ = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
            with lock:
                pending.append((tk, idx))
        with lock:
            pending.append("stop")
        th.join()
        wall_s = time.perf_counter() - start
        lat_s = [d[1] for d in done]
        out["stream"] = {"rate_per_s": args.rate, "seconds": args.stream_seconds,
                         "requests": len(done), "failed": sum(1 for d in done if d[0] != 0),
                         "state_steps_per_s": sum(d[3] for d in done) / wall_s,
                         "latency_ms": {"p50": float(np.percentile(lat_s, 50)),
                                        "p99": float(np.percentile(lat_s, 99))},
                         "queue_ms_p99": float(np.percentile([d[2] for d in done], 99))}
    srv.close()
    return out, systems, ys, y_plain


def batch_accuracy(systems, ys, y_plain, count, wide):
    """Max relative error vs the FP64 oracle over a fixed sample of the served
    results: the `count` cheapest systems to simulate (every system when count
    >= len(systems)) and, apart, the `wide` cheapest with n_a >= 768 (where the
    delta levels and, for ABM, the two-product chunk step run); and the plain
    results of the first systems vs their served ones."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from parode_amd import synth
    idx = sorted(range(len(systems)), key=lambda i: systems[i].steps * systems[i].n_scan ** 2)
    pick = idx[:count] if count < len(systems) else idx
    worst = 0.0
    f = lambda a: a.astype(np.float64)
    for i in pick:
        s = systems[i]
        y64 = orc.simulate_f64(s.method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0),
                               synth.samples_f64(s), s.steps, s.h)
        worst = max(worst, orc.max_rel_err(ys[i], y64))
    wide_idx = [i for i in idx if systems[i].n_scan >= 768][:max(0, wide)]
    worst_wide = 0.0
    for i in wide_idx:
        s = systems[i]
        y64 = orc.simulate_f64(s.method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0),
                               synth.samples_f64(s), s.steps, s.h)
        worst_wide = max(worst_wide, orc.max_rel_err(ys[i], y64))
    plain_vs_served = max(orc.max_rel_err(yp, f(ys[i])) for i, yp in enumerate(y_plain))
    return {"served_max_rel_err_vs_fp64": worst, "checked_systems": len(pick),
            "served_max_rel_err_vs_fp64_wide": worst_wide if wide_idx else None,
            "checked_wide_systems": [[systems[i].n, systems[i].steps] for i in wide_idx],
            "plain_vs_served_max_rel": plain_vs_served}


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
    ap.add_argument("--no-tune", action="store_true", help="skip the level autotuner")
    ap.add_argument("--graph", action="store_true",
                    help="time replays of the solve captured in a HIP graph instead of eager solves")
    ap.add_argument("--batch", type=int, default=512, help="systems in the batch legs (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=2)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--rate", type=float, default=200.0, help="Poisson arrivals per second")
    ap.add_argument("--stream-seconds", type=float, default=10.0)
    ap.add_argument("--check-systems", type=int, default=48)
    ap.add_argument("--check-wide", type=int, default=8,
                    help="served systems with n_a >= 768 also checked against the FP64 oracle")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device to device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)

    import torch
    from parode_amd import api, synth

    if args.copy_probe:
        n = args.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4}))
        return

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
    # the autotuner picks each level's kernel for this shape (outside the timed region)
    tuned = 0 if args.no_tune else plan.autotune(sp)
    torch.cuda.synchronize()
    # the solve is recorded once into a HIP graph (no memset nodes) and replayed
    if args.graph:
        plan.capture(*ptrs, sp)
        torch.cuda.synchronize()
    run = (lambda: plan.replay(sp)) if args.graph else (lambda: plan.execute(*ptrs, sp))
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()

    # timed region: K solves back to back on one stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        run()
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
    lc = plan.chunk_log
    flops, pairs = level_flops(na, n_scan, s.n if s.method == 1 else None, lc,
                               factored=prec == api.MIXED)
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
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": TRAFFIC_CHUNK_PASS2 if (args.config == 1 and not args.points
                                                        and lc == 7) else None,
                     "traffic_launch": "second chunk pass (k_abm_chunks)" if lc else None,
                     "kernel": ("k_abm_chunks (both passes) and the tree levels above them"
                                if lc else "k_level_mfma (every level launch of both sweeps)"),
                     "flops_per_solve": flops, "tree_pairs_per_sweep": pairs,
                     "sweep_ms": sweep_ms, "chunk_log": lc,
                     # second chunk pass: read each element's leaf block and each
                     # chunk's carry, write each element's state block
                     "algorithmic_bytes_traffic_launch":
                         (n_scan * 8 * s.n + (n_scan >> lc) * 4 * na) if lc else None},
        "phases_ms": ph,
        "autotuned_level_launches": tuned,
        "launch": "HIP graph replay" if args.graph else "eager",
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "plan_bytes": plan.footprint,
    }
    plan.close()
    del tA, tB, tC, tD, tx, tU
    torch.cuda.empty_cache()
    if args.batch > 0:
        bl, systems, ys, y_plain = batch_legs(args, torch, api, synth)
        if not args.no_check:
            bl.update(batch_accuracy(systems, ys, y_plain, args.check_systems, args.check_wide))
        line["batch"] = bl
    if not args.no_cpu or not args.no_check:
        cb, acc = cpu_baseline_leg(s, U64, y, full_reference=not args.no_check)
        if not args.no_cpu:
            line["cpu_baseline"] = cb
        line.update(acc)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
