"""Benchmark: one MI355X, the parallel-in-time LTI solve, state-steps/s.

Headline (`value`): the served batch -- 512 independent RK4/ABM LTI systems
(50/50), n log-uniform on [4, 256], m = p = max(1, n/8), N log-uniform on
[2^10, 2^18], h ~ U[1e-4, 1e-3], inputs sin(w t + phi) sampled in FP64 on the
device -- submitted at once to the continuous-batching server (mixed
precision, grouped launches: every stage of the scan one launch for all
systems), outputs written to device memory.  A step is one pass over the whole
batch, from the first submission to the last completion (host clock, GPU
synchronised on both sides); value = sum of N_i / time per pass.

Beside it, in the same JSON line:
  plain        the plain FP32 path, one system per solve, solves back to back on
               one stream (hipEvents), inputs and outputs resident in HBM;
               served_vs_plain = value / plain rate
  group        the same grouped solve through ParODE_GroupExecuteTimed: GPU
               time of each stage (hipEvents on its stream); `roofline` is the
               chunk-recurrence stage (the dominant kernels)
  served_host  one pass with outputs copied back to host memory
  stream       Poisson arrivals from the same distribution, p50/p99 latency
  accuracy     max relative error vs the FP64 sequential recurrence on a
               stratified sample (longest horizons with n >= 128 of both
               methods, the widest, the cheapest, a random draw)
  single       the single-system configurations: ABM n=64 N=2^20 (mixed and
               plain FP32), RK4 n=8 N=4096 and RK4 n=128 N=2^18 per-launch
               p50/p99
  cpu_baseline the reference's CPU paths (classic RK4 with float state / ABM4
               PECE, TestsSources/TestParODEStaticLibCInterface.cpp:158-202,
               :336-461, C port) over a thread pool on the host cores, on a
               bounded prefix of every system of the batch
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3e12, "fp64": 78.6e12, "hbm": 8.0e12}
METRIC = "state-steps/s at <=1e-6 rel err vs FP64, served vs plain scan; p50/p99 latency"


# ------------------------------------------------------------ work accounting
def _pow2(v):
    p = 1
    while p < v:
        p <<= 1
    return p


def scan_shape(method, n, points):
    """(n_scan, nPad, L, na, chunk log lc, top level lt) as the library plans a
    mixed-precision solve: chunks of 2^lc elements (lc = min(7, L - 10), at
    least 2), tree levels [lc, lt), the rest in one work-group."""
    n_scan = points if method == 0 else points - 3
    npad = _pow2(n_scan)
    L = npad.bit_length() - 1
    na = n if method == 0 else 4 * n
    pairs = lambda l: (n_scan + (2 << l) - 1) // (2 << l)
    if npad * na <= 65536:
        lt = 0
    elif na >= 64:
        lt = L
    else:
        lt = 0
        while lt < L and pairs(lt) > 32:
            lt += 1
    lc = min(7, L - 10)
    if lc < 2 or lt <= lc:
        lc = 0
    return n_scan, npad, L, na, lc, lt


def chunk_flops(method, n, points):
    """Algorithmic flops of both chunk-recurrence passes: per element and pass,
    RK4 2 n^2 (the dense map), ABM 4 n^2 (x3 + A w1 + A^2 w2 factored, or
    x3 + A (w1 + A w2))."""
    n_scan, _, _, _, lc, _ = scan_shape(method, n, points)
    if not lc:
        return 0.0
    return 2 * n_scan * (2.0 * n * n if method == 0 else 4.0 * n * n)


def level_flops(na, n_scan, n_leaf=None, chunk_log=0, factored=False):
    """Algorithmic flops of one solve's sweeps (both directions): 2 na^2 per
    pair of every tree level (pairs of level l = ceil(N / 2^(l+1))); levels
    below chunk_log as two chunk passes per element (ABM: the n x 4n block, or
    2 n (2n) factored; RK4: 2 na^2); ABM levels 0 and 1 without chunks need
    fewer products (M's shift rows, the leaves' zero rows).  Returns (flops,
    tree pairs per sweep)."""
    L = max(0, (n_scan - 1).bit_length())
    per = [(n_scan + (2 << l) - 1) // (2 << l) for l in range(L)]
    if chunk_log:
        per_step = (2.0 * n_leaf * (2 * n_leaf) if factored else 2.0 * n_leaf * na) \
            if n_leaf else 2.0 * na * na
        tree = per[chunk_log:]
        return 2 * per_step * n_scan + 2 * 2.0 * na * na * sum(tree), sum(tree)
    pairs = sum(per)
    flops = 2 * 2.0 * na * na * pairs
    if n_leaf and L:
        n = n_leaf
        flops -= per[0] * (2 * 2.0 * na * na - 2.0 * n * n - 2.0 * n * na)
        if L > 1:
            flops -= per[1] * (2 * 2.0 * na * na - 2.0 * (2 * n) * (2 * n) - 2.0 * (2 * n) * na)
    return flops, pairs


def system_bound_seconds(method, n, m, p, points):
    """Roofline time of one system's mixed-precision solve as built: FP32
    matrix-core flops (chunk passes and tree levels), FP64 matrix-core flops
    (the power squarings up to P_{L-1}: 2 n_a^3 each, 56 n^3 for the ABM
    companion; the leaves, 2 n k m per step with k = 3 for RK4 and 2 for the
    factored ABM; the outputs, 2 p (n + m) per point), the two kinds sharing
    the matrix cores, against the compulsory HBM bytes (FP64 samples on the h/2
    grid in, FP32 outputs out)."""
    n_scan, _, L, na, lc, _ = scan_shape(method, n, points)
    f32, _ = level_flops(na, n_scan, n_leaf=n if method else None, chunk_log=lc,
                         factored=bool(method))
    sq = 56.0 * n ** 3 if method else 2.0 * na ** 3
    f64 = max(L - 1, 0) * sq + 2.0 * n * (2 if method else 3) * m * n_scan \
        + 2.0 * p * (n + m) * points
    hbm = (2 * points + 1) * m * 8.0 + points * p * 4.0
    return max(f32 / PEAK["fp32"] + f64 / PEAK["fp64"], hbm / PEAK["hbm"])


# ------------------------------------------------------------------- helpers
def _sync_time(torch, fn, reps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def _dev_system(torch, s, dev, with_u=True, y=True):
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    t = dict(A=T(s.A), B=T(s.B), C=T(s.C), D=T(s.D), x0=T(s.x0), coef=T(s.ucoef))
    if with_u:
        t["U"] = torch.empty((2 * s.steps + 1, s.m), dtype=torch.float64, device=dev)
    if y:
        t["Y"] = torch.empty((s.steps, s.p), dtype=torch.float32, device=dev)
    return t


def _torch_samples(torch, s, dev, dtype):
    """u(t0 + k h/2), k = 0 .. 2N, of the system's sinusoid family, on the device."""
    k = torch.arange(2 * s.steps + 1, dtype=torch.float64, device=dev)
    t = (s.t0 + k * (s.h / 2.0))[:, None]
    c = torch.from_numpy(np.ascontiguousarray(s.ucoef)).to(dev)
    g = c[:, 4][None, :] * t + c[:, 5][None, :]
    U = c[:, 0] * torch.sin(g) + c[:, 1] * torch.cos(g) + c[:, 2] + c[:, 3] * t
    return U.to(dtype).contiguous()


def stratified(systems, seed=5):
    """Per method the two longest horizons with n >= 128 and the two widest
    systems, the eight cheapest, four at random."""
    idx = set()
    for method in (0, 1):
        mine = [i for i, s in enumerate(systems) if s.method == method]
        idx.update(sorted((i for i in mine if systems[i].n >= 128),
                          key=lambda i: -systems[i].steps)[:2])
        idx.update(sorted(mine, key=lambda i: -systems[i].n)[:2])
    cost = sorted(range(len(systems)), key=lambda i: systems[i].steps * systems[i].n_scan ** 2)
    idx.update(cost[:8])
    rest = [i for i in range(len(systems)) if i not in idx]
    idx.update(np.random.default_rng(seed).choice(rest, 4, replace=False).tolist())
    return sorted(idx)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# ---------------------------------------------------------------------- legs
def served_leg(args, torch, api, systems, dev):
    """All systems handed to the server in one batch submission; device outputs."""
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=args.lanes, device_bytes=int(need * 1.1) + (1 << 30),
                     pinned_bytes=1 << 30, precision=api.MIXED)
    outs = [torch.empty((s.steps, s.p), dtype=torch.float32, device=dev) for s in systems]
    lat = []

    items = [(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, o.data_ptr(), s.ucoef)
             for s, o in zip(systems, outs)]

    def one_pass():
        tk = srv.submit_device_batch(items)
        st, l, q = srv.wait_batch(tk)
        assert not st.any(), st
        lat.extend(l.tolist())

    for _ in range(args.warmup):
        one_pass()
    lat.clear()
    sec = _sync_time(torch, one_pass, args.steps)
    gs = srv.group_stats()
    stats = srv.stats()
    ys = {i: outs[i] for i in range(len(systems))}

    # one pass with host outputs (the copy back to pinned staging and out)
    t0 = time.perf_counter()
    got = [srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
           for s in systems]
    for t, _ in got:
        assert srv.wait(t)[0] == 0
    host_s = time.perf_counter() - t0
    srv.close()
    return sec, ys, {"latency_ms": {"p50": float(np.percentile(lat, 50)),
                                    "p99": float(np.percentile(lat, 99))},
                     "groups_per_pass": gs["groups"] / max(1, args.steps + args.warmup),
                     "largest_group": gs["largest_group"],
                     "arena_peak_bytes": stats["arena_peak"],
                     "host_outputs_ms": host_s * 1e3}


def plain_leg(args, torch, api, systems, dev):
    """FP32 plain path: one system per solve, back to back on one stream."""
    plans, bufs = [], []
    for s in systems:
        plans.append(api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, api.PLAIN32, False))
        t = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in (s.A, s.B, s.C, s.D, s.x0)]
        t.append(_torch_samples(torch, s, dev, torch.float32))
        t.append(torch.empty((s.steps, s.p), dtype=torch.float32, device=dev))
        bufs.append(t)
    stream = torch.cuda.Stream(device=dev)
    sp = stream.cuda_stream

    def run():
        for pl, t in zip(plans, bufs):
            pl.execute(*(x.data_ptr() for x in t), sp)

    run()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(1, args.plain_reps)
    ev0.record(stream)
    for _ in range(reps):
        run()
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    y_plain = {i: bufs[i][-1].cpu().numpy() for i in range(min(8, len(systems)))}
    for pl in plans:
        pl.close()
    del plans, bufs
    torch.cuda.empty_cache()
    return ms, y_plain


def group_leg(args, torch, api, systems, dev):
    """The grouped solve alone, GPU time per stage (hipEvents on its stream)."""
    keep = [_dev_system(torch, s, dev) for s in systems]
    recs = [dict(method=s.method, n=s.n, m=s.m, p=s.p, points=s.steps, h=s.h,
                 **{k: v.data_ptr() for k, v in t.items()}) for s, t in zip(systems, keep)]
    g = api.Group(recs)
    stream = torch.cuda.Stream(device=dev)
    sp = stream.cuda_stream
    g.execute(sp)
    torch.cuda.synchronize()
    st = [g.execute_timed(sp) for _ in range(max(3, args.steps))]
    stages = {k: float(np.median([x[k] for x in st])) for k in st[0]}
    total = float(np.median([sum(x.values()) for x in st]))
    # the same solve with the power tables and the start-up on the side stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps)
    ev0.record(stream)
    for _ in range(reps):
        g.execute(sp)
    ev1.record(stream)
    torch.cuda.synchronize()
    out = {"launches": g.launches, "bytes": g.footprint, "stage_ms": stages,
           "ms_stages_serial": total, "ms": ev0.elapsed_time(ev1) / reps}
    g.close()
    del keep
    torch.cuda.empty_cache()
    return out


def stream_leg(args, api, systems):
    import threading
    rng = np.random.default_rng(7)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / args.rate)
        if t > args.stream_seconds:
            break
        arrivals.append((t, int(rng.integers(len(systems)))))
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=args.lanes, device_bytes=min(int(need * 1.1), 96 << 30) + (1 << 30),
                     pinned_bytes=2 << 30, precision=api.MIXED)
    done, lock, pending = [], threading.Lock(), []

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
    last_report = start
    for ta, idx in arrivals:
        dt = start + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        s = systems[idx]
        tk, _ = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
        with lock:
            pending.append((tk, idx))
        if time.perf_counter() - last_report > 15.0:
            last_report = time.perf_counter()
            print(f"stream: {len(done)} of {len(arrivals)} requests done", file=sys.stderr,
                  flush=True)
    with lock:
        pending.append("stop")
    th.join()
    wall = time.perf_counter() - start
    gs = srv.group_stats()
    srv.close()
    lat = [d[1] for d in done]
    return {"rate_per_s": args.rate, "seconds": args.stream_seconds, "requests": len(done),
            "failed": sum(1 for d in done if d[0] != 0),
            "state_steps_per_s": sum(d[3] for d in done) / wall,
            "latency_ms": {"p50": float(np.percentile(lat, 50)),
                           "p99": float(np.percentile(lat, 99))},
            "queue_ms_p99": float(np.percentile([d[2] for d in done], 99)),
            "groups": gs["groups"], "largest_group": gs["largest_group"]}


def accuracy_leg(systems, ys, y_plain):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from parode_amd import synth
    pick = stratified(systems)
    f = lambda a: a.astype(np.float64)

    def err(i):
        s = systems[i]
        y64 = orc.simulate_f64(s.method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0),
                               synth.samples_f64(s), s.steps, s.h)
        return orc.max_rel_err(ys[i], y64)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 8)) as ex:
        errs = list(ex.map(err, pick))
    plain_vs_served = max(orc.max_rel_err(yp, f(ys[i])) for i, yp in y_plain.items())
    return {"served_max_rel_err_vs_fp64": max(errs),
            "checked": [[("ABM" if systems[i].method else "RK4"), systems[i].n, systems[i].steps,
                         e] for i, e in zip(pick, errs)],
            "plain_vs_served_max_rel": plain_vs_served}


def single_legs(args, torch, api, synth, dev):
    """The single-system configurations through device-resident plans."""
    out = {}
    stream = torch.cuda.Stream(device=dev)
    sp = stream.cuda_stream
    for name, k, prec in (("abm_n64_N1M_mixed", 1, api.MIXED), ("abm_n64_N1M_plain", 1, api.PLAIN32),
                          ("rk4_n8_N4096_plain", 0, api.PLAIN32),
                          ("rk4_n128_N256K_mixed", 2, api.MIXED)):
        s = synth.config(k)
        U64 = synth.samples_f64(s)
        U = U64 if prec != api.PLAIN32 else U64.astype(np.float32)
        t = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in (s.A, s.B, s.C, s.D, s.x0, U)]
        Y = torch.zeros((s.steps, s.p), dtype=torch.float32, device=dev)
        plan = api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, prec, U.dtype == np.float64)
        ptrs = [x.data_ptr() for x in t] + [Y.data_ptr()]
        tuned = plan.autotune(sp) if k != 0 else 0
        for _ in range(3):
            plan.execute(*ptrs, sp)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = args.single_reps
        ev0.record(stream)
        for _ in range(reps):
            plan.execute(*ptrs, sp)
        ev1.record(stream)
        torch.cuda.synchronize()
        ms = ev0.elapsed_time(ev1) / reps
        phases, lat = [], []
        for _ in range(args.single_reps):
            t0 = time.perf_counter()
            phases.append(plan.execute_timed(*ptrs, sp))
            lat.append((time.perf_counter() - t0) * 1e3)
        ph = {q: float(np.mean([p[q] for p in phases])) for q in phases[0]}
        rec = {"workload": f"{'ABM4' if s.method else 'RK4'} n={s.n} m={s.m} p={s.p} h={s.h} "
                           f"N={s.steps}", "precision": {0: "plain", 1: "mixed"}[prec],
               "ms": ms, "state_steps_per_s": s.steps / (ms * 1e-3),
               "latency_ms": {"p50": float(np.percentile(lat, 50)),
                              "p99": float(np.percentile(lat, 99))},
               "phases_ms": ph, "autotuned_level_launches": tuned}
        if k == 1 and prec == api.MIXED:
            na, n_scan, lc = s.n_scan, s.steps - 3, plan.chunk_log
            flops, _ = level_flops(na, n_scan, s.n, lc, factored=True)
            sweep_ms = ph["up_sweep"] + ph["down_sweep"]
            rec["sweeps_tflops"] = flops / (sweep_ms * 1e-3) / 1e12
        if not args.no_check:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as orc
            y = Y.cpu().numpy()
            if prec == api.PLAIN32:
                ye = orc.simulate_mirror_exact(s.method, s.A, s.B, s.C, s.D, s.x0, s.t0, s.t_end,
                                               s.h, Uhalf=U)
                rec["max_rel_err_vs_reference_algorithm"] = orc.max_rel_err(y, ye)
            f = lambda a: a.astype(np.float64)
            y64 = orc.simulate_f64(s.method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0), U64,
                                   s.steps, s.h)
            rec["max_rel_err_vs_fp64"] = orc.max_rel_err(y, y64)
        out[name] = rec
        plan.close()
        del t, Y
        torch.cuda.empty_cache()
    return out


def cpu_baseline_leg(systems, target_s=20.0):
    """The reference's CPU paths over all host cores (a thread pool, one system
    per task) on a bounded prefix of every system of the batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from parode_amd import synth
    cores = os.cpu_count() or 1
    threads = min(cores, 64)

    def run(s, pts):
        U = synth.samples_f64(s, count=2 * pts + 1).astype(np.float32)
        cpu = orc.cpu_abm_pece if s.method == 1 else orc.cpu_rk4_odeint
        t0 = time.perf_counter()
        cpu(s.A, s.B, s.C, s.D, s.x0, U, pts, s.h)
        return time.perf_counter() - t0

    # probe: 64 points of every system; then the prefix length whose modelled
    # cost (n^2 per point and system, the probe's rate) is about target_s
    probe = 64
    cost = lambda q: sum(s.n * s.n * min(q, s.steps) for s in systems)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda s: run(s, min(probe, s.steps)), systems))
        dt = time.perf_counter() - t0
    rate = cost(probe) / max(dt, 1e-3)
    lo, hi = probe, 1 << 18
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if cost(mid) / rate <= target_s:
            lo = mid
        else:
            hi = mid - 1
    pts = lo
    prefix = [min(pts, s.steps) for s in systems]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(lambda a: run(*a), zip(systems, prefix)))
        wall = time.perf_counter() - t0
    steps = float(sum(prefix))
    return {"value": steps / wall, "unit": "state-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "hardware_concurrency": cores,
            "sample": f"the first min(N, {pts}) points of each of the {len(systems)} systems "
                      f"({int(steps)} state-steps), classic RK4 with float state for RK4 systems "
                      f"and ABM4 PECE for ABM ones, one system per task on a pool of {threads} "
                      f"threads, {wall:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed passes over the served batch")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--plain-reps", type=int, default=2)
    ap.add_argument("--rate", type=float, default=200.0, help="Poisson arrivals per second")
    ap.add_argument("--stream-seconds", type=float, default=60.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--single-reps", type=int, default=20)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="copy BYTES once device to device and exit (calibrates the HBM "
                         "byte counters against a copy of known size)")
    ap.add_argument("--only", choices=["group", "served", "plain", "single"], default=None,
                    help="run one leg alone (for profilers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not run)"}))
        sys.exit(2)

    import torch
    from parode_amd import api, synth

    dev = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.ones(args.copy_probe // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return
    systems = synth.batch_config(args.batch)
    total = float(sum(s.steps for s in systems))
    weighted = float(sum(s.steps * s.n for s in systems))
    log = lambda msg: print(msg, file=sys.stderr, flush=True)

    if args.only == "group":
        g = group_leg(args, torch, api, systems, dev)
        print(json.dumps({"group": g}))
        return
    if args.only == "single":
        print(json.dumps({"single": single_legs(args, torch, api, synth, dev)}))
        return
    if args.only == "plain":
        ms, _ = plain_leg(args, torch, api, systems, dev)
        print(json.dumps({"plain_ms": ms}))
        return

    log("served leg")
    sec, ys, served = served_leg(args, torch, api, systems, dev)
    value = total / sec
    if args.only == "served":
        print(json.dumps({"served_ms": sec * 1e3, "value": value, **served}))
        return
    ys_host = {i: ys[i].cpu().numpy() for i in set(stratified(systems)) | set(range(8))}
    del ys
    torch.cuda.empty_cache()
    log("group leg")
    grp = group_leg(args, torch, api, systems, dev)
    log("plain leg")
    plain_ms, y_plain = plain_leg(args, torch, api, systems, dev)

    # roofline: the chunk-recurrence passes (dominant stages of the grouped solve)
    cflops = sum(chunk_flops(s.method, s.n, s.steps) for s in systems)
    chunk_ms = grp["stage_ms"]["chunks_pass1"] + grp["stage_ms"]["chunks_pass2"]
    achieved = cflops / (chunk_ms * 1e-3)
    bound = total / sum(system_bound_seconds(s.method, s.n, s.m, s.p, s.steps) for s in systems)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "state-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded systems; inputs sampled in FP64 on the device)",
        "config": {"workload": f"{len(systems)} independent RK4/ABM LTI systems (50/50), n "
                               f"log-uniform [4, 256], m = p = max(1, n/8), N log-uniform "
                               f"[2^10, 2^18], served at once, device outputs",
                   "systems": len(systems), "state_steps": total, "precision": "mixed",
                   "lanes": args.lanes},
        "served_vs_plain": value / (total / (plain_ms * 1e-3)),
        "served": {"ms": sec * 1e3, "state_steps_per_s": value,
                   "n_weighted_per_s": weighted / sec, **served},
        "plain": {"ms": plain_ms, "state_steps_per_s": total / (plain_ms * 1e-3),
                  "precision": "plain fp32", "launch": "one system per solve, one stream"},
        "group": grp,
        "batch_roofline": {
            "bound_state_steps_per_s": bound,
            "frac": value / bound,
            "model": "sum N_i / sum max(F32_i / 157.3 TF + F64_i / 78.6 TF, B_i / 8 TB/s) with "
                     "the as-built flops (chunk passes, tree levels, FP64 squarings, leaves, "
                     "outputs) and the compulsory bytes (FP64 samples in, FP32 outputs out)"},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"], "traffic": None,
                     "kernel": "grouped chunk recurrences (kg_chunks_fac / kg_chunks_two / "
                               "kg_chunks_dense), both passes",
                     "flops_per_pass_pair": cflops, "stage_ms": chunk_ms},
    }
    log("stream leg")
    if args.stream_seconds > 0:
        line["stream"] = stream_leg(args, api, systems)
    if not args.no_check:
        log("accuracy")
        line["accuracy"] = accuracy_leg(systems, ys_host, y_plain)
    if not args.no_single:
        log("single-system legs")
        line["single"] = single_legs(args, torch, api, synth, dev)
    if not args.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline_leg(systems)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
