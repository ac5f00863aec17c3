"""The benchmark configurations at their stated sizes on the GPU.

cfg0  RK4 n=8 m=p=2 N=4096, plain FP32 (test_gpu_parity.py covers it through
      the UFunc path; here through sampled inputs as well).
cfg1  ABM n=64 m=p=8 N=2^20, noisy inputs: plain FP32 against the reference's
      algorithm on its float inputs (bar 1e-5), mixed precision against the
      FP64 sequential recurrence (bar 1e-6).
cfg2  RK4 n=128 m=p=16 N=2^18, stiff + skew: mixed precision (FP64 map and top
      6 levels) against FP64 (bar 1e-6); plain FP32 against the reference's
      algorithm (bar 1e-5).
cfg3  512 systems served together: a stratified sample (the longest horizons
      with n >= 128 of both methods, the widest systems, the cheapest and a
      seeded random draw) against FP64 (bar 1e-6), and served == alone.
cfg4  a Poisson stream from the cfg3 distribution: every sampled request's
      outputs in the mix equal its outputs alone, bit for bit.
Error metric: max_k |y_k - y*_k|_inf / max_k |y*_k|_inf."""
from concurrent.futures import ThreadPoolExecutor
import time

import numpy as np
import pytest

import oracle as orc
from parode_amd import api, synth

pytestmark = pytest.mark.gpu

PARITY = 1e-5      # plain FP32 vs the reference's algorithm
MIXED_BAR = 1e-6   # mixed precision vs FP64


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ids, e = api.get_available_gpus()
    assert e == 0 and ids, "no GPU visible"
    assert api.initialize_selected_gpus(ids[:1]) == 0
    api.set_precision(api.PLAIN32)
    api.set_max_batch_steps(0)
    yield
    api.set_precision(api.PLAIN32)
    api.close_gpu()


def _f64(a):
    return np.asarray(a, np.float64)


def _fp64_reference(s, U=None):
    U = synth.samples_f64(s) if U is None else U
    return orc.simulate_f64(s.method, _f64(s.A), _f64(s.B), _f64(s.C), _f64(s.D), _f64(s.x0), U,
                            s.steps, s.h)


def _sampled(s, U, precision):
    api.set_precision(precision)
    try:
        tv, y, e = api.simulate_ode_sampled(s.method, s.A, s.B, s.C, s.D, U, s.t0, s.t_end, s.h,
                                            s.x0)
    finally:
        api.set_precision(api.PLAIN32)
    assert e == 0, api.last_error_message()
    assert y.shape == (s.steps, s.p)
    np.testing.assert_array_equal(tv, orc.time_vector(s.method, s.t0, s.h, s.steps))
    return y


def test_cfg0_sampled_plain():
    s = synth.config(0)
    U = synth.samples_f64(s).astype(np.float32)
    y = _sampled(s, U, api.PLAIN32)
    ye = orc.simulate_mirror_exact(0, s.A, s.B, s.C, s.D, s.x0, s.t0, s.t_end, s.h, Uhalf=U)
    err = orc.max_rel_err(y, ye)
    print(f"cfg0 plain: {err:.3e}")
    assert err <= PARITY


def test_cfg1_full_plain_fp32():
    s = synth.config(1)                          # ABM n=64 m=p=8 h=5e-4 N=2^20, noise 0.1
    U = synth.samples_f64(s).astype(np.float32)
    t0 = time.perf_counter()
    y = _sampled(s, U, api.PLAIN32)
    gpu_s = time.perf_counter() - t0
    ye = orc.simulate_mirror_exact(1, s.A, s.B, s.C, s.D, s.x0, s.t0, s.t_end, s.h, Uhalf=U)
    err = orc.max_rel_err(y, ye)
    print(f"cfg1 plain FP32 (N={s.steps}): {err:.3e} vs the reference's algorithm "
          f"(C ABI call {gpu_s:.2f} s)")
    assert err <= PARITY


def test_cfg1_full_mixed():
    s = synth.config(1)
    U = synth.samples_f64(s)
    y = _sampled(s, U, api.MIXED)
    err = orc.max_rel_err(y, _fp64_reference(s, U))
    print(f"cfg1 mixed: {err:.3e} vs FP64")
    assert err <= MIXED_BAR


def test_cfg2_full_mixed():
    s = synth.config(2)                          # RK4 n=128 m=p=16 h=2e-4 N=2^18, skew 0.1
    U = synth.samples_f64(s)
    y = _sampled(s, U, api.MIXED)
    err = orc.max_rel_err(y, _fp64_reference(s, U))
    print(f"cfg2 mixed (FP64 map, top 6 levels FP64): {err:.3e} vs FP64")
    assert err <= MIXED_BAR


def test_cfg2_full_plain_fp32():
    s = synth.config(2)
    U = synth.samples_f64(s).astype(np.float32)
    y = _sampled(s, U, api.PLAIN32)
    ye = orc.simulate_mirror_exact(0, s.A, s.B, s.C, s.D, s.x0, s.t0, s.t_end, s.h, Uhalf=U)
    err = orc.max_rel_err(y, ye)
    print(f"cfg2 plain FP32: {err:.3e} vs the reference's algorithm")
    assert err <= PARITY


# ------------------------------------------------------------ served (cfg3/4)
def _stratified(systems, rng_seed=5):
    """Indices checked against FP64: per method the two longest horizons with
    n >= 128 and the two widest systems, the eight cheapest, four at random."""
    idx = set()
    for method in (0, 1):
        mine = [i for i, s in enumerate(systems) if s.method == method]
        long_wide = sorted((i for i in mine if systems[i].n >= 128),
                           key=lambda i: -systems[i].steps)
        idx.update(long_wide[:2])
        idx.update(sorted(mine, key=lambda i: -systems[i].n)[:2])
    cost = sorted(range(len(systems)), key=lambda i: systems[i].steps * systems[i].n_scan ** 2)
    idx.update(cost[:8])
    rest = [i for i in range(len(systems)) if i not in idx]
    idx.update(np.random.default_rng(rng_seed).choice(rest, 4, replace=False).tolist())
    return sorted(idx)


def _serve(systems, lanes=8, order=None):
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=lanes, device_bytes=int(need * 1.1) + (1 << 30),
                     pinned_bytes=1 << 30, precision=api.MIXED)
    try:
        order = range(len(systems)) if order is None else order
        got = {}
        for i in order:
            s = systems[i]
            got[i] = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
        ys = {}
        for i, (t, Y) in got.items():
            st = srv.wait(t)[0]
            assert st == 0, (i, st)
            ys[i] = Y
    finally:
        srv.close()
    return ys


def test_cfg3_served_batch_stratified_against_fp64():
    systems = synth.batch_config(512)
    ys = _serve(systems)
    pick = _stratified(systems)
    with ThreadPoolExecutor(max_workers=8) as ex:
        refs = list(ex.map(lambda i: _fp64_reference(systems[i]), pick))
    worst = 0.0
    for i, y64 in zip(pick, refs):
        s = systems[i]
        err = orc.max_rel_err(ys[i], y64)
        print(f"cfg3[{i}] {'ABM' if s.method else 'RK4'} n={s.n} N={s.steps}: {err:.3e}")
        worst = max(worst, err)
        assert err <= MIXED_BAR, (i, s.method, s.n, s.steps, err)
    print(f"cfg3 stratified sample of {len(pick)}: worst {worst:.3e}")


def test_cfg3_served_equals_alone():
    """A request's outputs do not depend on what else is served with it: the
    whole batch in reversed arrival order, then a sample of requests each
    alone, give bitwise the outputs of the batch in arrival order."""
    systems = synth.batch_config(512)
    ys = _serve(systems)
    rev = _serve(systems, order=list(range(len(systems)))[::-1])
    for i in range(len(systems)):
        np.testing.assert_array_equal(ys[i], rev[i], err_msg=f"cfg3[{i}]")
    for i in _stratified(systems, rng_seed=9)[:10]:
        alone = _serve([systems[i]], lanes=1)[0]
        np.testing.assert_array_equal(ys[i], alone, err_msg=f"cfg3[{i}] alone")


def test_cfg4_poisson_stream_mix_equals_alone():
    """Poisson arrivals (200 per second, 2 s) drawn from the cfg3 distribution,
    submitted while earlier requests run; a sample of them rerun alone gives
    bitwise the same outputs, and a few are checked against FP64."""
    systems = synth.batch_config(512)
    rng = np.random.default_rng(11)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / 200.0)
        if t > 2.0:
            break
        arrivals.append((t, int(rng.integers(len(systems)))))
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=8, device_bytes=min(int(need * 1.1), 64 << 30) + (1 << 30),
                     pinned_bytes=1 << 30, precision=api.MIXED)
    got = []
    try:
        start = time.perf_counter()
        for ta, i in arrivals:
            dt = start + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            s = systems[i]
            got.append((i, srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                                      coef=s.ucoef)))
        lat = []
        for i, (tk, Y) in got:
            st, l, q = srv.wait(tk)
            assert st == 0, (i, st)
            lat.append(l)
        stats = srv.stats()
    finally:
        srv.close()
    print(f"cfg4: {len(got)} requests, p50 {np.percentile(lat, 50):.2f} ms, "
          f"p99 {np.percentile(lat, 99):.2f} ms, peak running {stats['peak_running']}")
    assert stats["failed"] == 0 and stats["completed"] == len(got)
    sample = np.random.default_rng(12).choice(len(got), 12, replace=False)
    for k in sample:
        i, (tk, Y) = got[k]
        alone = _serve([systems[i]], lanes=1)[0]
        np.testing.assert_array_equal(Y, alone, err_msg=f"request {k} (cfg3[{i}])")
    cheap = sorted(sample, key=lambda k: systems[got[k][0]].steps * systems[got[k][0]].n_scan ** 2)
    for k in cheap[:3]:
        i, (tk, Y) = got[k]
        assert orc.max_rel_err(Y, _fp64_reference(systems[i])) <= MIXED_BAR


# ---------------------------------------------------- batches (NSB) on the GPU
@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("nsb", [64, 256, 1024])
def test_forced_batches_match_the_batched_oracle(solver, nsb):
    """A horizon split into batches of NSB elements (the last one padded to a
    power of two, GPUFixedStepIterative.cpp:191-221) with UFunc inputs, whose
    float time grid restarts at float(t_b) in every batch
    (RungeKutta4thOrder.cpp:255-257, AdamsMoulton.cpp:426-428), against the
    oracle with the same NSB at the plain parity bar.  The horizon reaches
    t ~ 20, where the restart moves the result by more than the bar, so a
    single-batch oracle would not pass."""
    s = synth.make_system(solver, 8, 2, 2, -4.0, -0.2, 5e-3, 4001, seed=80 + solver)
    idx, e = api.register_input(synth.ufunc_source(s.ucoef))
    assert e == 0, api.last_error_message()
    api.set_max_batch_steps(nsb)
    try:
        tv, y, e = api.simulate_ode(solver, s.A, s.B, s.C, s.D, idx, 0.0, s.t_end, s.h, s.x0)
    finally:
        api.set_max_batch_steps(0)
    assert e == 0 and y.shape == (4001, 2)
    np.testing.assert_array_equal(tv, orc.time_vector(solver, 0.0, s.h, 4001))
    yb = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef, nsb=nsb)
    y1 = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef)
    err, moved = orc.max_rel_err(y, yb), orc.max_rel_err(yb, y1)
    print(f"NSB {nsb}: {err:.3e} vs batched oracle; batching moves the result {moved:.3e}")
    assert err <= PARITY


# ------------------------------------------------------ K-split 32 x 32 tiles
def test_ksplit_wide_tiles_meet_parity():
    """A level with 4096 pairs at n_a = 256 gives the K-split kernel
    ceil(4096 / 32) * ceil(256 / 32) = 1024 work-groups, the threshold from
    which it takes 32 x 32 tiles (RK4 n = 256, N = 8192, chunk recurrences
    off so that level 0 is a tree level)."""
    s = synth.make_system(0, 256, 4, 3, -6.0, -0.2, 1e-3, 8192, seed=77)
    U = synth.samples_f64(s).astype(np.float32)
    api.set_level_kernel(api.NO_CHUNKS)
    try:
        y = _sampled(s, U, api.PLAIN32)
    finally:
        api.set_level_kernel(0)
    ye = orc.simulate_mirror_exact(0, s.A, s.B, s.C, s.D, s.x0, s.t0, s.t_end, s.h, Uhalf=U)
    err = orc.max_rel_err(y, ye)
    print(f"K-split 32 x 32 tiles: {err:.3e}")
    assert err <= PARITY
