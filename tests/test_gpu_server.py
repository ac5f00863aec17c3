"""The continuous-batching request loop (ParODE_Server*): results in a mix of
concurrent requests equal the results of the same requests run alone, meet the
FP64 bar, and bad requests fail alone."""
import numpy as np
import pytest

import oracle as orc
from parode_amd import api, synth

pytestmark = pytest.mark.gpu


def _mix(count=24, seed=40, n_hi=48, N_hi=1 << 12):
    return synth.batch_config(count, seed=seed, n_lo=2, n_hi=n_hi, N_lo=1 << 4, N_hi=N_hi)


def _submit_all(srv, systems):
    out = []
    for s in systems:
        t, Y = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
        out.append((t, Y))
    return out


def test_mix_equals_alone_and_meets_fp64_bar():
    systems = _mix()
    srv = api.Server(lanes=4, device_bytes=2 << 30, pinned_bytes=256 << 20)
    got = _submit_all(srv, systems)
    ys = []
    for t, Y in got:
        st, lat, q = srv.wait(t)
        assert st == 0
        assert lat >= q >= 0.0
        ys.append(Y.copy())
    stats = srv.stats()
    assert stats["completed"] == len(systems) and stats["failed"] == 0
    print("peak running", stats["peak_running"])
    srv.close()

    alone = api.Server(lanes=1, device_bytes=2 << 30, pinned_bytes=256 << 20)
    for s, y in zip(systems, ys):
        t, Y = alone.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
        assert alone.wait(t)[0] == 0
        np.testing.assert_array_equal(Y, y)
    alone.close()

    for s, y in list(zip(systems, ys))[:8]:
        U = synth.samples_f64(s)
        f = lambda a: a.astype(np.float64)
        y64 = orc.simulate_f64(s.method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0), U, s.steps, s.h)
        assert orc.max_rel_err(y, y64) <= 1e-6, (s.tag, s.n, s.steps)


def test_host_samples_match_device_sampling():
    s = synth.make_system(1, 6, 2, 3, -5.0, -0.1, 1e-3, 700, seed=41)
    srv = api.Server(lanes=2, device_bytes=1 << 30, pinned_bytes=64 << 20)
    t1, Y1 = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
    t2, Y2 = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                        samples=synth.samples_f64(s))
    assert srv.wait(t1)[0] == 0 and srv.wait(t2)[0] == 0
    assert orc.max_rel_err(Y1, Y2.astype(np.float64)) <= 1e-7
    srv.close()


@pytest.mark.parametrize("solver", [0, 1])
def test_device_sampling_holds_over_a_long_horizon(solver):
    """Device samples come in runs of 16 steps (one sincos, then rotations);
    over 2^16 steps (8192 runs per input, the h/2 grid reaching t = 65) the
    outputs equal those of the host samples to 1e-7, as on short horizons."""
    s = synth.make_system(solver, 12, 3, 2, -5.0, -0.1, 1e-3, 1 << 16, seed=45)
    srv = api.Server(lanes=2, device_bytes=1 << 30, pinned_bytes=64 << 20)
    t1, Y1 = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
    t2, Y2 = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                        samples=synth.samples_f64(s))
    assert srv.wait(t1)[0] == 0 and srv.wait(t2)[0] == 0
    assert orc.max_rel_err(Y1, Y2.astype(np.float64)) <= 1e-7
    srv.close()


def test_noisy_inputs_sampled_on_device_match_the_oracle_generator():
    s = synth.make_system(0, 8, 2, 2, -5.0, -0.1, 1e-3, 1000, seed=42, noise_sigma=0.1)
    srv = api.Server(lanes=1, device_bytes=1 << 30, pinned_bytes=64 << 20)
    t, Y = srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef,
                      sigma=s.noise_sigma, seed=s.noise_seed)
    assert srv.wait(t)[0] == 0
    f = lambda a: a.astype(np.float64)
    y64 = orc.simulate_f64(0, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0), synth.samples_f64(s),
                           s.steps, s.h)
    assert orc.max_rel_err(Y, y64) <= 1e-6
    srv.close()


def test_bad_and_oversized_requests_fail_alone():
    s = synth.make_system(0, 4, 1, 1, -5.0, -0.1, 1e-3, 64, seed=43)
    srv = api.Server(lanes=2, device_bytes=64 << 20, pinned_bytes=16 << 20)
    ok1, Y1 = srv.submit(0, s.A, s.B, s.C, s.D, s.x0, 64, s.h, coef=s.ucoef)
    bad, _ = srv.submit(3, s.A, s.B, s.C, s.D, s.x0, 64, s.h, coef=s.ucoef)      # solver
    short, _ = srv.submit(1, s.A, s.B, s.C, s.D, s.x0, 3, s.h, coef=s.ucoef)     # ABM < 4
    big = synth.make_system(1, 64, 8, 8, -5.0, -0.1, 1e-3, 1 << 16, seed=44)
    huge, _ = srv.submit(1, big.A, big.B, big.C, big.D, big.x0, big.steps, big.h, coef=big.ucoef)
    ok2, Y2 = srv.submit(0, s.A, s.B, s.C, s.D, s.x0, 64, s.h, coef=s.ucoef)
    assert srv.wait(bad)[0] == -6
    assert srv.wait(short)[0] == -6
    assert srv.wait(huge)[0] == -4
    assert srv.wait(ok1)[0] == 0 and srv.wait(ok2)[0] == 0
    np.testing.assert_array_equal(Y1, Y2)
    st = srv.stats()
    assert st["completed"] == 2 and st["failed"] == 3
    srv.close()


def test_precision_modes_of_the_server():
    s = synth.make_system(0, 16, 2, 2, -10.0, -0.05, 5e-4, 1 << 13, seed=45)
    f = lambda a: a.astype(np.float64)
    y64 = orc.simulate_f64(0, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0), synth.samples_f64(s),
                           s.steps, s.h)
    for prec, bar in ((api.FP64, 1e-7), (api.MIXED, 1e-6)):
        srv = api.Server(lanes=1, device_bytes=1 << 30, pinned_bytes=64 << 20, precision=prec)
        t, Y = srv.submit(0, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
        assert srv.wait(t)[0] == 0
        err = orc.max_rel_err(Y, y64)
        print(f"server precision {prec}: {err:.2e}")
        assert err <= bar
        srv.close()
