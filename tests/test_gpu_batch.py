"""Grouped launches of the served path (BatchSolve): many systems of uneven
state size and horizon run each stage of the scan in one launch per kernel.
Every padded-size variant of the chunk recurrences (RK4 dense 16..256, ABM
factored 16..128 and two-product 160..256, in steps of 16 or 32), short horizons without chunks,
tiny systems whose whole tree runs in one work-group, and scan sizes below 16
(FP64 tree levels, separate copy buffers) are checked against the FP64
sequential recurrence at the mixed-precision bar, in one group; each result
equals the same request run alone bit for bit; device outputs equal host
outputs."""
import numpy as np
import pytest

# torch before the library starts the HIP runtime (its own start-up then finds the GPU)
torch = pytest.importorskip("torch")

import oracle as orc  # noqa: E402
from parode_amd import api, synth  # noqa: E402

pytestmark = pytest.mark.gpu

MIXED_BAR = 1e-6

# (method, n, m, p, points): one of each grouped kernel variant
SHAPES = [
    (0, 4, 1, 1, 300),        # whole tree in one work-group
    (0, 6, 1, 2, 40000),      # scan size < 16: FP64 tree levels, own copy buffers
    (0, 12, 2, 2, 70000),
    (0, 20, 2, 3, 9000),      # dense chunks, 32 padded states
    (0, 40, 5, 5, 30000),
    (0, 100, 12, 12, 20000),
    (0, 200, 25, 25, 15000),  # dense chunks, 256 padded states
    (1, 3, 1, 1, 60),         # ABM, no chunks
    (1, 10, 1, 2, 50000),     # factored chunks, 16
    (1, 30, 3, 3, 20000),
    (1, 50, 6, 6, 30000),
    (1, 100, 12, 12, 9000),   # factored, 128
    (1, 150, 18, 18, 12000),  # two-product chunks, 160
    (1, 190, 23, 23, 9000),
    (1, 210, 26, 26, 8000),
    (1, 250, 30, 30, 6000),   # two-product, 256
    (0, 64, 8, 8, 2000),      # short horizon: no chunks, K-split tree from level 0
    (1, 20, 2, 2, 3000),
    (0, 45, 5, 5, 20000),     # dense chunks, 48 padded states
    (0, 90, 11, 11, 20000),   # dense 96
    (0, 150, 18, 18, 12000),  # dense 160
    (0, 180, 22, 22, 10000),  # dense 192
    (1, 40, 5, 5, 20000),     # factored 48
    (1, 80, 10, 10, 12000),   # factored 96
]


def _systems():
    out = []
    for k, (method, n, m, p, pts) in enumerate(SHAPES):
        out.append(synth.make_system(method, n, m, p, -8.0, -0.05, 5e-4 if n > 64 else 1e-3, pts,
                                     seed=500 + k))
    return out


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ids, e = api.get_available_gpus()
    assert e == 0 and ids, "no GPU visible"
    assert api.initialize_selected_gpus(ids[:1]) == 0
    yield
    api.close_gpu()


def _serve(systems, lanes=4):
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=lanes, device_bytes=int(need * 1.2) + (256 << 20),
                     pinned_bytes=512 << 20, precision=api.MIXED)
    try:
        got = [srv.submit(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h, coef=s.ucoef)
               for s in systems]
        ys = []
        for t, Y in got:
            assert srv.wait(t)[0] == 0
            ys.append(Y)
        gs = srv.group_stats()
    finally:
        srv.close()
    return ys, gs


def test_grouped_variants_meet_the_fp64_bar_and_equal_alone():
    systems = _systems()
    ys, gs = _serve(systems)
    print("group stats", gs)
    assert gs["grouped_requests"] == len(systems)
    assert gs["largest_group"] > 1
    f = lambda a: np.asarray(a, np.float64)
    worst = 0.0
    for s, y in zip(systems, ys):
        y64 = orc.simulate_f64(s.method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0),
                               synth.samples_f64(s), s.steps, s.h)
        err = orc.max_rel_err(y, y64)
        print(f"{'ABM' if s.method else 'RK4'} n={s.n} N={s.steps}: {err:.3e}")
        worst = max(worst, err)
        assert err <= MIXED_BAR, (s.method, s.n, s.steps, err)
    print(f"worst {worst:.3e}")
    # a request alone (a group of one) gives the same bits as in the group
    for k in (1, 6, 9, 13, 15, 20, 23):
        alone, _ = _serve([systems[k]], lanes=1)
        np.testing.assert_array_equal(alone[0], ys[k], err_msg=f"shape {SHAPES[k]}")


def test_device_outputs_equal_host_outputs():
    systems = _systems()[:6] + _systems()[8:12]
    ys, _ = _serve(systems)
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    srv = api.Server(lanes=2, device_bytes=int(need * 1.2) + (256 << 20), pinned_bytes=64 << 20,
                     precision=api.MIXED)
    dev = torch.device("cuda:0")
    outs = [torch.full((s.steps, s.p), float("nan"), dtype=torch.float32, device=dev)
            for s in systems]
    try:
        tickets = [srv.submit_device(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                                     o.data_ptr(), coef=s.ucoef) for s, o in zip(systems, outs)]
        for t in tickets:
            assert srv.wait(t)[0] == 0
    finally:
        srv.close()
    for s, y, o in zip(systems, ys, outs):
        np.testing.assert_array_equal(o.cpu().numpy(), y, err_msg=f"n={s.n} N={s.steps}")


def test_batch_submission_equals_single_submissions():
    """ParODE_ServerSubmitDeviceBatch queues a whole batch in one call: every
    request's outputs equal those of the same request submitted on its own
    (bit for bit), and the tickets come back in order."""
    systems = _systems()[:12]
    need = sum(api.Server.footprint(s.method, s.n, s.m, s.p, s.steps) for s in systems)
    dev = torch.device("cuda:0")
    srv = api.Server(lanes=2, device_bytes=int(need * 1.2) + (256 << 20), pinned_bytes=64 << 20,
                     precision=api.MIXED)
    one = [torch.full((s.steps, s.p), float("nan"), dtype=torch.float32, device=dev)
           for s in systems]
    many = [torch.full_like(o, float("nan")) for o in one]
    try:
        single = [srv.submit_device(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                                    o.data_ptr(), coef=s.ucoef) for s, o in zip(systems, one)]
        for t in single:
            assert srv.wait(t)[0] == 0
        tickets = srv.submit_device_batch([(s.method, s.A, s.B, s.C, s.D, s.x0, s.steps, s.h,
                                            o.data_ptr(), s.ucoef)
                                           for s, o in zip(systems, many)])
        assert tickets == sorted(tickets) and len(set(tickets)) == len(systems)
        st, lat, que = srv.wait_batch(tickets)   # ParODE_ServerWaitBatch
        assert not st.any() and (lat >= que).all() and (que >= 0).all()
        gs = srv.group_stats()
    finally:
        srv.close()
    assert gs["largest_group"] == len(systems)   # the batch formed one group
    for s, a, b in zip(systems, one, many):
        np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy(),
                                      err_msg=f"n={s.n} N={s.steps}")

