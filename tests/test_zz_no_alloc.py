"""ParODE_SolveNoAlloc (scan buffers in the library's static device scratch
array) against ParODE_PlanExecute (buffers from hipMalloc) through the C ABI:
two calls in a row, the second on a smaller system, each bitwise equal to the
plan's outputs (the same kernels on the same inputs), and a system whose
buffers exceed the scratch array rejected with ParODE_NotEnoughResources."""
import numpy as np
import pytest

import oracle as orc
from parode_amd import api, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ids, e = api.get_available_gpus()
    assert e == 0 and ids, "no GPU visible"
    assert api.initialize_selected_gpus(ids[:1]) == 0
    yield
    api.close_gpu()


def _device_inputs(s, precision):
    dev = torch.device("cuda:0")
    U = synth.samples_f64(s)
    if precision == api.PLAIN32:
        U = U.astype(np.float32)
    t = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in (s.A, s.B, s.C, s.D, s.x0)]
    t.append(torch.from_numpy(U).to(dev))
    return t, U


def _both(s, precision):
    """Outputs of the plan and of the allocation-free entry point on one system."""
    t, U = _device_inputs(s, precision)
    dev = torch.device("cuda:0")
    y_plan = torch.full((s.steps, s.p), np.nan, dtype=torch.float32, device=dev)
    y_free = torch.full((s.steps, s.p), np.nan, dtype=torch.float32, device=dev)
    uD = U.dtype == np.float64
    pl = api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, precision, uD)
    try:
        pl.execute(*(x.data_ptr() for x in t), y_plan.data_ptr())
    finally:
        torch.cuda.synchronize()
        pl.close()
    st = api.solve_no_alloc(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, precision, uD,
                            *(x.data_ptr() for x in t), y_free.data_ptr())
    assert st == 0, api.last_error_message()
    torch.cuda.synchronize()
    return y_plan.cpu().numpy(), y_free.cpu().numpy(), U


@pytest.mark.parametrize("first,second", [
    # (method, n, m, p, steps, precision)
    ((1, 24, 3, 4, 30000, api.MIXED), (0, 8, 2, 2, 3001, api.PLAIN32)),
    ((0, 64, 4, 5, 20000, api.PLAIN32), (1, 6, 2, 3, 517, api.MIXED)),
])
def test_no_alloc_equals_plan_twice_second_smaller(first, second):
    for k, (method, n, m, p, steps, prec) in enumerate((first, second)):
        s = synth.make_system(method, n, m, p, -5.0, -0.1, 1e-3, steps, seed=300 + 7 * k + n)
        need = api.lib().ParODE_PlanEstimateBytes(method, n, m, steps, prec)
        assert 0 < need < api.no_alloc_scratch_bytes()
        y_plan, y_free, U = _both(s, prec)
        assert np.isfinite(y_free).all()
        np.testing.assert_array_equal(y_free, y_plan, err_msg=f"call {k + 1}")
        # and the plan itself meets its bar, so equality is not of two wrong results
        if prec == api.PLAIN32:
            ref = orc.simulate_mirror_exact(method, s.A, s.B, s.C, s.D, s.x0, s.t0, s.t_end, s.h,
                                            Uhalf=U)
            bar = 1e-5
        else:
            f = lambda a: np.asarray(a, np.float64)
            ref = orc.simulate_f64(method, f(s.A), f(s.B), f(s.C), f(s.D), f(s.x0), U, steps, s.h)
            bar = 1e-6
        assert orc.max_rel_err(y_free, ref) <= bar


def test_no_alloc_rejects_systems_beyond_the_scratch_array():
    # ABM n = 64 (scan size 256) over 2^18 points: 256 MiB of trajectory alone
    s = synth.make_system(1, 64, 2, 2, -5.0, -0.1, 1e-3, 1 << 18, seed=5)
    need = api.lib().ParODE_PlanEstimateBytes(1, 64, 2, 1 << 18, api.MIXED)
    assert need > api.no_alloc_scratch_bytes()
    dev = torch.device("cuda:0")
    t = [torch.zeros(8, dtype=torch.float32, device=dev) for _ in range(7)]
    st = api.solve_no_alloc(1, 64, 2, 2, 0, 0, 1 << 18, s.h, api.MIXED, True,
                            *(x.data_ptr() for x in t))
    assert st == -4
    assert "scratch" in api.last_error_message()
