"""Device-resident plans (ParODEExt.h) against the C ABI path and the oracle."""
import numpy as np
import pytest

import oracle as orc
from parode_amd import api, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a, dtype=None):
    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return torch.from_numpy(a).to("cuda:0")


def _run_plan(s, U, precision, p=None):
    zc = zd = 0
    plan = api.Plan(s.method, s.n, s.m, s.p, zc, zd, s.steps, s.h, precision,
                    U.dtype == np.float64)
    tA, tB, tC, tD, tx = (_dev(v) for v in (s.A, s.B, s.C, s.D, s.x0))
    tU = _dev(U)
    tY = torch.zeros((s.steps, s.p), dtype=torch.float32, device="cuda:0")
    plan.execute(tA.data_ptr(), tB.data_ptr(), tC.data_ptr(), tD.data_ptr(), tx.data_ptr(),
                 tU.data_ptr(), tY.data_ptr(), 0)
    torch.cuda.synchronize()
    y = tY.cpu().numpy()
    plan.close()
    return y


@pytest.mark.parametrize("solver", [0, 1])
def test_plan_plain_matches_oracle(solver):
    s = synth.make_system(solver, 12, 3, 4, -5.0, -0.1, 1e-3, 3000, seed=21)
    U = synth.samples_f64(s).astype(np.float32)
    y = _run_plan(s, U, api.PLAIN32)
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h, Uhalf=U)
    assert orc.max_rel_err(y, ye) <= 1e-5


@pytest.mark.parametrize("solver", [0, 1])
def test_plan_equals_c_abi_sampled_path(solver):
    ids, e = api.get_available_gpus()
    assert api.initialize_selected_gpus(ids[:1]) == 0
    s = synth.make_system(solver, 10, 2, 2, -5.0, -0.1, 1e-3, 1500, seed=22)
    U = synth.samples_f64(s).astype(np.float32)
    y = _run_plan(s, U, api.PLAIN32)
    tv, yc, e = api.simulate_ode_sampled(solver, s.A, s.B, s.C, s.D, U, 0.0, s.t_end, s.h, s.x0)
    assert e == 0
    # same scan kernels in the same order: identical scan outputs; the ABM
    # start-up outputs are formed on the device (FP64 sums) by the plan and on
    # the host (float sums, GPUFixedStepIterative.cpp:161-171) by the C ABI path
    lead = 3 if solver == 1 else 0
    np.testing.assert_array_equal(y[lead:], yc[lead:])
    assert orc.max_rel_err(y[:lead], yc[:lead]) <= 1e-6 if lead else True
    api.close_gpu()


@pytest.mark.parametrize("precision", [api.MIXED, api.FP64])
def test_plan_precision_modes(precision):
    s = synth.make_system(1, 16, 2, 2, -10.0, -0.05, 5e-4, 1 << 13, seed=23, noise_sigma=0.1)
    U = synth.samples_f64(s)
    y = _run_plan(s, U, precision)
    y64 = orc.simulate_f64(1, s.A.astype(np.float64), s.B.astype(np.float64),
                           s.C.astype(np.float64), s.D.astype(np.float64),
                           s.x0.astype(np.float64), U, s.steps, s.h)
    err = orc.max_rel_err(y, y64)
    print(f"precision {precision}: max rel err vs FP64 {err:.3e}")
    assert err <= (1e-7 if precision == api.FP64 else 1e-6)


def test_plan_repeated_execution_is_deterministic():
    s = synth.make_system(0, 32, 4, 4, -5.0, -0.1, 1e-3, 4096, seed=24)
    U = synth.samples_f64(s).astype(np.float32)
    y1 = _run_plan(s, U, api.PLAIN32)
    y2 = _run_plan(s, U, api.PLAIN32)
    np.testing.assert_array_equal(y1, y2)


def test_plan_graph_replay_equals_eager():
    """A captured solve replayed several times, with eager work between the
    replays, gives the eager result every time."""
    s = synth.make_system(1, 12, 2, 3, -5.0, -0.1, 1e-3, 5000, seed=25, noise_sigma=0.1)
    U = synth.samples_f64(s)
    plan = api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, api.MIXED, True)
    tA, tB, tC, tD, tx = (_dev(v) for v in (s.A, s.B, s.C, s.D, s.x0))
    tU = _dev(U)
    tY = torch.zeros((s.steps, s.p), dtype=torch.float32, device="cuda:0")
    ptrs = (tA.data_ptr(), tB.data_ptr(), tC.data_ptr(), tD.data_ptr(), tx.data_ptr(),
            tU.data_ptr(), tY.data_ptr())
    st = torch.cuda.Stream()
    plan.execute(*ptrs, st.cuda_stream)
    torch.cuda.synchronize()
    y_eager = tY.clone()
    plan.capture(*ptrs, st.cuda_stream)
    torch.cuda.synchronize()
    for r in range(3):
        tY.zero_()
        torch.cuda.synchronize()
        plan.replay(st.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(tY, y_eager), f"replay {r}"
        plan.execute(*ptrs, st.cuda_stream)      # eager work between replays
        torch.cuda.synchronize()
    plan.close()


@pytest.mark.parametrize("solver,n,precision", [(1, 40, api.MIXED), (0, 70, api.PLAIN32),
                                                (0, 24, api.FP64)])
def test_autotuned_schedule_meets_the_bars(solver, n, precision):
    """The autotuner times each level's kernels on the plan's own buffers and
    the later solves follow its choices; results stay within the bars (the
    choices change only the order of FP32 sums)."""
    s = synth.make_system(solver, n, 3, 2, -6.0, -0.05, 1e-3, 20000, seed=n)
    U = synth.samples_f64(s)
    if precision == api.PLAIN32:
        U = U.astype(np.float32)
    plan = api.Plan(s.method, s.n, s.m, s.p, 0, 0, s.steps, s.h, precision,
                    U.dtype == np.float64)
    tA, tB, tC, tD, tx = (_dev(v) for v in (s.A, s.B, s.C, s.D, s.x0))
    tU = _dev(U)
    tY = torch.zeros((s.steps, s.p), dtype=torch.float32, device="cuda:0")
    args = (tA.data_ptr(), tB.data_ptr(), tC.data_ptr(), tD.data_ptr(), tx.data_ptr(),
            tU.data_ptr(), tY.data_ptr(), 0)
    try:
        plan.execute(*args)
        torch.cuda.synchronize()
        y0 = tY.cpu().numpy()
        tuned = plan.autotune()
        assert tuned > 0
        tY.zero_()
        plan.execute(*args)
        torch.cuda.synchronize()
        y1 = tY.cpu().numpy()
    finally:
        plan.close()
        api.schedule_clear()
    if precision == api.PLAIN32:
        ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h,
                                       Uhalf=U)
        bound = 1e-5
    else:
        ye = orc.simulate_f64(solver, s.A.astype(np.float64), s.B.astype(np.float64),
                              s.C.astype(np.float64), s.D.astype(np.float64),
                              s.x0.astype(np.float64), U, s.steps, s.h)
        bound = 1e-6 if precision == api.MIXED else 1e-7
    print(f"levels tuned {tuned}: err before {orc.max_rel_err(y0, ye):.3e} "
          f"after {orc.max_rel_err(y1, ye):.3e}")
    assert orc.max_rel_err(y0, ye) <= bound
    assert orc.max_rel_err(y1, ye) <= bound
