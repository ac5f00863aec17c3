"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: the plain FP32 path within 1e-5 relative (max_k |y_k - y*_k|_inf /
max_k |y*_k|_inf) of the reference's algorithm on the reference's float inputs
(oracle.simulate_mirror_exact); tVect and nSteps bit-exact; the reference's own
sqrt(sum err^2) < 0.1 rule against its CPU path on its own scenarios; the
precision modes against the FP64 oracle with the bound stated in each test."""
import numpy as np
import pytest

import oracle as orc
import scenarios as sc
from parode_amd import api, synth

pytestmark = pytest.mark.gpu

PARITY = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu():
    ids, e = api.get_available_gpus()
    assert e == 0 and ids, "no GPU visible"
    assert api.initialize_selected_gpus(ids[:1]) == 0
    api.set_precision(api.PLAIN32)
    api.set_max_batch_steps(0)
    yield
    api.close_gpu()


def _register(src):
    idx, e = api.register_input(src)
    assert e == 0, api.last_error_message()
    return idx


def _run_ufunc(solver, s, src_coef=None, t_end=None):
    idx = _register(synth.ufunc_source(s.ucoef if src_coef is None else src_coef))
    t1 = s.t_end if t_end is None else t_end
    tv, y, e = api.simulate_ode(solver, s.A, s.B, s.C, s.D, idx, s.t0, t1, s.h, s.x0)
    assert e == 0, api.last_error_message()
    return tv, y


def _check_plain(solver, s, y, tv, t_end=None):
    t1 = s.t_end if t_end is None else t_end
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, s.t0, t1, s.h,
                                   ucoef=s.ucoef)
    assert y.shape == ye.shape
    np.testing.assert_array_equal(tv, orc.time_vector(solver, s.t0, s.h, ye.shape[0]))
    err = orc.max_rel_err(y, ye)
    print(f"plain FP32 vs reference algorithm: max rel err {err:.3e}")
    assert err <= PARITY
    return err


# ----------------------------------------------------------- configurations
def test_config0_rk4_plain_parity():
    s = synth.config(0)                       # RK4 n=8 m=2 p=2 h=1e-3 N=4096
    tv, y = _run_ufunc(0, s)
    assert y.shape == (4096, 2)
    _check_plain(0, s, y, tv)


@pytest.mark.parametrize("n,m,p,N", [(16, 2, 2, 3000), (5, 1, 3, 1025)])
def test_abm_plain_parity(n, m, p, N):
    s = synth.make_system(1, n, m, p, -8.0, -0.1, 1e-3, N, seed=n)
    tv, y = _run_ufunc(1, s)
    assert y.shape == (N, p)
    _check_plain(1, s, y, tv)


@pytest.mark.parametrize("solver", [0, 1])
def test_config2_shape_plain_parity(solver):
    s = synth.make_system(solver, 128 if solver == 0 else 32, 16, 16, -20.0, -0.01, 2e-4, 1 << 13,
                          seed=2, skew=0.1)
    tv, y = _run_ufunc(solver, s)
    _check_plain(solver, s, y, tv)


# -------------------------------------------------- the reference's scenarios
def _reference_rule(solver, d, y, t1, t0=0.0):
    """The reference's acceptance rule, sqrt(sum (y - y_cpu)^2) < 0.1 against
    its CPU path (TestParODEStaticLibCInterface.cpp:204, :466).  At long
    horizons the reference's own float time grid (Kernels/u.cl:8) breaks that
    rule by itself: the FP32-mirror restatement of its arithmetic exceeds 0.1
    at N >= 32000 (ABM) / 64000 (RK4).  There the bar is the mirror's own
    distance, +1%."""
    total = y.shape[0]
    Uh = sc.uhalf_from_double_sin(d["ucoef"], t0, d["h"], total)
    cpu = orc.cpu_rk4_odeint if solver == 0 else orc.cpu_abm_pece
    yc = cpu(d["A"], d["B"], d["C"], d["D"], d["x0"], Uh, total, d["h"])
    lo = 1 if solver == 0 else 4
    l2 = lambda a: float(np.sqrt(np.sum((a[lo:total - 3].astype(np.float64)
                                         - yc[lo:total - 3]) ** 2)))
    tv, ym = orc.simulate_mirror(solver, d["A"], d["B"], d["C"], d["D"], d["x0"], t0, t1, d["h"],
                                 ucoef=d["ucoef"])
    l2_gpu, l2_ref = l2(y), l2(ym)
    print(f"reference L2 rule: GPU {l2_gpu:.3e}, reference arithmetic {l2_ref:.3e}")
    assert l2_gpu < max(0.1, 1.01 * l2_ref)


@pytest.mark.parametrize("make", [sc.rk4_class, sc.abm_class])
def test_reference_class_scenarios(make):
    d = make()
    idx = _register(d["usrc"])
    t1 = (d["steps"] - 1) * d["h"]
    tv, y, e = api.simulate_ode(d["solver"], d["A"], d["B"], d["C"], d["D"], idx, 0.0, t1, d["h"],
                                d["x0"])
    assert e == 0
    ye = orc.simulate_mirror_exact(d["solver"], d["A"], d["B"], d["C"], d["D"], d["x0"], 0.0, t1,
                                   d["h"], ucoef=d["ucoef"])
    assert y.shape == ye.shape
    assert orc.max_rel_err(y, ye) <= PARITY
    _reference_rule(d["solver"], d, y, t1)


@pytest.mark.parametrize("solver", [0, 1])
def test_reference_cpp_api_scenario(solver):
    d = sc.cpp_api()
    idx = _register(d["usrc"])
    t1 = (d["steps"] - 1) * d["h"]
    tv, y, e = api.simulate_ode(solver, d["A"], d["B"], d["C"], d["D"], idx, 0.0, t1, d["h"],
                                d["x0"])
    assert e == 0
    ye = orc.simulate_mirror_exact(solver, d["A"], d["B"], d["C"], d["D"], d["x0"], 0.0, t1,
                                   d["h"], ucoef=d["ucoef"])
    assert orc.max_rel_err(y, ye) <= PARITY
    _reference_rule(solver, d, y, t1)


@pytest.mark.parametrize("steps", [100, 1000, 8000, 32000])
@pytest.mark.parametrize("solver", [0, 1])
def test_reference_c_abi_scenario(solver, steps):
    d = sc.c_abi(steps)
    idx = _register(d["usrc"])
    t1 = (steps - 1) * d["h"]
    tv, y, e = api.simulate_ode(solver, d["A"], d["B"], d["C"], d["D"], idx, 0.0, t1, d["h"],
                                d["x0"])
    assert e == 0
    ye = orc.simulate_mirror_exact(solver, d["A"], d["B"], d["C"], d["D"], d["x0"], 0.0, t1,
                                   d["h"], ucoef=d["ucoef"])
    np.testing.assert_array_equal(tv, orc.time_vector(solver, 0.0, d["h"], ye.shape[0]))
    assert orc.max_rel_err(y, ye) <= PARITY
    _reference_rule(solver, d, y, t1)


# ------------------------------------------------------------- edge cases
def test_single_point_rk4():
    s = synth.make_system(0, 4, 1, 2, -2.0, -1.0, 1e-2, 1, seed=3)
    tv, y = _run_ufunc(0, s)
    assert y.shape == (1, 2)
    _check_plain(0, s, y, tv)


def test_abm_minimum_horizon():
    s = synth.make_system(1, 4, 1, 2, -2.0, -1.0, 1e-2, 4, seed=4)
    tv, y = _run_ufunc(1, s)
    assert y.shape == (4, 2)
    _check_plain(1, s, y, tv)


def test_abm_horizon_shorter_than_startup_is_rejected():
    s = synth.make_system(1, 4, 1, 2, -2.0, -1.0, 1e-2, 3, seed=4)
    idx = _register(synth.ufunc_source(s.ucoef))
    tv, y, e = api.simulate_ode(1, s.A, s.B, s.C, s.D, idx, 0.0, s.t_end, s.h, s.x0)
    assert e == -4 and len(tv) == 0


@pytest.mark.parametrize("solver", [0, 1])
def test_state_output_without_c_and_d(solver):
    s = synth.make_system(solver, 6, 2, 2, -3.0, -0.5, 1e-2, 300, seed=5)
    idx = _register(synth.ufunc_source(s.ucoef))
    tv, y, e = api.simulate_ode(solver, s.A, s.B, None, None, idx, 0.0, s.t_end, s.h, s.x0)
    assert e == 0 and y.shape == (300, 6)
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, None, None, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef)
    assert orc.max_rel_err(y, ye) <= PARITY


@pytest.mark.parametrize("solver", [0, 1])
def test_no_inputs(solver):
    s = synth.make_system(solver, 6, 1, 3, -3.0, -0.5, 1e-2, 257, seed=6)
    tv, y, e = api.simulate_ode(solver, s.A, None, s.C, None, -1, 0.0, s.t_end, s.h, s.x0)
    assert e == 0 and y.shape == (257, 3)
    ye = orc.simulate_mirror_exact(solver, s.A, None, s.C, None, s.x0, 0.0, s.t_end, s.h)
    assert orc.max_rel_err(y, ye) <= PARITY


@pytest.mark.parametrize("solver", [0, 1])
def test_d_only_output(solver):
    s = synth.make_system(solver, 6, 2, 2, -3.0, -0.5, 1e-2, 100, seed=7)
    idx = _register(synth.ufunc_source(s.ucoef))
    tv, y, e = api.simulate_ode(solver, s.A, s.B, None, s.D, idx, 0.0, s.t_end, s.h, s.x0)
    assert e == 0 and y.shape == (100, 2)
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, None, s.D, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef)
    assert orc.max_rel_err(y, ye) <= PARITY


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("batch", [64, 256])
def test_batched_horizon_matches_single_batch(solver, batch):
    """Horizons longer than a batch are chained with the device-side root carry
    (GPUFixedStepIterative.cpp:294-337); the result equals one whole scan."""
    s = synth.make_system(solver, 8, 2, 2, -4.0, -0.2, 5e-3, 1000, seed=8)
    idx = _register(synth.ufunc_source(s.ucoef))
    api.set_max_batch_steps(batch)
    try:
        tv, y, e = api.simulate_ode(solver, s.A, s.B, s.C, s.D, idx, 0.0, s.t_end, s.h, s.x0)
    finally:
        api.set_max_batch_steps(0)
    assert e == 0 and y.shape == (1000, 2)
    # batches restart the float time grid at float(t_batch) (AdamsMoulton.cpp:426,
    # RungeKutta4thOrder.cpp:255): the oracle runs the same batches
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef, nsb=batch)
    assert orc.max_rel_err(y, ye) <= PARITY


def test_invalid_inputs():
    s = synth.make_system(0, 4, 1, 2, -2.0, -1.0, 1e-2, 10, seed=9)
    idx, e = api.register_input("fType UFunc(fType t, int i){ return sin(t) }")
    assert (idx, e) == (-3, -3)
    tv, y, e = api.simulate_ode(0, s.A, s.B, s.C, s.D, 999, 0.0, s.t_end, s.h, s.x0)
    assert e == -6
    good = _register(synth.ufunc_source(s.ucoef))
    tv, y, e = api.simulate_ode(5, s.A, s.B, s.C, s.D, good, 0.0, s.t_end, s.h, s.x0)
    assert e == -7


def test_timer_and_device_name():
    ids, _ = api.get_available_gpus()
    name, e = api.get_device_name(ids[0])
    assert e == 0 and name
    assert api.initialize_selected_gpus(ids[:1]) == 0
    assert api.start_timer() == 0
    t, e = api.stop_timer()
    assert e == 0 and t >= 0.0


# ------------------------------------------------------- sampled / precision
@pytest.mark.parametrize("solver", [0, 1])
def test_sampled_float_inputs_plain_parity(solver):
    s = synth.make_system(solver, 12, 3, 2, -5.0, -0.1, 1e-3, 2000, seed=10)
    U = synth.samples_f64(s).astype(np.float32)
    tv, y, e = api.simulate_ode_sampled(solver, s.A, s.B, s.C, s.D, U, 0.0, s.t_end, s.h, s.x0)
    assert e == 0, api.last_error_message()
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h, Uhalf=U)
    assert orc.max_rel_err(y, ye) <= PARITY


@pytest.mark.parametrize("mode,bound", [(api.FP64, 1e-9), (api.MIXED, 1e-6)])
@pytest.mark.parametrize("solver", [0, 1])
def test_precision_modes_against_fp64(solver, mode, bound):
    s = synth.make_system(solver, 16, 2, 2, -10.0, -0.05, 5e-4, 1 << 14, seed=11,
                          noise_sigma=0.1)
    U = synth.samples_f64(s)
    api.set_precision(mode)
    try:
        tv, y, e = api.simulate_ode_sampled(solver, s.A, s.B, s.C, s.D, U, 0.0, s.t_end, s.h,
                                            s.x0)
    finally:
        api.set_precision(api.PLAIN32)
    assert e == 0, api.last_error_message()
    y64 = orc.simulate_f64(solver, s.A.astype(np.float64), s.B.astype(np.float64),
                           s.C.astype(np.float64), s.D.astype(np.float64),
                           s.x0.astype(np.float64), U, s.steps, s.h)
    err = orc.max_rel_err(y, y64)
    print(f"mode {mode} solver {solver}: max rel err vs FP64 {err:.3e}")
    # FP64 mode: the only FP32 rounding is the output cast (y is fType)
    assert err <= (max(bound, 1e-7) if mode == api.FP64 else bound)


@pytest.mark.parametrize("mode,bound", [(api.FP64, 1e-7), (api.MIXED, 1e-6)])
@pytest.mark.parametrize("solver,n,m,steps,f32", [(1, 17, 1, 1000, False), (1, 40, 3, 3001, True),
                                                  (0, 5, 7, 999, False), (0, 70, 5, 2049, True),
                                                  (1, 64, 8, 4097, False),
                                                  (1, 100, 4, 5000, False),
                                                  (0, 128, 1, 3000, False),
                                                  (1, 48, 33, 2500, True),
                                                  (0, 200, 30, 1500, False)])
def test_fp64_leaves_shapes(solver, n, m, steps, f32, mode, bound):
    """The FP64 leaves run on the matrix cores in 16 x 16 tiles: ragged row
    tiles (n not a multiple of 16), input counts that do not fill a 4-column
    k-step, one input, staged samples too large for the factored kernel (the
    vector-ALU fallback), ragged horizons, FP32 and FP64 samples, against the
    FP64 oracle."""
    s = synth.make_system(solver, n, m, 2, -6.0, -0.05, 1e-3, steps, seed=n + m)
    U = synth.samples_f64(s)
    if f32:
        U = U.astype(np.float32)
    api.set_precision(mode)
    try:
        tv, y, e = api.simulate_ode_sampled(solver, s.A, s.B, s.C, s.D, U, 0.0, s.t_end, s.h,
                                            s.x0)
    finally:
        api.set_precision(api.PLAIN32)
    assert e == 0, api.last_error_message()
    y64 = orc.simulate_f64(solver, s.A.astype(np.float64), s.B.astype(np.float64),
                           s.C.astype(np.float64), s.D.astype(np.float64),
                           s.x0.astype(np.float64), U.astype(np.float64), s.steps, s.h)
    err = orc.max_rel_err(y, y64)
    print(f"mode {mode} solver {solver} n {n} m {m}: max rel err vs FP64 {err:.3e}")
    assert err <= bound


@pytest.mark.parametrize("mode,bound", [(api.FP64, 1e-7), (api.MIXED, 1e-6)])
@pytest.mark.parametrize("solver,n,p,zd", [(0, 200, 64, False), (1, 130, 40, False),
                                           (0, 90, 24, False), (1, 70, 17, False),
                                           (0, 300, 5, False), (0, 96, 33, True),
                                           (1, 64, 8, True)])
def test_wide_state_outputs_against_fp64(solver, n, p, zd, mode, bound):
    """Outputs of wide FP32 states (n > 64) and of FP64 states run as tiles of
    64 elements x 16, 32 or 64 outputs on the FP64 matrix cores, K over the
    states then the inputs (only the states when D is zero): each tile width,
    ragged output counts and K slices, against the FP64 oracle."""
    s = synth.make_system(solver, n, 3, p, -6.0, -0.2, 1e-3, 1500, seed=900 + n)
    if zd:
        s.D[:] = 0.0   # the oracle's D; the solver gets none (ZeroD)
    U = synth.samples_f64(s)
    api.set_precision(mode)
    try:
        tv, y, e = api.simulate_ode_sampled(solver, s.A, s.B, s.C, None if zd else s.D, U, 0.0,
                                            s.t_end, s.h, s.x0)
    finally:
        api.set_precision(api.PLAIN32)
    assert e == 0, api.last_error_message()
    assert y.shape == (1500, p)
    y64 = orc.simulate_f64(solver, s.A.astype(np.float64), s.B.astype(np.float64),
                           s.C.astype(np.float64), s.D.astype(np.float64),
                           s.x0.astype(np.float64), U, s.steps, s.h)
    err = orc.max_rel_err(y, y64)
    print(f"mode {mode} solver {solver} n {n} p {p}: max rel err vs FP64 {err:.3e}")
    assert err <= bound


@pytest.mark.parametrize("n,steps", [(150, 5000), (200, 65541), (256, 3000)])
def test_large_abm_chunk_recurrences(n, steps):
    """ABM systems with 128 < n <= 256 in mixed precision run the lowest
    levels as chunk recurrences whose step is two dependent products
    (x3 + A (w1 + A w2) + v); padded widths 160, 224 and 256, chunk lengths
    8, 128 and 4, against the FP64 oracle at the mixed-precision bar."""
    s = synth.make_system(1, n, 4, 3, -6.0, -0.05, 1e-3, steps, seed=500 + n)
    plan = api.Plan(1, n, 4, 3, 0, 0, steps, s.h, api.MIXED, True)
    try:
        lc = plan.chunk_log
    finally:
        plan.close()
    assert lc == (7 if steps > 65536 else 2 if steps < 4096 else 3)
    U = synth.samples_f64(s)
    api.set_precision(api.MIXED)
    try:
        tv, y, e = api.simulate_ode_sampled(1, s.A, s.B, s.C, s.D, U, 0.0, s.t_end, s.h, s.x0)
    finally:
        api.set_precision(api.PLAIN32)
    assert e == 0, api.last_error_message()
    y64 = orc.simulate_f64(1, s.A.astype(np.float64), s.B.astype(np.float64),
                           s.C.astype(np.float64), s.D.astype(np.float64),
                           s.x0.astype(np.float64), U, s.steps, s.h)
    err = orc.max_rel_err(y, y64)
    print(f"n {n} N {steps} lc {lc}: max rel err vs FP64 {err:.3e}")
    assert err <= 1e-6


# ------------------------------------------- matrix-core vs vector kernels
@pytest.mark.parametrize("solver,n", [(0, 20), (0, 48), (0, 100), (0, 256), (0, 300), (1, 17),
                                      (1, 40), (1, 64), (1, 130), (1, 256)])
def test_mfma_levels_match_vector_levels(solver, n):
    """The FP32 level update on MFMA (17 <= n_a <= 1024; rows split over
    work-groups above 256; ABM level 0 on the leaves' last block) and on the
    vector ALU give the same trajectory to FP32 rounding, and both meet the
    parity bar."""
    N = 2048 if n < 200 else 512
    s = synth.make_system(solver, n, 3, 3, -6.0, -0.2, 1e-3, N, seed=100 + n)
    idx = _register(synth.ufunc_source(s.ucoef))
    outs = []
    for variant in (api.VECTOR_LEVELS | api.VECTOR_GEMM | api.NO_FUSED_TOP, 0):
        api.set_level_kernel(variant)
        try:
            tv, y, e = api.simulate_ode(solver, s.A, s.B, s.C, s.D, idx, 0.0, s.t_end, s.h, s.x0)
        finally:
            api.set_level_kernel(0)
        assert e == 0, api.last_error_message()
        outs.append(y)
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef)
    e_vec, e_mfma = orc.max_rel_err(outs[0], ye), orc.max_rel_err(outs[1], ye)
    print(f"n_a={s.n_scan}: vector {e_vec:.2e}, mfma {e_mfma:.2e}")
    assert orc.max_rel_err(outs[1], outs[0]) <= PARITY
    assert e_mfma <= PARITY and e_vec <= PARITY


@pytest.mark.parametrize("variant", [1, 2, 4, 3, 5, 6, 8])
@pytest.mark.parametrize("solver", [0, 1])
def test_each_kernel_choice_meets_parity(solver, variant):
    """Every combination of the kernel switches (vector/MFMA levels, vector/MFMA
    FP64 GEMM, fused/unfused top of the tree) meets the parity bar."""
    s = synth.make_system(solver, 40, 3, 3, -6.0, -0.2, 1e-3, 3000, seed=7 + variant, skew=0.2)
    idx = _register(synth.ufunc_source(s.ucoef))
    api.set_level_kernel(variant)
    try:
        tv, y, e = api.simulate_ode(solver, s.A, s.B, s.C, s.D, idx, 0.0, s.t_end, s.h, s.x0)
    finally:
        api.set_level_kernel(0)
    assert e == 0, api.last_error_message()
    ye = orc.simulate_mirror_exact(solver, s.A, s.B, s.C, s.D, s.x0, 0.0, s.t_end, s.h,
                                   ucoef=s.ucoef)
    err = orc.max_rel_err(y, ye)
    print(f"variant {variant}: {err:.2e}")
    assert err <= PARITY


@pytest.mark.parametrize("solver,n,p", [(0, 200, 64), (1, 40, 96)])
def test_large_output_matrices(solver, n, p):
    """Output matrices too large to stage in LDS take the global-memory
    output kernel; parity as everywhere."""
    s = synth.make_system(solver, n, 3, p, -6.0, -0.2, 1e-3, 700, seed=300 + n)
    tv, y = _run_ufunc(solver, s)
    assert y.shape == (700, p)
    _check_plain(solver, s, y, tv)
