"""ctypes front end of the CPU oracle (liboracle.so built from parode_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py.  The product path never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_c_int, _c_long, _c_double = ctypes.c_int, ctypes.c_long, ctypes.c_double
_opt_f32 = ctypes.c_void_p


def build():
    """Compile liboracle.so with the recipe in oracle/Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "or_rk4_build_mirror": [_c_int, _c_int, _f32p, _f32p, _c_double, _f32p, _f32p],
        "or_abm_build_mirror": [_c_int, _c_int, _f32p, _f32p, _c_double, _f32p, _f32p],
        "or_rk4_build_f64": [_c_int, _c_int, _f64p, _f64p, _c_double, _f64p, _f64p],
        "or_abm_build_f64": [_c_int, _c_int, _f64p, _f64p, _c_double, _f64p, _f64p],
        "or_time_grid_f32": [_c_double, _c_double, _c_long, _c_long, _f32p],
        "or_ufunc_eval_f32": [_c_int, _f32p, _c_long, _f32p, _f32p],
        "or_ufunc_eval_f64": [_c_int, _f64p, _c_double, _c_double, _c_long, _c_long,
                              _c_double, ctypes.c_uint64, _f64p],
        "or_rk4_leaves_f32": [_c_int, _c_int, _f32p, _c_long, _f32p, _f32p],
        "or_abm_leaves_f32": [_c_int, _c_int, _f32p, _c_long, _f32p, _f32p],
        "or_rk4_leaves_f64": [_c_int, _c_int, _f64p, _c_long, _f64p, _f64p],
        "or_abm_leaves_f64": [_c_int, _c_int, _f64p, _c_long, _f64p, _f64p],
        "or_recurrence_f32": [_c_int, _f32p, _c_long, _f32p, _f32p, _f32p],
        "or_recurrence_f64": [_c_int, _f64p, _c_long, _f64p, _f64p, _f64p],
        "or_blelloch_mirror_f32": [_c_int, _c_long, _f32p, _f32p, _f32p, _c_int, _f32p],
        "or_abm_startup_f32": [_c_int, _c_int, _f32p, _f32p, _f32p, _f32p, _c_double, _f32p],
        "or_abm_startup_f64": [_c_int, _c_int, _f64p, _f64p, _f64p, _f64p, _c_double, _f64p],
        "or_step_count": [_c_int, _c_double, _c_double, _c_double],
        "or_time_vector": [_c_int, _c_double, _c_double, _c_long, _f32p],
        "or_simulate_mirror": [_c_int, _c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p,
                               _c_int, _c_int, _opt_f32, _opt_f32, _c_double, _c_double,
                               _c_double, _f32p, _f32p, _f32p],
        "or_simulate_mirror_exact": [_c_int, _c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p,
                                     _c_int, _c_int, _opt_f32, _opt_f32, _c_double, _c_double,
                                     _c_double, _f32p, _f64p],
        "or_simulate_mirror_nsb": [_c_int, _c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p,
                                   _c_int, _c_int, _opt_f32, _opt_f32, _c_double, _c_double,
                                   _c_double, _f32p, _c_long, _f32p, _f32p],
        "or_simulate_mirror_exact_nsb": [_c_int, _c_int, _c_int, _c_int, _f32p, _f32p, _f32p,
                                         _f32p, _c_int, _c_int, _opt_f32, _opt_f32, _c_double,
                                         _c_double, _c_double, _f32p, _c_long, _f64p],
        "or_simulate_f64": [_c_int, _c_int, _c_int, _c_int, _f64p, _f64p, _f64p, _f64p,
                            _c_int, _c_int, _f64p, _c_long, _c_double, _f64p, _f64p, _opt_f32],
        "or_cpu_rk4_odeint": [_c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p, _c_int, _c_int,
                              _f32p, _c_long, _c_double, _f32p, _f32p],
        "or_cpu_abm_pece": [_c_int, _c_int, _c_int, _f32p, _f32p, _f32p, _f32p, _c_int, _c_int,
                            _f32p, _c_long, _c_double, _f32p, _f32p],
        "or_gauss_noise": [ctypes.c_uint64, _c_long, _c_int],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = None
    lib.or_step_count.restype = _c_long
    lib.or_simulate_mirror.restype = _c_long
    lib.or_simulate_f64.restype = _c_long
    lib.or_simulate_mirror_exact.restype = _c_long
    lib.or_simulate_mirror_nsb.restype = _c_long
    lib.or_simulate_mirror_exact_nsb.restype = _c_long
    lib.or_cpu_rk4_odeint.restype = _c_long
    lib.or_cpu_abm_pece.restype = _c_long
    lib.or_gauss_noise.restype = _c_double
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _nz(a, dtype):
    """Non-empty contiguous buffer (a 1-element dummy for absent matrices)."""
    a = np.ascontiguousarray(a if a is not None else np.zeros(1), dtype=dtype)
    return a if a.size else np.zeros(1, dtype=dtype)


# -------------------------------------------------------------------- builders
def rk4_build_mirror(A, B, h):
    A, B = _f32(A), _f32(B)
    n, m = A.shape[0], (B.shape[1] if B.ndim == 2 else 0)
    M = np.empty((n, n), np.float32)
    Bb = np.empty((n, 3 * max(m, 1)), np.float32)
    lib().or_rk4_build_mirror(n, m, A, _nz(B, np.float32), h, M, Bb)
    return M, Bb[:, :3 * m]


def abm_build_mirror(A, B, h):
    A, B = _f32(A), _f32(B)
    n, m = A.shape[0], (B.shape[1] if B.ndim == 2 else 0)
    M = np.empty((4 * n, 4 * n), np.float32)
    Bb = np.empty((n, 5 * max(m, 1)), np.float32)
    lib().or_abm_build_mirror(n, m, A, _nz(B, np.float32), h, M, Bb)
    return M, Bb[:, :5 * m]


def rk4_build_f64(A, B, h):
    A, B = _f64(A), _f64(B)
    n, m = A.shape[0], (B.shape[1] if B.ndim == 2 else 0)
    M = np.empty((n, n))
    Bb = np.empty((n, 3 * max(m, 1)))
    lib().or_rk4_build_f64(n, m, A, _nz(B, np.float64), h, M, Bb)
    return M, Bb[:, :3 * m]


def abm_build_f64(A, B, h):
    A, B = _f64(A), _f64(B)
    n, m = A.shape[0], (B.shape[1] if B.ndim == 2 else 0)
    M = np.empty((4 * n, 4 * n))
    Bb = np.empty((n, 5 * max(m, 1)))
    lib().or_abm_build_f64(n, m, A, _nz(B, np.float64), h, M, Bb)
    return M, Bb[:, :5 * m]


# ---------------------------------------------------------------------- inputs
def time_grid_f32(t0, dt, count, s0=0):
    t = np.empty(count, np.float32)
    lib().or_time_grid_f32(t0, dt, s0, count, t)
    return t


def ufunc_eval_f32(coef, t):
    coef = _f32(coef)
    t = _f32(t)
    m = coef.shape[0]
    U = np.empty((t.size, m), np.float32)
    lib().or_ufunc_eval_f32(m, coef, t.size, t, U)
    return U


def ufunc_eval_f64(coef, t0, dt, S, s0=0, sigma=0.0, seed=0):
    coef = _f64(coef)
    m = coef.shape[0]
    U = np.empty((S, m))
    lib().or_ufunc_eval_f64(m, coef, t0, dt, s0, S, sigma, seed, U)
    return U


def gauss_noise(seed, s, j):
    return lib().or_gauss_noise(seed, s, j)


# ---------------------------------------------------------------- scan pieces
def rk4_leaves_f32(Bb, N, U):
    Bb, U = _f32(Bb), _f32(U)
    n, m = Bb.shape[0], Bb.shape[1] // 3
    V = np.empty((N, n), np.float32)
    lib().or_rk4_leaves_f32(n, m, Bb, N, U, V)
    return V


def abm_leaves_f32(Bb, N, U):
    Bb, U = _f32(Bb), _f32(U)
    n, m = Bb.shape[0], Bb.shape[1] // 5
    V = np.empty((N, 4 * n), np.float32)
    lib().or_abm_leaves_f32(n, m, Bb, N, U, V)
    return V


def recurrence_f32(M, V, x0):
    M, V, x0 = _f32(M), _f32(V), _f32(x0)
    na, N = M.shape[0], V.shape[0]
    X = np.empty((N, na), np.float32)
    lib().or_recurrence_f32(na, M, N, V, x0, X)
    return X


def recurrence_f64(M, V, x0):
    M, V, x0 = _f64(M), _f64(V), _f64(x0)
    na, N = M.shape[0], V.shape[0]
    X = np.empty((N, na))
    lib().or_recurrence_f64(na, M, N, V, x0, X)
    return X


def blelloch_mirror_f32(M, V, x0, cw):
    M, x0 = _f32(M), _f32(x0)
    V = np.array(V, dtype=np.float32, order="C", copy=True)
    na, N = M.shape[0], V.shape[0]
    tot = np.empty(na, np.float32)
    lib().or_blelloch_mirror_f32(na, N, M, V, x0, cw, tot)
    return V, tot


def abm_startup_f32(A, B, x0, Uh, h):
    A, B, x0, Uh = _f32(A), _f32(B), _f32(x0), _f32(Uh)
    n, m = A.shape[0], (B.shape[1] if B.ndim == 2 else 0)
    X0 = np.empty(4 * n, np.float32)
    lib().or_abm_startup_f32(n, m, A, _nz(B, np.float32), x0, _nz(Uh, np.float32), h, X0)
    return X0


def abm_startup_f64(A, B, x0, Uh, h):
    A, B, x0, Uh = _f64(A), _f64(B), _f64(x0), _f64(Uh)
    n, m = A.shape[0], (B.shape[1] if B.ndim == 2 else 0)
    X0 = np.empty(4 * n)
    lib().or_abm_startup_f64(n, m, A, _nz(B, np.float64), x0, _nz(Uh, np.float64), h, X0)
    return X0


def step_count(solver, t0, t1, h):
    return lib().or_step_count(solver, t0, t1, h)


def time_vector(solver, t0, h, total):
    tv = np.empty(max(total, 1), np.float32)
    lib().or_time_vector(solver, t0, h, total, tv)
    return tv[:total]


# ------------------------------------------------------------- whole solvers
def simulate_mirror(solver, A, B, C, D, x0, t0, t1, h, ucoef=None, Uhalf=None, nsb=0):
    """FP32-mirror restatement of the reference's SimulateODE, in batches of
    nsb elements (0: one batch; the input grid restarts per batch,
    GPUFixedStepIterative.cpp:191-234).

    Returns (tVect, xVect[total, nOut])."""
    A = _f32(A)
    n = A.shape[0]
    m = 0 if B is None or np.size(B) == 0 else np.shape(B)[1]
    zC = int(C is None or np.size(C) == 0)
    zD = int(D is None or np.size(D) == 0)
    p = 0 if zC and zD else (np.shape(C)[0] if not zC else np.shape(D)[0])
    nOut = n if (zC and zD) else p
    total = step_count(solver, t0, t1, h)
    if total < 1:
        raise ValueError("empty horizon")
    tv = np.empty(total, np.float32)
    xv = np.empty((total, nOut), np.float32)
    uc = _f32(ucoef) if ucoef is not None else None
    uh = _f32(Uhalf) if Uhalf is not None else None
    got = lib().or_simulate_mirror_nsb(
        solver, n, m, p, A, _nz(B, np.float32), _nz(C, np.float32), _nz(D, np.float32), zC, zD,
        uc.ctypes.data if uc is not None else None, uh.ctypes.data if uh is not None else None,
        t0, t1, h, _f32(x0), int(nsb), tv, xv)
    assert got == total
    return tv, xv


def simulate_mirror_exact(solver, A, B, C, D, x0, t0, t1, h, ucoef=None, Uhalf=None, nsb=0):
    """The reference's algorithm in exact arithmetic on the reference's float
    inputs (FP32-mirror M, Bbar, samples and start-up states; FP64 scan), in
    batches of nsb elements (0: one batch; see simulate_mirror).
    Returns xVect[total, nOut] in FP64."""
    A = _f32(A)
    n = A.shape[0]
    m = 0 if B is None or np.size(B) == 0 else np.shape(B)[1]
    zC = int(C is None or np.size(C) == 0)
    zD = int(D is None or np.size(D) == 0)
    p = 0 if zC and zD else (np.shape(C)[0] if not zC else np.shape(D)[0])
    nOut = n if (zC and zD) else p
    total = step_count(solver, t0, t1, h)
    if total < 1:
        raise ValueError("empty horizon")
    y = np.empty((total, nOut))
    uc = _f32(ucoef) if ucoef is not None else None
    uh = _f32(Uhalf) if Uhalf is not None else None
    got = lib().or_simulate_mirror_exact_nsb(
        solver, n, m, p, A, _nz(B, np.float32), _nz(C, np.float32), _nz(D, np.float32), zC, zD,
        uc.ctypes.data if uc is not None else None, uh.ctypes.data if uh is not None else None,
        t0, t1, h, _f32(x0), int(nsb), y)
    assert got == total
    return y


def simulate_f64(solver, A, B, C, D, x0, Uhalf, total, h, want_states=False):
    """FP64 precision reference.  Uhalf: FP64 samples u(t0 + s h/2), 2*total+1 rows.

    Returns y[total, nOut] (and the states when asked)."""
    A = _f64(A)
    n = A.shape[0]
    m = 0 if B is None or np.size(B) == 0 else np.shape(B)[1]
    zC = int(C is None or np.size(C) == 0)
    zD = int(D is None or np.size(D) == 0)
    p = 0 if zC and zD else (np.shape(C)[0] if not zC else np.shape(D)[0])
    nOut = n if (zC and zD) else p
    y = np.empty((total, nOut))
    nst = total if solver == 0 else total
    X = np.empty((nst, n)) if want_states else None
    got = lib().or_simulate_f64(solver, n, m, p, A, _nz(B, np.float64), _nz(C, np.float64),
                                _nz(D, np.float64), zC, zD, _nz(Uhalf, np.float64), total, h,
                                _f64(x0), y, X.ctypes.data if X is not None else None)
    assert got == total
    return (y, X) if want_states else y


def cpu_rk4_odeint(A, B, C, D, x0, Uhalf, total, h):
    return _cpu("or_cpu_rk4_odeint", A, B, C, D, x0, Uhalf, total, h)


def cpu_abm_pece(A, B, C, D, x0, Uhalf, total, h):
    return _cpu("or_cpu_abm_pece", A, B, C, D, x0, Uhalf, total, h)


def _cpu(name, A, B, C, D, x0, Uhalf, total, h):
    A = _f32(A)
    n = A.shape[0]
    m = 0 if B is None or np.size(B) == 0 else np.shape(B)[1]
    zC = int(C is None or np.size(C) == 0)
    zD = int(D is None or np.size(D) == 0)
    p = 0 if zC and zD else (np.shape(C)[0] if not zC else np.shape(D)[0])
    nOut = n if (zC and zD) else p
    y = np.empty((total, nOut), np.float32)
    getattr(lib(), name)(n, m, p, A, _nz(B, np.float32), _nz(C, np.float32), _nz(D, np.float32),
                         zC, zD, _nz(Uhalf, np.float32), total, h, _f32(x0), y)
    return y


def max_rel_err(y, yref):
    """max_k ||y_k - y*_k||_inf / max_k ||y*_k||_inf."""
    y = np.asarray(y, np.float64)
    yref = np.asarray(yref, np.float64)
    den = np.max(np.abs(yref))
    return float(np.max(np.abs(y - yref)) / (den if den > 0 else 1.0))
