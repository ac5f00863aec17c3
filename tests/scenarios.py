"""The reference's own test scenarios, regenerated bit-exactly with glibc rand().

The reference's tests fill their matrices from unseeded rand() (glibc seed 1),
row-major, in the order A, B, C, D (TestsSources/TestRungeKutta4thOrder.cpp:54-80,
TestAdamsMoulton.cpp:54-79, TestParODEInterface.cpp:312-335,
TestParODEStaticLibCInterface.cpp:527-555).  Each scenario starts from srand(1),
as each reference test is its own program.
"""
import ctypes

import numpy as np

_libc = ctypes.CDLL("libc.so.6")
_RAND_MAX_F = np.float32(2147483647.0)   # (fType)RAND_MAX


def _rf():
    return np.float32(np.float32(_libc.rand()) / _RAND_MAX_F)


def _fill(shape):
    a = np.empty(shape, np.float32)
    for r in range(shape[0]):
        for c in range(shape[1]):
            a[r, c] = _rf()
    return a


SIN_COS = np.array([[1, 0, 0, 0, 1, 0], [0, 1, 0, 0, 1, 0]], dtype=np.float64)
SIN_COS_SIN_COS = np.array([[1, 0, 0, 0, 1, 0], [0, 1, 0, 0, 1, 0],
                            [1, 0, 0, 0, 1, 0], [0, 1, 0, 0, 1, 0]], dtype=np.float64)

SIN_COS_SRC = ("fType UFunc(fType t, int nInputIndex){\n"
               "fType fVal = nInputIndex == 0 ? sin(t) :\n"
               "nInputIndex == 1 ? cos(t) :\n"
               "0.0;\n"
               "return fVal;\n"
               "}\n")                                # TestRungeKutta4thOrder.cpp:185-190
SIN_COS_SIN_COS_SRC = ("fType UFunc(fType t, int nInputIndex){\n"
                       "fType fVal = nInputIndex == 0 ? (fType)(sin(t)) :\n"
                       "nInputIndex == 1 ? (fType)(cos(t)) :\n"
                       "nInputIndex == 2 ? (fType)(sin(t)) :\n"
                       "nInputIndex == 3 ? (fType)(cos(t)) :\n"
                       "0.0;\n"
                       "return fVal;\n"
                       "}\n")                        # TestParODEStaticLibCInterface.cpp:506-514


def rk4_class():
    """TestRungeKutta4thOrder.cpp:54-97: n=m=p=2, A=-5I+rand, x0=10, h=0.1, 50 steps."""
    _libc.srand(1)
    A = np.float32(-5.0) * np.eye(2, dtype=np.float32)
    for r in range(2):
        for c in range(2):
            A[r, c] = np.float32(A[r, c] + _rf())
    B, C, D = _fill((2, 2)), _fill((2, 2)), _fill((2, 2))
    return dict(solver=0, A=A, B=B, C=C, D=D, x0=np.full(2, 10.0, np.float32),
                h=0.1, steps=50, ucoef=SIN_COS, usrc=SIN_COS_SRC)


def abm_class():
    """TestAdamsMoulton.cpp:54-97: n=m=p=2, A=-5I+2*rand (double product), x0=10."""
    _libc.srand(1)
    A = np.float32(-5.0) * np.eye(2, dtype=np.float32)
    for r in range(2):
        for c in range(2):
            A[r, c] = np.float32(np.float64(A[r, c]) + 2.0 * np.float64(_rf()))
    B, C, D = _fill((2, 2)), _fill((2, 2)), _fill((2, 2))
    return dict(solver=1, A=A, B=B, C=C, D=D, x0=np.full(2, 10.0, np.float32),
                h=0.1, steps=50, ucoef=SIN_COS, usrc=SIN_COS_SRC)


def _diag_plus_small(n, m, p):
    A = np.float32(-10.0) * np.eye(n, dtype=np.float32)
    for r in range(n):
        for c in range(n):
            A[r, c] = np.float32(A[r, c] + np.float32(np.float32(0.01) * _rf()))
    B, C, D = _fill((n, m)), _fill((p, n)), _fill((p, m))
    return A, B, C, D


def cpp_api():
    """TestParODEInterface.cpp:312-349: n=20, m=p=4, A=-10I+0.01 rand, x0=1, N=1000."""
    _libc.srand(1)
    A, B, C, D = _diag_plus_small(20, 4, 4)
    return dict(A=A, B=B, C=C, D=D, x0=np.ones(20, np.float32), h=0.1, steps=1000,
                ucoef=SIN_COS_SIN_COS, usrc=SIN_COS_SIN_COS_SRC)


def c_abi(steps):
    """TestParODEStaticLibCInterface.cpp:527-604: n=128, m=p=4, x0=0, h=0.1."""
    _libc.srand(1)
    A, B, C, D = _diag_plus_small(128, 4, 4)
    return dict(A=A, B=B, C=C, D=D, x0=np.zeros(128, np.float32), h=0.1, steps=steps,
                ucoef=SIN_COS_SIN_COS, usrc=SIN_COS_SIN_COS_SRC)


C_ABI_STEPS = [100, 250, 500, 1000, 2000, 4000, 8000, 16000, 32000, 64000, 128000]


def uhalf_from_double_sin(ucoef, t0, h, total):
    """Float inputs on the h/2 grid computed as (fType)f(double t), as the
    reference's CPU harness does (TestParODEStaticLibCInterface.cpp:17-24)."""
    s = np.arange(2 * total + 1, dtype=np.float64)
    t = t0 + s * (h / 2.0)
    a, b, c, d, w, phi = (np.asarray(ucoef)[:, i][None, :] for i in range(6))
    g = w * t[:, None] + phi
    return (a * np.sin(g) + b * np.cos(g) + c + d * t[:, None]).astype(np.float32)
