"""Writes the golden fixtures of this folder from the CPU oracle.

Each fixture is one small system (inputs and expected outputs) in a .npz file
(plain arrays, no pickles):
  A, B, C, D, x0 : float32 system, row-major
  U              : float32 input samples on the h/2 grid, 2*points+1 rows
  h, solver      : step and method (0 RK4, 1 ABM)
  y_exact        : reference arithmetic on these float inputs, FP64 scan
  y_f64          : FP64 sequential recurrence (inputs promoted to double)
  y_mirror       : FP32 mirror of the reference's per-element scan
  t              : the reference's float time vector

Run:  python tests/golden/make_golden.py   (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as orc  # noqa: E402
from parode_amd import synth  # noqa: E402

CASES = {
    "rk4_n8_m2_p2_N256": dict(method=0, n=8, m=2, p=2, lam=(-5.0, -0.1), h=1e-3, steps=256, seed=100),
    "abm_n8_m2_p2_N256": dict(method=1, n=8, m=2, p=2, lam=(-5.0, -0.1), h=1e-3, steps=256, seed=101),
    "rk4_n5_m3_p1_N37": dict(method=0, n=5, m=3, p=1, lam=(-20.0, -1.0), h=2e-2, steps=37, seed=102),
    "abm_n3_m1_p2_N9": dict(method=1, n=3, m=1, p=2, lam=(-3.0, -0.5), h=5e-2, steps=9, seed=103),
}


def build(case):
    c = CASES[case]
    s = synth.make_system(c["method"], c["n"], c["m"], c["p"], c["lam"][0], c["lam"][1], c["h"],
                          c["steps"], seed=c["seed"])
    U = synth.samples_f64(s).astype(np.float32)
    t1 = s.t_end
    y_exact = orc.simulate_mirror_exact(s.method, s.A, s.B, s.C, s.D, s.x0, 0.0, t1, s.h, Uhalf=U)
    t, y_mirror = orc.simulate_mirror(s.method, s.A, s.B, s.C, s.D, s.x0, 0.0, t1, s.h, Uhalf=U)
    f64 = lambda a: a.astype(np.float64)
    y_f64 = orc.simulate_f64(s.method, f64(s.A), f64(s.B), f64(s.C), f64(s.D), f64(s.x0),
                             f64(U), s.steps, s.h)
    return dict(A=s.A, B=s.B, C=s.C, D=s.D, x0=s.x0, U=U, h=np.float64(s.h),
                solver=np.int32(s.method), y_exact=y_exact, y_f64=y_f64, y_mirror=y_mirror, t=t)


def main():
    for case in CASES:
        path = os.path.join(HERE, case + ".npz")
        np.savez_compressed(path, **build(case))
        print("wrote", path)


if __name__ == "__main__":
    main()
