"""The C++ API (include/ParODECppInterface.h, matrixf/vectorf) driven by a
compiled program, as the reference's TestsSources/TestParODEInterface.cpp
drives it, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
import scenarios as sc

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "cpp_api_check")


def _run(solver, d, steps):
    n, m = d["B"].shape
    p = d["C"].shape[0]
    parts = [f"{solver} {n} {m} {p} {steps} {d['h']!r}"]
    for a in (d["A"], d["B"], d["C"], d["D"], d["x0"]):
        parts.append(" ".join(repr(float(v)) for v in np.asarray(a, np.float32).ravel()))
    parts.append(d["usrc"])
    out = subprocess.run([EXE], input="\n".join(parts), capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "ERR 0", lines[0]
    k = int(lines[1].split()[1])
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[2:2 + k]])
    return rows[:, 0], rows[:, 1:]


@pytest.mark.parametrize("solver", [0, 1])
def test_cpp_api_matches_oracle(solver):
    assert os.path.exists(EXE), "build with make -C parode_amd/csrc"
    d = sc.cpp_api()
    steps = 400
    t, y = _run(solver, d, steps)
    t1 = (steps - 1) * d["h"]
    ye = orc.simulate_mirror_exact(solver, d["A"], d["B"], d["C"], d["D"], d["x0"], 0.0, t1,
                                   d["h"], ucoef=d["ucoef"])
    assert y.shape == ye.shape
    assert orc.max_rel_err(y, ye) <= 1e-5
    np.testing.assert_array_equal(t.astype(np.float32), orc.time_vector(solver, 0.0, d["h"],
                                                                        ye.shape[0]))
