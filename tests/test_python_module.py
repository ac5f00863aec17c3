"""The Python 3 module "ParODE" (the reference's PythonInterface/ParODEmodule.c
call surface) over libParODEShared.so."""
import importlib
import os
import sys

import numpy as np
import pytest

import oracle as orc
import scenarios as sc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "parode_amd", "lib")
HAS_GPU = os.path.exists("/dev/kfd")

METHODS = ("GetErrorDescription", "InitializeAllGPUs", "GetAvailableGPUs", "GetDeviceName",
           "InitializeSelectedGPUs", "SimulateODE", "RegisterInput", "CloseGPU", "StartTimer",
           "StopTimer")


@pytest.fixture(scope="module")
def ParODE():
    if LIBDIR not in sys.path:
        sys.path.insert(0, LIBDIR)
    return importlib.import_module("ParODE")


def test_module_has_the_reference_methods(ParODE):
    # PythonInterface/ParODEmodule.c:450-462
    for name in METHODS:
        assert callable(getattr(ParODE, name)), name
    assert issubclass(ParODE.error, Exception)


def test_error_description_texts(ParODE):
    assert ParODE.GetErrorDescription(0) == "No Error"
    assert ParODE.GetErrorDescription(-2) == "ParODE not initialized. Please call InitializeGPU."
    assert ParODE.GetErrorDescription(-1).startswith("Sorry, no GPU available.")


def test_malformed_arguments_raise(ParODE):
    with pytest.raises(ValueError):
        ParODE.SimulateODE([[1.0, 2.0], [3.0]], [], [], [], 0, 0.0, 0.1, 10, [1.0, 1.0], 0)
    with pytest.raises(ValueError):
        ParODE.SimulateODE([[1.0, 2.0]], [], [], [], 0, 0.0, 0.1, 10, [1.0], 0)
    with pytest.raises(ValueError):
        ParODE.SimulateODE([[-1.0]], [[1.0], [1.0]], [], [], 0, 0.0, 0.1, 10, [1.0], 0)
    with pytest.raises(ValueError):
        ParODE.SimulateODE([[-1.0]], [], [], [], 0, 0.0, 0.1, 10, [1.0, 2.0], 0)
    with pytest.raises(TypeError):
        ParODE.SimulateODE([[-1.0]], [], [], [], 0, 0.0, 0.1, 10, [1.0])
    with pytest.raises(TypeError):
        ParODE.InitializeSelectedGPUs([0, "a"])


def test_uninitialized_paths(ParODE):
    ParODE.CloseGPU()
    assert ParODE.RegisterInput("fType UFunc(fType t, int i){ return t; }") == (-2, -2)
    assert ParODE.StartTimer() == -2
    assert ParODE.StopTimer()[1] == -2
    assert ParODE.SimulateODE([[-1.0]], [], [], [], 0, 0.0, 0.1, 10, [1.0], 0) == ([], [], -2)


@pytest.mark.skipif(HAS_GPU, reason="checks the no-GPU answers")
def test_no_gpu_answers(ParODE):
    assert ParODE.GetAvailableGPUs() == ([], 0)
    assert ParODE.InitializeAllGPUs() == -1


@pytest.mark.gpu
def test_reference_python_script_flow(ParODE):
    # PythonInterface/TestPythonBasicFunctions.py, with results checked
    gpus, e = ParODE.GetAvailableGPUs()
    assert e == 0 and len(gpus) >= 1
    name, e = ParODE.GetDeviceName(gpus[0])
    assert e == 0 and len(name) > 0
    assert ParODE.InitializeSelectedGPUs(gpus) == 0
    assert ParODE.StartTimer() == 0
    assert ParODE.GetErrorDescription(-1).startswith("Sorry")
    ms, e = ParODE.StopTimer()
    assert e == 0 and ms >= 0.0
    ParODE.CloseGPU()


@pytest.mark.gpu
@pytest.mark.parametrize("make", [sc.rk4_class, sc.abm_class])
def test_simulate_through_module_matches_oracle(ParODE, make):
    d = make()
    gpus, e = ParODE.GetAvailableGPUs()
    assert ParODE.InitializeSelectedGPUs(gpus[:1]) == 0
    idx, e = ParODE.RegisterInput(d["usrc"])
    assert e == 0
    tl, xl, e = ParODE.SimulateODE(d["A"].tolist(), d["B"].tolist(), d["C"].tolist(),
                                   d["D"].tolist(), idx, 0.0, d["h"], d["steps"],
                                   d["x0"].tolist(), d["solver"])
    assert e == 0
    # the reference module's horizon, tStart + (nSteps - 1) tStep (ParODEmodule.c:330),
    # and its truncating point count (RungeKutta4thOrder.cpp:20)
    t1 = (d["steps"] - 1) * d["h"]
    count = orc.step_count(d["solver"], 0.0, t1, d["h"])
    assert count in (d["steps"], d["steps"] - 1)
    assert len(tl) == count and len(xl) == count
    ye = orc.simulate_mirror_exact(d["solver"], d["A"], d["B"], d["C"], d["D"], d["x0"], 0.0, t1,
                                   d["h"], ucoef=d["ucoef"])
    assert ye.shape[0] == count
    assert orc.max_rel_err(np.array(xl, np.float32), ye) <= 1e-5
    np.testing.assert_array_equal(np.array(tl, np.float32),
                                  orc.time_vector(d["solver"], 0.0, d["h"], count))
    ParODE.CloseGPU()
