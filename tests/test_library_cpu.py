"""The C-ABI library without a GPU: it loads, exports every symbol the headers
declare, and its host-side error paths behave like the reference's."""
import os
import re

import pytest

from parode_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAS_GPU = os.path.exists("/dev/kfd")


def declared_symbols():
    names = set()
    for h in ("ParODECInterface.h", "ParODEExt.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(?:C_EXPORT\s+)?(?:void|int|long)\s+(\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = declared_symbols()
    assert set(api.C_ABI_SYMBOLS) <= names
    assert len(names) >= 10 + 12
    for name in sorted(names):
        assert hasattr(L, name), name


def test_error_descriptions_match_reference_texts():
    # Sources/ParODE.cpp:496-514
    assert api.get_error_description(0) == "No Error"
    assert api.get_error_description(-2) == "ParODE not initialized. Please call InitializeGPU."
    assert api.get_error_description(-3) == "Could not create GPU kernels; not enough resources."
    assert api.get_error_description(-1).startswith("Sorry, no GPU available.")
    assert api.get_error_description(-6) == "Unknown Error"


def test_uninitialized_calls_return_not_initialized():
    api.close_gpu()
    assert api.register_input("fType UFunc(fType t, int i){ return t; }") == (-2, -2)
    assert api.start_timer() == -2
    assert api.stop_timer()[1] == -2
    t, x, e = api.simulate_ode(0, [[-1.0]], None, None, None, 0, 0.0, 1.0, 0.1, [1.0])
    assert e == -2 and len(t) == 0


def test_invalid_system_is_rejected_before_the_gpu():
    # SimulateODE with nStates == 0 -> ParODE_InvalidSystem (ParODE.cpp:116-119)
    import ctypes
    import numpy as np
    L = api.lib()
    z = np.zeros(4, np.float32)
    fp = z.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    one = ctypes.c_int(1)
    zero = ctypes.c_int(0)
    e = ctypes.c_int(0)
    ns = ctypes.c_int(0)
    d = ctypes.c_double(0.1)
    L.SimulateODE(fp, fp, fp, fp, ctypes.byref(one), ctypes.byref(zero), ctypes.byref(one),
                  ctypes.byref(one), ctypes.byref(one), ctypes.byref(one), ctypes.byref(zero),
                  ctypes.byref(d), ctypes.byref(d), ctypes.byref(d), fp, ctypes.byref(zero), fp, fp,
                  ctypes.byref(ns), ctypes.byref(e))
    assert e.value == -6


def test_ufunc_compiles_for_gfx950_without_a_gpu():
    ok, log = api.compile_ufunc("fType UFunc(fType t, int nInputIndex){\n"
                                "fType fVal = nInputIndex == 0 ? sin(t) :\n"
                                "nInputIndex == 1 ? cos(t) :\n0.0;\nreturn fVal;\n}\n")
    assert ok == 0, log
    bad, log = api.compile_ufunc("fType UFunc(fType t, int i){ return sin(t) }")
    assert bad == -3 and "error" in log


def test_synthetic_ufunc_source_compiles():
    from parode_amd import synth
    s = synth.config(0)
    ok, log = api.compile_ufunc(synth.ufunc_source(s.ucoef))
    assert ok == 0, log


def test_plan_footprint_estimate():
    b = api.lib().ParODE_PlanEstimateBytes(1, 64, 8, 1 << 20, 1)
    # trajectory 2^20 x 256 floats dominates: 1 GiB, plus the 21-level FP64 table
    assert (1 << 30) < b < (1 << 30) + (64 << 20)
    assert api.lib().ParODE_PlanEstimateBytes(1, 64, 8, 3, 1) == -1     # ABM needs 4 points


@pytest.mark.skipif(HAS_GPU, reason="checks the no-GPU error path")
def test_no_gpu_paths():
    ids, e = api.get_available_gpus()
    assert ids == [] and e == 0
    assert api.initialize_all_gpus() == -1
    # InitializeGPU cleans first (ParODE.cpp:344): back to NOT_INITIALIZED
    assert api.get_device_name(0) == ("", -2)
    api.get_available_gpus()                      # PARTIAL
    assert api.get_device_name(0) == ("", -5)
    assert api.initialize_selected_gpus([0]) == -1
    api.close_gpu()


def test_server_footprint_and_no_gpu_create():
    f = api.Server.footprint(0, 8, 2, 2, 4096)
    assert f > 4096 * 8 * 4
    assert api.Server.footprint(1, 8, 2, 2, 3) == -1
    assert api.Server.footprint(1, 8, 2, 2, 4096) > f
    if not HAS_GPU:
        with pytest.raises(api.ParODEError):
            api.Server(lanes=1, device_bytes=1 << 20, pinned_bytes=1 << 20)
    with pytest.raises(api.ParODEError):
        api.Server(lanes=0)
