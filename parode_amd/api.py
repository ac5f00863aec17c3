"""ctypes bindings of libParODEShared.so (the C ABI of ParODECInterface.h and
the extensions of ParODEExt.h).

The library is the product: every call here goes to the HIP implementation.
A missing or unloadable library raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libParODEShared.so")

# the ten C symbols of the reference's ParODECInterface.h:23-99
C_ABI_SYMBOLS = ["GetErrorDescription", "InitializeAllGPUs", "GetAvailableGPUs", "GetDeviceName",
                 "InitializeSelectedGPUs", "SimulateODE", "RegisterInput", "CloseGPUC",
                 "StartTimer", "StopTimer"]
EXT_SYMBOLS = ["ParODE_SetLevelKernel", "ParODE_SetPrecision", "ParODE_GetPrecision", "ParODE_SetMaxBatchSteps",
               "ParODE_CompileUFunc", "ParODE_GetLastErrorMessage", "SimulateODESampled",
               "ParODE_PlanCreate", "ParODE_PlanExecute", "ParODE_PlanExecuteTimed",
               "ParODE_PlanFootprint",
               "ParODE_PlanLevels", "ParODE_PlanDestroy", "ParODE_PlanEstimateBytes",
               "ParODE_PlanCapture", "ParODE_PlanReplay", "ParODE_PlanAutotune",
               "ParODE_ScheduleClear", "ParODE_PlanChunkLog",
               "ParODE_ServerCreate", "ParODE_ServerSubmit", "ParODE_ServerWait",
               "ParODE_ServerWaitBatch",
               "ParODE_ServerDrain", "ParODE_ServerStats", "ParODE_ServerFootprint",
               "ParODE_ServerDestroy", "ParODE_SolveNoAlloc", "ParODE_NoAllocScratchBytes",
               "ParODE_ServerSubmitDevice", "ParODE_ServerGroupStats", "ParODE_GroupCreate",
               "ParODE_GroupExecute", "ParODE_GroupExecuteTimed", "ParODE_GroupStages",
               "ParODE_GroupLaunches", "ParODE_GroupFootprint", "ParODE_GroupEstimateBytes",
               "ParODE_GroupDestroy", "ParODE_ServerSubmitDeviceBatch"]

RK4, ABM = 0, 1
PLAIN32, MIXED, FP64 = 0, 1, 2

ERRORS = {0: "ParODE_OK", -1: "ParODE_NoGPU", -2: "ParODE_NotInitialized",
          -3: "ParODE_CouldNotCreateGPUKernels", -4: "ParODE_NotEnoughResources",
          -5: "ParODE_InvalidDeviceId", -6: "ParODE_InvalidSystem",
          -7: "ParODE_InvalidSolverType"}


class ParODEError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({code}) {last_error_message()}")


_lib = None

_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_F32, _F64 = np.dtype(np.float32), np.dtype(np.float64)


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is not built; run `make -C parode_amd/csrc` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.GetErrorDescription.argtypes = [_ip, ctypes.c_char_p]
        L.InitializeAllGPUs.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _ip]
        L.GetAvailableGPUs.argtypes = [_ip, _ip, _ip]
        L.GetDeviceName.argtypes = [_ip, ctypes.c_char_p, _ip]
        L.InitializeSelectedGPUs.argtypes = [_ip, _ip, ctypes.c_char_p, ctypes.c_char_p, _ip]
        L.SimulateODE.argtypes = [_fp, _fp, _fp, _fp, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp,
                                  _dp, _fp, _ip, _fp, _fp, _ip, _ip]
        L.RegisterInput.argtypes = [ctypes.c_char_p, _ip, _ip]
        L.CloseGPUC.argtypes = []
        L.StartTimer.argtypes = [_ip]
        L.StopTimer.argtypes = [_fp, _ip]
        L.ParODE_SetPrecision.argtypes = [ctypes.c_int]
        L.ParODE_GetPrecision.restype = ctypes.c_int
        L.ParODE_SetMaxBatchSteps.argtypes = [ctypes.c_long]
        L.ParODE_SetLevelKernel.argtypes = [ctypes.c_int]
        L.ParODE_CompileUFunc.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.c_int]
        L.ParODE_CompileUFunc.restype = ctypes.c_int
        L.ParODE_GetLastErrorMessage.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.SimulateODESampled.argtypes = [_fp, _fp, _fp, _fp, _ip, _ip, _ip, _ip, _ip, _ip, _vp,
                                         _ip, _dp, _dp, _dp, _fp, _ip, _fp, _fp, _ip, _ip]
        L.ParODE_PlanCreate.argtypes = [ctypes.c_int] * 6 + [ctypes.c_long, ctypes.c_double,
                                                             ctypes.c_int, ctypes.c_int,
                                                             ctypes.POINTER(_vp)]
        L.ParODE_PlanCreate.restype = ctypes.c_int
        L.ParODE_PlanExecute.argtypes = [_vp] * 9
        L.ParODE_PlanExecute.restype = ctypes.c_int
        L.ParODE_PlanExecuteTimed.argtypes = [_vp] * 9 + [_fp]
        L.ParODE_PlanExecuteTimed.restype = ctypes.c_int
        L.ParODE_PlanCapture.argtypes = [_vp] * 9
        L.ParODE_PlanCapture.restype = ctypes.c_int
        L.ParODE_PlanReplay.argtypes = [_vp, _vp]
        L.ParODE_PlanReplay.restype = ctypes.c_int
        L.ParODE_PlanFootprint.argtypes = [_vp]
        L.ParODE_PlanFootprint.restype = ctypes.c_long
        L.ParODE_PlanLevels.argtypes = [_vp]
        L.ParODE_PlanLevels.restype = ctypes.c_int
        L.ParODE_PlanAutotune.argtypes = [_vp, _vp]
        L.ParODE_PlanAutotune.restype = ctypes.c_int
        L.ParODE_ScheduleClear.argtypes = []
        L.ParODE_ScheduleClear.restype = None
        L.ParODE_PlanChunkLog.argtypes = [_vp]
        L.ParODE_PlanChunkLog.restype = ctypes.c_int
        L.ParODE_PlanDestroy.argtypes = [_vp]
        L.ParODE_PlanDestroy.restype = ctypes.c_int
        L.ParODE_PlanEstimateBytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_long, ctypes.c_int]
        L.ParODE_PlanEstimateBytes.restype = ctypes.c_long
        _l, _i32, _d = ctypes.c_long, ctypes.c_int, ctypes.c_double
        L.ParODE_SolveNoAlloc.argtypes = ([_i32] * 6 + [_l, _d, _i32, _i32] + [_vp] * 8)
        L.ParODE_SolveNoAlloc.restype = _i32
        L.ParODE_NoAllocScratchBytes.argtypes = []
        L.ParODE_NoAllocScratchBytes.restype = _l
        L.ParODE_ServerCreate.argtypes = [_i32, _l, _l, _i32, ctypes.POINTER(_vp)]
        L.ParODE_ServerCreate.restype = _i32
        L.ParODE_ServerSubmit.argtypes = ([_vp] + [_i32] * 6 + [_l, _d, _d] + [_fp] * 5 +
                                          [_dp, _d, ctypes.c_ulonglong, _dp, _fp,
                                           ctypes.POINTER(_l)])
        L.ParODE_ServerSubmit.restype = _i32
        L.ParODE_ServerSubmitDevice.argtypes = ([_vp] + [_i32] * 6 + [_l, _d, _d] + [_fp] * 5 +
                                                [_dp, _d, ctypes.c_ulonglong, _dp, _vp,
                                                 ctypes.POINTER(_l)])
        L.ParODE_ServerSubmitDevice.restype = _i32
        _pv = ctypes.POINTER(_vp)
        L.ParODE_ServerSubmitDeviceBatch.argtypes = (
            [_vp, _i32] + [_ip] * 6 + [ctypes.POINTER(_l), _dp, _dp] + [_pv] * 6 +
            [_dp, ctypes.POINTER(ctypes.c_ulonglong), _pv, _pv, ctypes.POINTER(_l)])
        L.ParODE_ServerSubmitDeviceBatch.restype = _i32
        L.ParODE_ServerGroupStats.argtypes = [_vp, ctypes.POINTER(_l)]
        L.ParODE_ServerGroupStats.restype = _i32
        _pp = ctypes.POINTER(_vp)
        L.ParODE_GroupCreate.argtypes = ([_i32] + [_ip] * 4 + [ctypes.POINTER(_l), _dp, _dp] +
                                         [_pp] * 6 + [_dp, ctypes.POINTER(ctypes.c_ulonglong)] +
                                         [_pp] * 2 + [_pp])
        L.ParODE_GroupCreate.restype = _i32
        L.ParODE_GroupExecute.argtypes = [_vp, _vp]
        L.ParODE_GroupExecute.restype = _i32
        L.ParODE_GroupExecuteTimed.argtypes = [_vp, _vp, _fp]
        L.ParODE_GroupExecuteTimed.restype = _i32
        L.ParODE_GroupStages.argtypes = []
        L.ParODE_GroupStages.restype = _i32
        L.ParODE_GroupStageName.argtypes = [_i32]
        L.ParODE_GroupStageName.restype = ctypes.c_char_p
        L.ParODE_GroupLaunches.argtypes = [_vp]
        L.ParODE_GroupLaunches.restype = _i32
        L.ParODE_GroupFootprint.argtypes = [_vp]
        L.ParODE_GroupFootprint.restype = _l
        L.ParODE_GroupEstimateBytes.argtypes = [_i32] * 4 + [_l]
        L.ParODE_GroupEstimateBytes.restype = _l
        L.ParODE_GroupDestroy.argtypes = [_vp]
        L.ParODE_GroupDestroy.restype = _i32
        L.ParODE_ServerWait.argtypes = [_vp, _l, _dp]
        L.ParODE_ServerWait.restype = _i32
        L.ParODE_ServerWaitBatch.argtypes = [_vp, _i32, ctypes.POINTER(_l),
                                             ctypes.POINTER(_i32), _dp]
        L.ParODE_ServerWaitBatch.restype = _i32
        L.ParODE_ServerDrain.argtypes = [_vp]
        L.ParODE_ServerDrain.restype = _i32
        L.ParODE_ServerStats.argtypes = [_vp, ctypes.POINTER(_l)]
        L.ParODE_ServerStats.restype = _i32
        L.ParODE_ServerFootprint.argtypes = [_i32] * 4 + [_l, _i32]
        L.ParODE_ServerFootprint.restype = _l
        L.ParODE_ServerDestroy.argtypes = [_vp]
        L.ParODE_ServerDestroy.restype = _i32
        _lib = L
    return _lib


def last_error_message():
    if _lib is None:
        return ""
    buf = ctypes.create_string_buffer(2048)
    _lib.ParODE_GetLastErrorMessage(buf, 2048)
    return buf.value.decode(errors="replace")


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_fp)


# ------------------------------------------------------------------ C ABI
def get_error_description(code):
    buf = ctypes.create_string_buffer(512)
    lib().GetErrorDescription(ctypes.byref(ctypes.c_int(code)), buf)
    return buf.value.decode()


def get_available_gpus():
    n = ctypes.c_int(0)
    ids = (ctypes.c_int * 64)()
    e = ctypes.c_int(0)
    lib().GetAvailableGPUs(ctypes.byref(n), ids, ctypes.byref(e))
    return [ids[i] for i in range(n.value)], e.value


def get_device_name(dev):
    buf = ctypes.create_string_buffer(256)
    e = ctypes.c_int(0)
    lib().GetDeviceName(ctypes.byref(ctypes.c_int(dev)), buf, ctypes.byref(e))
    return buf.value.decode(), e.value


def initialize_all_gpus(kernel_dir=b"./", include_dir=b"./"):
    e = ctypes.c_int(0)
    lib().InitializeAllGPUs(kernel_dir, include_dir, ctypes.byref(e))
    return e.value


def initialize_selected_gpus(ids, kernel_dir=b"./", include_dir=b"./"):
    arr = (ctypes.c_int * max(1, len(ids)))(*ids)
    e = ctypes.c_int(0)
    lib().InitializeSelectedGPUs(ctypes.byref(ctypes.c_int(len(ids))), arr, kernel_dir,
                                 include_dir, ctypes.byref(e))
    return e.value


def register_input(src):
    idx = ctypes.c_int(0)
    e = ctypes.c_int(0)
    lib().RegisterInput(src.encode() if isinstance(src, str) else src, ctypes.byref(idx),
                        ctypes.byref(e))
    return idx.value, e.value


def close_gpu():
    lib().CloseGPUC()


def start_timer():
    e = ctypes.c_int(0)
    lib().StartTimer(ctypes.byref(e))
    return e.value


def stop_timer():
    t = ctypes.c_float(0)
    e = ctypes.c_int(0)
    lib().StopTimer(ctypes.byref(t), ctypes.byref(e))
    return t.value, e.value


def set_precision(mode):
    lib().ParODE_SetPrecision(mode)


VECTOR_LEVELS, VECTOR_GEMM, NO_FUSED_TOP, NO_CHUNKS = 1, 2, 4, 8


def set_level_kernel(variant):
    """Bit mask (0 = automatic): VECTOR_LEVELS | VECTOR_GEMM | NO_FUSED_TOP | NO_CHUNKS."""
    lib().ParODE_SetLevelKernel(int(variant))


def set_max_batch_steps(n):
    lib().ParODE_SetMaxBatchSteps(int(n))


def compile_ufunc(src, arch="gfx950"):
    log = ctypes.create_string_buffer(16384)
    r = lib().ParODE_CompileUFunc(src.encode(), arch.encode(), log, 16384)
    return r, log.value.decode(errors="replace")


def _system_args(A, B, C, D, x0):
    A = _f32(A)
    n = A.shape[0]
    zB = B is None or np.size(B) == 0
    zC = C is None or np.size(C) == 0
    zD = D is None or np.size(D) == 0
    m = 0 if zB else np.shape(B)[1]
    if zB and not zD:
        m = np.shape(D)[1]
    p = n if (zC and zD) else (np.shape(C)[0] if not zC else np.shape(D)[0])
    Bf = None if zB else _f32(B)
    Cf = None if zC else _f32(C)
    Df = None if zD else _f32(D)
    return A, Bf, Cf, Df, _f32(x0), n, m, p, int(zB), int(zC), int(zD)


def simulate_ode(solver, A, B, C, D, u_index, t_start, t_end, t_step, x0, capacity=None):
    """SimulateODE through the C ABI.  Returns (tVect, xVect[nSteps, nOut], nErr)."""
    A, Bf, Cf, Df, x0f, n, m, p, zB, zC, zD = _system_args(A, B, C, D, x0)
    nOut = n if (zC and zD) else p
    cap = capacity or (int((t_end - t_start) / t_step) + 8)
    tv = np.zeros(cap, np.float32)
    xv = np.zeros((cap, max(nOut, n)), np.float32)
    nsteps = ctypes.c_int(0)
    e = ctypes.c_int(0)
    lib().SimulateODE(_fptr(A), _fptr(Bf), _fptr(Cf), _fptr(Df), _i(m), _i(n), _i(p), _i(zB),
                      _i(zC), _i(zD), _i(u_index), ctypes.byref(ctypes.c_double(t_start)),
                      ctypes.byref(ctypes.c_double(t_end)), ctypes.byref(ctypes.c_double(t_step)),
                      _fptr(x0f), _i(solver), _fptr(tv), _fptr(xv), ctypes.byref(nsteps),
                      ctypes.byref(e))
    k = nsteps.value
    return tv[:k], xv.reshape(-1)[:k * nOut].reshape(k, nOut), e.value


def simulate_ode_sampled(solver, A, B, C, D, samples, t_start, t_end, t_step, x0,
                         capacity=None):
    """SimulateODESampled: samples u(t_start + s h/2) (float32 or float64)."""
    A, Bf, Cf, Df, x0f, n, m, p, zB, zC, zD = _system_args(A, B, C, D, x0)
    nOut = n if (zC and zD) else p
    cap = capacity or (int((t_end - t_start) / t_step) + 8)
    tv = np.zeros(cap, np.float32)
    xv = np.zeros((cap, max(nOut, n)), np.float32)
    isd = samples is not None and np.asarray(samples).dtype == np.float64
    S = None if samples is None else np.ascontiguousarray(samples,
                                                          np.float64 if isd else np.float32)
    nsteps = ctypes.c_int(0)
    e = ctypes.c_int(0)
    lib().SimulateODESampled(_fptr(A), _fptr(Bf), _fptr(Cf), _fptr(Df), _i(m), _i(n), _i(p),
                             _i(zB), _i(zC), _i(zD),
                             None if S is None else S.ctypes.data_as(_vp), _i(int(isd)),
                             ctypes.byref(ctypes.c_double(t_start)),
                             ctypes.byref(ctypes.c_double(t_end)),
                             ctypes.byref(ctypes.c_double(t_step)), _fptr(x0f), _i(solver),
                             _fptr(tv), _fptr(xv), ctypes.byref(nsteps), ctypes.byref(e))
    k = nsteps.value
    return tv[:k], xv.reshape(-1)[:k * nOut].reshape(k, nOut), e.value


# ------------------------------------------------------------------- plans
class Plan:
    """A device-resident plan (ParODE_PlanCreate/Execute).  Arguments to
    execute() are device pointers (ints), e.g. torch tensors' data_ptr()."""

    def __init__(self, solver, n, m, p, zero_c, zero_d, points, h, precision, samples_double):
        self.handle = ctypes.c_void_p()
        r = lib().ParODE_PlanCreate(solver, n, m, p, int(zero_c), int(zero_d), points, h,
                                    precision, int(samples_double), ctypes.byref(self.handle))
        if r != 0:
            raise ParODEError(r, "ParODE_PlanCreate")
        self.solver, self.n, self.m, self.p, self.points = solver, n, m, p, points
        self.n_out = n if (zero_c and zero_d) else p

    def execute(self, dA, dB, dC, dD, dx0, dU, dY, stream=0):
        r = lib().ParODE_PlanExecute(self.handle, dA, dB, dC, dD, dx0, dU, dY, stream)
        if r != 0:
            raise ParODEError(r, "ParODE_PlanExecute")

    def capture(self, dA, dB, dC, dD, dx0, dU, dY, stream):
        """Record the solve with these pointers into a HIP graph (stream != 0)."""
        r = lib().ParODE_PlanCapture(self.handle, dA, dB, dC, dD, dx0, dU, dY, stream)
        if r != 0:
            raise ParODEError(r, "ParODE_PlanCapture")

    def replay(self, stream=0):
        r = lib().ParODE_PlanReplay(self.handle, stream)
        if r != 0:
            raise ParODEError(r, "ParODE_PlanReplay")

    PHASES = ("map_powers", "leaves", "up_sweep", "top", "down_sweep", "outputs")

    def execute_timed(self, dA, dB, dC, dD, dx0, dU, dY, stream=0):
        """Execute, wait, and return {phase: GPU ms} measured with hipEvents."""
        ms = (ctypes.c_float * 6)()
        r = lib().ParODE_PlanExecuteTimed(self.handle, dA, dB, dC, dD, dx0, dU, dY, stream, ms)
        if r != 0:
            raise ParODEError(r, "ParODE_PlanExecuteTimed")
        return dict(zip(self.PHASES, list(ms)))

    @property
    def footprint(self):
        return lib().ParODE_PlanFootprint(self.handle)

    @property
    def levels(self):
        return lib().ParODE_PlanLevels(self.handle)

    @property
    def chunk_log(self):
        """log2 of the chunk length of the ABM chunk recurrences (0: none)."""
        return lib().ParODE_PlanChunkLog(self.handle)

    def autotune(self, stream=0):
        """Times the kernels that can run each scan level of this shape and
        keeps the fastest (process-wide); returns the level launches tuned."""
        r = lib().ParODE_PlanAutotune(self.handle, ctypes.c_void_p(stream))
        if r < 0:
            raise RuntimeError(f"ParODE_PlanAutotune: {r} {last_error_message()}")
        return r

    def close(self):
        if self.handle:
            lib().ParODE_PlanDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (ParODE_Server*): submit() returns a
    ticket at once; the library's worker thread admits requests onto its lanes
    as device memory allows; wait() returns (status, latency_ms, queue_ms) and
    the outputs are in the array given to submit()."""

    STATS = ("submitted", "completed", "failed", "running", "queued", "peak_running",
             "arena_bytes", "arena_peak")

    def __init__(self, lanes=4, device_bytes=16 << 30, pinned_bytes=1 << 30, precision=MIXED):
        self.handle = ctypes.c_void_p()
        r = lib().ParODE_ServerCreate(lanes, device_bytes, pinned_bytes, precision,
                                      ctypes.byref(self.handle))
        if r != 0:
            raise ParODEError(r, "ParODE_ServerCreate")
        self._keep = {}

    def submit(self, solver, A, B, C, D, x0, points, h, t0=0.0, coef=None, sigma=0.0, seed=0,
               samples=None, Y=None):
        """Queues one solve; returns (ticket, Y) where Y (points x nOut float32)
        is filled when the ticket completes."""
        A, Bf, Cf, Df, x0f, n, m, p, zB, zC, zD = _system_args(A, B, C, D, x0)
        n_out = n if (zC and zD) else p
        if Y is None:
            Y = np.zeros((points, n_out), np.float32)
        assert Y.dtype == np.float32 and Y.flags.c_contiguous and Y.size >= points * n_out
        cf = None if coef is None else np.ascontiguousarray(coef, np.float64)
        sm = None if samples is None else np.ascontiguousarray(samples, np.float64)
        t = ctypes.c_long(0)
        r = lib().ParODE_ServerSubmit(
            self.handle, solver, n, m, p, zC, zD, points, t0, h, _fptr(A), _fptr(Bf), _fptr(Cf),
            _fptr(Df), _fptr(x0f),
            None if cf is None else cf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            sigma, seed,
            None if sm is None else sm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            Y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(t))
        if r != 0:
            raise ParODEError(r, "ParODE_ServerSubmit")
        self._keep[t.value] = (sm, Y)
        return t.value, Y

    def submit_device(self, solver, A, B, C, D, x0, points, h, dY, t0=0.0, coef=None, sigma=0.0,
                      seed=0, samples=None):
        """Queues one solve whose outputs go to device memory dY (an address of
        points x nOut floats, e.g. a torch tensor's data_ptr()); returns the ticket."""
        A, Bf, Cf, Df, x0f, n, m, p, zB, zC, zD = _system_args(A, B, C, D, x0)
        cf = None if coef is None else np.ascontiguousarray(coef, np.float64)
        sm = None if samples is None else np.ascontiguousarray(samples, np.float64)
        t = ctypes.c_long(0)
        r = lib().ParODE_ServerSubmitDevice(
            self.handle, solver, n, m, p, zC, zD, points, t0, h, _fptr(A), _fptr(Bf), _fptr(Cf),
            _fptr(Df), _fptr(x0f),
            None if cf is None else cf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            sigma, seed,
            None if sm is None else sm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            dY, ctypes.byref(t))
        if r != 0:
            raise ParODEError(r, "ParODE_ServerSubmitDevice")
        self._keep[t.value] = (sm, None)
        return t.value

    def submit_device_batch(self, items):
        """Queues many device-output solves in one call; items are tuples
        (solver, A, B, C, D, x0, points, h, dY, coef) with dY a device address
        of points x nOut floats and coef the m x 6 sinusoid coefficients.
        Returns the tickets, in order."""
        k = len(items)
        ints = [[0] * k for _ in range(6)]
        pts, hs = [0] * k, [0.0] * k
        ptr = [[None] * k for _ in range(7)]
        keep = []

        def addr(a, dtype):
            # the array's address, converted (and kept alive) only when needed
            if a is None:
                return None, None
            if not (type(a) is np.ndarray and a.dtype is dtype and a.flags.c_contiguous):
                a = np.ascontiguousarray(a, dtype)
                keep.append(a)
            if a.size == 0:
                return None, None
            try:   # (a third of the cost of a.ctypes.data; writable arrays only)
                return a, ctypes.addressof(ctypes.c_byte.from_buffer(a))
            except TypeError:
                return a, a.ctypes.data

        for i, (solver, A, B, C, D, x0, points, h, dY, coef) in enumerate(items):
            A, pA = addr(A, _F32)
            B, pB = addr(B, _F32)
            C, pC = addr(C, _F32)
            D, pD = addr(D, _F32)
            _, px = addr(x0, _F32)
            _, pc = addr(coef, _F64)
            n = A.shape[0]
            m = B.shape[1] if B is not None else (D.shape[1] if D is not None else 0)
            zC, zD = int(C is None), int(D is None)
            p = n if (zC and zD) else (C.shape[0] if C is not None else D.shape[0])
            for a, v in zip(ints, (solver, n, m, p, zC, zD)):
                a[i] = v
            pts[i], hs[i] = points, h
            for a, v in zip(ptr, (pA, pB, pC, pD, px, pc, dY)):
                a[i] = v
        ints = [np.array(a, np.int32) for a in ints]
        pts, t0s, hs = np.array(pts, np.int64), np.zeros(k), np.array(hs, np.float64)
        ptr = [(ctypes.c_void_p * k)(*a) for a in ptr]
        tickets = (ctypes.c_long * k)()
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        r = lib().ParODE_ServerSubmitDeviceBatch(
            self.handle, k, *(ip(a) for a in ints), pts.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
            dp(t0s), dp(hs), *ptr[:6], None, None, None, ptr[6], tickets)
        if r != 0:
            raise ParODEError(r, "ParODE_ServerSubmitDeviceBatch")
        del keep
        return list(tickets)

    def wait(self, ticket):
        ms = (ctypes.c_double * 2)()
        st = lib().ParODE_ServerWait(self.handle, ticket, ms)
        self._keep.pop(ticket, None)
        return st, ms[0], ms[1]

    def wait_batch(self, tickets):
        """Waits for every ticket (in order) in one call; returns the statuses,
        the submission-to-completion and the submission-to-admission times (ms)
        as arrays."""
        k = len(tickets)
        tk = (ctypes.c_long * k)(*tickets)
        st = (ctypes.c_int * k)()
        ms = np.zeros(2 * k)
        r = lib().ParODE_ServerWaitBatch(self.handle, k, tk, st,
                                         ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if r != 0:
            raise ParODEError(r, "ParODE_ServerWaitBatch")
        for t in tickets:
            self._keep.pop(t, None)
        return np.frombuffer(st, np.int32).copy(), ms[0::2].copy(), ms[1::2].copy()

    def drain(self):
        lib().ParODE_ServerDrain(self.handle)

    def stats(self):
        v = (ctypes.c_long * 8)()
        lib().ParODE_ServerStats(self.handle, v)
        return dict(zip(self.STATS, list(v)))

    def group_stats(self):
        """groups admitted, requests run in grouped launches, the largest group"""
        v = (ctypes.c_long * 3)()
        lib().ParODE_ServerGroupStats(self.handle, v)
        return dict(zip(("groups", "grouped_requests", "largest_group"), list(v)))

    @staticmethod
    def footprint(solver, n, m, p, points, precision=MIXED):
        return lib().ParODE_ServerFootprint(solver, n, m, p, points, precision)

    def close(self):
        if self.handle:
            lib().ParODE_ServerDestroy(self.handle)
            self.handle = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_no_alloc(solver, n, m, p, zero_c, zero_d, points, h, precision, samples_double,
                   dA, dB, dC, dD, dx0, dU, dY, stream=0):
    """ParODE_SolveNoAlloc: the plan's solve with its scan buffers in the
    library's static device scratch array (no device allocation).  Device
    pointers as for Plan.execute.  Returns the status (0 = OK)."""
    return lib().ParODE_SolveNoAlloc(solver, n, m, p, int(zero_c), int(zero_d), points, h,
                                     precision, int(samples_double), dA, dB, dC, dD, dx0, dU, dY,
                                     stream)


def no_alloc_scratch_bytes():
    return lib().ParODE_NoAllocScratchBytes()


class Group:
    """A device-resident grouped solve (ParODE_Group*): many systems, each stage
    of the scan one launch for all of them, mixed precision.  systems: list of
    dicts with keys method, n, m, p, points, h, t0 (optional), and device
    addresses A, B, C, D, x0, U, Y (ints), plus coef (device address or None:
    U then holds FP64 samples), sigma, seed (optional)."""

    def __init__(self, systems):
        k = len(systems)
        ints = lambda key: (ctypes.c_int * k)(*[int(s[key]) for s in systems])
        ptrs = lambda key: (_vp * k)(*[s.get(key) or None for s in systems])
        self.handle = ctypes.c_void_p()
        r = lib().ParODE_GroupCreate(
            k, ints("method"), ints("n"), ints("m"), ints("p"),
            (ctypes.c_long * k)(*[int(s["points"]) for s in systems]),
            (ctypes.c_double * k)(*[float(s.get("t0", 0.0)) for s in systems]),
            (ctypes.c_double * k)(*[float(s["h"]) for s in systems]),
            ptrs("A"), ptrs("B"), ptrs("C"), ptrs("D"), ptrs("x0"), ptrs("coef"),
            (ctypes.c_double * k)(*[float(s.get("sigma", 0.0)) for s in systems]),
            (ctypes.c_ulonglong * k)(*[int(s.get("seed", 0)) for s in systems]),
            ptrs("U"), ptrs("Y"), ctypes.byref(self.handle))
        if r != 0:
            raise ParODEError(r, "ParODE_GroupCreate")

    def execute(self, stream=0):
        r = lib().ParODE_GroupExecute(self.handle, stream)
        if r != 0:
            raise ParODEError(r, "ParODE_GroupExecute")

    def execute_timed(self, stream=0):
        """Execute, wait, and return {stage: GPU ms} measured with hipEvents."""
        k = lib().ParODE_GroupStages()
        ms = (ctypes.c_float * k)()
        r = lib().ParODE_GroupExecuteTimed(self.handle, stream, ms)
        if r != 0:
            raise ParODEError(r, "ParODE_GroupExecuteTimed")
        return {lib().ParODE_GroupStageName(i).decode(): ms[i] for i in range(k)}

    @property
    def launches(self):
        return lib().ParODE_GroupLaunches(self.handle)

    @property
    def footprint(self):
        return lib().ParODE_GroupFootprint(self.handle)

    @staticmethod
    def estimate_bytes(method, n, m, p, points):
        return lib().ParODE_GroupEstimateBytes(method, n, m, p, points)

    def close(self):
        if self.handle:
            lib().ParODE_GroupDestroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def schedule_clear():
    """Every scan level back to the built-in kernel rule."""
    lib().ParODE_ScheduleClear()
