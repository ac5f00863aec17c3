// A device-resident solve that allocates no device memory (ParODE_SolveNoAlloc).
//
// For callers that cannot call hipMalloc: code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation.  The solve is the one a
// plan runs (ParODE_PlanCreate + ParODE_PlanExecute, the same ScanCore and
// the same kernels), but its scan buffers (step map, FP64 power table, level
// tables, trajectory, K-split copies, seeds) are carved out of `scratch`, a
// file-scope device array of kScratchBytes that each GPU holds from the time
// the library's code is loaded on it until the process ends.
//
// Every call on one GPU uses that GPU's one copy of the array, so calls that
// overlap on one GPU (different streams, or several host threads) race on it;
// calls queued on one stream run one after the other and are safe.  The array
// keeps the last call's contents.  A solve writes every scan buffer before it
// reads it (the level tables' zero padding and the trajectory's padded tail
// included), as plans and served requests, which reuse buffers that hold old
// contents, rely on too; so nothing needs clearing between calls.
#include <ParODEEnums.h>
#include <ParODEExt.h>

#include <mutex>
#include <string>

#include "config.h"
#include "hip_util.h"
#include "scan_core.h"

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;   // 128 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};

// This device's copy of `scratch`: its address is looked up once per device,
// with that device current, and used from then on wherever host code needs
// the array's memory (the name itself is a host-side stand-in, not a device
// address).
void* scratch_on(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) throw std::out_of_range("ParODE_SolveNoAlloc: device id");
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) throw parode::HipError(e, "hipGetSymbolAddress(scratch)");
        g_scratchPtr[dev] = p;
    }
    return g_scratchPtr[dev];
}

}  // namespace

extern "C" {

long ParODE_NoAllocScratchBytes(void) { return (long)kScratchBytes; }

int ParODE_SolveNoAlloc(int solver, int nStates, int nInputs, int nOutputs, int zeroC, int zeroD,
                        long points, double tStep, int precision, int samplesDouble,
                        const float* dA, const float* dB, const float* dC, const float* dD,
                        const float* dx0, const void* dSamples, float* dY, void* stream)
{
    if (solver != 0 && solver != 1) return ParODE_InvalidSolverType;
    if (nStates <= 0 || nInputs < 0 || points < 1 || (solver == 1 && points < 4))
        return ParODE_InvalidSystem;
    if (precision < parode::kPlain32 || precision > parode::kFp64) return ParODE_InvalidSystem;
    if (!(zeroC && zeroD) && nOutputs <= 0) return ParODE_InvalidSystem;
    if (!dA || !dx0 || !dY || (nInputs > 0 && (!dB || !dSamples)) || (!zeroC && !dC) ||
        (!zeroD && !dD))
        return ParODE_InvalidSystem;
    try {
        parode::ScanOptions opt;
        opt.precision = precision;
        const long nScan = solver == 0 ? points : points - 3;
        const long nPad = parode::next_pow2(nScan);
        const size_t need =
            parode::ScanCore::estimate_footprint(solver, nStates, nInputs, nPad, opt, nOutputs);
        if (need > kScratchBytes) {
            parode::last_error() = "ParODE_SolveNoAlloc: this system needs " +
                                   std::to_string(need) + " bytes of scan buffers, more than the " +
                                   std::to_string(kScratchBytes) + "-byte scratch array";
            return ParODE_NotEnoughResources;
        }
        int dev = 0;
        PARODE_HIP(hipGetDevice(&dev));
        parode::Arena view(scratch_on(dev), kScratchBytes);   // bookkeeping only
        parode::ScanCore core(solver, nStates, nInputs, nOutputs, zeroC != 0, zeroD != 0, nPad,
                              opt, &view);
        core.solve(static_cast<hipStream_t>(stream), points, tStep, dA, dB, dC, dD, dx0, dSamples,
                   samplesDouble != 0, dY);
        return ParODE_OK;
    } catch (std::bad_alloc&) {
        parode::last_error() = "ParODE_SolveNoAlloc: the scan buffers do not fit the scratch array";
        return ParODE_NotEnoughResources;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

}  // extern "C"
