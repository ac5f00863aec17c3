// Device-resident plans (ParODEExt.h): one system shape whose matrices,
// samples and outputs already live in HBM; the whole solve is launched
// asynchronously on the caller's stream.
#include <ParODEEnums.h>
#include <ParODEExt.h>

#include <memory>

#include "config.h"
#include "hip_util.h"
#include "kernels.h"
#include "scan_core.h"

namespace {

struct Plan {
    int solver, n, m, p;
    bool zC, zD;
    long points, nScan, nPad;
    double h;
    bool uDouble;
    int device;
    std::unique_ptr<parode::ScanCore> core;
    hipEvent_t ev[8] = {};
    // the power table and the ABM start-up run here, beside the leaves
    parode::ScanCore::SideStream side;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~Plan()
    {
        drop_graph();
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (side.fork) (void)hipEventDestroy(side.fork);
        if (side.join) (void)hipEventDestroy(side.join);
        if (side.late) (void)hipEventDestroy(side.late);
        if (side.stream) (void)hipStreamDestroy(side.stream);
    }
    const parode::ScanCore::SideStream* side_stream()
    {
        if (!side.stream) {
            PARODE_HIP(hipStreamCreate(&side.stream));
            PARODE_HIP(hipEventCreateWithFlags(&side.fork, hipEventDisableTiming));
            PARODE_HIP(hipEventCreateWithFlags(&side.join, hipEventDisableTiming));
            PARODE_HIP(hipEventCreateWithFlags(&side.late, hipEventDisableTiming));
        }
        return &side;
    }
};

enum { kPhaseMap = 0, kPhaseLeaves, kPhaseUp, kPhaseRoot, kPhaseDown, kPhaseOutputs, kPhases };

// The whole solve on stream s; with mark, an event is recorded before each
// phase and after the last (phases as enumerated above).
void run_plan(Plan* pl, const float* dA, const float* dB, const float* dC, const float* dD,
              const float* dx0, const void* dSamples, float* dY, hipStream_t s, bool mark)
{
    pl->core->solve(s, pl->points, pl->h, dA, dB, dC, dD, dx0, dSamples, pl->uDouble, dY,
                    mark ? pl->ev : nullptr, pl->side_stream());
}

int shape_error(int solver, int n, int m, long points)
{
    if (solver != 0 && solver != 1) return ParODE_InvalidSolverType;
    if (n <= 0 || m < 0) return ParODE_InvalidSystem;
    if (points < 1 || (solver == 1 && points < 4)) return ParODE_InvalidSystem;
    return ParODE_OK;
}

}  // namespace

extern "C" {

long ParODE_PlanEstimateBytes(int solver, int nStates, int nInputs, long points, int precision)
{
    if (shape_error(solver, nStates, nInputs, points) != ParODE_OK) return -1;
    parode::ScanOptions opt;
    opt.precision = precision;
    long nScan = solver == 0 ? points : points - 3;
    return (long)parode::ScanCore::estimate_footprint(solver, nStates, nInputs, nScan, opt);
}

int ParODE_PlanCreate(int solver, int nStates, int nInputs, int nOutputs, int zeroC, int zeroD,
                      long points, double tStep, int precision, int samplesDouble, void** plan)
{
    *plan = nullptr;
    int err = shape_error(solver, nStates, nInputs, points);
    if (err != ParODE_OK) return err;
    if (precision < parode::kPlain32 || precision > parode::kFp64) return ParODE_InvalidSystem;
    if (!(zeroC && zeroD) && nOutputs <= 0) return ParODE_InvalidSystem;
    try {
        std::unique_ptr<Plan> pl(new Plan);
        pl->solver = solver;
        pl->n = nStates;
        pl->m = nInputs;
        pl->p = nOutputs;
        pl->zC = zeroC != 0;
        pl->zD = zeroD != 0;
        pl->points = points;
        pl->nScan = solver == 0 ? points : points - 3;
        pl->nPad = parode::next_pow2(pl->nScan);
        pl->h = tStep;
        pl->uDouble = samplesDouble != 0;
        PARODE_HIP(hipGetDevice(&pl->device));
        parode::ScanOptions opt;
        opt.precision = precision;
        pl->core.reset(new parode::ScanCore(solver, nStates, nInputs, nOutputs, pl->zC, pl->zD,
                                            pl->nPad, opt));
        *plan = pl.release();
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_PlanExecute(void* plan, const float* dA, const float* dB, const float* dC,
                       const float* dD, const float* dx0, const void* dSamples, float* dY,
                       void* stream)
{
    Plan* pl = static_cast<Plan*>(plan);
    if (!pl) return ParODE_InvalidSystem;
    if (!dA || !dx0 || !dY || (pl->m > 0 && (!dB || !dSamples)) || (!pl->zC && !dC) ||
        (!pl->zD && !dD))
        return ParODE_InvalidSystem;
    try {
        PARODE_HIP(hipSetDevice(pl->device));
        run_plan(pl, dA, dB, dC, dD, dx0, dSamples, dY, static_cast<hipStream_t>(stream), false);
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_PlanExecuteTimed(void* plan, const float* dA, const float* dB, const float* dC,
                            const float* dD, const float* dx0, const void* dSamples, float* dY,
                            void* stream, float* phaseMs)
{
    Plan* pl = static_cast<Plan*>(plan);
    if (!pl || !phaseMs) return ParODE_InvalidSystem;
    if (!dA || !dx0 || !dY || (pl->m > 0 && (!dB || !dSamples)) || (!pl->zC && !dC) ||
        (!pl->zD && !dD))
        return ParODE_InvalidSystem;
    try {
        PARODE_HIP(hipSetDevice(pl->device));
        for (hipEvent_t& e : pl->ev)
            if (!e) PARODE_HIP(hipEventCreate(&e));
        hipStream_t s = static_cast<hipStream_t>(stream);
        run_plan(pl, dA, dB, dC, dD, dx0, dSamples, dY, s, true);
        PARODE_HIP(hipEventSynchronize(pl->ev[kPhases]));
        for (int i = 0; i < kPhases; ++i)
            PARODE_HIP(hipEventElapsedTime(&phaseMs[i], pl->ev[i], pl->ev[i + 1]));
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_PlanCapture(void* plan, const float* dA, const float* dB, const float* dC,
                       const float* dD, const float* dx0, const void* dSamples, float* dY,
                       void* stream)
{
    Plan* pl = static_cast<Plan*>(plan);
    if (!pl || !stream) return ParODE_InvalidSystem;
    if (!dA || !dx0 || !dY || (pl->m > 0 && (!dB || !dSamples)) || (!pl->zC && !dC) ||
        (!pl->zD && !dD))
        return ParODE_InvalidSystem;
    hipStream_t s = static_cast<hipStream_t>(stream);
    try {
        PARODE_HIP(hipSetDevice(pl->device));
        pl->drop_graph();
        PARODE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            run_plan(pl, dA, dB, dC, dD, dx0, dSamples, dY, s, false);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        PARODE_HIP(hipStreamEndCapture(s, &pl->graph));
        PARODE_HIP(hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0));
        return ParODE_OK;
    } catch (std::exception& e) {
        pl->drop_graph();
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

int ParODE_PlanReplay(void* plan, void* stream)
{
    Plan* pl = static_cast<Plan*>(plan);
    if (!pl || !pl->exec) return ParODE_InvalidSystem;
    try {
        PARODE_HIP(hipSetDevice(pl->device));
        PARODE_HIP(hipGraphLaunch(pl->exec, static_cast<hipStream_t>(stream)));
        return ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

long ParODE_PlanFootprint(void* plan)
{
    Plan* pl = static_cast<Plan*>(plan);
    return pl ? (long)pl->core->footprint() : -1;
}

int ParODE_PlanLevels(void* plan)
{
    Plan* pl = static_cast<Plan*>(plan);
    return pl ? pl->core->levels() : -1;
}

int ParODE_PlanAutotune(void* plan, void* stream)
{
    Plan* pl = static_cast<Plan*>(plan);
    if (!pl) return ParODE_InvalidSystem;
    try {
        PARODE_HIP(hipSetDevice(pl->device));
        return pl->core->autotune(static_cast<hipStream_t>(stream), pl->nScan, pl->nPad);
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
}

void ParODE_ScheduleClear(void) { parode::ScanCore::clear_schedule(); }

int ParODE_PlanChunkLog(void* plan)
{
    Plan* pl = static_cast<Plan*>(plan);
    return pl ? pl->core->chunk_log(pl->nScan, pl->nPad) : -1;
}

int ParODE_PlanDestroy(void* plan)
{
    Plan* pl = static_cast<Plan*>(plan);
    if (!pl) return ParODE_InvalidSystem;
    (void)hipSetDevice(pl->device);
    (void)hipDeviceSynchronize();
    delete pl;
    return ParODE_OK;
}

}  // extern "C"
