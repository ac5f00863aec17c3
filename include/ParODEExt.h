/*
 * ParODEExt.h -- entry points beyond the reference's ten-function C ABI.
 * None of these has a counterpart in the reference; each says which caller it
 * is for.  All are C linkage, plain pointers and sizes; errors are the
 * ErrorCode values of ParODEEnums.h (0 = OK).
 */
#ifndef PARODEEXT_H_
#define PARODEEXT_H_

#include <FTypeDef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Precision of the C ABI's SimulateODE: 0 = plain FP32 with the reference's
 * rounding of M (default, parity mode), 1 = mixed (FP64 step map and powers,
 * FP64 top scan levels, FP32 lower levels), 2 = FP64 scan. */
void ParODE_SetPrecision(int mode);
int ParODE_GetPrecision(void);

/* Upper bound on the scan elements per batch (rounded down to a power of two);
 * 0 restores the automatic plan from free device memory.  For callers that
 * chain horizons explicitly and for tests of the batch carry. */
void ParODE_SetMaxBatchSteps(long maxSteps);

/* Kernel choices, for A/B checks and tuning (0 = automatic): bit 0 runs the
 * FP32 scan levels on the vector ALUs instead of the matrix cores, bit 1 the
 * FP64 GEMMs of the power table on the vector ALUs, bit 2 launches every level
 * separately instead of finishing the top of the tree in one work-group, bit 3
 * runs the lowest ABM levels as tree levels instead of chunk recurrences. */
void ParODE_SetLevelKernel(int variant);

/* Compiles a UFunc source for a gfx target (e.g. "gfx950") without a GPU.
 * Returns 0, or ParODE_CouldNotCreateGPUKernels with the compiler log. */
int ParODE_CompileUFunc(const char* uFunc, const char* arch, char* log, int logLen);

/* Text of the last exception caught at the ABI boundary. */
void ParODE_GetLastErrorMessage(char* buf, int len);

/* SimulateODE with sampled inputs instead of a registered UFunc: samples
 * holds u(tStart + s*tStep/2), s = 0 .. 2*points, nInputs per row, as float
 * (samplesDouble = 0) or double (1), on the host.  For inputs that are not a
 * pure function of float time (noise, measured signals) and for the FP64
 * precision path.  Other arguments as SimulateODE. */
void SimulateODESampled(fType A[], fType B[], fType C[], fType D[],
						int *nInputs, int *nStates, int *nOutputs,
						int *ZeroB, int *ZeroC, int *ZeroD,
						const void* samples, int *samplesDouble,
						double *tStart, double *tEnd, double *tStep,
						fType x0[], int *solver,
						fType tVect[], fType xVect[], int *nSteps, int nErr[]);

/* ---- device-resident plans: one system shape, inputs already in HBM ----
 * For serving loops and benchmarks that keep matrices, samples and outputs on
 * the GPU.  A plan owns its scan buffers; executing it launches the whole solve
 * (step map, FP64 powers, leaves, sweeps, outputs) asynchronously on `stream`
 * (a hipStream_t, or null for the default stream) with no host synchronisation.
 * Device arrays: A n*n, B n*m, C p*n, D p*m, x0 n (float, row-major);
 * samples (2*points+1)*m on the h/2 grid, float or double (fixed at creation);
 * Y points*nOut (float, time-major; nOut = n when C and D are both absent). */
int ParODE_PlanCreate(int solver, int nStates, int nInputs, int nOutputs,
					  int zeroC, int zeroD, long points, double tStep,
					  int precision, int samplesDouble, void** plan);
int ParODE_PlanExecute(void* plan, const float* dA, const float* dB, const float* dC,
					   const float* dD, const float* dx0, const void* dSamples,
					   float* dY, void* stream);
/* As ParODE_PlanExecute, then waits for the solve and writes the GPU time of
 * its phases to phaseMs[6]: step map + FP64 powers, leaves (and ABM start-up),
 * up-sweep launches, the top of the tree with the root (one work-group),
 * down-sweep launches, outputs (hipEvents on `stream`).  For
 * benchmarks and profilers. */
int ParODE_PlanExecuteTimed(void* plan, const float* dA, const float* dB, const float* dC,
							const float* dD, const float* dx0, const void* dSamples,
							float* dY, void* stream, float* phaseMs);
/* Records the whole solve with these device pointers into a HIP graph (on
 * `stream`, which must not be the null stream); ParODE_PlanReplay then
 * launches it as one graph.  For callers that repeat one solve (same shape,
 * same buffers, new contents) and want the launch cost of many small kernels
 * off the critical path.  A new capture replaces the previous one.  The graph
 * holds no memset nodes. */
int ParODE_PlanCapture(void* plan, const float* dA, const float* dB, const float* dC,
					   const float* dD, const float* dx0, const void* dSamples,
					   float* dY, void* stream);
int ParODE_PlanReplay(void* plan, void* stream);
long ParODE_PlanFootprint(void* plan);
int ParODE_PlanLevels(void* plan);
/* Autotuner: times each kernel that can run each scan level of this plan's
 * shape (K-split tiles, narrow or wide matrix-core tiles, SIMT) on the plan's
 * own trajectory buffer, on `stream`, and keeps the fastest in a process-wide
 * level schedule that every later solve of the same shape follows (plans,
 * the C ABI, served requests).  Returns the number of level launches tuned
 * (>= 0) or a negative ErrorCode.  Blocks until the timings are done; the
 * plan's previous outputs are not touched, its trajectory is overwritten.
 * For callers that solve one shape many times.  No counterpart in the
 * reference. */
int ParODE_PlanAutotune(void* plan, void* stream);
/* Forgets the autotuned schedule (every level back to the built-in rule). */
void ParODE_ScheduleClear(void);
/* Number of the plan's lowest scan levels that run as chunk recurrences
 * (0: all run as tree levels), -1 for a null plan.  For benchmarks that count
 * the work of each path. */
int ParODE_PlanChunkLog(void* plan);
int ParODE_PlanDestroy(void* plan);
/* Device memory the plan for this shape would hold, without allocating. */
long ParODE_PlanEstimateBytes(int solver, int nStates, int nInputs, long points, int precision);

/* ---- a solve that allocates no device memory ----
 * No counterpart in the reference.  For callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The same solve as ParODE_PlanCreate + ParODE_PlanExecute (same
 * kernels, same results), with the same arguments, but its scan buffers come
 * from a static device array of ParODE_NoAllocScratchBytes() bytes (128 MiB)
 * that each GPU holds until the process ends; a shape whose buffers
 * (ParODE_PlanEstimateBytes, outputs included) exceed it returns
 * ParODE_NotEnoughResources with the sizes in ParODE_GetLastErrorMessage.
 * Calls no hipMalloc or hipFree.  Every call on one GPU shares that GPU's
 * array: calls queued on one stream are safe, calls that overlap on one GPU
 * (other streams, other threads) race. */
int ParODE_SolveNoAlloc(int solver, int nStates, int nInputs, int nOutputs,
						int zeroC, int zeroD, long points, double tStep,
						int precision, int samplesDouble,
						const float* dA, const float* dB, const float* dC,
						const float* dD, const float* dx0, const void* dSamples,
						float* dY, void* stream);
long ParODE_NoAllocScratchBytes(void);

/* ---- grouped solves: many systems, each stage of the scan one launch ----
 * For batch callers with many independent systems of uneven state size and
 * horizon (and the engine of the server below).  Mixed precision.  Every
 * system's arrays live on the device: A n*n, B n*m, C p*n, D p*m, x0 n
 * (float); either dCoef[i] (m x 6 sinusoid-family coefficients, see the
 * server) with dU[i] receiving the FP64 samples, or dCoef NULL / dCoef[i]
 * NULL and dU[i] holding (2*points+1)*m FP64 samples on the h/2 grid; dY[i]
 * points*p outputs.  dC, dD, dCoef, tStart, sigma, seed may be NULL (absent,
 * zeros).  ParODE_GroupEstimateBytes gives -1 for a shape outside the grouped
 * kernels.  The group owns its scan buffers; executing it launches every
 * stage for all systems on `stream` without host synchronisation.
 * ExecuteTimed waits and writes the GPU time of each of ParODE_GroupStages()
 * stages (names from ParODE_GroupStageName). */
int ParODE_GroupCreate(int count, const int* solver, const int* nStates, const int* nInputs,
					   const int* nOutputs, const long* points, const double* tStart,
					   const double* tStep, const float* const* dA, const float* const* dB,
					   const float* const* dC, const float* const* dD, const float* const* dx0,
					   const double* const* dCoef, const double* sigma,
					   const unsigned long long* seed, double* const* dU, float* const* dY,
					   void** group);
int ParODE_GroupExecute(void* group, void* stream);
int ParODE_GroupExecuteTimed(void* group, void* stream, float* stageMs);
int ParODE_GroupStages(void);
const char* ParODE_GroupStageName(int stage);
int ParODE_GroupLaunches(void* group);
long ParODE_GroupFootprint(void* group);
long ParODE_GroupEstimateBytes(int solver, int nStates, int nInputs, int nOutputs, long points);
int ParODE_GroupDestroy(void* group);

/* ---- serving: continuous batching of independent solves on one GPU ----
 * For request loops with many systems of different shapes and horizons.  A
 * server owns a device arena (deviceBytes), a pinned staging area
 * (pinnedBytes), `lanes` HIP streams (1..16) and a worker thread.  Submitted
 * requests are admitted in arrival order whenever a lane has room and the
 * arena holds the request's footprint (ParODE_ServerFootprint).  With mixed
 * precision, the requests waiting at the head of the queue are admitted
 * together as one group whose scans run as grouped launches (each stage of
 * the scan one launch for all of them); a group's work runs on its lane with
 * no host synchronisation, so groups overlap on the device and arrivals are
 * admitted while others run.
 * The server copies A, B, C, D, x0 and coef at submission; samples (host,
 * (2*points+1) x nInputs doubles on the h/2 grid from tStart) and Y (host,
 * points x nOut floats) must stay valid until the ticket completes.  With
 * samples == NULL the inputs are the family u_j = a sin(w t + phi) +
 * b cos(w t + phi) + c + d t (+ sigma N(0,1) keyed by (seed, sample, j)),
 * coef row j = (a, b, c, d, w, phi), sampled in FP64 on the device.
 * precision as ParODE_SetPrecision (mixed = 1 is the serving default). */
int ParODE_ServerCreate(int lanes, long deviceBytes, long pinnedBytes, int precision,
						void** server);
int ParODE_ServerSubmit(void* server, int solver, int nStates, int nInputs, int nOutputs,
						int zeroC, int zeroD, long points, double tStart, double tStep,
						const float* A, const float* B, const float* C, const float* D,
						const float* x0, const double* coef, double sigma,
						unsigned long long seed, const double* samples, float* Y,
						long* ticket);
/* As ParODE_ServerSubmit, with the outputs written in place to dY, device
 * memory of points x nOut floats that stays valid until the ticket
 * completes: for serving loops whose consumers read the outputs on the GPU. */
int ParODE_ServerSubmitDevice(void* server, int solver, int nStates, int nInputs, int nOutputs,
							  int zeroC, int zeroD, long points, double tStart, double tStep,
							  const float* A, const float* B, const float* C, const float* D,
							  const float* x0, const double* coef, double sigma,
							  unsigned long long seed, const double* samples, float* dY,
							  long* ticket);
/* count device-output requests queued in one call, in order (element i of
 * each array describes request i; coef, sigma, seed and samples may be NULL
 * for all); tickets[i] receives request i's ticket.  For clients that hand
 * over a batch of systems at once: the worker sees the whole batch when it
 * next forms a group, and the per-request call overhead is paid once. */
int ParODE_ServerSubmitDeviceBatch(void* server, int count, const int* solver,
								   const int* nStates, const int* nInputs, const int* nOutputs,
								   const int* zeroC, const int* zeroD, const long* points,
								   const double* tStart, const double* tStep,
								   const float* const* A, const float* const* B,
								   const float* const* C, const float* const* D,
								   const float* const* x0, const double* const* coef,
								   const double* sigma, const unsigned long long* seed,
								   const double* const* samples, float* const* dY,
								   long* tickets);
/* Blocks until the ticket completes; returns its status.  timesMs (may be
 * NULL) receives [submission -> completion, submission -> admission]. */
int ParODE_ServerWait(void* server, long ticket, double* timesMs);
/* ParODE_ServerWait over count tickets, in order: status[i] receives ticket
 * i's status and timesMs[2 i], timesMs[2 i + 1] (timesMs may be NULL) its two
 * times.  Returns ParODE_OK once all have completed (each status says how), or
 * ParODE_InvalidSystem for bad arguments.  No counterpart in the reference;
 * for callers that hand a batch over with ParODE_ServerSubmitDeviceBatch. */
int ParODE_ServerWaitBatch(void* server, int count, const long* tickets, int* status,
						   double* timesMs);
int ParODE_ServerDrain(void* server);
/* stats[8]: submitted, completed, failed, running, queued, peak running,
 * arena bytes, peak arena bytes in use. */
int ParODE_ServerStats(void* server, long* stats);
/* stats[3]: groups admitted (one lane's work each), requests that ran in
 * grouped launches, the largest group. */
int ParODE_ServerGroupStats(void* server, long* stats);
long ParODE_ServerFootprint(int solver, int nStates, int nInputs, int nOutputs, long points,
							int precision);
int ParODE_ServerDestroy(void* server);

#ifdef __cplusplus
}
#endif

#endif /* PARODEEXT_H_ */
