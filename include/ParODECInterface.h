/*
 * ParODECInterface.h -- the drop-in C ABI of ParODE, rebuilt for one AMD
 * MI355X (gfx950).  Same ten symbols, same signatures, same conventions as the
 * reference's Include/ParODECInterface.h:23-99 (implemented there at
 * Sources/ParODE.cpp:23-174):
 *   - every scalar passed by pointer; results and errors through nErr[0]
 *     (0 = OK, negative = ErrorCode from ParODEEnums.h); no exception crosses;
 *   - matrices row-major, copied by the library; ZeroB/ZeroC/ZeroD = 1 means
 *     "absent"; with C and D both absent the output is the state (n wide);
 *   - tVect / xVect are caller-allocated; xVect is time-major,
 *     xVect[k*nOut + j];
 *   - the folder strings are accepted for compatibility (kernels are linked
 *     into the library; the folder is not read).
 * Define C_INCLUDE before including this header from C (as the reference's
 * Python binding does, ParODESetup.py:8).
 */
#ifndef PARODECINTERFACE_H_
#define PARODECINTERFACE_H_

#include <FTypeDef.h>

#ifdef C_INCLUDE
#define C_EXPORT extern
#else
#define C_EXPORT
#endif

#ifndef C_INCLUDE
extern "C"
{
#endif
	/* Error text for ErrNo (reference: ParODECInterface.h:27, ParODE.cpp:25-33). */
	C_EXPORT void GetErrorDescription(int *ErrNo, char strErrDescription[]);
	/* Initialise every visible GPU (reference: ParODECInterface.h:30-32, ParODE.cpp:35-42). */
	C_EXPORT void InitializeAllGPUs(char* pszKernelFolder,
			 char* pszIncludeFolder,
			 int nErr[]);
	/* List the GPU ids (reference: ParODECInterface.h:39, ParODE.cpp:44-58). */
	C_EXPORT void GetAvailableGPUs( int nGPUs[], int GPUIds[], int nErr[]);
	/* Device name of a listed GPU (reference: ParODECInterface.h:46, ParODE.cpp:60-73). */
	C_EXPORT void GetDeviceName( int *DeviceId, char strDeviceName[], int nErr[]);
	/* Initialise a subset of the listed GPUs (reference: ParODECInterface.h:54-56,
	 * ParODE.cpp:75-90). */
	C_EXPORT void InitializeSelectedGPUs( int *nGPUs, int GPUIds[],
				char* pszKernelFolder,
				char* pszIncludeFolder, int nErr[]);
	/* Simulate x' = Ax + Bu, y = Cx + Du with RK4 (solver 0) or ABM4 (solver 1)
	 * (reference: ParODECInterface.h:61-79, ParODE.cpp:95-152). */
	C_EXPORT void SimulateODE(	fType A[],
					fType B[],
					fType C[],
					fType D[],
					int *nInputs,
					int *nStates,
					int *nOutputs,
					int *ZeroB,
					int *ZeroC,
					int *ZeroD,
					int *uIndex,
					double *tStart, double *tEnd, double *tStep,
					fType x0[],
					int *solver,
					fType tVect[],
					fType xVect[],
					int *nSteps,
					int nErr[]);
	/* Register an input function `fType UFunc(fType t, int nInputIndex)` given
	 * as source text (reference: ParODECInterface.h:86, ParODE.cpp:154-158). */
	C_EXPORT void RegisterInput(char uFunc[], int uIndex[], int nErr[]);
	/* Release every GPU resource (reference: ParODECInterface.h:89, ParODE.cpp:161-163). */
	C_EXPORT void CloseGPUC();
	/* GPU timer in milliseconds (reference: ParODECInterface.h:94-99, ParODE.cpp:166-173). */
	C_EXPORT void StartTimer(int nErr[]);
	C_EXPORT void StopTimer(float fTime[], int nErr[]);
#ifndef C_INCLUDE
}
#endif

#endif /* PARODECINTERFACE_H_ */
