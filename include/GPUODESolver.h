/* Solver base class (the reference's Include/GPUODESolver.h:11-47).  The
 * constructor keeps its three-argument shape with HIP-backed types: the device
 * layer, the built-in kernel table and the registered input sampler (one per
 * device). */
#ifndef PARODE_GPUODESOLVER_H_
#define PARODE_GPUODESOLVER_H_

#include <FTypeDef.h>
#include <GPUManagement.h>
#include <ParODEEnums.h>
#include <ParODEMatrix.h>

#include <string>
#include <vector>

namespace parode { class UKernel; }

/* The reference's per-device kernel table (ParODE.h:105-115) names the ten
 * built-in kernels; here they are linked into the library, so the table only
 * records which device set they serve. */
struct KernelTable {
	int nDevices = 0;
};

class GPUODESolver {
public:
	GPUODESolver(GPUManagement* GPUM, KernelTable& kernels,
				 std::vector<parode::UKernel*>& vectUKernels);
	virtual ~GPUODESolver() {}

	virtual ErrorCode SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C, const matrixf& D,
								  double tStart, double tEnd, double tStep,
								  const vectorf& x0, vectorf& tVect, matrixf& xVect) = 0;
	/* array form: xVect[col * nOutputs + row] (GPUODESolver.cpp:25-44) */
	virtual ErrorCode SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C, const matrixf& D,
								  double tStart, double tEnd, double tStep,
								  const vectorf& x0, fType tVect[], fType xVect[], int& nSteps);

protected:
	GPUManagement* _pGPUM;
	KernelTable _kernels;
	std::vector<parode::UKernel*> _vectUKernels;
};

#endif /* PARODE_GPUODESOLVER_H_ */
