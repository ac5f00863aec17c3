/* Four-step Adams-Bashforth-Moulton PECE method plug-in (the reference's
 * Include/AdamsMoulton.h:6-53, Sources/AdamsMoulton.cpp): the 4n companion
 * state [x_{k-3}; x_{k-2}; x_{k-1}; x_k], started by three classic RK4 steps. */
#ifndef PARODE_ADAMSMOULTON_H_
#define PARODE_ADAMSMOULTON_H_

#include <GPUFixedStepIterative.h>

class AdamsMoulton : public GPUODESolverFixedStepIterative {
private:
	std::vector<vectorf> _UInit;   /* the 9 start-up input samples (h/2 grid) */
	double _tOrigin;               /* tStart before the start-up shift */
	void ComputeInitialState(const matrixf& A, const matrixf& B, const vectorf& x0,
							 double fDeltaT, double tStart, vectorf& x0Scan);

protected:
	virtual void InitializeSolverData(const matrixf& A, const matrixf& B,
									  double& tStart, double tEnd, double tStep, const vectorf& x0);
	virtual void InitializeBatchData(DeviceBuffer* pBufferMatrixSource,
									 DeviceBuffer* pBufferVectorSource,
									 int nDeviceIndex,
									 double tStart, double tStep, int nSteps);
	virtual void CleanSolverData();
	virtual void GetInitialStatesAndInputs(matrixf& xVectInit, matrixf& uVectInit);

public:
	AdamsMoulton(GPUManagement* GPUM, KernelTable& kernels, std::vector<parode::UKernel*>& vectU);
	virtual ~AdamsMoulton();
};

#endif /* PARODE_ADAMSMOULTON_H_ */
