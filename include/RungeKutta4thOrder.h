/* Classic RK4 method plug-in (the reference's Include/RungeKutta4thOrder.h:7-33,
 * Sources/RungeKutta4thOrder.cpp): x_{k+1} = M x_k + B1 u(t_k) + B2 u(t_k+h/2)
 * + B3 u(t_k+h). */
#ifndef PARODE_RUNGEKUTTA4THORDER_H_
#define PARODE_RUNGEKUTTA4THORDER_H_

#include <GPUFixedStepIterative.h>

class RungeKutta4thOrder : public GPUODESolverFixedStepIterative {
protected:
	virtual void InitializeSolverData(const matrixf& A, const matrixf& B,
									  double& tStart, double tEnd, double tStep, const vectorf& x0);
	virtual void InitializeBatchData(DeviceBuffer* pBufferMatrixSource,
									 DeviceBuffer* pBufferVectorSource,
									 int nDeviceIndex,
									 double tStart, double tStep, int nSteps);
	virtual void CleanSolverData();

public:
	RungeKutta4thOrder(GPUManagement* GPUM, KernelTable& kernels,
					   std::vector<parode::UKernel*>& vectU);
	virtual ~RungeKutta4thOrder();
};

#endif /* PARODE_RUNGEKUTTA4THORDER_H_ */
