/* The fixed-step scan driver (the reference's Include/GPUFixedStepIterative.h:
 * 13-129, Sources/GPUFixedStepIterative.cpp:131-466) with the same protected
 * hooks, called in the same order:
 *   InitializeSolverData -> BuildOutputObjects -> (GetInitialStatesAndInputs)
 *   -> per batch { InitializeBatchData, up-sweep, root carry, down-sweep,
 *   outputs } -> CleanSolverData.
 * Underneath, the per-element matrices of the reference are collapsed into one
 * power table P_l = M^(2^l) and the trajectory is scanned in place on one GPU;
 * the root carry between batches stays on the device. */
#ifndef PARODE_GPUFIXEDSTEPITERATIVE_H_
#define PARODE_GPUFIXEDSTEPITERATIVE_H_

#include <GPUODESolver.h>

#include <memory>
#include <vector>

namespace parode { class ScanCore; }

class GPUODESolverFixedStepIterative : public GPUODESolver {
protected:
	unsigned long _nTotalStepsSimulation;   /* output points */
	unsigned long _nSteps;                  /* scan elements (differs for multi-step methods) */
	std::vector<unsigned long> _nStepsPerDevice;   /* elements per batch (power of two) */
	unsigned int _nSystemSize;
	unsigned int _nMatrixSize;              /* scan state size n_a */
	unsigned int _nInputs;
	unsigned int _nOutputSize;
	int _bZeroC;
	int _bZeroD;
	/* launch-shape hints of the reference's planner (RungeKutta4thOrder.cpp:75-89);
	 * the HIP kernels size their own work-groups, these are kept for subclasses */
	size_t _nLocalSizeMM;
	size_t _nLocalSizeMVVColumn;
	size_t _nLocalSizeMVVRow;
	/* input / state strides and offsets of the output stage (SystemOutput.cl:18-21) */
	int _nInputStride;
	int _nInputOffset;
	int _nStateStride;
	int _nStateOffset;
	bool _bStrideSet;
	vectorf _X0;
	matrixf _M0;
	bool _bInitialConditionsSet;
	/* device state of the scan on the first selected device */
	std::unique_ptr<parode::ScanCore> _core;
	std::unique_ptr<DeviceBuffer> _BufferA, _BufferB, _BufferC, _BufferD, _BufferX0, _BufferYVectors;
	std::vector<DeviceBuffer*> _vectUBuffers;        /* input samples, one per device */
	long _nRealInBatch;                              /* real elements of the batch being built */
	long _nBatchIndex;
	matrixf _hostC, _hostD;                          /* host copies for the lead outputs */
	std::vector<double> _X0d;                        /* FP64 start-up states (precision modes) */
	std::vector<std::vector<double>> _UInitd;        /* FP64 start-up inputs (precision modes) */
	const void* _pBatchU;                            /* inputs of the batch being built */
	bool _bBatchUDouble;
	/* extensions: precision mode and sampled inputs (h/2 grid from tStart) */
	int _nPrecision;
	const void* _pSampledInput;
	bool _bSampledDouble;
	std::unique_ptr<DeviceBuffer> _BufferSampled;

	void SetStrides(int nInputStride, int nStateStride, int nInputOffset = 0, int nStateOffset = 0) {
		_bStrideSet = true;
		_nInputStride = nInputStride;
		_nStateStride = nStateStride;
		_nInputOffset = nInputOffset;
		_nStateOffset = nStateOffset;
	}
	void SetInitialConditions(vectorf& X0) {
		_X0 = X0;
		_M0.resize(_nMatrixSize, _nMatrixSize, false);
		for (unsigned r = 0; r < _nMatrixSize; ++r) _M0(r, r) = 1.0f;
		_bInitialConditionsSet = true;
	}
	void SetLocalSizeAttributes(size_t nLocalSizeMM, size_t nLocalSizeMVVColumn, size_t nLocalSizeMVVRow) {
		_nLocalSizeMM = nLocalSizeMM;
		_nLocalSizeMVVColumn = nLocalSizeMVVColumn;
		_nLocalSizeMVVRow = nLocalSizeMVVRow;
	}
	void SetNStepsPerDevice(std::vector<unsigned long>& nStepsPerDevice) { _nStepsPerDevice = nStepsPerDevice; }
	void SetNSteps(unsigned long nSteps) { _nSteps = nSteps; }
	void BuildOutputObjects(const matrixf& C, const matrixf& D);
	void DeleteGPUFSIObjects();
	/* shared planning: elements per batch from the HBM budget (power of two,
	 * shrunk to the horizon as RungeKutta4thOrder.cpp:126-142 does) */
	unsigned long PlanStepsPerBatch(unsigned long nSteps, std::size_t bytesPerElement,
									std::size_t fixedBytes) const;
	/* creates the device scan state and uploads A, B, x0 */
	void CreateScanCore(int method, const matrixf& A, const matrixf& B, double tStep,
						unsigned long nPadMax, const vectorf& x0);
	/* the input buffer of a batch: UFunc samples on the reference's float grid
	 * or a window of the sampled input */
	const void* PrepareInputs(double tFirst, double tStep, long nSamples, long firstHalfRow);

public:
	GPUODESolverFixedStepIterative(GPUManagement* GPUM, KernelTable& kernels,
								   std::vector<parode::UKernel*>& vectUKernels);
	virtual ~GPUODESolverFixedStepIterative();

	virtual ErrorCode SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C, const matrixf& D,
								  double tStart, double tEnd, double tStep,
								  const vectorf& x0, vectorf& tVect, matrixf& xVect);
	using GPUODESolver::SimulateODE;

	/* 0 = plain FP32 (reference parity), 1 = mixed precision, 2 = FP64 scan */
	void SetPrecision(int mode) { _nPrecision = mode; }
	/* host samples u(tStart + s h/2), s = 0..2*points, row-major by sample */
	void SetSampledInput(const void* hostSamples, bool isDouble) {
		_pSampledInput = hostSamples;
		_bSampledDouble = isDouble;
	}

protected:
	virtual void InitializeSolverData(const matrixf& A, const matrixf& B,
									  double& tStart, double tEnd, double tStep,
									  const vectorf& x0) = 0;
	virtual void InitializeBatchData(DeviceBuffer* pBufferMatrixSource,
									 DeviceBuffer* pBufferVectorSource,
									 int nDeviceIndex,
									 double tStart, double tStep, int nSteps) = 0;
	virtual void CleanSolverData() = 0;
	virtual void GetInitialStatesAndInputs(matrixf& xVectInit, matrixf& uVectInit);
};

#endif /* PARODE_GPUFIXEDSTEPITERATIVE_H_ */
