/* The ParODE facade singleton (the reference's Include/ParODE.h:28-120,
 * Sources/ParODE.cpp:244-572): initialisation state machine
 * NOT -> PARTIAL (GetAvailableGPUs) -> FULL, input registry, solver factory,
 * timers and error text. */
#ifndef PARODE_PARODE_H_
#define PARODE_PARODE_H_

#include <GPUODESolver.h>
#include <GPUManagement.h>
#include <ParODEEnums.h>
#include <ParODEMatrix.h>

#include <memory>
#include <string>
#include <vector>

class ParODE {
public:
	enum InitializationState { NOT_INITIALIZED, PARTIAL_INITIALIZATION, FULL_INITIALIZATION };

	static ParODE* Instance();
	static void Delete();

	ErrorCode InitializeGPU(std::string& strKernelsFolder, std::string& strIncludeFolder);
	ErrorCode GetAvailableGPUs(std::vector<unsigned int>& vectIds);
	ErrorCode GetDeviceName(unsigned int device, std::string& deviceName);
	ErrorCode InitializeSelectedGPUs(const std::vector<unsigned int>& selectedDevices,
									 std::string& strKernelsFolder, std::string& strIncludeFolder);
	/* returns the input index, or a negative ErrorCode */
	int RegisterUKernel(const std::string& strUProgram);
	ErrorCode SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C, const matrixf& D,
						  int uIndex, double tStart, double tEnd, double tStep, const vectorf& x0,
						  SolverType solver, vectorf& tVect, matrixf& xVect);
	ErrorCode SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C, const matrixf& D,
						  int uIndex, double tStart, double tEnd, double tStep, const vectorf& x0,
						  SolverType solver, fType tVect[], fType xVect[], int& nSteps);
	/* sampled-input form (extension): samples u(tStart + s h/2) on the host */
	ErrorCode SimulateODESampled(const matrixf& A, const matrixf& B, const matrixf& C,
								 const matrixf& D, const void* samples, bool samplesDouble,
								 double tStart, double tEnd, double tStep, const vectorf& x0,
								 SolverType solver, fType tVect[], fType xVect[], int& nSteps);
	ErrorCode StartTimer();
	ErrorCode StopTimer(float& fTime);
	void GetErrorDescription(ErrorCode ErrNo, std::string& strErrDescription);

	void SetPrecision(int mode) { _nPrecision = mode; }
	int GetPrecision() const { return _nPrecision; }
	InitializationState GetState() const { return _InitializationState; }

	/* accessors the reference exposes to its kernel tests (ParODE.h:80-87) */
	GPUManagement* GetGPUManagement() { return _pGPUManagement.get(); }
	KernelTable& GetKernelsvectorReference() { return _kernels; }
	std::vector<parode::UKernel*> GetUKernelsVectorReference(int index);

private:
	ParODE();
	~ParODE();
	GPUODESolver* CreateSolver(SolverType solver, int uIndex);
	void Clean();

	static ParODE* _Instance;
	InitializationState _InitializationState;
	std::unique_ptr<GPUManagement> _pGPUManagement;
	KernelTable _kernels;
	/* registered inputs: one compiled sampler per selected device */
	std::vector<std::vector<std::shared_ptr<parode::UKernel>>> _vectKernelsInputs;
	std::string _strKernelFolderName;
	std::string _strIncludeFolder;
	int _nPrecision;
};

#endif /* PARODE_PARODE_H_ */
