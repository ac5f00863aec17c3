/* The ParODE C++ API (the reference's Include/ParODECppInterface.h:17-57), with
 * the Boost-free matrixf / vectorf of ParODEMatrix.h. */
#ifndef PARODECPPINTERFACE_H_
#define PARODECPPINTERFACE_H_
#include <FTypeDef.h>
#include <ParODEMatrix.h>
#include <ParODEEnums.h>

#include <string>
#include <vector>

ErrorCode InitializeAllGPUs(std::string& strKernelFolder,
							std::string& strIncludeFolder);

ErrorCode GetAvailableGPUs(std::vector<unsigned int>& GPUIds);

ErrorCode GetDeviceName( unsigned int DeviceId, std::string& strDeviceName);

ErrorCode InitializeSelectedGPUs(const std::vector<unsigned int>& GPUIds,
								std::string& strKernelFolder,
								std::string& strIncludeFolder);

void CloseGPU();

ErrorCode StartTimer();
ErrorCode StopTimer(float &fTime);

void GetErrorDescription(ErrorCode ErrNo, std::string& strErrDescription);

ErrorCode SimulateODE(	const matrixf & A,
						const matrixf & B,
						const matrixf & C,
						const matrixf & D,
						int uIndex,
						double tStart, double tEnd, double tStep,
						const vectorf& x0,
						SolverType solver,
						vectorf& tVect,
						matrixf& xVect);

int RegisterInput(const std::string& uFunc);

#endif /* PARODECPPINTERFACE_H_ */
