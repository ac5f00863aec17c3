// The ParODE facade singleton, the ten-function C ABI and the C++ API
// (the reference's Sources/ParODE.cpp:1-572).
#include <ParODE.h>
#include <ParODECInterface.h>
#include <ParODECppInterface.h>
#include <ParODEExt.h>

#include <AdamsMoulton.h>
#include <RungeKutta4thOrder.h>

#include <cstring>
#include <iostream>

#include "config.h"
#include "hip_util.h"
#include "scan_core.h"
#include "ufunc.h"

namespace parode {
long& max_batch_steps()
{
    static long v = 0;
    return v;
}
// One text per thread: the server's worker threads and the callers of the
// C ABI set it concurrently, and each caller reads back its own.
std::string& last_error()
{
    static thread_local std::string s;
    return s;
}
int& level_kernel()
{
    static int v = 0;
    return v;
}
}  // namespace parode

// --------------------------------------------------------------- singleton
ParODE* ParODE::_Instance = nullptr;

ParODE* ParODE::Instance()
{
    if (_Instance == nullptr) _Instance = new ParODE;
    return _Instance;
}

void ParODE::Delete()
{
    if (_Instance != nullptr) {
        delete _Instance;
        _Instance = nullptr;
    }
}

ParODE::ParODE() : _InitializationState(NOT_INITIALIZED), _nPrecision(parode::kPlain32) {}

ParODE::~ParODE() { Clean(); }

void ParODE::Clean()
{
    _vectKernelsInputs.clear();
    _pGPUManagement.reset();
    _kernels = KernelTable();
    _InitializationState = NOT_INITIALIZED;
}

// ParODE.cpp:343-361
ErrorCode ParODE::InitializeGPU(std::string& strKernelsFolder, std::string& strIncludeFolder)
{
    Clean();
    _strKernelFolderName = strKernelsFolder;
    _strIncludeFolder = strIncludeFolder;
    try {
        _pGPUManagement.reset(new GPUManagement());
        if (!_pGPUManagement->Initialize(1)) {
            _pGPUManagement.reset();
            return ParODE_NoGPU;
        }
        _kernels.nDevices = _pGPUManagement->GetNumberOfDevices();
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        Clean();
        return ParODE_CouldNotCreateGPUKernels;
    }
    _InitializationState = FULL_INITIALIZATION;
    return ParODE_OK;
}

// ParODE.cpp:516-523
ErrorCode ParODE::GetAvailableGPUs(std::vector<unsigned int>& vectIds)
{
    Clean();
    _pGPUManagement.reset(new GPUManagement());
    _pGPUManagement->GetGPUDevices(vectIds);
    _InitializationState = PARTIAL_INITIALIZATION;
    return ParODE_OK;
}

// ParODE.cpp:525-532
ErrorCode ParODE::GetDeviceName(unsigned int device, std::string& deviceName)
{
    if (_InitializationState == NOT_INITIALIZED) return ParODE_NotInitialized;
    deviceName = _pGPUManagement->GetDeviceName(device);
    if (deviceName.empty()) return ParODE_InvalidDeviceId;
    return ParODE_OK;
}

// ParODE.cpp:534-553
ErrorCode ParODE::InitializeSelectedGPUs(const std::vector<unsigned int>& selectedDevices,
                                         std::string& strKernelsFolder,
                                         std::string& strIncludeFolder)
{
    if (_InitializationState != PARTIAL_INITIALIZATION) return ParODE_NotInitialized;
    try {
        _strKernelFolderName = strKernelsFolder;
        _strIncludeFolder = strIncludeFolder;
        if (!_pGPUManagement->Initialize(selectedDevices, 1)) return ParODE_NoGPU;
        _kernels.nDevices = _pGPUManagement->GetNumberOfDevices();
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        Clean();
        return ParODE_CouldNotCreateGPUKernels;
    }
    _InitializationState = FULL_INITIALIZATION;
    return ParODE_OK;
}

// ParODE.cpp:410-466: the input source is compiled once per device architecture
// and loaded on every selected device; the index is returned.
int ParODE::RegisterUKernel(const std::string& strUProgram)
{
    if (_InitializationState != FULL_INITIALIZATION) return ParODE_NotInitialized;
    std::vector<std::shared_ptr<parode::UKernel>> perDevice;
    try {
        std::vector<char> code;
        std::string log, lastArch;
        for (int i = 0; i < _pGPUManagement->GetNumberOfDevices(); ++i) {
            std::string arch = _pGPUManagement->Arch(i);
            if (arch != lastArch) {
                if (!parode::UKernel::compile(strUProgram, arch, code, log)) {
                    parode::last_error() = log;
                    std::cerr << "UFunc build failed:\n" << log << std::endl;
                    return ParODE_CouldNotCreateGPUKernels;
                }
                lastArch = arch;
            }
            _pGPUManagement->SetCurrent(i);
            perDevice.push_back(std::make_shared<parode::UKernel>(code));
        }
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_CouldNotCreateGPUKernels;
    }
    _vectKernelsInputs.push_back(perDevice);
    return (int)_vectKernelsInputs.size() - 1;
}

std::vector<parode::UKernel*> ParODE::GetUKernelsVectorReference(int index)
{
    std::vector<parode::UKernel*> v;
    if (index >= 0 && index < (int)_vectKernelsInputs.size())
        for (auto& p : _vectKernelsInputs[index]) v.push_back(p.get());
    return v;
}

// ParODE.cpp:281-304.  An index of -1 creates a solver without a UFunc (for
// sampled inputs).
GPUODESolver* ParODE::CreateSolver(SolverType solver, int uIndex)
{
    std::vector<parode::UKernel*> u = GetUKernelsVectorReference(uIndex);
    GPUODESolverFixedStepIterative* p = nullptr;
    switch (solver) {
    case RUNGE_KUTTA: p = new RungeKutta4thOrder(_pGPUManagement.get(), _kernels, u); break;
    case ADAMS_BASHFORTH_MOULTON: p = new AdamsMoulton(_pGPUManagement.get(), _kernels, u); break;
    default: return nullptr;
    }
    p->SetPrecision(_nPrecision);
    return p;
}

static bool valid_input_index(int uIndex, int count, const matrixf& B)
{
    if (B.size2() == 0) return true;          // no inputs: no UFunc needed
    return uIndex >= 0 && uIndex < count;
}

ErrorCode ParODE::SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C,
                              const matrixf& D, int uIndex, double tStart, double tEnd,
                              double tStep, const vectorf& x0, SolverType solver, fType tVect[],
                              fType xVect[], int& nSteps)
{
    if (_InitializationState != FULL_INITIALIZATION) return ParODE_NotInitialized;
    if (!valid_input_index(uIndex, (int)_vectKernelsInputs.size(), B)) return ParODE_InvalidSystem;
    std::unique_ptr<GPUODESolver> pSolver(CreateSolver(solver, uIndex));
    if (!pSolver) return ParODE_InvalidSolverType;
    return pSolver->SimulateODE(A, B, C, D, tStart, tEnd, tStep, x0, tVect, xVect, nSteps);
}

ErrorCode ParODE::SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C,
                              const matrixf& D, int uIndex, double tStart, double tEnd,
                              double tStep, const vectorf& x0, SolverType solver, vectorf& tVect,
                              matrixf& xVect)
{
    if (_InitializationState != FULL_INITIALIZATION) return ParODE_NotInitialized;
    if (!valid_input_index(uIndex, (int)_vectKernelsInputs.size(), B)) return ParODE_InvalidSystem;
    std::unique_ptr<GPUODESolver> pSolver(CreateSolver(solver, uIndex));
    if (!pSolver) return ParODE_InvalidSolverType;
    return pSolver->SimulateODE(A, B, C, D, tStart, tEnd, tStep, x0, tVect, xVect);
}

ErrorCode ParODE::SimulateODESampled(const matrixf& A, const matrixf& B, const matrixf& C,
                                     const matrixf& D, const void* samples, bool samplesDouble,
                                     double tStart, double tEnd, double tStep, const vectorf& x0,
                                     SolverType solver, fType tVect[], fType xVect[], int& nSteps)
{
    if (_InitializationState != FULL_INITIALIZATION) return ParODE_NotInitialized;
    std::unique_ptr<GPUODESolver> pSolver(CreateSolver(solver, -1));
    if (!pSolver) return ParODE_InvalidSolverType;
    if (B.size2() != 0) {
        if (!samples) return ParODE_InvalidSystem;
        static_cast<GPUODESolverFixedStepIterative*>(pSolver.get())
            ->SetSampledInput(samples, samplesDouble);
    }
    return pSolver->SimulateODE(A, B, C, D, tStart, tEnd, tStep, x0, tVect, xVect, nSteps);
}

// ParODE.cpp:252-276
ErrorCode ParODE::StartTimer()
{
    if (_InitializationState != FULL_INITIALIZATION) return ParODE_NotInitialized;
    try {
        _pGPUManagement->StartTimer();
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
    return ParODE_OK;
}

ErrorCode ParODE::StopTimer(float& fTime)
{
    if (_InitializationState != FULL_INITIALIZATION) return ParODE_NotInitialized;
    try {
        fTime = _pGPUManagement->StopTimer();
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_NotEnoughResources;
    }
    return ParODE_OK;
}

// ParODE.cpp:496-514 (same texts; codes below -3 read "Unknown Error" there too)
void ParODE::GetErrorDescription(ErrorCode ErrNo, std::string& strErrDescription)
{
    switch (ErrNo) {
    case ParODE_OK: strErrDescription = "No Error"; break;
    case ParODE_NoGPU:
        strErrDescription =
            "Sorry, no GPU available. Please make sure that you have installed a GPU with shared "
            "memory and you have installed the proper OpenCL drivers installed.";
        break;
    case ParODE_NotInitialized:
        strErrDescription = "ParODE not initialized. Please call InitializeGPU.";
        break;
    case ParODE_CouldNotCreateGPUKernels:
        strErrDescription = "Could not create GPU kernels; not enough resources.";
        break;
    default: strErrDescription = "Unknown Error"; break;
    }
}

// ------------------------------------------------------------------ marshal
static void to_matrix(const fType* src, int rows, int cols, matrixf& dst)
{
    dst.resize(rows, cols, false);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) dst(r, c) = src[(size_t)r * cols + c];
}

// ParODE.cpp:114-145
static bool marshal(fType A[], fType B[], fType C[], fType D[], int* nInputs, int* nStates,
                    int* nOutputs, int* ZeroB, int* ZeroC, int* ZeroD, fType x0[], matrixf& Ab,
                    matrixf& Bb, matrixf& Cb, matrixf& Db, vectorf& x0b)
{
    if (*nStates <= 0) return false;
    to_matrix(A, *nStates, *nStates, Ab);
    if (*ZeroB != 1) to_matrix(B, *nStates, *nInputs, Bb);
    if (*ZeroC != 1) to_matrix(C, *nOutputs, *nStates, Cb);
    if (*ZeroD != 1) to_matrix(D, *nOutputs, *nInputs, Db);
    x0b.resize(*nStates);
    for (int i = 0; i < *nStates; ++i) x0b[i] = x0[i];
    return true;
}

// --------------------------------------------------------------------- C ABI
extern "C" {

void GetErrorDescription(int* ErrNo, char strErrDescription[])
{
    std::string s;
    ParODE::Instance()->GetErrorDescription((ErrorCode)(*ErrNo), s);
    std::strcpy(strErrDescription, s.c_str());
}

void InitializeAllGPUs(char* pszKernelFolder, char* pszIncludeFolder, int nErr[])
{
    try {
        std::string k(pszKernelFolder ? pszKernelFolder : ""), i(pszIncludeFolder ? pszIncludeFolder : "");
        nErr[0] = ParODE::Instance()->InitializeGPU(k, i);
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        nErr[0] = ParODE_NoGPU;
    }
}

void GetAvailableGPUs(int nGPUs[], int GPUIds[], int nErr[])
{
    try {
        std::vector<unsigned int> ids;
        ErrorCode r = ParODE::Instance()->GetAvailableGPUs(ids);
        if (r != ParODE_OK) {
            nGPUs[0] = 0;
            nErr[0] = r;
            return;
        }
        nGPUs[0] = (int)ids.size();
        for (size_t i = 0; i < ids.size(); ++i) GPUIds[i] = (int)ids[i];
        nErr[0] = ParODE_OK;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        nGPUs[0] = 0;
        nErr[0] = ParODE_NoGPU;
    }
}

void GetDeviceName(int* DeviceId, char strDeviceName[], int nErr[])
{
    std::string name;
    ErrorCode r = (*DeviceId < 0) ? ParODE_InvalidDeviceId
                                  : ParODE::Instance()->GetDeviceName((unsigned)*DeviceId, name);
    if (r != ParODE_OK) {
        strDeviceName[0] = 0;
        nErr[0] = r;
        return;
    }
    std::strcpy(strDeviceName, name.c_str());
    nErr[0] = ParODE_OK;
}

void InitializeSelectedGPUs(int* nGPUs, int GPUIds[], char* pszKernelFolder,
                            char* pszIncludeFolder, int nErr[])
{
    try {
        std::vector<unsigned int> sel;
        for (int i = 0; i < *nGPUs; ++i) {
            if (GPUIds[i] < 0) {
                nErr[0] = ParODE_InvalidDeviceId;
                return;
            }
            sel.push_back((unsigned)GPUIds[i]);
        }
        std::string k(pszKernelFolder ? pszKernelFolder : ""), i(pszIncludeFolder ? pszIncludeFolder : "");
        nErr[0] = ParODE::Instance()->InitializeSelectedGPUs(sel, k, i);
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        nErr[0] = ParODE_CouldNotCreateGPUKernels;
    }
}

void SimulateODE(fType A[], fType B[], fType C[], fType D[], int* nInputs, int* nStates,
                 int* nOutputs, int* ZeroB, int* ZeroC, int* ZeroD, int* uIndex, double* tStart,
                 double* tEnd, double* tStep, fType x0[], int* solver, fType tVect[],
                 fType xVect[], int nSteps[], int nErr[])
{
    try {
        matrixf Ab, Bb, Cb, Db;
        vectorf x0b;
        if (!marshal(A, B, C, D, nInputs, nStates, nOutputs, ZeroB, ZeroC, ZeroD, x0, Ab, Bb, Cb,
                     Db, x0b)) {
            nErr[0] = ParODE_InvalidSystem;       // ParODE.cpp:116-119
            return;
        }
        nErr[0] = ParODE::Instance()->SimulateODE(Ab, Bb, Cb, Db, *uIndex, *tStart, *tEnd, *tStep,
                                                  x0b, (SolverType)(*solver), tVect, xVect,
                                                  nSteps[0]);
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        nErr[0] = ParODE_NotEnoughResources;
    }
}

void RegisterInput(char uFunc[], int uIndex[], int nErr[])
{
    try {
        uIndex[0] = ParODE::Instance()->RegisterUKernel(std::string(uFunc));
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        uIndex[0] = ParODE_CouldNotCreateGPUKernels;
    }
    nErr[0] = uIndex[0] >= 0 ? 0 : uIndex[0];
}

void CloseGPUC() { ParODE::Delete(); }

void StartTimer(int nErr[]) { nErr[0] = ParODE::Instance()->StartTimer(); }

void StopTimer(float fTime[], int nErr[]) { nErr[0] = ParODE::Instance()->StopTimer(fTime[0]); }

// ------------------------------------------------------------- extensions
void ParODE_SetPrecision(int mode)
{
    if (mode >= parode::kPlain32 && mode <= parode::kFp64) ParODE::Instance()->SetPrecision(mode);
}

int ParODE_GetPrecision(void) { return ParODE::Instance()->GetPrecision(); }

void ParODE_SetLevelKernel(int variant) { parode::level_kernel() = variant & 15; }

void ParODE_SetMaxBatchSteps(long maxSteps) { parode::max_batch_steps() = maxSteps > 0 ? maxSteps : 0; }

int ParODE_CompileUFunc(const char* uFunc, const char* arch, char* log, int logLen)
{
    try {
        std::vector<char> code;
        std::string lg;
        bool ok = parode::UKernel::compile(uFunc ? uFunc : "", arch ? arch : "gfx950", code, lg);
        if (log && logLen > 0) {
            std::strncpy(log, lg.c_str(), (size_t)logLen - 1);
            log[logLen - 1] = 0;
        }
        return ok ? ParODE_OK : ParODE_CouldNotCreateGPUKernels;
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        return ParODE_CouldNotCreateGPUKernels;
    }
}

void ParODE_GetLastErrorMessage(char* buf, int len)
{
    if (!buf || len <= 0) return;
    std::strncpy(buf, parode::last_error().c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
}

void SimulateODESampled(fType A[], fType B[], fType C[], fType D[], int* nInputs, int* nStates,
                        int* nOutputs, int* ZeroB, int* ZeroC, int* ZeroD, const void* samples,
                        int* samplesDouble, double* tStart, double* tEnd, double* tStep,
                        fType x0[], int* solver, fType tVect[], fType xVect[], int* nSteps,
                        int nErr[])
{
    try {
        matrixf Ab, Bb, Cb, Db;
        vectorf x0b;
        if (!marshal(A, B, C, D, nInputs, nStates, nOutputs, ZeroB, ZeroC, ZeroD, x0, Ab, Bb, Cb,
                     Db, x0b)) {
            nErr[0] = ParODE_InvalidSystem;
            return;
        }
        nErr[0] = ParODE::Instance()->SimulateODESampled(Ab, Bb, Cb, Db, samples, *samplesDouble != 0,
                                                         *tStart, *tEnd, *tStep, x0b,
                                                         (SolverType)(*solver), tVect, xVect,
                                                         nSteps[0]);
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        nErr[0] = ParODE_NotEnoughResources;
    }
}

}  // extern "C"

// ----------------------------------------------------------------- C++ API
void GetErrorDescription(ErrorCode ErrNo, std::string& strErrDescription)
{
    ParODE::Instance()->GetErrorDescription(ErrNo, strErrDescription);
}

ErrorCode InitializeAllGPUs(std::string& strKernelFolder, std::string& strIncludeFolder)
{
    return ParODE::Instance()->InitializeGPU(strKernelFolder, strIncludeFolder);
}

ErrorCode GetAvailableGPUs(std::vector<unsigned int>& GPUIds)
{
    return ParODE::Instance()->GetAvailableGPUs(GPUIds);
}

ErrorCode GetDeviceName(unsigned int DeviceId, std::string& strDeviceName)
{
    return ParODE::Instance()->GetDeviceName(DeviceId, strDeviceName);
}

ErrorCode InitializeSelectedGPUs(const std::vector<unsigned int>& GPUIds,
                                 std::string& strKernelFolder, std::string& strIncludeFolder)
{
    return ParODE::Instance()->InitializeSelectedGPUs(GPUIds, strKernelFolder, strIncludeFolder);
}

ErrorCode SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C, const matrixf& D,
                      int uIndex, double tStart, double tEnd, double tStep, const vectorf& x0,
                      SolverType solver, vectorf& tVect, matrixf& xVect)
{
    return ParODE::Instance()->SimulateODE(A, B, C, D, uIndex, tStart, tEnd, tStep, x0, solver,
                                           tVect, xVect);
}

int RegisterInput(const std::string& uFunc) { return ParODE::Instance()->RegisterUKernel(uFunc); }

void CloseGPU() { ParODE::Delete(); }

ErrorCode StartTimer() { return ParODE::Instance()->StartTimer(); }

ErrorCode StopTimer(float& fTime) { return ParODE::Instance()->StopTimer(fTime); }
