// ParODE for SciLab: link libParODEShared.so (MI355X build) and call its C ABI.
//
// Same SciLab-side functions as the reference's SciLabInterface
// (LoadParODELibrary.sci, ParODEInterfaceStubs.sci, UnloadParODE.sci) over the
// same ten C symbols (include/ParODECInterface.h).  One deliberate difference:
// tStart/tEnd/tStep are passed as 'd' (double), which is what
// SimulateODE(double*, double*, double*) reads; the reference passes them as
// 'r' (4-byte real), so its C side read 8 bytes from 4-byte buffers.
//
//   exec('parode.sci', -1);
//   lib = LoadParODELibrary('/path/to/parode_amd/lib/');
//   ...
//   UnloadParODE(lib);

function ParODELibIndex = LoadParODELibrary(strPathParODELibrary)
    entries = ['GetErrorDescription', 'InitializeAllGPUs', 'GetAvailableGPUs', ...
               'GetDeviceName', 'InitializeSelectedGPUs', 'SimulateODE', ...
               'RegisterInput', 'CloseGPUC', 'StartTimer', 'StopTimer'];
    ParODELibIndex = link(strPathParODELibrary + 'libParODEShared.so', entries, 'c');
endfunction

function UnloadParODE(ParODELibIndex)
    ulink(ParODELibIndex);
endfunction

// void GetErrorDescription(int* nErr, char strErrDescription[]);
function s = GetErrorDescription(nErr)
    s = call('GetErrorDescription', nErr, 1, 'i', 'out', [1, 512], 2, 'c');
endfunction

// void InitializeAllGPUs(char* kernelFolder, char* includeFolder, int nErr[]);
// The folders are accepted and unused: the kernels are compiled into the library.
function nErr = InitializeAllGPUs(strKFolder, strKIncludeFolder)
    nErr = call('InitializeAllGPUs', strKFolder, 1, 'c', strKIncludeFolder, 2, 'c', ...
                'out', [1, 1], 3, 'i');
endfunction

// void GetAvailableGPUs(int nGPUs[], int GPUIds[], int nErr[]);
function [nGPUs, GPUIds, nErr] = GetAvailableGPUs()
    [nGPUs, GPUIds, nErr] = call('GetAvailableGPUs', 'out', [1, 1], 1, 'i', ...
                                 [1, 64], 2, 'i', [1, 1], 3, 'i');
    GPUIds = GPUIds(1:nGPUs);
endfunction

// void GetDeviceName(int* deviceId, char strDeviceName[], int nErr[]);
function [strDeviceName, nErr] = GetDeviceName(DeviceID)
    [strDeviceName, nErr] = call('GetDeviceName', DeviceID, 1, 'i', ...
                                 'out', [1, 256], 2, 'c', [1, 1], 3, 'i');
endfunction

// void InitializeSelectedGPUs(int* nGPUs, int GPUIds[], char*, char*, int nErr[]);
function nErr = InitializeSelectedGPUs(nGPUs, GPUIds, strKFolder, strKIncludeFolder)
    nErr = call('InitializeSelectedGPUs', nGPUs, 1, 'i', GPUIds, 2, 'i', ...
                strKFolder, 3, 'c', strKIncludeFolder, 4, 'c', 'out', [1, 1], 5, 'i');
endfunction

// void RegisterInput(char uFunc[], int uIndex[], int nErr[]);
function [uIndex, nErr] = RegisterInput(uFunc)
    [uIndex, nErr] = call('RegisterInput', uFunc, 1, 'c', 'out', [1, 1], 2, 'i', [1, 1], 3, 'i');
endfunction

// void CloseGPUC();
function CloseGPU()
    call('CloseGPUC');
endfunction

// void StartTimer(int nErr[]);  void StopTimer(float fTime[], int nErr[]);
function nErr = StartTimer()
    nErr = call('StartTimer', 'out', [1, 1], 1, 'i');
endfunction

function [fTime, nErr] = StopTimer()
    [fTime, nErr] = call('StopTimer', 'out', [1, 1], 1, 'r', [1, 1], 2, 'i');
endfunction

// Simulates nSteps points of x' = A x + B u, y = C x + D u from tStart.
// SciLab matrices are column-major and the C ABI takes row-major ones, so the
// system matrices go in transposed; xVect comes back as [nOutputs, nSteps]
// (row-major [nSteps][nOutputs] read column-major).  An empty B, C or D is
// absent (ZeroB/ZeroC/ZeroD).  tEnd is placed 1e-4 steps past the last point
// so the solver's truncating step count yields exactly nSteps points.
function [tVect, xVect, nErr] = SimulateODE(A, B, C, D, uIndex, tStart, tStep, nSteps, x0, solver)
    nStates = size(A, 1);
    ZeroB = bool2s(size(B, 1) == 0);
    ZeroC = bool2s(size(C, 1) == 0);
    ZeroD = bool2s(size(D, 1) == 0);
    if ZeroB == 0 then
        nInputs = size(B, 2);
    else
        nInputs = size(D, 2);
    end
    if ZeroC == 0 then
        nOutputs = size(C, 1);
    elseif ZeroD == 0 then
        nOutputs = size(D, 1);
    else
        nOutputs = nStates;
    end
    tEnd = tStart + (nSteps - 1) * tStep + 1e-4 * tStep;
    [tVect, xVect, nStepsOut, nErr] = call('SimulateODE', ...
        A.', 1, 'r', B.', 2, 'r', C.', 3, 'r', D.', 4, 'r', ...
        nInputs, 5, 'i', nStates, 6, 'i', nOutputs, 7, 'i', ...
        ZeroB, 8, 'i', ZeroC, 9, 'i', ZeroD, 10, 'i', uIndex, 11, 'i', ...
        tStart, 12, 'd', tEnd, 13, 'd', tStep, 14, 'd', ...
        x0, 15, 'r', solver, 16, 'i', ...
        'out', [1, nSteps], 17, 'r', [nOutputs, nSteps], 18, 'r', [1, 1], 19, 'i', [1, 1], 20, 'i');
endfunction
