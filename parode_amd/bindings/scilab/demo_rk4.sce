// Two decoupled first-order lags driven by u = t, solved with RK4 on the GPU.
// Run from this folder:  scilab-cli -f demo_rk4.sce
// (set PARODE_LIB to the folder holding libParODEShared.so, with a trailing /)
libDir = getenv('PARODE_LIB', '../../lib/');
exec('parode.sci', -1);
lib = LoadParODELibrary(libDir);

[nGPUs, ids, nErr] = GetAvailableGPUs();
if nErr <> 0 then
    disp(GetErrorDescription(nErr));
    abort;
end
for k = 1:nGPUs
    [name, e] = GetDeviceName(ids(k));
    mprintf('GPU %d: %s\n', ids(k), name);
end
if InitializeSelectedGPUs(1, ids(1), libDir, libDir) <> 0 then
    disp('ParODE initialisation failed');
    abort;
end

uFunc = 'fType UFunc(fType t, int nInputIndex){ return nInputIndex < 2 ? t : 0.0f; }';
[uIndex, nErr] = RegisterInput(uFunc);

A = [-2, 0; 0, -2];
B = eye(2, 2);
C = eye(2, 2);
D = eye(2, 2);
x0 = [1, -1];
StartTimer();
[tVect, xVect, nErr] = SimulateODE(A, B, C, D, uIndex, 0.0, 0.1, 50, x0, 0);
[ms, e] = StopTimer();
if nErr == 0 then
    mprintf('RK4: %d points in %.3f ms, y(end) = [%g %g]\n', size(tVect, 2), ms, xVect(1, $), xVect(2, $));
else
    disp(GetErrorDescription(nErr));
end

CloseGPU();
UnloadParODE(lib);
