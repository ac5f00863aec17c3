// Solver classes: GPUODESolver, GPUODESolverFixedStepIterative (the scan
// driver) and the RK4 / ABM4 method plug-ins, on one GPU through ScanCore.
#include <AdamsMoulton.h>
#include <GPUFixedStepIterative.h>
#include <GPUODESolver.h>
#include <RungeKutta4thOrder.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>

#include "config.h"
#include "hip_util.h"
#include "kernels.h"
#include "scan_core.h"
#include "ufunc.h"

using parode::ScanCore;
using parode::ScanOptions;

// ------------------------------------------------------------- GPUODESolver
GPUODESolver::GPUODESolver(GPUManagement* GPUM, KernelTable& kernels,
                           std::vector<parode::UKernel*>& vectUKernels)
    : _pGPUM(GPUM), _kernels(kernels), _vectUKernels(vectUKernels)
{
}

// GPUODESolver.cpp:25-44: the array form transposes to xVect[col * nOut + row].
ErrorCode GPUODESolver::SimulateODE(const matrixf& A, const matrixf& B, const matrixf& C,
                                    const matrixf& D, double tStart, double tEnd, double tStep,
                                    const vectorf& x0, fType tVect[], fType xVect[], int& nSteps)
{
    matrixf xVectLocal;
    vectorf tVectLocal;
    ErrorCode retVal = SimulateODE(A, B, C, D, tStart, tEnd, tStep, x0, tVectLocal, xVectLocal);
    if (retVal != ParODE_OK) {
        nSteps = 0;
        return retVal;
    }
    nSteps = (int)tVectLocal.size();
    const size_t nOutputs = xVectLocal.size1();
    for (int col = 0; col < nSteps; ++col) {
        for (size_t row = 0; row < nOutputs; ++row)
            xVect[(size_t)col * nOutputs + row] = xVectLocal(row, col);
        tVect[col] = tVectLocal[col];
    }
    return ParODE_OK;
}

// ------------------------------------------------------- fixed-step driver
GPUODESolverFixedStepIterative::GPUODESolverFixedStepIterative(
    GPUManagement* GPUM, KernelTable& kernels, std::vector<parode::UKernel*>& vectUKernels)
    : GPUODESolver(GPUM, kernels, vectUKernels), _nTotalStepsSimulation(0), _nSteps(0),
      _nSystemSize(0), _nMatrixSize(0), _nInputs(0), _nOutputSize(0), _bZeroC(1), _bZeroD(1),
      _nLocalSizeMM(0), _nLocalSizeMVVColumn(0), _nLocalSizeMVVRow(0), _nInputStride(1),
      _nInputOffset(0), _nStateStride(1), _nStateOffset(0), _bStrideSet(false),
      _bInitialConditionsSet(false), _nRealInBatch(0), _nBatchIndex(0), _pBatchU(nullptr),
      _bBatchUDouble(false), _nPrecision(parode::kPlain32), _pSampledInput(nullptr),
      _bSampledDouble(false)
{
}

GPUODESolverFixedStepIterative::~GPUODESolverFixedStepIterative() { DeleteGPUFSIObjects(); }

void GPUODESolverFixedStepIterative::DeleteGPUFSIObjects()
{
    _BufferC.reset();
    _BufferD.reset();
    _BufferYVectors.reset();
    _BufferA.reset();
    _BufferB.reset();
    _BufferX0.reset();
    _BufferSampled.reset();
    for (DeviceBuffer* b : _vectUBuffers) delete b;
    _vectUBuffers.clear();
    _core.reset();
}

unsigned long GPUODESolverFixedStepIterative::PlanStepsPerBatch(unsigned long nSteps,
                                                                 std::size_t bytesPerElement,
                                                                 std::size_t fixedBytes) const
{
    // RungeKutta4thOrder.cpp:91-142 sizes batches from 0.8 of device memory and
    // then shrinks them to the horizon; in LTI form an element costs one
    // trajectory column plus its inputs and outputs, not an n_a x n_a matrix.
    double budget = 0.8 * (double)_pGPUM->FreeMemory(0) - (double)fixedBytes;
    unsigned long P = 1;
    while (budget > 0 && (double)(P * 2) * (double)bytesPerElement <= budget) P <<= 1;
    long cap = parode::max_batch_steps();
    if (cap > 0) {
        unsigned long c = 1;
        while (c * 2 <= (unsigned long)cap) c <<= 1;
        P = std::min(P, c);
    }
    if (P >= nSteps) P = (unsigned long)parode::next_pow2((long)nSteps);
    return std::max(P, 1UL);
}

void GPUODESolverFixedStepIterative::CreateScanCore(int method, const matrixf& A, const matrixf& B,
                                                    double tStep, unsigned long nPadMax,
                                                    const vectorf& x0)
{
    const int dev = _pGPUM->DeviceId(0);
    hipStream_t s = _pGPUM->Stream(0);
    _pGPUM->SetCurrent(0);
    const int n = (int)A.size1(), m = (int)B.size2();
    ScanOptions opt;
    opt.precision = _nPrecision;
    _core.reset(new ScanCore(method, n, m, (int)_nOutputSize, _bZeroC == 1, _bZeroD == 1,
                             (long)nPadMax, opt));
    _BufferA.reset(new DeviceBuffer(sizeof(float) * n * n, dev));
    _BufferA->MemWrite(A.data(), s);
    _BufferB.reset(new DeviceBuffer(sizeof(float) * n * (m > 0 ? m : 1), dev));
    if (m > 0) _BufferB->MemWrite(B.data(), s, 0, sizeof(float) * n * m);
    _BufferX0.reset(new DeviceBuffer(sizeof(float) * n, dev));
    _BufferX0->MemWrite(x0.data(), s);
    _core->build_map(s, (const float*)_BufferA->Get(), (const float*)_BufferB->Get(), tStep);
    if (_pSampledInput && m > 0) {
        size_t rows = 2 * (size_t)_nTotalStepsSimulation + 1;
        size_t bytes = rows * m * (_bSampledDouble ? sizeof(double) : sizeof(float));
        _BufferSampled.reset(new DeviceBuffer(bytes, dev));
        _BufferSampled->MemWrite(_pSampledInput, s);
    }
}

const void* GPUODESolverFixedStepIterative::PrepareInputs(double tFirst, double tStep,
                                                          long nSamples, long firstHalfRow)
{
    _pBatchU = nullptr;
    _bBatchUDouble = false;
    if (_nInputs == 0) return nullptr;
    if (_BufferSampled) {
        size_t es = _bSampledDouble ? sizeof(double) : sizeof(float);
        _pBatchU = (const char*)_BufferSampled->Get() + (size_t)firstHalfRow * _nInputs * es;
        _bBatchUDouble = _bSampledDouble;
        return _pBatchU;
    }
    if (_vectUKernels.empty() || !_vectUKernels[0])
        throw std::invalid_argument("no input function registered for this simulation");
    size_t bytes = (size_t)nSamples * _nInputs * sizeof(float);
    if (_vectUBuffers.empty() || _vectUBuffers[0]->Size() < bytes) {
        for (DeviceBuffer* b : _vectUBuffers) delete b;
        _vectUBuffers.assign(1, new DeviceBuffer(bytes, _pGPUM->DeviceId(0)));
    }
    // u.cl:6-11 with fType tStart = tFirst, fType tStep = tStep (host casts,
    // RungeKutta4thOrder.cpp:255-256, AdamsMoulton.cpp:426-427)
    _vectUKernels[0]->sample(_pGPUM->Stream(0), (float*)_vectUBuffers[0]->Get(), (float)tFirst,
                             (float)tStep, nSamples, (int)_nInputs);
    _pBatchU = _vectUBuffers[0]->Get();
    return _pBatchU;
}

// GPUFixedStepIterative.cpp:58-125
void GPUODESolverFixedStepIterative::BuildOutputObjects(const matrixf& C, const matrixf& D)
{
    _bZeroD = D.size1() == 0;
    _bZeroC = C.size1() == 0;
    const int dev = _pGPUM->DeviceId(0);
    hipStream_t s = _pGPUM->Stream(0);
    if (C.size1() != 0 || D.size1() != 0) {
        if (C.size1() != 0 && C.size2() != _nSystemSize)
            throw std::invalid_argument("C has the wrong number of columns");
        if (D.size1() != 0 && _nInputs != 0 && D.size2() != _nInputs)
            throw std::invalid_argument("D has the wrong number of columns");
        if (C.size1() != 0 && D.size1() != 0 && C.size1() != D.size1())
            throw std::invalid_argument("C and D have different row counts");
        _nOutputSize = (unsigned)std::max(C.size1(), D.size1());
        _hostC = C;
        _hostD = D;
        _BufferC.reset(new DeviceBuffer(std::max<size_t>(1, C.size1() * C.size2()) * 4, dev));
        if (C.size1()) _BufferC->MemWrite(C.data(), s, 0, C.size1() * C.size2() * 4);
        _BufferD.reset(new DeviceBuffer(std::max<size_t>(1, D.size1() * D.size2()) * 4, dev));
        if (D.size1()) _BufferD->MemWrite(D.data(), s, 0, D.size1() * D.size2() * 4);
    } else {
        _nOutputSize = _nSystemSize;
    }
    _BufferYVectors.reset(new DeviceBuffer(
        sizeof(float) * _nOutputSize * std::max<unsigned long>(1, _nStepsPerDevice[0]), dev));
}

void GPUODESolverFixedStepIterative::GetInitialStatesAndInputs(matrixf&, matrixf&) {}

// GPUFixedStepIterative.cpp:131-466
ErrorCode GPUODESolverFixedStepIterative::SimulateODE(const matrixf& A, const matrixf& B,
                                                      const matrixf& C, const matrixf& D,
                                                      double tStart, double tEnd, double tStep,
                                                      const vectorf& x0, vectorf& tVect,
                                                      matrixf& xVect)
{
    try {
        // the output shape is needed by the scan state, so C/D are checked first
        _bZeroC = C.size1() == 0;
        _bZeroD = D.size1() == 0;
        _nOutputSize = (unsigned)((_bZeroC && _bZeroD) ? A.size1() : std::max(C.size1(), D.size1()));
        InitializeSolverData(A, B, tStart, tEnd, tStep, x0);
        BuildOutputObjects(C, D);
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        DeleteGPUFSIObjects();
        return ParODE_NotEnoughResources;
    }
    const unsigned long total = _nTotalStepsSimulation, lead = total - _nSteps;
    tVect.resize(total);
    for (unsigned long ii = 0; ii < lead; ++ii) tVect[ii] = (fType)(tStart - (lead - ii) * tStep);
    tVect[lead] = (fType)tStart;
    for (unsigned long ii = 1; ii < _nSteps; ++ii)
        tVect[lead + ii] = (fType)((double)tVect[lead + ii - 1] + tStep);
    xVect.resize(_nOutputSize, total, false);
    try {
        if (lead) {
            matrixf xVectInit, uVectInit;
            GetInitialStatesAndInputs(xVectInit, uVectInit);
            for (unsigned long ii = 0; ii < lead; ++ii)
                for (unsigned jj = 0; jj < _nOutputSize; ++jj) {
                    if (_bZeroC == 1 && _bZeroD == 1) {
                        xVect(jj, ii) = xVectInit(jj, ii);
                        continue;
                    }
                    if (_nPrecision != parode::kPlain32) {
                        // precision modes: the lead outputs in FP64, rounded once
                        double v = 0.0;
                        if (!_bZeroC)
                            for (unsigned kk = 0; kk < _nSystemSize; ++kk)
                                v += (double)_hostC(jj, kk) * (double)_X0d[ii * _nSystemSize + kk];
                        if (!_bZeroD)
                            for (unsigned kk = 0; kk < _nInputs; ++kk)
                                v += (double)_hostD(jj, kk) * _UInitd[ii * 2][kk];
                        xVect(jj, ii) = (fType)v;
                        continue;
                    }
                    // GPUFixedStepIterative.cpp:161-171: float sums, C x then D u
                    fType val = 0.0f;
                    if (!_bZeroC)
                        for (unsigned kk = 0; kk < _nSystemSize; ++kk) {
                            fType pr = _hostC(jj, kk) * xVectInit(kk, ii);
                            val += pr;
                        }
                    if (!_bZeroD)
                        for (unsigned kk = 0; kk < _nInputs; ++kk) {
                            fType pr = _hostD(jj, kk) * uVectInit(kk, ii);
                            val += pr;
                        }
                    xVect(jj, ii) = val;
                }
        }
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        DeleteGPUFSIObjects();
        CleanSolverData();
        return ParODE_InvalidSolverType;
    }
    try {
        hipStream_t s = _pGPUM->Stream(0);
        _pGPUM->SetCurrent(0);
        const unsigned long NSB = _nStepsPerDevice[0];
        const unsigned long nBatches = (_nSteps - 1) / NSB + 1;
        double* seedCur = _core->seed();
        double* seedNext = _core->seedNext();
        std::vector<float> yHost;
        unsigned long globalResultTIndex = lead;
        const float* dC = _BufferC ? (const float*)_BufferC->Get() : nullptr;
        const float* dD = _BufferD ? (const float*)_BufferD->Get() : nullptr;
        for (unsigned long nB = 0; nB < nBatches; ++nB) {
            long nReal = (long)NSB, nb = (long)NSB;
            if (nB == nBatches - 1) {
                nReal = (long)(_nSteps - nB * NSB);
                nb = parode::next_pow2(nReal);   // the last batch is padded (:205-221)
            }
            _nRealInBatch = nReal;
            _nBatchIndex = (long)nB;
            const double tStartBatch = tStart + (double)(nB * NSB) * tStep;
            const size_t na2 = (size_t)_nMatrixSize * _nMatrixSize;
            DeviceBuffer matrices((void*)_core->powerT(0), na2 * (_core->levels() + 1) * 8,
                                  _pGPUM->DeviceId(0));
            DeviceBuffer vectors(_core->X(), (size_t)nb * _nMatrixSize *
                                                 (_core->xIsDouble() ? 8 : 4),
                                 _pGPUM->DeviceId(0));
            InitializeBatchData(&matrices, &vectors, 0, tStartBatch, tStep, (int)nb);
            _core->sweep_up(s, nReal, nb);
            _core->sweep_top(s, nReal, nb, seedCur, _core->tot());
            const bool more = nB + 1 < nBatches;
            if (more) _core->carry(s, nb, seedCur, _core->tot(), seedNext);
            _core->sweep_down(s, nReal, nb);
            _core->outputs(s, dC, dD, _pBatchU, _bBatchUDouble, _nInputStride, _nInputOffset,
                           nReal, (float*)_BufferYVectors->Get());
            yHost.resize((size_t)nReal * _nOutputSize);
            _BufferYVectors->MemRead(yHost.data(), s, 0, yHost.size() * sizeof(float));
            for (long ii = 0; ii < nReal; ++ii)
                if (globalResultTIndex + ii < total)
                    for (unsigned jj = 0; jj < _nOutputSize; ++jj)
                        xVect(jj, globalResultTIndex + ii) = yHost[(size_t)ii * _nOutputSize + jj];
            globalResultTIndex += (unsigned long)nb;
            if (more) std::swap(seedCur, seedNext);
        }
    } catch (std::exception& e) {
        parode::last_error() = e.what();
        DeleteGPUFSIObjects();
        CleanSolverData();
        return ParODE_NotEnoughResources;
    }
    CleanSolverData();
    DeleteGPUFSIObjects();
    return ParODE_OK;
}

// -------------------------------------------------------------------- RK4
RungeKutta4thOrder::RungeKutta4thOrder(GPUManagement* GPUM, KernelTable& kernels,
                                       std::vector<parode::UKernel*>& vectU)
    : GPUODESolverFixedStepIterative(GPUM, kernels, vectU)
{
}

RungeKutta4thOrder::~RungeKutta4thOrder() {}

static void check_shapes(const matrixf& A, const matrixf& B, const vectorf& x0)
{
    if (A.size1() == 0 || A.size1() != A.size2())
        throw std::invalid_argument("A must be square and non-empty");
    if (B.size1() != 0 && B.size1() != A.size1())
        throw std::invalid_argument("B has the wrong number of rows");
    if (x0.size() != A.size1()) throw std::invalid_argument("x0 has the wrong size");
}

// RungeKutta4thOrder.cpp:14-211
void RungeKutta4thOrder::InitializeSolverData(const matrixf& A, const matrixf& B, double& tStart,
                                              double tEnd, double tStep, const vectorf& x0)
{
    check_shapes(A, B, x0);
    _nMatrixSize = (unsigned)A.size1();
    _nInputs = (unsigned)B.size2();
    _nSystemSize = _nMatrixSize;
    const double q = (tEnd - tStart) / tStep;
    if (!(q >= 0.0)) throw std::invalid_argument("tEnd precedes tStart");
    _nSteps = (unsigned long)q + 1;   // :20
    _nTotalStepsSimulation = _nSteps;
    // the reference's launch-shape planner (:75-89), kept as information
    size_t tile = (_nMatrixSize < 8) ? 4 : ((_nMatrixSize < 16) ? 8 : 16);
    size_t col = 1;
    while (col < 3 * _nInputs) col <<= 1;
    SetLocalSizeAttributes(tile, col, std::min<size_t>(1024 / col, _nMatrixSize));
    const size_t xBytes = _nPrecision == parode::kFp64 ? 8 : 4;
    const size_t uBytes = _pSampledInput && _bSampledDouble ? 8 : 4;
    const size_t perElem = _nMatrixSize * xBytes + 2 * _nInputs * uBytes + _nOutputSize * 4;
    ScanOptions opt;
    opt.precision = _nPrecision;
    const size_t fixed = parode::ScanCore::estimate_footprint(parode::kRK4, (int)_nMatrixSize,
                                                              (int)_nInputs, 1L << 20, opt,
                                                              (int)_nOutputSize) -
                         (size_t)(1L << 20) * _nMatrixSize * xBytes;
    std::vector<unsigned long> per(1, PlanStepsPerBatch(_nSteps, perElem, fixed));
    SetNStepsPerDevice(per);
    vectorf x0L(x0);
    SetInitialConditions(x0L);
    CreateScanCore(parode::kRK4, A, B, tStep, per[0], x0);
    parode::k::to_f64<float>(_pGPUM->Stream(0), _nMatrixSize, (const float*)_BufferX0->Get(),
                             _core->seed());
    SetStrides(2, 1);   // :209
}

// RungeKutta4thOrder.cpp:234-288
void RungeKutta4thOrder::InitializeBatchData(DeviceBuffer*, DeviceBuffer*, int, double tStart,
                                             double tStep, int nSteps)
{
    hipStream_t s = _pGPUM->Stream(0);
    if (_nBatchIndex == 0) _core->build_powers(s);   // replaces DistributeA
    const void* U = PrepareInputs(tStart, tStep / 2.0, 2L * nSteps + 1,
                                  2L * _nBatchIndex * (long)_nStepsPerDevice[0]);
    _core->leaves(s, U, _bBatchUDouble, 2, 0, 1, _nRealInBatch, nSteps);
}

void RungeKutta4thOrder::CleanSolverData() {}

// -------------------------------------------------------------------- ABM
AdamsMoulton::AdamsMoulton(GPUManagement* GPUM, KernelTable& kernels,
                           std::vector<parode::UKernel*>& vectU)
    : GPUODESolverFixedStepIterative(GPUM, kernels, vectU), _tOrigin(0.0)
{
}

AdamsMoulton::~AdamsMoulton() {}

// AdamsMoulton.cpp:29-104: nine input samples on the h/2 grid from tStart and
// three classic RK4 steps give X0 = [x0; x1; x2; x3] (computed on the device).
void AdamsMoulton::ComputeInitialState(const matrixf& A, const matrixf& B, const vectorf& x0,
                                       double fDeltaT, double tStart, vectorf& x0Scan)
{
    hipStream_t s = _pGPUM->Stream(0);
    const int n = (int)A.size1(), m = (int)B.size2();
    const void* Uh = PrepareInputs(tStart, fDeltaT / 2.0, 9, 0);
    _core->abm_startup(s, (const float*)_BufferA->Get(), (const float*)_BufferB->Get(),
                       (const float*)_BufferX0->Get(), Uh, _bBatchUDouble, fDeltaT, _core->seed());
    std::vector<double> X0(4 * (size_t)n);
    PARODE_HIP(hipMemcpyAsync(X0.data(), _core->seed(), X0.size() * 8, hipMemcpyDeviceToHost, s));
    _UInit.assign(9, vectorf(m));
    _UInitd.assign(9, std::vector<double>(m, 0.0));
    if (m > 0) {
        std::vector<char> raw(9 * (size_t)m * (_bBatchUDouble ? 8 : 4));
        PARODE_HIP(hipMemcpyAsync(raw.data(), Uh, raw.size(), hipMemcpyDeviceToHost, s));
        PARODE_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < m; ++j)
            {
                double v = _bBatchUDouble ? ((double*)raw.data())[i * m + j]
                                          : (double)((float*)raw.data())[i * m + j];
                _UInit[i][j] = (fType)v;
                _UInitd[i][j] = v;
            }
    }
    PARODE_HIP(hipStreamSynchronize(s));
    _X0d = X0;
    x0Scan.resize(4 * (size_t)n);
    for (size_t i = 0; i < X0.size(); ++i) x0Scan[i] = (fType)X0[i];
    (void)x0;
}

void AdamsMoulton::GetInitialStatesAndInputs(matrixf& xVectInit, matrixf& uVectInit)
{
    xVectInit.resize(_nSystemSize, 3, false);
    uVectInit.resize(_nInputs, 3, false);
    for (unsigned ii = 0; ii < 3; ++ii) {
        for (unsigned jj = 0; jj < _nSystemSize; ++jj) xVectInit(jj, ii) = _X0[ii * _nSystemSize + jj];
        for (unsigned jj = 0; jj < _nInputs; ++jj) uVectInit(jj, ii) = _UInit[ii * 2][jj];
    }
}

// AdamsMoulton.cpp:341-401
void AdamsMoulton::InitializeSolverData(const matrixf& A, const matrixf& B, double& tStart,
                                        double tEnd, double tStep, const vectorf& x0)
{
    check_shapes(A, B, x0);
    _nMatrixSize = 4 * (unsigned)A.size1();
    _nInputs = (unsigned)B.size2();
    _nSystemSize = (unsigned)A.size1();
    _tOrigin = tStart;
    const double q = (tEnd - (tStart + 3 * tStep)) / tStep;
    if (!(q >= 0.0))
        throw std::invalid_argument("the horizon is shorter than the ABM start-up (4 points)");
    _nSteps = (unsigned long)q + 1;          // :356-357
    _nTotalStepsSimulation = _nSteps + 3ul;  // :358
    size_t col = 1;
    while (col < _nMatrixSize) col <<= 1;
    SetLocalSizeAttributes((_nMatrixSize < 8) ? 4 : ((_nMatrixSize < 16) ? 8 : 16), col,
                           std::min<size_t>(1024 / col, _nMatrixSize));
    const size_t xBytes = _nPrecision == parode::kFp64 ? 8 : 4;
    const size_t uBytes = _pSampledInput && _bSampledDouble ? 8 : 4;
    const size_t perElem = _nMatrixSize * xBytes + 2 * _nInputs * uBytes + _nOutputSize * 4;
    ScanOptions opt;
    opt.precision = _nPrecision;
    const size_t fixed = parode::ScanCore::estimate_footprint(parode::kABM, (int)_nSystemSize,
                                                              (int)_nInputs, 1L << 20, opt,
                                                              (int)_nOutputSize) -
                         (size_t)(1L << 20) * _nMatrixSize * xBytes;
    std::vector<unsigned long> per(1, PlanStepsPerBatch(_nSteps, perElem, fixed));
    SetNStepsPerDevice(per);
    CreateScanCore(parode::kABM, A, B, tStep, per[0], x0);
    vectorf x0Scan;
    ComputeInitialState(A, B, x0, tStep, tStart, x0Scan);
    SetInitialConditions(x0Scan);
    tStart += 3 * tStep;                     // :356
    if (_pSampledInput)
        SetStrides(2, 4, 6, 3);              // h/2-grid samples: u(t_{e+3}) is row 2e + 6
    else
        SetStrides(1, 4, 3, 3);              // :398
}

// AdamsMoulton.cpp:403-461
void AdamsMoulton::InitializeBatchData(DeviceBuffer*, DeviceBuffer*, int, double tStart,
                                       double tStep, int nSteps)
{
    hipStream_t s = _pGPUM->Stream(0);
    if (_nBatchIndex == 0) _core->build_powers(s);
    if (_pSampledInput) {
        const void* U = PrepareInputs(0, 0, 0, 2L * _nBatchIndex * (long)_nStepsPerDevice[0]);
        _core->leaves(s, U, _bBatchUDouble, 2, 0, 2, _nRealInBatch, nSteps);
    } else {
        // u on the h grid from t - 3h (:426-428)
        const void* U = PrepareInputs(tStart - 3 * tStep, tStep, (long)nSteps + 4, 0);
        _core->leaves(s, U, false, 1, 0, 1, _nRealInBatch, nSteps);
    }
}

void AdamsMoulton::CleanSolverData() {}
