/*
 * ParODE -- the Python 3 extension module over the ParODE C ABI.
 *
 * Same module name, method names, argument tuples and return tuples as the
 * reference's Python 2 binding (PythonInterface/ParODEmodule.c:450-462):
 *   GetErrorDescription(nErr) -> str
 *   InitializeAllGPUs() -> nErr
 *   GetAvailableGPUs() -> ([ids], nErr)
 *   GetDeviceName(id) -> (name, nErr)
 *   InitializeSelectedGPUs([ids]) -> nErr
 *   SimulateODE(A, B, C, D, uIndex, tStart, tStep, nSteps, x0, solver)
 *       -> (tVect, xVect[nSteps][nOut], nErr)
 *   RegisterInput(uFunc) -> (uIndex, nErr)
 *   CloseGPU() -> None
 *   StartTimer() -> nErr
 *   StopTimer() -> (fTime, nErr)
 * Matrices are lists (or any sequences) of rows; an empty matrix means
 * "absent" (ZeroB / ZeroC / ZeroD), as in ParODEmodule.c:240-290.  Unlike the
 * reference, the inputs are not echoed to stdout (:309-325), a malformed
 * argument raises ValueError instead of returning an unreferenced None
 * (:397-408).  tEnd = tStart + (nSteps - 1) tStep as the reference sets it
 * (:330), so the solver's truncating step count (RungeKutta4thOrder.cpp:20)
 * returns the same number of points as the reference's module: nSteps, or
 * nSteps - 1 where the floating-point quotient falls just below nSteps - 1.
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#define C_INCLUDE 1
#include <ParODECInterface.h>

#include <stdlib.h>
#include <string.h>

static char kFolder[] = "./";

static PyObject* ParODE_GetErrorDescription(PyObject* self, PyObject* args)
{
    int nErr;
    char strErr[512];
    if (!PyArg_ParseTuple(args, "i", &nErr)) return NULL;
    GetErrorDescription(&nErr, strErr);
    return Py_BuildValue("s", strErr);
}

static PyObject* ParODE_InitializeAllGPUs(PyObject* self, PyObject* args)
{
    int nErr = 0;
    InitializeAllGPUs(kFolder, kFolder, &nErr);
    return Py_BuildValue("i", nErr);
}

static PyObject* ParODE_GetAvailableGPUs(PyObject* self, PyObject* args)
{
    int nGPUs = 0, nErr = 0, ii;
    int GPUIds[64];
    GetAvailableGPUs(&nGPUs, GPUIds, &nErr);
    PyObject* list = PyList_New(nGPUs);
    if (!list) return NULL;
    for (ii = 0; ii < nGPUs; ii++) PyList_SET_ITEM(list, ii, PyLong_FromLong(GPUIds[ii]));
    return Py_BuildValue("(Ni)", list, nErr);
}

static PyObject* ParODE_GetDeviceName(PyObject* self, PyObject* args)
{
    int deviceID, nErr = 0;
    char name[256];
    if (!PyArg_ParseTuple(args, "i", &deviceID)) return NULL;
    name[0] = 0;
    GetDeviceName(&deviceID, name, &nErr);
    return Py_BuildValue("si", name, nErr);
}

static PyObject* ParODE_InitializeSelectedGPUs(PyObject* self, PyObject* args)
{
    PyObject *obj, *seq;
    int nErr = 0, n, ii;
    int* ids;
    if (!PyArg_ParseTuple(args, "O", &obj)) return NULL;
    seq = PySequence_Fast(obj, "InitializeSelectedGPUs expects a list of device ids");
    if (!seq) return NULL;
    n = (int)PySequence_Fast_GET_SIZE(seq);
    ids = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (ii = 0; ii < n; ii++) {
        long v = PyLong_AsLong(PySequence_Fast_GET_ITEM(seq, ii));
        if (v == -1 && PyErr_Occurred()) {
            free(ids);
            Py_DECREF(seq);
            return NULL;
        }
        ids[ii] = (int)v;
    }
    Py_DECREF(seq);
    InitializeSelectedGPUs(&n, ids, kFolder, kFolder, &nErr);
    free(ids);
    return Py_BuildValue("i", nErr);
}

/* rows x cols float copy of a sequence of row sequences (ParODEmodule.c:135-184);
 * returns 0 with an exception set on malformed input. */
static int disassemble(PyObject* obj, float** out, int* rows, int* cols, const char* name)
{
    PyObject* seq = PySequence_Fast(obj, "matrix must be a sequence of rows");
    Py_ssize_t r, c, nr, nc = -1;
    *out = NULL;
    *rows = *cols = 0;
    if (!seq) return 0;
    nr = PySequence_Fast_GET_SIZE(seq);
    for (r = 0; r < nr; r++) {
        PyObject* row = PySequence_Fast(PySequence_Fast_GET_ITEM(seq, r), "rows must be sequences");
        if (!row) goto fail;
        if (nc < 0) {
            nc = PySequence_Fast_GET_SIZE(row);
            *out = (float*)malloc(sizeof(float) * (nr * nc > 0 ? nr * nc : 1));
        } else if (PySequence_Fast_GET_SIZE(row) != nc) {
            Py_DECREF(row);
            PyErr_Format(PyExc_ValueError, "%s: rows of different lengths", name);
            goto fail;
        }
        for (c = 0; c < nc; c++) {
            double v = PyFloat_AsDouble(PySequence_Fast_GET_ITEM(row, c));
            if (v == -1.0 && PyErr_Occurred()) {
                Py_DECREF(row);
                goto fail;
            }
            (*out)[r * nc + c] = (float)v;
        }
        Py_DECREF(row);
    }
    Py_DECREF(seq);
    *rows = (int)nr;
    *cols = nr ? (int)nc : 0;
    return 1;
fail:
    Py_DECREF(seq);
    free(*out);
    *out = NULL;
    return 0;
}

static PyObject* ParODE_SimulateODE(PyObject* self, PyObject* args)
{
    PyObject *AL, *BL, *CL, *DL, *x0L, *ret = NULL, *tl = NULL, *xl = NULL, *seq = NULL;
    float *A = NULL, *B = NULL, *C = NULL, *D = NULL, *x0 = NULL, *tVect = NULL, *xVect = NULL;
    int uIndex, nSteps, solver, r, c, ii, jj;
    int nStates = 0, nInputs = 0, nOutputs = 0, ZeroB = 0, ZeroC = 0, ZeroD = 0;
    int nStepsOut = 0, nErr = 0;
    double tStart, tStep, tEnd;
    if (!PyArg_ParseTuple(args, "OOOOiddiOi", &AL, &BL, &CL, &DL, &uIndex, &tStart, &tStep,
                          &nSteps, &x0L, &solver))
        return NULL;
    if (!disassemble(AL, &A, &r, &c, "A")) goto done;
    if (r == 0 || r != c) {
        PyErr_SetString(PyExc_ValueError, "A must be square and non-empty");
        goto done;
    }
    nStates = r;
    if (!disassemble(BL, &B, &r, &c, "B")) goto done;
    if (r == 0 || c == 0) ZeroB = 1;
    else if (r == nStates) nInputs = c;
    else { PyErr_SetString(PyExc_ValueError, "B must have nStates rows"); goto done; }
    if (!disassemble(CL, &C, &r, &c, "C")) goto done;
    if (r == 0 || c == 0) ZeroC = 1;
    else if (c == nStates) nOutputs = r;
    else { PyErr_SetString(PyExc_ValueError, "C must have nStates columns"); goto done; }
    if (!disassemble(DL, &D, &r, &c, "D")) goto done;
    if (r == 0 || c == 0) ZeroD = 1;
    else {
        if (ZeroB) nInputs = c;
        if (ZeroC) nOutputs = r;
        if (nInputs != c || nOutputs != r) {
            PyErr_SetString(PyExc_ValueError, "D must be nOutputs x nInputs");
            goto done;
        }
    }
    if (ZeroC && ZeroD) nOutputs = nStates;
    seq = PySequence_Fast(x0L, "x0 must be a sequence");
    if (!seq) goto done;
    if (PySequence_Fast_GET_SIZE(seq) != nStates) {
        PyErr_SetString(PyExc_ValueError, "x0 must have nStates entries");
        goto done;
    }
    x0 = (float*)malloc(sizeof(float) * nStates);
    for (ii = 0; ii < nStates; ii++) {
        double v = PyFloat_AsDouble(PySequence_Fast_GET_ITEM(seq, ii));
        if (v == -1.0 && PyErr_Occurred()) goto done;
        x0[ii] = (float)v;
    }
    if (nSteps < 1) {
        PyErr_SetString(PyExc_ValueError, "nSteps must be positive");
        goto done;
    }
    tEnd = tStart + (nSteps - 1) * tStep;
    tVect = (float*)malloc(sizeof(float) * nSteps);
    xVect = (float*)malloc(sizeof(float) * (size_t)nSteps * nOutputs);
    Py_BEGIN_ALLOW_THREADS
    SimulateODE(A, B, C, D, &nInputs, &nStates, &nOutputs, &ZeroB, &ZeroC, &ZeroD, &uIndex,
                &tStart, &tEnd, &tStep, x0, &solver, tVect, xVect, &nStepsOut, &nErr);
    Py_END_ALLOW_THREADS
    if (nErr != 0) nStepsOut = 0;
    tl = PyList_New(nStepsOut);
    xl = PyList_New(nStepsOut);
    if (!tl || !xl) goto done;
    for (ii = 0; ii < nStepsOut; ii++) {
        PyObject* row = PyList_New(nOutputs);
        if (!row) goto done;
        PyList_SET_ITEM(tl, ii, PyFloat_FromDouble(tVect[ii]));
        for (jj = 0; jj < nOutputs; jj++)
            PyList_SET_ITEM(row, jj, PyFloat_FromDouble(xVect[(size_t)ii * nOutputs + jj]));
        PyList_SET_ITEM(xl, ii, row);
    }
    ret = Py_BuildValue("(NNi)", tl, xl, nErr);
    tl = xl = NULL;
done:
    Py_XDECREF(tl);
    Py_XDECREF(xl);
    Py_XDECREF(seq);
    free(A); free(B); free(C); free(D); free(x0); free(tVect); free(xVect);
    return ret;
}

static PyObject* ParODE_RegisterInput(PyObject* self, PyObject* args)
{
    const char* uFunc;
    int uIndex = 0, nErr = 0;
    char* copy;
    if (!PyArg_ParseTuple(args, "s", &uFunc)) return NULL;
    copy = strdup(uFunc);
    RegisterInput(copy, &uIndex, &nErr);
    free(copy);
    return Py_BuildValue("ii", uIndex, nErr);
}

static PyObject* ParODE_CloseGPU(PyObject* self, PyObject* args)
{
    CloseGPUC();
    Py_RETURN_NONE;
}

static PyObject* ParODE_StartTimer(PyObject* self, PyObject* args)
{
    int nErr = 0;
    StartTimer(&nErr);
    return Py_BuildValue("i", nErr);
}

static PyObject* ParODE_StopTimer(PyObject* self, PyObject* args)
{
    float fTime = 0.0f;
    int nErr = 0;
    StopTimer(&fTime, &nErr);
    return Py_BuildValue("fi", fTime, nErr);
}

static PyMethodDef ParODEMethods[] = {
    {"GetErrorDescription", ParODE_GetErrorDescription, METH_VARARGS, "Get Error Description String."},
    {"InitializeAllGPUs", ParODE_InitializeAllGPUs, METH_VARARGS, "Initialize all available GPUs."},
    {"GetAvailableGPUs", ParODE_GetAvailableGPUs, METH_VARARGS, "Retrieve all the available GPUs."},
    {"GetDeviceName", ParODE_GetDeviceName, METH_VARARGS, "Get the device name."},
    {"InitializeSelectedGPUs", ParODE_InitializeSelectedGPUs, METH_VARARGS, "Initialize Selected GPUs."},
    {"SimulateODE", ParODE_SimulateODE, METH_VARARGS, "Simulate an LTI system."},
    {"RegisterInput", ParODE_RegisterInput, METH_VARARGS, "Register an input kernel."},
    {"CloseGPU", ParODE_CloseGPU, METH_VARARGS, "Release GPU resources."},
    {"StartTimer", ParODE_StartTimer, METH_VARARGS, "Start GPU timer."},
    {"StopTimer", ParODE_StopTimer, METH_VARARGS, "Retrieve timer value."},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef ParODEModule = {PyModuleDef_HEAD_INIT, "ParODE",
                                          "GPU accelerated LTI simulation (ParODE, MI355X).", -1,
                                          ParODEMethods};

PyMODINIT_FUNC PyInit_ParODE(void)
{
    PyObject* m = PyModule_Create(&ParODEModule);
    if (!m) return NULL;
    PyObject* err = PyErr_NewException("ParODE.error", NULL, NULL);
    Py_XINCREF(err);
    if (PyModule_AddObject(m, "error", err) < 0) {
        Py_XDECREF(err);
        Py_DECREF(m);
        return NULL;
    }
    return m;
}
