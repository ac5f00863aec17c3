// Drives the ParODE C++ API (include/ParODECppInterface.h) the way the
// reference's TestsSources/TestParODEInterface.cpp does: query and select the
// GPUs, register a UFunc, simulate with matrixf/vectorf, print the results.
//
// Input on stdin (whitespace separated):
//   solver n m p steps h
//   A (n*n) B (n*m) C (p*n) D (p*m) x0 (n), row-major
//   then the UFunc source up to end of input
// Output on stdout: "ERR e", then "T k", then k lines "t y_0 ... y_{p-1}".
#include <ParODECppInterface.h>

#include <cstdio>
#include <iostream>
#include <iterator>
#include <sstream>
#include <string>

int main()
{
	int solver, n, m, p, steps;
	double h;
	if (!(std::cin >> solver >> n >> m >> p >> steps >> h)) return 2;
	matrixf A(n, n), B(n, m), C(p, n), D(p, m);
	vectorf x0(n);
	for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) std::cin >> A(i, j);
	for (int i = 0; i < n; ++i) for (int j = 0; j < m; ++j) std::cin >> B(i, j);
	for (int i = 0; i < p; ++i) for (int j = 0; j < n; ++j) std::cin >> C(i, j);
	for (int i = 0; i < p; ++i) for (int j = 0; j < m; ++j) std::cin >> D(i, j);
	for (int i = 0; i < n; ++i) std::cin >> x0(i);
	std::string src((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());

	std::vector<unsigned int> ids;
	ErrorCode e = GetAvailableGPUs(ids);
	if (e != ParODE_OK || ids.empty()) { std::printf("ERR %d\n", (int)e); return 1; }
	std::string name, folder = "./";
	GetDeviceName(ids[0], name);
	std::fprintf(stderr, "device %u: %s\n", ids[0], name.c_str());
	std::vector<unsigned int> one(1, ids[0]);
	e = InitializeSelectedGPUs(one, folder, folder);
	if (e != ParODE_OK) { std::printf("ERR %d\n", (int)e); return 1; }
	int uIndex = RegisterInput(src);
	if (uIndex < 0) { std::printf("ERR %d\n", uIndex); return 1; }
	vectorf t;
	matrixf x;
	const double tEnd = (steps - 1) * h;
	e = SimulateODE(A, B, C, D, uIndex, 0.0, tEnd, h, x0, (SolverType)solver, t, x);
	std::printf("ERR %d\n", (int)e);
	std::printf("T %zu\n", t.size());
	for (std::size_t k = 0; k < t.size(); ++k) {
		std::printf("%.9g", (double)t(k));
		// xVect is outputs x time (GPUFixedStepIterative.cpp:442-446 of the reference)
		for (std::size_t j = 0; j < x.size1(); ++j) std::printf(" %.9g", (double)x(j, k));
		std::printf("\n");
	}
	CloseGPU();
	return 0;
}
