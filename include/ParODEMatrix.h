/* Dense row-major float matrix and vector with the member surface the ParODE
 * C++ API uses from boost::numeric::ublas::matrix<fType> / vector<fType>
 * (Include/ParODECppInterface.h:11-14 of the reference): size1(), size2(),
 * size(), resize(), operator()(i, j), operator[](i).  Boost is not needed. */
#ifndef PARODE_MATRIX_H_
#define PARODE_MATRIX_H_

#include <FTypeDef.h>

#include <algorithm>
#include <cstddef>
#include <vector>

class matrixf {
public:
	matrixf() {}
	matrixf(std::size_t rows, std::size_t cols, fType v = 0) : _r(rows), _c(cols), _d(rows * cols, v) {}
	std::size_t size1() const { return _r; }
	std::size_t size2() const { return _c; }
	void resize(std::size_t rows, std::size_t cols, bool preserve = true) {
		std::vector<fType> d(rows * cols, fType(0));
		if (preserve)
			for (std::size_t i = 0; i < std::min(rows, _r); ++i)
				for (std::size_t j = 0; j < std::min(cols, _c); ++j) d[i * cols + j] = _d[i * _c + j];
		_d.swap(d);
		_r = rows;
		_c = cols;
	}
	fType& operator()(std::size_t i, std::size_t j) { return _d[i * _c + j]; }
	const fType& operator()(std::size_t i, std::size_t j) const { return _d[i * _c + j]; }
	fType* data() { return _d.data(); }
	const fType* data() const { return _d.data(); }

private:
	std::size_t _r = 0, _c = 0;
	std::vector<fType> _d;
};

class vectorf {
public:
	vectorf() {}
	explicit vectorf(std::size_t n, fType v = 0) : _d(n, v) {}
	std::size_t size() const { return _d.size(); }
	void resize(std::size_t n, bool preserve = true) {
		if (preserve) _d.resize(n, fType(0));
		else _d.assign(n, fType(0));
	}
	fType& operator()(std::size_t i) { return _d[i]; }
	const fType& operator()(std::size_t i) const { return _d[i]; }
	fType& operator[](std::size_t i) { return _d[i]; }
	const fType& operator[](std::size_t i) const { return _d[i]; }
	fType* data() { return _d.data(); }
	const fType* data() const { return _d.data(); }

private:
	std::vector<fType> _d;
};

#endif /* PARODE_MATRIX_H_ */
