// tabulated_function.cpp -- librt_common_hip.so: a TabulatedFunction for
// callers of the plugin API that do not link the reference's common library
// (the C++ API tests, tools built on include/rt/).
//
// Behaviour follows common-lib/tabulated_function.cpp: a single positive
// value is returned for every argument (lines 127-143); a table needs at
// least 4 points (148-150) and its second derivatives come from the engine's
// spline_init (162-168); evaluation outside [x0, x_last] throws OutOfRange,
// a knot hit exactly returns its y, otherwise the cubic spline on the
// bracketing interval (173-204).  Text files hold one "x y" pair a line;
// blank lines and lines starting with '#' are skipped, a line that does not
// begin with two numbers is an error (ascii_file.cpp:43-87).
#include "rt/rt_engine_func.h"
#include "rt/tabulated_function.h"

#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <new>
#include <string>

TabulatedFunction::bad_tab_func::bad_tab_func(TabulatedFunctionError err) : Error(err) {}

TabulatedFunction::TabulatedFunctionError TabulatedFunction::bad_tab_func::bad_tab_func_error() const {
    return Error;
}

TabulatedFunction::TabulatedFunction() : SingleValue(false) {}

TabulatedFunction::TabulatedFunction(const real_t& val) : SingleValue(true) { Set(val); }

TabulatedFunction::TabulatedFunction(const char* filename) : SingleValue(false) { Load(filename); }

// Like the reference (tabulated_function.cpp:40-48) this constructor only
// stores the table: no spline coefficients until Set or Load.
TabulatedFunction::TabulatedFunction(const std::vector<real_t>& xv, const std::vector<real_t>& yv)
    : SingleValue(false), x(xv), y(yv) {
    if (x.size() != y.size()) {
        x.clear();
        y.clear();
        throw bad_tab_func(BadInputValues);
    }
    dy2.assign(x.size(), 0.0);
}

void TabulatedFunction::Set(const real_t& val) {
    x.clear();
    y.clear();
    dy2.clear();
    SingleValue = true;
    if (val <= 0.0) throw bad_tab_func(BadValue);
    x.push_back(0.0);
    y.push_back(val);
    dy2.push_back(0.0);
}

void TabulatedFunction::Set(const size_t n, const real_t* xv, const real_t* yv) {
    if (n == 0) throw bad_tab_func(BadValue);
    if (n < 4) throw bad_tab_func(NotEnoughPoints);
    SingleValue = false;
    x.assign(xv, xv + n);
    y.assign(yv, yv + n);
    dy2.assign(n, 0.0);
    SplineInit();
}

void TabulatedFunction::Load(const char* filename) {
    x.clear();
    y.clear();
    dy2.clear();
    SingleValue = false;
    std::ifstream in(filename);
    if (!in) throw bad_tab_func(FileFailure);
    std::string line;
    std::vector<real_t> xs, ys;
    while (std::getline(in, line)) {
        size_t i = line.find_first_not_of(' ');
        if (i == std::string::npos || line[i] == '\r') continue;
        if (line[i] == '#') continue;
        const char* s = line.c_str() + i;
        char* end = nullptr;
        double a = strtod(s, &end);
        if (end == s) throw bad_tab_func(BadValue);
        s = end;
        double b = strtod(s, &end);
        if (end == s) throw bad_tab_func(BadValue);
        xs.push_back(a);
        ys.push_back(b);
    }
    if (xs.empty()) throw bad_tab_func(BadValue);
    if (xs.size() < 4) throw bad_tab_func(NotEnoughPoints);
    x.swap(xs);
    y.swap(ys);
    dy2.assign(x.size(), 0.0);
    SplineInit();
}

void TabulatedFunction::SplineInit() {
    RT_engine_error err = spline_init(x.size(), x.data(), y.data(), dy2.data());
    if (err == ENGINE_ERROR_BAD_ALLOC) throw bad_tab_func(MemAllocError);
    if (err != ENGINE_ERROR_OK) throw bad_tab_func(InitFailure);
}

real_t TabulatedFunction::operator[](const real_t& xval) {
    if (y.empty()) throw bad_tab_func(OutOfRange);
    if (SingleValue) return y[0];
    const size_t n = x.size();
    if (xval < x[0] || xval > x[n - 1]) throw bad_tab_func(OutOfRange);
    // bisection for lo < hi with x[lo] <= xval < x[hi]; a knot hit exactly returns its value
    size_t lo = 0, hi = n;
    while (hi - lo > 1) {
        size_t mid = (lo + hi) / 2;
        if (x[mid] == xval) return y[mid];
        if (x[mid] > xval)
            hi = mid;
        else
            lo = mid;
    }
    const real_t h = x[hi] - x[lo];
    if (h == 0.0) throw bad_tab_func(BadArgument);
    const real_t a = (x[hi] - xval) / h;
    const real_t b = (xval - x[lo]) / h;
    return a * y[lo] + b * y[hi] + ((a * a * a - a) * dy2[lo] + (b * b * b - b) * dy2[hi]) * (h * h) / 6.0;
}
