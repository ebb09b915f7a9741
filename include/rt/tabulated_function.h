/* tabulated_function.h -- the tabulated function type the plugin API takes
 * by reference (interface of common-lib/tabulated_function.h:18-54).  The
 * engine library calls only operator[]; its definition comes from the
 * caller's common library (or rt_amd's librt_common_hip.so). */
#ifndef TABULATED_FUNCTION_H
#define TABULATED_FUNCTION_H

#include "num_defs.h"

#include <exception>
#include <vector>

class TabulatedFunction {
public:
    enum TabulatedFunctionError {MemAllocError, BadValue, BadInputValues, FileFailure, NotEnoughPoints,
                                 InitFailure, BadArgument, OutOfRange};

    TabulatedFunction();
    explicit TabulatedFunction(const real_t &val);
    explicit TabulatedFunction(const char *filename);
    TabulatedFunction(const std::vector<real_t> &x, const std::vector<real_t> &yval);

    void Set(const real_t &val);
    void Set(const size_t Nelements, const real_t *xval, const real_t *yval);
    void Load(const char *filename);

    /* value at xval; throws bad_tab_func(OutOfRange) outside the table */
    real_t operator[](const real_t &xval);

    /* Read access for packing the function into the engine's spline pool
       (rt_amd's Scheme): not in common-lib's interface, inline so the
       class layout and exported symbols are unchanged. */
    bool IsSingleValue() const { return SingleValue; }
    const std::vector<real_t> &Abscissae() const { return x; }
    const std::vector<real_t> &Values() const { return y; }
    const std::vector<real_t> &SecondDerivatives() const { return dy2; }

    class bad_tab_func : public std::exception {
    public:
        bad_tab_func(TabulatedFunctionError err);
        TabulatedFunctionError bad_tab_func_error() const;
    private:
        TabulatedFunctionError Error;
    };

private:
    bool SingleValue;
    std::vector<real_t> x, y, dy2;
    void SplineInit();
};

#endif /* TABULATED_FUNCTION_H */
