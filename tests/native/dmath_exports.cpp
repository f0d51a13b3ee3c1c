// Host build of the FP64 step's deterministic elementary functions
// (landslidehysea_amd/csrc/lhysea_dmath.hpp) for tests/test_dmath.py.
// Build: g++ -std=c++17 -O2 -ffp-contract=off -fPIC -shared -I tests/native dmath_exports.cpp
#include <hip/hip_runtime.h>
#include "../../landslidehysea_amd/csrc/lhysea_dmath.hpp"

extern "C" void dm_eval(int fn, const double* x, const double* y, double* out, long n)
{
    using namespace lhysea;
    for (long i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = dm::exp(x[i]); break;
        case 1: out[i] = dm::log(x[i]); break;
        case 2: out[i] = dm::cbrt(x[i]); break;
        case 3: out[i] = dm::pow43(x[i]); break;
        default: out[i] = dm::powpos(x[i], y[i]); break;
        }
    }
}
