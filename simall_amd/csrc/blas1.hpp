// blas1.hpp — vector kernels and reduction scratch.
#pragma once

#include "mi_common.hpp"

#include <cmath>

namespace mi
{

constexpr int kMaxReduceValues = 4;

// Block-partial buffer of the two-launch reductions, one per stream user.
struct ReduceScratch
{
    DevBuf<double> partials;
    DevBuf<double> result;
    int ensure(hipStream_t s);
};

// mode: 0 axpy y+=a*x, 1 scal y*=a, 2 copy y=x, 3 axpby y=a*x+b*y, 4 xpby y=x+b*y
int launch_vec(int mode, int64_t n, double a, const double* x, double b, double* y, hipStream_t s);
// *out_dev = x . y
int launch_dot(int64_t n, const double* x, const double* y, ReduceScratch& rs, double* out_dev,
               hipStream_t s);

} // namespace mi
