// pfa_div.hpp -- the FP64 division of the resample passes' bracket fractions.
#pragma once

#include <hip/hip_runtime.h>

// n / d for 0 <= n <= d of ordinary magnitude, the fraction inside a bracket
// (xl <= Q < xh in the float passes, x0 < Q <= x1 in the fixed ones):
// reciprocal, two Newton steps, the quotient and one residual correction.
// These are the steps of the general correctly rounded FP64 division without
// its exponent pre-scaling and special-value fix-up, which leave such operands
// unchanged, so the quotient is the same correctly rounded one at four fewer
// FP64 instructions.
__device__ __forceinline__ double div_bracket(double n, double d) {
    const double r0 = __builtin_amdgcn_rcp(d);
    const double r1 = fma(fma(-d, r0, 1.0), r0, r0);
    const double r2 = fma(fma(-d, r1, 1.0), r1, r1);
    const double q = n * r2;
    return fma(fma(-d, q, n), r2, q);
}
