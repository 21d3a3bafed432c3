// ep_sched.h -- powers and logarithms of the optimiser's schedules (lambda's
// mu, gamma's 10^x, the step-model features), evaluated by the same sequence
// of IEEE double operations on the host and on the device: explicit fma, exact
// frexp/ldexp/rint, correctly rounded division, no library transcendental and
// no multiply-add contraction (the device compiler would otherwise fuse a*b - c
// where the host does not).  Host-controlled and device-controlled runs
// therefore take bit-identical decisions.  Relative error a few ulp.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace ep {

// 2^x for |x| < 1000.  x = k + f with |f| <= 1/2; e^(f ln 2) by its Taylor
// series to t^13/13! (|t| <= 0.35: truncation below 1e-17).
__host__ __device__ inline double sched_exp2(double x)
{
#pragma clang fp contract(off)
    const double k = rint(x);
    const double t = (x - k) * 0.69314718055994531;
    const double c[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                          1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,
                          1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,        0.5,
                          1.0,                1.0};
    double p = c[0];
    for (int i = 1; i < 14; ++i) p = fma(p, t, c[i]);
    return ldexp(p, (int)k);
}

// log2(x) for finite x > 0.  x = m 2^e with m in [sqrt(1/2), sqrt(2));
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.172, series to s^23.
__host__ __device__ inline double sched_log2(double x)
{
#pragma clang fp contract(off)
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double q = 1.0 / 23.0;
    for (int n = 21; n >= 1; n -= 2) q = fma(q, s2, 1.0 / n);
    return fma(2.0 * s * q, 1.4426950408889634, (double)e);
}

__host__ __device__ inline double sched_pow(double base, double y)
{
#pragma clang fp contract(off)
    return sched_exp2(y * sched_log2(base));
}
__host__ __device__ inline double sched_exp10(double x)
{
#pragma clang fp contract(off)
    return sched_exp2(x * 3.3219280948873622);
}
__host__ __device__ inline double sched_log10(double x)
{
#pragma clang fp contract(off)
    return sched_log2(x) * 0.30102999566398120;
}

}  // namespace ep
