// Cubic Hermite spline helpers shared by the force-field utilities and the
// relaxation kernels: Spline_Hermite::val and ::find_index
// (ppafm/cpp/spline_hermite.h:45-51, 159-168) and the radial tip force
// forceRSpline (ppafm/cpp/ProbeParticle.cpp:211-224).
#pragma once

#include <hip/hip_runtime.h>

namespace pp {

// Cubic Hermite interpolant on [0,1] from end values and end slopes.
__device__ __forceinline__ double hermite(double x, double y0, double y1, double dy0, double dy1)
{
    const double y01 = y0 - y1;
    return y0 + x * (dy0 + x * (-3 * y01 - 2 * dy0 - dy1 + x * (2 * y01 + dy0 + dy1)));
}

// Halving search from i over di knots for the last knot below x.
__device__ __forceinline__ int find_index(int i, int di, double x, const double *xs)
{
    while (di > 1) {
        di >>= 1;
        const int j = i + di;
        if (xs[j] < x) i = j;
    }
    return i;
}

// Radial tip force / distance at distance x from a spline of n knots xs with
// values and slopes ydys (2n): multiply the tip-to-probe vector by it.  The
// interval index is clamped to the last one, where the reference reads past
// the table.
__device__ __forceinline__ double tip_spline_force_scale(double x, int n, const double *xs, const double *ydys)
{
    int i = find_index(0, n, x, xs);
    if (i > n - 2) i = n - 2;
    const double x0 = xs[i], x1 = xs[i + 1];
    const double dx = x1 - x0;
    const double denom = 1 / dx;
    const double u = (x - x0) * denom;
    const int k = i << 1;
    const double fr = hermite(u, ydys[k], ydys[k + 2], ydys[k + 1] * dx, ydys[k + 3] * dx);
    return fr / x;
}

}  // namespace pp
