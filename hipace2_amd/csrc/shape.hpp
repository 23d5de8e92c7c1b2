// B-spline shape factors for deposition and gather on gfx950.
//
// Same piecewise polynomials as the reference's ShapeFactors.H (the table of
// compute_single_shape_factor :122-195 and single_derivative_shape_factor :211-466):
// value of the order-O spline at u = xmid - cell; derivative type 0 is its analytic
// derivative, type 1 (nodal) the derivative of the order-(O+1) spline on a one cell
// wider window, type 2 (centred) half the difference of the two neighbouring values.
// `ds` is returned negated, as the reference returns {s, -sdx, cell}.
// Order and type are template parameters so every stencil loop unrolls.
#pragma once
#include <hip/hip_runtime.h>

namespace hpq {

struct Sf { double s; int cell; };
struct Dsf { double s; double ds; int cell; };

template <int O>
__host__ __device__ __forceinline__ Sf shape(double xmid, int ix) {
    if constexpr (O == 0) {
        return {1.0, int(floor(xmid + 0.5))};
    } else if constexpr (O == 1) {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        return ix == 0 ? Sf{1.0 - t, int(xf)} : Sf{t, int(xf) + 1};
    } else if constexpr (O == 2) {
        const double xf = floor(xmid + 0.5);
        const double t = xmid - xf;
        if (ix == 0) return {0.5 * (0.5 - t) * (0.5 - t), int(xf) - 1};
        if (ix == 1) return {0.75 - t * t, int(xf)};
        return {0.5 * (0.5 + t) * (0.5 + t), int(xf) + 1};
    } else {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        if (ix == 0) return {1.0 / 6.0 * (1.0 - t) * (1.0 - t) * (1.0 - t), int(xf) - 1};
        if (ix == 1) return {2.0 / 3.0 - t * t * (1.0 - t / 2.0), int(xf)};
        if (ix == 2) return {2.0 / 3.0 - (1.0 - t) * (1.0 - t) * (1.0 - 0.5 * (1.0 - t)), int(xf) + 1};
        return {1.0 / 6.0 * t * t * t, int(xf) + 2};
    }
}

__host__ __device__ __forceinline__ Dsf dmk(double s, double sdx, int cell) { return {s, -sdx, cell}; }

// type 0: analytic derivative of the order-O spline, O+1 cells
template <int O>
__host__ __device__ __forceinline__ Dsf dshape_analytic(double xmid, int ix) {
    if constexpr (O == 0) {
        return dmk(1.0, 0.0, int(floor(xmid + 0.5)));
    } else if constexpr (O == 1) {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        return ix == 0 ? dmk(1 - t, -1.0, int(xf)) : dmk(t, 1.0, int(xf) + 1);
    } else if constexpr (O == 2) {
        const double xf = floor(xmid + 0.5);
        const double t = xmid + 0.5 - xf;
        const double t2 = t * t;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk(0.5 * t2 - 1.0 * t + 0.5, t - 1.0, c);
        if (ix == 1) return dmk(-t2 + 1.0 * t + 0.5, 1.0 - 2 * t, c);
        return dmk(0.5 * t2, t, c);
    } else {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        const double t2 = t * t, t3 = t2 * t;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk(-1.0 / 6.0 * t3 + 0.5 * t2 - 0.5 * t + 1.0 / 6.0, -0.5 * t2 + t - 0.5, c);
        if (ix == 1) return dmk(0.5 * t3 - t2 + 2.0 / 3.0, 1.5 * t2 - 2 * t, c);
        if (ix == 2) return dmk(-0.5 * t3 + 0.5 * t2 + 0.5 * t + 1.0 / 6.0, -1.5 * t2 + t + 0.5, c);
        return dmk(1.0 / 6.0 * t3, 0.5 * t2, c);
    }
}

// type 1: nodal (value of B_O, derivative of B_{O+1}), O+2 cells
template <int O>
__host__ __device__ __forceinline__ Dsf dshape_nodal(double xmid, int ix) {
    if constexpr (O == 0) {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        return ix == 0 ? dmk((t < 0.5) ? 1 : 0, -1, int(xf)) : dmk((t < 0.5) ? 0 : 1, 1, int(xf) + 1);
    } else if constexpr (O == 1) {
        const double xf = floor(xmid + 0.5);
        const double t = xmid + 0.5 - xf;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk((t < 0.5) ? 0.5 - t : 0, t - 1, c);
        if (ix == 1) return dmk((t < 0.5) ? t + 0.5 : 1.5 - t, 1 - 2 * t, c);
        return dmk((t < 0.5) ? 0 : t - 0.5, t, c);
    } else if constexpr (O == 2) {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        const double t2 = t * t;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk((t < 0.5) ? 0.5 * t2 - 0.5 * t + 1.0 / 8.0 : 0, -0.5 * t2 + t - 0.5, c);
        if (ix == 1) return dmk((t < 0.5) ? 3.0 / 4.0 - t2 : 0.5 * t2 - 1.5 * t + 9.0 / 8.0, 1.5 * t2 - 2 * t, c);
        if (ix == 2)
            return dmk((t < 0.5) ? 0.5 * t2 + 0.5 * t + 1.0 / 8.0 : -t2 + 2 * t - 1.0 / 4.0, -1.5 * t2 + t + 0.5, c);
        return dmk((t < 0.5) ? 0 : 0.5 * t2 - 0.5 * t + 1.0 / 8.0, 0.5 * t2, c);
    } else {
        const double xf = floor(xmid + 0.5);
        const double t = xmid + 0.5 - xf;
        const double t2 = t * t, t3 = t2 * t;
        const int c = int(xf) - 2 + ix;
        const bool lo = t < 0.5;
        if (ix == 0)
            return dmk(lo ? -1.0 / 6.0 * t3 + 0.25 * t2 - 1.0 / 8.0 * t + 1.0 / 48.0 : 0,
                       1.0 / 6.0 * t3 - 0.5 * t2 + 0.5 * t - 1.0 / 6.0, c);
        if (ix == 1)
            return dmk(lo ? 0.5 * t3 - 0.25 * t2 - 5.0 / 8.0 * t + 23.0 / 48.0
                          : -1.0 / 6.0 * t3 + 0.75 * t2 - 9.0 / 8.0 * t + 9.0 / 16.0,
                       -2.0 / 3.0 * t3 + 1.5 * t2 - 0.5 * t - 0.5, c);
        if (ix == 2)
            return dmk(lo ? -0.5 * t3 - 0.25 * t2 + 5.0 / 8.0 * t + 23.0 / 48.0
                          : 0.5 * t3 - 7.0 / 4.0 * t2 + 11.0 / 8.0 * t + 17.0 / 48.0,
                       t3 - 1.5 * t2 - 0.5 * t + 0.5, c);
        if (ix == 3)
            return dmk(lo ? 1.0 / 6.0 * t3 + 0.25 * t2 + 1.0 / 8.0 * t + 1.0 / 48.0
                          : -0.5 * t3 + 5.0 / 4.0 * t2 - 3.0 / 8.0 * t + 5.0 / 48.0,
                       -2.0 / 3.0 * t3 + 0.5 * t2 + 0.5 * t + 1.0 / 6.0, c);
        return dmk(lo ? 0 : 1.0 / 6.0 * t3 - 0.25 * t2 + 1.0 / 8.0 * t - 1.0 / 48.0, 1.0 / 6.0 * t3, c);
    }
}

// type 2: centred derivative, O+3 cells
template <int O>
__host__ __device__ __forceinline__ Dsf dshape_centred(double xmid, int ix) {
    if constexpr (O == 0) {
        const double xf = floor(xmid + 0.5);
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk(0, -0.5, c);
        if (ix == 1) return dmk(1, 0, c);
        return dmk(0, 0.5, c);
    } else if constexpr (O == 1) {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk(0, 0.5 * t - 0.5, c);
        if (ix == 1) return dmk(1 - t, -0.5 * t, c);
        if (ix == 2) return dmk(t, 0.5 - 0.5 * t, c);
        return dmk(0, 0.5 * t, c);
    } else if constexpr (O == 2) {
        const double xf = floor(xmid + 0.5);
        const double t = xmid + 0.5 - xf;
        const double t2 = t * t;
        const int c = int(xf) - 2 + ix;
        if (ix == 0) return dmk(0, -0.25 * t2 + 0.5 * t - 0.25, c);
        if (ix == 1) return dmk(0.5 * t2 - t + 0.5, 0.5 * t2 - 0.5 * t - 0.25, c);
        if (ix == 2) return dmk(-t2 + t + 0.5, 0.25 - 0.5 * t, c);
        if (ix == 3) return dmk(0.5 * t2, -0.5 * t2 + 0.5 * t + 0.25, c);
        return dmk(0, 0.25 * t2, c);
    } else {
        const double xf = floor(xmid);
        const double t = xmid - xf;
        const double t2 = t * t, t3 = t2 * t;
        const int c = int(xf) - 2 + ix;
        if (ix == 0) return dmk(0, 1.0 / 12.0 * t3 - 0.25 * t2 + 0.25 * t - 1.0 / 12.0, c);
        if (ix == 1) return dmk(-1.0 / 6.0 * t3 + 0.5 * t2 - 0.5 * t + 1.0 / 6.0, -0.25 * t3 + 0.5 * t2 - 1.0 / 3.0, c);
        if (ix == 2) return dmk(0.5 * t3 - t2 + 2.0 / 3.0, 1.0 / 6.0 * t3 - 0.5 * t, c);
        if (ix == 3) return dmk(-0.5 * t3 + 0.5 * t2 + 0.5 * t + 1.0 / 6.0, 1.0 / 6.0 * t3 - 0.5 * t2 + 1.0 / 3.0, c);
        if (ix == 4) return dmk(1.0 / 6.0 * t3, -0.25 * t3 + 0.25 * t2 + 0.25 * t + 1.0 / 12.0, c);
        return dmk(0, 1.0 / 12.0 * t3, c);
    }
}

template <int T, int O>
__host__ __device__ __forceinline__ Dsf dshape(double xmid, int ix) {
    if constexpr (T == 0) return dshape_analytic<O>(xmid, ix);
    else if constexpr (T == 1) return dshape_nodal<O>(xmid, ix);
    else return dshape_centred<O>(xmid, ix);
}

// number of cells of the derivative stencil
template <int T, int O>
constexpr int dshape_width() { return T == 0 ? O + 1 : (T == 1 ? O + 2 : O + 3); }

}  // namespace hpq
