// B-spline shape factors for deposition and gather on gfx950.
//
// Same piecewise polynomials as the reference's ShapeFactors.H (the table of
// compute_single_shape_factor :122-195 and single_derivative_shape_factor :211-466):
// value of the order-O spline at u = xmid - cell; derivative type 0 is its analytic
// derivative, type 1 (nodal) the derivative of the order-(O+1) spline on a one cell
// wider window, type 2 (centred) half the difference of the two neighbouring values.
// `ds` is returned negated, as the reference returns {s, -sdx, cell}.
// Order and type are template parameters so every stencil loop unrolls; the arithmetic
// type R of the polynomials is double for the FP64 path and float for the mixed-precision
// path. The position xmid stays FP64 in both: the cell index (floor), the offset inside the
// cell and the choice between polynomial pieces are taken in FP64, so both paths pick the
// same cells and pieces, and only the polynomial values are rounded to R.
#pragma once
#include <hip/hip_runtime.h>

namespace hpq {

// Value of a position: the position itself, or the value part of a forward-mode tangent
// (whose derivative parts carry d/d(inputs) through the polynomials; cells and pieces are
// always chosen from the value).
__host__ __device__ __forceinline__ double pos_value(double x) { return x; }

template <class R>
struct SfR { R s; int cell; };
template <class R>
struct DsfR { R s; R ds; int cell; };
using Sf = SfR<double>;
using Dsf = DsfR<double>;

template <int O, class R = double, class P = double>
__host__ __device__ __forceinline__ SfR<R> shape(P xmid, int ix) {
    if constexpr (O == 0) {
        return {R(1.0), int(floor(pos_value(xmid) + 0.5))};
    } else if constexpr (O == 1) {
        const double xf = floor(pos_value(xmid));
        const R t = R(xmid - xf);
        return ix == 0 ? SfR<R>{R(1.0) - t, int(xf)} : SfR<R>{t, int(xf) + 1};
    } else if constexpr (O == 2) {
        const double xf = floor(pos_value(xmid) + 0.5);
        const R t = R(xmid - xf);
        if (ix == 0) return {R(0.5) * (R(0.5) - t) * (R(0.5) - t), int(xf) - 1};
        if (ix == 1) return {R(0.75) - t * t, int(xf)};
        return {R(0.5) * (R(0.5) + t) * (R(0.5) + t), int(xf) + 1};
    } else {
        const double xf = floor(pos_value(xmid));
        const R t = R(xmid - xf);
        if (ix == 0) return {R(1.0) / R(6.0) * (R(1.0) - t) * (R(1.0) - t) * (R(1.0) - t), int(xf) - 1};
        if (ix == 1) return {R(2.0) / R(3.0) - t * t * (R(1.0) - t / R(2.0)), int(xf)};
        if (ix == 2) return {R(2.0) / R(3.0) - (R(1.0) - t) * (R(1.0) - t) * (R(1.0) - R(0.5) * (R(1.0) - t)), int(xf) + 1};
        return {R(1.0) / R(6.0) * t * t * t, int(xf) + 2};
    }
}

template <class R>
__host__ __device__ __forceinline__ DsfR<R> dmk(R s, R sdx, int cell) { return {s, -sdx, cell}; }

// type 0: analytic derivative of the order-O spline, O+1 cells
template <int O, class R = double, class P = double>
__host__ __device__ __forceinline__ DsfR<R> dshape_analytic(P xmid, int ix) {
    if constexpr (O == 0) {
        return dmk<R>(R(1.0), R(0.0), int(floor(pos_value(xmid) + 0.5)));
    } else if constexpr (O == 1) {
        const double xf = floor(pos_value(xmid));
        const R t = R(xmid - xf);
        return ix == 0 ? dmk<R>(1 - t, -R(1.0), int(xf)) : dmk<R>(t, R(1.0), int(xf) + 1);
    } else if constexpr (O == 2) {
        const double xf = floor(pos_value(xmid) + 0.5);
        const R t = R(xmid + 0.5 - xf);
        const R t2 = t * t;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk<R>(R(0.5) * t2 - R(1.0) * t + R(0.5), t - R(1.0), c);
        if (ix == 1) return dmk<R>(-t2 + R(1.0) * t + R(0.5), R(1.0) - 2 * t, c);
        return dmk<R>(R(0.5) * t2, t, c);
    } else {
        const double xf = floor(pos_value(xmid));
        const R t = R(xmid - xf);
        const R t2 = t * t, t3 = t2 * t;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk<R>(-R(1.0) / R(6.0) * t3 + R(0.5) * t2 - R(0.5) * t + R(1.0) / R(6.0), -R(0.5) * t2 + t - R(0.5), c);
        if (ix == 1) return dmk<R>(R(0.5) * t3 - t2 + R(2.0) / R(3.0), R(1.5) * t2 - 2 * t, c);
        if (ix == 2) return dmk<R>(-R(0.5) * t3 + R(0.5) * t2 + R(0.5) * t + R(1.0) / R(6.0), -R(1.5) * t2 + t + R(0.5), c);
        return dmk<R>(R(1.0) / R(6.0) * t3, R(0.5) * t2, c);
    }
}

// type 1: nodal (value of B_O, derivative of B_{O+1}), O+2 cells
template <int O, class R = double, class P = double>
__host__ __device__ __forceinline__ DsfR<R> dshape_nodal(P xmid, int ix) {
    if constexpr (O == 0) {
        const double xf = floor(pos_value(xmid));
        const bool lo = pos_value(xmid) - xf < 0.5;
        return ix == 0 ? dmk<R>(lo ? 1 : 0, -1, int(xf)) : dmk<R>(lo ? 0 : 1, 1, int(xf) + 1);
    } else if constexpr (O == 1) {
        const double xf = floor(pos_value(xmid) + 0.5);
        const P td = xmid + 0.5 - xf;
        const bool lo = pos_value(td) < 0.5;
        const R t = R(td);
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk<R>(lo ? R(0.5) - t : 0, t - 1, c);
        if (ix == 1) return dmk<R>(lo ? t + R(0.5) : R(1.5) - t, 1 - 2 * t, c);
        return dmk<R>(lo ? 0 : t - R(0.5), t, c);
    } else if constexpr (O == 2) {
        const double xf = floor(pos_value(xmid));
        const P td = xmid - xf;
        const bool lo = pos_value(td) < 0.5;
        const R t = R(td);
        const R t2 = t * t;
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk<R>(lo ? R(0.5) * t2 - R(0.5) * t + R(1.0) / R(8.0) : 0, -R(0.5) * t2 + t - R(0.5), c);
        if (ix == 1) return dmk<R>(lo ? R(3.0) / R(4.0) - t2 : R(0.5) * t2 - R(1.5) * t + R(9.0) / R(8.0), R(1.5) * t2 - 2 * t, c);
        if (ix == 2)
            return dmk<R>(lo ? R(0.5) * t2 + R(0.5) * t + R(1.0) / R(8.0) : -t2 + 2 * t - R(1.0) / R(4.0), -R(1.5) * t2 + t + R(0.5), c);
        return dmk<R>(lo ? 0 : R(0.5) * t2 - R(0.5) * t + R(1.0) / R(8.0), R(0.5) * t2, c);
    } else {
        const double xf = floor(pos_value(xmid) + 0.5);
        const P td = xmid + 0.5 - xf;
        const R t = R(td);
        const R t2 = t * t, t3 = t2 * t;
        const int c = int(xf) - 2 + ix;
        const bool lo = pos_value(td) < 0.5;
        if (ix == 0)
            return dmk<R>(lo ? -R(1.0) / R(6.0) * t3 + R(0.25) * t2 - R(1.0) / R(8.0) * t + R(1.0) / R(48.0) : 0,
                       R(1.0) / R(6.0) * t3 - R(0.5) * t2 + R(0.5) * t - R(1.0) / R(6.0), c);
        if (ix == 1)
            return dmk<R>(lo ? R(0.5) * t3 - R(0.25) * t2 - R(5.0) / R(8.0) * t + R(23.0) / R(48.0)
                          : -R(1.0) / R(6.0) * t3 + R(0.75) * t2 - R(9.0) / R(8.0) * t + R(9.0) / R(16.0),
                       -R(2.0) / R(3.0) * t3 + R(1.5) * t2 - R(0.5) * t - R(0.5), c);
        if (ix == 2)
            return dmk<R>(lo ? -R(0.5) * t3 - R(0.25) * t2 + R(5.0) / R(8.0) * t + R(23.0) / R(48.0)
                          : R(0.5) * t3 - R(7.0) / R(4.0) * t2 + R(11.0) / R(8.0) * t + R(17.0) / R(48.0),
                       t3 - R(1.5) * t2 - R(0.5) * t + R(0.5), c);
        if (ix == 3)
            return dmk<R>(lo ? R(1.0) / R(6.0) * t3 + R(0.25) * t2 + R(1.0) / R(8.0) * t + R(1.0) / R(48.0)
                          : -R(0.5) * t3 + R(5.0) / R(4.0) * t2 - R(3.0) / R(8.0) * t + R(5.0) / R(48.0),
                       -R(2.0) / R(3.0) * t3 + R(0.5) * t2 + R(0.5) * t + R(1.0) / R(6.0), c);
        return dmk<R>(lo ? 0 : R(1.0) / R(6.0) * t3 - R(0.25) * t2 + R(1.0) / R(8.0) * t - R(1.0) / R(48.0), R(1.0) / R(6.0) * t3, c);
    }
}

// type 2: centred derivative, O+3 cells
template <int O, class R = double, class P = double>
__host__ __device__ __forceinline__ DsfR<R> dshape_centred(P xmid, int ix) {
    if constexpr (O == 0) {
        const double xf = floor(pos_value(xmid) + 0.5);
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk<R>(0, -R(0.5), c);
        if (ix == 1) return dmk<R>(1, 0, c);
        return dmk<R>(0, R(0.5), c);
    } else if constexpr (O == 1) {
        const double xf = floor(pos_value(xmid));
        const R t = R(xmid - xf);
        const int c = int(xf) - 1 + ix;
        if (ix == 0) return dmk<R>(0, R(0.5) * t - R(0.5), c);
        if (ix == 1) return dmk<R>(1 - t, -R(0.5) * t, c);
        if (ix == 2) return dmk<R>(t, R(0.5) - R(0.5) * t, c);
        return dmk<R>(0, R(0.5) * t, c);
    } else if constexpr (O == 2) {
        const double xf = floor(pos_value(xmid) + 0.5);
        const R t = R(xmid + 0.5 - xf);
        const R t2 = t * t;
        const int c = int(xf) - 2 + ix;
        if (ix == 0) return dmk<R>(0, -R(0.25) * t2 + R(0.5) * t - R(0.25), c);
        if (ix == 1) return dmk<R>(R(0.5) * t2 - t + R(0.5), R(0.5) * t2 - R(0.5) * t - R(0.25), c);
        if (ix == 2) return dmk<R>(-t2 + t + R(0.5), R(0.25) - R(0.5) * t, c);
        if (ix == 3) return dmk<R>(R(0.5) * t2, -R(0.5) * t2 + R(0.5) * t + R(0.25), c);
        return dmk<R>(0, R(0.25) * t2, c);
    } else {
        const double xf = floor(pos_value(xmid));
        const R t = R(xmid - xf);
        const R t2 = t * t, t3 = t2 * t;
        const int c = int(xf) - 2 + ix;
        if (ix == 0) return dmk<R>(0, R(1.0) / R(12.0) * t3 - R(0.25) * t2 + R(0.25) * t - R(1.0) / R(12.0), c);
        if (ix == 1) return dmk<R>(-R(1.0) / R(6.0) * t3 + R(0.5) * t2 - R(0.5) * t + R(1.0) / R(6.0), -R(0.25) * t3 + R(0.5) * t2 - R(1.0) / R(3.0), c);
        if (ix == 2) return dmk<R>(R(0.5) * t3 - t2 + R(2.0) / R(3.0), R(1.0) / R(6.0) * t3 - R(0.5) * t, c);
        if (ix == 3) return dmk<R>(-R(0.5) * t3 + R(0.5) * t2 + R(0.5) * t + R(1.0) / R(6.0), R(1.0) / R(6.0) * t3 - R(0.5) * t2 + R(1.0) / R(3.0), c);
        if (ix == 4) return dmk<R>(R(1.0) / R(6.0) * t3, -R(0.25) * t3 + R(0.25) * t2 + R(0.25) * t + R(1.0) / R(12.0), c);
        return dmk<R>(0, R(1.0) / R(12.0) * t3, c);
    }
}

template <int T, int O, class R = double, class P = double>
__host__ __device__ __forceinline__ DsfR<R> dshape(P xmid, int ix) {
    if constexpr (T == 0) return dshape_analytic<O, R, P>(xmid, ix);
    else if constexpr (T == 1) return dshape_nodal<O, R, P>(xmid, ix);
    else return dshape_centred<O, R, P>(xmid, ix);
}

// number of cells of the derivative stencil
template <int T, int O>
constexpr int dshape_width() { return T == 0 ? O + 1 : (T == 1 ? O + 2 : O + 3); }

}  // namespace hpq
