// Oracle shape factors (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
//
// B-spline shape factors of order 0..3 and their derivative variants, restating
// src/particles/particles_utils/ShapeFactors.H:122-195 (compute_single_shape_factor,
// branchy variant) and :211-466 (single_derivative_shape_factor, types 0/1/2).
// The piecewise polynomials are the B-spline pieces B_o(u) evaluated at
// u = xmid - cell; derivative type 0 is dB_o/du, type 1 (nodal) is dB_{o+1}/du on a
// window one cell wider, type 2 (centred) is (B_o(u+1) - B_o(u-1))/2. The returned
// `ds` is the negated table derivative, as the reference returns {s, -sdx, cell}.
#include "oracle.hpp"

namespace orc {

SF shape(int order, double xmid, int ix) {
    if (order == 0) {
        return {1.0, int(std::floor(xmid + 0.5))};
    } else if (order == 1) {
        const double xf = std::floor(xmid);
        const double t = xmid - xf;
        return ix == 0 ? SF{1.0 - t, int(xf)} : SF{t, int(xf) + 1};
    } else if (order == 2) {
        const double xf = std::floor(xmid + 0.5);
        const double t = xmid - xf;
        if (ix == 0) return {0.5 * (0.5 - t) * (0.5 - t), int(xf) - 1};
        if (ix == 1) return {0.75 - t * t, int(xf)};
        return {0.5 * (0.5 + t) * (0.5 + t), int(xf) + 1};
    } else {
        const double xf = std::floor(xmid);
        const double t = xmid - xf;
        if (ix == 0) return {1.0 / 6.0 * (1.0 - t) * (1.0 - t) * (1.0 - t), int(xf) - 1};
        if (ix == 1) return {2.0 / 3.0 - t * t * (1.0 - t / 2.0), int(xf)};
        if (ix == 2) return {2.0 / 3.0 - (1.0 - t) * (1.0 - t) * (1.0 - 0.5 * (1.0 - t)), int(xf) + 1};
        return {1.0 / 6.0 * t * t * t, int(xf) + 2};
    }
}

static DSF mk(double s, double sdx, int cell) { return {s, -sdx, cell}; }

DSF dshape(int type, int order, double xmid, int ix) {
    if (type == 0) {  // analytic derivative of the order-o spline
        if (order == 0) {
            const double xf = std::floor(xmid + 0.5);
            return mk(1.0, 0.0, int(xf));
        } else if (order == 1) {
            const double xf = std::floor(xmid);
            const double t = xmid - xf;
            return ix == 0 ? mk(1 - t, -1.0, int(xf)) : mk(t, 1.0, int(xf) + 1);
        } else if (order == 2) {
            const double xf = std::floor(xmid + 0.5);
            const double t = xmid + 0.5 - xf;
            const double t2 = t * t;
            const int c = int(xf) - 1 + ix;
            if (ix == 0) return mk(0.5 * t2 - 1.0 * t + 0.5, t - 1.0, c);
            if (ix == 1) return mk(-t2 + 1.0 * t + 0.5, 1.0 - 2 * t, c);
            return mk(0.5 * t2, t, c);
        } else {
            const double xf = std::floor(xmid);
            const double t = xmid - xf;
            const double t2 = t * t, t3 = t2 * t;
            const int c = int(xf) - 1 + ix;
            if (ix == 0) return mk(-1.0 / 6.0 * t3 + 0.5 * t2 - 0.5 * t + 1.0 / 6.0, -0.5 * t2 + t - 0.5, c);
            if (ix == 1) return mk(0.5 * t3 - t2 + 2.0 / 3.0, 1.5 * t2 - 2 * t, c);
            if (ix == 2) return mk(-0.5 * t3 + 0.5 * t2 + 0.5 * t + 1.0 / 6.0, -1.5 * t2 + t + 0.5, c);
            return mk(1.0 / 6.0 * t3, 0.5 * t2, c);
        }
    } else if (type == 1) {  // nodal: value of B_o, derivative of B_{o+1}
        if (order == 0) {
            const double xf = std::floor(xmid);
            const double t = xmid - xf;
            return ix == 0 ? mk((t < 0.5) ? 1 : 0, -1, int(xf)) : mk((t < 0.5) ? 0 : 1, 1, int(xf) + 1);
        } else if (order == 1) {
            const double xf = std::floor(xmid + 0.5);
            const double t = xmid + 0.5 - xf;
            const int c = int(xf) - 1 + ix;
            if (ix == 0) return mk((t < 0.5) ? 0.5 - t : 0, t - 1, c);
            if (ix == 1) return mk((t < 0.5) ? t + 0.5 : 1.5 - t, 1 - 2 * t, c);
            return mk((t < 0.5) ? 0 : t - 0.5, t, c);
        } else if (order == 2) {
            const double xf = std::floor(xmid);
            const double t = xmid - xf;
            const double t2 = t * t;
            const int c = int(xf) - 1 + ix;
            if (ix == 0) return mk((t < 0.5) ? 0.5 * t2 - 0.5 * t + 1.0 / 8.0 : 0, -0.5 * t2 + t - 0.5, c);
            if (ix == 1) return mk((t < 0.5) ? 3.0 / 4.0 - t2 : 0.5 * t2 - 1.5 * t + 9.0 / 8.0, 1.5 * t2 - 2 * t, c);
            if (ix == 2) return mk((t < 0.5) ? 0.5 * t2 + 0.5 * t + 1.0 / 8.0 : -t2 + 2 * t - 1.0 / 4.0,
                                   -1.5 * t2 + t + 0.5, c);
            return mk((t < 0.5) ? 0 : 0.5 * t2 - 0.5 * t + 1.0 / 8.0, 0.5 * t2, c);
        } else {
            const double xf = std::floor(xmid + 0.5);
            const double t = xmid + 0.5 - xf;
            const double t2 = t * t, t3 = t2 * t;
            const int c = int(xf) - 2 + ix;
            const bool lo = t < 0.5;
            if (ix == 0) return mk(lo ? -1.0 / 6.0 * t3 + 0.25 * t2 - 1.0 / 8.0 * t + 1.0 / 48.0 : 0,
                                   1.0 / 6.0 * t3 - 0.5 * t2 + 0.5 * t - 1.0 / 6.0, c);
            if (ix == 1) return mk(lo ? 0.5 * t3 - 0.25 * t2 - 5.0 / 8.0 * t + 23.0 / 48.0
                                      : -1.0 / 6.0 * t3 + 0.75 * t2 - 9.0 / 8.0 * t + 9.0 / 16.0,
                                   -2.0 / 3.0 * t3 + 1.5 * t2 - 0.5 * t - 0.5, c);
            if (ix == 2) return mk(lo ? -0.5 * t3 - 0.25 * t2 + 5.0 / 8.0 * t + 23.0 / 48.0
                                      : 0.5 * t3 - 7.0 / 4.0 * t2 + 11.0 / 8.0 * t + 17.0 / 48.0,
                                   t3 - 1.5 * t2 - 0.5 * t + 0.5, c);
            if (ix == 3) return mk(lo ? 1.0 / 6.0 * t3 + 0.25 * t2 + 1.0 / 8.0 * t + 1.0 / 48.0
                                      : -0.5 * t3 + 5.0 / 4.0 * t2 - 3.0 / 8.0 * t + 5.0 / 48.0,
                                   -2.0 / 3.0 * t3 + 0.5 * t2 + 0.5 * t + 1.0 / 6.0, c);
            return mk(lo ? 0 : 1.0 / 6.0 * t3 - 0.25 * t2 + 1.0 / 8.0 * t - 1.0 / 48.0, 1.0 / 6.0 * t3, c);
        }
    } else {  // type 2: centred difference of the order-o spline
        if (order == 0) {
            const double xf = std::floor(xmid + 0.5);
            const int c = int(xf) - 1 + ix;
            if (ix == 0) return mk(0, -0.5, c);
            if (ix == 1) return mk(1, 0, c);
            return mk(0, 0.5, c);
        } else if (order == 1) {
            const double xf = std::floor(xmid);
            const double t = xmid - xf;
            const int c = int(xf) - 1 + ix;
            if (ix == 0) return mk(0, 0.5 * t - 0.5, c);
            if (ix == 1) return mk(1 - t, -0.5 * t, c);
            if (ix == 2) return mk(t, 0.5 - 0.5 * t, c);
            return mk(0, 0.5 * t, c);
        } else if (order == 2) {
            const double xf = std::floor(xmid + 0.5);
            const double t = xmid + 0.5 - xf;
            const double t2 = t * t;
            const int c = int(xf) - 2 + ix;
            if (ix == 0) return mk(0, -0.25 * t2 + 0.5 * t - 0.25, c);
            if (ix == 1) return mk(0.5 * t2 - t + 0.5, 0.5 * t2 - 0.5 * t - 0.25, c);
            if (ix == 2) return mk(-t2 + t + 0.5, 0.25 - 0.5 * t, c);
            if (ix == 3) return mk(0.5 * t2, -0.5 * t2 + 0.5 * t + 0.25, c);
            return mk(0, 0.25 * t2, c);
        } else {
            const double xf = std::floor(xmid);
            const double t = xmid - xf;
            const double t2 = t * t, t3 = t2 * t;
            const int c = int(xf) - 2 + ix;
            if (ix == 0) return mk(0, 1.0 / 12.0 * t3 - 0.25 * t2 + 0.25 * t - 1.0 / 12.0, c);
            if (ix == 1) return mk(-1.0 / 6.0 * t3 + 0.5 * t2 - 0.5 * t + 1.0 / 6.0,
                                   -0.25 * t3 + 0.5 * t2 - 1.0 / 3.0, c);
            if (ix == 2) return mk(0.5 * t3 - t2 + 2.0 / 3.0, 1.0 / 6.0 * t3 - 0.5 * t, c);
            if (ix == 3) return mk(-0.5 * t3 + 0.5 * t2 + 0.5 * t + 1.0 / 6.0,
                                   1.0 / 6.0 * t3 - 0.5 * t2 + 1.0 / 3.0, c);
            if (ix == 4) return mk(1.0 / 6.0 * t3, -0.25 * t3 + 0.25 * t2 + 0.25 * t + 1.0 / 12.0, c);
            return mk(0, 1.0 / 12.0 * t3, c);
        }
    }
}

}  // namespace orc
