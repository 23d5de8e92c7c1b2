// Vector-Jacobian products of the particle operations of one slice, the building blocks of
// the adjoint slice loop (no counterpart in the reference).
//
// Each kernel restates one forward kernel's per-particle arithmetic on forward-mode
// tangents (tangent.hpp): the per-particle inputs (position, momenta, psi) are seeded, the
// forward expressions are evaluated once, and the derivative parts of the adjoint-weighted
// outputs are the particle's share of the product. Field values enter linearly, so their
// adjoints are scattered to the cells with the forward weights. Branches (QSA removal,
// boundary conditions, shape-factor pieces) follow the values and contribute no derivative.
// Plain FP64 arithmetic throughout; one thread per particle in index order. The field
// adjoints of a workgroup's particles are summed in an LDS window over the bounding box of
// their stencils (consecutive plasma particles sit in neighbouring cells) and the window is
// added to the adjoint planes with one FP64 atomic per touched cell; a workgroup whose box
// does not fit adds every contribution with a global atomic.
#include <climits>
#include <cmath>

#include "internal.hpp"
#include "push.hpp"
#include "shape.hpp"
#include "tangent.hpp"

namespace hpq {

namespace {

constexpr int VT = 128;   // threads per workgroup of the adjoint kernels

__device__ __forceinline__ bool in_store(const Geom& g, int i, int j) {
    return i >= -g.ng && i <= g.nx - 1 + g.ng && j >= -g.ng && j <= g.ny - 1 + g.ng;
}
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

constexpr int VJ_T = 256;      // threads (particles) per workgroup of the scattering VJPs
constexpr int VJ_WIN = 1280;   // window cells per adjoint plane

// LDS window of cells [i0, i0 + w) x [j0, j0 + h) holding np adjoint planes; w == 0: none
struct AdjWin {
    int i0 = 0, j0 = 0, w = 0, h = 0;
    __device__ int cells() const { return w * h; }
};

// the bounding box of the S x S stencils starting at (ci, cj) of the workgroup's active
// particles, when it fits in cap cells (every thread of the workgroup calls this)
__device__ AdjWin adj_window(int ci, int cj, bool act, int S, int cap, int* red) {
    int mni = act ? ci : INT_MAX, mxi = act ? ci : INT_MIN, mnj = act ? cj : INT_MAX, mxj = act ? cj : INT_MIN;
    for (int o = 32; o > 0; o >>= 1) {
        mni = min(mni, __shfl_xor(mni, o)); mxi = max(mxi, __shfl_xor(mxi, o));
        mnj = min(mnj, __shfl_xor(mnj, o)); mxj = max(mxj, __shfl_xor(mxj, o));
    }
    constexpr int NW = VJ_T / 64;
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv] = mni; red[NW + wv] = mxi; red[2 * NW + wv] = mnj; red[3 * NW + wv] = mxj;
    }
    __syncthreads();
    mni = red[0]; mxi = red[NW]; mnj = red[2 * NW]; mxj = red[3 * NW];
    for (int w = 1; w < NW; ++w) {
        mni = min(mni, red[w]); mxi = max(mxi, red[NW + w]);
        mnj = min(mnj, red[2 * NW + w]); mxj = max(mxj, red[3 * NW + w]);
    }
    AdjWin W;
    if (mni > mxi) return W;
    const long bw = long(mxi) - mni + S, bh = long(mxj) - mnj + S;
    if (bw * bh > cap) return W;
    W.i0 = mni; W.j0 = mnj; W.w = int(bw); W.h = int(bh);
    return W;
}

// adds the window's np planes to the adjoint planes comp[0..np) with one atomic per nonzero
// cell (every thread of the workgroup calls this, after the accumulation's barrier)
template <int NP>
__device__ void adj_flush(const Geom& g, const AdjWin& W, const double* win, double* adj, const int* comp) {
    const int cells = W.cells();
    for (int t = threadIdx.x; t < NP * cells; t += VJ_T) {
        const double v = win[t];
        if (v == 0.0) continue;
        const int q = t / cells, c = t - q * cells;
        const int j = W.j0 + c / W.w, i = W.i0 + c % W.w;
        atomic_add(adj + long(comp[q]) * g.plane + g.idx(i, j), v);
    }
}

// ---- ::DepositCurrent (PlasmaDepositCurrent.cpp:182-251) ----------------------------------
// inputs seeded: x, y, ux, uy, psi
template <int O>
__global__ __launch_bounds__(VT) void k_deposit_vjp(Geom g, hpq_plasma p, hpq_plasma_adj a, const double* __restrict__ adj,
                                                    int cjx, int cjy, int cjz, int crho, int cchi, int crm,
                                                    double charge_invvol, double charge_mu0_mass_ratio,
                                                    double max_qsa) {
    using T = Tan<5>;
    constexpr int S = O + 1;
    const long ip = long(blockIdx.x) * VT + threadIdx.x;
    if (ip >= p.n || !p.valid[ip]) return;
    const double clight = 1.0, clightinv = 1.0 / 1.0;
    const T x = T::seed(p.x[ip], 0), y = T::seed(p.y[ip], 1);
    const T ux = T::seed(p.ux[ip], 2), uy = T::seed(p.uy[ip], 3), psi = T::seed(p.psi[ip], 4);
    const T psi_inv = 1.0 / psi;
    const T vx = ux * psi_inv, vy = uy * psi_inv;
    const double q_invvol = charge_invvol * p.w[ip];
    const T xmid = (x - g.xoff) * g.dxi, ymid = (y - g.yoff) * g.dyi;
    const double Aabssqp = 0.0;
    const T gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * psi_inv * psi_inv + vx * vx * clightinv * clightinv +
                               vy * vy * clightinv * clightinv + 1.0);
    if (gamma_psi.v < 0.0 || gamma_psi.v > max_qsa || psi_inv.v < 0.0) return;   // removed by the forward
    const int comp[6] = {cjx, cjy, cjz, crho, cchi, crm};
    T acc(0.0);
#pragma unroll
    for (int iy = 0; iy < S; ++iy) {
        const SfR<T> Y = shape<O, T, T>(ymid, iy);
#pragma unroll
        for (int ix = 0; ix < S; ++ix) {
            const SfR<T> X = shape<O, T, T>(xmid, ix);
            if (!in_store(g, X.cell, Y.cell)) continue;
            const long k = g.idx(X.cell, Y.cell);
            const T cd = q_invvol * X.s * Y.s;
            const T v[6] = {cd * vx, cd * vy, cd * (gamma_psi - 1.0) * clight, cd * gamma_psi,
                            cd * charge_mu0_mass_ratio * psi_inv, cd};
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (comp[q] >= 0) acc += adj[long(comp[q]) * g.plane + k] * v[q];
        }
    }
    a.x[ip] += acc.d[0];
    a.y[ip] += acc.d[1];
    a.ux[ip] += acc.d[2];
    a.uy[ip] += acc.d[3];
    a.psi[ip] += acc.d[4];
}

// ---- ::ExplicitDeposition (ExplicitDeposition.cpp:166-262) --------------------------------
// particle inputs seeded (x, y, ux, uy, psi); the Sy, Sx adjoints are read from their
// planes, the adjoints of Bz, Ez, ExmBy, EypBx are added to theirs.
// The per-particle coefficients depend on (ux, uy, psi) only and the shape factors on x (or
// y) only, so they carry three and one tangent directions instead of five: each product of a
// coefficient with a shape-factor product is expanded by hand into its value, its three
// momentum parts and its x and y parts. The terms dropped are the exact zeros a five-direction
// product would add, so the sums (taken in the same order) are those of the full tangents.
template <int O, int TD>
__global__ __launch_bounds__(VJ_T) void k_explicit_vjp(Geom g, hpq_plasma p, hpq_plasma_adj a,
                                                       const double* __restrict__ store, double* __restrict__ adj,
                                                       Layout L, double charge_invvol_mu0, double charge_mass_ratio) {
    using U = Tan<3>;   // directions ux, uy, psi
    using X1 = Tan<1>;  // direction x (or y)
    constexpr int S = dshape_width<TD, O>();
    constexpr int NP = 4;   // adjoint planes: Bz, Ez, ExmBy, EypBx
    __shared__ double win[NP * VJ_WIN];
    __shared__ int red[4 * VJ_T / 64];
    const long ip = long(blockIdx.x) * VJ_T + threadIdx.x;
    const bool act = ip < p.n && p.valid[ip];
    int ci = 0, cj = 0;
    if (act) {
        ci = dshape<TD, O>((p.x[ip] - g.xoff) * g.dxi, 0).cell;
        cj = dshape<TD, O>((p.y[ip] - g.yoff) * g.dyi, 0).cell;
    }
    const AdjWin W = adj_window(ci, cj, act, S, VJ_WIN, red);
    const int cells = W.cells();
    for (int t = threadIdx.x; t < NP * cells; t += VJ_T) win[t] = 0.0;
    __syncthreads();
    const int comp[NP] = {L.Bz, L.Ez, L.ExmBy, L.EypBx};
    auto scatter = [&](int q, int i, int j, double v) {
        if (W.w) lds_add(&win[q * cells + (j - W.j0) * W.w + (i - W.i0)], v);
        else atomic_add(adj + long(comp[q]) * g.plane + g.idx(i, j), v);
    };
    if (act) {
        const double clight = 1.0, clight_inv = 1.0 / 1.0;
        const U ux = U::seed(p.ux[ip], 0), uy = U::seed(p.uy[ip], 1), psi = U::seed(p.psi[ip], 2);
        const U pinv = 1.0 / psi;
        const U pvx = ux * pinv * clight_inv, pvy = uy * pinv * clight_inv;
        const double cdm = charge_invvol_mu0 * p.w[ip];
        const X1 xmid = (X1::seed(p.x[ip], 0) - g.xoff) * g.dxi, ymid = (X1::seed(p.y[ip], 0) - g.yoff) * g.dyi;
        const double Aabssqp = 0.0;
        const U gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * pinv * pinv + pvx * pvx + pvy * pvy + 1.0);
        const U vxvy = pvx * pvy;
        const U gpx = gamma_psi - pvx * pvx, gpy = gamma_psi - pvy * pvy;
        const U qp = charge_mass_ratio * pinv;
        const U ay = -cdm * qp, by = cdm * clight * g.dxi * vxvy, cy = -cdm * clight * g.dyi * (gpy - 1.0);
        const U ax = cdm * qp, bx = cdm * clight * g.dxi * (gpx - 1.0), cx = -cdm * clight * g.dyi * vxvy;
        const double* Bz = store + long(L.Bz) * g.plane;
        const double* Ez = store + long(L.Ez) * g.plane;
        const double* ExmBy = store + long(L.ExmBy) * g.plane;
        const double* EypBx = store + long(L.EypBx) * g.plane;
        const double* aSy = adj + long(L.Sy) * g.plane;
        const double* aSx = adj + long(L.Sx) * g.plane;
        // accumulated derivative parts: x, y, then ux, uy, psi
        double acc_x = 0.0, acc_y = 0.0, acc_u[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int iy = 0; iy < S; ++iy) {
            const DsfR<X1> Y = dshape<TD, O, X1, X1>(ymid, iy);
#pragma unroll
            for (int ix = 0; ix < S; ++ix) {
                if (TD == 2 && (ix == 0 || ix == O + 2) && (iy == 0 || iy == O + 2)) continue;
                const DsfR<X1> X = dshape<TD, O, X1, X1>(xmid, ix);
                if (!in_store(g, X.cell, Y.cell)) continue;
                const long k = g.idx(X.cell, Y.cell);
                const double sy_a = aSy[k], sx_a = aSx[k];
                if (sy_a == 0.0 && sx_a == 0.0) continue;
                const U A = -Bz[k] * pvx + (EypBx[k] * gpy - ExmBy[k] * vxvy + Ez[k] * pvy) * clight_inv;
                const U B = Bz[k] * pvy + (-EypBx[k] * vxvy + ExmBy[k] * gpx + Ez[k] * pvx) * clight_inv;
                // ss = X.s Y.s, ds = X.ds Y.s, sd = X.s Y.ds: values and x, y parts
                const double ss = X.s.v * Y.s.v, ds = X.ds.v * Y.s.v, sd = X.s.v * Y.ds.v;
                const double ss_x = X.s.d[0] * Y.s.v, ds_x = X.ds.d[0] * Y.s.v, sd_x = X.s.d[0] * Y.ds.v;
                const double ss_y = X.s.v * Y.s.d[0], ds_y = X.ds.v * Y.s.d[0], sd_y = X.s.v * Y.ds.d[0];
                const U ayA = ay * A, axB = ax * B;
                // sy_a (ayA ss + by ds + cy sd) + sx_a (axB ss + bx ds + cx sd), part by part
                acc_x += sy_a * (ayA.v * ss_x + by.v * ds_x + cy.v * sd_x) + sx_a * (axB.v * ss_x + bx.v * ds_x + cx.v * sd_x);
                acc_y += sy_a * (ayA.v * ss_y + by.v * ds_y + cy.v * sd_y) + sx_a * (axB.v * ss_y + bx.v * ds_y + cx.v * sd_y);
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    acc_u[q] += sy_a * (ayA.d[q] * ss + by.d[q] * ds + cy.d[q] * sd) +
                                sx_a * (axB.d[q] * ss + bx.d[q] * ds + cx.d[q] * sd);
                // the contributions are linear in the four fields of cell k
                const double wy = sy_a * ay.v * ss, wx = sx_a * ax.v * ss;
                scatter(0, X.cell, Y.cell, -wy * pvx.v + wx * pvy.v);
                scatter(1, X.cell, Y.cell, (wy * pvy.v + wx * pvx.v) * clight_inv);
                scatter(2, X.cell, Y.cell, (-wy * vxvy.v + wx * gpx.v) * clight_inv);
                scatter(3, X.cell, Y.cell, (wy * gpy.v - wx * vxvy.v) * clight_inv);
            }
        }
        a.x[ip] += acc_x;
        a.y[ip] += acc_y;
        a.ux[ip] += acc_u[0];
        a.uy[ip] += acc_u[1];
        a.psi[ip] += acc_u[2];
    }
    __syncthreads();
    if (W.w) adj_flush<NP>(g, W, win, adj, comp);
}

// ---- AdvancePlasmaParticles (PlasmaParticleAdvance.cpp:84-240), not a temporary slice -----
// The push of one particle as a function of 11 inputs: x_prev, y_prev, ux_half_step,
// uy_half_step, psi_half_step and the six gathered field values G (x_prev, y_prev also
// enter through G; that part is taken by the gather kernel below). One launch evaluates
// the tangents of NB of the inputs, starting at input k0: 11 inputs in three launches keeps
// the tangent arithmetic in registers. The adjoints of G go to a scratch array.
constexpr int PUSH_IN = 11;
constexpr int PUSH_T0 = 2;   // first input that needs a tangent direction

template <int N>
struct PushOut {
    Tan<N> x, y, uxh, uyh, psih, ux, uy, psi;
    bool removed;
    double rx, ry;   // d x / d x_prev, d y / d y_prev: -1 after a reflection, else 1
};

// EnforceBC (GetAndSetPosition.H:29-99) on tangents: folds and shifts by constants keep the
// derivative, a reflection negates it
template <int N>
__device__ bool enforce_bc_tan(const BC& bc, Tan<N>& x, Tan<N>& y, Tan<N>& ux, Tan<N>& uy, double& rx,
                               double& ry) {
    rx = 1.0;
    ry = 1.0;
    if (x.v < bc.plo0 || y.v < bc.plo1 || x.v > bc.phi0 || y.v > bc.phi1) {
        const double lx = bc.phi0 - bc.plo0, ly = bc.phi1 - bc.plo1;
        auto shift = [](Tan<N>& t, double nv) { t.v = nv; };
        if (bc.type == HPQ_BC_REFLECTING) {
            double xv = fmod(x.v - bc.plo0, 2 * lx); if (xv < 0) xv += 2 * lx; xv += bc.plo0;
            shift(x, xv);
            if (x.v > bc.phi0) { x = 2 * bc.phi0 - x; ux = -ux; rx = -1.0; }
            double yv = fmod(y.v - bc.plo1, 2 * ly); if (yv < 0) yv += 2 * ly; yv += bc.plo1;
            shift(y, yv);
            if (y.v > bc.phi1) { y = 2 * bc.phi1 - y; uy = -uy; ry = -1.0; }
            return false;
        } else if (bc.type == HPQ_BC_PERIODIC) {
            double xv = fmod(x.v - bc.plo0, lx); if (xv < 0) xv += lx; xv += bc.plo0;
            shift(x, xv);
            double yv = fmod(y.v - bc.plo1, ly); if (yv < 0) yv += ly; yv += bc.plo1;
            shift(y, yv);
            return false;
        }
        return true;
    }
    return false;
}

template <int O, int N>
__device__ PushOut<N> push_tan(const Geom& g, const hpq_plasma& p, long ip, const double* store, const Layout& L,
                               double qmc, const BC& bc, int k0) {
    using T = Tan<N>;
    auto seed = [&](double v, int input) { return (input >= k0 && input < k0 + N) ? T::seed(v, input - k0) : T(v); };
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double dz = g.dz / 1;
    T xp = seed(p.x_prev[ip], 0), yp = seed(p.y_prev[ip], 1);
    double Gv[6] = {0, 0, 0, 0, 0, 0};   // ExmBy, EypBx, Ez, Bx, By, Bz at the particle
    {
        const double xm = (p.x_prev[ip] - g.xoff) * g.dxi, ym = (p.y_prev[ip] - g.yoff) * g.dyi;
        constexpr int W = O + 2;
        Dsf sx[W], sy[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            sx[k] = dshape<1, O>(xm, k);
            sy[k] = dshape<1, O>(ym, k);
        }
        const double* Psi = store + long(L.Psi) * g.plane;
        const double* Ez = store + long(L.Ez) * g.plane;
        const double* Bx = store + long(L.Bx) * g.plane;
        const double* By = store + long(L.By) * g.plane;
        const double* Bz = store + long(L.Bz) * g.plane;
#pragma unroll
        for (int iy = 0; iy < W; ++iy)
#pragma unroll
            for (int ix = 0; ix < W; ++ix) {
                const Dsf& X = sx[ix];
                const Dsf& Y = sy[iy];
                if (!in_store(g, X.cell, Y.cell)) continue;
                const long k = g.idx(X.cell, Y.cell);
                Gv[0] += (X.ds * Y.s) * Psi[k] * g.dxi;
                Gv[1] += (X.s * Y.ds) * Psi[k] * g.dyi;
                const double ss = X.s * Y.s;
                Gv[2] += ss * Ez[k];
                Gv[3] += ss * Bx[k];
                Gv[4] += ss * By[k];
                Gv[5] += ss * Bz[k];
            }
    }
    Gv[3] *= clight;
    Gv[4] *= clight;
    T G[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) G[q] = seed(Gv[q], 5 + q);
    constexpr int nsub = 4;
    const double sdz = dz / nsub;
    T ux = seed(p.ux_half_step[ip], 2), uy = seed(p.uy_half_step[ip], 3), psi = seed(p.psi_half_step[ip], 4);
    const T cir(clight_inv), qmr(qmc);
    auto substep = [&]() {
        const T psi_inv = 1.0 / psi;
        T dux, duy, dpsi;
        momentum_rhs<T, T>(ux, uy, psi_inv, G[0], G[1], G[2], G[3], G[4], G[5], cir, qmr, dux, duy, dpsi);
        const Dual<T> uxd{ux, dux}, uyd{uy, duy}, pid{psi_inv, -psi_inv * psi_inv * dpsi};
        Dual<T> ddux(T(0.0)), dduy(T(0.0)), ddpsi(T(0.0));
        momentum_rhs<Dual<T>, T>(uxd, uyd, pid, G[0], G[1], G[2], G[3], G[4], G[5], cir, qmr, ddux, dduy, ddpsi);
        ux = ux + sdz * dux + 0.5 * sdz * sdz * ddux.e;
        uy = uy + sdz * duy + 0.5 * sdz * sdz * dduy.e;
        psi = psi + sdz * dpsi + 0.5 * sdz * sdz * ddpsi.e;
    };
#pragma unroll 1
    for (int s = 0; s < nsub; ++s) substep();
    xp = xp + dz * clight_inv * (ux * (1.0 / psi));
    yp = yp + dz * clight_inv * (uy * (1.0 / psi));
    PushOut<N> o;
    o.removed = enforce_bc_tan(bc, xp, yp, ux, uy, o.rx, o.ry);
    o.x = xp;
    o.y = yp;
    o.uxh = ux;
    o.uyh = uy;
    o.psih = psi;
#pragma unroll 1
    for (int s = 0; s < nsub / 2; ++s) substep();
    o.ux = ux;
    o.uy = uy;
    o.psi = psi;
    return o;
}

template <int O, int N>
__global__ __launch_bounds__(VT) void k_push_vjp(Geom g, hpq_plasma p, const double* __restrict__ store, Layout L,
                                                 double qmc, BC bc, hpq_plasma_adj out, hpq_plasma_adj in,
                                                 double* __restrict__ gadj, int k0) {
    const long ip = long(blockIdx.x) * VT + threadIdx.x;
    if (ip >= p.n || !p.valid[ip]) return;
    const PushOut<N> o = push_tan<O, N>(g, p, ip, store, L, qmc, bc, k0);
    double lam[N];
    if (o.removed) {   // absorbed: nothing after this slice depends on the particle
#pragma unroll
        for (int k = 0; k < N; ++k) lam[k] = 0.0;
    } else {
        const double lx = out.x[ip], ly = out.y[ip], luxh = out.ux_half_step[ip], luyh = out.uy_half_step[ip],
                     lph = out.psi_half_step[ip], lux = out.ux[ip], luy = out.uy[ip], lps = out.psi[ip];
#pragma unroll
        for (int k = 0; k < N; ++k)
            lam[k] = lx * o.x.d[k] + ly * o.y.d[k] + luxh * o.uxh.d[k] + luyh * o.uyh.d[k] + lph * o.psih.d[k] +
                     lux * o.ux.d[k] + luy * o.uy.d[k] + lps * o.psi.d[k];
        // x_prev, y_prev enter the outputs only through x = x_prev + ..., y = y_prev + ... (their
        // part through the gathered fields is k_push_vjp_gather's): no tangent pass needed
        if (k0 == PUSH_T0) {
            in.x[ip] += lx * o.rx;
            in.y[ip] += ly * o.ry;
        }
    }
    double* dst[5] = {in.x, in.y, in.ux_half_step, in.uy_half_step, in.psi_half_step};
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int input = k0 + k;
        if (input < 5) dst[input][ip] += lam[k];
        else if (input < PUSH_IN) gadj[long(input - 5) * p.n + ip] = lam[k];
    }
}

// The gathered values G are linear in the cell values and depend on (x_prev, y_prev)
// through the shape factors: scatter the G adjoints with the forward's weights and add
// their position part to the x_prev, y_prev adjoints (B*c scaled by c).
template <int O>
__global__ __launch_bounds__(VJ_T) void k_push_vjp_gather(Geom g, hpq_plasma p, const double* __restrict__ store,
                                                          Layout L, const double* __restrict__ gadj,
                                                          hpq_plasma_adj in, double* __restrict__ adj) {
    // the shape factors in x carry the x tangent only, those in y the y tangent only; the
    // products below are expanded by hand into value, x and y parts (the dropped terms are the
    // exact zeros of two-direction products, the sums keep their order)
    using X1 = Tan<1>;
    constexpr int W = O + 2;
    constexpr int NP = 5;   // adjoint planes: Psi, Ez, Bx, By, Bz
    __shared__ double win[NP * VJ_WIN];
    __shared__ int red[4 * VJ_T / 64];
    const long ip = long(blockIdx.x) * VJ_T + threadIdx.x;
    double lg[6] = {0, 0, 0, 0, 0, 0};
    bool act = ip < p.n && p.valid[ip];
    if (act) {
        bool any = false;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            lg[q] = gadj[long(q) * p.n + ip];
            any = any || lg[q] != 0.0;
        }
        act = any;
    }
    int ci = 0, cj = 0;
    if (act) {
        ci = dshape<1, O>((p.x_prev[ip] - g.xoff) * g.dxi, 0).cell;
        cj = dshape<1, O>((p.y_prev[ip] - g.yoff) * g.dyi, 0).cell;
    }
    const AdjWin Wn = adj_window(ci, cj, act, W, VJ_WIN, red);
    const int cells = Wn.cells();
    for (int t = threadIdx.x; t < NP * cells; t += VJ_T) win[t] = 0.0;
    __syncthreads();
    const int comp[NP] = {L.Psi, L.Ez, L.Bx, L.By, L.Bz};
    auto scatter = [&](int q, int i, int j, double v) {
        if (Wn.w) lds_add(&win[q * cells + (j - Wn.j0) * Wn.w + (i - Wn.i0)], v);
        else atomic_add(adj + long(comp[q]) * g.plane + g.idx(i, j), v);
    };
    if (act) {
        const double clight = 1.0;
        const X1 xm = (X1::seed(p.x_prev[ip], 0) - g.xoff) * g.dxi, ym = (X1::seed(p.y_prev[ip], 0) - g.yoff) * g.dyi;
        DsfR<X1> sx[W], sy[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            sx[k] = dshape<1, O, X1, X1>(xm, k);
            sy[k] = dshape<1, O, X1, X1>(ym, k);
        }
        const double* F[5] = {store + long(L.Psi) * g.plane, store + long(L.Ez) * g.plane,
                              store + long(L.Bx) * g.plane, store + long(L.By) * g.plane, store + long(L.Bz) * g.plane};
        double acc_x = 0.0, acc_y = 0.0;
#pragma unroll
        for (int iy = 0; iy < W; ++iy)
#pragma unroll
            for (int ix = 0; ix < W; ++ix) {
                const DsfR<X1>& X = sx[ix];
                const DsfR<X1>& Y = sy[iy];
                if (!in_store(g, X.cell, Y.cell)) continue;
                const long k = g.idx(X.cell, Y.cell);
                const double ss = X.s.v * Y.s.v, ss_x = X.s.d[0] * Y.s.v, ss_y = X.s.v * Y.s.d[0];
                const double ds = X.ds.v * Y.s.v, ds_x = X.ds.d[0] * Y.s.v, ds_y = X.ds.v * Y.s.d[0];
                const double sd = X.s.v * Y.ds.v, sd_x = X.s.d[0] * Y.ds.v, sd_y = X.s.v * Y.ds.d[0];
                // wpsi = lg0 (X.ds Y.s) dxi + lg1 (X.s Y.ds) dyi; wez, wbx, wby, wbz = lg * ss
                const double wpsi = lg[0] * ds * g.dxi + lg[1] * sd * g.dyi;
                const double wpsi_x = lg[0] * ds_x * g.dxi + lg[1] * sd_x * g.dyi;
                const double wpsi_y = lg[0] * ds_y * g.dxi + lg[1] * sd_y * g.dyi;
                const double c2 = lg[2], c3 = lg[3] * clight, c4 = lg[4] * clight, c5 = lg[5];
                const double f0 = F[0][k], f1 = F[1][k], f2 = F[2][k], f3 = F[3][k], f4 = F[4][k];
                acc_x += wpsi_x * f0 + (c2 * ss_x) * f1 + (c3 * ss_x) * f2 + (c4 * ss_x) * f3 + (c5 * ss_x) * f4;
                acc_y += wpsi_y * f0 + (c2 * ss_y) * f1 + (c3 * ss_y) * f2 + (c4 * ss_y) * f3 + (c5 * ss_y) * f4;
                const double w5[5] = {wpsi, c2 * ss, c3 * ss, c4 * ss, c5 * ss};
#pragma unroll
                for (int q = 0; q < 5; ++q)
                    if (w5[q] != 0.0) scatter(q, X.cell, Y.cell, w5[q]);
            }
        in.x[ip] += acc_x;
        in.y[ip] += acc_y;
    }
    __syncthreads();
    if (Wn.w) adj_flush<NP>(g, Wn, win, adj, comp);
}

}  // namespace

void deposit_vjp(hpq_ctx* c, const hpq_plasma* p, const double* adj, const int32_t* comps, hpq_plasma_adj* a) {
    if (p->n <= 0) return;
    const Geom& g = c->g;
    const double charge = c->p.plasma_charge;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double cmr = charge * 1.0 / c->p.plasma_mass;
    auto cm = [&](int k) { return comps[k] >= 0 ? comps[k] : -1; };
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_deposit_vjp<decltype(O)::value>), grid1d(p->n, VT), dim3(VT), 0, c->stream, g, *p, *a,
                           adj, cm(0), cm(1), cm(2), cm(3), cm(4), cm(5), charge * invvol, cmr, c->p.max_qsa);
    });
    HPQ_HIP(hipGetLastError());
}

void explicit_deposition_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store, double* adj, hpq_plasma_adj* a) {
    if (p->n <= 0) return;
    const Geom& g = c->g;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double qimu0 = c->p.plasma_charge * invvol * 1.0;
    const double qmr = c->p.plasma_charge / c->p.plasma_mass;
    const int T = c->p.derivative_type;
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid1d(p->n, VJ_T), dim3(VJ_T), 0, c->stream, g, *p, *a, store, adj, c->L, qimu0,
                               qmr);
        };
        if (T == 0) launch(k_explicit_vjp<o, 0>);
        else if (T == 1) launch(k_explicit_vjp<o, 1>);
        else launch(k_explicit_vjp<o, 2>);
    });
    HPQ_HIP(hipGetLastError());
}

void advance_plasma_vjp(hpq_ctx* c, const hpq_plasma* p, const double* store, const hpq_plasma_adj* out,
                        hpq_plasma_adj* in, double* adj) {
    if (p->n <= 0) return;
    const double qmc = c->p.plasma_charge / (c->p.plasma_mass * 1.0);
    BC bc{c->p.particle_bc, c->p.plo[0], c->p.plo[1], c->p.phi[0], c->p.phi[1]};
    constexpr int NB = 3;
    AdjointState& A = c->adj;
    if (A.push_g_cap < 6 * p->n) {
        if (A.push_g) (void)hipFree(A.push_g);
        A.push_g = nullptr;
        HPQ_HIP(hipMalloc(&A.push_g, sizeof(double) * 6 * p->n));
        A.push_g_cap = 6 * p->n;
    }
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        for (int k0 = PUSH_T0; k0 < PUSH_IN; k0 += NB)
            hipLaunchKernelGGL((k_push_vjp<o, NB>), grid1d(p->n, VT), dim3(VT), 0, c->stream, c->g, *p, store, c->L,
                               qmc, bc, *out, *in, A.push_g, k0);
        hipLaunchKernelGGL((k_push_vjp_gather<o>), grid1d(p->n, VJ_T), dim3(VJ_T), 0, c->stream, c->g, *p, store, c->L,
                           A.push_g, *in, adj);
    });
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq

// ---------------------------------------------------------------------------------------
// Linear parts of the adjoint slice step
// ---------------------------------------------------------------------------------------
namespace hpq {

namespace {

__device__ __forceinline__ bool in_valid(const Geom& g, int i, int j) { return i >= 0 && i < g.nx && j >= 0 && j < g.ny; }

// Staged right-hand sides of the adjoint Poisson solves: the Psi adjoint (valid cells) plus
// the transposed centred differences of ExmBy = -d_x Psi, EypBx = -d_y Psi (defined on the
// valid box plus one ring, Fields.cpp:931-956), and the Ez and Bz adjoints.
template <class T>
__global__ void k_adj_stage(Geom g, const double* __restrict__ apsi, const double* __restrict__ aexmby,
                            const double* __restrict__ aeypbx, const double* __restrict__ aez,
                            const double* __restrict__ abz, T* __restrict__ st, long ld, long fs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const double v = apsi[g.idx(i, j)] - aexmby[g.idx(i - 1, j)] * dxih + aexmby[g.idx(i + 1, j)] * dxih -
                     aeypbx[g.idx(i, j - 1)] * dyih + aeypbx[g.idx(i, j + 1)] * dyih;
    const long o = long(j) * ld + i;
    st[o] = T(v);
    st[fs + o] = T(aez[g.idx(i, j)]);
    st[2 * fs + o] = T(abz[g.idx(i, j)]);
}

// Transposes of the right-hand sides of Fields::SolvePoissonPsiExmByEypBxEzBz
// (Fields.cpp:886-917) for every cell of the guarded planes: r_* are the adjoint solves
// (valid cells), results are added to the rhomjz, jx, jy adjoints.
__global__ void k_adj_rhs(Geom g, const double* __restrict__ rpsi, const double* __restrict__ rez,
                          const double* __restrict__ rbz, double* __restrict__ arhomjz, double* __restrict__ ajx,
                          double* __restrict__ ajy) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x) - g.ng;
    const int j = int(blockIdx.y) - g.ng;
    if (i > g.nx - 1 + g.ng) return;
    const double ep0 = 1.0, mu0 = 1.0, cl = 1.0;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const double fe = 1.0 / (ep0 * cl);
    auto R = [&](const double* r, int a, int b) { return in_valid(g, a, b) ? r[g.idx(a, b)] : 0.0; };
    const long k = g.idx(i, j);
    if (in_valid(g, i, j)) arhomjz[k] += (-1.0 / ep0) * rpsi[k];
    ajx[k] += fe * (R(rez, i - 1, j) - R(rez, i + 1, j)) * dxih + mu0 * (R(rbz, i, j - 1) - R(rbz, i, j + 1)) * dyih;
    ajy[k] += fe * (R(rez, i, j - 1) - R(rez, i, j + 1)) * dyih - mu0 * (R(rbz, i - 1, j) - R(rbz, i + 1, j)) * dxih;
}

// chi adjoint of the Bx/By solve (Lap - chi) B = S: chi_adj += mu_x Bx + mu_y By
__global__ void k_adj_chi(Geom g, const double* __restrict__ Bx, const double* __restrict__ By,
                          const double* __restrict__ mux, const double* __restrict__ muy, double* __restrict__ achi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const long k = g.idx(i, j);
    achi[k] += mux[k] * Bx[k] + muy[k] * By[k];
}

// loss seed of Ez at this slice ([j][i] over the valid box) into the Ez adjoint
__global__ void k_adj_seed(Geom g, const double* __restrict__ seed, double* __restrict__ aez) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    aez[g.idx(i, j)] += seed[long(j) * g.nx + i];
}

// Hipace::InitializeSxSyWithBeam (Hipace.cpp:745-790) transposed for the beam jz: Sy has
// -d_y jz_b, Sx has +d_x jz_b on valid cells; the jz_b adjoint of every cell
__global__ void k_adj_sxsy_jzb(Geom g, const double* __restrict__ asy, const double* __restrict__ asx,
                               double* __restrict__ ajzb) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x) - g.ng;
    const int j = int(blockIdx.y) - g.ng;
    if (i > g.nx - 1 + g.ng) return;
    const double mu0 = 1.0;
    auto A = [&](const double* a, int p, int q) { return in_valid(g, p, q) ? a[g.idx(p, q)] : 0.0; };
    // Sy(i, j) = -(jz_b(i, j+1) - jz_b(i, j-1)) / 2dy + ..., Sx(i, j) = (jz_b(i+1, j) - jz_b(i-1, j)) / 2dx + ...
    const double v = mu0 * (A(asy, i, j + 1) - A(asy, i, j - 1)) / (2.0 * g.dy) +
                     mu0 * (A(asx, i - 1, j) - A(asx, i + 1, j)) / (2.0 * g.dx);
    ajzb[g.idx(i, j)] = v;
}

// ::DepositCurrentSlice jz_b (BeamDepositCurrent.cpp:133-193) transposed for the weights:
// w_adj += charge (uz / gamma) sum_cells S jz_b_adj
template <int O>
__global__ void k_beam_w_vjp(Geom g, hpq_beam b, long first, long count, const double* __restrict__ ajzb,
                             double* __restrict__ w_adj) {
    constexpr int S = O + 1;
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const long ip = first + t;
    if (!b.valid[ip]) return;
    const double clightsq = 1.0 / (1.0 * 1.0);
    const double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
    const double gaminv = 1.0 / std::sqrt(1.0 + ux * ux * clightsq + uy * uy * clightsq + uz * uz * clightsq);
    const double xmid = (b.x[ip] - g.xoff) * g.dxi, ymid = (b.y[ip] - g.yoff) * g.dyi;
    double acc = 0.0;
#pragma unroll
    for (int iy = 0; iy < S; ++iy) {
        const Sf Y = shape<O>(ymid, iy);
#pragma unroll
        for (int ix = 0; ix < S; ++ix) {
            const Sf X = shape<O>(xmid, ix);
            if (!in_store(g, X.cell, Y.cell)) continue;
            acc += X.s * Y.s * ajzb[g.idx(X.cell, Y.cell)];
        }
    }
    w_adj[ip] += b.charge * 1.0 * (uz * gaminv) * acc;
}

}  // namespace

void adj_poisson(hpq_ctx* c, double* adj) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    const Poisson& S = c->ps;
    auto pl = [&](int comp) { return adj + long(comp) * g.plane; };
    const dim3 grid((g.nx + 127) / 128, g.ny);
    if (c->p.poisson_solver == HPQ_POISSON_BF16X3)
        hipLaunchKernelGGL(k_adj_stage<float>, grid, dim3(128), 0, c->stream, g, pl(L.Psi), pl(L.ExmBy), pl(L.EypBx),
                           pl(L.Ez), pl(L.Bz), S.gS, long(S.npx), long(S.npx) * S.npy);
    else
        hipLaunchKernelGGL(k_adj_stage<double>, grid, dim3(128), 0, c->stream, g, pl(L.Psi), pl(L.ExmBy),
                           pl(L.EypBx), pl(L.Ez), pl(L.Bz), S.bufA, long(S.nx), long(S.nx) * S.ny);
    HPQ_HIP(hipGetLastError());
    // the solves overwrite the valid cells of the Psi, Ez, Bz adjoints, which the staging
    // has consumed
    double* lhs[3] = {pl(L.Psi), pl(L.Ez), pl(L.Bz)};
    poisson_solve_staged(c, 3, lhs);
    hipLaunchKernelGGL(k_adj_rhs, dim3((g.ldx + 127) / 128, g.ldy), dim3(128), 0, c->stream, g, pl(L.Psi), pl(L.Ez),
                       pl(L.Bz), pl(L.rhomjz), pl(L.jx), pl(L.jy));
    HPQ_HIP(hipGetLastError());
}

void adj_chi(hpq_ctx* c, const double* store, double* adj) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    hipLaunchKernelGGL(k_adj_chi, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g, store + L.Bx * g.plane,
                       store + L.By * g.plane, adj + L.Sy * g.plane, adj + L.Sx * g.plane, adj + L.chi * g.plane);
    HPQ_HIP(hipGetLastError());
}

void adj_seed(hpq_ctx* c, const double* seed, double* adj) {
    const Geom& g = c->g;
    hipLaunchKernelGGL(k_adj_seed, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g, seed,
                       adj + c->L.Ez * g.plane);
    HPQ_HIP(hipGetLastError());
}

void adj_beam_weights(hpq_ctx* c, const hpq_beam* b, int slice, double* adj, const double* jzb_seed,
                      double* w_adj) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    // the jz_b adjoint goes to the jz_beam plane of the adjoint store: through Sx/Sy, plus
    // the loss's own dependence on jz_beam (valid cells)
    hipLaunchKernelGGL(k_adj_sxsy_jzb, dim3((g.ldx + 127) / 128, g.ldy), dim3(128), 0, c->stream, g,
                       adj + L.Sy * g.plane, adj + L.Sx * g.plane, adj + L.jz_beam * g.plane);
    if (jzb_seed)
        hipLaunchKernelGGL(k_adj_seed, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g, jzb_seed,
                           adj + L.jz_beam * g.plane);
    HPQ_HIP(hipGetLastError());
    if (!b || !b->off || !w_adj) return;
    const long first = b->off[slice];
    const long count = b->off[slice + 1] - first;
    if (count <= 0) return;
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_beam_w_vjp<decltype(O)::value>), grid1d(count, 256), dim3(256), 0, c->stream, g, *b,
                           first, count, adj + L.jz_beam * g.plane, w_adj);
    });
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq
