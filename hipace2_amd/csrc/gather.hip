// Deposition in the reference CPU build's summation order, on gfx950, without atomics.
//
// The CPU build deposits with SharedMemoryDeposition's tiled loop
// (src/particles/deposition/DepositionUtil.H:204-253; tiling on, 32-cell tiles under OpenMP,
// src/Hipace.H:235-240): particles are binned by the tile of their lowest deposited cell
// (start cell), the tiles are visited in four colours (even/odd tile index in x, then in y),
// and inside a tile the particles go in bin order, which is index order (amrex::DenseBins is a
// stable counting sort). Tiles of one colour never share a cell, so every cell of the slice
// receives its contributions in one fixed sequence: the up to four tiles that reach it in
// colour order, and inside each tile the particles by index.
//
// Here every cell sums its own contributions in that sequence (one thread per cell, the
// running sum starts from the plane's value and each contribution is added in turn), so the
// planes come out bit for bit as the CPU build's loop leaves them, and no two threads write
// the same cell. The particles that reach a cell are those whose start cell lies in the
// S x S window below it (S = stencil width); they are found through per-start-cell lists:
//   keys    the start cell of every particle (and, for the plasma, the per-particle values
//           every contribution needs: q w, u/psi, gamma/psi, 1/psi, the cell-scaled position;
//           QSA violators are removed here, as the deposit functor would remove them);
//   count   particles per start cell, exclusive scan -> list offsets;
//   scatter particles into their lists (atomic slots, any order);
//   order   each list sorted by particle index (rank by counting inside the list);
//   gather  per cell, for each colour in turn, the lists of the window cells in that colour's
//           tile merged by particle index; each particle's contribution is the reference
//           expression evaluated for this cell's stencil offset.
// A start-cell list is keyed by the floor of the cell-scaled position the shape functions
// take their cell from (floor(xmid) or floor(xmid + 1/2)); a stencil's start cell is that
// floor plus a constant, so the explicit Sx/Sy deposition of the default derivative type
// reuses the lists the current deposition built on the same particles.
#include <climits>
#include <cmath>

#include "internal.hpp"
#include "shape.hpp"

namespace hpq {

namespace {

constexpr int GT_X = 16, GT_Y = 16;   // cells per gather workgroup

// start cell = floor(xmid + 0.5 * BASE) + DELTA, for the order-O shape factor and for the
// type-T derivative shape factor (ShapeFactors.H: compute_single_shape_factor,
// single_derivative_shape_factor)
template <int O> constexpr int shape_base() { return (O == 0 || O == 2) ? 1 : 0; }
template <int O> constexpr int shape_delta() { return O >= 2 ? -1 : 0; }
template <int T, int O> constexpr int dshape_base() {
    return T == 0 ? shape_base<O>() : (T == 1 ? ((O == 1 || O == 3) ? 1 : 0) : ((O == 0 || O == 2) ? 1 : 0));
}
template <int T, int O> constexpr int dshape_delta() {
    return T == 0 ? shape_delta<O>() : (T == 1 ? (O == 0 ? 0 : (O == 3 ? -2 : -1)) : (O >= 2 ? -2 : -1));
}

// lists of one particle set keyed by base cell: base cells [blo, blo + w) in each direction;
// tiles of the deposition order: T cells (hipace.tile_size), ntx x nty over the guarded box
struct CLView {
    int bxlo, bylo, wx, wy;
    int T, ntx, nty;
    const int32_t* __restrict__ off;    // wx * wy + 1 list offsets
    const int32_t* __restrict__ perm;   // particle indices, list by list, increasing in a list
};

// the tile of start cell s along one direction (C division truncates toward zero, as the
// reference's (cell - lo) / tile does; clamped to the tile grid)
__device__ __forceinline__ int dep_tile(int s, int ng, int T, int nt) { return min(max((s + ng) / T, 0), nt - 1); }

// Calls f(particle, ix, iy) for every particle of the lists whose start cell lies in the
// S x S window below cell (i, j), in the CPU build's order: colour by colour, and inside the
// window part of one colour by increasing particle index (a merge of the lists). (ix, iy) is
// the particle's stencil offset that lands on (i, j).
template <int S, int DELTA, class F>
__device__ __forceinline__ void walk_ordered(const Geom& g, const CLView& cl, int i, int j, F&& f) {
    const int x0 = i - S + 1, y0 = j - S + 1;
    int h[S * S], e[S * S];
#pragma unroll
    for (int ky = 0; ky < S; ++ky) {
        const int row = (y0 + ky - DELTA - cl.bylo) * cl.wx + (x0 - DELTA - cl.bxlo);
        int o[S + 1];
#pragma unroll
        for (int kx = 0; kx <= S; ++kx) o[kx] = cl.off[row + kx];
#pragma unroll
        for (int kx = 0; kx < S; ++kx) {
            h[ky * S + kx] = o[kx];
            e[ky * S + kx] = o[kx + 1];
        }
    }
    // window split at tile boundaries: columns [x0, xs) in tile txa, [xs, i] in tile txb
    // (S <= T: a window spans at most two tiles in each direction)
    const int txa = dep_tile(x0, g.ng, cl.T, cl.ntx), txb = dep_tile(i, g.ng, cl.T, cl.ntx);
    const int tya = dep_tile(y0, g.ng, cl.T, cl.nty), tyb = dep_tile(j, g.ng, cl.T, cl.nty);
    const int xs = txa == txb ? i + 1 : txb * cl.T - g.ng;
    const int ys = tya == tyb ? j + 1 : tyb * cl.T - g.ng;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        // the x part and the y part of this colour (empty when no window tile has it)
        int ax, bx, ay, by;
        if ((txa & 1) == (c >> 1)) { ax = x0; bx = xs - 1; }
        else if ((txb & 1) == (c >> 1)) { ax = xs; bx = i; }
        else { ax = 1; bx = 0; }
        if ((tya & 1) == (c & 1)) { ay = y0; by = ys - 1; }
        else if ((tyb & 1) == (c & 1)) { ay = ys; by = j; }
        else { ay = 1; by = 0; }
        if (ax > bx || ay > by) continue;
        int v[S * S];
#pragma unroll
        for (int k = 0; k < S * S; ++k) {
            const int sx = x0 + k % S, sy = y0 + k / S;
            const bool act = sx >= ax && sx <= bx && sy >= ay && sy <= by && h[k] < e[k];
            v[k] = act ? cl.perm[h[k]] : INT_MAX;
        }
        while (true) {
            int m1 = INT_MAX, m2 = INT_MAX, km = -1;
#pragma unroll
            for (int k = 0; k < S * S; ++k) {
                if (v[k] < m1) {
                    m2 = m1;
                    m1 = v[k];
                    km = k;
                } else if (v[k] < m2) {
                    m2 = v[k];
                }
            }
            if (km < 0) break;
            int hc = 0, ec = 0;
#pragma unroll
            for (int k = 0; k < S * S; ++k)
                if (k == km) { hc = h[k]; ec = e[k]; }
            const int ix = S - 1 - km % S, iy = S - 1 - km / S;
            // the list holding the smallest index goes on while it stays below every other head
            int p = m1;
            while (true) {
                f(p, ix, iy);
                ++hc;
                p = hc < ec ? cl.perm[hc] : INT_MAX;
                if (p >= m2) break;
            }
#pragma unroll
            for (int k = 0; k < S * S; ++k)
                if (k == km) { h[k] = hc; v[k] = p; }
        }
    }
}

// ---- list construction --------------------------------------------------------------------

// Plasma keys for the current deposition (PlasmaDepositCurrent.cpp:182-243): per particle the
// values of the deposit functor that do not depend on the cell, and the QSA removal.
// rec: [0] q_invvol, [1] vx_c, [2] vy_c, [3] gamma/psi, [4] 1/psi, [5] xmid, [6] ymid.
template <int O>
__global__ __launch_bounds__(256) void k_keys_plasma(Geom g, hpq_plasma p, CLView cl, double charge_invvol,
                                                     double max_qsa, int check_qsa, int32_t* __restrict__ key,
                                                     int32_t* __restrict__ slot, int32_t* __restrict__ count,
                                                     double* __restrict__ rec, int* __restrict__ qsa_count) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= p.n) return;
    int k = -1;
    if (p.valid[ip]) {
        const double clightinv = 1.0 / 1.0;
        const double psi_inv = 1.0 / p.psi[ip];
        const double vx_c = p.ux[ip] * psi_inv;
        const double vy_c = p.uy[ip] * psi_inv;
        const double q_invvol = charge_invvol * p.w[ip];
        const double xmid = (p.x[ip] - g.xoff) * g.dxi;
        const double ymid = (p.y[ip] - g.yoff) * g.dyi;
        const double Aabssqp = 0.0;
        const double gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * psi_inv * psi_inv +
                                        vx_c * vx_c * clightinv * clightinv +
                                        vy_c * vy_c * clightinv * clightinv + 1.0);
        if (check_qsa && (gamma_psi < 0.0 || gamma_psi > max_qsa || psi_inv < 0.0)) {
            atomicAdd(qsa_count, 1);
            p.w[ip] = 0.0;
            p.valid[ip] = 0;
        } else {
            const long n = p.n;
            rec[ip] = q_invvol;
            rec[n + ip] = vx_c;
            rec[2 * n + ip] = vy_c;
            rec[3 * n + ip] = gamma_psi;
            rec[4 * n + ip] = psi_inv;
            rec[5 * n + ip] = xmid;
            rec[6 * n + ip] = ymid;
            const int bx = shape<O>(xmid, 0).cell - shape_delta<O>() - cl.bxlo;
            const int by = shape<O>(ymid, 0).cell - shape_delta<O>() - cl.bylo;
            if (bx >= 0 && bx < cl.wx && by >= 0 && by < cl.wy) k = by * cl.wx + bx;
        }
    }
    key[ip] = k;
    if (k >= 0) slot[ip] = atomicAdd(&count[k], 1);
}

// Plasma keys for the explicit deposition with a start cell of its own
// (ExplicitDeposition.cpp:109-141), and the same per-particle values as k_keys_plasma.
template <int T, int O>
__global__ __launch_bounds__(256) void k_keys_explicit(Geom g, hpq_plasma p, CLView cl, double charge_invvol,
                                                       int32_t* __restrict__ key, int32_t* __restrict__ slot,
                                                       int32_t* __restrict__ count, double* __restrict__ rec) {
    const long ip = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= p.n) return;
    int k = -1;
    if (p.valid[ip]) {
        const double clight_inv = 1.0 / 1.0;
        const double psi_inv = 1.0 / p.psi[ip];
        const double vx = p.ux[ip] * psi_inv * clight_inv;
        const double vy = p.uy[ip] * psi_inv * clight_inv;
        const double xmid = (p.x[ip] - g.xoff) * g.dxi;
        const double ymid = (p.y[ip] - g.yoff) * g.dyi;
        const double Aabssqp = 0.0;
        const double gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * psi_inv * psi_inv + vx * vx + vy * vy + 1.0);
        const long n = p.n;
        rec[ip] = charge_invvol * p.w[ip];
        rec[n + ip] = vx;
        rec[2 * n + ip] = vy;
        rec[3 * n + ip] = gamma_psi;
        rec[4 * n + ip] = psi_inv;
        rec[5 * n + ip] = xmid;
        rec[6 * n + ip] = ymid;
        const int bx = dshape<T, O>(xmid, 0).cell - dshape_delta<T, O>() - cl.bxlo;
        const int by = dshape<T, O>(ymid, 0).cell - dshape_delta<T, O>() - cl.bylo;
        if (bx >= 0 && bx < cl.wx && by >= 0 && by < cl.wy) k = by * cl.wx + bx;
    }
    key[ip] = k;
    if (k >= 0) slot[ip] = atomicAdd(&count[k], 1);
}

// Beam keys (BeamDepositCurrent.cpp:109-160): particles first .. first + count - 1 of the
// beam, numbered from 0; rec: [0] wqx, [1] wqy, [2] wqz, [3] xmid, [4] ymid.
template <int O>
__global__ __launch_bounds__(256) void k_keys_beam(Geom g, hpq_beam b, long first, long count_p, CLView cl,
                                                   const int32_t* __restrict__ select, int select_value,
                                                   int32_t* __restrict__ key, int32_t* __restrict__ slot,
                                                   int32_t* __restrict__ count, double* __restrict__ rec) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count_p) return;
    const long ip = first + t;
    int k = -1;
    if (b.valid[ip] && !(select && select[ip] != select_value)) {
        const double clightsq = 1.0 / (1.0 * 1.0);
        const double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
        const double gaminv = 1.0 / sqrt(1.0 + ux * ux * clightsq + uy * uy * clightsq + uz * uz * clightsq);
        const double wq = b.charge * b.w[ip] * 1.0;
        const double xmid = (b.x[ip] - g.xoff) * g.dxi;
        const double ymid = (b.y[ip] - g.yoff) * g.dyi;
        rec[t] = wq * (ux * gaminv);
        rec[count_p + t] = wq * (uy * gaminv);
        rec[2 * count_p + t] = wq * (uz * gaminv);
        rec[3 * count_p + t] = xmid;
        rec[4 * count_p + t] = ymid;
        const int bx = shape<O>(xmid, 0).cell - shape_delta<O>() - cl.bxlo;
        const int by = shape<O>(ymid, 0).cell - shape_delta<O>() - cl.bylo;
        if (bx >= 0 && bx < cl.wx && by >= 0 && by < cl.wy) k = by * cl.wx + bx;
    }
    key[t] = k;
    if (k >= 0) slot[t] = atomicAdd(&count[k], 1);
}

__global__ __launch_bounds__(256) void k_zero_int(int32_t* __restrict__ a, long n) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0;
}

// particle t into its list at its atomic slot
__global__ __launch_bounds__(256) void k_lists_scatter(long n, const int32_t* __restrict__ key,
                                                       const int32_t* __restrict__ slot,
                                                       const int32_t* __restrict__ off, int32_t* __restrict__ unsorted) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || key[t] < 0) return;
    unsorted[off[key[t]] + slot[t]] = int32_t(t);
}

// particle t to its rank by index inside its list
__global__ __launch_bounds__(256) void k_lists_order(long n, const int32_t* __restrict__ key,
                                                     const int32_t* __restrict__ off,
                                                     const int32_t* __restrict__ unsorted, int32_t* __restrict__ perm) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || key[t] < 0) return;
    const int a = off[key[t]], b = off[key[t] + 1];
    int r = 0;
    for (int u = a; u < b; ++u) r += unsorted[u] < int32_t(t) ? 1 : 0;
    perm[a + r] = int32_t(t);
}

// ---- gathers ------------------------------------------------------------------------------

__device__ __forceinline__ bool gather_cell(const Geom& g, int& i, int& j) {
    i = int(blockIdx.x) * GT_X + int(threadIdx.x % GT_X) - g.ng;
    j = int(blockIdx.y) * GT_Y + int(threadIdx.x / GT_X) - g.ng;
    return i <= g.nx - 1 + g.ng && j <= g.ny - 1 + g.ng;
}

// ::DepositCurrent's per-cell sums, PlasmaDepositCurrent.cpp:199-283: comps jx, jy, jz, rho,
// chi, rhomjz (-1: skip; jx and jy go together)
template <int O>
__global__ __launch_bounds__(GT_X * GT_Y) void k_gather_current(Geom g, CLView cl,
                                                                const double* __restrict__ rec, long n,
                                                                double* __restrict__ store, int cjx, int cjy, int cjz,
                                                                int crho, int cchi, int crm,
                                                                double charge_mu0_mass_ratio) {
    int i, j;
    if (!gather_cell(g, i, j)) return;
    constexpr int S = O + 1;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    double ajx = cjx >= 0 ? store[cjx * ps + k] : 0.0, ajy = cjy >= 0 ? store[cjy * ps + k] : 0.0;
    double ajz = cjz >= 0 ? store[cjz * ps + k] : 0.0, arho = crho >= 0 ? store[crho * ps + k] : 0.0;
    double achi = cchi >= 0 ? store[cchi * ps + k] : 0.0, arm = crm >= 0 ? store[crm * ps + k] : 0.0;
    int hits = 0;
    const double clight = 1.0;
    const double q_mu0_mass_ratio = charge_mu0_mass_ratio;
    walk_ordered<S, shape_delta<O>()>(g, cl, i, j, [&](int p, int ix, int iy) {
        const double q_invvol = rec[p], vx_c = rec[n + p], vy_c = rec[2 * n + p];
        const double gamma_psi = rec[3 * n + p], psi_inv = rec[4 * n + p];
        const double sx = shape<O>(rec[5 * n + p], ix).s;
        const double sy = shape<O>(rec[6 * n + p], iy).s;
        const double charge_density = q_invvol * sx * sy;
        if (cjx >= 0) {
            ajx += charge_density * vx_c;
            ajy += charge_density * vy_c;
        }
        if (cjz >= 0) ajz += charge_density * (gamma_psi - 1.0) * clight;
        if (crho >= 0) arho += charge_density * gamma_psi;
        if (cchi >= 0) achi += charge_density * q_mu0_mass_ratio * psi_inv;
        if (crm >= 0) arm += charge_density;
        ++hits;
    });
    if (hits == 0) return;
    if (cjx >= 0) {
        store[cjx * ps + k] = ajx;
        store[cjy * ps + k] = ajy;
    }
    if (cjz >= 0) store[cjz * ps + k] = ajz;
    if (crho >= 0) store[crho * ps + k] = arho;
    if (cchi >= 0) store[cchi * ps + k] = achi;
    if (crm >= 0) store[crm * ps + k] = arm;
}

// ::ExplicitDeposition's per-cell sums, ExplicitDeposition.cpp:186-262 (no laser), with the
// cell's own Bz, Ez, ExmBy, EypBx
template <int T, int O>
__global__ __launch_bounds__(GT_X * GT_Y) void k_gather_explicit(Geom g, CLView cl,
                                                                 const double* __restrict__ rec, long n,
                                                                 double* __restrict__ store, int cBz, int cEz,
                                                                 int cExmBy, int cEypBx, int cSy, int cSx,
                                                                 double charge_mass_ratio) {
    int i, j;
    if (!gather_cell(g, i, j)) return;
    constexpr int S = dshape_width<T, O>();
    const long k = g.idx(i, j);
    const long ps = g.plane;
    const double Bz_v = store[cBz * ps + k], Ez_v = store[cEz * ps + k];
    const double ExmBy_v = store[cExmBy * ps + k], EypBx_v = store[cEypBx * ps + k];
    double aSy = store[cSy * ps + k], aSx = store[cSx * ps + k];
    int hits = 0;
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double dxi = g.dxi, dyi = g.dyi;
    const double q_mass_ratio = charge_mass_ratio;
    walk_ordered<S, dshape_delta<T, O>()>(g, cl, i, j, [&](int p, int ix, int iy) {
        if (T == 2 && (ix == 0 || ix == O + 2) && (iy == 0 || iy == O + 2)) return;
        const double charge_density_mu0 = rec[p], vx = rec[n + p], vy = rec[2 * n + p];
        const double gamma_psi = rec[3 * n + p], psi_inv = rec[4 * n + p];
        const Dsf sy = dshape<T, O>(rec[6 * n + p], iy);
        const Dsf sx = dshape<T, O>(rec[5 * n + p], ix);
        const double AabssqDxp = 0.0, AabssqDyp = 0.0;
        aSy += charge_density_mu0 * (
            -sx.s * sy.s * (
                -Bz_v * vx
                + (Ez_v * vy
                + ExmBy_v * (-vx * vy)
                + EypBx_v * (gamma_psi - vy * vy)) * clight_inv
                - 0.25 * AabssqDyp * q_mass_ratio * psi_inv
            ) * q_mass_ratio * psi_inv
            + (-sx.ds * sy.s * dxi * (-vx * vy)
               - sx.s * sy.ds * dyi * (gamma_psi - vy * vy - 1.0)) * clight);
        aSx += charge_density_mu0 * (
            +sx.s * sy.s * (
                +Bz_v * vy
                + (Ez_v * vx
                + ExmBy_v * (gamma_psi - vx * vx)
                + EypBx_v * (-vx * vy)) * clight_inv
                - 0.25 * AabssqDxp * q_mass_ratio * psi_inv
            ) * q_mass_ratio * psi_inv
            + (+sx.ds * sy.s * dxi * (gamma_psi - vx * vx - 1.0)
               + sx.s * sy.ds * dyi * (-vx * vy)) * clight);
        ++hits;
    });
    if (hits == 0) return;
    store[cSy * ps + k] = aSy;
    store[cSx * ps + k] = aSx;
}

// ::DepositCurrentSlice's per-cell sums, BeamDepositCurrent.cpp:160-193
template <int O>
__global__ __launch_bounds__(GT_X * GT_Y) void k_gather_beam(Geom g, CLView cl,
                                                             const double* __restrict__ rec, long n,
                                                             double* __restrict__ store, int cjx, int cjy, int cjz) {
    int i, j;
    if (!gather_cell(g, i, j)) return;
    constexpr int S = O + 1;
    const long k = g.idx(i, j);
    const long ps = g.plane;
    double ajx = cjx >= 0 ? store[cjx * ps + k] : 0.0, ajy = cjy >= 0 ? store[cjy * ps + k] : 0.0;
    double ajz = cjz >= 0 ? store[cjz * ps + k] : 0.0;
    int hits = 0;
    walk_ordered<S, shape_delta<O>()>(g, cl, i, j, [&](int p, int ix, int iy) {
        const double sx = shape<O>(rec[3 * n + p], ix).s;
        const double sy = shape<O>(rec[4 * n + p], iy).s;
        if (cjx >= 0) {
            ajx += sx * sy * rec[p];
            ajy += sx * sy * rec[n + p];
        }
        if (cjz >= 0) ajz += sx * sy * rec[2 * n + p];
        ++hits;
    });
    if (hits == 0) return;
    if (cjx >= 0) {
        store[cjx * ps + k] = ajx;
        store[cjy * ps + k] = ajy;
    }
    if (cjz >= 0) store[cjz * ps + k] = ajz;
}

template <class T>
void grow(T*& p, long& cap, long need) {
    if (need <= cap) return;
    if (p) HPQ_HIP(hipFree(p));
    p = nullptr;
    HPQ_HIP(hipMalloc(&p, sizeof(T) * size_t(std::max(need, 1L))));
    cap = need;
}

// base-cell grid of a context: every start cell whose stencil (at most 6 wide, start at most 2
// below the base cell) reaches the guarded plane; the deposition tiles
void lists_geometry(const hpq_ctx* c, CellLists& L) {
    const Geom& g = c->g;
    L.bxlo = -g.ng - 6;
    L.bylo = -g.ng - 6;
    L.wx = g.nx + 2 * g.ng + 9;
    L.wy = g.ny + 2 * g.ng + 9;
    L.T = c->p.tile_size > 0 ? c->p.tile_size : 32;
    L.ntx = (g.ldx + L.T - 1) / L.T;
    L.nty = (g.ldy + L.T - 1) / L.T;
}

CLView view(const CellLists& L) {
    return CLView{L.bxlo, L.bylo, L.wx, L.wy, L.T, L.ntx, L.nty, L.off, L.perm};
}

void lists_reserve(CellLists& L, long n, int nrec) {
    const long ns = long(L.wx) * L.wy;
    grow(L.off, L.off_cap, ns + 1);
    grow(L.tmp, L.tmp_cap, scan_tmp_size(ns + 1));
    grow(L.key, L.key_cap, n);
    grow(L.slot, L.slot_cap, n);
    grow(L.unsorted, L.unsorted_cap, n);
    grow(L.perm, L.perm_cap, n);
    grow(L.rec, L.rec_cap, long(nrec) * n);
}

// count (already in L.off) -> offsets -> lists
void lists_finish(hipStream_t s, CellLists& L, long n) {
    const long ns = long(L.wx) * L.wy;
    scan_exclusive(s, L.off, ns + 1, L.tmp);
    hipLaunchKernelGGL(k_lists_scatter, grid1d(n, 256), dim3(256), 0, s, n, L.key, L.slot, L.off, L.unsorted);
    hipLaunchKernelGGL(k_lists_order, grid1d(n, 256), dim3(256), 0, s, n, L.key, L.off, L.unsorted, L.perm);
    HPQ_HIP(hipGetLastError());
}

void lists_clear(hipStream_t s, CellLists& L) {
    const long ns = long(L.wx) * L.wy;
    hipLaunchKernelGGL(k_zero_int, grid1d(ns + 1, 256), dim3(256), 0, s, L.off, ns + 1);
}

dim3 gather_grid(const Geom& g) { return dim3((g.ldx + GT_X - 1) / GT_X, (g.ldy + GT_Y - 1) / GT_Y); }

int shape_base_of(int order) {
    switch (order) {
        case 0: return shape_base<0>();
        case 1: return shape_base<1>();
        case 2: return shape_base<2>();
        default: return shape_base<3>();
    }
}

}  // namespace

void cell_lists_free(CellLists& L) {
    for (void* q : {(void*)L.off, (void*)L.tmp, (void*)L.key, (void*)L.slot, (void*)L.unsorted, (void*)L.perm,
                    (void*)L.rec})
        if (q) (void)hipFree(q);
    L = CellLists{};
}

// ::DepositCurrent in the CPU build's order; builds the plasma lists (L.state names the
// particle state they describe, for the explicit deposition that follows on the same state)
void deposit_current_ordered(hpq_ctx* c, const hpq_plasma* p, double* store, const int32_t* comps, bool ions) {
    const Geom& g = c->g;
    CellLists& L = c->cl_plasma;
    L.valid = false;
    if (p->n <= 0) return;
    if (p->n > INT32_MAX) throw HpqError(HPQ_E_INVALID, "ordered deposition: more than 2^31-1 particles");
    lists_geometry(c, L);
    lists_reserve(L, p->n, 7);
    const double charge = ions ? -c->p.plasma_charge : c->p.plasma_charge;
    const double mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double charge_invvol = charge * invvol;
    const double cmr = charge * mu0 / c->p.plasma_mass;
    TraceScope ts(c, HPQ_TRACE_DEPOSIT);
    lists_clear(c->stream, L);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        hipLaunchKernelGGL(k_keys_plasma<o>, grid1d(p->n, 256), dim3(256), 0, c->stream, g, *p, view(L), charge_invvol,
                           c->p.max_qsa, 1, L.key, L.slot, L.off, L.rec, c->d_count);
        HPQ_HIP(hipGetLastError());
        lists_finish(c->stream, L, p->n);
        hipLaunchKernelGGL(k_gather_current<o>, gather_grid(g), dim3(GT_X * GT_Y), 0, c->stream, g, view(L),
                           L.rec, p->n, store, comps[0], comps[1], comps[2], comps[3], comps[4], comps[5], cmr);
    });
    HPQ_HIP(hipGetLastError());
    L.valid = !ions;
    L.base = shape_base_of(c->p.depos_order);
    L.x = p->x;
    L.n = p->n;
}

// ::ExplicitDeposition in the CPU build's order; reuses the lists of the current deposition
// on the same particle state when the start cells share their floor
void explicit_deposition_ordered(hpq_ctx* c, const hpq_plasma* p, double* store) {
    const Geom& g = c->g;
    const Layout& Ly = c->L;
    CellLists& L = c->cl_plasma;
    if (p->n <= 0) return;
    if (p->n > INT32_MAX) throw HpqError(HPQ_E_INVALID, "ordered deposition: more than 2^31-1 particles");
    const double mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double qimu0 = c->p.plasma_charge * invvol * mu0;
    const double qmr = c->p.plasma_charge / c->p.plasma_mass;
    const int T = c->p.derivative_type;
    TraceScope ts(c, HPQ_TRACE_EXPLICIT_DEPOSIT);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        auto run = [&](auto TT) {
            constexpr int t = decltype(TT)::value;
            const bool reuse = L.valid && L.x == p->x && L.n == p->n && L.base == dshape_base<t, o>();
            if (!reuse) {
                lists_geometry(c, L);
                lists_reserve(L, p->n, 7);
                lists_clear(c->stream, L);
                hipLaunchKernelGGL((k_keys_explicit<t, o>), grid1d(p->n, 256), dim3(256), 0, c->stream, g, *p, view(L),
                                   qimu0, L.key, L.slot, L.off, L.rec);
                HPQ_HIP(hipGetLastError());
                lists_finish(c->stream, L, p->n);
                L.valid = true;
                L.base = dshape_base<t, o>();
                L.x = p->x;
                L.n = p->n;
            }
            hipLaunchKernelGGL((k_gather_explicit<t, o>), gather_grid(g), dim3(GT_X * GT_Y), 0, c->stream, g, view(L),
                               L.rec, p->n, store, Ly.Bz, Ly.Ez, Ly.ExmBy, Ly.EypBx, Ly.Sy, Ly.Sx, qmr);
        };
        if (T == 0) run(std::integral_constant<int, 0>{});
        else if (T == 1) run(std::integral_constant<int, 1>{});
        else run(std::integral_constant<int, 2>{});
    });
    HPQ_HIP(hipGetLastError());
}

// ::DepositCurrentSlice in the CPU build's order (particles of `slice` whose select entry
// equals select_value; select NULL: all)
void deposit_beam_ordered(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps,
                          const int32_t* select, int select_value) {
    const Geom& g = c->g;
    CellLists& L = c->cl_beam;
    const long first = b->off[slice];
    const long count = b->off[slice + 1] - first;
    if (count <= 0) return;
    if (count > INT32_MAX) throw HpqError(HPQ_E_INVALID, "ordered deposition: more than 2^31-1 particles");
    lists_geometry(c, L);
    lists_reserve(L, count, 5);
    TraceScope ts(c, HPQ_TRACE_BEAM_DEPOSIT);
    lists_clear(c->stream, L);
    dispatch_order(c->p.depos_order, [&](auto O) {
        constexpr int o = decltype(O)::value;
        hipLaunchKernelGGL(k_keys_beam<o>, grid1d(count, 256), dim3(256), 0, c->stream, g, *b, first, count, view(L),
                           select, select_value, L.key, L.slot, L.off, L.rec);
        HPQ_HIP(hipGetLastError());
        lists_finish(c->stream, L, count);
        hipLaunchKernelGGL(k_gather_beam<o>, gather_grid(g), dim3(GT_X * GT_Y), 0, c->stream, g, view(L),
                           L.rec, count, store, comps[0], comps[1], comps[2]);
    });
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq
