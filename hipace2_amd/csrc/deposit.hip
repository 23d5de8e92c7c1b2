// Current, chi and Sx/Sy deposition on gfx950 with LDS-private windows.
//
// Each workgroup takes a chunk of DCHUNK particles (in tile-sorted order when the
// deposition order from bins.hip is in use), reduces the bounding box of their stencils
// and, when the box fits in LDS, accumulates the whole chunk there with ds_add_f64 and
// adds the window to the slice store with one global atomic per touched cell and
// component. A chunk whose box is too large gets a square window around the mean start
// cell; its particles outside the window add to the store directly. The per-particle
// contributions are the reference's expressions in its operation order; only the order
// of the floating-point sums differs (the reference's GPU path uses global atomics), and in
// the explicit Sx/Sy deposition the per-cell terms are regrouped around per-particle
// coefficients (a few ulp per contribution).
#include <climits>
#include <cmath>

#include "internal.hpp"
#include "shape.hpp"

namespace hpq {

namespace {

constexpr int DT = 256;                 // threads per workgroup
// particles per thread: four for the current and beam depositions; two for the explicit one,
// whose per-particle work (21 cells, four window reads and two LDS adds each) is longest, so
// twice the workgroups hide its latency better (72 -> 59 us at 511^2, 2 x 2 ppc, while the
// current deposition runs 12 % slower with two)
constexpr int DPPT = 4;
constexpr int DPPT_EXPLICIT = 2;
constexpr int DCHUNK = DT * DPPT;       // particles per workgroup
constexpr int DCHUNK_EXPLICIT = DT * DPPT_EXPLICIT;
// LDS window sizes (doubles). A chunk of 1024 tile-sorted particles at 2 x 2 per cell
// covers one 32 x 8-cell tile: 35 x 11 cells for a 3-point stencil, 37 x 13 for the
// 5-point explicit one; the windows leave room for particles that drifted since the sort
// while keeping four workgroups per CU resident.
constexpr int WIN_DEPOSIT = 3072;       // 4 planes x 768 cells, 24 KiB
constexpr int WIN_EXPLICIT = 4608;      // 6 planes x 768 cells, 36 KiB
constexpr int WIN_BEAM = 7680;          // 3 planes x 2560 cells, 60 KiB: a beam core of +-25 cells

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void lds_add(double* p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ bool in_store(const Geom& g, int i, int j) {
    return i >= -g.ng && i <= g.nx - 1 + g.ng && j >= -g.ng && j <= g.ny - 1 + g.ng;
}

// the S x S stencil starting at cell (i, j) lies inside the guarded plane
__device__ __forceinline__ bool stencil_in_store(const Geom& g, int i, int j, int S) {
    return i >= -g.ng && i + S - 1 <= g.nx - 1 + g.ng && j >= -g.ng && j + S - 1 <= g.ny - 1 + g.ng;
}

// window of cells [i0, i0 + w) x [j0, j0 + h) in LDS; w == 0: none
struct Win {
    int i0, j0, w, h;
    __device__ bool holds(int ci, int cj, int S) const {
        return w > 0 && ci >= i0 && cj >= j0 && ci + S <= i0 + w && cj + S <= j0 + h;
    }
    __device__ int at(int i, int j) const { return (j - j0) * w + (i - i0); }
};

// Block-wide choice of the window for stencils of S x S cells starting at (ci, cj) for
// the particles with act set; cap is the number of cells that fit. red: 7 * DT/64 ints.
template <int PPT = DPPT>
__device__ Win choose_window(const int* ci, const int* cj, const bool* act, int S, int cap, int* red) {
    int mni = INT_MAX, mnj = INT_MAX, mxi = INT_MIN, mxj = INT_MIN, si = 0, sj = 0, cnt = 0;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        if (!act[q]) continue;
        mni = min(mni, ci[q]); mxi = max(mxi, ci[q]);
        mnj = min(mnj, cj[q]); mxj = max(mxj, cj[q]);
        si += ci[q]; sj += cj[q]; ++cnt;
    }
    for (int o = 32; o > 0; o >>= 1) {
        mni = min(mni, __shfl_xor(mni, o)); mxi = max(mxi, __shfl_xor(mxi, o));
        mnj = min(mnj, __shfl_xor(mnj, o)); mxj = max(mxj, __shfl_xor(mxj, o));
        si += __shfl_xor(si, o); sj += __shfl_xor(sj, o); cnt += __shfl_xor(cnt, o);
    }
    constexpr int NW = DT / 64;
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0 * NW + wv] = mni; red[1 * NW + wv] = mxi; red[2 * NW + wv] = mnj; red[3 * NW + wv] = mxj;
        red[4 * NW + wv] = si; red[5 * NW + wv] = sj; red[6 * NW + wv] = cnt;
    }
    __syncthreads();
    mni = red[0]; mxi = red[NW]; mnj = red[2 * NW]; mxj = red[3 * NW];
    si = red[4 * NW]; sj = red[5 * NW]; cnt = red[6 * NW];
    for (int w = 1; w < NW; ++w) {
        mni = min(mni, red[w]); mxi = max(mxi, red[NW + w]);
        mnj = min(mnj, red[2 * NW + w]); mxj = max(mxj, red[3 * NW + w]);
        si += red[4 * NW + w]; sj += red[5 * NW + w]; cnt += red[6 * NW + w];
    }
    if (cnt == 0) return {0, 0, 0, 0};
    const long bw = long(mxi) - mni + S, bh = long(mxj) - mnj + S;
    if (bw * bh <= cap) return {mni, mnj, int(bw), int(bh)};
    // square window around the mean start cell, no larger than the box
    int side = int(sqrt(double(cap)));
    if (side < S) return {0, 0, 0, 0};
    const int w = int(min<long>(side, bw)), h = int(min<long>(cap / w, bh));
    if (h < S) return {0, 0, 0, 0};
    const int mi = int(floor(double(si) / cnt)), mj = int(floor(double(sj) / cnt));
    const int i0 = min(max(mi + S / 2 - w / 2, mni), int(mxi + S - w));
    const int j0 = min(max(mj + S / 2 - h / 2, mnj), int(mxj + S - h));
    return {i0, j0, w, h};
}

// Accumulation in the window. SUM = HPQ_SUM_FP64: every contribution, FP64 or the FP32
// value of the mixed path, is added to an FP64 LDS word (an FP64 sum of FP32 terms is exact
// to FP64 rounding). SUM = HPQ_SUM_KAHAN32 (mixed path): each window word holds two FP32
// values, the sum and its compensation; a contribution is added to the sum with an atomic
// that returns the old value, and the rounding error of that addition (TwoSum: old + v equals
// the new sum plus the error exactly) is added to the compensation; the flush adds the two.
template <class R, int SUM = HPQ_SUM_FP64>
struct WinAcc {
    static __device__ __forceinline__ void add(double* win, int cells, int s, int k, R v) {
        if constexpr (SUM == HPQ_SUM_KAHAN32) {
            float* f = reinterpret_cast<float*>(&win[s * cells + k]);
            const float x = float(v);
            const float old = __hip_atomic_fetch_add(&f[0], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const float sum = old + x;
            const float bv = sum - old;
            const float err = (old - (sum - bv)) + (x - bv);
            if (err != 0.0f) __hip_atomic_fetch_add(&f[1], err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            lds_add(&win[s * cells + k], double(v));
        }
    }
    static __device__ __forceinline__ double get(const double* win, int cells, int s, int k) {
        if constexpr (SUM == HPQ_SUM_KAHAN32) {
            const float* f = reinterpret_cast<const float*>(&win[s * cells + k]);
            return double(f[0]) + double(f[1]);
        } else {
            return win[s * cells + k];
        }
    }
};

// adds the window to the store: quantity k (slot[k] >= 0) goes to component comp[k];
// cells holding zero are skipped
template <class R, int NK, int SUM = HPQ_SUM_FP64>
__device__ void flush_window(const Geom& g, const Win& W, const double* win, const int (&slot)[NK],
                             const int (&comp)[NK], double* store) {
    const int cells = W.w * W.h;
    for (int t = threadIdx.x; t < cells; t += DT) {
        const int j = W.j0 + t / W.w, i = W.i0 + t % W.w;
        if (!in_store(g, i, j)) continue;
        const long k = g.idx(i, j);
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            if (slot[q] < 0) continue;
            const double v = WinAcc<R, SUM>::get(win, cells, slot[q], t);
            if (v != 0.0) atomic_add(store + long(comp[q]) * g.plane + k, v);
        }
    }
}

// window slots of the quantities whose component is not -1, in order
template <int NK>
__device__ __forceinline__ int assign_slots(const int (&comp)[NK], int (&slot)[NK]) {
    int ns = 0;
#pragma unroll
    for (int q = 0; q < NK; ++q) {
        slot[q] = comp[q] >= 0 ? ns : -1;
        ns += comp[q] >= 0 ? 1 : 0;
    }
    return ns;
}

// ::DepositCurrent deposit functor, src/particles/deposition/PlasmaDepositCurrent.cpp:182-251,
// arithmetic in R. comp order: jx, jy, jz, rho, chi, rhomjz (-1: skip; jx and jy go together).
template <int O, class R, int SUM = HPQ_SUM_FP64>
__global__ __launch_bounds__(DT) void k_deposit_plasma(Geom g, hpq_plasma p, const int32_t* __restrict__ order,
                                                       double* __restrict__ store, int cjx, int cjy, int cjz, int crho,
                                                       int cchi, int crm, double charge_invvol,
                                                       double charge_mu0_mass_ratio, double max_qsa,
                                                       int* __restrict__ qsa_count) {
    constexpr int S = O + 1;
    __shared__ double win[WIN_DEPOSIT];
    __shared__ int red[7 * DT / 64];
    const R clight = 1.0, clightinv = R(1.0) / R(1.0);
    const int comp[6] = {cjx, cjy, cjz, crho, cchi, crm};
    int slot[6];
    const int ns = assign_slots(comp, slot);

    R q_invvol[DPPT], vx_c[DPPT], vy_c[DPPT], gamma_psi[DPPT], psi_inv[DPPT];
    double xmid[DPPT], ymid[DPPT];
    int ci[DPPT], cj[DPPT];
    bool act[DPPT];
    const long base = long(blockIdx.x) * DCHUNK;
#pragma unroll
    for (int q = 0; q < DPPT; ++q) {
        const long t = base + q * DT + threadIdx.x;
        act[q] = false;
        if (t >= p.n) continue;
        const long ip = order ? long(order[t]) : t;
        if (!p.valid[ip]) continue;
        psi_inv[q] = R(1.0) / R(p.psi[ip]);
        const double xp = p.x[ip], yp = p.y[ip];
        vx_c[q] = R(p.ux[ip]) * psi_inv[q];
        vy_c[q] = R(p.uy[ip]) * psi_inv[q];
        q_invvol[q] = R(charge_invvol) * R(p.w[ip]);
        xmid[q] = (xp - g.xoff) * g.dxi;
        ymid[q] = (yp - g.yoff) * g.dyi;
        const R Aabssqp = 0.0;
        gamma_psi[q] = R(0.5) * ((R(1.0) + R(0.5) * Aabssqp) * psi_inv[q] * psi_inv[q] +
                                 vx_c[q] * vx_c[q] * clightinv * clightinv +
                                 vy_c[q] * vy_c[q] * clightinv * clightinv + R(1.0));
        if (gamma_psi[q] < R(0.0) || gamma_psi[q] > R(max_qsa) || psi_inv[q] < R(0.0)) {
            atomicAdd(qsa_count, 1);
            p.w[ip] = 0.0;
            p.valid[ip] = 0;
            continue;
        }
        act[q] = true;
        ci[q] = shape<O, R>(xmid[q], 0).cell;
        cj[q] = shape<O, R>(ymid[q], 0).cell;
    }
    const Win W = choose_window(ci, cj, act, S, ns > 0 ? WIN_DEPOSIT / ns : 0, red);
    const int cells = W.w * W.h;
    for (int t = threadIdx.x; t < cells * ns; t += DT) win[t] = 0.0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < DPPT; ++q) {
        if (!act[q]) continue;
        SfR<R> sx[S], sy[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            sx[k] = shape<O, R>(xmid[q], k);
            sy[k] = shape<O, R>(ymid[q], k);
        }
        const bool local = W.holds(ci[q], cj[q], S);
        const R cmmr = R(charge_mu0_mass_ratio);
        if (local && stencil_in_store(g, sx[0].cell, sy[0].cell, S)) {
            // the stencil's cells are consecutive and all in the window: no per-cell checks
            const int k0 = W.at(sx[0].cell, sy[0].cell);
#pragma unroll
            for (int iy = 0; iy < S; ++iy) {
#pragma unroll
                for (int ix = 0; ix < S; ++ix) {
                    const R cd = q_invvol[q] * sx[ix].s * sy[iy].s;
                    const R v[6] = {cd * vx_c[q], cd * vy_c[q], cd * (gamma_psi[q] - R(1.0)) * clight,
                                    cd * gamma_psi[q], cd * cmmr * psi_inv[q], cd};
                    const int k = k0 + iy * W.w + ix;
#pragma unroll
                    for (int a = 0; a < 6; ++a)
                        if (slot[a] >= 0) WinAcc<R, SUM>::add(win, cells, slot[a], k, v[a]);
                }
            }
            continue;
        }
#pragma unroll
        for (int iy = 0; iy < S; ++iy) {
#pragma unroll
            for (int ix = 0; ix < S; ++ix) {
                const int i = sx[ix].cell, j = sy[iy].cell;
                if (!in_store(g, i, j)) continue;
                const R cd = q_invvol[q] * sx[ix].s * sy[iy].s;
                const R v[6] = {cd * vx_c[q], cd * vy_c[q], cd * (gamma_psi[q] - R(1.0)) * clight,
                                cd * gamma_psi[q], cd * cmmr * psi_inv[q], cd};
                if (local) {
                    const int k = W.at(i, j);
#pragma unroll
                    for (int a = 0; a < 6; ++a)
                        if (slot[a] >= 0) WinAcc<R, SUM>::add(win, cells, slot[a], k, v[a]);
                } else {
                    const long k = g.idx(i, j);
#pragma unroll
                    for (int a = 0; a < 6; ++a)
                        if (slot[a] >= 0) atomic_add(store + long(comp[a]) * g.plane + k, double(v[a]));
                }
            }
        }
    }
    __syncthreads();
    flush_window<R, 6, SUM>(g, W, win, slot, comp, store);
}

// ::ExplicitDeposition deposit functor, src/particles/deposition/ExplicitDeposition.cpp:166-262,
// arithmetic in R (the fields are read from the FP64 store and rounded to R)
template <int O, int T, class R, int SUM = HPQ_SUM_FP64>
__global__ __launch_bounds__(DT) void k_explicit_deposit(Geom g, hpq_plasma p, const int32_t* __restrict__ order,
                                                         double* __restrict__ store, int cBz, int cEz, int cExmBy,
                                                         int cEypBx, int cSy, int cSx, double charge_invvol_mu0,
                                                         double charge_mass_ratio) {
    constexpr int S = dshape_width<T, O>();
    __shared__ double win[WIN_EXPLICIT];
    __shared__ int red[7 * DT / 64];
    const R clight = 1.0, clight_inv = R(1.0) / R(1.0);
    // Only the start cells of a thread's particles stay in registers across the window choice;
    // each particle's momenta, weight and position are read again (cache hits) when its turn
    // comes, which keeps the FP64 kernel within 128 VGPRs (four waves per SIMD).
    long ipq[DPPT_EXPLICIT];
    int ci[DPPT_EXPLICIT], cj[DPPT_EXPLICIT];
    bool act[DPPT_EXPLICIT];
    const long base = long(blockIdx.x) * DCHUNK_EXPLICIT;
#pragma unroll
    for (int q = 0; q < DPPT_EXPLICIT; ++q) {
        const long t = base + q * DT + threadIdx.x;
        act[q] = false;
        ipq[q] = 0;
        if (t >= p.n) continue;
        const long ip = order ? long(order[t]) : t;
        if (!p.valid[ip]) continue;
        act[q] = true;
        ipq[q] = ip;
        ci[q] = dshape<T, O, R>((p.x[ip] - g.xoff) * g.dxi, 0).cell;
        cj[q] = dshape<T, O, R>((p.y[ip] - g.yoff) * g.dyi, 0).cell;
    }
    // window planes: Sy, Sx (accumulated), then copies of Bz, Ez, ExmBy, EypBx
    const Win W = choose_window<DPPT_EXPLICIT>(ci, cj, act, S, WIN_EXPLICIT / 6, red);
    const int cells = W.w * W.h;
    const double* Bz = store + long(cBz) * g.plane;
    const double* Ez = store + long(cEz) * g.plane;
    const double* ExmBy = store + long(cExmBy) * g.plane;
    const double* EypBx = store + long(cEypBx) * g.plane;
    double* Sy = store + long(cSy) * g.plane;
    double* Sx = store + long(cSx) * g.plane;
    for (int t = threadIdx.x; t < cells; t += DT) {
        const int j = W.j0 + t / W.w, i = W.i0 + t - (t / W.w) * W.w;
        win[t] = 0.0;
        win[cells + t] = 0.0;
        if (in_store(g, i, j)) {
            const long k = g.idx(i, j);
            win[2 * cells + t] = Bz[k];
            win[3 * cells + t] = Ez[k];
            win[4 * cells + t] = ExmBy[k];
            win[5 * cells + t] = EypBx[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < DPPT_EXPLICIT; ++q) {
        if (!act[q]) continue;
        // Per particle, the reference's Sy/Sx contribution of a cell
        //   Sy += cdm (-XsYs A q/psi + (-XdsYs dxi (-vx vy) - XsYds dyi (gamma/psi - vy^2 - 1)) c)
        //   Sx += cdm ( XsYs B q/psi + ( XdsYs dxi (gamma/psi - vx^2 - 1) + XsYds dyi (-vx vy)) c)
        // with A = -Bz vx + (Ez vy - ExmBy vx vy + EypBx (gamma/psi - vy^2)) / c and
        // B = Bz vy + (Ez vx + ExmBy (gamma/psi - vx^2) - EypBx vx vy) / c (no laser term), is
        // regrouped into per-particle coefficients times three shape-factor products.
        const long ip = ipq[q];
        const R pinv = R(1.0) / R(p.psi[ip]);
        const R pvx = R(p.ux[ip]) * pinv * clight_inv, pvy = R(p.uy[ip]) * pinv * clight_inv;
        const R cdm = R(charge_invvol_mu0) * R(p.w[ip]);
        const double xmid_q = (p.x[ip] - g.xoff) * g.dxi, ymid_q = (p.y[ip] - g.yoff) * g.dyi;
        const R q_mass_ratio = R(charge_mass_ratio);
        const R Aabssqp = 0.0;
        const R gamma_psi =
            R(0.5) * ((R(1.0) + R(0.5) * Aabssqp) * pinv * pinv + pvx * pvx + pvy * pvy + R(1.0));
        const R vxvy = pvx * pvy;
        const R gpx = gamma_psi - pvx * pvx, gpy = gamma_psi - pvy * pvy;
        const R qp = q_mass_ratio * pinv;
        const R dxi = R(g.dxi), dyi = R(g.dyi);
        const R ay = -cdm * qp, by = cdm * clight * dxi * vxvy, cy = -cdm * clight * dyi * (gpy - R(1.0));
        const R ax = cdm * qp, bx = cdm * clight * dxi * (gpx - R(1.0)), cx = -cdm * clight * dyi * vxvy;
        DsfR<R> sx[S], sy[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            sx[k] = dshape<T, O, R>(xmid_q, k);
            sy[k] = dshape<T, O, R>(ymid_q, k);
        }
        const bool local = W.holds(ci[q], cj[q], S);
        // the contribution of cell (ix, iy) from the four field values there
        auto contrib = [&](int ix, int iy, R Bz_v, R Ez_v, R ExmBy_v, R EypBx_v, R& sy_add, R& sx_add) {
            const DsfR<R> X = sx[ix], Y = sy[iy];
            const R A = fma(-Bz_v, pvx, fma(EypBx_v, gpy, fma(-ExmBy_v, vxvy, Ez_v * pvy)) * clight_inv);
            const R B = fma(Bz_v, pvy, fma(-EypBx_v, vxvy, fma(ExmBy_v, gpx, Ez_v * pvx)) * clight_inv);
            const R ss = X.s * Y.s, ds = X.ds * Y.s, sd = X.s * Y.ds;
            sy_add = fma(ay * A, ss, fma(by, ds, cy * sd));
            sx_add = fma(ax * B, ss, fma(bx, ds, cx * sd));
        };
        if (local && stencil_in_store(g, sx[0].cell, sy[0].cell, S)) {
            // the stencil's cells are consecutive and all in the window: no per-cell checks
            const int k0 = W.at(sx[0].cell, sy[0].cell);
            const double* wf = win + 2 * cells;
#pragma unroll
            for (int iy = 0; iy < S; ++iy) {
#pragma unroll
                for (int ix = 0; ix < S; ++ix) {
                    if (T == 2 && (ix == 0 || ix == O + 2) && (iy == 0 || iy == O + 2)) continue;
                    const int kw = k0 + iy * W.w + ix;
                    R sy_add, sx_add;
                    contrib(ix, iy, R(wf[kw]), R(wf[cells + kw]), R(wf[2 * cells + kw]), R(wf[3 * cells + kw]),
                            sy_add, sx_add);
                    WinAcc<R, SUM>::add(win, cells, 0, kw, sy_add);
                    WinAcc<R, SUM>::add(win, cells, 1, kw, sx_add);
                }
            }
            continue;
        }
#pragma unroll
        for (int iy = 0; iy < S; ++iy) {
#pragma unroll
            for (int ix = 0; ix < S; ++ix) {
                if (T == 2 && (ix == 0 || ix == O + 2) && (iy == 0 || iy == O + 2)) continue;
                const DsfR<R> X = sx[ix], Y = sy[iy];
                if (!in_store(g, X.cell, Y.cell)) continue;
                const long k = g.idx(X.cell, Y.cell);
                const int kw = local ? W.at(X.cell, Y.cell) : 0;
                R Bz_v, Ez_v, ExmBy_v, EypBx_v;
                if (local) {
                    Bz_v = R(win[2 * cells + kw]); Ez_v = R(win[3 * cells + kw]);
                    ExmBy_v = R(win[4 * cells + kw]); EypBx_v = R(win[5 * cells + kw]);
                } else {
                    Bz_v = R(Bz[k]); Ez_v = R(Ez[k]); ExmBy_v = R(ExmBy[k]); EypBx_v = R(EypBx[k]);
                }
                R sy_add, sx_add;
                contrib(ix, iy, Bz_v, Ez_v, ExmBy_v, EypBx_v, sy_add, sx_add);
                if (local) {
                    WinAcc<R, SUM>::add(win, cells, 0, kw, sy_add);
                    WinAcc<R, SUM>::add(win, cells, 1, kw, sx_add);
                } else {
                    atomic_add(Sy + k, double(sy_add));
                    atomic_add(Sx + k, double(sx_add));
                }
            }
        }
    }
    __syncthreads();
    const int comp[2] = {cSy, cSx};
    const int slot[2] = {0, 1};
    flush_window<R, 2, SUM>(g, W, win, slot, comp, store);
}

// ::DepositCurrentSlice deposit functor, src/particles/deposition/BeamDepositCurrent.cpp:133-193
template <int O>
__global__ __launch_bounds__(DT) void k_deposit_beam(Geom g, hpq_beam b, long first, long count,
                                                     double* __restrict__ store, int cjx, int cjy, int cjz,
                                                     const int32_t* __restrict__ select, int select_value) {
    constexpr int S = O + 1;
    __shared__ double win[WIN_BEAM];
    __shared__ int red[7 * DT / 64];
    const double clightsq = 1.0 / (1.0 * 1.0);
    const int comp[3] = {cjx, cjy, cjz};
    int slot[3];
    const int ns = assign_slots(comp, slot);
    double wqx[DPPT], wqy[DPPT], wqz[DPPT], xmid[DPPT], ymid[DPPT];
    int ci[DPPT], cj[DPPT];
    bool act[DPPT];
    const long base = long(blockIdx.x) * DCHUNK;
#pragma unroll
    for (int q = 0; q < DPPT; ++q) {
        const long t = base + q * DT + threadIdx.x;
        act[q] = false;
        if (t >= count) continue;
        const long ip = first + t;
        if (!b.valid[ip]) continue;
        if (select && select[ip] != select_value) continue;
        act[q] = true;
        const double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
        const double gaminv = 1.0 / sqrt(1.0 + ux * ux * clightsq + uy * uy * clightsq + uz * uz * clightsq);
        const double wq = b.charge * b.w[ip] * 1.0;
        wqx[q] = wq * (ux * gaminv);
        wqy[q] = wq * (uy * gaminv);
        wqz[q] = wq * (uz * gaminv);
        xmid[q] = (b.x[ip] - g.xoff) * g.dxi;
        ymid[q] = (b.y[ip] - g.yoff) * g.dyi;
        ci[q] = shape<O>(xmid[q], 0).cell;
        cj[q] = shape<O>(ymid[q], 0).cell;
    }
    const Win W = choose_window(ci, cj, act, S, ns > 0 ? WIN_BEAM / ns : 0, red);
    const int cells = W.w * W.h;
    for (int t = threadIdx.x; t < cells * ns; t += DT) win[t] = 0.0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < DPPT; ++q) {
        if (!act[q]) continue;
        Sf sx[S], sy[S];
#pragma unroll
        for (int k = 0; k < S; ++k) {
            sx[k] = shape<O>(xmid[q], k);
            sy[k] = shape<O>(ymid[q], k);
        }
        const bool local = W.holds(ci[q], cj[q], S);
#pragma unroll
        for (int iy = 0; iy < S; ++iy) {
#pragma unroll
            for (int ix = 0; ix < S; ++ix) {
                const int i = sx[ix].cell, j = sy[iy].cell;
                if (!in_store(g, i, j)) continue;
                const double s = sx[ix].s * sy[iy].s;
                const double v[3] = {s * wqx[q], s * wqy[q], s * wqz[q]};
                if (local) {
                    const int k = W.at(i, j);
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        if (slot[a] >= 0) lds_add(&win[slot[a] * cells + k], v[a]);
                } else {
                    const long k = g.idx(i, j);
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        if (slot[a] >= 0) atomic_add(store + long(comp[a]) * g.plane + k, v[a]);
                }
            }
        }
    }
    __syncthreads();
    flush_window<double>(g, W, win, slot, comp, store);
}

}  // namespace

// the FP32 depositions, in a translation unit of their own (deposit_f32.hip) built with FMA
// contraction; this one keeps every FP64 multiply and add separately rounded
void deposit_plasma_f32(hpq_ctx* c, const hpq_plasma* p, const int32_t* order, double* store, const int* cm,
                        double charge_invvol, double cmr);
void explicit_deposit_f32(hpq_ctx* c, const hpq_plasma* p, const int32_t* order, double* store, double qimu0,
                          double qmr);

#ifndef HPQ_F32_TU
// the FP64 path in the reference CPU build's summation order (gather.hip)
static bool ordered(const hpq_ctx* c) {
    return c->p.deposition == HPQ_DEPOSIT_ORDERED && c->p.precision == HPQ_PREC_FP64;
}

void deposit_current(hpq_ctx* c, const hpq_plasma* p, double* store, const int32_t* comps, bool ions) {
    if (p->n <= 0) return;
    if ((comps[0] >= 0) != (comps[1] >= 0)) throw HpqError(HPQ_E_INVALID, "deposit_current: jx and jy go together");
    if (ordered(c)) {
        deposit_current_ordered(c, p, store, comps, ions);
        return;
    }
    const Geom& g = c->g;
    const double charge = ions ? -c->p.plasma_charge : c->p.plasma_charge;
    const double mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double charge_invvol = charge * invvol;
    const double cmr = charge * mu0 / c->p.plasma_mass;
    auto cm = [&](int k) { return comps[k] >= 0 ? comps[k] : -1; };
    if ((cm(0) >= 0) != (cm(1) >= 0)) throw HpqError(HPQ_E_INVALID, "deposit_current: jx and jy go together");
    // the first deposition of every slice counts towards the re-sort period
    if (!ions && c->p.bin_period > 0 && ++c->bins.calls >= c->p.bin_period) {
        c->bins.calls = 0;
        c->bins.stale = true;
    }
    const int32_t* order = deposit_order(c, p);
    TraceScope ts(c, HPQ_TRACE_DEPOSIT);
    if (c->p.precision == HPQ_PREC_MIXED) {
        const int cms[6] = {cm(0), cm(1), cm(2), cm(3), cm(4), cm(5)};
        deposit_plasma_f32(c, p, order, store, cms, charge_invvol, cmr);
    } else {
        dispatch_order(c->p.depos_order, [&](auto O) {
            hipLaunchKernelGGL((k_deposit_plasma<decltype(O)::value, double>), grid1d(p->n, DCHUNK), dim3(DT), 0,
                               c->stream, g, *p, order, store, cm(0), cm(1), cm(2), cm(3), cm(4), cm(5), charge_invvol,
                               cmr, c->p.max_qsa, c->d_count);
        });
    }
    HPQ_HIP(hipGetLastError());
}

void explicit_deposition(hpq_ctx* c, const hpq_plasma* p, double* store) {
    if (p->n <= 0) return;
    const Geom& g = c->g;
    const Layout& L = c->L;
    const double mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double qimu0 = c->p.plasma_charge * invvol * mu0;
    const double qmr = c->p.plasma_charge / c->p.plasma_mass;
    const int T = c->p.derivative_type;
    if (T < 0 || T > 2) throw HpqError(HPQ_E_UNSUPPORTED, "depos_derivative_type must be 0..2");
    if (ordered(c)) {
        explicit_deposition_ordered(c, p, store);
        return;
    }
    const int32_t* order = deposit_order(c, p);
    TraceScope ts(c, HPQ_TRACE_EXPLICIT_DEPOSIT);
    if (c->p.precision == HPQ_PREC_MIXED) {
        explicit_deposit_f32(c, p, order, store, qimu0, qmr);
    } else {
        dispatch_order(c->p.depos_order, [&](auto O) {
            constexpr int o = decltype(O)::value;
            auto launch = [&](auto kern) {
                hipLaunchKernelGGL(kern, grid1d(p->n, DCHUNK_EXPLICIT), dim3(DT), 0, c->stream, g, *p, order, store,
                                   L.Bz, L.Ez, L.ExmBy, L.EypBx, L.Sy, L.Sx, qimu0, qmr);
            };
            if (T == 0) launch(k_explicit_deposit<o, 0, double>);
            else if (T == 1) launch(k_explicit_deposit<o, 1, double>);
            else launch(k_explicit_deposit<o, 2, double>);
        });
    }
    HPQ_HIP(hipGetLastError());
}

void deposit_beam(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps) {
    deposit_beam_select(c, b, slice, store, comps, nullptr, 0);
}

// only the particles whose select entry equals select_value deposit (select NULL: all)
void deposit_beam_select(hpq_ctx* c, const hpq_beam* b, int slice, double* store, const int32_t* comps,
                         const int32_t* select, int select_value) {
    if (!b || slice < 0 || slice >= c->g.nz || !b->off) return;
    const long first = b->off[slice];
    const long count = b->off[slice + 1] - first;
    if (count <= 0) return;
    if ((comps[0] >= 0) != (comps[1] >= 0))
        throw HpqError(HPQ_E_INVALID, "deposit_current_slice: jx and jy go together");
    if (c->p.deposition == HPQ_DEPOSIT_ORDERED) {
        deposit_beam_ordered(c, b, slice, store, comps, select, select_value);
        return;
    }
    TraceScope ts(c, HPQ_TRACE_BEAM_DEPOSIT);
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_deposit_beam<decltype(O)::value>), grid1d(count, DCHUNK), dim3(DT), 0, c->stream, c->g,
                           *b, first, count, store, comps[0], comps[1], comps[2], select, select_value);
    });
    HPQ_HIP(hipGetLastError());
}

#endif  // HPQ_F32_TU

}  // namespace hpq
