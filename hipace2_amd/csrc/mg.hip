// Geometric multigrid for the explicit Bx/By system on gfx950 (hpmg type I):
//   -acf * sol + Lap(sol) = rhs, homogeneous Dirichlet, two components sharing acf.
//
// Semantics are those of the reference's CPU build of hpmg::MultiGrid
// (src/mg_solver/HpMultiGrid.cpp): node-centred levels for odd n, cell-centred with the
// 4/3 boundary stencil for even n (ctor :1043-1163); V-cycle :1430-1512 with four
// red-black Gauss-Seidel sweeps (+ residual) per level, full-weighting restriction and
// bilinear (node) / piecewise-constant (cell) prolongation; bottom solve of
// max(16, box length) sweeps on the coarsest level (:1583-1593); convergence on the max
// norm of the residual (:1330-1416). Every point update uses the reference's operation
// order, so the iterates equal the CPU build's for equal inputs.
//
// Mapping to the GPU: the four sweeps (+ residual) of one level are one launch of an
// LDS-tiled kernel: each workgroup loads its tile plus a five-point halo, runs the four
// colour sweeps in LDS (points four or more away from the tile edge see exactly the
// values a global sweep gives them) and writes back the interior. Prolongation is fused
// into the load of the up-sweep. Levels of at most 33x33 points run as one workgroup
// that does the whole coarse part of the V-cycle with barriers between sweeps.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace hpq {

namespace {

constexpr int TX = 64, TY = 16, HALO = 5;
constexpr int LX = TX + 2 * HALO, LY = TY + 2 * HALO;
constexpr int COARSE_MAX_POINTS = 33 * 33;
constexpr int COARSE_THREADS = 1024;

__host__ __device__ __forceinline__ long at(const MGLevel& L, int i, int j) {
    return long(j - L.jlo + 1) * L.ldx + (i - L.ilo + 1);
}
__device__ __forceinline__ bool in_valid(const MGLevel& L, int i, int j) {
    return i >= L.vilo && i <= L.vihi && j >= L.vjlo && j <= L.vjhi;
}
__host__ __device__ __forceinline__ int cdiv2(int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

// gs1<is_cell_centered> (HpMultiGrid.cpp:267-292); P(a, b) reads phi at offset (a, b)
template <class PF>
__device__ __forceinline__ double gs1(const MGLevel& L, int i, int j, PF P, double rhs, double acf, double facx,
                                      double facy) {
    const bool cc = L.cc;
    double l;
    double c0 = -(acf + 2. * (facx + facy));
    if (cc && i == L.ilo) { l = facx * (4. / 3.) * P(1, 0); c0 -= 2. * facx; }
    else if (cc && i == L.ihi) { l = facx * (4. / 3.) * P(-1, 0); c0 -= 2. * facx; }
    else { l = facx * (P(-1, 0) + P(1, 0)); }
    if (cc && j == L.jlo) { l += facy * (4. / 3.) * P(0, 1); c0 -= 2. * facy; }
    else if (cc && j == L.jhi) { l += facy * (4. / 3.) * P(0, -1); c0 -= 2. * facy; }
    else { l += facy * (P(0, -1) + P(0, 1)); }
    const double c0_inv = 1. / c0;
    return (rhs - l) * c0_inv;
}

// laplacian + residual1 (HpMultiGrid.cpp:163-190)
template <class PF>
__device__ __forceinline__ double residual1(const MGLevel& L, int i, int j, PF P, double rhs, double acf,
                                            double facx, double facy) {
    double r = -2. * (facx + facy) * P(0, 0);
    if (i == L.ilo) r += facx * (4. / 3. * P(1, 0) - 2. * P(0, 0));
    else if (i == L.ihi) r += facx * (4. / 3. * P(-1, 0) - 2. * P(0, 0));
    else r += facx * (P(-1, 0) + P(1, 0));
    if (j == L.jlo) r += facy * (4. / 3. * P(0, 1) - 2. * P(0, 0));
    else if (j == L.jhi) r += facy * (4. / 3. * P(0, -1) - 2. * P(0, 0));
    else r += facy * (P(0, -1) + P(0, 1));
    return rhs + acf * P(0, 0) - r;
}

// fine_in + I(crse) at fine point (i, j) (interpcpy_cc / interpcpy_nd, :88-121)
__device__ __forceinline__ double interp_add(const MGLevel& F, const MGLevel& C, const double* cor_c, int i, int j,
                                             double fine_in) {
    const int ic = cdiv2(i), jc = cdiv2(j);
    auto c = [&](int a, int b) { return cor_c[at(C, a, b)]; };
    if (F.cc) return fine_in + c(ic, jc);
    const bool io = (ic * 2 != i), jo = (jc * 2 != j);
    if (io && jo) return fine_in + (c(ic, jc) + c(ic + 1, jc) + c(ic, jc + 1) + c(ic + 1, jc + 1)) * 0.25;
    if (io) return fine_in + (c(ic, jc) + c(ic + 1, jc)) * 0.5;
    if (jo) return fine_in + (c(ic, jc) + c(ic, jc + 1)) * 0.5;
    return fine_in + c(ic, jc);
}

// restrict_cc / restrict_nd (:31-52) at coarse point (i, j)
__device__ __forceinline__ double restrict_at(const MGLevel& C, const MGLevel& F, const double* fine, int i, int j) {
    auto f = [&](int a, int b) { return fine[at(F, a, b)]; };
    if (C.cc) return 0.25 * (f(2 * i, 2 * j) + f(2 * i + 1, 2 * j) + f(2 * i, 2 * j + 1) + f(2 * i + 1, 2 * j + 1));
    return (1. / 16.) * (f(2 * i - 1, 2 * j - 1) + 2. * f(2 * i, 2 * j - 1) + f(2 * i + 1, 2 * j - 1) +
                         2. * f(2 * i - 1, 2 * j) + 4. * f(2 * i, 2 * j) + 2. * f(2 * i + 1, 2 * j) +
                         f(2 * i - 1, 2 * j + 1) + 2. * f(2 * i, 2 * j + 1) + f(2 * i + 1, 2 * j + 1));
}

// Four red-black sweeps (+ residual) of one level, LDS-tiled.
// Start: zero (ZERO), phi_in + I(cor_c) (INTERP) or phi_in; result -> phi_out; residual -> res.
template <bool ZERO, bool RES, bool INTERP>
__global__ __launch_bounds__(256) void k_gsrb4(MGLevel L, MGLevel C, double* __restrict__ phi_out,
                                               const double* __restrict__ rhs, long rcs,
                                               const double* __restrict__ acf, const double* __restrict__ phi_in,
                                               const double* __restrict__ cor_c, double* __restrict__ res) {
    __shared__ double s[2][LY][LX];
    const int ti0 = L.vilo + int(blockIdx.x) * TX - HALO;
    const int tj0 = L.vjlo + int(blockIdx.y) * TY - HALO;
    const double facx = 1. / (L.dx * L.dx), facy = 1. / (L.dy * L.dy);
    for (int e = threadIdx.x; e < LX * LY; e += blockDim.x) {
        const int ly = e / LX, lx = e - ly * LX;
        const int i = ti0 + lx, j = tj0 + ly;
        double v0 = 0.0, v1 = 0.0;
        if (!ZERO && in_valid(L, i, j)) {
            const long k = at(L, i, j);
            if (INTERP) {
                v0 = interp_add(L, C, cor_c, i, j, phi_in[k]);
                v1 = interp_add(L, C, cor_c + C.cs, i, j, phi_in[k + L.cs]);
            } else {
                v0 = phi_in[k];
                v1 = phi_in[k + L.cs];
            }
        }
        s[0][ly][lx] = v0;
        s[1][ly][lx] = v1;
    }
    __syncthreads();
    constexpr int IW = LX - 2, IH = LY - 2;
    for (int color = 0; color < 4; ++color) {
        for (int e = threadIdx.x; e < IW * IH; e += blockDim.x) {
            const int ly = 1 + e / IW, lx = 1 + e - (ly - 1) * IW;
            const int i = ti0 + lx, j = tj0 + ly;
            if (((i + j + color) & 1) == 0 && in_valid(L, i, j)) {
                const long k = at(L, i, j);
                const double a = acf[k];
                const double r0 = rhs[k], r1 = rhs[k + rcs];
                auto P0 = [&](int di, int dj) { return s[0][ly + dj][lx + di]; };
                auto P1 = [&](int di, int dj) { return s[1][ly + dj][lx + di]; };
                const double n0 = gs1(L, i, j, P0, r0, a, facx, facy);
                const double n1 = gs1(L, i, j, P1, r1, a, facx, facy);
                s[0][ly][lx] = n0;
                s[1][ly][lx] = n1;
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < TX * TY; e += blockDim.x) {
        const int oy = e / TX, ox = e - oy * TX;
        const int lx = HALO + ox, ly = HALO + oy;
        const int i = ti0 + lx, j = tj0 + ly;
        if (!in_valid(L, i, j)) continue;
        const long k = at(L, i, j);
        if (RES) {
            const double a = acf[k];
            auto P0 = [&](int di, int dj) { return s[0][ly + dj][lx + di]; };
            auto P1 = [&](int di, int dj) { return s[1][ly + dj][lx + di]; };
            res[k] = residual1(L, i, j, P0, rhs[k], a, facx, facy);
            res[k + L.cs] = residual1(L, i, j, P1, rhs[k + rcs], a, facx, facy);
        }
        phi_out[k] = s[0][ly][lx];
        phi_out[k + L.cs] = s[1][ly][lx];
    }
}

// crse (valid points) = R(fine), ncomp components
__global__ void k_restrict(MGLevel C, MGLevel F, double* __restrict__ crse, const double* __restrict__ fine,
                           int ncomp, long ccs, long fcs) {
    const int nvx = C.vihi - C.vilo + 1, nvy = C.vjhi - C.vjlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvx * nvy) return;
    const int j = C.vjlo + t / nvx, i = C.vilo + t % nvx;
    for (int n = 0; n < ncomp; ++n) crse[n * ccs + at(C, i, j)] = restrict_at(C, F, fine + n * fcs, i, j);
}

// ---------------------------------------------------------------------------------
// coarse levels in one workgroup (global memory, barriers between sweeps)
// ---------------------------------------------------------------------------------
__device__ void wg_gsrb(const MGLevel& L, int color, double* phi, const double* rhs) {
    const double facx = 1. / (L.dx * L.dx), facy = 1. / (L.dy * L.dy);
    const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
    for (int t = threadIdx.x; t < nvx * nvy; t += blockDim.x) {
        const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
        if (((i + j + color) & 1) != 0) continue;
        const long k = at(L, i, j);
        const double a = L.acf[k];
        for (int n = 0; n < 2; ++n) {
            double* ph = phi + n * L.cs;
            auto P = [&](int di, int dj) { return ph[at(L, i + di, j + dj)]; };
            ph[k] = gs1(L, i, j, P, rhs[k + n * L.cs], a, facx, facy);
        }
    }
    __syncthreads();
}

// phi (= cor) starts from start (zero when null), four sweeps, optional residual into res
__device__ void wg_gsrb4(const MGLevel& L, const double* start, bool do_res) {
    const long tot = 2 * L.cs;
    for (long t = threadIdx.x; t < tot; t += blockDim.x) L.cor[t] = 0.0;
    __syncthreads();
    if (start) {
        const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
        for (int t = threadIdx.x; t < nvx * nvy; t += blockDim.x) {
            const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
            const long k = at(L, i, j);
            L.cor[k] = start[k];
            L.cor[k + L.cs] = start[k + L.cs];
        }
        __syncthreads();
    }
    for (int color = 0; color < 4; ++color) wg_gsrb(L, color, L.cor, L.res);
    if (do_res) {
        const double facx = 1. / (L.dx * L.dx), facy = 1. / (L.dy * L.dy);
        const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
        for (int t = threadIdx.x; t < nvx * nvy; t += blockDim.x) {
            const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
            const long k = at(L, i, j);
            const double a = L.acf[k];
            for (int n = 0; n < 2; ++n) {
                const double* ph = L.cor + n * L.cs;
                auto P = [&](int di, int dj) { return ph[at(L, i + di, j + dj)]; };
                L.rescor[k + n * L.cs] = residual1(L, i, j, P, L.res[k + n * L.cs], a, facx, facy);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(COARSE_THREADS) void k_coarse_cycle(const MGLevel* __restrict__ lev, int cb, int maxl) {
    for (int l = cb; l < maxl; ++l) {
        const MGLevel L = lev[l];
        const MGLevel Cn = lev[l + 1];
        wg_gsrb4(L, nullptr, true);
        const int nvx = Cn.vihi - Cn.vilo + 1, nvy = Cn.vjhi - Cn.vjlo + 1;
        for (int t = threadIdx.x; t < nvx * nvy; t += blockDim.x) {
            const int j = Cn.vjlo + t / nvx, i = Cn.vilo + t % nvx;
            for (int n = 0; n < 2; ++n) Cn.res[n * Cn.cs + at(Cn, i, j)] = restrict_at(Cn, L, L.rescor + n * L.cs, i, j);
        }
        __syncthreads();
    }
    {  // bottom solve, CPU branch of MultiGrid::bottomsolve (:1583-1593)
        const MGLevel B = lev[maxl];
        for (long t = threadIdx.x; t < 2 * B.cs; t += blockDim.x) B.cor[t] = 0.0;
        __syncthreads();
        const int len = max(B.ihi - B.ilo + 1, B.jhi - B.jlo + 1);
        const int nsw = max(16, (len + 1) / 2 * 2);
        for (int sw = 0; sw < nsw; ++sw) wg_gsrb(B, sw, B.cor, B.res);
    }
    for (int l = maxl - 1; l >= cb; --l) {
        const MGLevel L = lev[l];
        const MGLevel Cn = lev[l + 1];
        const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
        for (int t = threadIdx.x; t < nvx * nvy; t += blockDim.x) {
            const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
            const long k = at(L, i, j);
            for (int n = 0; n < 2; ++n)
                L.rescor[k + n * L.cs] = interp_add(L, Cn, Cn.cor + n * Cn.cs, i, j, L.cor[k + n * L.cs]);
        }
        __syncthreads();
        wg_gsrb4(L, L.rescor, false);
    }
}

// level-0 load: acf over the whole box, sol and rhs (2 comps) over the whole box
__global__ void k_mg_load(MGLevel L, int sh, Geom g, const double* __restrict__ sol, const double* __restrict__ rhs,
                          const double* __restrict__ acf, long cstride, double* __restrict__ msol,
                          double* __restrict__ mrhs) {
    const int nbx = L.ihi - L.ilo + 1, nby = L.jhi - L.jlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbx * nby) return;
    const int j = L.jlo + t / nbx, i = L.ilo + t % nbx;
    const long gk = g.idx(i - sh, j - sh);
    const long k = at(L, i, j);
    L.acf[k] = acf[gk];
    msol[k] = sol[gk];
    msol[k + L.cs] = sol[gk + cstride];
    mrhs[k] = rhs[gk];
    mrhs[k + L.cs] = rhs[gk + cstride];
}

__global__ void k_mg_store(MGLevel L, int sh, Geom g, double* __restrict__ sol, long cstride) {
    const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvx * nvy) return;
    const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
    const long gk = g.idx(i - sh, j - sh);
    const long k = at(L, i, j);
    sol[gk] = L.cor[k];
    sol[gk + cstride] = L.cor[k + L.cs];
}

// max |a| over points (valid box when only_valid, else whole box), 2 comps, into out (as bits)
__global__ void k_absmax(MGLevel L, const double* __restrict__ a, bool only_valid, unsigned long long* out) {
    __shared__ double red[256];
    const int bx0 = only_valid ? L.vilo : L.ilo, bx1 = only_valid ? L.vihi : L.ihi;
    const int by0 = only_valid ? L.vjlo : L.jlo, by1 = only_valid ? L.vjhi : L.jhi;
    const int nx = bx1 - bx0 + 1, ny = by1 - by0 + 1;
    double m = 0.0;
    for (long t = blockIdx.x * long(blockDim.x) + threadIdx.x; t < 2L * nx * ny; t += long(gridDim.x) * blockDim.x) {
        const int n = int(t / (long(nx) * ny));
        const long r = t - n * long(nx) * ny;
        const int j = by0 + int(r / nx), i = bx0 + int(r % nx);
        m = fmax(m, fabs(a[n * L.cs + at(L, i, j)]));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

bool coarsenable(bool cc, int lo, int hi, int min_width) {
    const int size = hi - lo + 1;
    if (size < 2 * min_width) return false;
    if (cc) return (cdiv2(lo) * 2 == lo) && (cdiv2(hi) * 2 + 1 == hi);
    return (cdiv2(lo) * 2 == lo) && (cdiv2(hi) * 2 == hi);
}

}  // namespace

void mg_init(hpq_ctx* c) {
    MultiGrid& M = c->mg;
    const int nx = c->g.nx, ny = c->g.ny;
    if (nx % 2 != ny % 2) throw HpqError(HPQ_E_UNSUPPORTED, "hpmg: nx and ny must have the same parity");
    M.cc = (nx % 2 == 0);
    M.dx = c->g.dx;
    M.dy = c->g.dy;
    MGLevel l0;
    l0.cc = M.cc;
    l0.ilo = 0; l0.jlo = 0;
    l0.ihi = M.cc ? nx - 1 : nx + 1;
    l0.jhi = M.cc ? ny - 1 : ny + 1;
    M.lev.clear();
    M.lev.push_back(l0);
    const int mw = M.cc ? 2 : 4;
    for (int i = 0; i < 30; ++i) {
        const MGLevel& b = M.lev.back();
        if (coarsenable(M.cc, b.ilo, b.ihi, mw) && coarsenable(M.cc, b.jlo, b.jhi, mw)) {
            MGLevel n;
            n.cc = M.cc;
            n.ilo = b.ilo / 2; n.jlo = b.jlo / 2; n.ihi = b.ihi / 2; n.jhi = b.jhi / 2;
            M.lev.push_back(n);
        } else {
            break;
        }
    }
    M.max_level = int(M.lev.size()) - 1;
    if (M.max_level < 1)
        throw HpqError(HPQ_E_UNSUPPORTED, "hpmg: the slice cannot be coarsened at least once");
    for (int l = 0; l <= M.max_level; ++l) {
        MGLevel& L = M.lev[l];
        if (M.cc) { L.vilo = L.ilo; L.vihi = L.ihi; L.vjlo = L.jlo; L.vjhi = L.jhi; }
        else { L.vilo = L.ilo + 1; L.vihi = L.ihi - 1; L.vjlo = L.jlo + 1; L.vjhi = L.jhi - 1; }
        L.ldx = (L.ihi - L.ilo + 1) + 2;
        L.ldy = (L.jhi - L.jlo + 1) + 2;
        L.cs = long(L.ldx) * L.ldy;
        L.dx = M.dx * double(1 << l);
        L.dy = M.dy * double(1 << l);
        const size_t b1 = sizeof(double) * L.cs;
        HPQ_HIP(hipMalloc(&L.acf, b1));
        HPQ_HIP(hipMalloc(&L.res, 2 * b1));
        HPQ_HIP(hipMalloc(&L.cor, 2 * b1));
        HPQ_HIP(hipMalloc(&L.rescor, 2 * b1));
        HPQ_HIP(hipMemset(L.acf, 0, b1));
        HPQ_HIP(hipMemset(L.res, 0, 2 * b1));
        HPQ_HIP(hipMemset(L.cor, 0, 2 * b1));
        HPQ_HIP(hipMemset(L.rescor, 0, 2 * b1));
    }
    M.coarse_begin = M.max_level;
    for (int l = 1; l <= M.max_level; ++l) {
        const MGLevel& L = M.lev[l];
        if ((L.ihi - L.ilo + 1) * (L.jhi - L.jlo + 1) <= COARSE_MAX_POINTS) { M.coarse_begin = l; break; }
    }
    HPQ_HIP(hipMalloc(&M.d_lev, sizeof(MGLevel) * M.lev.size()));
    HPQ_HIP(hipMemcpy(M.d_lev, M.lev.data(), sizeof(MGLevel) * M.lev.size(), hipMemcpyHostToDevice));
    const size_t b0 = sizeof(double) * 2 * M.lev[0].cs;
    HPQ_HIP(hipMalloc(&M.msol, b0));
    HPQ_HIP(hipMalloc(&M.mrhs, b0));
    HPQ_HIP(hipMemset(M.msol, 0, b0));
    HPQ_HIP(hipMemset(M.mrhs, 0, b0));
    HPQ_HIP(hipMalloc(&M.d_norm, 4 * sizeof(unsigned long long)));
    HPQ_HIP(hipHostMalloc(&M.h_norm, 4 * sizeof(unsigned long long), hipHostMallocDefault));
}

void mg_free(hpq_ctx* c) {
    MultiGrid& M = c->mg;
    for (auto& L : M.lev)
        for (void* p : {(void*)L.acf, (void*)L.res, (void*)L.cor, (void*)L.rescor})
            if (p) (void)hipFree(p);
    if (M.d_lev) (void)hipFree(M.d_lev);
    if (M.msol) (void)hipFree(M.msol);
    if (M.mrhs) (void)hipFree(M.mrhs);
    if (M.d_norm) (void)hipFree(M.d_norm);
    if (M.h_norm) (void)hipHostFree(M.h_norm);
    M = MultiGrid{};
}

namespace {

template <bool ZERO, bool RES, bool INTERP>
void launch_gsrb4(hpq_ctx* c, const MGLevel& L, const MGLevel& C, double* phi_out, const double* rhs, long rcs,
                  const double* acf, const double* phi_in, const double* cor_c, double* res) {
    TraceScope ts(c, HPQ_TRACE_MG_SMOOTH);
    const dim3 grid((L.vihi - L.vilo + 1 + TX - 1) / TX, (L.vjhi - L.vjlo + 1 + TY - 1) / TY);
    hipLaunchKernelGGL((k_gsrb4<ZERO, RES, INTERP>), grid, dim3(256), 0, c->stream, L, C, phi_out, rhs, rcs, acf,
                       phi_in, cor_c, res);
}

dim3 tile_grid(const MGLevel& L) {
    return dim3((L.vihi - L.vilo + 1 + TX - 1) / TX, (L.vjhi - L.vjlo + 1 + TY - 1) / TY);
}

void launch_restrict(hpq_ctx* c, const MGLevel& C, const MGLevel& F, double* crse, const double* fine, int ncomp) {
    const int n = (C.vihi - C.vilo + 1) * (C.vjhi - C.vjlo + 1);
    hipLaunchKernelGGL(k_restrict, grid1d(n, 256), dim3(256), 0, c->stream, C, F, crse, fine, ncomp, C.cs, F.cs);
}

// MultiGrid::vcycle (:1430-1512), CPU level split: levels [0, cb) tiled, [cb, max] in one workgroup
void vcycle(hpq_ctx* c) {
    MultiGrid& M = c->mg;
    const int cb = M.coarse_begin;
    for (int l = 0; l < cb; ++l) {
        MGLevel& L = M.lev[l];
        if (l > 0) launch_gsrb4<true, true, false>(c, L, L, L.cor, L.res, L.cs, L.acf, nullptr, nullptr, L.rescor);
        launch_restrict(c, M.lev[l + 1], L, M.lev[l + 1].res, L.rescor, 2);
    }
    {
        TraceScope ts(c, HPQ_TRACE_MG_COARSE);
        hipLaunchKernelGGL(k_coarse_cycle, dim3(1), dim3(COARSE_THREADS), 0, c->stream, M.d_lev, cb, M.max_level);
    }
    // Up-sweep: the result of level l goes to rescor[l] (dead after the restriction), so
    // the kernel never overwrites cor[l], which neighbouring tiles read as their halo.
    // The coarse kernel leaves its top level in cor[cb].
    for (int l = cb - 1; l >= 0; --l) {
        MGLevel& L = M.lev[l];
        MGLevel& C = M.lev[l + 1];
        const double* corc = (l + 1 == cb) ? C.cor : C.rescor;
        if (l == 0)
            launch_gsrb4<false, false, true>(c, L, C, M.msol, M.mrhs, L.cs, L.acf, L.cor, corc, nullptr);
        else
            launch_gsrb4<false, false, true>(c, L, C, L.rescor, L.res, L.cs, L.acf, L.cor, corc, nullptr);
    }
    MGLevel& L0 = M.lev[0];
    launch_gsrb4<false, true, false>(c, L0, L0, L0.cor, M.mrhs, L0.cs, L0.acf, M.msol, nullptr, L0.rescor);
    HPQ_HIP(hipGetLastError());
}

}  // namespace

// MultiGrid::solve1 + solve_doit (HpMultiGrid.cpp:1170-1190, 1308-1427)
int mg_solve1(hpq_ctx* c, double* sol, const double* rhs, const double* acf, long cstride, double tol_rel,
              double tol_abs, int maxiter, int* iters) {
    MultiGrid& M = c->mg;
    MGLevel& L0 = M.lev[0];
    const int sh = M.cc ? 0 : 1;
    {
        const int n = (L0.ihi - L0.ilo + 1) * (L0.jhi - L0.jlo + 1);
        hipLaunchKernelGGL(k_mg_load, grid1d(n, 256), dim3(256), 0, c->stream, L0, sh, c->g, sol, rhs, acf, cstride,
                           M.msol, M.mrhs);
    }
    for (int l = 1; l <= M.max_level; ++l) launch_restrict(c, M.lev[l], M.lev[l - 1], M.lev[l].acf, M.lev[l - 1].acf, 1);
    launch_gsrb4<false, true, false>(c, L0, L0, L0.cor, M.mrhs, L0.cs, L0.acf, M.msol, nullptr, L0.rescor);
    HPQ_HIP(hipMemsetAsync(M.d_norm, 0, 4 * sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_absmax, dim3(64), dim3(256), 0, c->stream, L0, L0.rescor, true, M.d_norm);
    hipLaunchKernelGGL(k_absmax, dim3(64), dim3(256), 0, c->stream, L0, M.mrhs, true, M.d_norm + 1);
    HPQ_HIP(hipMemcpyAsync(M.h_norm, M.d_norm, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    HPQ_HIP(hipStreamSynchronize(c->stream));
    double resnorm0, rhsnorm0;
    std::memcpy(&resnorm0, &M.h_norm[0], 8);
    std::memcpy(&rhsnorm0, &M.h_norm[1], 8);
    const double max_norm = (rhsnorm0 >= resnorm0) ? rhsnorm0 : resnorm0;
    const double res_target = std::max(tol_abs, std::max(tol_rel, 1.e-16) * max_norm);
    int status = 0, it = 0;
    if (!(resnorm0 <= res_target)) {
        bool converged = false;
        for (it = 0; it < maxiter; ++it) {
            vcycle(c);
            HPQ_HIP(hipMemsetAsync(M.d_norm + 2, 0, sizeof(unsigned long long), c->stream));
            hipLaunchKernelGGL(k_absmax, dim3(64), dim3(256), 0, c->stream, L0, L0.rescor, false, M.d_norm + 2);
            HPQ_HIP(hipMemcpyAsync(M.h_norm + 2, M.d_norm + 2, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   c->stream));
            HPQ_HIP(hipStreamSynchronize(c->stream));
            double norminf;
            std::memcpy(&norminf, &M.h_norm[2], 8);
            if (norminf <= res_target) { converged = true; ++it; break; }
            if (norminf > 1.e20 * max_norm) { status = HPQ_E_MG_DIVERGED; ++it; break; }
        }
        if (!converged && status == 0) status = HPQ_E_MG_NOT_CONVERGED;
    }
    if (iters) *iters = it;
    {
        const int n = (L0.vihi - L0.vilo + 1) * (L0.vjhi - L0.vjlo + 1);
        hipLaunchKernelGGL(k_mg_store, grid1d(n, 256), dim3(256), 0, c->stream, L0, sh, c->g, sol, cstride);
    }
    HPQ_HIP(hipGetLastError());
    return status;
}

}  // namespace hpq
