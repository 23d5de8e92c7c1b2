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
// order, so the iterates equal the CPU build's for equal inputs. The diagonal inverse
// 1/c0 of a point depends only on acf and the point's position, so it is computed once
// per launch (fine levels) or per V-cycle (coarse levels) instead of once per sweep.
//
// Mapping to the GPU: the sweeps (+ residual) of one level are one launch of an LDS-tiled
// kernel: each workgroup loads its tile plus a halo of NSW + 2 points, runs the NSW
// colour sweeps in LDS (points NSW or more away from the frozen outer ring see exactly
// the values a global sweep gives them) and writes back the interior; the residual, its
// restriction to the next level and its max norm come out of the same launch.
// Prolongation is fused into the load of the up-sweep. On level 0 the up-sweep and the
// final smoothing of the cycle are one launch of eight sweeps. Levels from the first one
// of at most 33x33 points down run as one workgroup that does the coarse part of the
// V-cycle with barriers between sweeps, its arrays in LDS.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace hpq {

namespace {

constexpr int TX = 32, TY = 16;                // output tile of the fine-level kernel
constexpr int NT = 256;                        // threads per workgroup
constexpr int RW = TX + 2, RH = TY + 2;        // residual region: tile + 1 ring
constexpr int COARSE_MAX_POINTS = 33 * 33;
constexpr int CTX = 16, CTY = 32;              // coarse workgroup: CTX x CTY threads
constexpr int COARSE_THREADS = CTX * CTY;
constexpr int COARSE_ARRAYS = 8;               // acf, 1/c0, res (2), cor (2), rescor (2)
constexpr size_t COARSE_LDS_MAX = 150 * 1024;
constexpr int COARSE_MAX_LEVELS = 16;

__host__ __device__ __forceinline__ long at(const MGLevel& L, int i, int j) {
    return long(j - L.jlo + 1) * L.ldx + (i - L.ilo + 1);
}
template <class LV>
__device__ __forceinline__ bool in_valid(const LV& L, int i, int j) {
    return i >= L.vilo && i <= L.vihi && j >= L.vjlo && j <= L.vjhi;
}
__host__ __device__ __forceinline__ int cdiv2(int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

// 1/c0 of gs1<is_cell_centered> (HpMultiGrid.cpp:267-292)
template <bool CC, class LV>
__device__ __forceinline__ double c0_inverse(const LV& L, int i, int j, double acf, double facx, double facy) {
    double c0 = -(acf + 2. * (facx + facy));
    if (CC && (i == L.ilo || i == L.ihi)) c0 -= 2. * facx;
    if (CC && (j == L.jlo || j == L.jhi)) c0 -= 2. * facy;
    return 1. / c0;
}

// the update of gs1<is_cell_centered> (HpMultiGrid.cpp:267-292); P(a, b) reads phi at
// offset (a, b)
template <bool CC, class LV, class PF>
__device__ __forceinline__ double gs_point(const LV& L, int i, int j, PF P, double rhs, double c0_inv, double facx,
                                           double facy) {
    double l;
    if (CC && i == L.ilo) l = facx * (4. / 3.) * P(1, 0);
    else if (CC && i == L.ihi) l = facx * (4. / 3.) * P(-1, 0);
    else l = facx * (P(-1, 0) + P(1, 0));
    if (CC && j == L.jlo) l += facy * (4. / 3.) * P(0, 1);
    else if (CC && j == L.jhi) l += facy * (4. / 3.) * P(0, -1);
    else l += facy * (P(0, -1) + P(0, 1));
    return (rhs - l) * c0_inv;
}

// laplacian + residual1 (HpMultiGrid.cpp:163-190). The box-edge branches only fire on
// cell-centred levels: on node-centred ones the edge nodes are not valid points.
template <bool CC, class LV, class PF>
__device__ __forceinline__ double residual1(const LV& L, int i, int j, PF P, double rhs, double acf, double facx,
                                            double facy) {
    double r = -2. * (facx + facy) * P(0, 0);
    if (CC && i == L.ilo) r += facx * (4. / 3. * P(1, 0) - 2. * P(0, 0));
    else if (CC && i == L.ihi) r += facx * (4. / 3. * P(-1, 0) - 2. * P(0, 0));
    else r += facx * (P(-1, 0) + P(1, 0));
    if (CC && j == L.jlo) r += facy * (4. / 3. * P(0, 1) - 2. * P(0, 0));
    else if (CC && j == L.jhi) r += facy * (4. / 3. * P(0, -1) - 2. * P(0, 0));
    else r += facy * (P(0, -1) + P(0, 1));
    return rhs + acf * P(0, 0) - r;
}

// fine_in + I(crse) at fine point (i, j) (interpcpy_cc / interpcpy_nd, :88-121);
// c(a, b) reads the coarse correction
template <bool CC, class CF>
__device__ __forceinline__ double interp_add(int i, int j, double fine_in, CF c) {
    const int ic = cdiv2(i), jc = cdiv2(j);
    if (CC) return fine_in + c(ic, jc);
    const bool io = (ic * 2 != i), jo = (jc * 2 != j);
    if (io && jo) return fine_in + (c(ic, jc) + c(ic + 1, jc) + c(ic, jc + 1) + c(ic + 1, jc + 1)) * 0.25;
    if (io) return fine_in + (c(ic, jc) + c(ic + 1, jc)) * 0.5;
    if (jo) return fine_in + (c(ic, jc) + c(ic, jc + 1)) * 0.5;
    return fine_in + c(ic, jc);
}

// restrict_cc / restrict_nd (:31-52) at coarse point (i, j); f(a, b) reads the fine array
template <bool CC, class FA>
__device__ __forceinline__ double restrict_from(FA f, int i, int j) {
    if (CC) return 0.25 * (f(2 * i, 2 * j) + f(2 * i + 1, 2 * j) + f(2 * i, 2 * j + 1) + f(2 * i + 1, 2 * j + 1));
    return (1. / 16.) * (f(2 * i - 1, 2 * j - 1) + 2. * f(2 * i, 2 * j - 1) + f(2 * i + 1, 2 * j - 1) +
                         2. * f(2 * i - 1, 2 * j) + 4. * f(2 * i, 2 * j) + 2. * f(2 * i + 1, 2 * j) +
                         f(2 * i - 1, 2 * j + 1) + 2. * f(2 * i, 2 * j + 1) + f(2 * i + 1, 2 * j + 1));
}

// NSW red-black sweeps (colours 0,1,2,3,0,1,...) of one level, LDS-tiled, optionally
// followed by the residual, its restriction to the coarse points this tile owns (res_c)
// and the max norm of the residual over the tile (atomicMax into *norm).
// Start: zero (ZERO), phi_in + I(cor_c) (INTERP) or phi_in. Result -> phi_out.
// PP (ping-pong): on odd V-cycles (ctl->iters odd) phi_in and phi_out swap roles, so a
// level-0 cycle never sweeps in place while neighbouring tiles read their halo.
// Each thread keeps rhs, acf and 1/c0 of the points it sweeps in registers.
template <int NSW, bool CC, bool ZERO, bool RES, bool INTERP, bool PP>
__global__ __launch_bounds__(NT) void k_smooth(MGLevel L, MGLevel C, double* __restrict__ phi_out,
                                               const double* __restrict__ rhs, long rcs,
                                               const double* __restrict__ acf, double* __restrict__ phi_in,
                                               const double* __restrict__ cor_c, double* __restrict__ res_c,
                                               unsigned long long* __restrict__ norm, const MGCtl* __restrict__ ctl) {
    constexpr int HALO = NSW + 2;
    constexpr int LX = TX + 2 * HALO, LY = TY + 2 * HALO;
    constexpr int SW = LX - 2, SH = LY - 2;        // swept region (inside the frozen ring)
    constexpr int KPT = (SW * SH + NT - 1) / NT;   // swept points per thread
    if (ctl && ctl->done) return;
    if (PP && (ctl->iters & 1)) {
        double* t = phi_out;
        phi_out = phi_in;
        phi_in = t;
    }
    __shared__ double s[2][LY][LX];
    __shared__ double r[2][RH][RW];
    __shared__ double red[NT / 64];
    const int ti0 = L.vilo + int(blockIdx.x) * TX - HALO;
    const int tj0 = L.vjlo + int(blockIdx.y) * TY - HALO;
    const double facx = 1. / (L.dx * L.dx), facy = 1. / (L.dy * L.dy);
    // own swept points: rhs, acf and 1/c0 into registers
    double r0[KPT], r1[KPT], ac[KPT], ci[KPT];
    int pl[KPT];   // lx | ly << 8, or -1 when outside the valid box or the swept region
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int e = threadIdx.x + q * NT;
        pl[q] = -1;
        r0[q] = r1[q] = ac[q] = ci[q] = 0.0;
        if (e < SW * SH) {
            const int ly = 1 + e / SW, lx = 1 + e - (ly - 1) * SW;
            const int i = ti0 + lx, j = tj0 + ly;
            if (in_valid(L, i, j)) {
                const long k = at(L, i, j);
                pl[q] = lx | (ly << 8);
                r0[q] = rhs[k];
                r1[q] = rhs[k + rcs];
                ac[q] = acf[k];
                ci[q] = c0_inverse<CC>(L, i, j, ac[q], facx, facy);
            }
        }
    }
    for (int e = threadIdx.x; e < LX * LY; e += NT) {
        const int ly = e / LX, lx = e - ly * LX;
        const int i = ti0 + lx, j = tj0 + ly;
        double v0 = 0.0, v1 = 0.0;
        if (!ZERO && in_valid(L, i, j)) {
            const long k = at(L, i, j);
            if (INTERP) {
                auto c0 = [&](int a, int b) { return cor_c[at(C, a, b)]; };
                auto c1 = [&](int a, int b) { return cor_c[at(C, a, b) + C.cs]; };
                v0 = interp_add<CC>(i, j, phi_in[k], c0);
                v1 = interp_add<CC>(i, j, phi_in[k + L.cs], c1);
            } else {
                v0 = phi_in[k];
                v1 = phi_in[k + L.cs];
            }
        }
        s[0][ly][lx] = v0;
        s[1][ly][lx] = v1;
    }
    __syncthreads();
    for (int color = 0; color < NSW; ++color) {
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            if (pl[q] < 0) continue;
            const int lx = pl[q] & 255, ly = pl[q] >> 8;
            const int i = ti0 + lx, j = tj0 + ly;
            if (((i + j + color) & 1) != 0) continue;
            auto P0 = [&](int di, int dj) { return s[0][ly + dj][lx + di]; };
            auto P1 = [&](int di, int dj) { return s[1][ly + dj][lx + di]; };
            const double n0 = gs_point<CC>(L, i, j, P0, r0[q], ci[q], facx, facy);
            const double n1 = gs_point<CC>(L, i, j, P1, r1[q], ci[q], facx, facy);
            s[0][ly][lx] = n0;
            s[1][ly][lx] = n1;
        }
        __syncthreads();
    }
    double m = 0.0;
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e >= SW * SH) continue;
        const int ly = 1 + e / SW, lx = 1 + e - (ly - 1) * SW;
        const bool out = lx >= HALO && lx < HALO + TX && ly >= HALO && ly < HALO + TY;
        if (RES) {
            const bool ring = lx >= HALO - 1 && lx <= HALO + TX && ly >= HALO - 1 && ly <= HALO + TY;
            if (ring) {
                double v0 = 0.0, v1 = 0.0;
                if (pl[q] >= 0) {
                    const int i = ti0 + lx, j = tj0 + ly;
                    auto P0 = [&](int di, int dj) { return s[0][ly + dj][lx + di]; };
                    auto P1 = [&](int di, int dj) { return s[1][ly + dj][lx + di]; };
                    v0 = residual1<CC>(L, i, j, P0, r0[q], ac[q], facx, facy);
                    v1 = residual1<CC>(L, i, j, P1, r1[q], ac[q], facx, facy);
                    if (out) m = fmax(m, fmax(fabs(v0), fabs(v1)));
                }
                r[0][ly - HALO + 1][lx - HALO + 1] = v0;
                r[1][ly - HALO + 1][lx - HALO + 1] = v1;
            }
        }
        if (out && pl[q] >= 0) {
            const long k = at(L, ti0 + lx, tj0 + ly);
            phi_out[k] = s[0][ly][lx];
            phi_out[k + L.cs] = s[1][ly][lx];
        }
    }
    if (RES) {
        __syncthreads();
        // coarse points whose centre fine point 2*ic lies in this tile
        const int ox0 = ti0 + HALO, oy0 = tj0 + HALO;
        const int ic0 = cdiv2(ox0 + 1), jc0 = cdiv2(oy0 + 1);
        for (int t = threadIdx.x; t < (TX / 2 + 1) * (TY / 2 + 1); t += NT) {
            const int ic = ic0 + t % (TX / 2 + 1), jc = jc0 + t / (TX / 2 + 1);
            if (2 * ic < ox0 || 2 * ic >= ox0 + TX || 2 * jc < oy0 || 2 * jc >= oy0 + TY) continue;
            if (!in_valid(C, ic, jc)) continue;
            // fine residual at (fi, fj) lives in r[.][fj - oy0 + 1][fi - ox0 + 1]
            auto f0 = [&](int fi, int fj) { return r[0][fj - oy0 + 1][fi - ox0 + 1]; };
            auto f1 = [&](int fi, int fj) { return r[1][fj - oy0 + 1][fi - ox0 + 1]; };
            res_c[at(C, ic, jc)] = restrict_from<CC>(f0, ic, jc);
            res_c[at(C, ic, jc) + C.cs] = restrict_from<CC>(f1, ic, jc);
        }
        if (norm) {
            for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
            __syncthreads();
            if (threadIdx.x == 0) {
                double mm = red[0];
                for (int w = 1; w < NT / 64; ++w) mm = fmax(mm, red[w]);
                atomicMax(norm, (unsigned long long)__double_as_longlong(mm));
            }
        }
    }
}

// crse (valid points) = R(fine), ncomp components
template <bool CC>
__global__ void k_restrict(MGLevel C, MGLevel F, double* __restrict__ crse, const double* __restrict__ fine,
                           int ncomp, long ccs, long fcs) {
    const int nvx = C.vihi - C.vilo + 1, nvy = C.vjhi - C.vjlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvx * nvy) return;
    const int j = C.vjlo + t / nvx, i = C.vilo + t % nvx;
    for (int n = 0; n < ncomp; ++n) {
        const double* f = fine + n * fcs;
        auto fa = [&](int a, int b) { return f[at(F, a, b)]; };
        crse[n * ccs + at(C, i, j)] = restrict_from<CC>(fa, i, j);
    }
}

// ---------------------------------------------------------------------------------
// coarse levels in one workgroup
// ---------------------------------------------------------------------------------
// Geometry of one coarse level and the offsets of its arrays in the work array (LDS, or
// global memory when the levels do not fit). Every array covers the level box plus one
// ghost ring, x fastest, as MGLevel's arrays do.
struct CLev {
    int ilo, ihi, jlo, jhi, vilo, vihi, vjlo, vjhi, ldx, cs;
    int acf, c0i, res, cor, rescor;
    double facx, facy;
};
struct CArgs {
    int nlev;           // levels cb .. max_level
    int bottom_sweeps;
    int work_doubles;
    CLev lv[COARSE_MAX_LEVELS];
    const double* acf;  // level cb in global memory
    const double* res;
    double* cor;
    long gcs;
};

__device__ __forceinline__ int lat(const CLev& L, int i, int j) { return (j - L.jlo + 1) * L.ldx + (i - L.ilo + 1); }

// thread (tx, ty) visits the valid points (vilo + tx + CTX a, vjlo + ty + CTY b)
template <class F>
__device__ __forceinline__ void for_valid(const CLev& L, F f) {
    const int tx = threadIdx.x % CTX, ty = threadIdx.x / CTX;
    for (int j = L.vjlo + ty; j <= L.vjhi; j += CTY)
        for (int i = L.vilo + tx; i <= L.vihi; i += CTX) f(i, j);
}
// the valid points with (i + j + color) even, two per CTX columns
template <class F>
__device__ __forceinline__ void for_color(const CLev& L, int color, F f) {
    const int tx = threadIdx.x % CTX, ty = threadIdx.x / CTX;
    for (int j = L.vjlo + ty; j <= L.vjhi; j += CTY) {
        const int i0 = L.vilo + ((L.vilo + j + color) & 1) + 2 * tx;
        for (int i = i0; i <= L.vihi; i += 2 * CTX) f(i, j);
    }
}

template <bool CC>
__device__ __forceinline__ void c_sweep(double* w, const CLev& L, int color) {
    const int n1 = L.cs;
    for_color(L, color, [&](int i, int j) {
        const int k = lat(L, i, j);
        const double* p0 = w + L.cor;
        const double* p1 = w + L.cor + n1;
        auto P0 = [&](int a, int b) { return p0[k + b * L.ldx + a]; };
        auto P1 = [&](int a, int b) { return p1[k + b * L.ldx + a]; };
        const double ci = w[L.c0i + k];
        const double v0 = gs_point<CC>(L, i, j, P0, w[L.res + k], ci, L.facx, L.facy);
        const double v1 = gs_point<CC>(L, i, j, P1, w[L.res + n1 + k], ci, L.facx, L.facy);
        w[L.cor + k] = v0;
        w[L.cor + n1 + k] = v1;
    });
    __syncthreads();
}

// Levels [cb, maxl] of one V-cycle: down-sweep with residual restriction, bottom solve
// (CPU branch of MultiGrid::bottomsolve, :1583-1593), up-sweep. acf and res of level cb
// are loaded; acf of the coarser levels is averaged down here; cor of level cb is written
// back for the finer up-sweep.
template <bool CC, bool IN_LDS>
__global__ __launch_bounds__(COARSE_THREADS) void k_coarse(const CArgs A, double* __restrict__ gwork,
                                                           const MGCtl* __restrict__ ctl) {
    if (ctl && ctl->done) return;
    extern __shared__ double sm[];
    double* w = IN_LDS ? sm : gwork;
    const int nl = A.nlev;
    {
        const CLev& T = A.lv[0];
        for (int l = 0; l < nl; ++l) {
            const CLev& L = A.lv[l];
            for (int t = threadIdx.x; t < 2 * L.cs; t += COARSE_THREADS) w[L.cor + t] = 0.0;
        }
        for (int t = threadIdx.x; t < T.cs; t += COARSE_THREADS) {
            w[T.acf + t] = A.acf[t];
            w[T.res + t] = A.res[t];
            w[T.res + T.cs + t] = A.res[t + A.gcs];
        }
        __syncthreads();
    }
    // average_down_acoef (:1641-1661) and the diagonal inverses
    for (int l = 0; l < nl; ++l) {
        const CLev& L = A.lv[l];
        for_valid(L, [&](int i, int j) {
            const int k = lat(L, i, j);
            w[L.c0i + k] = c0_inverse<CC>(L, i, j, w[L.acf + k], L.facx, L.facy);
        });
        if (l + 1 < nl) {
            const CLev& C = A.lv[l + 1];
            for_valid(C, [&](int i, int j) {
                auto fa = [&](int a, int b) { return w[L.acf + lat(L, a, b)]; };
                w[C.acf + lat(C, i, j)] = restrict_from<CC>(fa, i, j);
            });
        }
        __syncthreads();
    }
    // down-sweep
    for (int l = 0; l + 1 < nl; ++l) {
        const CLev& L = A.lv[l];
        const CLev& C = A.lv[l + 1];
        for (int color = 0; color < 4; ++color) c_sweep<CC>(w, L, color);
        for_valid(L, [&](int i, int j) {
            const int k = lat(L, i, j);
            const double a = w[L.acf + k];
            for (int n = 0; n < 2; ++n) {
                const double* p = w + L.cor + n * L.cs;
                auto P = [&](int di, int dj) { return p[k + dj * L.ldx + di]; };
                w[L.rescor + n * L.cs + k] = residual1<CC>(L, i, j, P, w[L.res + n * L.cs + k], a, L.facx, L.facy);
            }
        });
        __syncthreads();
        for_valid(C, [&](int i, int j) {
            for (int n = 0; n < 2; ++n) {
                auto fa = [&](int a, int b) { return w[L.rescor + n * L.cs + lat(L, a, b)]; };
                w[C.res + n * C.cs + lat(C, i, j)] = restrict_from<CC>(fa, i, j);
            }
        });
        __syncthreads();
    }
    // bottom solve from zero
    for (int sw = 0; sw < A.bottom_sweeps; ++sw) c_sweep<CC>(w, A.lv[nl - 1], sw);
    // up-sweep: cor += I(cor_coarse) in place (each point reads only itself on this level),
    // then four sweeps
    for (int l = nl - 2; l >= 0; --l) {
        const CLev& L = A.lv[l];
        const CLev& C = A.lv[l + 1];
        for_valid(L, [&](int i, int j) {
            const int k = lat(L, i, j);
            for (int n = 0; n < 2; ++n) {
                auto c = [&](int a, int b) { return w[C.cor + n * C.cs + lat(C, a, b)]; };
                w[L.cor + n * L.cs + k] = interp_add<CC>(i, j, w[L.cor + n * L.cs + k], c);
            }
        });
        __syncthreads();
        for (int color = 0; color < 4; ++color) c_sweep<CC>(w, L, color);
    }
    {
        const CLev& T = A.lv[0];
        for (int t = threadIdx.x; t < T.cs; t += COARSE_THREADS) {
            A.cor[t] = w[T.cor + t];
            A.cor[t + A.gcs] = w[T.cor + T.cs + t];
        }
    }
}

// level-0 load: acf over the whole box, sol and rhs (2 comps) over the whole box; max |rhs|
// over the valid box into ctl->rhs0_bits
__global__ void k_mg_load(MGLevel L, int sh, Geom g, const double* __restrict__ sol, const double* __restrict__ rhs,
                          const double* __restrict__ acf, long cstride, double* __restrict__ msol,
                          double* __restrict__ mrhs, MGCtl* __restrict__ ctl) {
    __shared__ double red[4];
    const int nbx = L.ihi - L.ilo + 1, nby = L.jhi - L.jlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < nbx * nby) {
        const int j = L.jlo + t / nbx, i = L.ilo + t % nbx;
        const long gk = g.idx(i - sh, j - sh);
        const long k = at(L, i, j);
        L.acf[k] = acf[gk];
        msol[k] = sol[gk];
        msol[k + L.cs] = sol[gk + cstride];
        const double a = rhs[gk], b = rhs[gk + cstride];
        mrhs[k] = a;
        mrhs[k + L.cs] = b;
        if (in_valid(L, i, j)) m = fmax(fabs(a), fabs(b));
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double mm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        atomicMax(&ctl->rhs0_bits, (unsigned long long)__double_as_longlong(mm));
    }
}

// sol (valid points) = the current level-0 iterate: cor after an even number of
// V-cycles, alt after an odd number (the level-0 cycle kernel alternates between them)
__global__ void k_mg_store(MGLevel L, int sh, Geom g, double* __restrict__ sol, long cstride,
                           const double* __restrict__ alt, const MGCtl* __restrict__ ctl) {
    const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvx * nvy) return;
    const double* src = (ctl->iters & 1) ? alt : L.cor;
    const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
    const long gk = g.idx(i - sh, j - sh);
    const long k = at(L, i, j);
    sol[gk] = src[k];
    sol[gk + cstride] = src[k + L.cs];
}

__global__ void k_ctl_reset(MGCtl* ctl, int maxiter) {
    MGCtl c;
    c.maxiter = maxiter;
    *ctl = c;
}

__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }

// convergence target after the initial residual (solve_doit :1330-1366)
__global__ void k_target(MGCtl* ctl, double tol_rel, double tol_abs) {
    const double resnorm0 = bits2d(ctl->res0_bits), rhsnorm0 = bits2d(ctl->rhs0_bits);
    const double max_norm = (rhsnorm0 >= resnorm0) ? rhsnorm0 : resnorm0;
    ctl->max_norm = max_norm;
    ctl->res_target = fmax(tol_abs, fmax(tol_rel, 1.e-16) * max_norm);
    ctl->norm_bits = 0;
    if (resnorm0 <= ctl->res_target) ctl->done = 1;
}

// end of one V-cycle (solve_doit :1371-1417)
__global__ void k_check(MGCtl* ctl) {
    if (ctl->done) return;
    ctl->iters += 1;
    const double norminf = bits2d(ctl->norm_bits);
    ctl->norm_bits = 0;
    if (norminf <= ctl->res_target) {
        ctl->done = 1;
    } else if (norminf > 1.e20 * ctl->max_norm) {
        ctl->done = 1;
        ctl->status = HPQ_E_MG_DIVERGED;
    } else if (ctl->iters >= ctl->maxiter) {
        ctl->done = 1;
        ctl->status = HPQ_E_MG_NOT_CONVERGED;
    }
}

bool coarsenable(bool cc, int lo, int hi, int min_width) {
    const int size = hi - lo + 1;
    if (size < 2 * min_width) return false;
    if (cc) return (cdiv2(lo) * 2 == lo) && (cdiv2(hi) * 2 + 1 == hi);
    return (cdiv2(lo) * 2 == lo) && (cdiv2(hi) * 2 == hi);
}

// the single-workgroup description of levels cb..max_level
void coarse_args(MultiGrid& M) {
    CArgs& A = *reinterpret_cast<CArgs*>(M.coarse_args);
    std::memset(&A, 0, sizeof(CArgs));
    const int cb = M.coarse_begin;
    A.nlev = M.max_level - cb + 1;
    int off = 0;
    for (int l = cb; l <= M.max_level; ++l) {
        const MGLevel& G = M.lev[l];
        CLev& L = A.lv[l - cb];
        L.ilo = G.ilo; L.ihi = G.ihi; L.jlo = G.jlo; L.jhi = G.jhi;
        L.vilo = G.vilo; L.vihi = G.vihi; L.vjlo = G.vjlo; L.vjhi = G.vjhi;
        L.ldx = G.ldx;
        L.cs = int(G.cs);
        L.acf = off; off += L.cs;
        L.c0i = off; off += L.cs;
        L.res = off; off += 2 * L.cs;
        L.cor = off; off += 2 * L.cs;
        L.rescor = off; off += 2 * L.cs;
        L.facx = 1. / (G.dx * G.dx);
        L.facy = 1. / (G.dy * G.dy);
    }
    A.work_doubles = off;
    const MGLevel& B = M.lev[M.max_level];
    const int len = std::max(B.ihi - B.ilo + 1, B.jhi - B.jlo + 1);
    A.bottom_sweeps = std::max(16, (len + 1) / 2 * 2);
    const MGLevel& T = M.lev[cb];
    A.acf = T.acf;
    A.res = T.res;
    A.cor = T.cor;
    A.gcs = T.cs;
}

}  // namespace

void mg_init(hpq_ctx* c) {
    static_assert(sizeof(CArgs) <= sizeof(MultiGrid::coarse_args), "coarse argument block too small");
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
    if (M.max_level - M.coarse_begin + 1 > COARSE_MAX_LEVELS)
        throw HpqError(HPQ_E_UNSUPPORTED, "hpmg: too many coarse levels");
    coarse_args(M);
    const CArgs& A = *reinterpret_cast<const CArgs*>(M.coarse_args);
    M.coarse_lds_bytes = sizeof(double) * size_t(A.work_doubles);
    M.coarse_in_lds = M.coarse_lds_bytes <= COARSE_LDS_MAX;
    if (!M.coarse_in_lds) HPQ_HIP(hipMalloc(&M.coarse_work, M.coarse_lds_bytes));
    const size_t b0 = sizeof(double) * 2 * M.lev[0].cs;
    HPQ_HIP(hipMalloc(&M.msol, b0));
    HPQ_HIP(hipMalloc(&M.mrhs, b0));
    HPQ_HIP(hipMemset(M.msol, 0, b0));
    HPQ_HIP(hipMemset(M.mrhs, 0, b0));
    HPQ_HIP(hipMalloc(&M.d_ctl, sizeof(MGCtl)));
    HPQ_HIP(hipHostMalloc(&M.h_ctl, sizeof(MGCtl), hipHostMallocDefault));
    if (M.coarse_in_lds) {
        const void* f = M.cc ? (const void*)k_coarse<true, true> : (const void*)k_coarse<false, true>;
        HPQ_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(M.coarse_lds_bytes)));
    }
}

void mg_free(hpq_ctx* c) {
    MultiGrid& M = c->mg;
    for (auto& L : M.lev)
        for (void* p : {(void*)L.acf, (void*)L.res, (void*)L.cor, (void*)L.rescor})
            if (p) (void)hipFree(p);
    if (M.coarse_work) (void)hipFree(M.coarse_work);
    if (M.msol) (void)hipFree(M.msol);
    if (M.mrhs) (void)hipFree(M.mrhs);
    if (M.d_ctl) (void)hipFree(M.d_ctl);
    if (M.h_ctl) (void)hipHostFree(M.h_ctl);
    M = MultiGrid{};
}

namespace {

template <int NSW, bool ZERO, bool RES, bool INTERP, bool PP>
void smooth(hpq_ctx* c, const MGLevel& L, const MGLevel& C, double* phi_out, const double* rhs, long rcs,
            const double* acf, double* phi_in, const double* cor_c, double* res_c, unsigned long long* norm,
            const MGCtl* ctl) {
    TraceScope ts(c, HPQ_TRACE_MG_SMOOTH);
    const dim3 grid((L.vihi - L.vilo + 1 + TX - 1) / TX, (L.vjhi - L.vjlo + 1 + TY - 1) / TY);
    if (L.cc)
        hipLaunchKernelGGL((k_smooth<NSW, true, ZERO, RES, INTERP, PP>), grid, dim3(NT), 0, c->stream, L, C, phi_out,
                           rhs, rcs, acf, phi_in, cor_c, res_c, norm, ctl);
    else
        hipLaunchKernelGGL((k_smooth<NSW, false, ZERO, RES, INTERP, PP>), grid, dim3(NT), 0, c->stream, L, C,
                           phi_out, rhs, rcs, acf, phi_in, cor_c, res_c, norm, ctl);
}

void launch_restrict(hpq_ctx* c, const MGLevel& C, const MGLevel& F, double* crse, const double* fine, int ncomp) {
    const int n = (C.vihi - C.vilo + 1) * (C.vjhi - C.vjlo + 1);
    if (C.cc)
        hipLaunchKernelGGL(k_restrict<true>, grid1d(n, 256), dim3(256), 0, c->stream, C, F, crse, fine, ncomp, C.cs,
                           F.cs);
    else
        hipLaunchKernelGGL(k_restrict<false>, grid1d(n, 256), dim3(256), 0, c->stream, C, F, crse, fine, ncomp, C.cs,
                           F.cs);
}

// MultiGrid::vcycle (:1430-1512). Level 0's residual and its restriction to level 1 come
// from the end of the previous cycle (or the initial residual); every kernel returns at
// once when the device-side control block says the solve is finished.
void vcycle(hpq_ctx* c) {
    MultiGrid& M = c->mg;
    MGCtl* ctl = M.d_ctl;
    const int cb = M.coarse_begin;
    for (int l = 1; l < cb; ++l) {
        MGLevel& L = M.lev[l];
        MGLevel& C = M.lev[l + 1];
        smooth<4, true, true, false, false>(c, L, C, L.cor, L.res, L.cs, L.acf, nullptr, nullptr, C.res, nullptr,
                                            ctl);
    }
    {
        TraceScope ts(c, HPQ_TRACE_MG_COARSE);
        const CArgs& A = *reinterpret_cast<const CArgs*>(M.coarse_args);
        if (M.coarse_in_lds) {
            if (M.cc)
                hipLaunchKernelGGL((k_coarse<true, true>), dim3(1), dim3(COARSE_THREADS), M.coarse_lds_bytes,
                                   c->stream, A, nullptr, ctl);
            else
                hipLaunchKernelGGL((k_coarse<false, true>), dim3(1), dim3(COARSE_THREADS), M.coarse_lds_bytes,
                                   c->stream, A, nullptr, ctl);
        } else {
            if (M.cc)
                hipLaunchKernelGGL((k_coarse<true, false>), dim3(1), dim3(COARSE_THREADS), 0, c->stream, A,
                                   M.coarse_work, ctl);
            else
                hipLaunchKernelGGL((k_coarse<false, false>), dim3(1), dim3(COARSE_THREADS), 0, c->stream, A,
                                   M.coarse_work, ctl);
        }
    }
    // Up-sweep: the result of level l >= 1 goes to rescor[l] (dead after the restriction),
    // so the kernel never overwrites cor[l], which neighbouring tiles read as their halo.
    // The coarse kernel leaves its top level in cor[cb].
    for (int l = cb - 1; l >= 1; --l) {
        MGLevel& L = M.lev[l];
        MGLevel& C = M.lev[l + 1];
        const double* corc = (l + 1 == cb) ? C.cor : C.rescor;
        smooth<4, false, false, true, false>(c, L, C, L.rescor, L.res, L.cs, L.acf, L.cor, corc, nullptr, nullptr,
                                             ctl);
    }
    // Level 0: the up-sweep's four sweeps and the cycle's final four sweeps + residual are
    // eight consecutive sweeps from x + I(cor_1), where x is the current iterate; the
    // intermediate after four sweeps is not needed. x alternates between cor[0] and msol.
    MGLevel& L0 = M.lev[0];
    MGLevel& L1 = M.lev[1];
    const double* corc = (cb == 1) ? L1.cor : L1.rescor;
    smooth<8, false, true, true, true>(c, L0, L1, M.msol, M.mrhs, L0.cs, L0.acf, L0.cor, corc, L1.res,
                                       &ctl->norm_bits, ctl);
    hipLaunchKernelGGL(k_check, dim3(1), dim3(1), 0, c->stream, ctl);
}

}  // namespace

// MultiGrid::solve1 + solve_doit (HpMultiGrid.cpp:1170-1190, 1308-1427)
int mg_solve1(hpq_ctx* c, double* sol, const double* rhs, const double* acf, long cstride, double tol_rel,
              double tol_abs, int maxiter, int* iters) {
    MultiGrid& M = c->mg;
    MGLevel& L0 = M.lev[0];
    MGLevel& L1 = M.lev[1];
    MGCtl* ctl = M.d_ctl;
    const int sh = M.cc ? 0 : 1;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, c->stream, ctl, maxiter);
    {
        const int n = (L0.ihi - L0.ilo + 1) * (L0.jhi - L0.jlo + 1);
        hipLaunchKernelGGL(k_mg_load, grid1d(n, 256), dim3(256), 0, c->stream, L0, sh, c->g, sol, rhs, acf, cstride,
                           M.msol, M.mrhs, ctl);
    }
    // average_down_acoef (:1641-1661): the tiled levels here, the coarse ones inside k_coarse
    for (int l = 1; l <= M.coarse_begin; ++l)
        launch_restrict(c, M.lev[l], M.lev[l - 1], M.lev[l].acf, M.lev[l - 1].acf, 1);
    // initial smoothing from the loaded guess (msol) into cor[0], which is the iterate
    // after zero V-cycles
    smooth<4, false, true, false, false>(c, L0, L1, L0.cor, M.mrhs, L0.cs, L0.acf, M.msol, nullptr, L1.res,
                                         &ctl->res0_bits, nullptr);
    hipLaunchKernelGGL(k_target, dim3(1), dim3(1), 0, c->stream, ctl, tol_rel, tol_abs);
    HPQ_HIP(hipGetLastError());
    int batch = std::max(1, std::min(M.last_iters, maxiter));
    int issued = 0;
    while (true) {
        for (int b = 0; b < batch && issued < maxiter; ++b, ++issued) vcycle(c);
        HPQ_HIP(hipGetLastError());
        HPQ_HIP(hipMemcpyAsync(M.h_ctl, ctl, sizeof(MGCtl), hipMemcpyDeviceToHost, c->stream));
        HPQ_HIP(hipStreamSynchronize(c->stream));
        if (M.h_ctl->done || issued >= maxiter) break;
        batch = 1;
    }
    const int it = M.h_ctl->iters;
    const int status = M.h_ctl->status;
    M.last_iters = std::max(1, it);
    if (iters) *iters = it;
    {
        const int n = (L0.vihi - L0.vilo + 1) * (L0.vjhi - L0.vjlo + 1);
        hipLaunchKernelGGL(k_mg_store, grid1d(n, 256), dim3(256), 0, c->stream, L0, sh, c->g, sol, cstride, M.msol,
                           ctl);
    }
    HPQ_HIP(hipGetLastError());
    return status;
}

}  // namespace hpq
