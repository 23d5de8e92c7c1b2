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
// final smoothing of the cycle are one launch of eight sweeps that leaves every
// workgroup's residual max norm; a one-workgroup launch after it reduces them, evaluates
// convergence and publishes the result to pinned host memory, which the host polls
// instead of synchronising the stream. Levels from the first one of at most 33x33
// points down run as one workgroup that does the coarse part of the V-cycle with barriers
// between sweeps, its arrays in LDS; the levels of at most 64 points (and the bottom
// solve) run in its first wavefront alone, without workgroup barriers. For odd n the
// level-0 arrays are the caller's planes themselves (the node-centred level-0 box with
// its ghost ring has the slice store's guarded shape).
#include <algorithm>
#include <chrono>
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
constexpr size_t COARSE_LDS_MAX = 150 * 1024;
constexpr int COARSE_MAX_LEVELS = 16;

__host__ __device__ __forceinline__ long at(const MGLevel& L, int i, int j) {
    return long(j - L.jlo + 1) * L.ldx + (i - L.ilo + 1);
}

// Tile of this workgroup. Workgroups go to the eight XCDs in turn (linear id mod 8), and
// each XCD has its own L2: the tiles are renumbered so that XCD k takes the k-th eighth of
// the tiles in row-major order, a band of whole tile rows, and the halos that neighbouring
// tiles read twice are mostly read from that XCD's L2 (ids beyond the last multiple of
// eight keep their tile).
__device__ __forceinline__ void xcd_tile(int& bx, int& by) {
    const int gx = int(gridDim.x), n = gx * int(gridDim.y);
    const int id = int(blockIdx.x) + int(blockIdx.y) * gx;
    const int m = n & ~7;
    const int t = id < m ? (id & 7) * (m >> 3) + (id >> 3) : id;
    by = t / gx;
    bx = t - by * gx;
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


__device__ void clear_ctl(MGCtl* c) {
    c->done = 0; c->iters = 0; c->status = 0;
    c->max_norm = 0; c->res_target = 0;
}

// one relaxed 64-bit store: the host sees the whole word or none of it
__device__ void publish(const MGCtl* c, MGMirror* m, int seq) {
    const unsigned long long w = (unsigned long long)(unsigned)seq | ((unsigned long long)(c->iters & 0xffffff) << 32) |
                                 ((unsigned long long)(c->status & 0x7f) << 56) |
                                 ((unsigned long long)(c->done ? 1 : 0) << 63);
    __hip_atomic_store(&m->word, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// end of one V-cycle (solve_doit :1371-1417), run by k_check with the max norm of the
// cycle's residual
__device__ void end_cycle(MGCtl* ctl, MGCtl* other, MGMirror* mirror, int seq, int maxiter, double norminf) {
    ctl->iters += 1;
    if (norminf <= ctl->res_target) {
        ctl->done = 1;
    } else if (norminf > 1.e20 * ctl->max_norm) {
        ctl->done = 1;
        ctl->status = HPQ_E_MG_DIVERGED;
    } else if (ctl->iters >= maxiter) {
        ctl->done = 1;
        ctl->status = HPQ_E_MG_NOT_CONVERGED;
    }
    if (ctl->done) clear_ctl(other);
    publish(ctl, mirror, seq);
}

// convergence target after the initial residual (solve_doit :1330-1366)
__device__ void end_initial(MGCtl* ctl, MGCtl* other, double tol_rel, double tol_abs, double resnorm0,
                            double rhsnorm0) {
    const double max_norm = (rhsnorm0 >= resnorm0) ? rhsnorm0 : resnorm0;
    ctl->max_norm = max_norm;
    ctl->res_target = fmax(tol_abs, fmax(tol_rel, 1.e-16) * max_norm);
    if (resnorm0 <= ctl->res_target) {
        ctl->done = 1;
        clear_ctl(other);
    }
}

// Arguments of the convergence bookkeeping after a level-0 launch.
struct Check {
    MGCtl* ctl;
    MGCtl* other;
    MGMirror* mirror;
    const double* norms;   // max |residual|, max |rhs| of every workgroup of the launch
    int nblk, seq, maxiter;
    double tol_rel, tol_abs;
};

// max over the workgroups' entries norms[2 b + k], one workgroup of NT threads
__device__ double max_norms(const double* norms, int nblk, int k, double* red) {
    double m = 0.0;
    for (int b = threadIdx.x; b < nblk; b += NT) m = fmax(m, norms[2 * b + k]);
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    double mm = red[0];
    for (int w = 1; w < NT / 64; ++w) mm = fmax(mm, red[w]);
    return mm;
}

// end of one V-cycle; a cycle enqueued after convergence only reports that it drained
__global__ __launch_bounds__(NT) void k_check(Check chk) {
    __shared__ double red[NT / 64];
    if (chk.ctl->done) {
        if (threadIdx.x == 0) publish(chk.ctl, chk.mirror, chk.seq);
        return;
    }
    const double norminf = max_norms(chk.norms, chk.nblk, 0, red);
    if (threadIdx.x == 0) end_cycle(chk.ctl, chk.other, chk.mirror, chk.seq, chk.maxiter, norminf);
}

// the convergence target; with seq != 0 the result (done after zero V-cycles or not) is
// published to the host as well
__global__ __launch_bounds__(NT) void k_target(Check chk) {
    __shared__ double red[NT / 64];
    const double resnorm0 = max_norms(chk.norms, chk.nblk, 0, red);
    const double rhsnorm0 = max_norms(chk.norms, chk.nblk, 1, red);
    if (threadIdx.x == 0) {
        end_initial(chk.ctl, chk.other, chk.tol_rel, chk.tol_abs, resnorm0, rhsnorm0);
        if (chk.seq) publish(chk.ctl, chk.mirror, chk.seq);
    }
}

// Convergence bookkeeping carried by the first launch of a V-cycle instead of a launch of
// its own. TARGET (before cycle 0): the convergence target from the initial residual and
// rhs norms, as k_target sets it. CHECK (before cycle m >= 1): the end of cycle m - 1, as
// k_check evaluates it, with m cycles done. Every workgroup of the launch reduces the
// level-0 norms and reaches the same decision; workgroup 0 records it in the control block.
struct Fold {
    enum : int { NONE = 0, TARGET = 1, CHECK = 2 };
    int mode = NONE;
    int m = 0, maxiter = 0, nblk = 0;
    double tol_rel = 0.0, tol_abs = 0.0;
    const double* norms = nullptr;
    MGCtl* ctl = nullptr;
    MGCtl* other = nullptr;
};

// the folded decision: whether the solve is finished before this launch's work
__device__ bool fold_decide(const Fold& f, const MGCtl* gate, double* red) {
    if (f.mode == Fold::NONE) return gate ? gate->done != 0 : false;
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0;
    if (f.mode == Fold::TARGET) {
        const double resnorm0 = max_norms(f.norms, f.nblk, 0, red);
        const double rhsnorm0 = max_norms(f.norms, f.nblk, 1, red);
        const double max_norm = (rhsnorm0 >= resnorm0) ? rhsnorm0 : resnorm0;
        const double target = fmax(f.tol_abs, fmax(f.tol_rel, 1.e-16) * max_norm);
        const bool done = resnorm0 <= target;
        if (lead) {
            f.ctl->max_norm = max_norm;
            f.ctl->res_target = target;
            if (done) {
                f.ctl->done = 1;
                clear_ctl(f.other);
            }
        }
        return done;
    }
    // CHECK: the control block's done flag is set by this solve's earlier launches, or by
    // workgroup 0 of this one with the decision every workgroup reaches here; one thread
    // reads it for the whole workgroup (after the barriers of the reduction), so all of its
    // threads take the same branch
    __shared__ int done_before;
    const double norminf = max_norms(f.norms, f.nblk, 0, red);
    if (threadIdx.x == 0) done_before = f.ctl->done;
    __syncthreads();
    if (done_before) return true;
    int status = 0;
    bool done = false;
    if (norminf <= f.ctl->res_target) {
        done = true;
    } else if (norminf > 1.e20 * f.ctl->max_norm) {
        done = true;
        status = HPQ_E_MG_DIVERGED;
    } else if (f.m >= f.maxiter) {
        done = true;
        status = HPQ_E_MG_NOT_CONVERGED;
    }
    if (lead) {
        f.ctl->iters = f.m;
        if (done) {
            f.ctl->status = status;
            f.ctl->done = 1;
            clear_ctl(f.other);
        }
    }
    return done;
}

// NSW red-black sweeps (colours 0,1,2,3,0,1,...) of one level, LDS-tiled, optionally
// followed by the residual and its restriction to the coarse points this tile owns (res_c).
// Start: zero (ZERO), phi_in + I(cor_c) (INTERP) or phi_in. Result -> phi_out.
// NORMS: the workgroup's max |residual| over its tile and (RHSMAX) max |rhs| over its
// points go to norms[2 b], norms[2 b + 1] for the convergence check that follows.
// ACF: acf is restricted to the coarse points the tile owns as well (average_down_acoef,
// :1641-1661), so the first pass over each level prepares the next level's coefficient.
// Every launch returns at once when the control block says the solve is finished (the
// data loads are issued before that flag is tested, so its latency is hidden).
//
// Every thread owns up to KP swept points of each parity of (i + j) and keeps their rhs,
// acf and 1/c0 in registers. A sweep of one colour touches one parity class only; the
// thread reads all neighbours of its points first, then updates, then writes (points of
// one colour never neighbour each other, so this is the order a global sweep has).
template <int NSW, bool CC, bool ZERO, bool RES, bool INTERP, bool RHSMAX, bool NORMS, bool ACF>
__global__ __launch_bounds__(NT) void k_smooth(MGLevel L, MGLevel C, double* __restrict__ phi_out,
                                               const double* __restrict__ rhs, long rcs,
                                               const double* __restrict__ acf, const double* __restrict__ phi_in,
                                               const double* __restrict__ cor_c, double* __restrict__ res_c,
                                               const MGCtl* __restrict__ gate, double* __restrict__ norms,
                                               Fold fold) {
    static_assert(NSW % 2 == 0, "sweeps come in red-black pairs");
    constexpr int HALO = NSW + 2;
    constexpr int LX = TX + 2 * HALO, LY = TY + 2 * HALO;
    constexpr int SW = LX - 2, SH = LY - 2;        // swept region (inside the frozen ring)
    constexpr int HW = SW / 2;                     // points of one parity per swept row
    constexpr int NPAR = HW * SH;                  // points of one parity
    constexpr int KP = (NPAR + NT - 1) / NT;       // points of one parity per thread
    constexpr int NLD = (LX * LY + NT - 1) / NT;   // start values loaded per thread
    // LDS holds the two colours of the tile apart: point (lx, ly) lives in
    // s[c][(base + lx + ly) & 1][ly][lx >> 1], so a colour sweep and its neighbour reads walk
    // consecutive doubles across the lanes (no two-way bank conflict from stride-2 rows)
    constexpr int LXH = LX / 2;
    static_assert(LX % 2 == 0, "a tile row holds whole point pairs");
    __shared__ double s[2][2][LY][LXH];
    __shared__ double r[2][RH][RW];
    __shared__ double red[2][NT / 64];
    __shared__ double ra[ACF ? RH : 1][ACF ? RW : 1];
    static_assert(!NORMS || RES, "the norms come from the residual");
    static_assert(!ACF || RES, "acf is restricted with the residual");
    const int gate_done = (fold.mode == Fold::NONE) ? (gate ? gate->done : 0) : 0;
    int bx, by;
    xcd_tile(bx, by);
    const int ti0 = L.vilo + bx * TX - HALO;
    const int tj0 = L.vjlo + by * TY - HALO;
    const int base = (ti0 + tj0) & 1;
    auto S = [&](int c, int ly, int lx) -> double& { return s[c][(base + lx + ly) & 1][ly][lx >> 1]; };
    const double facx = 1. / (L.dx * L.dx), facy = 1. / (L.dy * L.dy);
    // owned points: [parity][q]; pl = lx | ly << 8, or -1 outside the valid box / region;
    // acf stays in registers for the residual (ai), ci becomes 1/c0
    double r0[2][KP], r1[2][KP], ci[2][KP], ai[2][RES ? KP : 1];
    int pl[2][KP];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            const int e = threadIdx.x + q * NT;
            pl[par][q] = -1;
            r0[par][q] = r1[par][q] = ci[par][q] = 0.0;
            if (RES) ai[par][q] = 0.0;
            if (e < NPAR) {
                const int ly = 1 + e / HW;
                const int lx = 1 + 2 * (e - (ly - 1) * HW) + ((par + base + ly + 1) & 1);
                const int i = ti0 + lx, j = tj0 + ly;
                if (in_valid(L, i, j)) {
                    const long k = at(L, i, j);
                    pl[par][q] = lx | (ly << 8);
                    r0[par][q] = rhs[k];
                    r1[par][q] = rhs[k + rcs];
                    ci[par][q] = acf[k];   // acf for now, 1/c0 below
                    if (RES) ai[par][q] = ci[par][q];
                }
            }
        }
    }
    // start values: phi_in (and the coarse correction under the tile) into registers with
    // all loads in flight at once, then into LDS; the prolongation reads the coarse tile
    // from LDS
    constexpr int CLX = INTERP ? LX / 2 + 2 : 1, CLY = INTERP ? LY / 2 + 2 : 1;
    constexpr int NCL = INTERP ? (CLX * CLY + NT - 1) / NT : 1;
    __shared__ double cc[2][CLY][CLX];
    const int ic0 = cdiv2(ti0), jc0 = cdiv2(tj0);
    double v0[NLD], v1[NLD], cv0[NCL], cv1[NCL];
#pragma unroll
    for (int it = 0; it < NLD; ++it) {
        const int e = threadIdx.x + it * NT;
        const int ly = e / LX, lx = e - ly * LX;
        const int i = ti0 + lx, j = tj0 + ly;
        v0[it] = v1[it] = 0.0;
        if (!ZERO && e < LX * LY && in_valid(L, i, j)) {
            const long k = at(L, i, j);
            v0[it] = phi_in[k];
            v1[it] = phi_in[k + L.cs];
        }
    }
    if (INTERP) {
#pragma unroll
        for (int it = 0; it < NCL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int cy = e / CLX, cx = e - cy * CLX;
            const int ic = ic0 + cx, jc = jc0 + cy;
            cv0[it] = cv1[it] = 0.0;
            if (e < CLX * CLY && ic >= C.ilo - 1 && ic <= C.ihi + 1 && jc >= C.jlo - 1 && jc <= C.jhi + 1) {
                const long k = at(C, ic, jc);
                cv0[it] = cor_c[k];
                cv1[it] = cor_c[k + C.cs];
            }
        }
    }
    if (gate_done) return;
    // a folded convergence decision: its norm loads are issued after the tile's own
    if (fold.mode != Fold::NONE && fold_decide(fold, nullptr, &red[0][0])) return;
    if (INTERP) {
#pragma unroll
        for (int it = 0; it < NCL; ++it) {
            const int e = threadIdx.x + it * NT;
            if (e < CLX * CLY) {
                const int cy = e / CLX, cx = e - cy * CLX;
                cc[0][cy][cx] = cv0[it];
                cc[1][cy][cx] = cv1[it];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < NLD; ++it) {
        const int e = threadIdx.x + it * NT;
        if (e < LX * LY) {
            const int ly = e / LX, lx = e - ly * LX;
            double a0 = v0[it], a1 = v1[it];
            if (INTERP && in_valid(L, ti0 + lx, tj0 + ly)) {
                auto c0 = [&](int a, int b) { return cc[0][b - jc0][a - ic0]; };
                auto c1 = [&](int a, int b) { return cc[1][b - jc0][a - ic0]; };
                a0 = interp_add<CC>(ti0 + lx, tj0 + ly, a0, c0);
                a1 = interp_add<CC>(ti0 + lx, tj0 + ly, a1, c1);
            }
            S(0, ly, lx) = a0;
            S(1, ly, lx) = a1;
        }
    }
    double rmax = 0.0;
#pragma unroll
    for (int par = 0; par < 2; ++par) {
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            if (pl[par][q] < 0) continue;
            const int lx = pl[par][q] & 255, ly = pl[par][q] >> 8;
            ci[par][q] = c0_inverse<CC>(L, ti0 + lx, tj0 + ly, ci[par][q], facx, facy);
            if (RHSMAX) rmax = fmax(rmax, fmax(fabs(r0[par][q]), fabs(r1[par][q])));
        }
    }
    __syncthreads();
    // one colour: reads, then updates, then writes, in batches of up to KB points
    constexpr int KB = KP > 2 ? (KP + 1) / 2 : KP;
    auto sweep = [&](auto PARC) {
        constexpr int par = decltype(PARC)::value;
        // a point of colour par has (base + lx + ly) & 1 == par; its neighbours the other one
        constexpr int np = 1 - par;
#pragma unroll
        for (int q0 = 0; q0 < KP; q0 += KB) {
            double w0[KB], e0[KB], s0[KB], n0[KB], w1[KB], e1[KB], s1[KB], n1[KB];
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                const int q = q0 + b;
                if (q >= KP || pl[par][q] < 0) continue;
                const int lx = pl[par][q] & 255, ly = pl[par][q] >> 8;
                const int xw = (lx - 1) >> 1, xe = (lx + 1) >> 1, xc = lx >> 1;
                w0[b] = s[0][np][ly][xw]; e0[b] = s[0][np][ly][xe]; s0[b] = s[0][np][ly - 1][xc]; n0[b] = s[0][np][ly + 1][xc];
                w1[b] = s[1][np][ly][xw]; e1[b] = s[1][np][ly][xe]; s1[b] = s[1][np][ly - 1][xc]; n1[b] = s[1][np][ly + 1][xc];
            }
            double u0[KB], u1[KB];
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                const int q = q0 + b;
                if (q >= KP || pl[par][q] < 0) continue;
                const int lx = pl[par][q] & 255, ly = pl[par][q] >> 8;
                const int i = ti0 + lx, j = tj0 + ly;
                auto P0 = [&](int di, int dj) { return di < 0 ? w0[b] : di > 0 ? e0[b] : dj < 0 ? s0[b] : n0[b]; };
                auto P1 = [&](int di, int dj) { return di < 0 ? w1[b] : di > 0 ? e1[b] : dj < 0 ? s1[b] : n1[b]; };
                u0[b] = gs_point<CC>(L, i, j, P0, r0[par][q], ci[par][q], facx, facy);
                u1[b] = gs_point<CC>(L, i, j, P1, r1[par][q], ci[par][q], facx, facy);
            }
#pragma unroll
            for (int b = 0; b < KB; ++b) {
                const int q = q0 + b;
                if (q >= KP || pl[par][q] < 0) continue;
                const int lx = pl[par][q] & 255, ly = pl[par][q] >> 8;
                s[0][par][ly][lx >> 1] = u0[b];
                s[1][par][ly][lx >> 1] = u1[b];
            }
        }
        __syncthreads();
    };
    // colour c updates the points with (i + j + c) even, i.e. parity c & 1
    for (int c2 = 0; c2 < NSW; c2 += 2) {
        sweep(std::integral_constant<int, 0>{});
        sweep(std::integral_constant<int, 1>{});
    }
    double m = 0.0;
#pragma unroll
    for (int par = 0; par < 2; ++par) {
#pragma unroll
        for (int q = 0; q < KP; ++q) {
            if (pl[par][q] < 0) continue;
            const int lx = pl[par][q] & 255, ly = pl[par][q] >> 8;
            const bool out = lx >= HALO && lx < HALO + TX && ly >= HALO && ly < HALO + TY;
            if (RES) {
                const bool ring = lx >= HALO - 1 && lx <= HALO + TX && ly >= HALO - 1 && ly <= HALO + TY;
                if (ring) {
                    const int i = ti0 + lx, j = tj0 + ly;
                    const double a = ai[par][RES ? q : 0];
                    // the point has colour par, its four neighbours the other one
                    auto P0 = [&](int di, int dj) {
                        return (di | dj) ? s[0][1 - par][ly + dj][(lx + di) >> 1] : s[0][par][ly][lx >> 1];
                    };
                    auto P1 = [&](int di, int dj) {
                        return (di | dj) ? s[1][1 - par][ly + dj][(lx + di) >> 1] : s[1][par][ly][lx >> 1];
                    };
                    const double g0 = residual1<CC>(L, i, j, P0, r0[par][q], a, facx, facy);
                    const double g1 = residual1<CC>(L, i, j, P1, r1[par][q], a, facx, facy);
                    if (out) m = fmax(m, fmax(fabs(g0), fabs(g1)));
                    r[0][ly - HALO + 1][lx - HALO + 1] = g0;
                    r[1][ly - HALO + 1][lx - HALO + 1] = g1;
                    if (ACF) ra[ly - HALO + 1][lx - HALO + 1] = a;
                }
            }
            if (out) {
                const long k = at(L, ti0 + lx, tj0 + ly);
                phi_out[k] = s[0][par][ly][lx >> 1];
                phi_out[k + L.cs] = s[1][par][ly][lx >> 1];
            }
        }
    }
    if (RES) {
        // ring points outside the valid box hold zero residual
        for (int e = threadIdx.x; e < RW * RH; e += NT) {
            const int ry = e / RW, rx = e - ry * RW;
            if (!in_valid(L, ti0 + HALO - 1 + rx, tj0 + HALO - 1 + ry)) {
                r[0][ry][rx] = 0.0;
                r[1][ry][rx] = 0.0;
                if (ACF) ra[ry][rx] = 0.0;
            }
        }
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
            if (ACF) {
                auto fa = [&](int fi, int fj) { return ra[fj - oy0 + 1][fi - ox0 + 1]; };
                C.acf[at(C, ic, jc)] = restrict_from<CC>(fa, ic, jc);
            }
        }
    }
    if (NORMS) {
        for (int o = 32; o > 0; o >>= 1) {
            m = fmax(m, __shfl_xor(m, o));
            if (RHSMAX) rmax = fmax(rmax, __shfl_xor(rmax, o));
        }
        if ((threadIdx.x & 63) == 0) {
            red[0][threadIdx.x >> 6] = m;
            red[1][threadIdx.x >> 6] = rmax;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double mm = red[0][0], rr = red[1][0];
            for (int w = 1; w < NT / 64; ++w) {
                mm = fmax(mm, red[0][w]);
                rr = fmax(rr, red[1][w]);
            }
            const unsigned b = by * gridDim.x + bx;
            norms[2 * b] = mm;
            norms[2 * b + 1] = rr;
        }
    }
}

// ---------------------------------------------------------------------------------
// coarse levels in one workgroup
// ---------------------------------------------------------------------------------
// Geometry of one coarse level and the offsets of its arrays in the work array (LDS, or
// global memory when the levels do not fit). Every array covers the level box plus one
// ghost ring, x fastest, as MGLevel's arrays do. The two components are independent
// systems with the same acf, so each has a workgroup of its own: the work array holds one
// component (a second one follows it when the arrays live in global memory).
struct CLev {
    int ilo, ihi, jlo, jhi, vilo, vihi, vjlo, vjhi, ldx, cs;
    int acf, c0i, res, cor, rescor;
    double facx, facy;
};
struct CArgs {
    int nlev;           // levels cb .. max_level
    int nblock;         // levels [0, nblock) run on the whole workgroup, the rest on wave 0
    int bottom_sweeps;
    int work_doubles;
    CLev lv[COARSE_MAX_LEVELS];
    const double* acf;  // level cb in global memory
    const double* res;
    double* cor;
    long gcs;
};

__device__ __forceinline__ int lat(const CLev& L, int i, int j) { return (j - L.jlo + 1) * L.ldx + (i - L.ilo + 1); }

// Thread layouts of the coarse kernel. Block: thread (tx, ty) of CTX x CTY visits the
// valid points (vilo + tx + CTX a, vjlo + ty + CTY b); a colour sweep gives each thread
// two columns of CTX. Wave: the 64 lanes of wavefront 0 walk the valid points in order;
// LDS traffic of one wavefront is ordered, so a wait for the LDS counter (plus a compiler
// barrier) separates phases.
struct BlockMode {
    template <class F>
    static __device__ __forceinline__ void valid(const CLev& L, F f) {
        const int tx = threadIdx.x % CTX, ty = threadIdx.x / CTX;
        for (int j = L.vjlo + ty; j <= L.vjhi; j += CTY)
            for (int i = L.vilo + tx; i <= L.vihi; i += CTX) f(i, j);
    }
    template <class F>
    static __device__ __forceinline__ void color(const CLev& L, int c, F f) {
        const int tx = threadIdx.x % CTX, ty = threadIdx.x / CTX;
        for (int j = L.vjlo + ty; j <= L.vjhi; j += CTY) {
            const int i0 = L.vilo + ((L.vilo + j + c) & 1) + 2 * tx;
            for (int i = i0; i <= L.vihi; i += 2 * CTX) f(i, j);
        }
    }
    static __device__ __forceinline__ void sync() { __syncthreads(); }
};
struct WaveMode {
    template <class F>
    static __device__ __forceinline__ void valid(const CLev& L, F f) {
        const int nvx = L.vihi - L.vilo + 1, n = nvx * (L.vjhi - L.vjlo + 1);
        for (int e = threadIdx.x; e < n; e += 64) f(L.vilo + e % nvx, L.vjlo + e / nvx);
    }
    template <class F>
    static __device__ __forceinline__ void color(const CLev& L, int c, F f) {
        valid(L, [&](int i, int j) {
            if (((i + j + c) & 1) == 0) f(i, j);
        });
    }
    static __device__ __forceinline__ void sync() {
        __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
        __builtin_amdgcn_wave_barrier();
    }
};

template <class Mode, bool CC>
__device__ __forceinline__ void c_sweep(double* w, const CLev& L, int color) {
    Mode::color(L, color, [&](int i, int j) {
        const int k = lat(L, i, j);
        const double* p = w + L.cor;
        auto P = [&](int a, int b) { return p[k + b * L.ldx + a]; };
        w[L.cor + k] = gs_point<CC>(L, i, j, P, w[L.res + k], w[L.c0i + k], L.facx, L.facy);
    });
    Mode::sync();
}

// down-sweep step of level L: four sweeps from zero, residual, restriction into C.res
template <class Mode, bool CC>
__device__ void c_down(double* w, const CLev& L, const CLev& C) {
    for (int color = 0; color < 4; ++color) c_sweep<Mode, CC>(w, L, color);
    Mode::valid(L, [&](int i, int j) {
        const int k = lat(L, i, j);
        const double* p = w + L.cor;
        auto P = [&](int di, int dj) { return p[k + dj * L.ldx + di]; };
        w[L.rescor + k] = residual1<CC>(L, i, j, P, w[L.res + k], w[L.acf + k], L.facx, L.facy);
    });
    Mode::sync();
    Mode::valid(C, [&](int i, int j) {
        auto fa = [&](int a, int b) { return w[L.rescor + lat(L, a, b)]; };
        w[C.res + lat(C, i, j)] = restrict_from<CC>(fa, i, j);
    });
    Mode::sync();
}

// up-sweep step of level L: cor += I(C.cor) in place (each point reads only itself on this
// level), then four sweeps
template <class Mode, bool CC>
__device__ void c_up(double* w, const CLev& L, const CLev& C) {
    Mode::valid(L, [&](int i, int j) {
        const int k = lat(L, i, j);
        auto c = [&](int a, int b) { return w[C.cor + lat(C, a, b)]; };
        w[L.cor + k] = interp_add<CC>(i, j, w[L.cor + k], c);
    });
    Mode::sync();
    for (int color = 0; color < 4; ++color) c_sweep<Mode, CC>(w, L, color);
}

// Levels [cb, maxl] of one V-cycle for component blockIdx.x: down-sweep with residual
// restriction, bottom solve (CPU branch of MultiGrid::bottomsolve, :1583-1593), up-sweep.
// acf and res of level cb are loaded; acf of the coarser levels is averaged down here;
// cor of level cb is written back for the finer up-sweep.
template <bool CC, bool IN_LDS>
__global__ __launch_bounds__(COARSE_THREADS) void k_coarse(const CArgs A, double* __restrict__ gwork,
                                                           const MGCtl* __restrict__ ctl) {
    constexpr int NB = 8;   // loads in flight per thread
    extern __shared__ double sm[];
    const int comp = blockIdx.x;
    double* w = IN_LDS ? sm : gwork + long(comp) * A.work_doubles;
    const int nl = A.nlev;
    const CLev& T = A.lv[0];
    const int done = ctl ? ctl->done : 0;
    const double* gres = A.res + long(comp) * A.gcs;
    // acf, res of level cb: up to NB loads per thread in flight at once
    const int nload = 2 * T.cs;
    for (int b0 = 0; b0 < nload; b0 += NB * COARSE_THREADS) {
        double v[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int t = b0 + q * COARSE_THREADS + threadIdx.x;
            v[q] = 0.0;
            if (t < nload) v[q] = t < T.cs ? A.acf[t] : gres[t - T.cs];
        }
        if (done) return;
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            const int t = b0 + q * COARSE_THREADS + threadIdx.x;
            if (t < nload) w[t < T.cs ? T.acf + t : T.res + (t - T.cs)] = v[q];
        }
    }
    for (int l = 0; l < nl; ++l) {
        const CLev& L = A.lv[l];
        for (int t = threadIdx.x; t < L.cs; t += COARSE_THREADS) w[L.cor + t] = 0.0;
    }
    __syncthreads();
    // average_down_acoef (:1641-1661) and the diagonal inverses
    for (int l = 0; l < nl; ++l) {
        const CLev& L = A.lv[l];
        BlockMode::valid(L, [&](int i, int j) {
            const int k = lat(L, i, j);
            w[L.c0i + k] = c0_inverse<CC>(L, i, j, w[L.acf + k], L.facx, L.facy);
        });
        if (l + 1 < nl) {
            const CLev& C = A.lv[l + 1];
            BlockMode::valid(C, [&](int i, int j) {
                auto fa = [&](int a, int b) { return w[L.acf + lat(L, a, b)]; };
                w[C.acf + lat(C, i, j)] = restrict_from<CC>(fa, i, j);
            });
        }
        __syncthreads();
    }
    const int nb = A.nblock;
    for (int l = 0; l < nb && l + 1 < nl; ++l) c_down<BlockMode, CC>(w, A.lv[l], A.lv[l + 1]);
    if (threadIdx.x < 64) {
        for (int l = nb; l + 1 < nl; ++l) c_down<WaveMode, CC>(w, A.lv[l], A.lv[l + 1]);
        // bottom solve from zero (the bottom level is always a wavefront level)
        for (int sw = 0; sw < A.bottom_sweeps; ++sw) c_sweep<WaveMode, CC>(w, A.lv[nl - 1], sw);
        for (int l = nl - 2; l >= nb; --l) c_up<WaveMode, CC>(w, A.lv[l], A.lv[l + 1]);
    }
    __syncthreads();
    for (int l = min(nb, nl - 1) - 1; l >= 0; --l) c_up<BlockMode, CC>(w, A.lv[l], A.lv[l + 1]);
    double* gcor = A.cor + long(comp) * A.gcs;
    for (int t = threadIdx.x; t < T.cs; t += COARSE_THREADS) gcor[t] = w[T.cor + t];
}

// level-0 load (copy path): acf, sol and rhs (2 comps) over the whole box
__global__ void k_mg_load(MGLevel L, int sh, Geom g, const double* __restrict__ sol, const double* __restrict__ rhs,
                          const double* __restrict__ acf, long cstride, double* __restrict__ acf0,
                          double* __restrict__ msol, double* __restrict__ mrhs) {
    const int nbx = L.ihi - L.ilo + 1, nby = L.jhi - L.jlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbx * nby) return;
    const int j = L.jlo + t / nbx, i = L.ilo + t % nbx;
    const long gk = g.idx(i - sh, j - sh);
    const long k = at(L, i, j);
    acf0[k] = acf[gk];
    msol[k] = sol[gk];
    msol[k + L.cs] = sol[gk + cstride];
    mrhs[k] = rhs[gk];
    mrhs[k + L.cs] = rhs[gk + cstride];
}

// sol (valid points) = src, the level-0 iterate
__global__ void k_mg_store(MGLevel L, int sh, Geom g, double* __restrict__ sol, long cstride,
                           const double* __restrict__ src) {
    const int nvx = L.vihi - L.vilo + 1, nvy = L.vjhi - L.vjlo + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvx * nvy) return;
    const int j = L.vjlo + t / nvx, i = L.vilo + t % nvx;
    const long gk = g.idx(i - sh, j - sh);
    const long k = at(L, i, j);
    sol[gk] = src[k];
    sol[gk + cstride] = src[k + L.cs];
}

// initial guess of the next solve: b = 2 b - h and h = old b (EXTRAP), or h = b (record only)
template <bool EXTRAP>
__global__ void k_guess_linear(double* __restrict__ b, double* __restrict__ h, long n) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double b1 = b[t];
    if (EXTRAP) b[t] = 2.0 * b1 - h[t];
    h[t] = b1;
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
        L.res = off; off += L.cs;
        L.cor = off; off += L.cs;
        L.rescor = off; off += L.cs;
        L.facx = 1. / (G.dx * G.dx);
        L.facy = 1. / (G.dy * G.dy);
    }
    A.work_doubles = off;
    // levels of more than 64 valid points run on the whole workgroup; the bottom level
    // always on one wavefront
    A.nblock = 0;
    while (A.nblock < A.nlev - 1) {
        const CLev& L = A.lv[A.nblock];
        if ((L.vihi - L.vilo + 1) * (L.vjhi - L.vjlo + 1) <= 64) break;
        ++A.nblock;
    }
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
    if (!M.coarse_in_lds) HPQ_HIP(hipMalloc(&M.coarse_work, 2 * M.coarse_lds_bytes));
    const size_t b0 = sizeof(double) * 2 * M.lev[0].cs;
    HPQ_HIP(hipMalloc(&M.msol, b0));
    HPQ_HIP(hipMalloc(&M.mrhs, b0));
    HPQ_HIP(hipMemset(M.msol, 0, b0));
    HPQ_HIP(hipMemset(M.mrhs, 0, b0));
    HPQ_HIP(hipMalloc(&M.d_ctl, 2 * sizeof(MGCtl)));
    HPQ_HIP(hipMemset(M.d_ctl, 0, 2 * sizeof(MGCtl)));
    {
        const MGLevel& L0 = M.lev[0];
        const long nblk = long((L0.vihi - L0.vilo + 1 + TX - 1) / TX) * ((L0.vjhi - L0.vjlo + 1 + TY - 1) / TY);
        HPQ_HIP(hipMalloc(&M.d_norms, sizeof(double) * 2 * nblk));
    }
    HPQ_HIP(hipHostMalloc(&M.h_mirror, sizeof(MGMirror), hipHostMallocMapped | hipHostMallocCoherent));
    M.h_mirror->word = 0;
    HPQ_HIP(hipHostGetDevicePointer((void**)&M.d_mirror, M.h_mirror, 0));
    if (M.coarse_in_lds) {
        const void* f = M.cc ? (const void*)k_coarse<true, true> : (const void*)k_coarse<false, true>;
        HPQ_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(M.coarse_lds_bytes)));
    }
}

void mg_guess_linear(hpq_ctx* c, double* bxby, int islice) {
    MultiGrid& M = c->mg;
    const long n = 2 * c->g.plane;
    if (!M.hist) HPQ_HIP(hipMalloc(&M.hist, sizeof(double) * n));
    const dim3 grid(unsigned((n + 255) / 256));
    if (M.hist_slice == islice + 2)
        hipLaunchKernelGGL(k_guess_linear<true>, grid, dim3(256), 0, c->stream, bxby, M.hist, n);
    else
        hipLaunchKernelGGL(k_guess_linear<false>, grid, dim3(256), 0, c->stream, bxby, M.hist, n);
    HPQ_HIP(hipGetLastError());
    M.hist_slice = islice + 1;
}

void mg_free(hpq_ctx* c) {
    MultiGrid& M = c->mg;
    if (M.hist) (void)hipFree(M.hist);
    for (auto& L : M.lev)
        for (void* p : {(void*)L.acf, (void*)L.res, (void*)L.cor, (void*)L.rescor})
            if (p) (void)hipFree(p);
    if (M.coarse_work) (void)hipFree(M.coarse_work);
    if (M.msol) (void)hipFree(M.msol);
    if (M.mrhs) (void)hipFree(M.mrhs);
    if (M.d_ctl) (void)hipFree(M.d_ctl);
    if (M.d_norms) (void)hipFree(M.d_norms);
    if (M.h_mirror) (void)hipHostFree(M.h_mirror);
    M = MultiGrid{};
}

namespace {

int level0_blocks(const MGLevel& L) {
    return ((L.vihi - L.vilo + 1 + TX - 1) / TX) * ((L.vjhi - L.vjlo + 1 + TY - 1) / TY);
}

// the caller's view of one solve: level-0 rhs, acf and the two iterate buffers
struct Solve {
    const double* rhs0;   // comp stride rcs
    long rcs;
    MGLevel L0;           // level 0 with acf pointing at the level-0 acf
    double* buf[2];       // iterate after an even / odd number of V-cycles
    MGCtl* ctl;
    MGCtl* other;
    double tol_rel, tol_abs;
};

template <int NSW, bool ZERO, bool RES, bool INTERP, bool RHSMAX = false, bool NORMS = false, bool ACF = false>
void smooth(hpq_ctx* c, const MGLevel& L, const MGLevel& C, double* phi_out, const double* rhs, long rcs,
            const double* acf, const double* phi_in, const double* cor_c, double* res_c, const MGCtl* gate,
            double* norms = nullptr, const Fold& fold = Fold{}) {
    TraceScope ts(c, HPQ_TRACE_MG_SMOOTH);
    const dim3 grid((L.vihi - L.vilo + 1 + TX - 1) / TX, (L.vjhi - L.vjlo + 1 + TY - 1) / TY);
    if (L.cc)
        hipLaunchKernelGGL((k_smooth<NSW, true, ZERO, RES, INTERP, RHSMAX, NORMS, ACF>), grid, dim3(NT), 0, c->stream, L,
                           C, phi_out, rhs, rcs, acf, phi_in, cor_c, res_c, gate, norms, fold);
    else
        hipLaunchKernelGGL((k_smooth<NSW, false, ZERO, RES, INTERP, RHSMAX, NORMS, ACF>), grid, dim3(NT), 0, c->stream,
                           L, C, phi_out, rhs, rcs, acf, phi_in, cor_c, res_c, gate, norms, fold);
}

// MultiGrid::vcycle (:1430-1512), V-cycle number m of the solve. Level 0's residual and
// its restriction to level 1 come from the end of the previous cycle (or the initial
// residual); every kernel returns at once when the control block says the solve is
// finished. Cycle m runs only if cycles 0..m-1 ran, so the host knows that it starts from
// buf[m & 1] and ends in buf[(m + 1) & 1]. fold: the convergence target (m = 0) or the end
// of cycle m - 1 is evaluated by the cycle's first launch (level 1 down); check: the cycle
// ends with a k_check launch that publishes its result to the host.
void vcycle(hpq_ctx* c, const Solve& S, int m, int maxiter, bool fold, bool check) {
    MultiGrid& M = c->mg;
    MGCtl* ctl = S.ctl;
    const int cb = M.coarse_begin;
    for (int l = 1; l < cb; ++l) {
        MGLevel& L = M.lev[l];
        MGLevel& C = M.lev[l + 1];
        Fold f;
        if (fold && l == 1) {
            f.mode = m == 0 ? Fold::TARGET : Fold::CHECK;
            f.m = m;
            f.maxiter = maxiter;
            f.nblk = level0_blocks(S.L0);
            f.tol_rel = S.tol_rel;
            f.tol_abs = S.tol_abs;
            f.norms = M.d_norms;
            f.ctl = ctl;
            f.other = S.other;
        }
        if (m == 0)
            smooth<4, true, true, false, false, false, true>(c, L, C, L.cor, L.res, L.cs, L.acf, nullptr, nullptr,
                                                             C.res, ctl, nullptr, f);
        else
            smooth<4, true, true, false>(c, L, C, L.cor, L.res, L.cs, L.acf, nullptr, nullptr, C.res, ctl, nullptr, f);
    }
    {
        TraceScope ts(c, HPQ_TRACE_MG_COARSE);
        const CArgs& A = *reinterpret_cast<const CArgs*>(M.coarse_args);
        if (M.coarse_in_lds) {
            if (M.cc)
                hipLaunchKernelGGL((k_coarse<true, true>), dim3(2), dim3(COARSE_THREADS), M.coarse_lds_bytes,
                                   c->stream, A, nullptr, ctl);
            else
                hipLaunchKernelGGL((k_coarse<false, true>), dim3(2), dim3(COARSE_THREADS), M.coarse_lds_bytes,
                                   c->stream, A, nullptr, ctl);
        } else {
            if (M.cc)
                hipLaunchKernelGGL((k_coarse<true, false>), dim3(2), dim3(COARSE_THREADS), 0, c->stream, A,
                                   M.coarse_work, ctl);
            else
                hipLaunchKernelGGL((k_coarse<false, false>), dim3(2), dim3(COARSE_THREADS), 0, c->stream, A,
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
        smooth<4, false, false, true>(c, L, C, L.rescor, L.res, L.cs, L.acf, L.cor, corc, nullptr, ctl);
    }
    // Level 0: the up-sweep's four sweeps and the cycle's final four sweeps + residual are
    // eight consecutive sweeps from x + I(cor_1), where x is the current iterate; the
    // intermediate after four sweeps is not needed.
    MGLevel& L1 = M.lev[1];
    const double* corc = (cb == 1) ? L1.cor : L1.rescor;
    smooth<8, false, true, true, false, true>(c, S.L0, L1, S.buf[(m + 1) & 1], S.rhs0, S.rcs, S.L0.acf, S.buf[m & 1],
                                              corc, L1.res, ctl, M.d_norms);
    if (check) {
        const Check chk{ctl, S.other, M.d_mirror, M.d_norms, level0_blocks(S.L0), ++M.seq, maxiter, 0.0, 0.0};
        hipLaunchKernelGGL(k_check, dim3(1), dim3(NT), 0, c->stream, chk);
    }
}

struct Published {
    int seq, iters, status;
    bool done;
};

Published decode(unsigned long long w) {
    return {int(unsigned(w & 0xffffffffull)), int((w >> 32) & 0xffffff), int((w >> 56) & 0x7f), (w >> 63) != 0};
}

// waits until the device has published sequence number seq
Published wait_mirror(hpq_ctx* c, int seq) {
    MultiGrid& M = c->mg;
    unsigned long long* word = &M.h_mirror->word;
    const auto t0 = std::chrono::steady_clock::now();
    for (long spin = 0;; ++spin) {
        const Published p = decode(__atomic_load_n(word, __ATOMIC_ACQUIRE));
        if (p.seq == seq) return p;
        if ((spin & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(30)) break;
    }
    // the device has not reported: let the stream drain, then look once more
    HPQ_HIP(hipStreamSynchronize(c->stream));
    const Published p = decode(__atomic_load_n(word, __ATOMIC_ACQUIRE));
    if (p.seq != seq) throw HpqError(HPQ_E_HIP, "hpmg: the device did not report convergence");
    return p;
}

}  // namespace

// MultiGrid::solve1 + solve_doit (HpMultiGrid.cpp:1170-1190, 1308-1427)
int mg_solve1(hpq_ctx* c, double* sol, const double* rhs, const double* acf, long cstride, double tol_rel,
              double tol_abs, int maxiter, int* iters) {
    MultiGrid& M = c->mg;
    MGLevel& L1 = M.lev[1];
    if (M.ctl_dirty) {
        HPQ_HIP(hipMemsetAsync(M.d_ctl, 0, 2 * sizeof(MGCtl), c->stream));
        M.ctl_dirty = false;
    }
    Solve S;
    S.ctl = M.d_ctl + M.ctl_pick;
    S.other = M.d_ctl + (M.ctl_pick ^ 1);
    S.L0 = M.lev[0];
    const int sh = M.cc ? 0 : 1;
    // node-centred level 0 (odd n) has the guarded plane's shape: work on the caller's
    // planes directly when the two components are one plane apart
    const bool alias = !M.cc && cstride == S.L0.cs;
    if (alias) {
        S.rhs0 = rhs;
        S.rcs = cstride;
        S.L0.acf = const_cast<double*>(acf);
        S.buf[0] = S.L0.cor;
        S.buf[1] = sol;
    } else {
        const int n = (S.L0.ihi - S.L0.ilo + 1) * (S.L0.jhi - S.L0.jlo + 1);
        hipLaunchKernelGGL(k_mg_load, grid1d(n, 256), dim3(256), 0, c->stream, S.L0, sh, c->g, sol, rhs, acf, cstride,
                           M.lev[0].acf, M.msol, M.mrhs);
        S.rhs0 = M.mrhs;
        S.rcs = S.L0.cs;
        S.buf[0] = S.L0.cor;
        S.buf[1] = M.msol;
    }
    M.ctl_dirty = true;   // until the solve ends normally
    // average_down_acoef (:1641-1661): level 1 in the initial smoothing, levels 2..cb in the
    // first V-cycle's down-sweep, the coarse ones inside k_coarse
    // initial smoothing from the guess (buf[1]) into buf[0], the iterate after zero cycles;
    // k_target then sets the convergence target from its residual and rhs max norms
    smooth<4, false, true, false, true, true, true>(c, S.L0, L1, S.buf[0], S.rhs0, S.rcs, S.L0.acf, S.buf[1], nullptr,
                                                    L1.res, nullptr, M.d_norms);
    S.tol_rel = tol_rel;
    S.tol_abs = tol_abs;
    // the convergence bookkeeping rides on the first launch of the next V-cycle when there
    // is a tiled level 1 to carry it; the last cycle of every batch publishes with k_check.
    // After a solve that needed no V-cycle (an initial guess within the tolerance), the
    // target is evaluated and published first, so that no cycle is enqueued speculatively.
    const bool can_fold = M.coarse_begin > 1 && maxiter > 0;
    const bool probe = M.last_zero && maxiter > 0;
    if (!can_fold || probe) {
        const int seq = probe ? ++M.seq : 0;
        const Check chk0{S.ctl, S.other, M.d_mirror, M.d_norms, level0_blocks(S.L0), seq, maxiter, tol_rel, tol_abs};
        hipLaunchKernelGGL(k_target, dim3(1), dim3(NT), 0, c->stream, chk0);
    }
    HPQ_HIP(hipGetLastError());
    int batch = std::max(1, std::min(M.last_iters, maxiter));
    int issued = 0;
    bool checked = !can_fold || probe;   // the previous cycle's end (or the target) has been evaluated
    Published pub{};
    const bool done0 = probe && (pub = wait_mirror(c, M.seq)).done;
    while (!done0) {
        for (int b = 0; b < batch && issued < maxiter; ++b, ++issued) {
            const bool last = b + 1 == batch || issued + 1 == maxiter;
            vcycle(c, S, issued, maxiter, !checked, last || !can_fold);
            checked = last || !can_fold;
        }
        HPQ_HIP(hipGetLastError());
        if (issued == 0) break;   // maxiter == 0

        pub = wait_mirror(c, M.seq);
        if (pub.done || issued >= maxiter) break;
        batch = 1;
    }
    int it = 0, status = 0;
    if (issued || done0) {
        it = pub.iters;
        status = pub.status;
    } else {
        // no V-cycle allowed: converged only if the initial residual met the target
        MGCtl h;
        HPQ_HIP(hipMemcpyAsync(&h, S.ctl, sizeof(MGCtl), hipMemcpyDeviceToHost, c->stream));
        HPQ_HIP(hipStreamSynchronize(c->stream));
        status = h.done ? HPQ_OK : HPQ_E_MG_NOT_CONVERGED;
    }
    M.last_iters = std::max(1, it);
    M.last_zero = (issued || done0) && it == 0 && status == HPQ_OK;
    if (iters) *iters = it;
    // the result is buf[it & 1]; in the aliased layout buf[1] is sol itself
    if (!(alias && (it & 1))) {
        const int n = (S.L0.vihi - S.L0.vilo + 1) * (S.L0.vjhi - S.L0.vjlo + 1);
        hipLaunchKernelGGL(k_mg_store, grid1d(n, 256), dim3(256), 0, c->stream, S.L0, sh, c->g, sol, cstride,
                           S.buf[it & 1]);
    }
    HPQ_HIP(hipGetLastError());
    if (issued || done0) {
        M.ctl_pick ^= 1;
        M.ctl_dirty = false;
    }
    return status;
}

}  // namespace hpq
