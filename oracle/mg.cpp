// Oracle multigrid (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
//
// Restates hpmg::MultiGrid type I (`-acoef*sol + Lap(sol) = rhs`, homogeneous
// Dirichlet) as the reference runs it on CPU: src/mg_solver/HpMultiGrid.cpp
//   ctor :1043-1163 (node-centred for odd n, cell-centred for even n, coarsening rule),
//   solve1 :1170-1190, solve_doit :1308-1427, vcycle :1430-1512,
//   bottomsolve CPU branch :1583-1593, average_down_acoef :1641-1661,
//   gsrb_cached :596-740 (== four global red-black sweeps, then the residual),
//   gs1 :267-292, laplacian/residual1 :163-190, restrict/interp :31-121.
#include "oracle.hpp"
#include <algorithm>
#include <stdexcept>

namespace orc {
namespace {

struct Level {
    bool cc = true;   // cell centred
    int lo = 0, hi = 0;      // index range of the box (same in x and y extents per dim)
    int nlx = 0, nly = 0;    // box length in x and y
    int ilo = 0, ihi = 0, jlo = 0, jhi = 0;      // box bounds
    int vilo = 0, vihi = 0, vjlo = 0, vjhi = 0;  // valid loop bounds
    // arrays over the box with 1 ghost layer, 2 components (acf: 1)
    int ldx = 0, ldy = 0;
    std::vector<double> acf, res, cor, rescor;
    long at(int i, int j) const { return long(j - jlo + 1) * ldx + (i - ilo + 1); }
    long cs() const { return long(ldx) * ldy; }
    void alloc() {
        ldx = (ihi - ilo + 1) + 2; ldy = (jhi - jlo + 1) + 2;
        acf.assign(cs(), 0.0); res.assign(2 * cs(), 0.0);
        cor.assign(2 * cs(), 0.0); rescor.assign(2 * cs(), 0.0);
    }
};

// amrex::Box::coarsenable(ratio 2, min_width): size >= 2*min_width and refine(coarsen(b)) == b
bool coarsenable(bool cc, int lo, int hi, int min_width) {
    const int size = hi - lo + 1;
    if (size < 2 * min_width) return false;
    // cell: coarsen lo>>1, hi>>1 ; refine lo*2, hi*2+1. node: coarsen lo/2,hi/2 ; refine *2
    auto fl = [](int v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); };
    if (cc) return (fl(lo) * 2 == lo) && (fl(hi) * 2 + 1 == hi);
    return (fl(lo) * 2 == lo) && (fl(hi) * 2 == hi);
}

struct MG {
    std::vector<Level> lev;
    double dx = 0, dy = 0;
    bool cc = true;
    int max_level = 0;

    MG(int nx, int ny, double dx_, double dy_) : dx(dx_), dy(dy_) {
        if (nx % 2 != ny % 2) throw std::runtime_error("hpmg: nx and ny parity must match");
        cc = (nx % 2 == 0);
        Level l0;
        l0.cc = cc;
        if (cc) { l0.ilo = 0; l0.ihi = nx - 1; l0.jlo = 0; l0.jhi = ny - 1; }
        else { l0.ilo = 0; l0.ihi = nx + 1; l0.jlo = 0; l0.jhi = ny + 1; }
        lev.push_back(l0);
        const int mw = cc ? 2 : 4;
        for (int i = 0; i < 30; ++i) {
            const Level& b = lev.back();
            if (coarsenable(cc, b.ilo, b.ihi, mw) && coarsenable(cc, b.jlo, b.jhi, mw)) {
                Level c;
                c.cc = cc;
                c.ilo = b.ilo / 2; c.jlo = b.jlo / 2;
                c.ihi = b.ihi / 2;  // cell: hi>>1 ; node: hi/2 (all indices >= 0)
                c.jhi = b.jhi / 2;
                lev.push_back(c);
            } else {
                break;
            }
        }
        max_level = int(lev.size()) - 1;
        for (auto& l : lev) {
            if (cc) { l.vilo = l.ilo; l.vihi = l.ihi; l.vjlo = l.jlo; l.vjhi = l.jhi; }
            else { l.vilo = l.ilo + 1; l.vihi = l.ihi - 1; l.vjlo = l.jlo + 1; l.vjhi = l.jhi - 1; }
            l.alloc();
        }
    }
};

inline double lap(const Level& L, const double* phi, int i, int j, double facx, double facy) {
    const auto P = [&](int a, int b) { return phi[L.at(a, b)]; };
    double r = -2. * (facx + facy) * P(i, j);
    if (i == L.ilo) r += facx * (4. / 3. * P(i + 1, j) - 2. * P(i, j));
    else if (i == L.ihi) r += facx * (4. / 3. * P(i - 1, j) - 2. * P(i, j));
    else r += facx * (P(i - 1, j) + P(i + 1, j));
    if (j == L.jlo) r += facy * (4. / 3. * P(i, j + 1) - 2. * P(i, j));
    else if (j == L.jhi) r += facy * (4. / 3. * P(i, j - 1) - 2. * P(i, j));
    else r += facy * (P(i, j - 1) + P(i, j + 1));
    return r;
}

// gs1<is_cell_centered> (:267-292)
inline void gs1(const Level& L, double* phi, int i, int j, double rhs, double acf, double facx, double facy) {
    const auto P = [&](int a, int b) { return phi[L.at(a, b)]; };
    double l;
    double c0 = -(acf + 2. * (facx + facy));
    if (L.cc && i == L.ilo) { l = facx * (4. / 3.) * P(i + 1, j); c0 -= 2. * facx; }
    else if (L.cc && i == L.ihi) { l = facx * (4. / 3.) * P(i - 1, j); c0 -= 2. * facx; }
    else { l = facx * (P(i - 1, j) + P(i + 1, j)); }
    if (L.cc && j == L.jlo) { l += facy * (4. / 3.) * P(i, j + 1); c0 -= 2. * facy; }
    else if (L.cc && j == L.jhi) { l += facy * (4. / 3.) * P(i, j - 1); c0 -= 2. * facy; }
    else { l += facy * (P(i, j - 1) + P(i, j + 1)); }
    const double c0_inv = 1. / c0;
    phi[L.at(i, j)] = (rhs - l) * c0_inv;
}

// one red-black colour sweep over the valid loop box (gsrb :382-412)
void gsrb(const Level& L, int color, double* phi, const double* rhs, long rcs, const double* acf,
          double dxl, double dyl) {
    const double facx = 1. / (dxl * dxl), facy = 1. / (dyl * dyl);
    const long cs = L.cs();
#pragma omp parallel for
    for (int j = L.vjlo; j <= L.vjhi; ++j) {
        for (int i = L.vilo; i <= L.vihi; ++i) {
            if ((i + j + color) % 2 == 0) {
                const double a = acf[L.at(i, j)];
                gs1(L, phi, i, j, rhs[L.at(i, j)], a, facx, facy);
                gs1(L, phi + cs, i, j, rhs[L.at(i, j) + rcs], a, facx, facy);
            }
        }
    }
}

// gsrb_4_residual<zero_init, do_residual>, CPU semantics of gsrb_cached (:596-740):
// phi_out = 4 red-black sweeps starting from phi_in (or 0) with zero outside valid;
// res = rhs + acf*phi - Lap(phi) on the valid box.
void gsrb4(const Level& L, double* phi_out, const double* rhs, long rcs, const double* acf, double* res,
           const double* phi_in, bool zero_init, bool do_res, double dxl, double dyl) {
    const long cs = L.cs();
    std::vector<double> phi(2 * cs, 0.0);
    if (!zero_init) {
        for (int n = 0; n < 2; ++n)
            for (int j = L.vjlo; j <= L.vjhi; ++j)
                for (int i = L.vilo; i <= L.vihi; ++i) phi[n * cs + L.at(i, j)] = phi_in[n * cs + L.at(i, j)];
    }
    for (int c = 0; c < 4; ++c) gsrb(L, c, phi.data(), rhs, rcs, acf, dxl, dyl);
    const double facx = 1. / (dxl * dxl), facy = 1. / (dyl * dyl);
#pragma omp parallel for
    for (int j = L.vjlo; j <= L.vjhi; ++j)
        for (int i = L.vilo; i <= L.vihi; ++i) {
            const long k = L.at(i, j);
            if (do_res) {
                for (int n = 0; n < 2; ++n)
                    res[n * cs + k] = rhs[k + n * rcs] + acf[k] * phi[n * cs + k] -
                                      lap(L, phi.data() + n * cs, i, j, facx, facy);
            }
            for (int n = 0; n < 2; ++n) phi_out[n * cs + k] = phi[n * cs + k];
        }
}

void restriction(const Level& C, const Level& Fl, double* crse, const double* fine, int ncomp) {
    const long ccs = C.cs(), fcs = Fl.cs();
    for (int n = 0; n < ncomp; ++n)
        for (int j = C.vjlo; j <= C.vjhi; ++j)
            for (int i = C.vilo; i <= C.vihi; ++i) {
                const auto f = [&](int a, int b) { return fine[n * fcs + Fl.at(a, b)]; };
                double v;
                if (C.cc) {
                    v = 0.25 * (f(2 * i, 2 * j) + f(2 * i + 1, 2 * j) + f(2 * i, 2 * j + 1) + f(2 * i + 1, 2 * j + 1));
                } else {
                    v = (1. / 16.) * (f(2 * i - 1, 2 * j - 1) + 2. * f(2 * i, 2 * j - 1) + f(2 * i + 1, 2 * j - 1) +
                                      2. * f(2 * i - 1, 2 * j) + 4. * f(2 * i, 2 * j) + 2. * f(2 * i + 1, 2 * j) +
                                      f(2 * i - 1, 2 * j + 1) + 2. * f(2 * i, 2 * j + 1) + f(2 * i + 1, 2 * j + 1));
                }
                crse[n * ccs + C.at(i, j)] = v;
            }
}

// interpolation_outofplace (:134-150): fine_out = fine_in + I(crse)
void interp(const Level& Fl, const Level& C, const double* fine_in, const double* crse, double* fine_out) {
    const long ccs = C.cs(), fcs = Fl.cs();
    for (int n = 0; n < 2; ++n)
        for (int j = Fl.vjlo; j <= Fl.vjhi; ++j)
            for (int i = Fl.vilo; i <= Fl.vihi; ++i) {
                const auto c = [&](int a, int b) { return crse[n * ccs + C.at(a, b)]; };
                const long k = n * fcs + Fl.at(i, j);
                const int ic = (i >= 0) ? i / 2 : -((-i + 1) / 2);
                const int jc = (j >= 0) ? j / 2 : -((-j + 1) / 2);
                if (Fl.cc) {
                    fine_out[k] = fine_in[k] + c(ic, jc);
                } else {
                    const bool io = (ic * 2 != i), jo = (jc * 2 != j);
                    if (io && jo) fine_out[k] = fine_in[k] + (c(ic, jc) + c(ic + 1, jc) + c(ic, jc + 1) + c(ic + 1, jc + 1)) * 0.25;
                    else if (io) fine_out[k] = fine_in[k] + (c(ic, jc) + c(ic + 1, jc)) * 0.5;
                    else if (jo) fine_out[k] = fine_in[k] + (c(ic, jc) + c(ic, jc + 1)) * 0.5;
                    else fine_out[k] = fine_in[k] + c(ic, jc);
                }
            }
}

}  // namespace

// solve1: sol/rhs are 2 planes of the guarded slice store (comp_stride apart), acf is
// the chi plane. Node-centred MG node k <-> slice cell k-1; cell-centred: same index.
// Returns 0 on convergence, 1 if not converged within maxiter, 2 on divergence.
int mg_solve1(const Geom& g, double* sol, const double* rhs, const double* acf, long comp_stride,
              double tol_rel, double tol_abs, int maxiter, int* iters) {
    MG mg(g.nx, g.ny, g.dx, g.dy);
    Level& L0 = mg.lev[0];
    const int sh = mg.cc ? 0 : 1;  // MG index = slice index + sh
    const long cs0 = L0.cs();
    // copy acf, sol (into a level-0-shaped buffer), rhs
    std::vector<double> msol(2 * cs0, 0.0), mrhs(2 * cs0, 0.0);
    for (int j = L0.jlo; j <= L0.jhi; ++j)
        for (int i = L0.ilo; i <= L0.ihi; ++i) {
            const long gk = g.idx(i - sh, j - sh);
            L0.acf[L0.at(i, j)] = acf[gk];
            for (int n = 0; n < 2; ++n) {
                msol[n * cs0 + L0.at(i, j)] = sol[gk + n * comp_stride];
                mrhs[n * cs0 + L0.at(i, j)] = rhs[gk + n * comp_stride];
            }
        }
    // average_down_acoef: levels 1..max_level (single_block_level_begin = max_level on CPU)
    for (int l = 1; l <= mg.max_level; ++l) restriction(mg.lev[l], mg.lev[l - 1], mg.lev[l].acf.data(), mg.lev[l - 1].acf.data(), 1);

    auto dxl = [&](int l) { return mg.dx * double(1 << l); };
    auto dyl = [&](int l) { return mg.dy * double(1 << l); };

    gsrb4(L0, L0.cor.data(), mrhs.data(), cs0, L0.acf.data(), L0.rescor.data(), msol.data(), false, true, mg.dx, mg.dy);
    double resnorm0 = 0, rhsnorm0 = 0;
    for (int n = 0; n < 2; ++n)
        for (int j = L0.vjlo; j <= L0.vjhi; ++j)
            for (int i = L0.vilo; i <= L0.vihi; ++i) {
                resnorm0 = std::max(resnorm0, std::abs(L0.rescor[n * cs0 + L0.at(i, j)]));
                rhsnorm0 = std::max(rhsnorm0, std::abs(mrhs[n * cs0 + L0.at(i, j)]));
            }
    const double max_norm = (rhsnorm0 >= resnorm0) ? rhsnorm0 : resnorm0;
    const double res_target = std::max(tol_abs, std::max(tol_rel, 1.e-16) * max_norm);
    int status = 0, it = 0;
    if (!(resnorm0 <= res_target)) {
        bool converged = false;
        for (it = 0; it < maxiter; ++it) {
            // vcycle
            for (int l = 0; l < mg.max_level; ++l) {
                Level& Ll = mg.lev[l];
                if (l > 0) gsrb4(Ll, Ll.cor.data(), Ll.res.data(), Ll.cs(), Ll.acf.data(), Ll.rescor.data(), nullptr, true, true, dxl(l), dyl(l));
                restriction(mg.lev[l + 1], Ll, mg.lev[l + 1].res.data(), Ll.rescor.data(), 2);
            }
            {  // bottomsolve, CPU branch
                Level& Lb = mg.lev[mg.max_level];
                std::fill(Lb.cor.begin(), Lb.cor.end(), 0.0);
                const int len = std::max(Lb.ihi - Lb.ilo + 1, Lb.jhi - Lb.jlo + 1);
                const int nsw = std::max(16, (len + 1) / 2 * 2);
                for (int s = 0; s < nsw; ++s) gsrb(Lb, s, Lb.cor.data(), Lb.res.data(), Lb.cs(), Lb.acf.data(), dxl(mg.max_level), dyl(mg.max_level));
            }
            for (int l = mg.max_level - 1; l >= 0; --l) {
                Level& Ll = mg.lev[l];
                interp(Ll, mg.lev[l + 1], Ll.cor.data(), mg.lev[l + 1].cor.data(), Ll.rescor.data());
                if (l == 0) gsrb4(Ll, msol.data(), mrhs.data(), cs0, Ll.acf.data(), nullptr, Ll.rescor.data(), false, false, dxl(0), dyl(0));
                else gsrb4(Ll, Ll.cor.data(), Ll.res.data(), Ll.cs(), Ll.acf.data(), nullptr, Ll.rescor.data(), false, false, dxl(l), dyl(l));
            }
            gsrb4(L0, L0.cor.data(), mrhs.data(), cs0, L0.acf.data(), L0.rescor.data(), msol.data(), false, true, mg.dx, mg.dy);
            double norminf = 0;
            for (int n = 0; n < 2; ++n)
                for (int j = L0.jlo; j <= L0.jhi; ++j)
                    for (int i = L0.ilo; i <= L0.ihi; ++i) norminf = std::max(norminf, std::abs(L0.rescor[n * cs0 + L0.at(i, j)]));
            if (norminf <= res_target) { converged = true; ++it; break; }
            if (norminf > 1.e20 * max_norm) { status = 2; ++it; break; }
        }
        if (!converged && status == 0) status = 1;
    }
    if (iters) *iters = it;
    for (int n = 0; n < 2; ++n)
        for (int j = L0.vjlo; j <= L0.vjhi; ++j)
            for (int i = L0.vilo; i <= L0.vihi; ++i) sol[g.idx(i - sh, j - sh) + n * comp_stride] = L0.cor[n * cs0 + L0.at(i, j)];
    return status;
}

}  // namespace orc
