// Oracle slice loop and C API (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
#include "oracle.hpp"
#include <algorithm>
#include <cstring>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

struct PlasmaProfile {
    int ppc[2] = {1, 1};
    double density = 1.0;   // n0
    double ramp_a = 0.0;    // n(x,y) = n0 (1 + ramp_a (x^2+y^2))
    double radius = INFINITY;
    double u_mean[3] = {0, 0, 0};
    // the density at every lattice position, [(j*nx + i)*nppc + i_part], or empty
    std::vector<double> table;
};

struct Sim {
    Ctx c;
    std::vector<double> F;
    Plasma plasma;
    Beam beam;
    bool has_plasma = true;
    PlasmaProfile prof;
    std::vector<double> checks;              // sum |F| per component
    std::vector<std::pair<int, double*>> captures;  // comp -> nz*nx*ny buffer
    long n_qsa = 0;
    long mg_iters = 0;
    int mg_status = 0;
    long pc_iters = 0;
    double pushes = 0;
    // SALAME settings and state (Hipace::m_salame_*, src/Hipace.H:284-301)
    bool salame_on = false;
    int salame_n_iter = 5;
    bool salame_do_advance = true;
    double salame_tol = 1e-4;
    double (*salame_target)(double, double, double, void*) = nullptr;
    void* salame_user = nullptr;
    double* salame_weight_factor = nullptr;
    int salame_last_slice = -1;
    bool salame_overloaded = false;
    double salame_zeta_initial = 0.0;
};

static void set_plane(Sim& s, int comp, double v) {
    if (comp < 0) return;
    std::fill(s.F.begin() + comp * s.c.g.plane, s.F.begin() + (comp + 1) * s.c.g.plane, v);
}
static void copy_plane(Sim& s, int dst, int src) {
    std::copy(s.F.begin() + src * s.c.g.plane, s.F.begin() + (src + 1) * s.c.g.plane, s.F.begin() + dst * s.c.g.plane);
}
static void add_plane(Sim& s, int dst, int src) {
    double* d = s.F.data() + dst * s.c.g.plane;
    const double* a = s.F.data() + src * s.c.g.plane;
    for (long k = 0; k < s.c.g.plane; ++k) d[k] += a[k];
}

// PlasmaParticleContainer::InitParticles, src/particles/plasma/PlasmaParticleContainerInit.cpp:18-378,
// without fine patch / symmetrisation / hollow core; u_std = 0. The lattice is injected with
// i_part outer and cells inner (:189-192); the CPU build then sorts the particles by cell,
// x fastest, keeping the i_part order inside a cell (SortParticlesByCell, :372-377, a stable
// amrex::DenseBins sort). That is the order here: cells (x fastest), then i_part.
static void init_plasma(Sim& s) {
    const Geom& g = s.c.g;
    const PlasmaProfile& pr = s.prof;
    const int nppc = pr.ppc[0] * pr.ppc[1];
    const double scale_fac = nppc <= 0 ? 0. : 1. / nppc;
    const double* blo = s.c.o.plo;
    const double* bhi = s.c.o.phi;
    std::vector<double> X, Y, W;
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i)
            for (int ipart = 0; ipart < nppc; ++ipart) {
                const int ixp = ipart % pr.ppc[0], iyp = ipart / pr.ppc[0];
                const double r0 = (0.5 + ixp) / pr.ppc[0], r1 = (0.5 + iyp) / pr.ppc[1];
                const double x = g.xlo + (i + r0 + 0.0) * g.dx;
                const double y = g.ylo + (j + r1 + 0.0) * g.dy;
                const double rsq = x * x + y * y;
                const double dens = pr.table.empty() ? pr.density * (1.0 + pr.ramp_a * rsq)
                                                     : pr.table[(size_t(j) * g.nx + i) * nppc + ipart];
                if (x >= bhi[0] || x < blo[0] || y >= bhi[1] || y < blo[1] || rsq > pr.radius * pr.radius ||
                    dens <= 0.0)
                    continue;
                X.push_back(x); Y.push_back(y); W.push_back(dens * scale_fac);
            }
    Plasma& p = s.plasma;
    p.resize(X.size());
    const double u0 = pr.u_mean[0], u1 = pr.u_mean[1], u2 = pr.u_mean[2];
    const double psi0 = std::sqrt(1. + u0 * u0 + u1 * u1 + u2 * u2) - u2;
    for (size_t k = 0; k < X.size(); ++k) {
        p.x[k] = X[k]; p.y[k] = Y[k]; p.w[k] = W[k];
        p.ux[k] = u0 * 1.0; p.uy[k] = u1 * 1.0; p.psi[k] = psi0;
        p.x_prev[k] = X[k]; p.y_prev[k] = Y[k];
        p.ux_h[k] = u0 * 1.0; p.uy_h[k] = u1 * 1.0; p.psi_h[k] = psi0;
        p.valid[k] = 1;
    }
}

static void accumulate_diag(Sim& s, int islice) {
    const Geom& g = s.c.g;
    for (int comp = 0; comp < s.c.L.ncomp; ++comp) {
        const double* P = s.F.data() + comp * g.plane;
        double acc = 0;
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) acc += std::abs(P[g.idx(i, j)]);
        s.checks[comp] += acc;
    }
    for (auto& cap : s.captures) {
        const double* P = s.F.data() + cap.first * g.plane;
        double* dst = cap.second + long(islice) * g.nx * g.ny;
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) dst[long(j) * g.nx + i] = P[g.idx(i, j)];
    }
}

// Sum of 64 lane values in the order of a xor butterfly (offsets 32, 16, ..., 1): every lane
// ends with the same value; this is lane 0's.
static double butterfly64(double v[64]) {
    for (int off = 32; off > 0; off >>= 1) {
        double t[64];
        for (int l = 0; l < 64; ++l) t[l] = v[l] + v[l ^ off];
        for (int l = 0; l < 64; ++l) v[l] = t[l];
    }
    return v[0];
}

// Fields::ComputeRelBFieldError, src/fields/Fields.cpp:1229-1286: the two sums over valid
// cells in a fixed order (the reference's ReduceOps order depends on its thread count): per
// row, lane l of 64 adds the cells i = l, l + 64, ... in turn, the lanes are combined by
// butterfly64; the row sums are combined the same way, lane l adding rows l, l + 64, ...
static double rel_b_error(Sim& s, int bx, int by, int bxi, int byi) {
    const Geom& g = s.c.g;
    const double* Bx = s.F.data() + bx * g.plane; const double* By = s.F.data() + by * g.plane;
    const double* Bxi = s.F.data() + bxi * g.plane; const double* Byi = s.F.data() + byi * g.plane;
    const double factor = g.dx * g.dy / (g.dx * g.dy);
    std::vector<double> rb(g.ny), rd(g.ny);
    for (int j = 0; j < g.ny; ++j) {
        double a[64], b[64];
        for (int l = 0; l < 64; ++l) {
            a[l] = 0.0;
            b[l] = 0.0;
            for (int i = l; i < g.nx; i += 64) {
                const long k = g.idx(i, j);
                a[l] += factor * std::sqrt(Bx[k] * Bx[k] + By[k] * By[k]);
                b[l] += factor * std::sqrt((Bx[k] - Bxi[k]) * (Bx[k] - Bxi[k]) + (By[k] - Byi[k]) * (By[k] - Byi[k]));
            }
        }
        rb[j] = butterfly64(a);
        rd[j] = butterfly64(b);
    }
    double a[64], b[64];
    for (int l = 0; l < 64; ++l) {
        a[l] = 0.0;
        b[l] = 0.0;
        for (int j = l; j < g.ny; j += 64) {
            a[l] += rb[j];
            b[l] += rd[j];
        }
    }
    const double nB = butterfly64(a), nD = butterfly64(b);
    return (nB > 0.) ? nD / nB : 0.;
}

static void lincomb(Sim& s, int dst, double a, int sa, double b, int sb) {
    double* d = s.F.data() + dst * s.c.g.plane;
    const double* A = s.F.data() + sa * s.c.g.plane;
    const double* B = s.F.data() + sb * s.c.g.plane;
    for (long k = 0; k < s.c.g.plane; ++k) d[k] = a * A[k] + b * B[k];
}

// Hipace::PredictorCorrectorLoopToSolveBxBy, src/Hipace.cpp:936-1031, with
// Fields::InitialBfieldGuess (:1150-1171) and MixAndShiftBfields (:1174-1226).
static void pc_loop(Sim& s, int islice) {
    const Layout& L = s.c.L;
    const Options& o = s.c.o;
    double err_prev = 1.0;
    double err = rel_b_error(s, L.prev_Bx, L.prev_By, L.pcprev_Bx, L.pcprev_By);
    const double m = std::exp(-0.5 * std::pow(err / (2.5 * o.pc_tol), 2));
    lincomb(s, L.Bx, 1. + m, L.prev_Bx, -m, L.pcprev_Bx);
    lincomb(s, L.By, 1. + m, L.prev_By, -m, L.pcprev_By);
    set_plane(s, L.pciter_Bx, 0.); set_plane(s, L.pciter_By, 0.);
    copy_plane(s, L.pcprev_Bx, L.Bx); copy_plane(s, L.pcprev_By, L.By);
    int it = 0;
    err = 1.0;
    while (err > o.pc_tol && it < o.pc_maxiter) {
        ++it;
        if (s.has_plasma) {
            push_plasma(s.c, s.plasma, s.F.data(), true, L.Psi, L.Ez, L.Bx, L.By, L.Bz);
            deposit_plasma(s.c, s.plasma, s.F.data(), L.next_jx, L.next_jy, -1, -1, -1, -1, false, nullptr);
        }
        if (o.do_beam_jx_jy) deposit_beam(s.c, s.beam, islice - 1, s.F.data(), L.next_jx, L.next_jy, -1);
        solve_poisson_bxby(s.c, s.F.data());
        err = rel_b_error(s, L.Bx, L.By, L.pciter_Bx, L.pciter_By);
        if (it == 1) err_prev = err;
        double wi, wp;
        if (err != 0. || err_prev != 0.) { wi = err_prev / (err + err_prev); wp = err / (err + err_prev); }
        else { wi = 0.5; wp = 0.5; }
        lincomb(s, L.pcprev_Bx, wi, L.pciter_Bx, wp, L.pcprev_Bx);
        lincomb(s, L.pcprev_By, wi, L.pciter_By, wp, L.pcprev_By);
        lincomb(s, L.Bx, 1. - o.pc_mix, L.Bx, o.pc_mix, L.pcprev_Bx);
        lincomb(s, L.By, 1. - o.pc_mix, L.By, o.pc_mix, L.pcprev_By);
        copy_plane(s, L.pcprev_Bx, L.pciter_Bx); copy_plane(s, L.pcprev_By, L.pciter_By);
        set_plane(s, L.next_jx, 0.); set_plane(s, L.next_jy, 0.);
        err_prev = err;
    }
    s.pc_iters += it;
}

static void mg_solve_bxby(Sim& s, int bx, int sy) {
    const Options& o = s.c.o;
    double* F = s.F.data();
    int it = 0;
    const int st = mg_solve1(s.c.g, F + bx * s.c.g.plane, F + sy * s.c.g.plane, F + s.c.L.chi * s.c.g.plane,
                             s.c.g.plane, o.mg_tol_rel, o.mg_tol_abs, o.mg_maxiter, &it);
    s.mg_iters += it;
    if (st) s.mg_status = st;
}

// SalameModule, src/salame/Salame.cpp:14-184 (level 0, no grid current), with
// SalameInitializeSxSyWithBeam (:187-219), SalameGetJxJyFromBxBy (:222-256),
// SalameGetW (:332-407) and SalameMultiplyBeamWeight (:409-437)
static void salame_module(Sim& s, int islice) {
    const Layout& L = s.c.L;
    const Geom& g = s.c.g;
    double* F = s.F.data();
    Beam& b = s.beam;
    const int32_t* mask = b.salame.data();
    long nq = 0;
    if (islice + 1 != s.salame_last_slice) {
        copy_plane(s, L.sal_Ez_target, L.Ez);
        s.salame_overloaded = false;
        s.salame_zeta_initial = islice * g.dz + g.zoff;
    }
    s.salame_last_slice = islice;
    set_plane(s, L.Sy, 0.); set_plane(s, L.Sx, 0.);
    if (s.has_plasma) explicit_deposit(s.c, s.plasma, F);
    copy_plane(s, L.sal_Sy_back, L.Sy); copy_plane(s, L.sal_Sx_back, L.Sx);
    double W = 1.0;
    for (int iter = 0; iter < s.salame_n_iter; ++iter) {
        // STEP 1
        if (s.has_plasma) push_plasma(s.c, s.plasma, F, true, L.Psi, L.Ez, L.Bx, L.By, L.Bz);
        copy_plane(s, L.sal_jx, L.next_jx_beam); copy_plane(s, L.sal_jy, L.next_jy_beam);
        if (s.has_plasma) deposit_plasma(s.c, s.plasma, F, L.sal_jx, L.sal_jy, -1, -1, -1, -1, false, &nq);
        for (int comp : {L.sal_Ez, L.sal_jz_beam, L.sal_Sy, L.sal_Sx, L.sal_Bx, L.sal_By}) set_plane(s, comp, 0.);
        solve_poisson_ez(s.c, F, L.sal_jx, L.sal_jy, L.sal_Ez);
        copy_plane(s, L.sal_Ez_no_salame, L.sal_Ez);
        // STEP 2
        deposit_beam(s.c, b, islice, F, -1, -1, L.sal_jz_beam, mask, 1);
        {
            const double mu0 = 1.0;
            const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
            const double* jzb = F + L.sal_jz_beam * g.plane;
            double* Sy = F + L.sal_Sy * g.plane;
            double* Sx = F + L.sal_Sx * g.plane;
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const double dx_jzb = (jzb[g.idx(i + 1, j)] - jzb[g.idx(i - 1, j)]) * dxih;
                    const double dy_jzb = (jzb[g.idx(i, j + 1)] - jzb[g.idx(i, j - 1)]) * dyih;
                    Sy[g.idx(i, j)] = mu0 * (-dy_jzb);
                    Sx[g.idx(i, j)] = -mu0 * (-dx_jzb);
                }
        }
        mg_solve_bxby(s, L.sal_Bx, L.sal_Sy);
        for (int comp : {L.sal_Ez, L.sal_jx, L.sal_jy}) set_plane(s, comp, 0.);
        if (s.salame_do_advance) {
            if (s.has_plasma) {
                salame_only_advance(s.c, s.plasma, F);
                deposit_plasma(s.c, s.plasma, F, L.sal_jx, L.sal_jy, -1, -1, -1, -1, false, &nq);
            }
        } else {
            const double dz = 1.5 * g.dz;
            const double mu0_inv = 1.0 / 1.0;
            const double* chi = F + L.chi * g.plane;
            const double* Bx = F + L.sal_Bx * g.plane;
            const double* By = F + L.sal_By * g.plane;
            double* jx = F + L.sal_jx * g.plane;
            double* jy = F + L.sal_jy * g.plane;
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const long k = g.idx(i, j);
                    jx[k] = dz * chi[k] * By[k] * mu0_inv;
                    jy[k] = -dz * chi[k] * Bx[k] * mu0_inv;
                }
        }
        solve_poisson_ez(s.c, F, L.sal_jx, L.sal_jy, L.sal_Ez);
        // STEP 3
        set_plane(s, L.sal_jz_beam, 0.);
        deposit_beam(s.c, b, islice, F, -1, -1, L.sal_jz_beam, mask, 1);
        double sum_Ez_target = 0, sum_Ez_no_salame = 0, sum_Ez_only_salame = 0, sum_jz = 0;
        {
            const double factor = g.dx * g.dy / (g.dx * g.dy);
            const double* jz = F + L.sal_jz_beam * g.plane;
            const double* et = F + L.sal_Ez_target * g.plane;
            const double* en = F + L.sal_Ez_no_salame * g.plane;
            const double* ez = F + L.sal_Ez * g.plane;
            for (int j = 0; j < g.ny; ++j)
                for (int i = 0; i < g.nx; ++i) {
                    const long k = g.idx(i, j);
                    sum_Ez_target += factor * jz[k] * et[k];
                    sum_Ez_no_salame += factor * jz[k] * en[k];
                    sum_Ez_only_salame += factor * jz[k] * ez[k];
                    sum_jz += factor * jz[k];
                }
        }
        sum_Ez_target /= sum_jz;
        sum_Ez_no_salame /= sum_jz;
        sum_Ez_only_salame /= sum_jz;
        const double zeta = (islice - 1) * g.dz + g.zoff;
        if (s.salame_target) sum_Ez_target = s.salame_target(zeta, s.salame_zeta_initial, sum_Ez_target, s.salame_user);
        W = (sum_Ez_target - sum_Ez_no_salame) / sum_Ez_only_salame + 1.0;
        if (W < 0 || s.salame_overloaded) {
            W = 0;
            iter = s.salame_n_iter - 1;
            s.salame_overloaded = true;
        }
        if (!s.salame_overloaded && iter >= 1 && std::abs(W - 1.) < s.salame_tol) iter = s.salame_n_iter - 1;
        for (int64_t ip = b.off[islice]; ip < b.off[islice + 1]; ++ip) {
            if (mask[ip] != 1 || !b.valid[ip]) continue;
            if (W == 0) {
                b.valid[ip] = 0;
                continue;
            }
            b.w[ip] *= W;
        }
        // STEP 4
        for (int comp : {L.jz_beam, L.Sy, L.Sx}) set_plane(s, comp, 0.);
        deposit_beam(s.c, b, islice, F, -1, -1, L.jz_beam);
        initialize_sxsy_with_beam(s.c, F);
        add_plane(s, L.Sy, L.sal_Sy_back); add_plane(s, L.Sx, L.sal_Sx_back);
        mg_solve_bxby(s, L.Bx, L.Sy);
    }
    s.n_qsa += nq;
    if (s.salame_weight_factor) s.salame_weight_factor[islice] = W;
}

// Hipace::SolveOneSlice, src/Hipace.cpp:557-728 (level 0, no laser/ionisation).
static void solve_one_slice(Sim& s, int islice) {
    const Layout& L = s.c.L;
    const Options& o = s.c.o;
    double* F = s.F.data();
    long nq = 0;
    if (o.explicit_solver) {
        // Fields::InitializeSlices, Fields.cpp:536-586
        for (int comp : {L.chi, L.Sy, L.Sx, L.ExmBy, L.EypBx, L.jz_beam, L.rhomjz, L.next_jx_beam, L.next_jy_beam, L.rho})
            set_plane(s, comp, 0.);
        if (s.has_plasma)
            deposit_plasma(s.c, s.plasma, F, L.jx, L.jy, -1, L.rho, L.chi, L.rhomjz, false, &nq);
        deposit_beam(s.c, s.beam, islice, F, -1, -1, L.jz_beam);
    } else {
        for (int comp : {L.ExmBy, L.EypBx, L.jx, L.jy, L.jz, L.rhomjz, L.rho}) set_plane(s, comp, 0.);
        if (s.has_plasma)
            deposit_plasma(s.c, s.plasma, F, L.jx, L.jy, L.jz, L.rho, -1, L.rhomjz, false, &nq);
        deposit_beam(s.c, s.beam, islice, F, o.do_beam_jx_jy ? L.jx : -1, o.do_beam_jx_jy ? L.jy : -1, L.jz);
    }
    s.n_qsa += nq;
    if (L.ions_rhomjz >= 0) {  // Fields::AddRhoIons, Fields.cpp:607-615
        add_plane(s, L.rhomjz, L.ions_rhomjz);
        if (L.rho >= 0) add_plane(s, L.rho, L.ions_rhomjz);
    }
    solve_poisson_psi_ez_bz(s.c, F);
    if (o.explicit_solver) {
        // SALAME beams deposit no jx, jy while SALAME runs (src/particles/beam/MultiBeam.cpp:50-56)
        const int32_t* sal = s.salame_on ? s.beam.salame.data() : nullptr;
        if (o.do_beam_jx_jy) deposit_beam(s.c, s.beam, islice - 1, F, L.next_jx_beam, L.next_jy_beam, -1, sal, 0);
        initialize_sxsy_with_beam(s.c, F);
        if (s.has_plasma) explicit_deposit(s.c, s.plasma, F);
        mg_solve_bxby(s, L.Bx, L.Sy);
        // MultiBeam::isSalameNow (MultiBeam.cpp:128-140): particles of a SALAME beam on this slice
        if (s.salame_on) {
            long n = 0;
            for (int64_t ip = s.beam.off[islice]; ip < s.beam.off[islice + 1]; ++ip) n += sal[ip] == 1;
            if (n > 0) salame_module(s, islice);
        }
    } else {
        pc_loop(s, islice);
    }
    accumulate_diag(s, islice);
    if (s.has_plasma) {
        push_plasma(s.c, s.plasma, F, false, L.Psi, L.Ez, L.Bx, L.By, L.Bz);
        s.pushes += double(s.plasma.size());
    }
    // Fields::ShiftSlices, Fields.cpp:589-604
    if (o.explicit_solver) {
        copy_plane(s, L.prev_jx_beam, L.jx_beam); copy_plane(s, L.prev_jy_beam, L.jy_beam);
        copy_plane(s, L.jx_beam, L.next_jx_beam); copy_plane(s, L.jy_beam, L.next_jy_beam);
        copy_plane(s, L.jx, L.next_jx_beam); copy_plane(s, L.jy, L.next_jy_beam);
    } else {
        copy_plane(s, L.pcprev_Bx, L.prev_Bx); copy_plane(s, L.pcprev_By, L.prev_By);
        copy_plane(s, L.prev_Bx, L.Bx); copy_plane(s, L.prev_By, L.By);
        copy_plane(s, L.prev_jx, L.jx); copy_plane(s, L.prev_jy, L.jy);
    }
}

// One time step of Hipace::Evolve (src/Hipace.cpp:394-480) with dt = 0: reset, plasma
// init, neutralising background, slice loop from the head (nz-1) to the tail (0).
static void begin_step(Sim& s) {
    std::fill(s.F.begin(), s.F.end(), 0.0);
    std::fill(s.checks.begin(), s.checks.end(), 0.0);
    s.n_qsa = 0; s.mg_iters = 0; s.mg_status = 0; s.pc_iters = 0; s.pushes = 0;
    if (s.has_plasma) {
        init_plasma(s);
        if (s.plasma.neutralize && s.c.L.ions_rhomjz >= 0)
            deposit_plasma(s.c, s.plasma, s.F.data(), -1, -1, -1, -1, -1, s.c.L.ions_rhomjz, true, nullptr);
    }
}

}  // namespace orc

using namespace orc;

extern "C" {

struct orc_config {
    int nx, ny, nz;
    double xlo, xhi, ylo, yhi, zlo, zhi;
    int depos_order, derivative_type, explicit_solver, deposit_rho, do_beam_jx_jy;
    double mg_tol_rel, mg_tol_abs;
    int mg_maxiter;
    int particle_bc;
    double plo[2], phi[2];
    double pc_tol;
    int pc_maxiter;
    double pc_mix;
    int has_plasma;
    int ppc[2];
    double density, ramp_a, plasma_radius;
    double u_mean[3];
    double plasma_charge, plasma_mass, max_qsa;
    int neutralize;
    int nthreads;
    int salame;
    int tile_size;
};

static Geom geom_from(const orc_config* cf) {
    Geom g;
    g.nx = cf->nx; g.ny = cf->ny; g.nz = cf->nz; g.ng = 2;
    g.xlo = cf->xlo; g.xhi = cf->xhi; g.ylo = cf->ylo; g.yhi = cf->yhi; g.zlo = cf->zlo; g.zhi = cf->zhi;
    g.finalize();
    return g;
}

static Ctx ctx_from(const orc_config* cf) {
    Ctx c;
    c.g = geom_from(cf);
    Options& o = c.o;
    o.depos_order = cf->depos_order; o.derivative_type = cf->derivative_type;
    o.explicit_solver = cf->explicit_solver != 0; o.deposit_rho = cf->deposit_rho != 0;
    o.do_beam_jx_jy = cf->do_beam_jx_jy != 0;
    o.mg_tol_rel = cf->mg_tol_rel; o.mg_tol_abs = cf->mg_tol_abs; o.mg_maxiter = cf->mg_maxiter;
    o.particle_bc = cf->particle_bc;
    o.plo[0] = cf->plo[0]; o.plo[1] = cf->plo[1]; o.phi[0] = cf->phi[0]; o.phi[1] = cf->phi[1];
    o.pc_tol = cf->pc_tol; o.pc_maxiter = cf->pc_maxiter; o.pc_mix = cf->pc_mix;
    o.nthreads = cf->nthreads;
    o.tile_size = cf->tile_size > 0 ? cf->tile_size : 32;
    c.L = make_layout(o.explicit_solver, o.deposit_rho, cf->has_plasma && cf->neutralize, cf->salame != 0);
    return c;
}

void* orc_sim_create(const orc_config* cf) {
    Sim* s = new Sim();
    s->c = ctx_from(cf);
    s->F.assign(size_t(s->c.L.ncomp) * s->c.g.plane, 0.0);
    s->checks.assign(s->c.L.ncomp, 0.0);
    s->has_plasma = cf->has_plasma != 0;
    s->prof.ppc[0] = cf->ppc[0]; s->prof.ppc[1] = cf->ppc[1];
    s->prof.density = cf->density; s->prof.ramp_a = cf->ramp_a; s->prof.radius = cf->plasma_radius;
    for (int k = 0; k < 3; ++k) s->prof.u_mean[k] = cf->u_mean[k];
    s->plasma.charge = cf->plasma_charge; s->plasma.mass = cf->plasma_mass; s->plasma.max_qsa = cf->max_qsa;
    s->plasma.neutralize = cf->neutralize != 0;
    s->beam.off.assign(s->c.g.nz + 1, 0);
#ifdef _OPENMP
    if (cf->nthreads > 0) omp_set_num_threads(cf->nthreads);
#endif
    return s;
}

void orc_sim_destroy(void* h) { delete static_cast<Sim*>(h); }

// <plasma>.density(x,y,z) at every lattice position (nx*ny*nppc doubles, cell (i, j) particle
// i_part at (j*nx + i)*nppc + i_part); NULL: the analytic profile
void orc_sim_set_density_table(void* h, const double* table) {
    Sim* s = static_cast<Sim*>(h);
    const Geom& g = s->c.g;
    const size_t n = size_t(g.nx) * g.ny * s->prof.ppc[0] * s->prof.ppc[1];
    if (table) s->prof.table.assign(table, table + n);
    else s->prof.table.clear();
}

int orc_sim_ncomp(void* h) { return static_cast<Sim*>(h)->c.L.ncomp; }

void orc_sim_layout(void* h, int* out /* 44 ints */) {
    const Layout& L = static_cast<Sim*>(h)->c.L;
    const int v[] = {L.ncomp, L.next_jx_beam, L.next_jy_beam, L.chi, L.Sy, L.Sx, L.ExmBy, L.EypBx, L.Ez, L.Bx,
                     L.By, L.Bz, L.Psi, L.jx_beam, L.jy_beam, L.jz_beam, L.jx, L.jy, L.rhomjz, L.rho,
                     L.prev_jx_beam, L.prev_jy_beam, L.ions_rhomjz, L.next_jx, L.next_jy, L.jz, L.prev_Bx,
                     L.prev_By, L.prev_jx, L.prev_jy, L.pciter_Bx, L.pciter_By, L.sal_Ez_target,
                     L.sal_Ez_no_salame, L.sal_Ez, L.sal_jx, L.sal_jy, L.sal_jz_beam, L.sal_Bx, L.sal_By,
                     L.sal_Sy, L.sal_Sx, L.sal_Sy_back, L.sal_Sx_back};
    std::memcpy(out, v, sizeof(v));
}

void orc_sim_set_beam(void* h, int64_t n, const double* x, const double* y, const double* z, const double* ux,
                      const double* uy, const double* uz, const double* w, const int32_t* valid,
                      const int64_t* off, double charge, double mass) {
    Sim* s = static_cast<Sim*>(h);
    Beam& b = s->beam;
    b.x.assign(x, x + n); b.y.assign(y, y + n); b.z.assign(z, z + n);
    b.ux.assign(ux, ux + n); b.uy.assign(uy, uy + n); b.uz.assign(uz, uz + n);
    b.w.assign(w, w + n); b.valid.assign(valid, valid + n);
    b.off.assign(off, off + s->c.g.nz + 1);
    b.charge = charge; b.mass = mass;
    b.ext.clear();
}

// SALAME on the following slices (mask: n per-particle flags of the present beam; NULL: off)
void orc_sim_set_salame(void* h, const int32_t* mask, int n_iter, int do_advance, double tol,
                        double (*target)(double, double, double, void*), void* user, double* weight_factor) {
    Sim* s = static_cast<Sim*>(h);
    s->salame_last_slice = -1;
    s->salame_overloaded = false;
    s->salame_on = mask != nullptr;
    if (!mask) return;
    s->beam.salame.assign(mask, mask + s->beam.x.size());
    s->salame_n_iter = n_iter;
    s->salame_do_advance = do_advance != 0;
    s->salame_tol = tol;
    s->salame_target = target;
    s->salame_user = user;
    s->salame_weight_factor = weight_factor;
}

// beams.external_E / external_B as affine functions of position (24 coefficients; NULL: none)
void orc_sim_set_beam_external(void* h, const double* ext) {
    Beam& b = static_cast<Sim*>(h)->beam;
    if (ext) b.ext.assign(ext, ext + 24);
    else b.ext.clear();
}

void orc_sim_add_capture(void* h, int comp, double* buf) { static_cast<Sim*>(h)->captures.push_back({comp, buf}); }
void orc_sim_clear_captures(void* h) { static_cast<Sim*>(h)->captures.clear(); }

void orc_sim_begin_step(void* h) { begin_step(*static_cast<Sim*>(h)); }
void orc_sim_solve_slice(void* h, int islice) { solve_one_slice(*static_cast<Sim*>(h), islice); }

// Runs one full step; returns the wall time of the slice loop in seconds.
double orc_sim_run_step(void* h, int first_slice, int last_slice) {
    Sim* s = static_cast<Sim*>(h);
    begin_step(*s);
    const auto t0 = std::chrono::steady_clock::now();
    for (int isl = first_slice; isl >= last_slice; --isl) solve_one_slice(*s, isl);
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

void orc_sim_checksums(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(out, s->checks.data(), sizeof(double) * s->checks.size());
}

void orc_sim_stats(void* h, double* out /* 6 */) {
    Sim* s = static_cast<Sim*>(h);
    out[0] = double(s->n_qsa); out[1] = double(s->mg_iters); out[2] = double(s->mg_status);
    out[3] = double(s->pc_iters); out[4] = s->pushes; out[5] = double(s->plasma.size());
}

void orc_sim_get_store(void* h, double* out) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(out, s->F.data(), sizeof(double) * s->F.size());
}
void orc_sim_set_store(void* h, const double* in) {
    Sim* s = static_cast<Sim*>(h);
    std::memcpy(s->F.data(), in, sizeof(double) * s->F.size());
}

int64_t orc_sim_plasma_size(void* h) { return int64_t(static_cast<Sim*>(h)->plasma.size()); }

// 11 real arrays in PlasmaIdx order x,y,w,ux,uy,psi,x_prev,y_prev,ux_h,uy_h,psi_h + valid
void orc_sim_get_plasma(void* h, double* reals /* 11*n */, int32_t* valid) {
    Sim* s = static_cast<Sim*>(h);
    Plasma& p = s->plasma;
    const size_t n = p.size();
    const std::vector<double>* arr[] = {&p.x, &p.y, &p.w, &p.ux, &p.uy, &p.psi, &p.x_prev, &p.y_prev, &p.ux_h, &p.uy_h, &p.psi_h};
    for (int k = 0; k < 11; ++k) std::memcpy(reals + k * n, arr[k]->data(), sizeof(double) * n);
    std::memcpy(valid, p.valid.data(), sizeof(int32_t) * n);
}
void orc_sim_set_plasma(void* h, int64_t n, const double* reals, const int32_t* valid) {
    Sim* s = static_cast<Sim*>(h);
    Plasma& p = s->plasma;
    p.resize(size_t(n));
    std::vector<double>* arr[] = {&p.x, &p.y, &p.w, &p.ux, &p.uy, &p.psi, &p.x_prev, &p.y_prev, &p.ux_h, &p.uy_h, &p.psi_h};
    for (int k = 0; k < 11; ++k) std::memcpy(arr[k]->data(), reals + k * n, sizeof(double) * n);
    std::memcpy(p.valid.data(), valid, sizeof(int32_t) * n);
}

// ---- kernel-level entry points on one slice store (for unit parity tests) ----
void orc_deposit_plasma(void* h, const int* comps /* jx,jy,jz,rho,chi,rhomjz */, int ions, int64_t* nq) {
    Sim* s = static_cast<Sim*>(h);
    long q = 0;
    deposit_plasma(s->c, s->plasma, s->F.data(), comps[0], comps[1], comps[2], comps[3], comps[4], comps[5], ions != 0, &q);
    if (nq) *nq = q;
}
void orc_explicit_deposit(void* h) { Sim* s = static_cast<Sim*>(h); explicit_deposit(s->c, s->plasma, s->F.data()); }
void orc_deposit_beam(void* h, int slice, int cjx, int cjy, int cjz) {
    Sim* s = static_cast<Sim*>(h);
    deposit_beam(s->c, s->beam, slice, s->F.data(), cjx, cjy, cjz);
}
void orc_push_plasma(void* h, int temp_slice) {
    Sim* s = static_cast<Sim*>(h);
    const Layout& L = s->c.L;
    push_plasma(s->c, s->plasma, s->F.data(), temp_slice != 0, L.Psi, L.Ez, L.Bx, L.By, L.Bz);
}
void orc_solve_poisson_psi_ez_bz(void* h) { Sim* s = static_cast<Sim*>(h); solve_poisson_psi_ez_bz(s->c, s->F.data()); }
void orc_solve_poisson_bxby(void* h) { Sim* s = static_cast<Sim*>(h); solve_poisson_bxby(s->c, s->F.data()); }
void orc_initialize_sxsy_with_beam(void* h) { Sim* s = static_cast<Sim*>(h); initialize_sxsy_with_beam(s->c, s->F.data()); }
int orc_mg_solve1_store(void* h, double tol_rel, double tol_abs, int maxiter, int* iters) {
    Sim* s = static_cast<Sim*>(h);
    const Layout& L = s->c.L;
    double* F = s->F.data();
    return mg_solve1(s->c.g, F + L.Bx * s->c.g.plane, F + L.Sy * s->c.g.plane, F + L.chi * s->c.g.plane,
                     s->c.g.plane, tol_rel, tol_abs, maxiter, iters);
}

// plain DST-I (RODFT00) of a contiguous vector
void orc_dst1(const double* in, double* out, int n) { dst1(in, out, n, 1, 1); }
// Poisson solve of a staging array (nx*ny) into a guarded plane
void orc_poisson(const orc_config* cf, const double* staging, double* lhs) {
    const Geom g = geom_from(cf);
    poisson_dirichlet(g, staging, lhs);
}
// shape factors (for closed-form tests): out = {s, cell} or {s, ds, cell}
void orc_shape(int order, double xmid, int ix, double* out) {
    const SF r = shape(order, xmid, ix);
    out[0] = r.s; out[1] = r.cell;
}
void orc_dshape(int type, int order, double xmid, int ix, double* out) {
    const DSF r = dshape(type, order, xmid, ix);
    out[0] = r.s; out[1] = r.ds; out[2] = r.cell;
}

// AdvanceBeamParticlesSlice on caller arrays with the store of the simulation; returns the
// number of particles listed in slip_out (capacity n_slip_cap)
long orc_push_beam_slice(void* h, double* x, double* y, double* z, double* ux, double* uy, double* uz, double* w,
                         int32_t* valid, int32_t* nsub, double charge, double mass, long first, long count,
                         const int32_t* slip_in, long n_in, int32_t* slip_out, long n_slip_cap, int slice, double dt,
                         int n_subcycles, int do_z_push, const double* ext) {
    Sim* s = static_cast<Sim*>(h);
    BeamRef b{x, y, z, ux, uy, uz, w, valid, nsub, charge, mass, ext};
    std::vector<int32_t> out;
    push_beam_slice(s->c, s->F.data(), b, first, count, slip_in, n_in, out, slice, dt, n_subcycles, do_z_push != 0);
    const long n = std::min<long>(long(out.size()), n_slip_cap);
    for (long k = 0; k < n; ++k) slip_out[k] = out[k];
    return long(out.size());
}
// One time step with the beam pushed after every slice (src/Hipace.cpp:557-728): counters
// reset, then SolveOneSlice and AdvanceBeamParticlesSlice per slice, slipped particles handed
// to the next slice. After orc_sim_begin_step.
void orc_sim_run_step_with_beam(void* h, double dt, int n_subcycles, int do_z_push) {
    Sim* s = static_cast<Sim*>(h);
    Beam& b = s->beam;
    const long n = long(b.x.size());
    std::vector<int32_t> nsub(n, 0), slip_in, slip_out;
    BeamRef r{b.x.data(), b.y.data(), b.z.data(), b.ux.data(), b.uy.data(), b.uz.data(), b.w.data(), b.valid.data(),
              nsub.data(), b.charge, b.mass, b.ext.empty() ? nullptr : b.ext.data()};
    for (int isl = s->c.g.nz - 1; isl >= 0; --isl) {
        solve_one_slice(*s, isl);
        slip_out.clear();
        const long first = b.off[isl], count = b.off[isl + 1] - b.off[isl];
        push_beam_slice(s->c, s->F.data(), r, first, count, slip_in.data(), long(slip_in.size()), slip_out, isl, dt,
                        n_subcycles, do_z_push != 0);
        slip_in.swap(slip_out);
    }
}

// BoxSorter::sortParticlesByBox (src/particles/sorting/BoxSort.cpp:20-60): stable sort of the
// beam by zeta-slice; invalid or outside the domain last, made invalid
void orc_sim_sort_beam(void* h) {
    Sim* s = static_cast<Sim*>(h);
    Beam& b = s->beam;
    const Geom& g = s->c.g;
    const long n = long(b.x.size());
    std::vector<int> key(n);
    for (long i = 0; i < n; ++i) {
        int k = g.nz;
        if (b.valid[i]) {
            const double sl = std::floor((b.z[i] - g.zlo) * g.dzi);
            if (sl >= 0.0 && sl < double(g.nz)) k = int(sl);
        }
        key[i] = k;
    }
    std::vector<long> order(n);
    for (long i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](long a, long c) { return key[a] < key[c]; });
    auto perm = [&](std::vector<double>& v) {
        std::vector<double> t(n);
        for (long i = 0; i < n; ++i) t[i] = v[order[i]];
        v.swap(t);
    };
    for (auto* v : {&b.x, &b.y, &b.z, &b.ux, &b.uy, &b.uz, &b.w}) perm(*v);
    std::vector<int32_t> vv(n);
    for (long i = 0; i < n; ++i) vv[i] = b.valid[order[i]];
    b.valid.swap(vv);
    b.off.assign(g.nz + 1, 0);
    for (long i = 0; i < n; ++i)
        if (key[order[i]] < g.nz) b.off[key[order[i]] + 1]++;
    for (int sl = 0; sl < g.nz; ++sl) b.off[sl + 1] += b.off[sl];
    for (long i = b.off[g.nz]; i < n; ++i) {
        b.valid[i] = 0;
        b.w[i] = 0.0;
    }
}

// copies the simulation's beam out (arrays of n, off of nz+1)
void orc_sim_get_beam(void* h, double* x, double* y, double* z, double* ux, double* uy, double* uz, double* w,
                      int32_t* valid, int64_t* off) {
    Sim* s = static_cast<Sim*>(h);
    const Beam& b = s->beam;
    const size_t n = b.x.size();
    std::copy(b.x.begin(), b.x.end(), x); std::copy(b.y.begin(), b.y.end(), y); std::copy(b.z.begin(), b.z.end(), z);
    std::copy(b.ux.begin(), b.ux.end(), ux); std::copy(b.uy.begin(), b.uy.end(), uy);
    std::copy(b.uz.begin(), b.uz.end(), uz); std::copy(b.w.begin(), b.w.end(), w);
    std::copy(b.valid.begin(), b.valid.begin() + n, valid);
    std::copy(b.off.begin(), b.off.end(), off);
}
}  // extern "C"
