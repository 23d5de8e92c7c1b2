// Oracle: CPU FP64 restatement of HiPACE++'s per-slice quasi-static loop.
//
// TEST INFRASTRUCTURE ONLY. This code is the checker for the HIP product path in
// hipace2_amd/. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it. The product library never links, calls or falls back to it.
//
// Each function cites the reference file:line whose algorithm it restates.
// Arithmetic is written so that every floating-point operation happens in the same
// order as the reference expression (compiled with -ffp-contract=off), so results
// agree with the reference CPU build to rounding of the summation order only.
//
// Floating-point sums run in fixed orders: deposition in the CPU build's tiled four-colour order
// (particles.cpp), the DST with rows paired as the product pairs them (fields.cpp), reductions
// in the blocked orders the product uses. The product's FP64 path runs the same arithmetic, so
// the two agree bit for bit.
//
// Parity pins (see tests/): golden checksums of the reference's own CPU/double CI
// (tests/checksum/benchmarks_json/*.json), closed forms (DST eigenmodes, B-spline
// moments, linear-field gather, uniform-field push), brute force on tiny inputs.
#pragma once
#include <cstdint>
#include <cmath>
#include <vector>
#include <string>

namespace orc {

// Geometry of one transverse slice plus the longitudinal grid.
// Positions follow GetPosOffset (src/fields/Fields.H:72-77): x_i = i*dx + xoff.
struct Geom {
    int nx = 0, ny = 0, nz = 0, ng = 2;
    double xlo = 0, xhi = 0, ylo = 0, yhi = 0, zlo = 0, zhi = 0;
    double dx = 0, dy = 0, dz = 0, dxi = 0, dyi = 0, dzi = 0;
    double xoff = 0, yoff = 0, zoff = 0;
    int ldx = 0, ldy = 0;
    long plane = 0;
    void finalize() {
        // amrex::Geometry: CellSize = length/ncell, InvCellSize = 1/CellSize
        dx = (xhi - xlo) / double(nx); dy = (yhi - ylo) / double(ny); dz = (zhi - zlo) / double(nz);
        dxi = 1.0 / dx; dyi = 1.0 / dy; dzi = 1.0 / dz;
        // src/fields/Fields.H:72-77 with the box of valid (+guard) cells: smallEnd+bigEnd = n-1
        xoff = 0.5 * (xlo + xhi - dx * double(nx - 1));
        yoff = 0.5 * (ylo + yhi - dy * double(ny - 1));
        zoff = 0.5 * (zlo + zhi - dz * double(nz - 1));
        ldx = nx + 2 * ng; ldy = ny + 2 * ng; plane = long(ldx) * ldy;
    }
    inline long idx(int i, int j) const { return long(j + ng) * ldx + (i + ng); }
};

// Component layout of the slice store, src/fields/Fields.cpp:70-165.
// -1 means "not allocated".
struct Layout {
    int ncomp = 0;
    // explicit solver
    int next_jx_beam = -1, next_jy_beam = -1;
    int chi = -1, Sy = -1, Sx = -1, ExmBy = -1, EypBx = -1, Ez = -1, Bx = -1, By = -1, Bz = -1,
        Psi = -1, jx_beam = -1, jy_beam = -1, jz_beam = -1, jx = -1, jy = -1, rhomjz = -1, rho = -1;
    int prev_jx_beam = -1, prev_jy_beam = -1;
    int ions_rhomjz = -1;
    // predictor-corrector only
    int next_jx = -1, next_jy = -1, jz = -1;
    int prev_Bx = -1, prev_By = -1, prev_jx = -1, prev_jy = -1;
    int pciter_Bx = -1, pciter_By = -1, pcprev_Bx = -1, pcprev_By = -1;
    // WhichSlice::Salame (src/fields/Fields.cpp:111-115)
    int sal_Ez_target = -1, sal_Ez_no_salame = -1, sal_Ez = -1, sal_jx = -1, sal_jy = -1, sal_jz_beam = -1,
        sal_Bx = -1, sal_By = -1, sal_Sy = -1, sal_Sx = -1, sal_Sy_back = -1, sal_Sx_back = -1;
};
Layout make_layout(bool explicit_solver, bool deposit_rho, bool neutralize, bool salame = false);

// Plasma species SoA (src/particles/plasma/PlasmaParticleContainer.H:21-46).
struct Plasma {
    std::vector<double> x, y, w, ux, uy, psi, x_prev, y_prev, ux_h, uy_h, psi_h;
    std::vector<int32_t> valid;
    double charge = -1.0, mass = 1.0;
    double max_qsa = 35.0;
    bool neutralize = true;
    size_t size() const { return x.size(); }
    void resize(size_t n) {
        for (auto* v : {&x, &y, &w, &ux, &uy, &psi, &x_prev, &y_prev, &ux_h, &uy_h, &psi_h})
            v->assign(n, 0.0);
        valid.assign(n, 0);
    }
};

// Beam particles sorted by slice (CSR): particles of slice s are [off[s], off[s+1]).
struct Beam {
    std::vector<double> x, y, z, ux, uy, uz, w;
    std::vector<int32_t> valid;
    std::vector<int64_t> off;   // nz+1
    double charge = -1.0, mass = 1.0;
    std::vector<double> ext;    // 24 affine external-field coefficients, or empty
    std::vector<int32_t> salame;  // 1 for particles of a SALAME beam, or empty
};

enum ParticleBC { BC_REFLECTING = 0, BC_PERIODIC = 1, BC_ABSORBING = 2 };

struct Options {
    int depos_order = 2;          // hipace.depos_order_xy (default 2, src/Hipace.H:73)
    int derivative_type = 2;      // hipace.depos_derivative_type (default 2)
    bool explicit_solver = true;  // hipace.bxby_solver (default explicit)
    bool deposit_rho = false;
    bool do_beam_jx_jy = true;    // hipace.do_beam_jx_jy_deposition
    double mg_tol_rel = 1e-4;     // hipace.MG_tolerance_rel
    double mg_tol_abs = 2.2250738585072014e-308;  // numeric_limits<double>::min()
    int mg_maxiter = 200;
    int particle_bc = BC_PERIODIC;
    double plo[2] = {0, 0}, phi[2] = {0, 0};  // particle boundary box
    double pc_tol = 4e-2;         // predcorr_B_error_tolerance
    int pc_maxiter = 30;          // predcorr_max_iterations
    double pc_mix = 0.05;         // predcorr_B_mixing_factor
    int nthreads = 0;             // OpenMP threads (0: default)
    int tile_size = 32;           // hipace.tile_size (src/Hipace.H:240)
};

// ---- shape factors: src/particles/particles_utils/ShapeFactors.H ----
struct SF { double s; int cell; };
struct DSF { double s; double ds; int cell; };  // ds is the returned dx_factor (= -table sdx)
SF shape(int order, double xmid, int ix);
DSF dshape(int type, int order, double xmid, int ix);

// ---- kernels (plain FP64, serial particle order) ----
struct Ctx {
    Geom g;
    Options o;
    Layout L;
};

void deposit_plasma(const Ctx& c, Plasma& p, double* F, int cjx, int cjy, int cjz, int crho,
                    int cchi, int crhomjz, bool ions, long* n_qsa);
void explicit_deposit(const Ctx& c, const Plasma& p, double* F);
// select (may be null): only particles whose entry equals select_value deposit
void deposit_beam(const Ctx& c, const Beam& b, int slice, double* F, int cjx, int cjy, int cjz,
                  const int32_t* select = nullptr, int select_value = 0);
// SalameOnlyAdvancePlasma, src/salame/Salame.cpp:258-330
void salame_only_advance(const Ctx& c, Plasma& p, const double* F);
// Fields::SolvePoissonEz, src/fields/Fields.cpp:960-1006: Ez from jx, jy (components of F)
void solve_poisson_ez(const Ctx& c, double* F, int cjx, int cjy, int cez);
void push_plasma(const Ctx& c, Plasma& p, const double* F, bool temp_slice, int cPsi, int cEz,
                 int cBx, int cBy, int cBz);
// mutable beam arrays for the push (caller-owned)
struct BeamRef {
    double *x, *y, *z, *ux, *uy, *uz, *w;
    int32_t *valid, *nsub;
    double charge, mass;
    // external fields affine in position, (Ex, Ey, Ez, Bx, By, Bz) x (1, x, y, z); null: none
    const double* ext = nullptr;
};
void push_beam_slice(const Ctx& c, const double* F, const BeamRef& b, long first, long count, const int32_t* slip_in,
                     long n_in, std::vector<int32_t>& slip_out, int slice, double dt_step, int n_subcycles,
                     bool do_z_push);
void dst1(const double* in, double* out, int n, int stride_in, int stride_out);
void poisson_dirichlet(const Geom& g, const double* staging, double* lhs);  // staging: nx*ny
// nf right-hand sides solved together (their rows paired as one list)
void poisson_dirichlet_n(const Geom& g, int nf, const double* const* staging, double* const* lhs);
void solve_poisson_psi_ez_bz(const Ctx& c, double* F);
void solve_poisson_bxby(const Ctx& c, double* F);
void initialize_sxsy_with_beam(const Ctx& c, double* F);
int mg_solve1(const Geom& g, double* sol, const double* rhs, const double* acf, long comp_stride,
              double tol_rel, double tol_abs, int maxiter, int* iters);

}  // namespace orc
