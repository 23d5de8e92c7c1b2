// Oracle field operations (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
#include "oracle.hpp"
#include <complex>
#include <algorithm>

namespace orc {

// Component layout, restating Fields::AllocData src/fields/Fields.cpp:70-165 for the
// configurations this path covers (no laser, no SALAME, no MLMG, no individual rho).
Layout make_layout(bool explicit_solver, bool deposit_rho, bool neutralize, bool salame) {
    Layout L;
    int n = 0;
    if (explicit_solver) {
        L.next_jx_beam = n++; L.next_jy_beam = n++;
        L.chi = n++; L.Sy = n++; L.Sx = n++; L.ExmBy = n++; L.EypBx = n++; L.Ez = n++;
        L.Bx = n++; L.By = n++; L.Bz = n++; L.Psi = n++;
        L.jx_beam = n++; L.jy_beam = n++; L.jz_beam = n++; L.jx = n++; L.jy = n++; L.rhomjz = n++;
        if (deposit_rho) L.rho = n++;
        L.prev_jx_beam = n++; L.prev_jy_beam = n++;
        if (neutralize) L.ions_rhomjz = n++;
        if (salame) {
            L.sal_Ez_target = n++; L.sal_Ez_no_salame = n++; L.sal_Ez = n++; L.sal_jx = n++; L.sal_jy = n++;
            L.sal_jz_beam = n++; L.sal_Bx = n++; L.sal_By = n++; L.sal_Sy = n++; L.sal_Sx = n++;
            L.sal_Sy_back = n++; L.sal_Sx_back = n++;
        }
    } else {
        L.next_jx = n++; L.next_jy = n++;
        L.ExmBy = n++; L.EypBx = n++; L.Ez = n++; L.Bx = n++; L.By = n++; L.Bz = n++; L.Psi = n++;
        L.jx = n++; L.jy = n++; L.jz = n++; L.rhomjz = n++;
        if (deposit_rho) L.rho = n++;
        L.prev_Bx = n++; L.prev_By = n++; L.prev_jx = n++; L.prev_jy = n++;
        if (neutralize) L.ions_rhomjz = n++;
        L.pciter_Bx = n++; L.pciter_By = n++;
        L.pcprev_Bx = n++; L.pcprev_By = n++;
    }
    L.ncomp = n;
    return L;
}

// ---------------------------------------------------------------------------------
// DST-I with FFTW RODFT00 normalisation: out_k = 2 sum_j in_j sin(pi (j+1)(k+1)/(n+1)).
// This is the transform the reference uses on CPU (WrapFFTW.cpp:121-124, RODFT00).
// For n+1 a power of two an in-repo radix-2 FFT of the odd extension is used;
// otherwise a direct sum with an exactly reduced sine table.
// ---------------------------------------------------------------------------------
static void fft_inplace(std::vector<std::complex<double>>& a, const std::vector<std::complex<double>>& w) {
    const int n = int(a.size());
    for (int i = 1, j = 0; i < n; ++i) {  // bit reversal
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; ++k) {
                const std::complex<double> u = a[i + k];
                const std::complex<double> v = a[i + k + len / 2] * w[k * step];
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
        }
    }
}

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

static double sin_pi_frac(long m, long N) {
    // sin(pi*m/N) with m reduced mod 2N and symmetric folding for accuracy
    m %= 2 * N;
    if (m < 0) m += 2 * N;
    double sgn = 1.0;
    if (m >= N) { m -= N; sgn = -1.0; }
    if (2 * m > N) m = N - m;
    return sgn * std::sin(M_PI * double(m) / double(N));
}

void dst1(const double* in, double* out, int n, int si, int so) {
    const int N = n + 1;
    if (is_pow2(N) && N >= 2) {
        const int M = 2 * N;
        std::vector<std::complex<double>> a(M), w(M / 2);
        for (int k = 0; k < M / 2; ++k) w[k] = std::complex<double>(sin_pi_frac(k + N / 2, N), -sin_pi_frac(k, N));
        // w[k] = exp(-2 pi i k / M) = cos(pi k/N) - i sin(pi k/N); cos(a) = sin(a + pi/2)
        a[0] = 0.0; a[N] = 0.0;
        for (int j = 0; j < n; ++j) { a[j + 1] = in[long(j) * si]; a[M - 1 - j] = -in[long(j) * si]; }
        fft_inplace(a, w);
        // A_k = -2i sum_m x_m sin(pi m k / N)  =>  out_{k-1} = 2 * (-Im A_k / 2) = -Im A_k
        for (int k = 1; k <= n; ++k) out[long(k - 1) * so] = -a[k].imag();
    } else {
        std::vector<double> tab(2 * N);
        for (int m = 0; m < 2 * N; ++m) tab[m] = sin_pi_frac(m, N);
        std::vector<double> tmp(n);
        for (int k = 0; k < n; ++k) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += in[long(j) * si] * tab[(long(j + 1) * (k + 1)) % (2 * N)];
            tmp[k] = 2.0 * s;
        }
        for (int k = 0; k < n; ++k) out[long(k) * so] = tmp[k];
    }
}

// FFTPoissonSolverDirichletDirect::SolvePoissonEquation, src/fields/fft_poisson_solver/
// FFTPoissonSolverDirichletDirect.cpp:43-139: forward 2-D DST-I, multiply by the
// eigenvalue matrix (with RODFT00 normalisation folded in), inverse 2-D DST-I.
// staging is nx*ny (x fastest, no guards); lhs is a guarded plane, valid cells written.
void poisson_dirichlet(const Geom& g, const double* staging, double* lhs) {
    const int nx = g.nx, ny = g.ny;
    std::vector<double> a(size_t(nx) * ny), b(size_t(nx) * ny);
    const double dxsq = g.dx * g.dx, dysq = g.dy * g.dy;
    const double sxf = M_PI / (2. * (nx + 1)), syf = M_PI / (2. * (ny + 1));
    const double norm_fac = 0.5 / (2 * ((nx + 1) * (ny + 1)));
#pragma omp parallel for
    for (int j = 0; j < ny; ++j) dst1(staging + long(j) * nx, a.data() + long(j) * nx, nx, 1, 1);
#pragma omp parallel for
    for (int i = 0; i < nx; ++i) dst1(a.data() + i, b.data() + i, ny, nx, nx);
#pragma omp parallel for
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const double sx = std::sin((i + 1) * sxf) * std::sin((i + 1) * sxf);
            const double sy = std::sin((j + 1) * syf) * std::sin((j + 1) * syf);
            double ev = 0.0;
            if (sx != 0 && sy != 0) ev = norm_fac / (-4.0 * (sx / dxsq + sy / dysq));
            b[long(j) * nx + i] *= ev;
        }
    }
#pragma omp parallel for
    for (int i = 0; i < nx; ++i) dst1(b.data() + i, a.data() + i, ny, nx, nx);
#pragma omp parallel for
    for (int j = 0; j < ny; ++j) {
        std::vector<double> row(nx);
        dst1(a.data() + long(j) * nx, row.data(), nx, 1, 1);
        for (int i = 0; i < nx; ++i) lhs[g.idx(i, j)] = row[i];
    }
}

// Fields::SolvePoissonPsiExmByEypBxEzBz, src/fields/Fields.cpp:841-957, level 0,
// Dirichlet, normalised units (ep0 = mu0 = c = 1). RHS builders: Multiply
// (:397-411), LinCombination (:370-387) of derivative<> (:223-277).
void solve_poisson_psi_ez_bz(const Ctx& c, double* F) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const double ep0 = 1.0, mu0 = 1.0, cl = 1.0;
    const int nx = g.nx, ny = g.ny;
    std::vector<double> st(size_t(nx) * ny);
    const double* rhomjz = F + L.rhomjz * g.plane;
    const double* jx = F + L.jx * g.plane;
    const double* jy = F + L.jy * g.plane;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    // Psi
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) st[long(j) * nx + i] = (-1.0 / ep0) * rhomjz[g.idx(i, j)];
    poisson_dirichlet(g, st.data(), F + L.Psi * g.plane);
    // Ez
    const double fe = 1.0 / (ep0 * cl);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            st[long(j) * nx + i] = fe * ((jx[g.idx(i + 1, j)] - jx[g.idx(i - 1, j)]) * dxih) +
                                   fe * ((jy[g.idx(i, j + 1)] - jy[g.idx(i, j - 1)]) * dyih);
    poisson_dirichlet(g, st.data(), F + L.Ez * g.plane);
    // Bz
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            st[long(j) * nx + i] = mu0 * ((jx[g.idx(i, j + 1)] - jx[g.idx(i, j - 1)]) * dyih) +
                                   (-mu0) * ((jy[g.idx(i + 1, j)] - jy[g.idx(i - 1, j)]) * dxih);
    poisson_dirichlet(g, st.data(), F + L.Bz * g.plane);
    // ExmBy = -d/dx Psi, EypBx = -d/dy Psi on valid + 1 guard ring (Fields.cpp:931-956)
    const double* psi = F + L.Psi * g.plane;
    double* exmby = F + L.ExmBy * g.plane;
    double* eypbx = F + L.EypBx * g.plane;
    for (int j = -1; j <= ny; ++j)
        for (int i = -1; i <= nx; ++i) {
            exmby[g.idx(i, j)] = -(psi[g.idx(i + 1, j)] - psi[g.idx(i - 1, j)]) * dxih;
            eypbx[g.idx(i, j)] = -(psi[g.idx(i, j + 1)] - psi[g.idx(i, j - 1)]) * dyih;
        }
}

// Fields::SolvePoissonEz, src/fields/Fields.cpp:960-1006 (level 0, Dirichlet): the Ez part
// of solve_poisson_psi_ez_bz on other components
void solve_poisson_ez(const Ctx& c, double* F, int cjx, int cjy, int cez) {
    const Geom& g = c.g;
    const double ep0 = 1.0, cl = 1.0;
    const int nx = g.nx, ny = g.ny;
    std::vector<double> st(size_t(nx) * ny);
    const double* jx = F + cjx * g.plane;
    const double* jy = F + cjy * g.plane;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const double fe = 1.0 / (ep0 * cl);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
            st[long(j) * nx + i] = fe * ((jx[g.idx(i + 1, j)] - jx[g.idx(i - 1, j)]) * dxih) +
                                   fe * ((jy[g.idx(i, j + 1)] - jy[g.idx(i, j - 1)]) * dyih);
    poisson_dirichlet(g, st.data(), F + cez * g.plane);
}

// Fields::SolvePoissonBxBy, src/fields/Fields.cpp:1009-1078 (predictor-corrector path).
// Solves into the PCIter Bx/By planes.
void solve_poisson_bxby(const Ctx& c, double* F) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const double mu0 = 1.0;
    const int nx = g.nx, ny = g.ny;
    std::vector<double> st(size_t(nx) * ny);
    const double* jz = F + L.jz * g.plane;
    const double* pjx = F + L.prev_jx * g.plane;
    const double* pjy = F + L.prev_jy * g.plane;
    const double* njx = F + L.next_jx * g.plane;
    const double* njy = F + L.next_jy * g.plane;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi, dzih = 0.5 * g.dzi;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const long k = g.idx(i, j);
            st[long(j) * nx + i] = -mu0 * ((jz[g.idx(i, j + 1)] - jz[g.idx(i, j - 1)]) * dyih) +
                                   mu0 * ((pjy[k] - njy[k]) * dzih);
        }
    poisson_dirichlet(g, st.data(), F + L.pciter_Bx * g.plane);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const long k = g.idx(i, j);
            st[long(j) * nx + i] = mu0 * ((jz[g.idx(i + 1, j)] - jz[g.idx(i - 1, j)]) * dxih) +
                                   (-mu0) * ((pjx[k] - njx[k]) * dzih);
        }
    poisson_dirichlet(g, st.data(), F + L.pciter_By * g.plane);
}

// Hipace::InitializeSxSyWithBeam, src/Hipace.cpp:745-790 (valid cells).
void initialize_sxsy_with_beam(const Ctx& c, double* F) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const double mu0 = 1.0;
    const double* jzb = F + L.jz_beam * g.plane;
    const double* pjx = F + L.prev_jx_beam * g.plane;
    const double* pjy = F + L.prev_jy_beam * g.plane;
    const double* njx = F + L.next_jx_beam * g.plane;
    const double* njy = F + L.next_jy_beam * g.plane;
    double* Sx = F + L.Sx * g.plane;
    double* Sy = F + L.Sy * g.plane;
    for (int j = 0; j < g.ny; ++j)
        for (int i = 0; i < g.nx; ++i) {
            const long k = g.idx(i, j);
            const double dx_jzb = (jzb[g.idx(i + 1, j)] - jzb[g.idx(i - 1, j)]) / (2.0 * g.dx);
            const double dy_jzb = (jzb[g.idx(i, j + 1)] - jzb[g.idx(i, j - 1)]) / (2.0 * g.dy);
            const double dz_jxb = (pjx[k] - njx[k]) / (2.0 * g.dz);
            const double dz_jyb = (pjy[k] - njy[k]) / (2.0 * g.dz);
            Sy[k] = mu0 * (-dy_jzb + dz_jyb);
            Sx[k] = -mu0 * (-dx_jzb + dz_jxb);
        }
}

}  // namespace orc
