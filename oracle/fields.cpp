// Oracle field operations (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
#include "oracle.hpp"
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
// For n+1 a power of two (and 2(n+1) <= 4096) two real rows a, b are transformed together:
// the odd extension of a + i b, of length M = 2(n+1), goes through a radix-2
// decimation-in-time FFT F, and RODFT00(a)_{k-1} = -Im F_k, RODFT00(b)_{k-1} = Re F_k (the
// transforms of odd real sequences are imaginary). Otherwise a direct sum with an exactly
// reduced sine table, row by row. Any DST-I algorithm agrees with FFTW to rounding; this one
// fixes every rounding, so a GPU that runs the same arithmetic gives the same bits.
// ---------------------------------------------------------------------------------
struct Cplx { double re, im; };
static inline Cplx cmul(Cplx x, Cplx w) { return {x.re * w.re - x.im * w.im, x.re * w.im + x.im * w.re}; }

// in-place radix-2 DIT FFT of a (input in natural order), twiddles w[k] = exp(-2 pi i k / M)
static void fft_inplace(std::vector<Cplx>& a, const std::vector<Cplx>& w) {
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
                const Cplx u = a[i + k];
                const Cplx v = cmul(a[i + k + len / 2], w[k * step]);
                a[i + k] = {u.re + v.re, u.im + v.im};
                a[i + k + len / 2] = {u.re - v.re, u.im - v.im};
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

static bool dst_fft_path(int n) { return is_pow2(n + 1) && n + 1 >= 2 && 2 * (n + 1) <= 4096; }

static std::vector<Cplx> twiddles(int N) {
    std::vector<Cplx> w(N);
    // w[k] = exp(-2 pi i k / M) = cos(pi k/N) - i sin(pi k/N); cos(a) = sin(a + pi/2)
    for (int k = 0; k < N; ++k) w[k] = {sin_pi_frac(k + N / 2, N), -sin_pi_frac(k, N)};
    return w;
}

// RODFT00 of rows a and b (b may be null) on the FFT path, outputs times sa[k], sb[k] when
// given; strides in elements
static void dst1_pair(const double* a, const double* b, double* oa, double* ob, int n, int si, int so,
                      const std::vector<Cplx>& w, const double* sa, const double* sb) {
    const int N = n + 1, M = 2 * N;
    std::vector<Cplx> x(M, Cplx{0.0, 0.0});
    for (int m = 1; m <= n; ++m) {
        const Cplx v{a[long(m - 1) * si], b ? b[long(m - 1) * si] : 0.0};
        x[m] = v;
        x[M - m] = {-v.re, -v.im};
    }
    fft_inplace(x, w);
    for (int k = 1; k <= n; ++k) {
        double o0 = -x[k].im, o1 = x[k].re;
        if (sa) o0 *= sa[k - 1];
        if (sb) o1 *= sb[k - 1];
        oa[long(k - 1) * so] = o0;
        if (ob) ob[long(k - 1) * so] = o1;
    }
}

// direct sum of one row, times s[k] when given
static void dst1_direct(const double* in, double* out, int n, int si, int so, const std::vector<double>& tab,
                        const double* s) {
    const int N = n + 1;
    std::vector<double> tmp(n);
    for (int k = 0; k < n; ++k) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += in[long(j) * si] * tab[(long(j + 1) * (k + 1)) % (2 * N)];
        double o = 2.0 * acc;
        if (s) o *= s[k];
        tmp[k] = o;
    }
    for (int k = 0; k < n; ++k) out[long(k) * so] = tmp[k];
}

void dst1(const double* in, double* out, int n, int si, int so) {
    const int N = n + 1;
    if (dst_fft_path(n)) {
        dst1_pair(in, nullptr, out, nullptr, n, si, so, twiddles(N), nullptr, nullptr);
    } else {
        std::vector<double> tab(2 * N);
        for (int m = 0; m < 2 * N; ++m) tab[m] = sin_pi_frac(m, N);
        dst1_direct(in, out, n, si, so, tab, nullptr);
    }
}

// RODFT00 of nrows contiguous rows of length n: row r at in[r * ld_in], result at
// out[r * ld_out]; on the FFT path rows 2b and 2b+1 share one transform (an odd row count
// leaves the last alone). scale (nrows x n, may be null): result times scale[r][k].
// twice: the result is transformed once more with the same pairing, without scale.
static void dst_rows(const double* in, long ld_in, double* out, long ld_out, int nrows, int n, const double* scale,
                     bool twice) {
    const int N = n + 1;
    if (dst_fft_path(n)) {
        const std::vector<Cplx> w = twiddles(N);
#pragma omp parallel for
        for (int b = 0; b < (nrows + 1) / 2; ++b) {
            const int r0 = 2 * b, r1 = r0 + 1;
            const bool has1 = r1 < nrows;
            std::vector<double> t0(n), t1(n);
            dst1_pair(in + r0 * ld_in, has1 ? in + r1 * ld_in : nullptr, t0.data(), t1.data(), n, 1, 1, w,
                      scale ? scale + long(r0) * n : nullptr, (scale && has1) ? scale + long(r1) * n : nullptr);
            if (twice) {
                dst1_pair(t0.data(), has1 ? t1.data() : nullptr, out + r0 * ld_out, has1 ? out + r1 * ld_out : nullptr,
                          n, 1, 1, w, nullptr, nullptr);
            } else {
                std::copy(t0.begin(), t0.end(), out + r0 * ld_out);
                if (has1) std::copy(t1.begin(), t1.end(), out + r1 * ld_out);
            }
        }
    } else {
        std::vector<double> tab(2 * N);
        for (int m = 0; m < 2 * N; ++m) tab[m] = sin_pi_frac(m, N);
#pragma omp parallel for
        for (int r = 0; r < nrows; ++r) {
            std::vector<double> t(n);
            dst1_direct(in + r * ld_in, t.data(), n, 1, 1, tab, scale ? scale + long(r) * n : nullptr);
            if (twice) dst1_direct(t.data(), out + r * ld_out, n, 1, 1, tab, nullptr);
            else std::copy(t.begin(), t.end(), out + r * ld_out);
        }
    }
}

// FFTPoissonSolverDirichletDirect::SolvePoissonEquation, src/fields/fft_poisson_solver/
// FFTPoissonSolverDirichletDirect.cpp:43-139, for nf right-hand sides solved together:
// DST-I along x of every row of every field, DST-I along y with the eigenvalue matrix (with
// the RODFT00 normalisation folded in) applied and the inverse DST-I along y, inverse DST-I
// along x. The rows of all fields form one list, paired across field boundaries.
// staging[f] is nx*ny (x fastest, no guards); lhs[f] is a guarded plane, valid cells written.
void poisson_dirichlet_n(const Geom& g, int nf, const double* const* staging, double* const* lhs) {
    const int nx = g.nx, ny = g.ny;
    const long fs = long(nx) * ny;
    std::vector<double> a(size_t(nf) * fs), b(size_t(nf) * fs);
    for (int f = 0; f < nf; ++f) std::copy(staging[f], staging[f] + fs, a.begin() + f * fs);
    // x pass: rows j of every field
    dst_rows(a.data(), nx, b.data(), nx, nf * ny, nx, nullptr, false);
    // transpose to rows i
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) a[f * fs + long(i) * ny + j] = b[f * fs + long(j) * nx + i];
    // eigenvalue rows [i][j] (FFTPoissonSolverDirichletDirect.cpp:60-82), repeated per field
    const double dxsq = g.dx * g.dx, dysq = g.dy * g.dy;
    const double sxf = M_PI / (2. * (nx + 1)), syf = M_PI / (2. * (ny + 1));
    const double norm_fac = 0.5 / (2 * ((nx + 1) * (ny + 1)));
    std::vector<double> ev(size_t(nf) * fs);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const double sx = std::sin((i + 1) * sxf) * std::sin((i + 1) * sxf);
            const double sy = std::sin((j + 1) * syf) * std::sin((j + 1) * syf);
            double e = 0.0;
            if (sx != 0 && sy != 0) e = norm_fac / (-4.0 * (sx / dxsq + sy / dysq));
            for (int f = 0; f < nf; ++f) ev[f * fs + long(i) * ny + j] = e;
        }
    // y pass, scaling, inverse y pass
    dst_rows(a.data(), ny, b.data(), ny, nf * nx, ny, ev.data(), true);
    // transpose back to rows j
    for (int f = 0; f < nf; ++f)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) a[f * fs + long(j) * nx + i] = b[f * fs + long(i) * ny + j];
    // inverse x pass
    dst_rows(a.data(), nx, b.data(), nx, nf * ny, nx, nullptr, false);
    for (int f = 0; f < nf; ++f)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) lhs[f][g.idx(i, j)] = b[f * fs + long(j) * nx + i];
}

void poisson_dirichlet(const Geom& g, const double* staging, double* lhs) {
    poisson_dirichlet_n(g, 1, &staging, &lhs);
}

// Fields::SolvePoissonPsiExmByEypBxEzBz, src/fields/Fields.cpp:841-957, level 0,
// Dirichlet, normalised units (ep0 = mu0 = c = 1). RHS builders: Multiply
// (:397-411), LinCombination (:370-387) of derivative<> (:223-277).
void solve_poisson_psi_ez_bz(const Ctx& c, double* F) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const double ep0 = 1.0, mu0 = 1.0, cl = 1.0;
    const int nx = g.nx, ny = g.ny;
    const long fs = long(nx) * ny;
    std::vector<double> st(size_t(3) * fs);
    double* sp = st.data();
    double* se = sp + fs;
    double* sb = se + fs;
    const double* rhomjz = F + L.rhomjz * g.plane;
    const double* jx = F + L.jx * g.plane;
    const double* jy = F + L.jy * g.plane;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const double fe = 1.0 / (ep0 * cl);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            // Psi, Ez, Bz right-hand sides (Fields.cpp:886-917)
            sp[long(j) * nx + i] = (-1.0 / ep0) * rhomjz[g.idx(i, j)];
            se[long(j) * nx + i] = fe * ((jx[g.idx(i + 1, j)] - jx[g.idx(i - 1, j)]) * dxih) +
                                   fe * ((jy[g.idx(i, j + 1)] - jy[g.idx(i, j - 1)]) * dyih);
            sb[long(j) * nx + i] = mu0 * ((jx[g.idx(i, j + 1)] - jx[g.idx(i, j - 1)]) * dyih) +
                                   (-mu0) * ((jy[g.idx(i + 1, j)] - jy[g.idx(i - 1, j)]) * dxih);
        }
    const double* in[3] = {sp, se, sb};
    double* out[3] = {F + L.Psi * g.plane, F + L.Ez * g.plane, F + L.Bz * g.plane};
    poisson_dirichlet_n(g, 3, in, out);
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
    const long fs = long(nx) * ny;
    std::vector<double> st(size_t(2) * fs);
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
            st[fs + long(j) * nx + i] = mu0 * ((jz[g.idx(i + 1, j)] - jz[g.idx(i - 1, j)]) * dxih) +
                                        (-mu0) * ((pjx[k] - njx[k]) * dzih);
        }
    const double* in[2] = {st.data(), st.data() + fs};
    double* out[2] = {F + L.pciter_Bx * g.plane, F + L.pciter_By * g.plane};
    poisson_dirichlet_n(g, 2, in, out);
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
