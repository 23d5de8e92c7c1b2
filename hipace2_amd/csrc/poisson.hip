// Transverse Dirichlet Poisson solver on gfx950: batched DST-I along rows in LDS,
// LDS-tiled transposes between the x and y passes, eigenvalue scaling fused into the
// middle pass. Solves up to three fields (Psi, Ez, Bz) per call so each launch has
// 3 x n rows of work.
//
// Semantics follow FFTPoissonSolverDirichletDirect (src/fields/fft_poisson_solver/
// FFTPoissonSolverDirichletDirect.cpp:43-139): forward 2-D DST-I (FFTW RODFT00
// normalisation), multiply by the eigenvalue matrix (:60-82), inverse 2-D DST-I,
// copy the valid region into the destination plane.
//
// Row DST, power-of-two n+1: two real rows a, b are packed as a + i b, odd-extended to
// length M = 2(n+1) and transformed with a radix-2 FFT held in LDS (M <= 4096,
// 64 KiB). With F = FFT(a + i b), RODFT00(a)_{k-1} = -Im F_k and RODFT00(b)_{k-1} = Re F_k.
// Other n: direct sum with an exactly reduced sine table, one row per workgroup.
#include <cmath>

#include <type_traits>

#include "internal.hpp"

namespace hpq {

struct Rows {
    double* base[3];
    long stride;
    int rows_per_field;
};

__device__ __forceinline__ double* row_ptr(const Rows& R, int r) {
    const int f = r / R.rows_per_field;
    const int j = r - f * R.rows_per_field;
    return R.base[f] + long(j) * R.stride;
}

constexpr int DST_T = 256;

// complex a * w
__device__ __forceinline__ double2 cmul(double2 x, double2 w) {
    return make_double2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
}

// LDS slot of FFT entry i: the low four bits (16 double2 = the 64 banks) are XORed with a
// function of the higher bits, chosen so that the bit-reversed scatter of the load and every
// butterfly stage of fft_lds touch 16 distinct bank groups per 16 lanes for M in [64, 4096]
// (unswizzled, the scatter is a 16-way and the first two stage pairs 4-way conflicts). A
// bijection on [0, M) for every power of two M.
__device__ __forceinline__ int sw(int i) {
    const int h = i >> 4;
    return i ^ ((h * 5 ^ (h >> 4)) & 15);
}

// In-place radix-2 DIT FFT of the bit-reversed sequence a[0..M): the stages are taken in
// pairs (a 4-point butterfly per thread and one barrier per pair); the arithmetic is that
// of the radix-2 stages. tw[k] = exp(-2 pi i k / M).
__device__ void fft_lds(double2* a, int M, int logM, const double2* __restrict__ tw) {
    int s = 1;
    for (; s + 1 <= logM; s += 2) {
        const int h = 1 << (s - 1);
        const int st1 = M >> s, st2 = M >> (s + 1);
        for (int b = threadIdx.x; b < (M >> 2); b += DST_T) {
            const int k = b & (h - 1);
            const int i0 = ((b >> (s - 1)) << (s + 1)) + k;
            const int i1 = i0 + h, i2 = i0 + 2 * h, i3 = i0 + 3 * h;
            const double2 w1 = tw[k * st1], w2 = tw[k * st2], w3 = tw[(k + h) * st2];
            const double2 a0 = a[sw(i0)], a1 = a[sw(i1)], a2 = a[sw(i2)], a3 = a[sw(i3)];
            const double2 v1 = cmul(a1, w1), v3 = cmul(a3, w1);
            const double2 x0 = make_double2(a0.x + v1.x, a0.y + v1.y), x1 = make_double2(a0.x - v1.x, a0.y - v1.y);
            const double2 x2 = make_double2(a2.x + v3.x, a2.y + v3.y), x3 = make_double2(a2.x - v3.x, a2.y - v3.y);
            const double2 u2 = cmul(x2, w2), u3 = cmul(x3, w3);
            a[sw(i0)] = make_double2(x0.x + u2.x, x0.y + u2.y);
            a[sw(i2)] = make_double2(x0.x - u2.x, x0.y - u2.y);
            a[sw(i1)] = make_double2(x1.x + u3.x, x1.y + u3.y);
            a[sw(i3)] = make_double2(x1.x - u3.x, x1.y - u3.y);
        }
        __syncthreads();
    }
    if (s == logM) {   // odd number of stages: the last one alone
        const int half = 1 << (s - 1);
        const int step = M >> s;
        for (int b = threadIdx.x; b < (M >> 1); b += DST_T) {
            const int k = b & (half - 1);
            const int i0 = ((b >> (s - 1)) << s) + k;
            const int i1 = i0 + half;
            const double2 u = a[sw(i0)];
            const double2 v = cmul(a[sw(i1)], tw[k * step]);
            a[sw(i0)] = make_double2(u.x + v.x, u.y + v.y);
            a[sw(i1)] = make_double2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
}

// Two rows per workgroup: a + i b odd-extended to length M, bit-reversed into LDS.
// v(m) gives the pair of values at m in [1, n].
template <class V>
__device__ __forceinline__ void load_odd_extension(double2* a, int n, int M, int logM, V v) {
    const int N = M >> 1;
    for (int m = threadIdx.x; m < M; m += DST_T) {
        double2 x = make_double2(0.0, 0.0);
        if (m >= 1 && m <= n) {
            x = v(m);
        } else if (m > N) {
            const double2 y = v(M - m);
            x = make_double2(-y.x, -y.y);
        }
        a[sw(__brev(unsigned(m)) >> (32 - logM))] = x;
    }
    __syncthreads();
}

// Rows r0, r0 + 1: RODFT00 of each (with F = FFT of the packed odd extension,
// RODFT00(a)_{k-1} = -Im F_k and RODFT00(b)_{k-1} = Re F_k), optionally times the
// eigenvalue rows (scale != null). SYM: the result is transformed once more in LDS (the
// inverse pass of the same direction), so the middle of the 2-D solve is one launch.
// shift (parts != null): the shift sigma is the sum of the n parts times inv_count, added
// in a fixed order by every workgroup; the scale entries s = f / lambda become
// f / (lambda - sigma) = s / (1 - sigma s / f), f = 1 / inv_f, the solve of
// (Lap - sigma) u = rhs. add: the output is added to the destination.
struct Shift {
    const double* parts = nullptr;
    int n = 0;
    double inv_count = 0.0, inv_f = 0.0;
};

template <bool SYM>
__global__ __launch_bounds__(DST_T) void k_dst_rows_fft(Rows in, Rows out, int nrows, int n, int M, int logM,
                                                        const double2* __restrict__ tw,
                                                        const double* __restrict__ scale, Shift shift, int add) {
    extern __shared__ double2 a[];
    __shared__ double sigma_s;
    double sh = 0.0;
    if (shift.parts) {
        if (threadIdx.x < 64) {
            double t = 0.0;
            for (int q = threadIdx.x; q < shift.n; q += 64) t += shift.parts[q];
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
            if (threadIdx.x == 0) sigma_s = t * shift.inv_count;
        }
        __syncthreads();
        sh = sigma_s;
    }
    auto shifted = [&](double s) { return shift.parts ? s / (1.0 - sh * s * shift.inv_f) : s; };
    const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
    const bool has1 = r1 < nrows;
    const double* p0 = row_ptr(in, r0);
    const double* p1 = has1 ? row_ptr(in, r1) : p0;
    load_odd_extension(a, n, M, logM, [&](int m) { return make_double2(p0[m - 1], has1 ? p1[m - 1] : 0.0); });
    fft_lds(a, M, logM, tw);
    const double* s0 = scale ? scale + long(r0 % in.rows_per_field) * n : nullptr;
    const double* s1 = (scale && has1) ? scale + long(r1 % in.rows_per_field) * n : nullptr;
    if (SYM) {
        // RODFT00 of the scaled rows: pack them again; F_m is read for m in [1, n] only,
        // so the new sequence goes through registers (each thread M / DST_T entries)
        constexpr int MAXQ = 4096 / DST_T;
        double2 nv[MAXQ];
        const int N = M >> 1;
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
            const int m = threadIdx.x + q * DST_T;
            nv[q] = make_double2(0.0, 0.0);
            if (m < M) {
                const int src = (m >= 1 && m <= n) ? m : (m > N ? M - m : 0);
                if (src) {
                    const double2 F = a[sw(src)];
                    double o0 = -F.y, o1 = F.x;
                    if (s0) o0 *= shifted(s0[src - 1]);
                    if (s1) o1 *= shifted(s1[src - 1]);
                    if (!has1) o1 = 0.0;
                    nv[q] = m <= n ? make_double2(o0, o1) : make_double2(-o0, -o1);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
            const int m = threadIdx.x + q * DST_T;
            if (m < M) a[sw(__brev(unsigned(m)) >> (32 - logM))] = nv[q];
        }
        __syncthreads();
        fft_lds(a, M, logM, tw);
        s0 = s1 = nullptr;
    }
    double* q0 = row_ptr(out, r0);
    double* q1 = has1 ? row_ptr(out, r1) : nullptr;
    // entries m = k + 1 in [1, n], lanes on aligned runs of m
    for (int m = threadIdx.x; m <= n; m += DST_T) {
        if (m == 0) continue;
        const int k = m - 1;
        const double2 F = a[sw(m)];
        double o0 = -F.y, o1 = F.x;
        if (s0) o0 *= shifted(s0[k]);
        if (s1) o1 *= shifted(s1[k]);
        if (add) {
            q0[k] += o0;
            if (has1) q1[k] += o1;
        } else {
            q0[k] = o0;
            if (has1) q1[k] = o1;
        }
    }
}

__global__ __launch_bounds__(256) void k_dst_rows_direct(Rows in, Rows out, int n, const double* __restrict__ tab,
                                                         const double* __restrict__ scale) {
    extern __shared__ double row[];
    const int r = blockIdx.x;
    const double* p = row_ptr(in, r);
    for (int j = threadIdx.x; j < n; j += blockDim.x) row[j] = p[j];
    __syncthreads();
    const int N2 = 2 * (n + 1);
    double* q = row_ptr(out, r);
    const double* sc = scale ? scale + long(r % in.rows_per_field) * n : nullptr;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double s = 0.0;
        int idx = k + 1;
        for (int j = 0; j < n; ++j) {
            s += row[j] * tab[idx];
            idx += k + 1;
            if (idx >= N2) idx -= N2;
        }
        double o = 2.0 * s;
        if (sc) o *= sc[k];
        q[k] = o;
    }
}

// out[f][c][r] = in[f][r][c] for f in [0, nf), in: R x C rows
__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ in, double* __restrict__ out, int R,
                                                   int C) {
    __shared__ double t[32][33];
    const long fo = long(blockIdx.z) * R * C;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + threadIdx.x;
        if (r < R && c < C) t[k][threadIdx.x] = in[fo + long(r) * C + c];
    }
    __syncthreads();
    for (int k = threadIdx.y; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + threadIdx.x;
        if (r < R && c < C) out[fo + long(c) * R + r] = t[threadIdx.x][k];
    }
}

// RODFT00 of nrows rows (times the eigenvalue rows when scale != null); sym: followed by
// the inverse transform of the same rows (fused in LDS on the FFT path)
static void dst_rows(hpq_ctx* c, const DstPlan& P, const Rows& in, const Rows& out, int nrows, const double* scale,
                     bool sym = false, const Shift& shift = Shift{}, bool add = false) {
    TraceScope ts(c, HPQ_TRACE_DST);
    if (P.fft) {
        const size_t lds = size_t(P.M) * sizeof(double2);
        if (sym)
            hipLaunchKernelGGL(k_dst_rows_fft<true>, dim3((nrows + 1) / 2), dim3(DST_T), lds, c->stream, in, out,
                               nrows, P.n, P.M, P.logM, P.tw, scale, shift, int(add));
        else
            hipLaunchKernelGGL(k_dst_rows_fft<false>, dim3((nrows + 1) / 2), dim3(DST_T), lds, c->stream, in, out,
                               nrows, P.n, P.M, P.logM, P.tw, scale, shift, int(add));
    } else {
        if (shift.parts || add) throw HpqError(HPQ_E_UNSUPPORTED, "shifted DST solve: n + 1 must be a power of two");
        const size_t lds = size_t(P.n) * sizeof(double);
        hipLaunchKernelGGL(k_dst_rows_direct, dim3(nrows), dim3(256), lds, c->stream, in, out, P.n, P.sintab, scale);
        if (sym) {
            HPQ_HIP(hipGetLastError());
            // the direct path writes in place row by row: transform out again into in
            hipLaunchKernelGGL(k_dst_rows_direct, dim3(nrows), dim3(256), lds, c->stream, out, in, P.n, P.sintab,
                               nullptr);
        }
    }
    HPQ_HIP(hipGetLastError());
}

static double sin_pi_frac(long m, long N) {
    m %= 2 * N;
    if (m < 0) m += 2 * N;
    double sgn = 1.0;
    if (m >= N) { m -= N; sgn = -1.0; }
    if (2 * m > N) m = N - m;
    return sgn * std::sin(M_PI * double(m) / double(N));
}

static void plan_init(DstPlan& P, int n) {
    P.n = n;
    P.N = n + 1;
    P.M = 2 * P.N;
    P.fft = (P.N & (P.N - 1)) == 0 && P.M <= 4096 && P.N >= 2;
    if (P.fft) {
        P.logM = 0;
        while ((1 << P.logM) < P.M) ++P.logM;
        std::vector<double2> tw(P.M / 2);
        for (int k = 0; k < P.M / 2; ++k)
            tw[k] = make_double2(sin_pi_frac(k + P.N / 2, P.N), -sin_pi_frac(k, P.N));
        HPQ_HIP(hipMalloc(&P.tw, sizeof(double2) * tw.size()));
        HPQ_HIP(hipMemcpy(P.tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice));
    } else {
        std::vector<double> tab(2 * P.N);
        for (int m = 0; m < 2 * P.N; ++m) tab[m] = sin_pi_frac(m, P.N);
        HPQ_HIP(hipMalloc(&P.sintab, sizeof(double) * tab.size()));
        HPQ_HIP(hipMemcpy(P.sintab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    }
}

void poisson_init(hpq_ctx* c) {
    Poisson& S = c->ps;
    const int nx = c->g.nx, ny = c->g.ny;
    S.nx = nx;
    S.ny = ny;
    plan_init(S.px, nx);
    plan_init(S.py, ny);
    // eigenvalue matrix (FFTPoissonSolverDirichletDirect.cpp:60-82), stored [i][j]
    std::vector<double> ev(size_t(nx) * ny);
    const double dxsq = c->g.dx * c->g.dx, dysq = c->g.dy * c->g.dy;
    const double sxf = M_PI / (2. * (nx + 1)), syf = M_PI / (2. * (ny + 1));
    const double norm_fac = 0.5 / (2 * ((nx + 1) * (ny + 1)));
    S.inv_norm = 1.0 / norm_fac;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const double sx = std::sin((i + 1) * sxf) * std::sin((i + 1) * sxf);
            const double sy = std::sin((j + 1) * syf) * std::sin((j + 1) * syf);
            double e = 0.0;
            if (sx != 0 && sy != 0) e = norm_fac / (-4.0 * (sx / dxsq + sy / dysq));
            ev[size_t(i) * ny + j] = e;
        }
    HPQ_HIP(hipMalloc(&S.eigT, sizeof(double) * ev.size()));
    HPQ_HIP(hipMemcpy(S.eigT, ev.data(), sizeof(double) * ev.size(), hipMemcpyHostToDevice));
    HPQ_HIP(hipMalloc(&S.bufA, sizeof(double) * 3 * size_t(nx) * ny));
    HPQ_HIP(hipMalloc(&S.bufB, sizeof(double) * 3 * size_t(nx) * ny));
    if (c->p.poisson_solver == HPQ_POISSON_BF16X3) dst_mfma_init(c);
}

void poisson_free(hpq_ctx* c) {
    Poisson& S = c->ps;
    for (void* p : {(void*)S.px.tw, (void*)S.px.sintab, (void*)S.py.tw, (void*)S.py.sintab, (void*)S.eigT,
                    (void*)S.bufA, (void*)S.bufB, (void*)S.chi_part})
        if (p) (void)hipFree(p);
    dst_mfma_free(c);
    S = Poisson{};
}

// Solves nfields staged right-hand sides held in ps.bufA as [f][j][i] and writes the
// valid region of field f into lhs_planes[f] (guarded planes).
// the FFT path of the staged solve; shift (parts of the shift, or none) solves
// (Lap - shift) u = rhs, add adds u to the destination planes
static void solve_fft(hpq_ctx* c, int nf, double* const* lhs_planes, Shift shift = Shift{}, bool add = false) {
    Poisson& S = c->ps;
    const int nx = S.nx, ny = S.ny;
    const long fs = long(nx) * ny;
    auto rows = [&](double* base, long stride, int rpf) {
        Rows R;
        for (int f = 0; f < 3; ++f) R.base[f] = base + f * fs;
        R.stride = stride;
        R.rows_per_field = rpf;
        return R;
    };
    // x pass on rows j
    dst_rows(c, S.px, rows(S.bufA, nx, ny), rows(S.bufB, nx, ny), nf * ny, nullptr);
    hipLaunchKernelGGL(k_transpose, dim3((nx + 31) / 32, (ny + 31) / 32, nf), dim3(32, 8), 0, c->stream, S.bufB,
                       S.bufA, ny, nx);
    // y pass, eigenvalue scaling, inverse y pass on rows i (one launch on the FFT path;
    // the direct path leaves its result in bufA as well)
    if (S.py.fft)
        dst_rows(c, S.py, rows(S.bufA, ny, nx), rows(S.bufA, ny, nx), nf * nx, S.eigT, true, shift);
    else
        dst_rows(c, S.py, rows(S.bufA, ny, nx), rows(S.bufB, ny, nx), nf * nx, S.eigT, true, shift);
    hipLaunchKernelGGL(k_transpose, dim3((ny + 31) / 32, (nx + 31) / 32, nf), dim3(32, 8), 0, c->stream, S.bufA,
                       S.bufB, nx, ny);
    // inverse x pass straight into the destination planes
    Rows out;
    for (int f = 0; f < 3; ++f) out.base[f] = lhs_planes[f < nf ? f : 0] + c->g.idx(0, 0);
    out.stride = c->g.ldx;
    out.rows_per_field = ny;
    dst_rows(c, S.px, rows(S.bufB, nx, ny), out, nf * ny, nullptr, false, Shift{}, add);
}

// Solves nfields staged right-hand sides held in ps.bufA as [f][j][i] and writes the
// valid region of field f into lhs_planes[f] (guarded planes).
void poisson_solve_staged(hpq_ctx* c, int nf, double* const* lhs_planes) {
    if (c->p.poisson_solver == HPQ_POISSON_BF16X3) {
        poisson_solve_mfma(c, nf, lhs_planes);
        return;
    }
    solve_fft(c, nf, lhs_planes);
}

// where the staged right-hand sides go: FP64 [f][j][i] for the FFT path, FP32 padded rows
// for the sine-matrix path
template <class F>
void with_staging(hpq_ctx* c, F&& f) {
    const Poisson& S = c->ps;
    if (c->p.poisson_solver == HPQ_POISSON_BF16X3)
        f(S.gS, long(S.npx), long(S.npx) * S.npy);
    else
        f(S.bufA, long(S.nx), long(S.nx) * S.ny);
}

// copies a caller's nx x ny FP64 right-hand side into the staging of field 0
template <class T>
__global__ void k_stage_rhs(Geom g, const double* __restrict__ src, T* __restrict__ st, long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    st[long(j) * ld + i] = T(src[long(j) * g.nx + i]);
}

void stage_rhs(hpq_ctx* c, const double* src) {
    const Geom& g = c->g;
    with_staging(c, [&](auto* st, long ld, long) {
        using T = std::remove_pointer_t<decltype(st)>;
        hipLaunchKernelGGL(k_stage_rhs<T>, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g, src, st, ld);
    });
    HPQ_HIP(hipGetLastError());
}

// Right-hand sides of Fields::SolvePoissonPsiExmByEypBxEzBz (src/fields/Fields.cpp:886-917):
// Psi: -rhomjz/ep0; Ez: (d_x jx + d_y jy)/(ep0 c); Bz: mu0 (d_y jx - d_x jy), centred
// differences (derivative<>, :223-277) that read the guard cells of jx, jy.
// st: [f][j][i] with row pitch ld and field stride fs, FP64 (FFT path) or FP32 (bf16x3 path)
template <class T>
__global__ void k_rhs_psi_ez_bz(Geom g, const double* __restrict__ rhomjz, const double* __restrict__ jx,
                                const double* __restrict__ jy, T* __restrict__ st, long ld, long fs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double ep0 = 1.0, mu0 = 1.0, cl = 1.0;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const long o = long(j) * ld + i;
    st[o] = T((-1.0 / ep0) * rhomjz[g.idx(i, j)]);
    const double fe = 1.0 / (ep0 * cl);
    st[fs + o] = T(fe * ((jx[g.idx(i + 1, j)] - jx[g.idx(i - 1, j)]) * dxih) +
                   fe * ((jy[g.idx(i, j + 1)] - jy[g.idx(i, j - 1)]) * dyih));
    st[2 * fs + o] = T(mu0 * ((jx[g.idx(i, j + 1)] - jx[g.idx(i, j - 1)]) * dyih) +
                       (-mu0) * ((jy[g.idx(i + 1, j)] - jy[g.idx(i - 1, j)]) * dxih));
}

void build_rhs_psi_ez_bz(hpq_ctx* c, const double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    with_staging(c, [&](auto* st, long ld, long fs) {
        using T = std::remove_pointer_t<decltype(st)>;
        hipLaunchKernelGGL(k_rhs_psi_ez_bz<T>, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                           store + L.rhomjz * g.plane, store + L.jx * g.plane, store + L.jy * g.plane, st, ld, fs);
    });
    HPQ_HIP(hipGetLastError());
}

// Right-hand side of Fields::SolvePoissonEz (src/fields/Fields.cpp:988-993) into the staging of
// field 0: (d_x jx + d_y jy)/(ep0 c), the Ez term of k_rhs_psi_ez_bz on other planes
template <class T>
__global__ void k_rhs_ez(Geom g, const double* __restrict__ jx, const double* __restrict__ jy, T* __restrict__ st,
                         long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double ep0 = 1.0, cl = 1.0;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const double fe = 1.0 / (ep0 * cl);
    st[long(j) * ld + i] = T(fe * ((jx[g.idx(i + 1, j)] - jx[g.idx(i - 1, j)]) * dxih) +
                             fe * ((jy[g.idx(i, j + 1)] - jy[g.idx(i, j - 1)]) * dyih));
}

// SolvePoissonEz on the SALAME planes: RHS from their jx, jy
void salame_rhs_ez(hpq_ctx* c, const double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    with_staging(c, [&](auto* st, long ld, long) {
        using T = std::remove_pointer_t<decltype(st)>;
        hipLaunchKernelGGL(k_rhs_ez<T>, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                           store + L.sal_jx * g.plane, store + L.sal_jy * g.plane, st, ld);
    });
    HPQ_HIP(hipGetLastError());
}

// Right-hand sides of Fields::SolvePoissonBxBy (src/fields/Fields.cpp:1044-1068).
template <class T>
__global__ void k_rhs_bxby(Geom g, const double* __restrict__ jz, const double* __restrict__ pjx,
                           const double* __restrict__ pjy, const double* __restrict__ njx,
                           const double* __restrict__ njy, T* __restrict__ st, long ld, long fs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double mu0 = 1.0;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi, dzih = 0.5 * g.dzi;
    const long o = long(j) * ld + i;
    const long k = g.idx(i, j);
    st[o] = T(-mu0 * ((jz[g.idx(i, j + 1)] - jz[g.idx(i, j - 1)]) * dyih) + mu0 * ((pjy[k] - njy[k]) * dzih));
    st[fs + o] =
        T(mu0 * ((jz[g.idx(i + 1, j)] - jz[g.idx(i - 1, j)]) * dxih) + (-mu0) * ((pjx[k] - njx[k]) * dzih));
}

void build_rhs_bxby(hpq_ctx* c, const double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    with_staging(c, [&](auto* st, long ld, long fs) {
        using T = std::remove_pointer_t<decltype(st)>;
        hipLaunchKernelGGL(k_rhs_bxby<T>, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                           store + L.jz * g.plane, store + L.prev_jx * g.plane, store + L.prev_jy * g.plane,
                           store + L.next_jx * g.plane, store + L.next_jy * g.plane, st, ld, fs);
    });
    HPQ_HIP(hipGetLastError());
}

// ExmBy = -d_x Psi, EypBx = -d_y Psi on the valid box plus one guard ring (Fields.cpp:931-956)
__global__ void k_exmby_eypbx(Geom g, const double* __restrict__ psi, double* __restrict__ exmby,
                              double* __restrict__ eypbx) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x) - 1;
    const int j = int(blockIdx.y) - 1;
    if (i > g.nx) return;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    exmby[g.idx(i, j)] = -(psi[g.idx(i + 1, j)] - psi[g.idx(i - 1, j)]) * dxih;
    eypbx[g.idx(i, j)] = -(psi[g.idx(i, j + 1)] - psi[g.idx(i, j - 1)]) * dyih;
}

void exmby_eypbx(hpq_ctx* c, double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    hipLaunchKernelGGL(k_exmby_eypbx, dim3((g.nx + 2 + 127) / 128, g.ny + 2), dim3(128), 0, c->stream, g,
                       store + L.Psi * g.plane, store + L.ExmBy * g.plane, store + L.EypBx * g.plane);
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq

// ---------------------------------------------------------------------------------------
// DST-preconditioned initial guess of the explicit Bx/By multigrid (HPQ_MG_GUESS_DST)
// ---------------------------------------------------------------------------------------
namespace hpq {
namespace {

constexpr int RES_T = 256;   // threads of the residual kernel, one row per workgroup

// residual of the explicit Bx/By system, (Sy, Sx) - (Lap - chi) (Bx, By) on the valid
// cells (the guard cells of Bx, By hold the homogeneous Dirichlet values), staged as the
// right-hand sides [f][j][i] of the shifted DST solve. The workgroups of component 0 also
// write the sum of chi over their row (a fixed order) to part[j]; the shifted solve adds
// the rows' sums.
__global__ __launch_bounds__(RES_T) void k_bxby_residual(Geom g, const double* __restrict__ bx,
                                                         const double* __restrict__ sy,
                                                         const double* __restrict__ chi, double* __restrict__ st,
                                                         double* __restrict__ part) {
    __shared__ double red[RES_T / 64];
    const int j = blockIdx.x, f = blockIdx.y;
    const double* b = bx + long(f) * g.plane;
    const double* s = sy + long(f) * g.plane;
    const double fx = 1. / (g.dx * g.dx), fy = 1. / (g.dy * g.dy);
    double cs = 0.0;
    for (int i = threadIdx.x; i < g.nx; i += RES_T) {
        const long k = g.idx(i, j);
        const double c = b[k], cv = chi[k];
        cs += cv;
        const double lap = fx * (b[k - 1] + b[k + 1] - 2. * c) + fy * (b[k - g.ldx] + b[k + g.ldx] - 2. * c);
        st[long(f) * g.nx * g.ny + long(j) * g.nx + i] = s[k] - (lap - cv * c);
    }
    if (f != 0) return;
    for (int o = 32; o > 0; o >>= 1) cs += __shfl_xor(cs, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cs;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < RES_T / 64; ++w) t += red[w];
        part[j] = t;
    }
}

}  // namespace

void bxby_guess_dst(hpq_ctx* c, double* store, int iters) {
    Poisson& S = c->ps;
    const Geom& g = c->g;
    const Layout& L = c->L;
    if (!S.px.fft || !S.py.fft) throw HpqError(HPQ_E_UNSUPPORTED, "mg_guess DST: nx + 1 and ny + 1 must be powers of two");
    if (L.By != L.Bx + 1 || L.Sx != L.Sy + 1) throw HpqError(HPQ_E_UNSUPPORTED, "mg_guess DST: Bx, By / Sy, Sx not adjacent");
    if (!S.chi_part) HPQ_HIP(hipMalloc(&S.chi_part, sizeof(double) * g.ny));
    const double* chi = store + long(L.chi) * g.plane;
    double* bx = store + long(L.Bx) * g.plane;
    double* const lhs[3] = {bx, bx + g.plane, nullptr};
    Shift sh;
    sh.parts = S.chi_part;
    sh.n = g.ny;
    sh.inv_count = 1.0 / (double(g.nx) * g.ny);
    sh.inv_f = S.inv_norm;
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_bxby_residual, dim3(g.ny, 2), dim3(RES_T), 0, c->stream, g, bx,
                           store + long(L.Sy) * g.plane, chi, S.bufA, S.chi_part);
        HPQ_HIP(hipGetLastError());
        solve_fft(c, 2, lhs, sh, true);
    }
}

}  // namespace hpq
