// Dirichlet Poisson solve with DST-I as dense sine-matrix products on the gfx950 matrix
// cores (poisson_solver = HPQ_POISSON_BF16X3).
//
// The solve is F = Sy (L o (Sy X Sx)) Sx with S_jk = 2 sin(pi (j+1)(k+1) / (n+1)) (the RODFT00
// normalisation) and L the normalised eigenvalue matrix (the same operator as
// FFTPoissonSolverDirichletDirect,
// src/fields/fft_poisson_solver/FFTPoissonSolverDirichletDirect.cpp:60-139). S is
// symmetric, so one kernel shape serves all four products: out = (in * S)^T, i.e. every
// row of `in` is transformed and the result is written transposed; two applications give
// S X S. Each FP32 operand value v is split into v_hi = bf16(v) and v_lo = bf16(v - v_hi),
// and a product uses the three terms hi*hi + hi*lo + lo*hi with FP32 accumulation on
// v_mfma_f32_16x16x32_bf16 (the lo*lo term is below FP32 resolution of the sum).
// Operands are zero-padded to multiples of 64: intermediates live in padded FP32 buffers
// whose padding the products keep at zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "internal.hpp"

namespace hpq {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;        // workgroup tile: 64 rows of `in` x 64 columns of S
constexpr int GK = 64;        // K chunk staged in LDS
constexpr int PK = GK + 8;    // LDS row pitch in bf16 (16 B of padding against bank conflicts)
constexpr int GEMM_T = 256;   // 4 wavefronts, each a 32 x 32 sub-tile

template <class T>
struct Ptr3 { T* p[3]; };

struct GemmArgs {
    long lda;          // row pitch of `in` (a multiple of 4)
    int kpad;          // K extent (a multiple of GK)
    const __bf16* s_hi;  // S split, [j][k] with pitch kpad (S is symmetric)
    const __bf16* s_lo;
    long ldo;          // row pitch of `out`
    int out_rows;      // FP64 output: rows of `out` written (columns j of the product)
    int out_cols;      // FP64 output: columns of `out` written (rows i of `in`)
    const double* eig; // optional scale: out(r, c) *= eig[c * eig_ld + r] for r < eig_rows,
    long eig_ld;       // c < eig_cols, and 0 outside (keeps the padding of `out` zero)
    int eig_rows, eig_cols;
};

__device__ __forceinline__ void split(float v, __bf16& hi, __bf16& lo) {
    hi = __bf16(v);
    lo = __bf16(v - float(hi));
}

// One workgroup computes a 64 x 64 tile of (in * S) over the whole K extent, K chunks of 64
// double-buffered in LDS (one barrier per chunk). Inputs are FP32, padded to multiples of
// 64 with zeros outside the data, rows 16-byte aligned: vector loads without bounds.
template <class TOUT, bool SCALE>
__global__ __launch_bounds__(GEMM_T) void k_dst_gemm(Ptr3<const float> in, Ptr3<TOUT> out, GemmArgs a) {
    constexpr int STAGE = 4 * GT * PK;   // ah, al, bh, bl of one K chunk
    __shared__ __attribute__((aligned(16))) __bf16 lds[2 * STAGE];
    const int f = blockIdx.z;
    const float* A = in.p[f];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.y * GT;   // rows of `in` (output columns)
    const int n0 = blockIdx.x * GT;   // columns of S (output rows)

    // staging: `in` as 4 consecutive k per thread over 4 row groups; S as 8 consecutive k.
    // Chunk ck+1 is loaded into registers while chunk ck is multiplied, then split into the
    // other LDS buffer.
    const int ar = t >> 4, ak = (t & 15) * 4;
    const int br = t >> 3, bk = (t & 7) * 8;
    struct Regs {
        f32x4 av[4];
        bf16x8 bvh[2], bvl[2];
    };
    Regs R;
    auto load = [&](int kc, Regs& R) {
#pragma unroll
        for (int it = 0; it < 4; ++it)
            R.av[it] = *reinterpret_cast<const f32x4*>(A + long(m0 + ar + 16 * it) * a.lda + kc + ak);
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const long o = long(n0 + br + 32 * it) * a.kpad + kc + bk;
            R.bvh[it] = *reinterpret_cast<const bf16x8*>(a.s_hi + o);
            R.bvl[it] = *reinterpret_cast<const bf16x8*>(a.s_lo + o);
        }
    };
    auto stage = [&](const Regs& R, __bf16* st) {
        __bf16 *ah = st, *al = st + GT * PK, *bh = st + 2 * GT * PK, *bl = st + 3 * GT * PK;
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            bf16x4 h, l;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                __bf16 x, y;
                split(R.av[it][e], x, y);
                h[e] = x;
                l[e] = y;
            }
            const int o = (ar + 16 * it) * PK + ak;
            *reinterpret_cast<bf16x4*>(ah + o) = h;
            *reinterpret_cast<bf16x4*>(al + o) = l;
        }
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int o = (br + 32 * it) * PK + bk;
            *reinterpret_cast<bf16x8*>(bh + o) = R.bvh[it];
            *reinterpret_cast<bf16x8*>(bl + o) = R.bvl[it];
        }
    };

    f32x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};

    // fragment of v_mfma_f32_16x16x32_bf16: lane l holds row l%16, k = 8*(l/16) .. +7
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    auto multiply = [&](const __bf16* st) {
        const __bf16 *ah = st, *al = st + GT * PK, *bh = st + 2 * GT * PK, *bl = st + 3 * GT * PK;
#pragma unroll
        for (int kk = 0; kk < GK; kk += 32) {
            bf16x8 xh[2], xl[2], yh[2], yl[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const int o = (wm * 32 + x * 16 + fr) * PK + kk + fk;
                xh[x] = *reinterpret_cast<const bf16x8*>(ah + o);
                xl[x] = *reinterpret_cast<const bf16x8*>(al + o);
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int o = (wn * 32 + y * 16 + fr) * PK + kk + fk;
                yh[y] = *reinterpret_cast<const bf16x8*>(bh + o);
                yl[y] = *reinterpret_cast<const bf16x8*>(bl + o);
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[x], yh[y], acc[x][y], 0, 0, 0);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[x], yl[y], acc[x][y], 0, 0, 0);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[x], yh[y], acc[x][y], 0, 0, 0);
        }
    };
    // the other LDS buffer was last read in chunk ck-1, before the barrier that ended it
    const int nchunk = a.kpad / GK;
    load(0, R);
    stage(R, lds);
    __syncthreads();
    for (int ck = 0; ck < nchunk; ++ck) {
        if (ck + 1 < nchunk) load((ck + 1) * GK, R);
        multiply(lds + (ck & 1) * STAGE);
        if (ck + 1 < nchunk) stage(R, lds + ((ck + 1) & 1) * STAGE);
        __syncthreads();
    }

    // D fragment: lane l holds rows 4*(l/16) + r (r = 0..3) of column l%16; stored transposed,
    // the four rows are four consecutive output columns
    TOUT* O = out.p[f];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int orow = n0 + wn * 32 + y * 16 + fr;
            const int ocol = m0 + wm * 32 + x * 16 + 4 * (lane >> 4);
            f32x4 v = acc[x][y];
            if constexpr (SCALE) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int c = ocol + r;
                    v[r] = (orow < a.eig_rows && c < a.eig_cols) ? float(double(v[r]) * a.eig[long(c) * a.eig_ld + orow])
                                                                 : 0.0f;
                }
            }
            if constexpr (std::is_same_v<TOUT, float>) {
                // padded FP32 intermediate: the whole tile is inside the buffer
                *reinterpret_cast<f32x4*>(O + long(orow) * a.ldo + ocol) = v;
            } else {
                if (orow >= a.out_rows) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (ocol + r < a.out_cols) O[long(orow) * a.ldo + ocol + r] = TOUT(v[r]);
            }
        }
}

uint16_t bf16_bits(float v) {   // round to nearest even (finite values)
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

float bf16_value(uint16_t b) {
    const uint32_t u = uint32_t(b) << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

int pad64(int n) { return (n + GT - 1) / GT * GT; }

// S_jk = 2 sin(pi (j+1)(k+1) / (n+1)) split into hi and lo BF16 planes, zero outside n x n
void sine_split(int n, int np, __bf16** hi, __bf16** lo) {
    std::vector<uint16_t> h(size_t(np) * np, 0), l(size_t(np) * np, 0);
    const long N = n + 1;
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
            long m = (long(j + 1) * (k + 1)) % (2 * N);
            double sgn = 1.0;
            if (m >= N) { m -= N; sgn = -1.0; }
            if (2 * m > N) m = N - m;
            const double s = 2.0 * sgn * std::sin(M_PI * double(m) / double(N));
            const float sf = float(s);
            const uint16_t bh = bf16_bits(sf);
            h[size_t(j) * np + k] = bh;
            l[size_t(j) * np + k] = bf16_bits(float(s - double(bf16_value(bh))));
        }
    HPQ_HIP(hipMalloc(hi, sizeof(uint16_t) * h.size()));
    HPQ_HIP(hipMalloc(lo, sizeof(uint16_t) * l.size()));
    HPQ_HIP(hipMemcpy(*hi, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice));
    HPQ_HIP(hipMemcpy(*lo, l.data(), sizeof(uint16_t) * l.size(), hipMemcpyHostToDevice));
}

template <class TOUT, bool SCALE>
void gemm(hpq_ctx* c, int nf, Ptr3<const float> in, Ptr3<TOUT> out, const GemmArgs& a, int mpad) {
    const dim3 grid(a.kpad / GT, mpad / GT, nf);
    hipLaunchKernelGGL((k_dst_gemm<TOUT, SCALE>), grid, dim3(GEMM_T), 0, c->stream, in, out, a);
    HPQ_HIP(hipGetLastError());
}

}  // namespace

void dst_mfma_init(hpq_ctx* c) {
    Poisson& S = c->ps;
    S.npx = pad64(S.nx);
    S.npy = pad64(S.ny);
    sine_split(S.nx, S.npx, &S.sx_hi, &S.sx_lo);
    sine_split(S.ny, S.npy, &S.sy_hi, &S.sy_lo);
    const size_t fs = size_t(S.npx) * S.npy;
    HPQ_HIP(hipMalloc(&S.gA, sizeof(float) * 3 * fs));
    HPQ_HIP(hipMalloc(&S.gB, sizeof(float) * 3 * fs));
    HPQ_HIP(hipMalloc(&S.gS, sizeof(float) * 3 * fs));
    // the padding of the staged right-hand sides is never written: zero once
    HPQ_HIP(hipMemset(S.gS, 0, sizeof(float) * 3 * fs));
}

void dst_mfma_free(hpq_ctx* c) {
    Poisson& S = c->ps;
    for (void* p : {(void*)S.sx_hi, (void*)S.sx_lo, (void*)S.sy_hi, (void*)S.sy_lo, (void*)S.gA, (void*)S.gB,
                    (void*)S.gS})
        if (p) (void)hipFree(p);
    S.sx_hi = S.sx_lo = S.sy_hi = S.sy_lo = nullptr;
    S.gA = S.gB = S.gS = nullptr;
}

// The staged right-hand sides ps.gS ([f][npy][npx], FP32, zero padding) through four
// products into the valid region of lhs_planes[f].
void poisson_solve_mfma(hpq_ctx* c, int nf, double* const* lhs_planes) {
    Poisson& S = c->ps;
    const int nx = S.nx, ny = S.ny, npx = S.npx, npy = S.npy;
    const long fs = long(npx) * npy;
    TraceScope ts(c, HPQ_TRACE_DST);
    auto fbuf = [&](float* b) {
        Ptr3<float> p;
        for (int f = 0; f < 3; ++f) p.p[f] = b + f * fs;
        return p;
    };
    auto cfbuf = [&](const float* b) {
        Ptr3<const float> p;
        for (int f = 0; f < 3; ++f) p.p[f] = b + f * fs;
        return p;
    };
    // 1: rows j of X times Sx -> P1 [i][j]
    {
        GemmArgs a{npx, npx, S.sx_hi, S.sx_lo, npy, 0, 0, nullptr, 0, 0, 0};
        gemm<float, false>(c, nf, cfbuf(S.gS), fbuf(S.gA), a, npy);
    }
    // 2: rows i of P1 times Sy -> P2 [j][i] = Sy X Sx, scaled by the eigenvalues (ev[i][j]),
    // zero in the padding
    {
        GemmArgs a{npy, npy, S.sy_hi, S.sy_lo, npx, 0, 0, S.eigT, ny, ny, nx};
        gemm<float, true>(c, nf, cfbuf(S.gA), fbuf(S.gB), a, npx);
    }
    // 3: rows j of P2 times Sx -> P3 [i][j]
    {
        GemmArgs a{npx, npx, S.sx_hi, S.sx_lo, npy, 0, 0, nullptr, 0, 0, 0};
        gemm<float, false>(c, nf, cfbuf(S.gB), fbuf(S.gA), a, npy);
    }
    // 4: rows i of P3 times Sy -> F [j][i] into the guarded planes
    {
        Ptr3<double> out;
        for (int f = 0; f < 3; ++f) out.p[f] = lhs_planes[f < nf ? f : 0] + c->g.idx(0, 0);
        GemmArgs a{npy, npy, S.sy_hi, S.sy_lo, c->g.ldx, ny, nx, nullptr, 0, 0, 0};
        gemm<double, false>(c, nf, cfbuf(S.gA), out, a, npx);
    }
}

}  // namespace hpq
