// Kernels of SALAME, the reference's beam-loading algorithm (src/salame/Salame.cpp):
//   SalameInitializeSxSyWithBeam (:187-219), SalameGetJxJyFromBxBy (:222-256),
//   SalameOnlyAdvancePlasma (:258-330), SalameGetW (:332-407), SalameMultiplyBeamWeight (:409-437),
// and the count of a slice's SALAME particles (MultiBeam::isSalameNow,
// src/particles/beam/MultiBeam.cpp:128-140). The iteration itself is in hipace.cpp.
#include "internal.hpp"
#include "shape.hpp"

namespace hpq {

namespace {

constexpr int ST = 256;

__device__ __forceinline__ bool in_plane(const Geom& g, int i, int j) {
    return i >= -g.ng && i < g.nx + g.ng && j >= -g.ng && j < g.ny + g.ng;
}

// one workgroup: number of particles in [first, first + n) with mask 1
__global__ __launch_bounds__(ST) void k_sal_count(const int32_t* __restrict__ mask, long first, long n,
                                                  double* __restrict__ out) {
    __shared__ double red[ST / 64];
    double v = 0.0;
    for (long t = threadIdx.x; t < n; t += ST) v += mask[first + t] == 1 ? 1.0 : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[4] = red[0] + red[1] + red[2] + red[3];
}

// Salame.cpp:203-216: Sy = mu0 (-d_y jz_beam), Sx = -mu0 (-d_x jz_beam) on the SALAME planes
__global__ void k_sal_sxsy(Geom g, const double* __restrict__ jzb, double* __restrict__ Sx, double* __restrict__ Sy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double mu0 = 1.0;
    const double dxih = 0.5 * g.dxi, dyih = 0.5 * g.dyi;
    const double dx_jzb = (jzb[g.idx(i + 1, j)] - jzb[g.idx(i - 1, j)]) * dxih;
    const double dy_jzb = (jzb[g.idx(i, j + 1)] - jzb[g.idx(i, j - 1)]) * dyih;
    const long k = g.idx(i, j);
    Sy[k] = mu0 * (-dy_jzb);
    Sx[k] = -mu0 * (-dx_jzb);
}

// Salame.cpp:246-253: jx = 1.5 dz chi By / mu0, jy = -1.5 dz chi Bx / mu0
__global__ void k_sal_jxjy(Geom g, const double* __restrict__ chi, const double* __restrict__ Bx,
                           const double* __restrict__ By, double* __restrict__ jx, double* __restrict__ jy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double dz = 1.5 * g.dz;
    const double mu0_inv = 1.0 / 1.0;
    const long k = g.idx(i, j);
    jx[k] = dz * chi[k] * By[k] * mu0_inv;
    jy[k] = -dz * chi[k] * Bx[k] * mu0_inv;
}

// Salame.cpp:293-323: ux, uy of every valid plasma particle from the SALAME-only Bx, By
// gathered at (x_prev, y_prev) (doBxByGatherShapeN, FieldGather.H:183-216)
template <int O>
__global__ __launch_bounds__(ST) void k_sal_only_advance(Geom g, hpq_plasma p, const double* __restrict__ Bx,
                                                         const double* __restrict__ By, double qm) {
    const long ip = long(blockIdx.x) * ST + threadIdx.x;
    if (ip >= p.n || !p.valid[ip]) return;
    const double xmid = (p.x_prev[ip] - g.xoff) * g.dxi;
    const double ymid = (p.y_prev[ip] - g.yoff) * g.dyi;
    double Bxp = 0.0, Byp = 0.0;
#pragma unroll
    for (int iy = 0; iy <= O; ++iy)
#pragma unroll
        for (int ix = 0; ix <= O; ++ix) {
            const Sf sx = shape<O>(xmid, ix), sy = shape<O>(ymid, iy);
            if (!in_plane(g, sx.cell, sy.cell)) continue;
            const long k = g.idx(sx.cell, sy.cell);
            Bxp += sx.s * sy.s * Bx[k];
            Byp += sx.s * sy.s * By[k];
        }
    const double dz = g.dz;
    p.ux[ip] = 1.5 * dz * qm * Byp;
    p.uy[ip] = -1.5 * dz * qm * Bxp;
}

// Salame.cpp:357-374: the four sums of SalameGetW over the valid cells, in one workgroup
// (a fixed order), into out[0..3]
__global__ __launch_bounds__(1024) void k_sal_sums(Geom g, const double* __restrict__ jz,
                                                   const double* __restrict__ ez_target,
                                                   const double* __restrict__ ez_no, const double* __restrict__ ez,
                                                   double* __restrict__ out) {
    __shared__ double red[1024 / 64][4];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    const double factor = g.dx * g.dy / (g.dx * g.dy);
    const long n = long(g.nx) * g.ny;
    for (long t = threadIdx.x; t < n; t += 1024) {
        const int j = int(t / g.nx), i = int(t - long(j) * g.nx);
        const long k = g.idx(i, j);
        a[0] += factor * jz[k] * ez_target[k];
        a[1] += factor * jz[k] * ez_no[k];
        a[2] += factor * jz[k] * ez[k];
        a[3] += factor * jz[k];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double v = a[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = red[0][threadIdx.x];
        for (int w = 1; w < 1024 / 64; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// Salame.cpp:421-434: SALAME particles of the slice: W = 0 invalidates, otherwise w *= W
__global__ void k_sal_multiply(const int32_t* __restrict__ mask, double* __restrict__ w, int32_t* __restrict__ valid,
                               long first, long n, double W) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long ip = first + t;
    if (mask[ip] != 1 || !valid[ip]) return;
    if (W == 0) {
        valid[ip] = 0;
        return;
    }
    w[ip] *= W;
}

}  // namespace

void salame_alloc(hpq_ctx* c) {
    if (c->sal_dev) return;
    HPQ_HIP(hipMalloc(&c->sal_dev, 8 * sizeof(double)));
    HPQ_HIP(hipHostMalloc(&c->sal_host, 8 * sizeof(double), hipHostMallocDefault));
}

void salame_free(hpq_ctx* c) {
    if (c->sal_dev) (void)hipFree(c->sal_dev);
    if (c->sal_host) (void)hipHostFree(c->sal_host);
    c->sal_dev = nullptr;
    c->sal_host = nullptr;
}

long salame_count(hpq_ctx* c, const hpq_beam* b, int slice) {
    if (!b || b->n <= 0 || !b->off) return 0;
    const long first = long(b->off[slice]), n = long(b->off[slice + 1]) - first;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_sal_count, dim3(1), dim3(ST), 0, c->stream, c->salame.mask, first, n, c->sal_dev);
    HPQ_HIP(hipGetLastError());
    HPQ_HIP(hipMemcpyAsync(c->sal_host + 4, c->sal_dev + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPQ_HIP(hipStreamSynchronize(c->stream));
    return long(c->sal_host[4]);
}

void salame_sxsy(hpq_ctx* c, double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    hipLaunchKernelGGL(k_sal_sxsy, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                       store + L.sal_jz_beam * g.plane, store + L.sal_Sx * g.plane, store + L.sal_Sy * g.plane);
    HPQ_HIP(hipGetLastError());
}

void salame_jxjy_from_b(hpq_ctx* c, double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    hipLaunchKernelGGL(k_sal_jxjy, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                       store + L.chi * g.plane, store + L.sal_Bx * g.plane, store + L.sal_By * g.plane,
                       store + L.sal_jx * g.plane, store + L.sal_jy * g.plane);
    HPQ_HIP(hipGetLastError());
}

void salame_only_advance(hpq_ctx* c, const hpq_plasma* p, double* store) {
    c->cl_plasma.valid = false;
    if (p->n <= 0) return;
    const Geom& g = c->g;
    const Layout& L = c->L;
    const double qm = c->p.plasma_charge / c->p.plasma_mass;
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_sal_only_advance<decltype(O)::value>), grid1d(p->n, ST), dim3(ST), 0, c->stream, g, *p,
                           store + L.sal_Bx * g.plane, store + L.sal_By * g.plane, qm);
    });
    HPQ_HIP(hipGetLastError());
}

void salame_sums(hpq_ctx* c, const double* store, double out[4]) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    hipLaunchKernelGGL(k_sal_sums, dim3(1), dim3(1024), 0, c->stream, g, store + L.sal_jz_beam * g.plane,
                       store + L.sal_Ez_target * g.plane, store + L.sal_Ez_no_salame * g.plane,
                       store + L.sal_Ez * g.plane, c->sal_dev);
    HPQ_HIP(hipGetLastError());
    HPQ_HIP(hipMemcpyAsync(c->sal_host, c->sal_dev, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPQ_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) out[k] = c->sal_host[k];
}

void salame_multiply(hpq_ctx* c, const hpq_beam* b, int slice, double W) {
    if (!b || b->n <= 0 || !b->off) return;
    const long first = long(b->off[slice]), n = long(b->off[slice + 1]) - first;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sal_multiply, grid1d(n, 256), dim3(256), 0, c->stream, c->salame.mask, c->salame.w,
                       c->salame.valid, first, n, W);
    HPQ_HIP(hipGetLastError());
}

}  // namespace hpq
