// Slice-store bookkeeping kernels: fill/copy/add/linear combination of planes
// (Fields::setVal/duplicate/add, MultiFab::LinComb; src/fields/Fields.H:143-287),
// Sx/Sy from the beam (Hipace::InitializeSxSyWithBeam, src/Hipace.cpp:745-790),
// the relative B-field error of the predictor-corrector loop
// (Fields::ComputeRelBFieldError, src/fields/Fields.cpp:1229-1286) and the per-slice
// field diagnostics (sum |F| over valid cells, tests/checksum/backend/openpmd_backend.py).
// Up to 8 planes are handled per launch so a slice-store update is one kernel.
#include "internal.hpp"

namespace hpq {

namespace {

struct PlaneSet {
    int n;
    int dst[8];
    int src[8];
    double v;
};

__global__ void k_fill(double* __restrict__ store, long plane, PlaneSet s) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    for (int k = 0; k < s.n; ++k) store[long(s.dst[k]) * plane + t] = s.v;
}

__global__ void k_copy(double* __restrict__ store, long plane, PlaneSet s) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    double v[8];
    for (int k = 0; k < s.n; ++k) v[k] = store[long(s.src[k]) * plane + t];
    for (int k = 0; k < s.n; ++k) store[long(s.dst[k]) * plane + t] = v[k];
}

__global__ void k_add(double* __restrict__ store, long plane, int dst, int src) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    store[long(dst) * plane + t] += store[long(src) * plane + t];
}

__global__ void k_lincomb(double* __restrict__ store, long plane, int dst, double a, int sa, double b, int sb) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    store[long(dst) * plane + t] = a * store[long(sa) * plane + t] + b * store[long(sb) * plane + t];
}

__global__ void k_sxsy(Geom g, const double* __restrict__ jzb, const double* __restrict__ pjx,
                       const double* __restrict__ pjy, const double* __restrict__ njx, const double* __restrict__ njy,
                       double* __restrict__ Sx, double* __restrict__ Sy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    const double mu0 = 1.0;
    const long k = g.idx(i, j);
    const double dx_jzb = (jzb[g.idx(i + 1, j)] - jzb[g.idx(i - 1, j)]) / (2.0 * g.dx);
    const double dy_jzb = (jzb[g.idx(i, j + 1)] - jzb[g.idx(i, j - 1)]) / (2.0 * g.dy);
    const double dz_jxb = (pjx[k] - njx[k]) / (2.0 * g.dz);
    const double dz_jyb = (pjy[k] - njy[k]) / (2.0 * g.dz);
    Sy[k] = mu0 * (-dy_jzb + dz_jyb);
    Sx[k] = -mu0 * (-dx_jzb + dz_jxb);
}

// sum |F| over valid cells of every component: one block per (component, 8 rows)
__global__ void k_diag(Geom g, const double* __restrict__ store, double* __restrict__ sums) {
    __shared__ double red[256];
    const int comp = blockIdx.y;
    const double* P = store + long(comp) * g.plane;
    double acc = 0.0;
    const int j0 = blockIdx.x * 8;
    for (int jj = 0; jj < 8; ++jj) {
        const int j = j0 + jj;
        if (j >= g.ny) break;
        for (int i = threadIdx.x; i < g.nx; i += blockDim.x) acc += fabs(P[g.idx(i, j)]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(sums + comp, red[0]);
}

__global__ void k_capture(Geom g, const double* __restrict__ P, double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    dst[long(j) * g.nx + i] = P[g.idx(i, j)];
}

// partial sums of |B| and |B - B_iter| over valid cells
__global__ void k_relb(Geom g, const double* __restrict__ Bx, const double* __restrict__ By,
                       const double* __restrict__ Bxi, const double* __restrict__ Byi, double* __restrict__ out) {
    __shared__ double r0[256], r1[256];
    double a = 0.0, b = 0.0;
    const long n = long(g.nx) * g.ny;
    for (long t = long(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += long(gridDim.x) * blockDim.x) {
        const int j = int(t / g.nx), i = int(t - long(j) * g.nx);
        const long k = g.idx(i, j);
        const double factor = g.dx * g.dy / (g.dx * g.dy);
        a += factor * sqrt(Bx[k] * Bx[k] + By[k] * By[k]);
        b += factor * sqrt((Bx[k] - Bxi[k]) * (Bx[k] - Bxi[k]) + (By[k] - Byi[k]) * (By[k] - Byi[k]));
    }
    r0[threadIdx.x] = a;
    r1[threadIdx.x] = b;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(out, r0[0]);
        atomicAdd(out + 1, r1[0]);
    }
}

}  // namespace

void plane_fill(hpq_ctx* c, double* store, const int* comps, int ncomps, double v) {
    PlaneSet s{};
    s.v = v;
    for (int k = 0; k < ncomps; ++k)
        if (comps[k] >= 0) s.dst[s.n++] = comps[k];
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_fill, grid1d(c->g.plane, 256), dim3(256), 0, c->stream, store, c->g.plane, s);
    HPQ_HIP(hipGetLastError());
}

void plane_copy(hpq_ctx* c, double* store, const int* dst, const int* src, int ncomps) {
    PlaneSet s{};
    for (int k = 0; k < ncomps; ++k) {
        s.dst[s.n] = dst[k];
        s.src[s.n] = src[k];
        ++s.n;
    }
    hipLaunchKernelGGL(k_copy, grid1d(c->g.plane, 256), dim3(256), 0, c->stream, store, c->g.plane, s);
    HPQ_HIP(hipGetLastError());
}

void plane_add(hpq_ctx* c, double* store, int dst, int src) {
    hipLaunchKernelGGL(k_add, grid1d(c->g.plane, 256), dim3(256), 0, c->stream, store, c->g.plane, dst, src);
    HPQ_HIP(hipGetLastError());
}

void plane_lincomb(hpq_ctx* c, double* store, int dst, double a, int sa, double b, int sb) {
    hipLaunchKernelGGL(k_lincomb, grid1d(c->g.plane, 256), dim3(256), 0, c->stream, store, c->g.plane, dst, a, sa, b,
                       sb);
    HPQ_HIP(hipGetLastError());
}

void sxsy_from_beam(hpq_ctx* c, double* store) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    hipLaunchKernelGGL(k_sxsy, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g, store + L.jz_beam * g.plane,
                       store + L.prev_jx_beam * g.plane, store + L.prev_jy_beam * g.plane,
                       store + L.next_jx_beam * g.plane, store + L.next_jy_beam * g.plane, store + L.Sx * g.plane,
                       store + L.Sy * g.plane);
    HPQ_HIP(hipGetLastError());
}

void diag_accumulate(hpq_ctx* c, const double* store, int islice) {
    const Geom& g = c->g;
    if (c->checksums) {
        hipLaunchKernelGGL(k_diag, dim3((g.ny + 7) / 8, c->L.ncomp), dim3(256), 0, c->stream, g, store, c->checksums);
        HPQ_HIP(hipGetLastError());
    }
    for (const Capture& cap : c->captures) {
        hipLaunchKernelGGL(k_capture, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                           store + long(cap.comp) * g.plane, cap.dst + long(islice) * g.nx * g.ny);
        HPQ_HIP(hipGetLastError());
    }
}

double rel_b_error(hpq_ctx* c, const double* store, int bx, int by, int bxi, int byi) {
    const Geom& g = c->g;
    HPQ_HIP(hipMemsetAsync(c->d_red, 0, 2 * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_relb, dim3(64), dim3(256), 0, c->stream, g, store + long(bx) * g.plane,
                       store + long(by) * g.plane, store + long(bxi) * g.plane, store + long(byi) * g.plane, c->d_red);
    HPQ_HIP(hipGetLastError());
    HPQ_HIP(hipMemcpyAsync(c->h_red, c->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPQ_HIP(hipStreamSynchronize(c->stream));
    const double nB = c->h_red[0], nD = c->h_red[1];
    return (nB > 0.) ? nD / nB : 0.;
}

}  // namespace hpq
