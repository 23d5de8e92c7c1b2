// Slice-store bookkeeping kernels: fill/copy/add/linear combination of planes
// (Fields::setVal/duplicate/add, MultiFab::LinComb; src/fields/Fields.H:143-287),
// Sx/Sy from the beam (Hipace::InitializeSxSyWithBeam, src/Hipace.cpp:745-790),
// the relative B-field error of the predictor-corrector loop
// (Fields::ComputeRelBFieldError, src/fields/Fields.cpp:1229-1286) and the per-slice
// field diagnostics (sum |F| over valid cells, tests/checksum/backend/openpmd_backend.py).
// Up to 16 planes are handled per launch so a slice-store update is one kernel; the end of
// a slice (diagnostics, ShiftSlices and the next slice's InitializeSlices) is one pass.
#include <vector>

#include "internal.hpp"

namespace hpq {

namespace {

struct PlaneSet {
    int n;
    int dst[16];
    int src[16];
    double v;
};

__global__ void k_fill(double* __restrict__ store, long plane, PlaneSet s) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    for (int k = 0; k < s.n; ++k) store[long(s.dst[k]) * plane + t] = s.v;
}

// copies applied in list order at every cell (a later copy may read an earlier one's dst)
__global__ void k_copy(double* __restrict__ store, long plane, PlaneSet s) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    for (int k = 0; k < s.n; ++k) store[long(s.dst[k]) * plane + t] = store[long(s.src[k]) * plane + t];
}

// dst[k] += src[k]
__global__ void k_add(double* __restrict__ store, long plane, PlaneSet s) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= plane) return;
    for (int k = 0; k < s.n; ++k) store[long(s.dst[k]) * plane + t] += store[long(s.src[k]) * plane + t];
}

// End of a slice in one pass over the guarded plane: sum |F| of the diagnostic
// components over valid cells (per-workgroup partial sums, accumulated over slices in
// partial[block][k], no atomics), then the copies of ShiftSlices in order, then the zeroing
// of InitializeSlices for the next slice. Per cell the reads come before the writes.
constexpr int SE_T = 256, SE_DIAG = 24;
struct SliceEnd {
    int ndiag;
    int diag[SE_DIAG];
    int ins[7];       // in-situ planes: ExmBy, EypBx, Ez, Bx, By, Bz, jz_beam
    PlaneSet shift;   // dst <- src, in order
    PlaneSet zero;    // dst <- 0
};
constexpr int SE_INSITU = HPQ_INSITU_FIELD_NR;

template <bool DIAG, bool SHIFT, bool INSITU>
__global__ __launch_bounds__(SE_T) void k_slice_end(Geom g, double* __restrict__ store, SliceEnd e,
                                                    double* __restrict__ partial, double* __restrict__ ipartial) {
    __shared__ double red[SE_T / 64][SE_DIAG];
    __shared__ double ired[SE_T / 64][SE_INSITU];
    const int plane = int(g.plane);
    const int t = blockIdx.x * SE_T + threadIdx.x;
    bool inside = false;
    if (t < plane) {
        const int jj = t / g.ldx, ii = t - jj * g.ldx;
        inside = ii >= g.ng && ii < g.ng + g.nx && jj >= g.ng && jj < g.ng + g.ny;
    }
    double acc[SE_DIAG];
#pragma unroll
    for (int k = 0; k < SE_DIAG; ++k) acc[k] = 0.0;
    if (DIAG && inside) {
#pragma unroll
        for (int k = 0; k < SE_DIAG; ++k)
            if (k < e.ndiag) acc[k] = fabs(store[long(e.diag[k]) * plane + t]);
    }
    // Fields::InSituComputeDiags, src/fields/Fields.cpp:1323-1336 (c = 1)
    double q[SE_INSITU];
#pragma unroll
    for (int k = 0; k < SE_INSITU; ++k) q[k] = 0.0;
    if (INSITU && inside) {
        const double clight = 1.0;
        const double exmby = store[long(e.ins[0]) * plane + t], eypbx = store[long(e.ins[1]) * plane + t];
        const double ez = store[long(e.ins[2]) * plane + t], bx = store[long(e.ins[3]) * plane + t];
        const double by = store[long(e.ins[4]) * plane + t], bz = store[long(e.ins[5]) * plane + t];
        const double jzb = store[long(e.ins[6]) * plane + t];
        const double ex = exmby + by * clight, ey = eypbx - bx * clight;
        q[0] = ex * ex;
        q[1] = ey * ey;
        q[2] = ez * ez;
        q[3] = bx * bx;
        q[4] = by * by;
        q[5] = bz * bz;
        q[6] = exmby * exmby;
        q[7] = eypbx * eypbx;
        q[8] = jzb;
        q[9] = ez * jzb;
    }
    if (SHIFT && t < plane) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < e.shift.n) store[long(e.shift.dst[k]) * plane + t] = store[long(e.shift.src[k]) * plane + t];
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < e.zero.n) store[long(e.zero.dst[k]) * plane + t] = 0.0;
    }
    if (DIAG) {
#pragma unroll
        for (int k = 0; k < SE_DIAG; ++k) {
            if (k < e.ndiag) {
                double v = acc[k];
                for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
                if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;
            }
        }
    }
    if (INSITU) {
#pragma unroll
        for (int k = 0; k < SE_INSITU; ++k) {
            double v = q[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if ((threadIdx.x & 63) == 0) ired[threadIdx.x >> 6][k] = v;
        }
    }
    if (DIAG || INSITU) __syncthreads();
    if (DIAG && threadIdx.x < e.ndiag) {
        double v = red[0][threadIdx.x];
        for (int w = 1; w < SE_T / 64; ++w) v += red[w][threadIdx.x];
        partial[long(blockIdx.x) * SE_DIAG + threadIdx.x] += v;
    }
    // this slice's row of in-situ sums (overwritten every slice)
    if (INSITU && threadIdx.x < SE_INSITU) {
        double v = ired[0][threadIdx.x];
        for (int w = 1; w < SE_T / 64; ++w) v += ired[w][threadIdx.x];
        ipartial[long(blockIdx.x) * INSITU_STRIDE + threadIdx.x] = v;
    }
}

// checksums[k] += sum over workgroups of partial[b][k]; partial cleared
__global__ __launch_bounds__(256) void k_diag_finalize(double* __restrict__ partial, int nblk, int ndiag,
                                                       const int* __restrict__ comp_of, double* __restrict__ sums) {
    __shared__ double red[256 / 64];
    for (int k = 0; k < ndiag; ++k) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nblk; b += 256) {
            v += partial[long(b) * SE_DIAG + k];
            partial[long(b) * SE_DIAG + k] = 0.0;
        }
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) sums[comp_of[k]] += red[0] + red[1] + red[2] + red[3];
    }
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

__global__ void k_capture(Geom g, const double* __restrict__ P, double* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= g.nx) return;
    dst[long(j) * g.nx + i] = P[g.idx(i, j)];
}

// The two sums of ComputeRelBFieldError over valid cells in a fixed order, so that the
// predictor-corrector loop repeats bit for bit: one wavefront per row, lane l adding the cells
// i = l, l + 64, ... in turn, the lanes combined by a xor butterfly (offsets 32, 16, ..., 1;
// every lane ends with the same value); then one wavefront does the same over the row sums.
__device__ __forceinline__ double butterfly64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(256) void k_relb_rows(Geom g, const double* __restrict__ Bx, const double* __restrict__ By,
                                                   const double* __restrict__ Bxi, const double* __restrict__ Byi,
                                                   double* __restrict__ rows) {
    const int j = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= g.ny) return;
    const double factor = g.dx * g.dy / (g.dx * g.dy);
    double a = 0.0, b = 0.0;
    for (int i = lane; i < g.nx; i += 64) {
        const long k = g.idx(i, j);
        a += factor * sqrt(Bx[k] * Bx[k] + By[k] * By[k]);
        b += factor * sqrt((Bx[k] - Bxi[k]) * (Bx[k] - Bxi[k]) + (By[k] - Byi[k]) * (By[k] - Byi[k]));
    }
    a = butterfly64(a);
    b = butterfly64(b);
    if (lane == 0) {
        rows[j] = a;
        rows[g.ny + j] = b;
    }
}

__global__ __launch_bounds__(64) void k_relb_total(int ny, const double* __restrict__ rows, double* __restrict__ out) {
    double a = 0.0, b = 0.0;
    for (int j = threadIdx.x; j < ny; j += 64) {
        a += rows[j];
        b += rows[ny + j];
    }
    a = butterfly64(a);
    b = butterfly64(b);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
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

void plane_add(hpq_ctx* c, double* store, const int* dst, const int* src, int ncomps) {
    PlaneSet s{};
    for (int k = 0; k < ncomps; ++k)
        if (dst[k] >= 0) {
            s.dst[s.n] = dst[k];
            s.src[s.n] = src[k];
            ++s.n;
        }
    if (s.n == 0) return;
    hipLaunchKernelGGL(k_add, grid1d(c->g.plane, 256), dim3(256), 0, c->stream, store, c->g.plane, s);
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

// Components summed by the field diagnostics: every plane of the layout except the
// scratch planes of the slice shift, of the predictor-corrector loop and of SALAME.
void diag_comps(const hpq_ctx* c, std::vector<int>& out) {
    const Layout& L = c->L;
    const int skip[] = {L.next_jx_beam, L.next_jy_beam, L.prev_jx_beam, L.prev_jy_beam, L.ions_rhomjz, L.next_jx,
                        L.next_jy,      L.prev_Bx,      L.prev_By,      L.prev_jx,      L.prev_jy,     L.pciter_Bx,
                        L.pciter_By,    L.pcprev_Bx,    L.pcprev_By,   L.sal_Ez_target, L.sal_Ez_no_salame,
                        L.sal_Ez,       L.sal_jx,       L.sal_jy,      L.sal_jz_beam, L.sal_Bx,
                        L.sal_By,       L.sal_Sy,       L.sal_Sx,      L.sal_Sy_back, L.sal_Sx_back};
    out.clear();
    for (int k = 0; k < L.ncomp; ++k) {
        bool s = false;
        for (int q : skip) s = s || q == k;
        if (!s) out.push_back(k);
    }
}

void captures(hpq_ctx* c, const double* store, int islice) {
    const Geom& g = c->g;
    for (const Capture& cap : c->captures) {
        hipLaunchKernelGGL(k_capture, dim3((g.nx + 127) / 128, g.ny), dim3(128), 0, c->stream, g,
                           store + long(cap.comp) * g.plane, cap.dst + long(islice) * g.nx * g.ny);
        HPQ_HIP(hipGetLastError());
    }
}

void slice_end(hpq_ctx* c, double* store, int islice, const int* shift_dst, const int* shift_src, int nshift,
               const int* zero, int nzero) {
    const Geom& g = c->g;
    const Layout& L = c->L;
    const bool diag = c->checksums != nullptr;
    const bool insitu = c->insitu.field_rdata != nullptr;
    const bool shift = nshift > 0 || nzero > 0;
    if (!diag && !shift && !insitu) return;
    SliceEnd e{};
    std::vector<int> dc;
    diag_comps(c, dc);
    if (int(dc.size()) > SE_DIAG) throw HpqError(HPQ_E_UNSUPPORTED, "slice_end: too many diagnostic components");
    e.ndiag = int(dc.size());
    for (int k = 0; k < e.ndiag; ++k) e.diag[k] = dc[k];
    const int ins[7] = {L.ExmBy, L.EypBx, L.Ez, L.Bx, L.By, L.Bz, L.jz_beam};
    for (int k = 0; k < 7; ++k) {
        if (insitu && ins[k] < 0) throw HpqError(HPQ_E_UNSUPPORTED, "Must use explicit solver for field insitu diagnostic");
        e.ins[k] = ins[k];
    }
    for (int k = 0; k < nshift; ++k) {
        e.shift.dst[e.shift.n] = shift_dst[k];
        e.shift.src[e.shift.n] = shift_src[k];
        ++e.shift.n;
    }
    for (int k = 0; k < nzero; ++k)
        if (zero[k] >= 0) e.zero.dst[e.zero.n++] = zero[k];
    const long nblk = (g.plane + SE_T - 1) / SE_T;
    if (diag && !c->diag_partial) {
        HPQ_HIP(hipMalloc(&c->diag_partial, sizeof(double) * SE_DIAG * nblk));
        HPQ_HIP(hipMemset(c->diag_partial, 0, sizeof(double) * SE_DIAG * nblk));
        HPQ_HIP(hipMalloc(&c->diag_comp_of, sizeof(int) * SE_DIAG));
        HPQ_HIP(hipMemcpy(c->diag_comp_of, dc.data(), sizeof(int) * dc.size(), hipMemcpyHostToDevice));
    }
    double* ipart = insitu ? insitu_scratch(c, nblk) : nullptr;
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(unsigned(nblk)), dim3(SE_T), 0, c->stream, g, store, e, c->diag_partial, ipart);
    };
    if (insitu) {
        if (diag && shift) launch(k_slice_end<true, true, true>);
        else if (diag) launch(k_slice_end<true, false, true>);
        else if (shift) launch(k_slice_end<false, true, true>);
        else launch(k_slice_end<false, false, true>);
    } else {
        if (diag && shift) launch(k_slice_end<true, true, false>);
        else if (diag) launch(k_slice_end<true, false, false>);
        else launch(k_slice_end<false, true, false>);
    }
    HPQ_HIP(hipGetLastError());
    c->diag_pending = c->diag_pending || diag;
    if (insitu)
        insitu_finalize(c, int(nblk), HPQ_INSITU_FIELD_NR, false, g.dx * g.dy * g.dz, 0u, c->insitu.field_rdata,
                        c->insitu.field_sum, nullptr, nullptr, islice);
}

// adds the pending per-workgroup diagnostic sums into the caller's checksums
void diag_finalize(hpq_ctx* c) {
    if (!c->diag_pending || !c->checksums || !c->diag_partial) return;
    std::vector<int> dc;
    diag_comps(c, dc);
    const long nblk = (c->g.plane + SE_T - 1) / SE_T;
    hipLaunchKernelGGL(k_diag_finalize, dim3(1), dim3(256), 0, c->stream, c->diag_partial, int(nblk), int(dc.size()),
                       c->diag_comp_of, c->checksums);
    HPQ_HIP(hipGetLastError());
    c->diag_pending = false;
}

double rel_b_error(hpq_ctx* c, const double* store, int bx, int by, int bxi, int byi) {
    const Geom& g = c->g;
    if (c->relb_cap < 2L * g.ny) {
        if (c->relb_rows) HPQ_HIP(hipFree(c->relb_rows));
        c->relb_rows = nullptr;
        HPQ_HIP(hipMalloc(&c->relb_rows, sizeof(double) * 2 * g.ny));
        c->relb_cap = 2L * g.ny;
    }
    hipLaunchKernelGGL(k_relb_rows, dim3((g.ny + 3) / 4), dim3(256), 0, c->stream, g, store + long(bx) * g.plane,
                       store + long(by) * g.plane, store + long(bxi) * g.plane, store + long(byi) * g.plane,
                       c->relb_rows);
    hipLaunchKernelGGL(k_relb_total, dim3(1), dim3(64), 0, c->stream, g.ny, c->relb_rows, c->d_red);
    HPQ_HIP(hipGetLastError());
    HPQ_HIP(hipMemcpyAsync(c->h_red, c->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HPQ_HIP(hipStreamSynchronize(c->stream));
    const double nB = c->h_red[0], nD = c->h_red[1];
    return (nB > 0.) ? nD / nB : 0.;
}

}  // namespace hpq
