// In-situ reduced diagnostics of the particles and the per-slice finalisation shared with
// the field diagnostics of the slice-end pass:
//   PlasmaParticleContainer::InSituComputeDiags (src/particles/plasma/PlasmaParticleContainer.cpp:442-527),
//   BeamParticleContainer::InSituComputeDiags (src/particles/beam/BeamParticleContainer.cpp:475-594),
//   Fields::InSituComputeDiags (src/fields/Fields.cpp:1289-1346).
// A reduction kernel leaves one row of sums per workgroup in insitu_partial; one workgroup
// then adds the rows in a fixed order, so a slice's values do not depend on timing.
#include "internal.hpp"

namespace hpq {

namespace {

constexpr int IS_T = 256;
constexpr int IS_MAXBLK = 1024;

// per-thread sums a[0..NQ) -> out[0..NQ) for this workgroup
template <int NQ>
__device__ __forceinline__ void block_row(const double (&a)[NQ], double* __restrict__ out) {
    __shared__ double red[IS_T / 64][NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double v = a[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double v = red[0][threadIdx.x];
        for (int w = 1; w < IS_T / 64; ++w) v += red[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// plasma: 14 weighted sums and the particle count (PlasmaParticleContainer.cpp:471-498)
__global__ __launch_bounds__(IS_T) void k_insitu_plasma(const double* __restrict__ x, const double* __restrict__ y,
                                                        const double* __restrict__ w, const double* __restrict__ ux,
                                                        const double* __restrict__ uy, const double* __restrict__ psi,
                                                        const int32_t* __restrict__ valid, long n, double r2,
                                                        double* __restrict__ partial) {
    double a[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) a[k] = 0.0;
    for (long i = long(blockIdx.x) * IS_T + threadIdx.x; i < n; i += long(gridDim.x) * IS_T) {
        if (!valid[i]) continue;
        const double xi = x[i], yi = y[i];
        if (xi * xi + yi * yi > r2) continue;
        const double u_x = ux[i], u_y = uy[i], ps = psi[i];
        const double ga = (1.0 + u_x * u_x + u_y * u_y + ps * ps) / (2.0 * ps);
        const double uz = ga - ps;
        const double wq = w[i] * ga / ps;
        const double en = w[i] * (ga - 1.0);
        a[0] += wq;
        a[1] += wq * xi;
        a[2] += wq * xi * xi;
        a[3] += wq * yi;
        a[4] += wq * yi * yi;
        a[5] += wq * u_x;
        a[6] += wq * u_x * u_x;
        a[7] += wq * u_y;
        a[8] += wq * u_y * u_y;
        a[9] += wq * uz;
        a[10] += wq * uz * uz;
        a[11] += wq * ga;
        a[12] += wq * ga * ga;
        a[13] += en;
        a[14] += 1.0;
    }
    block_row<15>(a, partial + long(blockIdx.x) * INSITU_STRIDE);
}

// beam: 22 weighted sums and the particle count (BeamParticleContainer.cpp:491-528)
__global__ __launch_bounds__(IS_T) void k_insitu_beam(const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, const double* __restrict__ ux,
                                                      const double* __restrict__ uy, const double* __restrict__ uz,
                                                      const double* __restrict__ w, const int32_t* __restrict__ valid,
                                                      long first, long n, double r2, double* __restrict__ partial) {
    double a[23];
#pragma unroll
    for (int k = 0; k < 23; ++k) a[k] = 0.0;
    for (long t = long(blockIdx.x) * IS_T + threadIdx.x; t < n; t += long(gridDim.x) * IS_T) {
        const long i = first + t;
        if (!valid[i]) continue;
        const double xi = x[i], yi = y[i];
        if (xi * xi + yi * yi > r2) continue;
        const double zi = z[i], u_x = ux[i], u_y = uy[i], u_z = uz[i], wi = w[i];
        const double uz_inv = u_z == 0.0 ? 0.0 : 1.0 / u_z;
        const double ga = sqrt(1.0 + u_x * u_x + u_y * u_y + u_z * u_z);
        a[0] += wi;
        a[1] += wi * xi;
        a[2] += wi * xi * xi;
        a[3] += wi * yi;
        a[4] += wi * yi * yi;
        a[5] += wi * zi;
        a[6] += wi * zi * zi;
        a[7] += wi * u_x;
        a[8] += wi * u_x * u_x;
        a[9] += wi * u_y;
        a[10] += wi * u_y * u_y;
        a[11] += wi * u_z;
        a[12] += wi * u_z * u_z;
        a[13] += wi * xi * u_x;
        a[14] += wi * yi * u_y;
        a[15] += wi * zi * u_z;
        a[16] += wi * xi * u_y;
        a[17] += wi * yi * u_x;
        a[18] += wi * u_x * uz_inv;
        a[19] += wi * u_y * uz_inv;
        a[20] += wi * ga;
        a[21] += wi * ga * ga;
        a[22] += 1.0;
    }
    block_row<23>(a, partial + long(blockIdx.x) * INSITU_STRIDE);
}

struct InsituOut {
    double* rdata;
    double* sum;
    int32_t* idata;
    int32_t* isum;
};

// S[k] = sum over workgroups of partial[b][k], b in increasing order per thread, then a
// fixed tree; rdata[k*nz + islice] = S[k]*scale (times 1/S[0] where norm_mask has bit k),
// sum[k] += S[k]*scale; with count, S[nreal] is the particle count.
__global__ __launch_bounds__(IS_T) void k_insitu_finalize(const double* __restrict__ partial, int nblk, int nreal,
                                                          int count, double scale, uint32_t norm_mask, int nz,
                                                          int islice, InsituOut o) {
    __shared__ double red[IS_T / 64][INSITU_STRIDE];
    __shared__ double S[INSITU_STRIDE];
    const int nq = nreal + (count ? 1 : 0);
    for (int k = 0; k < nq; ++k) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nblk; b += IS_T) v += partial[long(b) * INSITU_STRIDE + k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < nq) {
        double v = red[0][threadIdx.x];
        for (int w = 1; w < IS_T / 64; ++w) v += red[w][threadIdx.x];
        S[threadIdx.x] = v;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < nreal) {
        const double inv_w = S[0] <= 0.0 ? 0.0 : 1.0 / S[0];
        const double v = S[k] * scale;
        o.rdata[long(k) * nz + islice] = ((norm_mask >> k) & 1u) ? v * inv_w : v;
        o.sum[k] += v;
    } else if (count && k == nreal) {
        const int32_t np = int32_t(S[nreal]);
        o.idata[islice] = np;
        o.isum[0] += np;
    }
}

}  // namespace

// workgroup rows for the largest in-situ reduction: the particle kernels and the
// slice-end pass over the guarded plane
void insitu_reserve(hpq_ctx* c) {
    const long se = (c->g.plane + 255) / 256;
    const long need = se > IS_MAXBLK ? se : IS_MAXBLK;
    if (need <= c->insitu_cap) return;
    HPQ_HIP(hipStreamSynchronize(c->stream));
    if (c->insitu_partial) HPQ_HIP(hipFree(c->insitu_partial));
    c->insitu_partial = nullptr;
    c->insitu_cap = 0;
    HPQ_HIP(hipMalloc(&c->insitu_partial, sizeof(double) * INSITU_STRIDE * need));
    c->insitu_cap = need;
}

double* insitu_scratch(hpq_ctx* c, long nblk) {
    if (nblk > c->insitu_cap) throw HpqError(HPQ_E_INVALID, "insitu: reduction larger than its scratch");
    return c->insitu_partial;
}

void insitu_free(hpq_ctx* c) {
    if (c->insitu_partial) (void)hipFree(c->insitu_partial);
    c->insitu_partial = nullptr;
    c->insitu_cap = 0;
}

void insitu_finalize(hpq_ctx* c, int nblk, int nreal, bool count, double scale, uint32_t norm_mask, double* rdata,
                     double* sum, int32_t* idata, int32_t* isum, int islice) {
    if (nreal + (count ? 1 : 0) > INSITU_STRIDE || nblk > c->insitu_cap)
        throw HpqError(HPQ_E_INVALID, "insitu: reduction larger than its scratch");
    const InsituOut o{rdata, sum, idata, isum};
    hipLaunchKernelGGL(k_insitu_finalize, dim3(1), dim3(IS_T), 0, c->stream, c->insitu_partial, nblk, nreal,
                       count ? 1 : 0, scale, norm_mask, c->g.nz, islice, o);
    HPQ_HIP(hipGetLastError());
}

static int particle_blocks(long n) {
    const long b = (n + IS_T - 1) / IS_T;
    return int(b < 1 ? 1 : (b > IS_MAXBLK ? IS_MAXBLK : b));
}

void insitu_plasma(hpq_ctx* c, const hpq_plasma* p, int islice) {
    const hpq_insitu& d = c->insitu;
    if (!d.plasma_rdata) return;
    const int nblk = particle_blocks(p->n);
    double* part = insitu_scratch(c, nblk);
    hipLaunchKernelGGL(k_insitu_plasma, dim3(nblk), dim3(IS_T), 0, c->stream, p->x, p->y, p->w, p->ux, p->uy, p->psi,
                       p->valid, long(p->n), d.plasma_radius * d.plasma_radius, part);
    HPQ_HIP(hipGetLastError());
    // sum(w) (0) and [(ga-1)*(1-vz)] (13) are not divided by sum(w)
    const uint32_t mask = ((1u << HPQ_INSITU_PLASMA_NR) - 1u) & ~1u & ~(1u << 13);
    insitu_finalize(c, nblk, HPQ_INSITU_PLASMA_NR, true, 1.0, mask, d.plasma_rdata, d.plasma_sum, d.plasma_idata,
                    d.plasma_isum, islice);
}

void insitu_beam(hpq_ctx* c, const hpq_beam* b, int islice) {
    const hpq_insitu& d = c->insitu;
    if (!d.beam_rdata) return;
    long first = 0, n = 0;
    if (b && b->n > 0) {
        first = long(b->off[islice]);
        n = long(b->off[islice + 1]) - first;
    }
    const int nblk = particle_blocks(n);
    double* part = insitu_scratch(c, nblk);
    // with no particles the one workgroup writes a row of zeros and reads nothing
    hipLaunchKernelGGL(k_insitu_beam, dim3(nblk), dim3(IS_T), 0, c->stream, n > 0 ? b->x : nullptr,
                       n > 0 ? b->y : nullptr, n > 0 ? b->z : nullptr, n > 0 ? b->ux : nullptr,
                       n > 0 ? b->uy : nullptr, n > 0 ? b->uz : nullptr, n > 0 ? b->w : nullptr,
                       n > 0 ? b->valid : nullptr, first, n, d.beam_radius * d.beam_radius, part);
    HPQ_HIP(hipGetLastError());
    // sum(w) (0) is not divided by sum(w)
    const uint32_t mask = ((1u << HPQ_INSITU_BEAM_NR) - 1u) & ~1u;
    insitu_finalize(c, nblk, HPQ_INSITU_BEAM_NR, true, 1.0, mask, d.beam_rdata, d.beam_sum, d.beam_idata, d.beam_isum,
                    islice);
}

}  // namespace hpq
