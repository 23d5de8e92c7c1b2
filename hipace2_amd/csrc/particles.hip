// Particle kernels on gfx950: field gather + leapfrog push of the plasma, plasma injection.
// (Deposition is in deposit.hip.)
//
// One thread per particle over coalesced SoA arrays; stencils unrolled at compile time
// (shape order is a template parameter). Every expression keeps the reference's operation
// order, so FP64 results equal the reference CPU build's bit for bit. The arithmetic type R
// of the gather and of the momentum sub-steps is double, or float on the mixed-precision
// path, where the FP64 particle state accumulates the FP32 increments.
#include <cmath>

#include "internal.hpp"
#include "push.hpp"
#include "shape.hpp"

namespace hpq {

namespace {

__device__ __forceinline__ bool in_store(const Geom& g, int i, int j) {
    return i >= -g.ng && i <= g.nx - 1 + g.ng && j >= -g.ng && j <= g.ny - 1 + g.ng;
}

// EnforceBC, src/particles/pusher/GetAndSetPosition.H:29-99; returns true if removed
__device__ __forceinline__ bool enforce_bc(const BC& bc, double& x, double& y, double& ux, double& uy) {
    if (x < bc.plo0 || y < bc.plo1 || x > bc.phi0 || y > bc.phi1) {
        const double lx = bc.phi0 - bc.plo0, ly = bc.phi1 - bc.plo1;
        if (bc.type == HPQ_BC_REFLECTING) {
            x = fmod(x - bc.plo0, 2 * lx); if (x < 0) x += 2 * lx; x += bc.plo0;
            if (x > bc.phi0) { x = 2 * bc.phi0 - x; ux = -ux; }
            y = fmod(y - bc.plo1, 2 * ly); if (y < 0) y += 2 * ly; y += bc.plo1;
            if (y > bc.phi1) { y = 2 * bc.phi1 - y; uy = -uy; }
            return false;
        } else if (bc.type == HPQ_BC_PERIODIC) {
            x = fmod(x - bc.plo0, lx); if (x < 0) x += lx; x += bc.plo0;
            y = fmod(y - bc.plo1, ly); if (y < 0) y += ly; y += bc.plo1;
            return false;
        }
        return true;
    }
    return false;
}

// AdvancePlasmaParticles, src/particles/pusher/PlasmaParticleAdvance.cpp:84-240 (leapfrog)
template <int O, class R>
__global__ __launch_bounds__(256) void k_push_plasma(Geom g, hpq_plasma p, const double* __restrict__ store, int cPsi,
                                                     int cEz, int cBx, int cBy, int cBz, double qmc, BC bc,
                                                     int temp_slice) {
    // XCD-contiguous particle ranges: the plasma is stored as ppc sweeps over the plane, so
    // one XCD's particles read a band of field rows that stays in its L2 (HBM fetch per
    // launch 222 MB -> 89 MB at 511^2, 2x2 ppc). The tile-sorted deposition kernels keep
    // the round-robin order: their window flushes contend less when neighbouring chunks sit
    // on different XCDs.
    const long ip = long(xcd_block(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (ip >= p.n || !p.valid[ip]) return;
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double dz = g.dz / 1;
    double xp = p.x_prev[ip], yp = p.y_prev[ip];
    R ExmByp = 0, EypBxp = 0, Ezp = 0, Bxp = 0, Byp = 0, Bzp = 0;
    {
        const double xm = (xp - g.xoff) * g.dxi, ym = (yp - g.yoff) * g.dyi;
        const double* Psi = store + long(cPsi) * g.plane;
        const double* Ez = store + long(cEz) * g.plane;
        const double* Bx = store + long(cBx) * g.plane;
        const double* By = store + long(cBy) * g.plane;
        const double* Bz = store + long(cBz) * g.plane;
        constexpr int W = O + 2;
        const R dxi = R(g.dxi), dyi = R(g.dyi);
        DsfR<R> sx[W], sy[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            sx[k] = dshape<1, O, R>(xm, k);
            sy[k] = dshape<1, O, R>(ym, k);
        }
        auto accumulate = [&](const DsfR<R>& X, const DsfR<R>& Y, R ps, R ez, R bx, R by, R bz) {
            ExmByp += (X.ds * Y.s) * ps * dxi;
            EypBxp += (X.s * Y.ds) * ps * dyi;
            Ezp += (X.s * Y.s) * ez;
            Bxp += (X.s * Y.s) * bx;
            Byp += (X.s * Y.s) * by;
            Bzp += (X.s * Y.s) * bz;
        };
        if (in_store(g, sx[0].cell, sy[0].cell) && in_store(g, sx[W - 1].cell, sy[W - 1].cell)) {
            // The whole stencil lies in the guarded plane: one field at a time, its W x W
            // loads issued together. Every accumulator still sums its cells in (iy, ix)
            // order, as the reference's cell loop does.
            const long k0 = g.idx(sx[0].cell, sy[0].cell);
            auto tile = [&](const double* F, R (&v)[W][W]) {
#pragma unroll
                for (int iy = 0; iy < W; ++iy)
#pragma unroll
                    for (int ix = 0; ix < W; ++ix) v[iy][ix] = R(F[k0 + long(iy) * g.ldx + ix]);
            };
            R v[W][W];
            tile(Psi, v);
#pragma unroll
            for (int iy = 0; iy < W; ++iy)
#pragma unroll
                for (int ix = 0; ix < W; ++ix) {
                    ExmByp += (sx[ix].ds * sy[iy].s) * v[iy][ix] * dxi;
                    EypBxp += (sx[ix].s * sy[iy].ds) * v[iy][ix] * dyi;
                }
            auto plain = [&](const double* F, R& acc) {
                tile(F, v);
#pragma unroll
                for (int iy = 0; iy < W; ++iy)
#pragma unroll
                    for (int ix = 0; ix < W; ++ix) acc += (sx[ix].s * sy[iy].s) * v[iy][ix];
            };
            plain(Ez, Ezp);
            plain(Bx, Bxp);
            plain(By, Byp);
            plain(Bz, Bzp);
        } else {
#pragma unroll
            for (int iy = 0; iy < W; ++iy) {
#pragma unroll
                for (int ix = 0; ix < W; ++ix) {
                    const DsfR<R> X = sx[ix], Y = sy[iy];
                    if (!in_store(g, X.cell, Y.cell)) continue;
                    const long k = g.idx(X.cell, Y.cell);
                    accumulate(X, Y, R(Psi[k]), R(Ez[k]), R(Bx[k]), R(By[k]), R(Bz[k]));
                }
            }
        }
    }
    Bxp *= R(clight);
    Byp *= R(clight);
    constexpr int nsub = 4;
    const double sdz = dz / nsub;
    // the (FP64) state; the increments are computed in R
    double ux = p.ux_half_step[ip], uy = p.uy_half_step[ip], psi = p.psi_half_step[ip];
    const R sdr = R(sdz), cir = R(clight_inv), qmr = R(qmc);
    auto substep = [&]() {
        const R psi_inv = R(1.0) / R(psi);
        R dux, duy, dpsi;
        momentum_rhs<R, R>(R(ux), R(uy), psi_inv, ExmByp, EypBxp, Ezp, Bxp, Byp, Bzp, cir, qmr, dux, duy, dpsi);
        const Dual<R> uxd{R(ux), dux}, uyd{R(uy), duy}, pid{psi_inv, -psi_inv * psi_inv * dpsi};
        Dual<R> ddux(R(0.0)), dduy(R(0.0)), ddpsi(R(0.0));
        momentum_rhs<Dual<R>, R>(uxd, uyd, pid, ExmByp, EypBxp, Ezp, Bxp, Byp, Bzp, cir, qmr, ddux, dduy, ddpsi);
        ux += double(sdr * dux + R(0.5) * sdr * sdr * ddux.e);
        uy += double(sdr * duy + R(0.5) * sdr * sdr * dduy.e);
        psi += double(sdr * dpsi + R(0.5) * sdr * sdr * ddpsi.e);
    };
#pragma unroll 1
    for (int s = 0; s < nsub; ++s) substep();
    xp += double(R(dz) * cir * (R(ux) * (R(1.0) / R(psi))));
    yp += double(R(dz) * cir * (R(uy) * (R(1.0) / R(psi))));
    if (enforce_bc(bc, xp, yp, ux, uy)) {
        p.w[ip] = 0.0;
        p.valid[ip] = 0;
        return;
    }
    p.x[ip] = xp;
    p.y[ip] = yp;
    if (!temp_slice) {
        p.ux_half_step[ip] = ux;
        p.uy_half_step[ip] = uy;
        p.psi_half_step[ip] = psi;
        p.x_prev[ip] = xp;
        p.y_prev[ip] = yp;
    }
#pragma unroll 1
    for (int s = 0; s < nsub / 2; ++s) substep();
    p.ux[ip] = ux;
    p.uy[ip] = uy;
    p.psi[ip] = psi;
}

// AdvanceBeamParticlesSlice (BeamParticleAdvance.cpp:128-336; no radiation reaction, spin
// or external fields) for particle ip; a particle that ends below the slice with sub-cycles
// left is appended to slip_out (shiftSlippedParticles, SliceSort.cpp:13-70).
template <int O>
__device__ void push_beam_particle(const Geom& g, const hpq_beam_state& b, long ip, const double* __restrict__ store,
                                   int cPsi, int cEz, int cBx, int cBy, int cBz, double dt, int n_subcycles,
                                   bool do_z_push, double qm, double min_z, const BC& bc, int32_t* slip_out,
                                   int32_t* slip_out_cnt) {
    if (!b.valid[ip]) return;
    const double clight = 1.0, inv_c2 = 1.0 / (1.0 * 1.0);
    double xp = b.x[ip], yp = b.y[ip], zp = b.z[ip];
    double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
    int i = b.nsubcycles[ip];
    const double* Psi = store + long(cPsi) * g.plane;
    const double* Ez = store + long(cEz) * g.plane;
    const double* Bx = store + long(cBx) * g.plane;
    const double* By = store + long(cBy) * g.plane;
    const double* Bz = store + long(cBz) * g.plane;
    constexpr int W = O + 2;
    for (; i < n_subcycles; i++) {
        if (zp < min_z) break;
        const double gammap_inv = 1.0 / sqrt(1.0 + (ux * ux + uy * uy + uz * uz) * inv_c2);
        xp += dt * 0.5 * ux * gammap_inv;
        yp += dt * 0.5 * uy * gammap_inv;
        if (enforce_bc(bc, xp, yp, ux, uy)) {
            b.w[ip] = 0.0;
            b.valid[ip] = 0;
            return;
        }
        double ExmByp = 0, EypBxp = 0, Ezp = 0, Bxp = 0, Byp = 0, Bzp = 0;
        const double xm = (xp - g.xoff) * g.dxi, ym = (yp - g.yoff) * g.dyi;
        Dsf sx[W], sy[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            sx[k] = dshape<1, O>(xm, k);
            sy[k] = dshape<1, O>(ym, k);
        }
#pragma unroll
        for (int iy = 0; iy < W; ++iy)
#pragma unroll
            for (int ix = 0; ix < W; ++ix) {
                const Dsf X = sx[ix], Y = sy[iy];
                if (!in_store(g, X.cell, Y.cell)) continue;
                const long k = g.idx(X.cell, Y.cell);
                ExmByp += (X.ds * Y.s) * Psi[k] * g.dxi;
                EypBxp += (X.s * Y.ds) * Psi[k] * g.dyi;
                Ezp += (X.s * Y.s) * Ez[k];
                Bxp += (X.s * Y.s) * Bx[k];
                Byp += (X.s * Y.s) * By[k];
                Bzp += (X.s * Y.s) * Bz[k];
            }
        if (b.use_external_fields) {
            // ApplyExternalField, src/particles/pusher/ExternalFields.H:43-55
            const double* e = b.external_fields;
            const double Ex = ((e[0] + e[1] * xp) + e[2] * yp) + e[3] * zp;
            const double Ey = ((e[4] + e[5] * xp) + e[6] * yp) + e[7] * zp;
            const double Eze = ((e[8] + e[9] * xp) + e[10] * yp) + e[11] * zp;
            const double Bxe = ((e[12] + e[13] * xp) + e[14] * yp) + e[15] * zp;
            const double Bye = ((e[16] + e[17] * xp) + e[18] * yp) + e[19] * zp;
            const double Bze = ((e[20] + e[21] * xp) + e[22] * yp) + e[23] * zp;
            ExmByp += Ex - clight * Bye;
            EypBxp += Ey + clight * Bxe;
            Ezp += Eze;
            Bxp += Bxe;
            Byp += Bye;
            Bzp += Bze;
        }
        const double ux_next = ux + dt * qm * (ExmByp + (clight - uz * gammap_inv) * Byp + uy * gammap_inv * Bzp);
        const double uy_next = uy + dt * qm * (EypBxp + (uz * gammap_inv - clight) * Bxp - ux * gammap_inv * Bzp);
        const double ux_i = (ux_next + ux) * 0.5;
        const double uy_i = (uy_next + uy) * 0.5;
        const double uz_i = uz + dt * 0.5 * qm * Ezp;
        const double gamma_i_inv = 1.0 / sqrt(1.0 + (ux_i * ux_i + uy_i * uy_i + uz_i * uz_i) * inv_c2);
        const double uz_next = uz + dt * qm * (Ezp + (ux_i * Byp - uy_i * Bxp) * gamma_i_inv);
        const double gamma_next_inv =
            1.0 / sqrt(1.0 + (ux_next * ux_next + uy_next * uy_next + uz_next * uz_next) * inv_c2);
        xp += dt * 0.5 * ux_next * gamma_next_inv;
        yp += dt * 0.5 * uy_next * gamma_next_inv;
        if (do_z_push) zp += dt * (uz_next * gamma_next_inv - clight);
        ux = ux_next;
        uy = uy_next;
        uz = uz_next;
    }
    if (enforce_bc(bc, xp, yp, ux, uy)) {
        b.w[ip] = 0.0;
        b.valid[ip] = 0;
        return;
    }
    b.x[ip] = xp;
    b.y[ip] = yp;
    b.z[ip] = zp;
    b.nsubcycles[ip] = i;
    b.ux[ip] = ux;
    b.uy[ip] = uy;
    b.uz[ip] = uz;
    if (zp < min_z && i < n_subcycles) {
        const int at = atomicAdd(slip_out_cnt, 1);
        slip_out[at] = int32_t(ip);
    }
}

// the particles [first, first + count) one per thread, then the slipped-in list grid-strided
template <int O>
__global__ __launch_bounds__(256) void k_push_beam(Geom g, hpq_beam_state b, long first, long count,
                                                   const double* __restrict__ store, int cPsi, int cEz, int cBx, int cBy,
                                                   int cBz, double dt, int n_subcycles, int do_z_push, double qm,
                                                   double min_z, BC bc, const int32_t* __restrict__ slip_in,
                                                   const int32_t* __restrict__ slip_in_cnt, int32_t* slip_out,
                                                   int32_t* slip_out_cnt) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long total = long(gridDim.x) * blockDim.x;
    if (t < count)
        push_beam_particle<O>(g, b, first + t, store, cPsi, cEz, cBx, cBy, cBz, dt, n_subcycles, do_z_push != 0, qm,
                              min_z, bc, slip_out, slip_out_cnt);
    const long n_in = *slip_in_cnt;
    for (long j = t; j < n_in; j += total)
        push_beam_particle<O>(g, b, long(slip_in[j]), store, cPsi, cEz, cBx, cBy, cBz, dt, n_subcycles,
                              do_z_push != 0, qm, min_z, bc, slip_out, slip_out_cnt);
}

__global__ void k_fill_i32(int32_t* a, long n, int32_t v) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

}  // namespace

// the FP32 push, in a translation unit of its own (particles_f32.hip) built with FMA
// contraction; this one keeps every FP64 multiply and add separately rounded
void push_plasma_f32(hpq_ctx* c, const hpq_plasma* p, const double* store, double qmc, const BC& bc, bool temp_slice);

#ifndef HPQ_F32_TU
void advance_plasma(hpq_ctx* c, const hpq_plasma* p, const double* store, bool temp_slice) {
    c->cl_plasma.valid = false;   // positions change: the deposition lists are out of date
    if (p->n <= 0) return;
    const Layout& L = c->L;
    const double qmc = c->p.plasma_charge / (c->p.plasma_mass * 1.0);
    BC bc{c->p.particle_bc, c->p.plo[0], c->p.plo[1], c->p.phi[0], c->p.phi[1]};
    TraceScope ts(c, HPQ_TRACE_PUSH);
    if (c->p.precision == HPQ_PREC_MIXED) {
        push_plasma_f32(c, p, store, qmc, bc, temp_slice);
    } else {
        dispatch_order(c->p.depos_order, [&](auto O) {
            hipLaunchKernelGGL((k_push_plasma<decltype(O)::value, double>), grid1d(p->n, 256), dim3(256), 0, c->stream,
                               c->g, *p, store, L.Psi, L.Ez, L.Bx, L.By, L.Bz, qmc, bc, int(temp_slice));
        });
    }
    HPQ_HIP(hipGetLastError());
}

void beam_slip_free(hpq_ctx* c) {
    BeamSlip& S = c->slip;
    for (int32_t* p : {S.idx[0], S.idx[1], S.cnt})
        if (p) (void)hipFree(p);
    S = BeamSlip{};
}

void beam_begin_step(hpq_ctx* c, hpq_beam_state* b) {
    BeamSlip& S = c->slip;
    if (S.cap < b->n) {
        beam_slip_free(c);
        const long cap = std::max<long>(b->n, 1);
        HPQ_HIP(hipMalloc(&S.idx[0], sizeof(int32_t) * cap));
        HPQ_HIP(hipMalloc(&S.idx[1], sizeof(int32_t) * cap));
        HPQ_HIP(hipMalloc(&S.cnt, sizeof(int32_t) * 2));
        S.cap = cap;
    }
    S.cur = 0;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, S.cnt, 2L, 0);
    if (b->n > 0)
        hipLaunchKernelGGL(k_fill_i32, grid1d(b->n, 256), dim3(256), 0, c->stream, b->nsubcycles, long(b->n), 0);
    HPQ_HIP(hipGetLastError());
}

void advance_beam(hpq_ctx* c, hpq_beam_state* b, int slice, const double* store, double dt_step, int n_subcycles,
                  bool do_z_push) {
    BeamSlip& S = c->slip;
    if (S.cap < b->n || !S.cnt) throw HpqError(HPQ_E_INVALID, "hpq_beam_begin_step must precede the beam pushes");
    const Layout& L = c->L;
    const Geom& g = c->g;
    const long first = b->off[slice];
    const long count = b->off[slice + 1] - first;
    const int in = S.cur, out = 1 - S.cur;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, S.cnt + out, 1L, 0);
    const double qm = b->charge / b->mass;
    const double dt = dt_step / n_subcycles;
    const double min_z = g.zlo + slice * g.dz;
    BC bc{c->p.particle_bc, c->p.plo[0], c->p.plo[1], c->p.phi[0], c->p.phi[1]};
    TraceScope ts(c, HPQ_TRACE_PUSH);
    dispatch_order(c->p.depos_order, [&](auto O) {
        hipLaunchKernelGGL((k_push_beam<decltype(O)::value>), dim3(unsigned(std::max<long>(1, (count + 255) / 256))),
                           dim3(256), 0, c->stream, g, *b, first, count, store, L.Psi, L.Ez, L.Bx, L.By, L.Bz, dt,
                           n_subcycles, int(do_z_push), qm, min_z, bc, S.idx[in], S.cnt + in, S.idx[out], S.cnt + out);
    });
    HPQ_HIP(hipGetLastError());
    S.cur = out;
}

// ---------------------------------------------------------------------------------
// plasma injection, PlasmaParticleContainer::InitParticles
// (src/particles/plasma/PlasmaParticleContainerInit.cpp:18-378): lattice positions
// (0.5 + k)/ppc per cell, weight n(x,y)/ppc, in the CPU build's order (by cell, then i_part).
// ---------------------------------------------------------------------------------
namespace {

struct Prof {
    int ppc0, ppc1;
    double density, ramp_a, radius;
    double u0, u1, u2;
    double blo0, blo1, bhi0, bhi1;
    const double* table;   // density at every lattice position, or null
};

__device__ __forceinline__ bool inject(const Geom& g, const Prof& pr, int i, int j, int ipart, double& x, double& y,
                                       double& w) {
    const int ixp = ipart % pr.ppc0, iyp = ipart / pr.ppc0;
    const double r0 = (0.5 + ixp) / pr.ppc0, r1 = (0.5 + iyp) / pr.ppc1;
    x = g.xlo + (i + r0 + 0.0) * g.dx;
    y = g.ylo + (j + r1 + 0.0) * g.dy;
    const double rsq = x * x + y * y;
    const double dens = pr.table ? pr.table[(long(j) * g.nx + i) * (pr.ppc0 * pr.ppc1) + ipart]
                                 : pr.density * (1.0 + pr.ramp_a * rsq);
    if (x >= pr.bhi0 || x < pr.blo0 || y >= pr.bhi1 || y < pr.blo1 || rsq > pr.radius * pr.radius || dens <= 0.0)
        return false;
    w = dens * (1.0 / (pr.ppc0 * pr.ppc1));
    return true;
}

// particles the profile injects into cell (i, j)
__global__ void k_init_count(Geom g, Prof pr, int32_t* __restrict__ cnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int j = t / g.nx, i = t - j * g.nx;
    int n = 0;
    double x, y, w;
    for (int ip = 0; ip < pr.ppc0 * pr.ppc1; ++ip) n += inject(g, pr, i, j, ip, x, y, w) ? 1 : 0;
    cnt[t] = n;
}

// one thread per cell writes its particles, i_part in order, from the cell's first slot
__global__ void k_init_plasma(Geom g, Prof pr, const int32_t* __restrict__ off, hpq_plasma p) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nx * g.ny) return;
    const int j = t / g.nx, i = t - j * g.nx;
    long k = off[t];
    const double psi0 = sqrt(1. + pr.u0 * pr.u0 + pr.u1 * pr.u1 + pr.u2 * pr.u2) - pr.u2;
    for (int ip = 0; ip < pr.ppc0 * pr.ppc1; ++ip) {
        double x, y, w;
        if (!inject(g, pr, i, j, ip, x, y, w)) continue;
        p.x[k] = x; p.y[k] = y; p.w[k] = w;
        p.ux[k] = pr.u0 * 1.0; p.uy[k] = pr.u1 * 1.0; p.psi[k] = psi0;
        p.x_prev[k] = x; p.y_prev[k] = y;
        p.ux_half_step[k] = pr.u0 * 1.0; p.uy_half_step[k] = pr.u1 * 1.0; p.psi_half_step[k] = psi0;
        p.valid[k] = 1;
        ++k;
    }
}

Prof make_prof(const hpq_ctx* c, const hpq_plasma_profile* prof) {
    Prof pr;
    pr.ppc0 = prof->ppc[0]; pr.ppc1 = prof->ppc[1];
    pr.density = prof->density; pr.ramp_a = prof->ramp_a; pr.radius = prof->radius;
    pr.u0 = prof->u_mean[0]; pr.u1 = prof->u_mean[1]; pr.u2 = prof->u_mean[2];
    pr.blo0 = c->p.plo[0]; pr.blo1 = c->p.plo[1]; pr.bhi0 = c->p.phi[0]; pr.bhi1 = c->p.phi[1];
    pr.table = prof->density_table;
    return pr;
}

}  // namespace

// The lattice in the CPU build's order: cells x fastest, the i_part of a cell in order (the
// GPU build's order, i_part outer, sorted by cell: SortParticlesByCell,
// PlasmaParticleContainerInit.cpp:372-377). Slots: count per cell, exclusive scan, fill; the
// total is read back once to check the capacity.
// per-cell counts and their exclusive scan in init.off (ncell + 1 entries, the last the total)
static void init_slots(hpq_ctx* c, const Prof& pr) {
    const Geom& g = c->g;
    const long ncell = long(g.nx) * g.ny;
    InitScratch& S = c->init;
    if (S.cap < ncell + 1) {
        if (S.off) HPQ_HIP(hipFree(S.off));
        if (S.tmp) HPQ_HIP(hipFree(S.tmp));
        S.off = S.tmp = nullptr;
        HPQ_HIP(hipMalloc(&S.off, sizeof(int32_t) * (ncell + 1)));
        HPQ_HIP(hipMalloc(&S.tmp, sizeof(int32_t) * scan_tmp_size(ncell + 1)));
        S.cap = ncell + 1;
    }
    hipLaunchKernelGGL(k_init_count, grid1d(ncell, 256), dim3(256), 0, c->stream, g, pr, S.off);
    HPQ_HIP(hipMemsetAsync(S.off + ncell, 0, sizeof(int32_t), c->stream));
    scan_exclusive(c->stream, S.off, ncell + 1, S.tmp);
    HPQ_HIP(hipMemcpyAsync(c->h_count + 1, S.off + ncell, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HPQ_HIP(hipStreamSynchronize(c->stream));
}

int64_t plasma_count(hpq_ctx* c, const hpq_plasma_profile* prof) {
    init_slots(c, make_prof(c, prof));
    return int64_t(c->h_count[1]);
}

void init_plasma(hpq_ctx* c, const hpq_plasma_profile* prof, const hpq_plasma* p) {
    c->cl_plasma.valid = false;
    const Prof pr = make_prof(c, prof);
    const Geom& g = c->g;
    const long ncell = long(g.nx) * g.ny;
    init_slots(c, pr);
    if (long(c->h_count[1]) != p->n)
        throw HpqError(HPQ_E_INVALID, "hpq_init_plasma: plasma capacity does not match the profile");
    hipLaunchKernelGGL(k_init_plasma, grid1d(ncell, 256), dim3(256), 0, c->stream, g, pr, c->init.off, *p);
    HPQ_HIP(hipGetLastError());
    // a cell's particles are consecutive: the atomic deposition walks them in that many runs
    c->inj_ppc = prof->ppc[0] * prof->ppc[1];
    c->bins.stale = true;
}

void init_scratch_free(hpq_ctx* c) {
    if (c->init.off) (void)hipFree(c->init.off);
    if (c->init.tmp) (void)hipFree(c->init.tmp);
    c->init = InitScratch{};
}
#endif  // HPQ_F32_TU

}  // namespace hpq
