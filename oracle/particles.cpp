// Oracle particle operations (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
#include "oracle.hpp"
#include <algorithm>

namespace orc {

// The CPU build's deposition engine: SharedMemoryDeposition with tiling
// (src/particles/deposition/DepositionUtil.H:204-253; Hipace::m_do_tiling is true in an OpenMP
// build, tile width T = hipace.tile_size, 32 by default, src/Hipace.H:235-240). Particles are
// binned by the tile of the lowest cell they deposit into, key ((cx - lo)/T) * ntile_y + (cy - lo)/T
// with lo the first guard cell and C integer division, invalid particles last
// (amrex::DenseBins::build, a stable counting sort); then the tiles are visited in four colours,
// (even, even), (even, odd), (odd, even), (odd, odd) tile indices, and inside a tile the particles
// in bin order. Tiles of one colour are at least T cells apart (T is at least the stencil width),
// so they run in parallel without touching the same cell, and every cell receives its
// contributions in one fixed order whatever the thread count. The tile index is clamped to the
// tile grid (the reference indexes out of it for particles outside the guarded box, whose
// stencils deposit nothing into the box).
template <class StartF, class DepF>
static void tiled_deposit(const Geom& g, int T, long n, StartF start, DepF dep) {
    const int lo = -g.ng;
    const int ntx = (g.nx + 2 * g.ng + T - 1) / T, nty = (g.ny + 2 * g.ng + T - 1) / T;
    const int nbins = ntx * nty + 1;
    std::vector<int> key(size_t(std::max(n, 0L)));
    for (long ip = 0; ip < n; ++ip) {
        int cx = 0, cy = 0;
        if (start(ip, cx, cy)) {
            const int tx = std::min(std::max((cx - lo) / T, 0), ntx - 1);
            const int ty = std::min(std::max((cy - lo) / T, 0), nty - 1);
            key[ip] = tx * nty + ty;
        } else {
            key[ip] = ntx * nty;
        }
    }
    std::vector<long> off(size_t(nbins) + 1, 0), perm(size_t(std::max(n, 0L)));
    for (long ip = 0; ip < n; ++ip) ++off[key[ip] + 1];
    for (int b = 0; b < nbins; ++b) off[b + 1] += off[b];
    {
        std::vector<long> nxt(off.begin(), off.end() - 1);
        for (long ip = 0; ip < n; ++ip) perm[nxt[key[ip]]++] = ip;
    }
    for (int px = 0; px < 2; ++px)
        for (int py = 0; py < 2; ++py) {
#pragma omp parallel for collapse(2) schedule(dynamic)
            for (int tx = px; tx < ntx; tx += 2)
                for (int ty = py; ty < nty; ty += 2) {
                    const int t = tx * nty + ty;
                    for (long k = off[t]; k < off[t + 1]; ++k) dep(perm[k]);
                }
        }
}

// ::DepositCurrent, src/particles/deposition/PlasmaDepositCurrent.cpp:22-259: start cell
// (get_start_cell, :137-155) and deposit functor (:192-251) in the tiled order above. QSA
// violators get w=0 and are invalidated (:234-241). `ions` flips the charge for the RhomJzIons
// background (:41). Component indices < 0 are skipped.
void deposit_plasma(const Ctx& c, Plasma& p, double* F, int cjx, int cjy, int cjz, int crho,
                    int cchi, int crhomjz, bool ions, long* n_qsa) {
    const Geom& g = c.g;
    const int order = c.o.depos_order;
    const double charge = ions ? -p.charge : p.charge;
    const double mass = p.mass;
    const double clight = 1.0, clightinv = 1.0 / 1.0, mu0 = 1.0;
    // normalised units: invvol = dx*dy*dxi*dyi at level 0 (:70-72)
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double charge_invvol = charge * invvol;
    const double charge_mu0_mass_ratio = charge * mu0 / mass;
    long nq = 0;
    double* pjx = cjx >= 0 ? F + cjx * g.plane : nullptr;
    double* pjy = cjy >= 0 ? F + cjy * g.plane : nullptr;
    double* pjz = cjz >= 0 ? F + cjz * g.plane : nullptr;
    double* prho = crho >= 0 ? F + crho * g.plane : nullptr;
    double* pchi = cchi >= 0 ? F + cchi * g.plane : nullptr;
    double* prm = crhomjz >= 0 ? F + crhomjz * g.plane : nullptr;
    const int lo = -g.ng, hix = g.nx - 1 + g.ng, hiy = g.ny - 1 + g.ng;
    auto start = [&](long ip, int& cx, int& cy) {
        if (!p.valid[ip]) return false;
        cx = shape(order, (p.x[ip] - g.xoff) * g.dxi, 0).cell;
        cy = shape(order, (p.y[ip] - g.yoff) * g.dyi, 0).cell;
        return true;
    };
    auto dep = [&](long ip) {
        const double psi_inv = 1.0 / p.psi[ip];
        const double xp = p.x[ip], yp = p.y[ip];
        const double vx_c = p.ux[ip] * psi_inv;
        const double vy_c = p.uy[ip] * psi_inv;
        const double q_invvol = charge_invvol * p.w[ip];
        const double q_mu0_mass_ratio = charge_mu0_mass_ratio;
        const double xmid = (xp - g.xoff) * g.dxi;
        const double ymid = (yp - g.yoff) * g.dyi;
        const double Aabssqp = 0.0;
        const double gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * psi_inv * psi_inv +
                                        vx_c * vx_c * clightinv * clightinv +
                                        vy_c * vy_c * clightinv * clightinv + 1.0);
        if (gamma_psi < 0.0 || gamma_psi > p.max_qsa || psi_inv < 0.0) {
#pragma omp atomic
            ++nq;
            p.w[ip] = 0.0;
            p.valid[ip] = 0;
            return;
        }
        for (int iy = 0; iy <= order; ++iy) {
            for (int ix = 0; ix <= order; ++ix) {
                const SF sx = shape(order, xmid, ix);
                const SF sy = shape(order, ymid, iy);
                if (sx.cell < lo || sx.cell > hix || sy.cell < lo || sy.cell > hiy) continue;
                const long k = g.idx(sx.cell, sy.cell);
                const double charge_density = q_invvol * sx.s * sy.s;
                const double wqx = charge_density * vx_c;
                const double wqy = charge_density * vy_c;
                const double wqz = charge_density * (gamma_psi - 1.0) * clight;
                const double wq = charge_density * gamma_psi;
                const double wchi = charge_density * q_mu0_mass_ratio * psi_inv;
                const double wrhomjz = charge_density;
                if (pjx) { pjx[k] += wqx; pjy[k] += wqy; }
                if (pjz) pjz[k] += wqz;
                if (prho) prho[k] += wq;
                if (pchi) pchi[k] += wchi;
                if (prm) prm[k] += wrhomjz;
            }
        }
    };
    tiled_deposit(g, c.o.tile_size, long(p.size()), start, dep);
    if (n_qsa) *n_qsa = nq;
}

// ::ExplicitDeposition, src/particles/deposition/ExplicitDeposition.cpp:20-265: start cell
// (get_cell, :109-141) and deposit functor (:143-262, no laser) in the tiled order above: Sy,
// Sx from plasma using the cell-local Bz, Ez, ExmBy, EypBx and derivative shape factors of type
// depos_derivative_type.
void explicit_deposit(const Ctx& c, const Plasma& p, double* F) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const int order = c.o.depos_order, dtype = c.o.derivative_type;
    const double clight = 1.0, clight_inv = 1.0 / 1.0, mu0 = 1.0;
    const double invvol = g.dx * g.dy * g.dxi * g.dyi;
    const double charge_invvol_mu0 = p.charge * invvol * mu0;
    const double charge_mass_ratio = p.charge / p.mass;
    const double* Bz = F + L.Bz * g.plane;
    const double* Ez = F + L.Ez * g.plane;
    const double* ExmBy = F + L.ExmBy * g.plane;
    const double* EypBx = F + L.EypBx * g.plane;
    double* Sy = F + L.Sy * g.plane;
    double* Sx = F + L.Sx * g.plane;
    const int ns = order + dtype;
    const int lo = -g.ng, hix = g.nx - 1 + g.ng, hiy = g.ny - 1 + g.ng;
    auto start = [&](long ip, int& cx, int& cy) {
        if (!p.valid[ip]) return false;
        cy = dshape(dtype, order, (p.y[ip] - g.yoff) * g.dyi, 0).cell;
        cx = dshape(dtype, order, (p.x[ip] - g.xoff) * g.dxi, 0).cell;
        return true;
    };
    auto dep = [&](long ip) {
        const double psi_inv = 1.0 / p.psi[ip];
        const double xp = p.x[ip], yp = p.y[ip];
        const double vx = p.ux[ip] * psi_inv * clight_inv;
        const double vy = p.uy[ip] * psi_inv * clight_inv;
        const double q_invvol_mu0 = charge_invvol_mu0;
        const double q_mass_ratio = charge_mass_ratio;
        const double charge_density_mu0 = q_invvol_mu0 * p.w[ip];
        const double xmid = (xp - g.xoff) * g.dxi;
        const double ymid = (yp - g.yoff) * g.dyi;
        const double Aabssqp = 0.0;
        const double gamma_psi = 0.5 * ((1.0 + 0.5 * Aabssqp) * psi_inv * psi_inv + vx * vx + vy * vy + 1.0);
        for (int iy = 0; iy <= ns; ++iy) {
            for (int ix = 0; ix <= ns; ++ix) {
                if (dtype == 2 && (ix == 0 || ix == order + 2) && (iy == 0 || iy == order + 2)) continue;
                const DSF sy = dshape(dtype, order, ymid, iy);
                const DSF sx = dshape(dtype, order, xmid, ix);
                if (sx.cell < lo || sx.cell > hix || sy.cell < lo || sy.cell > hiy) continue;
                const long k = g.idx(sx.cell, sy.cell);
                const double Bz_v = Bz[k], Ez_v = Ez[k], ExmBy_v = ExmBy[k], EypBx_v = EypBx[k];
                const double AabssqDxp = 0.0, AabssqDyp = 0.0;
                Sy[k] += charge_density_mu0 * (
                    -sx.s * sy.s * (
                        -Bz_v * vx
                        + (Ez_v * vy
                        + ExmBy_v * (-vx * vy)
                        + EypBx_v * (gamma_psi - vy * vy)) * clight_inv
                        - 0.25 * AabssqDyp * q_mass_ratio * psi_inv
                    ) * q_mass_ratio * psi_inv
                    + (-sx.ds * sy.s * g.dxi * (-vx * vy)
                       - sx.s * sy.ds * g.dyi * (gamma_psi - vy * vy - 1.0)) * clight);
                Sx[k] += charge_density_mu0 * (
                    +sx.s * sy.s * (
                        +Bz_v * vy
                        + (Ez_v * vx
                        + ExmBy_v * (gamma_psi - vx * vx)
                        + EypBx_v * (-vx * vy)) * clight_inv
                        - 0.25 * AabssqDxp * q_mass_ratio * psi_inv
                    ) * q_mass_ratio * psi_inv
                    + (+sx.ds * sy.s * g.dxi * (gamma_psi - vx * vx - 1.0)
                       + sx.s * sy.ds * g.dyi * (-vx * vy)) * clight);
            }
        }
    };
    tiled_deposit(g, c.o.tile_size, long(p.size()), start, dep);
}

// ::DepositCurrentSlice, src/particles/deposition/BeamDepositCurrent.cpp:22-196: the beam
// particles of one slice deposit jx, jy, jz (rhomjz of the beam not modelled), start cell
// (get_cell, :109-127) and deposit functor (:128-193) in the tiled order above.
void deposit_beam(const Ctx& c, const Beam& b, int slice, double* F, int cjx, int cjy, int cjz,
                  const int32_t* select, int select_value) {
    const Geom& g = c.g;
    const int order = c.o.depos_order;
    const double clightinv = 1.0 / 1.0, clightsq = 1.0 / (1.0 * 1.0);
    const double invvol = 1.0;
    const double q = b.charge;
    double* pjx = cjx >= 0 ? F + cjx * g.plane : nullptr;
    double* pjy = cjy >= 0 ? F + cjy * g.plane : nullptr;
    double* pjz = cjz >= 0 ? F + cjz * g.plane : nullptr;
    (void)clightinv;
    const int lo = -g.ng, hix = g.nx - 1 + g.ng, hiy = g.ny - 1 + g.ng;
    if (slice < 0 || slice + 1 >= int(b.off.size())) return;
    const int64_t first = b.off[slice];
    auto start = [&](long t, int& cx, int& cy) {
        const int64_t ip = first + t;
        if (!b.valid[ip]) return false;
        if (select && select[ip] != select_value) return false;
        cx = shape(order, (b.x[ip] - g.xoff) * g.dxi, 0).cell;
        cy = shape(order, (b.y[ip] - g.yoff) * g.dyi, 0).cell;
        return true;
    };
    auto dep = [&](long t) {
        const int64_t ip = first + t;
        const double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
        const double gaminv = 1.0 / std::sqrt(1.0 + ux * ux * clightsq + uy * uy * clightsq + uz * uz * clightsq);
        const double wq = q * b.w[ip] * invvol;
        const double vx = ux * gaminv, vy = uy * gaminv, vz = uz * gaminv;
        const double wqx = wq * vx, wqy = wq * vy, wqz = wq * vz;
        const double xmid = (b.x[ip] - g.xoff) * g.dxi;
        const double ymid = (b.y[ip] - g.yoff) * g.dyi;
        for (int iy = 0; iy <= order; ++iy) {
            for (int ix = 0; ix <= order; ++ix) {
                const SF sx = shape(order, xmid, ix);
                const SF sy = shape(order, ymid, iy);
                if (sx.cell < lo || sx.cell > hix || sy.cell < lo || sy.cell > hiy) continue;
                const long k = g.idx(sx.cell, sy.cell);
                if (pjx) { pjx[k] += sx.s * sy.s * wqx; pjy[k] += sx.s * sy.s * wqy; }
                if (pjz) pjz[k] += sx.s * sy.s * wqz;
            }
        }
    };
    tiled_deposit(g, c.o.tile_size, long(b.off[slice + 1] - first), start, dep);
}

// Forward-mode dual number (value, derivative) for the second-order Taylor step of
// the leapfrog pusher (src/utils/DualNumbers.H:13-43 uses the same product rule).
struct Dual {
    double v, e;
    Dual(double a) : v(a), e(0) {}
    Dual(double a, double b) : v(a), e(b) {}
};
static inline Dual operator+(const Dual& a, const Dual& b) { return {a.v + b.v, a.e + b.e}; }
static inline Dual operator-(const Dual& a, const Dual& b) { return {a.v - b.v, a.e - b.e}; }
static inline Dual operator*(const Dual& a, const Dual& b) { return {a.v * b.v, a.e * b.v + a.v * b.e}; }

// PlasmaMomentumPush, src/particles/pusher/PushPlasmaParticles.H:37-75 (no laser terms)
template <class T>
static inline void momentum_rhs(const T& ux, const T& uy, const T& psi_inv, double ExmBy, double EypBx,
                                double Ez, double Bx_c, double By_c, double Bz, double ci, double qmc,
                                T& dux, T& duy, T& dpsi) {
    const double A = 0.0, ADx = 0.0, ADy = 0.0;
    const T gamma_psi = T(0.5) * psi_inv * psi_inv * (T(1.0) + T(A) + ux * ux * T(ci * ci) + uy * uy * T(ci * ci)) + T(0.5);
    dux = (T(qmc) * (gamma_psi * T(ExmBy) + T(By_c) + (uy * T(Bz)) * psi_inv) - T(ADx) * psi_inv);
    duy = (T(qmc) * (gamma_psi * T(EypBx) - T(Bx_c) - (ux * T(Bz)) * psi_inv) - T(ADy) * psi_inv);
    dpsi = (T(qmc) * T(ci) * ((ux * T(ExmBy) + uy * T(EypBx)) * T(ci) * psi_inv - T(Ez)));
}

static inline bool enforce_bc(const Options& o, double& x, double& y, double& ux, double& uy) {
    if (x < o.plo[0] || y < o.plo[1] || x > o.phi[0] || y > o.phi[1]) {
        const double lx = o.phi[0] - o.plo[0], ly = o.phi[1] - o.plo[1];
        if (o.particle_bc == BC_REFLECTING) {
            x = std::fmod(x - o.plo[0], 2 * lx); if (x < 0) x += 2 * lx; x += o.plo[0];
            if (x > o.phi[0]) { x = 2 * o.phi[0] - x; ux = -ux; }
            y = std::fmod(y - o.plo[1], 2 * ly); if (y < 0) y += 2 * ly; y += o.plo[1];
            if (y > o.phi[1]) { y = 2 * o.phi[1] - y; uy = -uy; }
            return false;
        } else if (o.particle_bc == BC_PERIODIC) {
            x = std::fmod(x - o.plo[0], lx); if (x < 0) x += lx; x += o.plo[0];
            y = std::fmod(y - o.plo[1], ly); if (y < 0) y += ly; y += o.plo[1];
            return false;
        }
        return true;
    }
    return false;
}

// AdvancePlasmaParticles, src/particles/pusher/PlasmaParticleAdvance.cpp:29-240 (leapfrog,
// n_subcycles = 1, no laser): gather with the nodal derivative shape (FieldGather.H:45-96),
// 4 momentum sub-steps, position push, EnforceBC (GetAndSetPosition.H:29-99), 2 half sub-steps.
void push_plasma(const Ctx& c, Plasma& p, const double* F, bool temp_slice, int cPsi, int cEz, int cBx,
                 int cBy, int cBz) {
    const Geom& g = c.g;
    const int order = c.o.depos_order;
    const double clight = 1.0, clight_inv = 1.0 / 1.0;
    const double qmc = p.charge / (p.mass * clight);
    const double dz = g.dz / 1;
    const double* Psi = F + cPsi * g.plane;
    const double* Ez = F + cEz * g.plane;
    const double* Bx = F + cBx * g.plane;
    const double* By = F + cBy * g.plane;
    const double* Bz = F + cBz * g.plane;
    const int lo = -g.ng, hix = g.nx - 1 + g.ng, hiy = g.ny - 1 + g.ng;
    const long np = long(p.size());
#pragma omp parallel for
    for (long ip = 0; ip < np; ++ip) {
        if (!p.valid[ip]) continue;
        double xp = p.x_prev[ip], yp = p.y_prev[ip];
        double ExmByp = 0, EypBxp = 0, Ezp = 0, Bxp = 0, Byp = 0, Bzp = 0;
        const double xm = (xp - g.xoff) * g.dxi, ym = (yp - g.yoff) * g.dyi;
        for (int iy = 0; iy <= order + 1; ++iy) {
            for (int ix = 0; ix <= order + 1; ++ix) {
                const DSF sy = dshape(1, order, ym, iy);
                const DSF sx = dshape(1, order, xm, ix);
                if (sx.cell < lo || sx.cell > hix || sy.cell < lo || sy.cell > hiy) continue;
                const long k = g.idx(sx.cell, sy.cell);
                ExmByp += (sx.ds * sy.s) * Psi[k] * g.dxi;
                EypBxp += (sx.s * sy.ds) * Psi[k] * g.dyi;
                Ezp += (sx.s * sy.s) * Ez[k];
                Bxp += (sx.s * sy.s) * Bx[k];
                Byp += (sx.s * sy.s) * By[k];
                Bzp += (sx.s * sy.s) * Bz[k];
            }
        }
        Bxp *= clight;
        Byp *= clight;
        const int nsub = 4;
        const double sdz = dz / nsub;
        double ux = p.ux_h[ip], uy = p.uy_h[ip], psi = p.psi_h[ip];
        auto substep = [&]() {
            const double psi_inv = 1.0 / psi;
            double dux, duy, dpsi;
            momentum_rhs<double>(ux, uy, psi_inv, ExmByp, EypBxp, Ezp, Bxp, Byp, Bzp, clight_inv, qmc, dux, duy, dpsi);
            const Dual uxd{ux, dux}, uyd{uy, duy}, pid{psi_inv, -psi_inv * psi_inv * dpsi};
            Dual ddux(0), dduy(0), ddpsi(0);
            momentum_rhs<Dual>(uxd, uyd, pid, ExmByp, EypBxp, Ezp, Bxp, Byp, Bzp, clight_inv, qmc, ddux, dduy, ddpsi);
            ux += sdz * dux + 0.5 * sdz * sdz * ddux.e;
            uy += sdz * duy + 0.5 * sdz * sdz * dduy.e;
            psi += sdz * dpsi + 0.5 * sdz * sdz * ddpsi.e;
        };
        for (int s = 0; s < nsub; ++s) substep();
        xp += dz * clight_inv * (ux * (1.0 / psi));
        yp += dz * clight_inv * (uy * (1.0 / psi));
        if (enforce_bc(c.o, xp, yp, ux, uy)) {
            p.w[ip] = 0.0;
            p.valid[ip] = 0;
            continue;
        }
        p.x[ip] = xp;
        p.y[ip] = yp;
        if (!temp_slice) {
            p.ux_h[ip] = ux; p.uy_h[ip] = uy; p.psi_h[ip] = psi;
            p.x_prev[ip] = xp; p.y_prev[ip] = yp;
        }
        for (int s = 0; s < nsub / 2; ++s) substep();
        p.ux[ip] = ux; p.uy[ip] = uy; p.psi[ip] = psi;
    }
}

// SalameOnlyAdvancePlasma, src/salame/Salame.cpp:258-330: ux, uy of every valid plasma particle
// from the SALAME-only Bx, By gathered at (x_prev, y_prev) (doBxByGatherShapeN,
// src/particles/particles_utils/FieldGather.H:183-216)
void salame_only_advance(const Ctx& c, Plasma& p, const double* F) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const int order = c.o.depos_order;
    const double* Bx = F + L.sal_Bx * g.plane;
    const double* By = F + L.sal_By * g.plane;
    const double qm = p.charge / p.mass;
    const double dz = g.dz;
    const int lo = -g.ng, hix = g.nx - 1 + g.ng, hiy = g.ny - 1 + g.ng;
    for (size_t ip = 0; ip < p.size(); ++ip) {
        if (!p.valid[ip]) continue;
        const double xmid = (p.x_prev[ip] - g.xoff) * g.dxi;
        const double ymid = (p.y_prev[ip] - g.yoff) * g.dyi;
        double Bxp = 0.0, Byp = 0.0;
        for (int iy = 0; iy <= order; ++iy)
            for (int ix = 0; ix <= order; ++ix) {
                const SF sx = shape(order, xmid, ix), sy = shape(order, ymid, iy);
                if (sx.cell < lo || sx.cell > hix || sy.cell < lo || sy.cell > hiy) continue;
                const long k = g.idx(sx.cell, sy.cell);
                Bxp += sx.s * sy.s * Bx[k];
                Byp += sx.s * sy.s * By[k];
            }
        p.ux[ip] = 1.5 * dz * qm * Byp;
        p.uy[ip] = -1.5 * dz * qm * Bxp;
    }
}

// AdvanceBeamParticlesSlice, src/particles/pusher/BeamParticleAdvance.cpp:20-336 (no
// radiation reaction or spin tracking; external fields affine in position, ApplyExternalField
// src/particles/pusher/ExternalFields.H:28-56), for the particles [first, first+count)
// and then the listed ones that slipped into this slice. A particle below min_z (the lower end
// of the slice) stops sub-cycling (:152-155); with sub-cycles left it is listed for the next
// slice, the role of shiftSlippedParticles (src/particles/sorting/SliceSort.cpp:13-70).
void push_beam_slice(const Ctx& c, const double* F, const BeamRef& b, long first, long count, const int32_t* slip_in,
                     long n_in, std::vector<int32_t>& slip_out, int slice, double dt_step, int n_subcycles,
                     bool do_z_push) {
    const Geom& g = c.g;
    const Layout& L = c.L;
    const int order = c.o.depos_order;
    const double clight = 1.0, inv_c2 = 1.0 / (1.0 * 1.0);
    const double qm = b.charge / b.mass;
    const double dt = dt_step / n_subcycles;
    const double min_z = g.zlo + slice * g.dz;
    const double* Psi = F + L.Psi * g.plane;
    const double* Ez = F + L.Ez * g.plane;
    const double* Bx = F + L.Bx * g.plane;
    const double* By = F + L.By * g.plane;
    const double* Bz = F + L.Bz * g.plane;
    const int lo = -g.ng, hix = g.nx - 1 + g.ng, hiy = g.ny - 1 + g.ng;
    auto push = [&](long ip) {
        if (!b.valid[ip]) return;
        double xp = b.x[ip], yp = b.y[ip], zp = b.z[ip];
        double ux = b.ux[ip], uy = b.uy[ip], uz = b.uz[ip];
        int i = b.nsub[ip];
        for (; i < n_subcycles; i++) {
            if (zp < min_z) break;
            const double gammap_inv = 1.0 / std::sqrt(1.0 + (ux * ux + uy * uy + uz * uz) * inv_c2);
            xp += dt * 0.5 * ux * gammap_inv;
            yp += dt * 0.5 * uy * gammap_inv;
            if (enforce_bc(c.o, xp, yp, ux, uy)) {
                b.w[ip] = 0.0;
                b.valid[ip] = 0;
                return;
            }
            double ExmByp = 0, EypBxp = 0, Ezp = 0, Bxp = 0, Byp = 0, Bzp = 0;
            const double xm = (xp - g.xoff) * g.dxi, ym = (yp - g.yoff) * g.dyi;
            for (int iy = 0; iy <= order + 1; ++iy) {
                for (int ix = 0; ix <= order + 1; ++ix) {
                    const DSF sy = dshape(1, order, ym, iy);
                    const DSF sx = dshape(1, order, xm, ix);
                    if (sx.cell < lo || sx.cell > hix || sy.cell < lo || sy.cell > hiy) continue;
                    const long k = g.idx(sx.cell, sy.cell);
                    ExmByp += (sx.ds * sy.s) * Psi[k] * g.dxi;
                    EypBxp += (sx.s * sy.ds) * Psi[k] * g.dyi;
                    Ezp += (sx.s * sy.s) * Ez[k];
                    Bxp += (sx.s * sy.s) * Bx[k];
                    Byp += (sx.s * sy.s) * By[k];
                    Bzp += (sx.s * sy.s) * Bz[k];
                }
            }
            if (b.ext) {
                double f[6];
                for (int q = 0; q < 6; ++q) {
                    const double* e = b.ext + 4 * q;
                    f[q] = ((e[0] + e[1] * xp) + e[2] * yp) + e[3] * zp;
                }
                ExmByp += f[0] - clight * f[4];
                EypBxp += f[1] + clight * f[3];
                Ezp += f[2];
                Bxp += f[3];
                Byp += f[4];
                Bzp += f[5];
            }
            const double ux_next = ux + dt * qm * (ExmByp + (clight - uz * gammap_inv) * Byp + uy * gammap_inv * Bzp);
            const double uy_next = uy + dt * qm * (EypBxp + (uz * gammap_inv - clight) * Bxp - ux * gammap_inv * Bzp);
            const double ux_i = (ux_next + ux) * 0.5;
            const double uy_i = (uy_next + uy) * 0.5;
            const double uz_i = uz + dt * 0.5 * qm * Ezp;
            const double gamma_i_inv = 1.0 / std::sqrt(1.0 + (ux_i * ux_i + uy_i * uy_i + uz_i * uz_i) * inv_c2);
            const double uz_next = uz + dt * qm * (Ezp + (ux_i * Byp - uy_i * Bxp) * gamma_i_inv);
            const double gamma_next_inv =
                1.0 / std::sqrt(1.0 + (ux_next * ux_next + uy_next * uy_next + uz_next * uz_next) * inv_c2);
            xp += dt * 0.5 * ux_next * gamma_next_inv;
            yp += dt * 0.5 * uy_next * gamma_next_inv;
            if (do_z_push) zp += dt * (uz_next * gamma_next_inv - clight);
            ux = ux_next;
            uy = uy_next;
            uz = uz_next;
        }
        if (enforce_bc(c.o, xp, yp, ux, uy)) {
            b.w[ip] = 0.0;
            b.valid[ip] = 0;
            return;
        }
        b.x[ip] = xp;
        b.y[ip] = yp;
        b.z[ip] = zp;
        b.nsub[ip] = i;
        b.ux[ip] = ux;
        b.uy[ip] = uy;
        b.uz[ip] = uz;
        if (zp < min_z && i < n_subcycles) slip_out.push_back(int32_t(ip));
    };
    for (long t = 0; t < count; ++t) push(first + t);
    for (long t = 0; t < n_in; ++t) push(slip_in[t]);
}

}  // namespace orc
