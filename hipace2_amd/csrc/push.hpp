// Plasma push arithmetic shared by the push kernel and its adjoint: the dual number of the
// second-order Taylor sub-step, the momentum right-hand side and the particle boundary
// parameters. Templates on the number type, so the adjoint instantiates the same
// expressions with forward-mode tangents.
#pragma once
#include <hip/hip_runtime.h>

namespace hpq {

// Forward-mode dual number for the second-order Taylor step (src/utils/DualNumbers.H:13-43)
template <class R>
struct Dual {
    R v, e;
    __host__ __device__ Dual(R a) : v(a), e(0) {}
    __host__ __device__ Dual(R a, R b) : v(a), e(b) {}
};
template <class R>
__host__ __device__ __forceinline__ Dual<R> operator+(const Dual<R>& a, const Dual<R>& b) { return {a.v + b.v, a.e + b.e}; }
template <class R>
__host__ __device__ __forceinline__ Dual<R> operator-(const Dual<R>& a, const Dual<R>& b) { return {a.v - b.v, a.e - b.e}; }
template <class R>
__host__ __device__ __forceinline__ Dual<R> operator*(const Dual<R>& a, const Dual<R>& b) {
    return {a.v * b.v, a.e * b.v + a.v * b.e};
}

// PlasmaMomentumPush, src/particles/pusher/PushPlasmaParticles.H:37-75 (no laser terms)
template <class T, class R>
__host__ __device__ __forceinline__ void momentum_rhs(const T& ux, const T& uy, const T& psi_inv, R ExmBy, R EypBx, R Ez,
                                             R Bx_c, R By_c, R Bz, R ci, R qmc, T& dux, T& duy, T& dpsi) {
    const R A = 0.0, ADx = 0.0, ADy = 0.0;
    const T gamma_psi =
        T(0.5) * psi_inv * psi_inv * (T(1.0) + T(A) + ux * ux * T(ci * ci) + uy * uy * T(ci * ci)) + T(0.5);
    dux = (T(qmc) * (gamma_psi * T(ExmBy) + T(By_c) + (uy * T(Bz)) * psi_inv) - T(ADx) * psi_inv);
    duy = (T(qmc) * (gamma_psi * T(EypBx) - T(Bx_c) - (ux * T(Bz)) * psi_inv) - T(ADy) * psi_inv);
    dpsi = (T(qmc) * T(ci) * ((ux * T(ExmBy) + uy * T(EypBx)) * T(ci) * psi_inv - T(Ez)));
}

struct BC {
    int type;
    double plo0, plo1, phi0, phi1;
};

}  // namespace hpq
