// Forward-mode derivatives with N directions at once, for the per-particle Jacobians of the
// adjoint kernels: a value and its partial derivatives with respect to N chosen inputs.
// Arithmetic follows the usual rules (product, quotient and square-root rules); comparisons
// and floors act on the value only, so a computation takes the same branches as its
// plain FP64 counterpart.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace hpq {

template <int N>
struct Tan {
    double v;
    double d[N];
    __host__ __device__ Tan() : v(0) {
        for (int k = 0; k < N; ++k) d[k] = 0;
    }
    __host__ __device__ Tan(double a) : v(a) {
        for (int k = 0; k < N; ++k) d[k] = 0;
    }
    // the input number `k` with value a
    static __host__ __device__ Tan seed(double a, int k) {
        Tan t(a);
        t.d[k] = 1.0;
        return t;
    }
};

template <int N>
__host__ __device__ __forceinline__ double pos_value(const Tan<N>& x) { return x.v; }

template <int N>
__host__ __device__ __forceinline__ Tan<N> operator+(const Tan<N>& a, const Tan<N>& b) {
    Tan<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator-(const Tan<N>& a, const Tan<N>& b) {
    Tan<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator-(const Tan<N>& a) {
    Tan<N> r;
    r.v = -a.v;
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator*(const Tan<N>& a, const Tan<N>& b) {
    Tan<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator/(const Tan<N>& a, const Tan<N>& b) {
    Tan<N> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator+(const Tan<N>& a, double b) {
    Tan<N> r = a;
    r.v += b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator+(double a, const Tan<N>& b) { return b + a; }
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator-(const Tan<N>& a, double b) {
    Tan<N> r = a;
    r.v -= b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator-(double a, const Tan<N>& b) {
    Tan<N> r = -b;
    r.v += a;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator*(const Tan<N>& a, double b) {
    Tan<N> r;
    r.v = a.v * b;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator*(double a, const Tan<N>& b) { return b * a; }
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator/(const Tan<N>& a, double b) { return a * (1.0 / b); }
template <int N>
__host__ __device__ __forceinline__ Tan<N> operator/(double a, const Tan<N>& b) { return Tan<N>(a) / b; }
template <int N>
__host__ __device__ __forceinline__ Tan<N>& operator+=(Tan<N>& a, const Tan<N>& b) { return a = a + b; }
template <int N>
__host__ __device__ __forceinline__ Tan<N>& operator-=(Tan<N>& a, const Tan<N>& b) { return a = a - b; }
template <int N>
__host__ __device__ __forceinline__ Tan<N>& operator*=(Tan<N>& a, const Tan<N>& b) { return a = a * b; }
template <int N>
__host__ __device__ __forceinline__ Tan<N> fma(const Tan<N>& a, const Tan<N>& b, const Tan<N>& c) {
    return a * b + c;
}
template <int N>
__host__ __device__ __forceinline__ Tan<N> sqrt(const Tan<N>& a) {
    Tan<N> r;
    r.v = std::sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int N>
__host__ __device__ __forceinline__ bool operator<(const Tan<N>& a, double b) { return a.v < b; }
template <int N>
__host__ __device__ __forceinline__ bool operator>(const Tan<N>& a, double b) { return a.v > b; }
template <int N>
__host__ __device__ __forceinline__ bool operator<(const Tan<N>& a, const Tan<N>& b) { return a.v < b.v; }
template <int N>
__host__ __device__ __forceinline__ bool operator>(const Tan<N>& a, const Tan<N>& b) { return a.v > b.v; }

}  // namespace hpq
