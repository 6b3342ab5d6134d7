// Double-precision complex element of the complex twins (Complex64 of the crate, which is
// num-complex 0.4's Complex<f64>: two doubles {re, im}, so complex matrices are interleaved
// re/im pairs and a column-major complex matrix with leading dimension ld is, viewed as
// doubles, a 2m x n real matrix with leading dimension 2ld whose even rows are the real parts).
//
// The operators are num-complex's published formulas evaluated unfused, as rustc compiles
// them (no contraction): each body carries its own `fp contract(off)`, so the products inside
// are never fused with anything, wherever the operator is inlined.
//   a * b = (a.re b.re - a.im b.im, a.re b.im + a.im b.re)
//   a / b = ((a.re b.re + a.im b.im) / q, (a.im b.re - a.re b.im) / q), q = b.re^2 + b.im^2
//   a / r = (a.re / r, a.im / r) for real r
#pragma once

#include <hip/hip_runtime.h>

namespace al {

struct __align__(16) zd {
    double re, im;
    zd() = default;
    __host__ __device__ constexpr zd(double r, double i = 0.0) : re(r), im(i) {}
};

__host__ __device__ __forceinline__ zd operator*(zd a, zd b) {
#pragma clang fp contract(off)
    return zd(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ __forceinline__ zd operator/(zd a, zd b) {
#pragma clang fp contract(off)
    const double q = b.re * b.re + b.im * b.im;
    return zd((a.re * b.re + a.im * b.im) / q, (a.im * b.re - a.re * b.im) / q);
}
__host__ __device__ __forceinline__ zd operator/(zd a, double r) { return zd(a.re / r, a.im / r); }
__host__ __device__ __forceinline__ zd operator+(zd a, zd b) { return zd(a.re + b.re, a.im + b.im); }
__host__ __device__ __forceinline__ zd operator-(zd a, zd b) { return zd(a.re - b.re, a.im - b.im); }
__host__ __device__ __forceinline__ zd& operator-=(zd& a, zd b) {
    a = a - b;
    return a;
}
__host__ __device__ __forceinline__ zd conj(zd a) { return zd(a.re, -a.im); }
__host__ __device__ __forceinline__ double norm_sqr(zd a) {
#pragma clang fp contract(off)
    return a.re * a.re + a.im * a.im;
}
// IZAMAX's measure, |Re| + |Im| (src/lapack/complex.rs:52-54)
__host__ __device__ __forceinline__ double abs1(zd a) { return fabs(a.re) + fabs(a.im); }
__host__ __device__ __forceinline__ bool is_nan(zd a) { return a.re != a.re || a.im != a.im; }

}  // namespace al
