// Forward-mode dual numbers with P tangent directions for the derivative
// kernels.  Dual<P> carries a value and its derivatives along P parameter
// directions; every operation propagates them exactly (chain rule), so a
// kernel written once over a scalar type T yields the FP64 values (T =
// double) or values plus P directional derivatives (T = Dual<P>).
// Branches test values only, as automatic differentiation of the reference
// code does.
#pragma once
#include <hip/hip_runtime.h>

namespace mh {

template <int P>
struct Dual {
    double v;
    double d[P];
};

// 1/x for the tangent arithmetic: the hardware reciprocal estimate refined
// by two Newton steps (error within a few ulp for normal, finite x; the
// quotient's special-case handling of a full division is not needed for
// derivative terms).
__device__ __forceinline__ double drcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double val(double x) { return x; }
template <int P>
__device__ __forceinline__ double val(const Dual<P>& x) {
    return x.v;
}

template <int P>
__device__ __forceinline__ Dual<P> dconst(double c) {
    Dual<P> r;
    r.v = c;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = 0.0;
    return r;
}

// constant of scalar type T: Cst<T>::make(c)
template <class T>
struct Cst;
template <>
struct Cst<double> {
    __device__ __forceinline__ static double make(double c) { return c; }
};
template <int P>
struct Cst<Dual<P>> {
    __device__ __forceinline__ static Dual<P> make(double c) { return dconst<P>(c); }
};

// chain rule helper: r = f(x) with f'(x) = fp
template <int P>
__device__ __forceinline__ Dual<P> chain(const Dual<P>& x, double f, double fp) {
    Dual<P> r;
    r.v = f;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = fp * x.d[m];
    return r;
}

template <int P>
__device__ __forceinline__ Dual<P> operator+(const Dual<P>& a, const Dual<P>& b) {
    Dual<P> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = a.d[m] + b.d[m];
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator-(const Dual<P>& a, const Dual<P>& b) {
    Dual<P> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = a.d[m] - b.d[m];
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator-(const Dual<P>& a) {
    Dual<P> r;
    r.v = -a.v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = -a.d[m];
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator*(const Dual<P>& a, const Dual<P>& b) {
    Dual<P> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = a.d[m] * b.v + a.v * b.d[m];
    return r;
}
// Quotients take one division (the reciprocal): the gradient kernels use
// dual numbers for derivatives, which a one-rounding-different value does
// not change beyond the last bits.
template <int P>
__device__ __forceinline__ Dual<P> operator/(const Dual<P>& a, const Dual<P>& b) {
    Dual<P> r;
    const double ib = drcp(b.v);
    const double q = a.v * ib;
    r.v = q;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = (a.d[m] - q * b.d[m]) * ib;
    return r;
}
// mixed with double
template <int P>
__device__ __forceinline__ Dual<P> operator+(const Dual<P>& a, double b) {
    Dual<P> r = a;
    r.v += b;
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator+(double a, const Dual<P>& b) {
    return b + a;
}
template <int P>
__device__ __forceinline__ Dual<P> operator-(const Dual<P>& a, double b) {
    Dual<P> r = a;
    r.v -= b;
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator-(double a, const Dual<P>& b) {
    Dual<P> r;
    r.v = a - b.v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = -b.d[m];
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator*(const Dual<P>& a, double b) {
    Dual<P> r;
    r.v = a.v * b;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = a.d[m] * b;
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator*(double a, const Dual<P>& b) {
    return b * a;
}
template <int P>
__device__ __forceinline__ Dual<P> operator/(const Dual<P>& a, double b) {
    Dual<P> r;
    const double ib = drcp(b);
    r.v = a.v * ib;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = a.d[m] * ib;
    return r;
}
template <int P>
__device__ __forceinline__ Dual<P> operator/(double a, const Dual<P>& b) {
    const double ib = drcp(b.v);
    const double q = a * ib;
    return chain(b, q, -q * ib);
}
template <int P>
__device__ __forceinline__ Dual<P>& operator+=(Dual<P>& a, const Dual<P>& b) {
    a = a + b;
    return a;
}
template <int P>
__device__ __forceinline__ Dual<P>& operator-=(Dual<P>& a, const Dual<P>& b) {
    a = a - b;
    return a;
}

// elementary functions: double overloads forward to the device libm (so that
// unqualified calls inside namespace mh resolve for both scalar types)
__device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ double exp(double x) { return ::exp(x); }
__device__ __forceinline__ double log(double x) { return ::log(x); }
__device__ __forceinline__ double fabs(double x) { return ::fabs(x); }
__device__ __forceinline__ double atan2(double y, double x) { return ::atan2(y, x); }

template <int P>
__device__ __forceinline__ Dual<P> sqrt(const Dual<P>& x) {
    const double s = ::sqrt(x.v);
    return chain(x, s, 0.5 * drcp(s));
}
template <int P>
__device__ __forceinline__ Dual<P> exp(const Dual<P>& x) {
    const double e = ::exp(x.v);
    return chain(x, e, e);
}
template <int P>
__device__ __forceinline__ Dual<P> log(const Dual<P>& x) {
    return chain(x, ::log(x.v), drcp(x.v));
}
template <int P>
__device__ __forceinline__ Dual<P> fabs(const Dual<P>& x) {
    const double s = (x.v > 0.0) ? 1.0 : ((x.v < 0.0) ? -1.0 : 0.0);
    return chain(x, ::fabs(x.v), s);
}
template <int P>
__device__ __forceinline__ Dual<P> atan2(const Dual<P>& y, const Dual<P>& x) {
    Dual<P> r;
    r.v = ::atan2(y.v, x.v);
    const double den = drcp(x.v * x.v + y.v * y.v);
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = (x.v * y.d[m] - y.v * x.d[m]) * den;
    return r;
}

// Bessel functions of the first and second kind, orders 0 and 1, together
// (each derivative needs the other order): J0' = -J1, J1' = J0 - J1/x,
// Y0' = -Y1, Y1' = Y0 - Y1/x.
//
// For |x| <= 3, J0 and J1 come together from their power series in
// q = x^2/4 (Abramowitz and Stegun 9.1.10), nested from the 14th term:
//   J0 = 1 - q/1^2 (1 - q/2^2 (1 - q/3^2 (...))),
//   J1 = x/2 (1 - q/(1 2) (1 - q/(2 3) (...))),
// within 6e-16 of the library functions on [0, 3] (checked against scipy's
// j0 and j1 on 2e6 points); larger arguments use the library.
constexpr double kBesselSeriesMax = 3.0;
__device__ __forceinline__ void bessel_j01(double x, double& J0, double& J1) {
    if (::fabs(x) <= kBesselSeriesMax) {
        const double q = 0.25 * x * x;
        double s0 = 1.0, s1 = 1.0;
#pragma unroll
        for (int k = 14; k >= 1; --k) {
            s0 = 1.0 - (q * (1.0 / (double)(k * k))) * s0;
            s1 = 1.0 - (q * (1.0 / (double)(k * (k + 1)))) * s1;
        }
        J0 = s0;
        J1 = 0.5 * x * s1;
    } else {
        J0 = ::j0(x);
        J1 = ::j1(x);
    }
}
__device__ __forceinline__ void bessel_y01(double x, double& Y0, double& Y1) {
    Y0 = ::y0(x);
    Y1 = ::y1(x);
}
template <int P>
__device__ __forceinline__ void bessel_j01(const Dual<P>& x, Dual<P>& J0, Dual<P>& J1) {
    double a, b;
    bessel_j01(x.v, a, b);
    J0 = chain(x, a, -b);
    J1 = chain(x, b, (x.v != 0.0) ? a - b * drcp(x.v) : 0.5);
}
template <int P>
__device__ __forceinline__ void bessel_y01(const Dual<P>& x, Dual<P>& Y0, Dual<P>& Y1) {
    const double a = ::y0(x.v), b = ::y1(x.v);
    Y0 = chain(x, a, -b);
    Y1 = chain(x, b, a - b * drcp(x.v));
}

// lift a double with given tangents
template <int P>
__device__ __forceinline__ Dual<P> lift(double v, const double* d, int stride) {
    Dual<P> r;
    r.v = v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = d[m * stride];
    return r;
}

}  // namespace mh
