// Exact Guevel-Delhommeau wave term, reference
// src/green_functions/exact_Guevel_Delhommeau.jl:15-59:
//   G(r, z)     = 4 int_0^{pi/2} [Re J + i Re e^zeta] dtheta
//   dG/dr(r, z) = 4 int_0^{pi/2} [Re(i cos t (J - 1/(pi zeta))) + i Re(i cos t e^zeta)] dtheta
//   dG/dz       = G + 2 / hypot(r, z)
//   J = e^zeta (E1(zeta)/pi + i),  zeta = z + i r cos(theta),  z < 0.
// The reference integrates adaptively (HCubature); here the integral is a
// fixed composite Gauss-Legendre rule whose panels are graded geometrically
// toward theta = pi/2, where zeta approaches the real axis and the integrand
// has its logarithmic near-singularity (strength set by |z| / r).  The
// imaginary parts are closed forms: 2 pi e^z J0(r) and -2 pi e^z J1(r).
// Written for host and device so the same code is checked on the CPU.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace mh {
namespace delh {

struct zc {
    double re, im;
};
__host__ __device__ inline zc zadd(zc a, zc b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zc zsub(zc a, zc b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zc zmul(zc a, zc b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zc zscale(zc a, double s) { return {a.re * s, a.im * s}; }
__host__ __device__ inline zc zdiv(zc a, zc b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
__host__ __device__ inline zc zexp(zc a) {
    const double e = exp(a.re);
    double s, c;
    sincos(a.im, &s, &c);
    return {e * c, e * s};
}
__host__ __device__ inline zc zlog(zc a) { return {log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }

// e^z E1(z) for z in the closed upper half plane (principal branch).
__host__ __device__ inline zc exp_e1(zc z) {
    const double az = hypot(z.re, z.im);
    const bool series = az <= 5.0 || (z.re < 0.0 && fabs(z.im) < 0.5 * az && az <= 40.0);
    if (series) {
        // E1(z) = -gamma - log z - sum_{n>=1} (-z)^n / (n n!)
        zc term = {1.0, 0.0}, sum = {0.0, 0.0};
        const zc mz = {-z.re, -z.im};
        for (int n = 1; n < 200; ++n) {
            term = zscale(zmul(term, mz), 1.0 / n);
            const zc add = zscale(term, 1.0 / n);
            sum = zadd(sum, add);
            if (hypot(add.re, add.im) <= 1e-17 * hypot(sum.re, sum.im)) break;
        }
        const zc lz = zlog(z);
        const zc e1 = {-0.57721566490153286061 - lz.re - sum.re, -lz.im - sum.im};
        return zmul(zexp(z), e1);
    }
    // e^z E1(z) = 1/(z+1- 1/(z+3- 4/(z+5- ...))), modified Lentz
    const double tiny = 1e-300;
    zc f = {z.re + 1.0, z.im};
    zc C = f, D = {0.0, 0.0};
    for (int n = 1; n < 2000; ++n) {
        const double a = -(double)n * (double)n;
        const zc b = {z.re + 2.0 * n + 1.0, z.im};
        D = zadd(b, zscale(D, a));
        if (D.re == 0.0 && D.im == 0.0) D.re = tiny;
        C = zadd(b, zdiv({a, 0.0}, C));
        if (C.re == 0.0 && C.im == 0.0) C.re = tiny;
        D = zdiv({1.0, 0.0}, D);
        const zc delta = zmul(C, D);
        f = zmul(f, delta);
        if (fabs(delta.re - 1.0) + fabs(delta.im) < 1e-16) break;
    }
    return zdiv({1.0, 0.0}, f);
}

// 16-point Gauss-Legendre nodes and weights on [-1, 1] (positive half).
__host__ __device__ inline void gl16(int i, double& x, double& w) {
    const double X[8] = {0.0950125098376374401853193, 0.2816035507792589132304605,
                         0.4580167776572273863424194, 0.6178762444026437484466718,
                         0.7554044083550030338951012, 0.8656312023878317438804679,
                         0.9445750230732325760779884, 0.9894009349916499325961542};
    const double W[8] = {0.1894506104550684962853967, 0.1826034150449235888667637,
                         0.1691565193950025381893121, 0.1495959888165767320815017,
                         0.1246289712555338720524763, 0.0951585116824927848099251,
                         0.0622535239386478928628438, 0.0271524594117540948517806};
    const int k = i < 8 ? 7 - i : i - 8;
    x = (i < 8) ? -X[k] : X[k];
    w = W[k];
}

// Re parts of the theta-integrands at theta: re(J) and re(i cos t J') with
// J' = J - 1/(pi zeta); with SECOND, also the integrands of d/dr and d/dz of
// the latter: re(-cos^2 t J'') and re(i cos t J''), J'' = J' + 1/(pi zeta^2).
template <bool SECOND>
__host__ __device__ inline void integrand(double r, double z, double th, double& g, double& gr,
                                          double& grr, double& grz) {
    const double ct = cos(th);
    const zc zeta = {z, r * ct};
    const zc ez = zexp(zeta);
    const zc e1 = exp_e1(zeta);
    const double ip = 0.31830988618379067154;  // 1/pi
    const zc J = {e1.re * ip - ez.im, e1.im * ip + ez.re};  // e^z E1/pi + i e^z
    g = J.re;
    const zc iz = zdiv({1.0, 0.0}, zeta);
    const zc q = zsub(J, zscale(iz, ip));
    gr = -ct * q.im;  // Re(i ct q)
    if (SECOND) {
        const zc q2 = zadd(q, zscale(zmul(iz, iz), ip));
        grr = -ct * ct * q2.re;
        grz = -ct * q2.im;
    }
}

// G, dG/dr, dG/dz of the exact term (dimensionless, z < 0).  With SECOND,
// also d(dG/dr)/dr and d(dG/dr)/dz (for the gradient kernels).
template <bool SECOND = false>
__host__ __device__ inline void delhommeau_exact(double r, double z, double G[2], double Gr[2],
                                                 double Gz[2], double Grr[2] = nullptr,
                                                 double Grz[2] = nullptr) {
    const double half_pi = 1.57079632679489661923;
    // Panels: each at most 6/r wide (the phase of e^{i r cos t} turns by at
    // most 6 rad across one 16-point panel) and at most 3/4 of the distance
    // left to pi/2 (geometric grading toward the near-singularity, which
    // sits |z|/r off the real axis at pi/2); the last panel starts below
    // |z| / (8 (r + |z|)) of pi/2.
    const double eps = fabs(z) / (r + fabs(z) + 1e-300);
    double stop = eps * 0.125;
    if (stop < 1e-12) stop = 1e-12;
    const double wmax = (r > 1e-300) ? 6.0 / r : half_pi;
    double sg = 0.0, sr = 0.0, srr = 0.0, srz = 0.0, a = 0.0;
    for (int panel = 0; panel < 400; ++panel) {
        const double left = half_pi - a;
        double b;
        if (left <= stop || panel == 399) {
            b = half_pi;
        } else {
            double w = 0.75 * left;
            if (w > wmax) w = wmax;
            if (w > 0.5 * half_pi) w = 0.5 * half_pi;
            b = a + w;
        }
        const double m = 0.5 * (a + b), h = 0.5 * (b - a);
        for (int i = 0; i < 16; ++i) {
            double x, w;
            gl16(i, x, w);
            double g, gr, grr = 0.0, grz = 0.0;
            integrand<SECOND>(r, z, m + h * x, g, gr, grr, grz);
            sg += w * h * g;
            sr += w * h * gr;
            if (SECOND) {
                srr += w * h * grr;
                srz += w * h * grz;
            }
        }
        if (b >= half_pi) break;
        a = b;
    }
    const double ez = exp(z);
    const double pi = 3.14159265358979323846;
    G[0] = 4.0 * sg;
    G[1] = 2.0 * pi * ez * j0(r);  // 4 int Re e^zeta = 2 pi e^z J0(r)
    Gr[0] = 4.0 * sr;
    Gr[1] = -2.0 * pi * ez * j1(r);  // 4 int Re(i cos t e^zeta) = -2 pi e^z J1(r)
    Gz[0] = G[0] + 2.0 / hypot(r, z);
    Gz[1] = G[1];
    if (SECOND) {
        const double J0 = j0(r), J1 = j1(r);
        const double dJ1 = (r != 0.0) ? J0 - J1 / r : 0.5;
        Grr[0] = 4.0 * srr;
        Grr[1] = -2.0 * pi * ez * dJ1;
        Grz[0] = 4.0 * srz;
        Grz[1] = Gr[1];
    }
}

}  // namespace delh
}  // namespace mh
