/*
 * FP64 CPU oracle for the MarineHydro boundary-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm
 * (MarineHydro.jl), used to check the HIP library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product path never does.
 *
 * Every function cites the reference file:line it restates.  Paths are
 * relative to the reference repository root.
 *
 * Pinning (what checks this file independently of its own formulas):
 * tests/test_oracle.py -- mpmath vectors at 30 digits (written by
 * tests/golden/make_golden.py), the reference's own closed forms and
 * constants, quadrature, finite differences, LAPACK, the Hulme heave and
 * surge tables, Froude-Krylov closed forms and the Haskind relation.
 */
#include "mh_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* small vector helpers                                                      */
/* ------------------------------------------------------------------------ */
static double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static double norm3(const double a[3]) { return sqrt(dot3(a, a)); }
static void sub3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static void cross3(const double a[3], const double b[3], double o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double sgn(double x) { return (x > 0) - (x < 0); } /* Julia sign(0)=0 */

/* ------------------------------------------------------------------------ */
/* Rankine panel integral, reference src/green_functions/rankine.jl:24-63    */
/* ------------------------------------------------------------------------ */
double orc_rankine_integral(const double x[3], const orc_panel* p) {
    double rb[3];
    sub3(x, p->center, rb);
    double r = norm3(rb);
    if (r > 7.0 * p->radius) return p->area / r; /* rankine.jl:33-34 */
    double integral = 0.0;
    double GZ = dot3(rb, p->normal); /* :37 */
    double RR[4];
    for (int v = 0; v < 4; ++v) {
        double d[3];
        sub3(x, p->vertices[v], d);
        RR[v] = norm3(d); /* :38 */
    }
    for (int v = 0; v < 4; ++v) { /* :39-60 */
        int nx = (v + 1) % 4;
        double ctn[3];
        sub3(p->vertices[nx], p->vertices[v], ctn);
        double L = norm3(ctn);
        if (L >= 1e-3 * p->radius) { /* :45 */
            double u[3] = {ctn[0] / L, ctn[1] / L, ctn[2] / L};
            double gyx[3], pv[3];
            cross3(p->normal, u, gyx);
            sub3(x, p->vertices[v], pv);
            double GY = dot3(pv, gyx); /* :46 */
            if (fabs(GZ) >= 1e-4 * p->radius) { /* :47-52 */
                double ANT = 2 * GY * L;
                double s = RR[nx] + RR[v];
                double DNT = s * s - L * L + 2 * fabs(GZ) * s;
                integral -= 2 * fabs(GZ) * atan2(ANT, DNT);
            }
            if (fabs(GY) > 1e-5) { /* :53-58 */
                double ANL = RR[nx] + RR[v] + L;
                double DNL = RR[nx] + RR[v] - L;
                integral += GY * log(ANL / DNL);
            }
        }
    }
    return integral;
}

/* Gradient part of both_integral_and_integral_gradient(::Rankine),
 * reference src/green_functions/rankine.jl:65-123 (used by integral_gradient
 * :126-128).  Returns the gradient w.r.t. the source point (wrt_first=0) or
 * w.r.t. the collocation point (wrt_first=1, negated, :118-122). */
void orc_rankine_integral_gradient(const double x[3], const orc_panel* p,
                                   int wrt_first, double g[3]) {
    double rb[3];
    sub3(x, p->center, rb);
    double r = norm3(rb);
    if (r > 7.0 * p->radius) { /* :74-77 */
        double r3 = r * r * r;
        for (int c = 0; c < 3; ++c) g[c] = (1.0 / r3 * rb[c]) * p->area;
    } else {
        g[0] = g[1] = g[2] = 0.0;
        double GZ = dot3(rb, p->normal);
        double RR[4], DRX[4][3];
        for (int v = 0; v < 4; ++v) {
            double d[3];
            sub3(x, p->vertices[v], d);
            RR[v] = norm3(d);
            for (int c = 0; c < 3; ++c) DRX[v][c] = d[c] / RR[v]; /* :83 */
        }
        for (int v = 0; v < 4; ++v) {
            int nx = (v + 1) % 4;
            double ctn[3];
            sub3(p->vertices[nx], p->vertices[v], ctn);
            double L = norm3(ctn);
            if (L >= 1e-3 * p->radius) { /* :90 */
                double u[3] = {ctn[0] / L, ctn[1] / L, ctn[2] / L};
                double GYX[3], pv[3];
                cross3(p->normal, u, GYX);
                sub3(x, p->vertices[v], pv);
                double GY = dot3(pv, GYX);
                double s = RR[nx] + RR[v];
                double ANT = 2 * GY * L;
                double DNT = s * s - L * L + 2 * fabs(GZ) * s;
                double ANL = s + L;
                double DNL = s - L;
                double ALDEN = log(ANL / DNL);
                double AT = (fabs(GZ) >= 1e-4 * p->radius) ? atan2(ANT, DNT)
                                                           : 0.0; /* :99-103 */
                double sg = sgn(GZ);
                double den = ANT * ANT + DNT * DNT;
                for (int c = 0; c < 3; ++c) { /* :105-114 */
                    double ANLX = DRX[nx][c] + DRX[v][c];
                    double ANTX = 2 * L * GYX[c];
                    double DNTX = 2 * (s + fabs(GZ)) * ANLX +
                                  2 * sg * s * p->normal[c];
                    g[c] -= ALDEN * GYX[c] - 2 * sg * AT * p->normal[c] +
                            GY * (DNL - ANL) / (ANL * DNL) * ANLX -
                            2 * fabs(GZ) * (ANTX * DNT - DNTX * ANT) / den;
                }
            }
        }
    }
    if (wrt_first)
        for (int c = 0; c < 3; ++c) g[c] = -g[c];
}

/* ------------------------------------------------------------------------ */
/* Wu approximation, reference src/green_functions/wu.jl                     */
/* The coefficient tables below are the published fits the reference uses;  */
/* they must be identical for parity.                                         */
/* ------------------------------------------------------------------------ */
static double horner(const double* c, int deg, double t) {
    /* c[0] + c[1] t + ... + c[deg] t^deg, evaluated as the reference does:
     * (c[deg-1] + c[deg] t) t, then (c[i] + acc) t, ..., c[0] + acc. */
    double acc = (c[deg - 1] + c[deg] * t) * t;
    for (int i = deg - 2; i >= 1; --i) acc = (c[i] + acc) * t;
    return c[0] + acc;
}
static const double WU_dC[6] = {14.19, -148.24, 847.8, -2318.58, 3168.35,
                                -1590.27}; /* wu.jl:51-66 */
static const double WU_dB[10] = {1.11,      2.894,     -76.765,  1565.35,
                                 -11336.19, 44270.15,  -97014.11, 118879.26,
                                 -76209.82, 19923.28}; /* wu.jl:69-92 */
static const double WU_dA[10] = {2.948,    -24.53,    249.69,  -754.85,
                                 -1187.71, 16370.75,  -48811.41, 68220.87,
                                 -46688.0, 12622.25}; /* wu.jl:94-117 */
static const double WU_D[10] = {0.632,     -40.97,    667.16,    -6072.07,
                                31127.39,  -96293.05, 181856.75, -205690.43,
                                128170.2,  -33744.6}; /* wu.jl:159-182 */
static const double WU_C[8] = {1.268,   -9.747,   209.653, -1397.89,
                               5155.67, -9844.35, 9136.4,  -3272.62}; /* :184-203 */
static const double WU_B[10] = {0.938,     5.737,     -67.92,    796.534,
                                -4780.77,  17137.74,  -36618.81, 44894.06,
                                -29030.24, 7671.22}; /* wu.jl:205-228 */
static const double WU_A[10] = {1.21,      -13.328,  215.896, -1763.96,
                                8418.94,   -24314.21, 42002.57, -41592.9,
                                21859.0,   -4838.6}; /* wu.jl:230-253 */

/* StruveH0, reference wu.jl:295-327 */
double orc_struve_h0(double xx) {
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        double h = 1.909859164 +
                   (-1.909855001 +
                    (0.687514637 +
                     (-0.126164557 + (0.013828813 + -0.000876918 * yy) * yy) *
                         yy) *
                        yy) *
                       yy;
        return h * (xx / 3.0);
    }
    double yy = (3.0 / xx) * (3.0 / xx);
    double c1 = 2.0 * (0.99999906 +
                       (4.77228920 + (3.85542044 + 0.32303607 * yy) * yy) * yy);
    double c2 = PI * xx *
                (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * yy) * yy) * yy);
    return c1 / c2 + y0(xx);
}

/* StruveH1, reference wu.jl:331-362 */
double orc_struve_h1(double xx) {
    if (xx <= 3.0) {
        double yy = (xx / 3.0) * (xx / 3.0);
        return (1.909859286 +
                (-1.145914713 +
                 (0.294656958 +
                  (-0.042070508 + (0.003785727 + -0.000207183 * yy) * yy) *
                      yy) *
                     yy) *
                    yy) *
               yy;
    }
    double yy = (3.0 / xx) * (3.0 / xx);
    double c1 = 2.0 * (1.00000004 +
                       (3.92205313 + (2.64893033 + 0.27450895 * yy) * yy) * yy);
    double c2 =
        PI * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * yy) * yy) * yy);
    return c1 / c2 + y1(xx);
}

/* GF_Func_Lp, wu.jl:266-280 */
static double wu_Lp(double alpha, double beta, double rho) {
    double A = horner(WU_A, 9, rho), B = horner(WU_B, 9, rho);
    double C = horner(WU_C, 7, rho), D = horner(WU_D, 9, rho);
    double RR = (1.0 - beta) * A;
    RR -= beta * B;
    RR -= alpha * C / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * D;
    double om = 1.0 - rho;
    return rho * (om * om * om) * RR;
}
/* GF_Func_L0, wu.jl:282-293 */
static double wu_L0(double vv, double dd, double alpha, double beta,
                    double rho) {
    double gama = 0.5772156649;
    double PP = log(0.5 * (dd - vv)) + gama - 2.0 * (dd * dd);
    PP = exp(vv) * PP;
    PP = PP + dd * dd - vv;
    return 2.0 * PP / (1.0 + dd * dd * dd) + 2.0 * wu_Lp(alpha, beta, rho);
}
/* GF_Func_Lsp, wu.jl:131-143 */
static double wu_Lsp(double alpha, double beta, double rho) {
    double A = horner(WU_dA, 9, rho), B = horner(WU_dB, 9, rho);
    double C = horner(WU_dC, 5, rho);
    double RR = beta * A;
    RR -= (1.0 - alpha) * B;
    RR += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * C;
    double om = 1.0 - rho;
    return rho * (om * om * om) * RR;
}
/* GF_Func_Ls, wu.jl:145-157 */
static double wu_Ls(double hh, double vv, double dd, double alpha, double beta,
                    double rho) {
    double PS = (beta + hh) / (dd - vv);
    PS = PS - 2.0 * beta + 2.0 * exp(vv) * dd - hh;
    double d3 = dd * dd * dd;
    double QS = exp(-dd) * (1.0 - beta);
    QS = QS * (1.0 + dd / (1.0 + d3));
    return 2.0 * PS / (1.0 + d3) - 4.0 * QS + 2.0 * wu_Lsp(alpha, beta, rho);
}

/* greens(::GFWu) / both_greens_and_gradient_greens(::GFWu) kernels,
 * wu.jl:5-13 and :19-30; W and Wh from wu.jl:257-264 and :122-129. */
void orc_wu_dimless(double hh, double vv, double G[2], double dGdh[2],
                    double dGdv[2]) {
    double dd = sqrt(hh * hh + vv * vv);
    double alpha = -vv / dd;
    double beta = hh / dd;
    double rho = dd / (1.0 + dd);
    double ev = exp(vv);
    double twopi = 2.0 * PI;
    /* W = 2 pi (H0 - i J0) e^v */
    double H0 = orc_struve_h0(hh), J0 = j0(hh);
    double Wr = twopi * H0 * ev, Wi = twopi * (-J0) * ev;
    /* Wh = 2 pi (2/pi - H1 + i J1) e^v */
    double H1 = orc_struve_h1(hh), J1 = j1(hh);
    double Whr = twopi * (2.0 / PI - H1) * ev, Whi = twopi * J1 * ev;
    double L0 = wu_L0(vv, dd, alpha, beta, rho);
    double Ls = wu_Ls(hh, vv, dd, alpha, beta, rho);
    G[0] = -L0 - Wr;
    G[1] = -Wi;
    dGdh[0] = -Ls - Whr;
    dGdh[1] = -Whi;
    dGdv[0] = G[0] + 2.0 / dd;
    dGdv[1] = G[1];
}

/* ------------------------------------------------------------------------ */
/* Exact Guevel-Delhommeau, reference exact_Guevel_Delhommeau.jl:15-59       */
/* ------------------------------------------------------------------------ */
typedef long double complex lcplx;
#define EULER_GAMMA_L 0.577215664901532860606512090082402431L
#define PI_L 3.141592653589793238462643383279502884L

/* e^z E1(z), principal branch, for z in the closed upper half plane.
 * Series for small |z| (and along the negative real axis, where the series
 * has no cancellation), Lentz continued fraction elsewhere.  Checked against
 * mpmath at 30 digits in tests/test_oracle.py (tests/golden/delhommeau_mpmath.json). */
static lcplx exp_e1(lcplx z) {
    long double az = cabsl(z);
    long double re = creall(z), im = cimagl(z);
    int use_series = (az <= 8.0L) || (re < 0 && fabsl(im) < 0.6L * az &&
                                      az <= 60.0L);
    if (use_series) {
        /* E1(z) = -gamma - log z - sum_{n>=1} (-z)^n / (n n!) */
        lcplx term = 1.0L, sum = 0.0L;
        for (int n = 1; n < 400; ++n) {
            term *= -z / (long double)n;
            lcplx add = term / (long double)n;
            sum += add;
            if (cabsl(add) <= 1e-22L * cabsl(sum) && n > 2) break;
        }
        lcplx e1 = -EULER_GAMMA_L - clogl(z) - sum;
        return cexpl(z) * e1;
    }
    /* e^z E1(z) = 1/(z+1- 1/(z+3- 4/(z+5- 9/(z+7- ...)))), modified Lentz */
    const long double tiny = 1e-300L;
    lcplx f = z + 1.0L;
    if (cabsl(f) == 0) f = tiny;
    lcplx C = f, D = 0.0L;
    for (int n = 1; n < 5000; ++n) {
        lcplx a = -(long double)n * (long double)n;
        lcplx b = z + (long double)(2 * n + 1);
        D = b + a * D;
        if (cabsl(D) == 0) D = tiny;
        C = b + a / C;
        if (cabsl(C) == 0) C = tiny;
        D = 1.0L / D;
        lcplx delta = C * D;
        f *= delta;
        if (cabsl(delta - 1.0L) < 1e-21L) break;
    }
    return 1.0L / f;
}

void orc_exp_e1(double zr, double zi, double out[2]) {
    lcplx v = exp_e1((long double)zr + I * (long double)zi);
    out[0] = (double)creall(v);
    out[1] = (double)cimagl(v);
}

/* integrands of _dimless_wave_term (:16-23) and _d_dimless_wave_term_dr
 * (:36-43), evaluated together (they share e^zeta E1(zeta)). */
static void delh_integrand(long double r, long double z, long double theta,
                           long double out[4]) {
    long double ct = cosl(theta);
    lcplx zeta = z + I * r * ct;
    lcplx ez = cexpl(zeta);
    lcplx j = exp_e1(zeta) / PI_L + I * ez; /* e^z (E1/pi + i) */
    out[0] = creall(j);
    out[1] = creall(ez);
    lcplx dj = I * ct * (j - 1.0L / (PI_L * zeta));
    out[2] = creall(dj);
    out[3] = creall(I * ct * ez);
}

/* Gauss-Legendre rule of order GLN on [-1,1], built at first use by Newton
 * iteration on P_n. */
#define GLN 24
static long double gl_x[GLN], gl_w[GLN];
static int gl_ready = 0;
static long double legendre_ratio(long double x, long double* dp) {
    long double p0 = 1.0L, p1 = x;
    for (int k = 2; k <= GLN; ++k) {
        long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    *dp = GLN * (x * p1 - p0) / (x * x - 1.0L);
    return p1;
}
static void gl_init(void) {
    if (gl_ready) return;
    for (int i = 0; i < GLN; ++i) {
        long double x = cosl(PI_L * (i + 0.75L) / (GLN + 0.5L)), dp;
        for (int it = 0; it < 40; ++it) x -= legendre_ratio(x, &dp) / dp;
        legendre_ratio(x, &dp);
        gl_x[i] = x;
        gl_w[i] = 2.0L / ((1.0L - x * x) * dp * dp);
    }
    gl_ready = 1;
}
static void gl_panel(long double r, long double z, long double a, long double b,
                     long double out[4]) {
    long double m = 0.5L * (a + b), h = 0.5L * (b - a);
    out[0] = out[1] = out[2] = out[3] = 0;
    for (int i = 0; i < GLN; ++i) {
        long double f[4];
        delh_integrand(r, z, m + h * gl_x[i], f);
        for (int c = 0; c < 4; ++c) out[c] += gl_w[i] * f[c];
    }
    for (int c = 0; c < 4; ++c) out[c] *= h;
}
/* Adaptive bisection: a panel is accepted when its 2-half estimate agrees
 * with the whole-panel estimate to tol_density * width (a local criterion,
 * so rounding, which also scales with the width, never stalls it). */
static void gl_adapt(long double r, long double z, long double a, long double b,
                     const long double whole[4], int depth,
                     long double tol_density, long double out[4]) {
    long double m = 0.5L * (a + b), L[4], R[4];
    gl_panel(r, z, a, m, L);
    gl_panel(r, z, m, b, R);
    long double err = 0;
    for (int c = 0; c < 4; ++c) {
        long double e = fabsl(L[c] + R[c] - whole[c]);
        if (e > err) err = e;
    }
    if (err <= tol_density * (b - a) || depth >= 30) {
        for (int c = 0; c < 4; ++c) out[c] = L[c] + R[c];
        return;
    }
    long double o1[4], o2[4];
    gl_adapt(r, z, a, m, L, depth + 1, tol_density, o1);
    gl_adapt(r, z, m, b, R, depth + 1, tol_density, o2);
    for (int c = 0; c < 4; ++c) out[c] = o1[c] + o2[c];
}

/* 4 * integral over [0, pi/2] (reference :25 and :45) */
void orc_delhommeau_dimless(double r, double z, double G[2], double dGdr[2],
                            double dGdz[2]) {
#pragma omp critical(orc_gl_init)
    gl_init();
    long double whole[4], out[4];
    gl_panel(r, z, 0.0L, PI_L / 2, whole);
    long double scale = fabsl(whole[0]) + fabsl(whole[1]) + fabsl(whole[2]) +
                        fabsl(whole[3]) + 1e-3L;
    gl_adapt(r, z, 0.0L, PI_L / 2, whole, 0, 1e-16L * scale / (PI_L / 2), out);
    G[0] = (double)(4 * out[0]);
    G[1] = (double)(4 * out[1]);
    dGdr[0] = (double)(4 * out[2]);
    dGdr[1] = (double)(4 * out[3]);
    /* both_greens_and_gradient_greens: dG/dz = G + 2/hypot(r,z) (:56) */
    dGdz[0] = G[0] + 2.0 / hypot(r, z);
    dGdz[1] = G[1];
}

/* ------------------------------------------------------------------------ */
/* reduced coordinates, abstract_greens_function.jl:88-155                  */
/* ------------------------------------------------------------------------ */
static double horiz_dist(const double x[3], const double xi[3]) {
    double d2 = (x[0] - xi[0]) * (x[0] - xi[0]) + (x[1] - xi[1]) * (x[1] - xi[1]);
    const double purt = 1e-18;
    if (d2 == 0) return sqrt(d2 + purt * purt) - purt; /* :93-99 */
    return sqrt(d2);
}

typedef void (*dimless_fn)(double, double, double*, double*, double*);

static void wave_value(dimless_fn fn, const double x[3], const double xi[3],
                       double k, double out[2]) {
    double r = k * horiz_dist(x, xi);
    double z = k * (x[2] + xi[2]);
    double G[2], dr[2], dz[2];
    fn(r, z, G, dr, dz);
    out[0] = k * G[0];
    out[1] = k * G[1];
}

static void wave_gradient(dimless_fn fn, const double x[3], const double xi[3],
                          double k, int wrt_first, double gout[6]) {
    double r = k * horiz_dist(x, xi);
    double z = k * (x[2] + xi[2]);
    double G[2], dGdr[2], dGdz[2];
    fn(r, z, G, dGdr, dGdz);
    double d1 = 0.0, d2 = 0.0;
    if (fabs(r) > 1e-6) { /* :133-139, :144-150 */
        if (wrt_first) {
            d1 = k * k / r * (x[0] - xi[0]);
            d2 = k * k / r * (x[1] - xi[1]);
        } else {
            d1 = k * k / r * (xi[0] - x[0]);
            d2 = k * k / r * (xi[1] - x[1]);
        }
    }
    double dzz = k;
    gout[0] = k * (d1 * dGdr[0]);
    gout[1] = k * (d1 * dGdr[1]);
    gout[2] = k * (d2 * dGdr[0]);
    gout[3] = k * (d2 * dGdr[1]);
    gout[4] = k * (dzz * dGdz[0]);
    gout[5] = k * (dzz * dGdz[1]);
}

/* ------------------------------------------------------------------------ */
/* sums over terms, abstract_greens_function.jl:18-83                        */
/* ------------------------------------------------------------------------ */
static void reflect(const double x[3], double o[3]) {
    o[0] = x[0];
    o[1] = x[1];
    o[2] = -x[2];
}

void orc_greens(int m, const double x[3], const double xi[3], double k,
                double out[2]) {
    out[0] = out[1] = 0.0;
    double d[3], xr[3];
    reflect(x, xr);
    if (m & ORC_GF_RANKINE) { /* rankine.jl:8-11 */
        sub3(x, xi, d);
        out[0] += 1.0 / norm3(d);
    }
    if (m & (ORC_GF_RANKINE_REFLECTED | ORC_GF_RANKINE_REFLECTED_NEG)) {
        sub3(xr, xi, d); /* rankine_reflected.jl:1-4 */
        double v = 1.0 / norm3(d);
        if (m & ORC_GF_RANKINE_REFLECTED) out[0] += v;
        if (m & ORC_GF_RANKINE_REFLECTED_NEG) out[0] -= v;
    }
    double w[2];
    if (m & ORC_GF_WU) {
        wave_value(orc_wu_dimless, x, xi, k, w);
        out[0] += w[0];
        out[1] += w[1];
    }
    if (m & ORC_GF_DELHOMMEAU) {
        wave_value(orc_delhommeau_dimless, x, xi, k, w);
        out[0] += w[0];
        out[1] += w[1];
    }
}

/* rankine.jl:13-22 (point gradient), reflected variants
 * rankine_reflected.jl:6-14 */
static void rankine_point_gradient(const double x[3], const double xi[3],
                                   int wrt_first, double g[3]) {
    double d[3];
    sub3(x, xi, d);
    double r = norm3(d);
    double r3 = r * r * r;
    for (int c = 0; c < 3; ++c) {
        double gx = 1.0 / r3 * (xi[c] - x[c]);
        g[c] = wrt_first ? gx : -gx;
    }
}

void orc_gradient_greens(int m, const double x[3], const double xi[3],
                         double k, int wrt_first, double gout[6]) {
    for (int c = 0; c < 6; ++c) gout[c] = 0.0;
    double g[3];
    if (m & ORC_GF_RANKINE) {
        rankine_point_gradient(x, xi, wrt_first, g);
        for (int c = 0; c < 3; ++c) gout[2 * c] += g[c];
    }
    if (m & (ORC_GF_RANKINE_REFLECTED | ORC_GF_RANKINE_REFLECTED_NEG)) {
        /* wrt_first: gradient_greens(Rankine, e1, reflect(e2)); else
         * gradient_greens(Rankine, reflect(e1), e2) */
        double a[3], b[3];
        if (wrt_first) {
            a[0] = x[0]; a[1] = x[1]; a[2] = x[2];
            reflect(xi, b);
        } else {
            reflect(x, a);
            b[0] = xi[0]; b[1] = xi[1]; b[2] = xi[2];
        }
        rankine_point_gradient(a, b, wrt_first, g);
        double s = 0.0;
        if (m & ORC_GF_RANKINE_REFLECTED) s += 1.0;
        if (m & ORC_GF_RANKINE_REFLECTED_NEG) s -= 1.0;
        for (int c = 0; c < 3; ++c) gout[2 * c] += s * g[c];
    }
    double w[6];
    if (m & ORC_GF_WU) {
        wave_gradient(orc_wu_dimless, x, xi, k, wrt_first, w);
        for (int c = 0; c < 6; ++c) gout[c] += w[c];
    }
    if (m & ORC_GF_DELHOMMEAU) {
        wave_gradient(orc_delhommeau_dimless, x, xi, k, wrt_first, w);
        for (int c = 0; c < 6; ++c) gout[c] += w[c];
    }
}

void orc_integral(int m, const double x[3], const orc_panel* p, double k,
                  double out[2]) {
    out[0] = out[1] = 0.0;
    if (m & ORC_GF_RANKINE) out[0] += orc_rankine_integral(x, p);
    if (m & (ORC_GF_RANKINE_REFLECTED | ORC_GF_RANKINE_REFLECTED_NEG)) {
        double xr[3];
        reflect(x, xr); /* rankine_reflected.jl:19-22 */
        double v = orc_rankine_integral(xr, p);
        if (m & ORC_GF_RANKINE_REFLECTED) out[0] += v;
        if (m & ORC_GF_RANKINE_REFLECTED_NEG) out[0] -= v;
    }
    double w[2];
    if (m & ORC_GF_WU) { /* wu.jl:33-36: one-point rule */
        wave_value(orc_wu_dimless, x, p->center, k, w);
        out[0] += w[0] * p->area;
        out[1] += w[1] * p->area;
    }
    if (m & ORC_GF_DELHOMMEAU) { /* exact_Guevel_Delhommeau.jl:61-64 */
        wave_value(orc_delhommeau_dimless, x, p->center, k, w);
        out[0] += w[0] * p->area;
        out[1] += w[1] * p->area;
    }
}

void orc_integral_gradient(int m, const double x[3], const orc_panel* p,
                           double k, int wrt_first, double gout[6]) {
    for (int c = 0; c < 6; ++c) gout[c] = 0.0;
    double g[3];
    if (m & ORC_GF_RANKINE) {
        orc_rankine_integral_gradient(x, p, wrt_first, g);
        for (int c = 0; c < 3; ++c) gout[2 * c] += g[c];
    }
    if (m & (ORC_GF_RANKINE_REFLECTED | ORC_GF_RANKINE_REFLECTED_NEG)) {
        double xr[3];
        reflect(x, xr); /* rankine_reflected.jl:28-35 */
        orc_rankine_integral_gradient(xr, p, wrt_first, g);
        if (wrt_first) g[2] = -g[2];
        double s = 0.0;
        if (m & ORC_GF_RANKINE_REFLECTED) s += 1.0;
        if (m & ORC_GF_RANKINE_REFLECTED_NEG) s -= 1.0;
        for (int c = 0; c < 3; ++c) gout[2 * c] += s * g[c];
    }
    double w[6];
    if (m & ORC_GF_WU) { /* wu.jl:38-41 */
        wave_gradient(orc_wu_dimless, x, p->center, k, wrt_first, w);
        for (int c = 0; c < 6; ++c) gout[c] += w[c] * p->area;
    }
    if (m & ORC_GF_DELHOMMEAU) { /* exact_Guevel_Delhommeau.jl:66-69 */
        wave_gradient(orc_delhommeau_dimless, x, p->center, k, wrt_first, w);
        for (int c = 0; c < 6; ++c) gout[c] += w[c] * p->area;
    }
}

/* ------------------------------------------------------------------------ */
/* assembly, reference src/matrix_assembly.jl:4-42 (comprehension variant,   */
/* the default `assemble_matrices`, :97)                                     */
/* ------------------------------------------------------------------------ */
static void load_panel(int64_t nf, const double* V, int64_t nv,
                       const int64_t* F, const double* C, const double* N,
                       const double* A, const double* R, int64_t j,
                       orc_panel* p) {
    for (int c = 0; c < 3; ++c) {
        p->center[c] = C[j + c * nf];
        p->normal[c] = N[j + c * nf];
    }
    for (int v = 0; v < 4; ++v) {
        int64_t iv = F[j + v * nf]; /* 0-based, src/meshes.jl:179 */
        for (int c = 0; c < 3; ++c) p->vertices[v][c] = V[iv + c * nv];
    }
    p->area = A[j];
    p->radius = R[j];
}

void orc_assemble(int m, int64_t nf, const double* V, int64_t nv,
                  const int64_t* F, const double* C, const double* N,
                  const double* A, const double* R, double k, int direct,
                  double* S, double* DK, int nthreads) {
    const double c4 = -1.0 / (4.0 * PI); /* -1/2tau, :14 */
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
    for (int64_t j = 0; j < nf; ++j) {
        orc_panel pj;
        load_panel(nf, V, nv, F, C, N, A, R, j, &pj);
        for (int64_t i = 0; i < nf; ++i) {
            double x[3] = {C[i], C[i + nf], C[i + 2 * nf]};
            double s[2], g[6];
            orc_integral(m, x, &pj, k, s);
            orc_integral_gradient(m, x, &pj, k, !direct, g);
            const double* n;
            double ni[3];
            if (direct) {
                n = pj.normal;
            } else {
                ni[0] = N[i]; ni[1] = N[i + nf]; ni[2] = N[i + 2 * nf];
                n = ni;
            }
            double dr = n[0] * g[0] + n[1] * g[2] + n[2] * g[4];
            double di = n[0] * g[1] + n[1] * g[3] + n[2] * g[5];
            int64_t idx = 2 * (i + j * nf);
            S[idx] = c4 * s[0];
            S[idx + 1] = c4 * s[1];
            DK[idx] = (i == j ? 0.5 : 0.0) + c4 * dr; /* :16-26 */
            DK[idx + 1] = c4 * di;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* dense complex LU (zgetrf / zgetrs semantics)                              */
/* ------------------------------------------------------------------------ */
typedef double complex zc;

int orc_zgetrf(int64_t n, double* Ad, int64_t* ipiv) {
    zc* A = (zc*)Ad;
    int info = 0;
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = k;
        double best = -1.0;
        for (int64_t i = k; i < n; ++i) {
            zc a = A[i + k * n];
            double v = fabs(creal(a)) + fabs(cimag(a));
            if (v > best) {
                best = v;
                p = i;
            }
        }
        ipiv[k] = p;
        if (best == 0.0) {
            if (!info) info = (int)(k + 1);
            continue;
        }
        if (p != k)
            for (int64_t j = 0; j < n; ++j) {
                zc t = A[k + j * n];
                A[k + j * n] = A[p + j * n];
                A[p + j * n] = t;
            }
        zc inv = 1.0 / A[k + k * n];
        for (int64_t i = k + 1; i < n; ++i) A[i + k * n] *= inv;
#pragma omp parallel for schedule(static) if (n - k > 256)
        for (int64_t j = k + 1; j < n; ++j) {
            zc u = A[k + j * n];
            if (u == 0) continue;
            zc* cj = A + j * n;
            const zc* ck = A + k * n;
            for (int64_t i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
        }
    }
    return info;
}

void orc_zgetrs(int64_t n, const double* LUd, const int64_t* ipiv, int trans,
                int64_t nrhs, double* Bd) {
    const zc* LU = (const zc*)LUd;
    zc* B = (zc*)Bd;
    for (int64_t r = 0; r < nrhs; ++r) {
        zc* b = B + r * n;
        if (!trans) {
            for (int64_t k = 0; k < n; ++k)
                if (ipiv[k] != k) {
                    zc t = b[k];
                    b[k] = b[ipiv[k]];
                    b[ipiv[k]] = t;
                }
            for (int64_t j = 0; j < n; ++j) /* L y = b, unit lower */
                for (int64_t i = j + 1; i < n; ++i) b[i] -= LU[i + j * n] * b[j];
            for (int64_t j = n - 1; j >= 0; --j) { /* U x = y */
                b[j] /= LU[j + j * n];
                for (int64_t i = 0; i < j; ++i) b[i] -= LU[i + j * n] * b[j];
            }
        } else {
            for (int64_t i = 0; i < n; ++i) { /* U^T y = b */
                zc s = b[i];
                for (int64_t k = 0; k < i; ++k) s -= LU[k + i * n] * b[k];
                b[i] = s / LU[i + i * n];
            }
            for (int64_t i = n - 1; i >= 0; --i) { /* L^T z = y */
                zc s = b[i];
                for (int64_t k = i + 1; k < n; ++k) s -= LU[k + i * n] * b[k];
                b[i] = s;
            }
            for (int64_t k = n - 1; k >= 0; --k)
                if (ipiv[k] != k) {
                    zc t = b[k];
                    b[k] = b[ipiv[k]];
                    b[ipiv[k]] = t;
                }
        }
    }
}

static void zgemv(int64_t n, const zc* A, const zc* x, zc* y) {
    for (int64_t i = 0; i < n; ++i) y[i] = 0;
    for (int64_t j = 0; j < n; ++j) {
        zc xj = x[j];
        const zc* c = A + j * n;
        for (int64_t i = 0; i < n; ++i) y[i] += c[i] * xj;
    }
}

/* solve(D, S, bc; direct), reference src/matrix_assembly.jl:105-133 */
int orc_solve(int64_t n, double* DKd, const double* Sd, const double* bcd,
              int64_t nrhs, int direct, double* phid) {
    int64_t* ipiv = (int64_t*)malloc(sizeof(int64_t) * n);
    int info = orc_zgetrf(n, DKd, ipiv);
    const zc* S = (const zc*)Sd;
    const zc* bc = (const zc*)bcd;
    zc* phi = (zc*)phid;
    zc* tmp = (zc*)malloc(sizeof(zc) * n * nrhs);
    if (direct) { /* phi = D \ (S*bc) */
        for (int64_t r = 0; r < nrhs; ++r) zgemv(n, S, bc + r * n, phi + r * n);
        orc_zgetrs(n, DKd, ipiv, 0, nrhs, phid);
    } else { /* sigma = K \ bc; phi = S*sigma */
        memcpy(tmp, bc, sizeof(zc) * n * nrhs);
        orc_zgetrs(n, DKd, ipiv, 0, nrhs, (double*)tmp);
        for (int64_t r = 0; r < nrhs; ++r) zgemv(n, S, tmp + r * n, phi + r * n);
    }
    free(tmp);
    free(ipiv);
    return info;
}

/* ------------------------------------------------------------------------ */
/* one frequency end to end: src/waves.jl:4-59 and :112-142                 */
/* ------------------------------------------------------------------------ */
int orc_freq_solve(int m, int direct, int64_t nf, const double* V, int64_t nv,
                   const int64_t* F, const double* C, const double* N,
                   const double* A, const double* R, double omega, double g,
                   double rho, int64_t ndof, const double* dofs,
                   int64_t nbeta, const double* betas, double* AB, double* Fd,
                   int nthreads) {
    double k = omega * omega / g; /* waves.jl:30 */
    zc* S = (zc*)malloc(sizeof(zc) * nf * nf);
    zc* DK = (zc*)malloc(sizeof(zc) * nf * nf);
    orc_assemble(m, nf, V, nv, F, C, N, A, R, k, direct, (double*)S,
                 (double*)DK, nthreads);
    int64_t nrhs = ndof + nbeta;
    zc* bc = (zc*)malloc(sizeof(zc) * nf * nrhs);
    zc* phi = (zc*)malloc(sizeof(zc) * nf * nrhs);
    for (int64_t d = 0; d < ndof; ++d) { /* radiation_bc, waves.jl:89-102 */
        const double* f = dofs + d * nf * 3;
        for (int64_t i = 0; i < nf; ++i) {
            double nd = N[i] * f[i] + N[i + nf] * f[i + nf] +
                        N[i + 2 * nf] * f[i + 2 * nf];
            bc[i + d * nf] = -I * omega * nd;
        }
    }
    for (int64_t b = 0; b < nbeta; ++b) { /* AiryBC, waves.jl:121-142 */
        double cb = cos(betas[b]), sb = sin(betas[b]);
        for (int64_t i = 0; i < nf; ++i) {
            double x = C[i], y = C[i + nf], z = C[i + 2 * nf];
            double wbar = x * cb + y * sb;
            double cih = exp(k * z);
            zc pre = g * k / omega * cexp(I * k * wbar);
            zc v0 = pre * (cb * cih), v1 = pre * (sb * cih),
               v2 = pre * (-I * cih);
            bc[i + (ndof + b) * nf] =
                -(v0 * N[i] + v1 * N[i + nf] + v2 * N[i + 2 * nf]);
        }
    }
    int info = orc_solve(nf, (double*)DK, (double*)S, (double*)bc, nrhs,
                         direct, (double*)phi);
    /* pressure and integrate_pressure, waves.jl:21-37, :55-59, :72-85 */
    for (int64_t rj = 0; rj < nrhs; ++rj) {
        for (int64_t di = 0; di < ndof; ++di) {
            const double* f = dofs + di * nf * 3;
            zc force = 0;
            for (int64_t i = 0; i < nf; ++i) {
                zc p = I * rho * omega * phi[i + rj * nf];
                double amp = -(f[i] * N[i] + f[i + nf] * N[i + nf] +
                               f[i + 2 * nf] * N[i + 2 * nf]);
                force += p * amp * A[i];
            }
            if (rj < ndof) {
                AB[di + rj * ndof] = creal(force) / (omega * omega);
                AB[ndof * ndof + di + rj * ndof] = cimag(force) / omega;
            } else {
                int64_t b = rj - ndof;
                Fd[2 * (di + b * ndof)] = creal(force);
                Fd[2 * (di + b * ndof) + 1] = cimag(force);
            }
        }
    }
    free(S);
    free(DK);
    free(bc);
    free(phi);
    return info;
}

/* ------------------------------------------------------------------------ */
/* Froude-Krylov force, src/waves.jl:62-66 / :150-164: the incident-wave     */
/* pressure p = i omega rho phi_Airy (airy_waves_pressure, :150-154, with    */
/* airy_waves_potential :112-119) at the panel centres, integrated by        */
/* integrate_pressure (:21-25): F_FK = sum_i p_i (-n_i . dof_i) a_i.          */
/* FK[2*(d + b*ndof)] = Re, [+1] = Im for influenced dof d and heading b.    */
/* ------------------------------------------------------------------------ */
void orc_froude_krylov(int64_t nf, const double* C, const double* N,
                       const double* A, double omega, double g, double rho,
                       int64_t ndof, const double* dofs, int64_t nbeta,
                       const double* betas, double* FK) {
    double k = omega * omega / g;
    for (int64_t b = 0; b < nbeta; ++b) {
        double cb = cos(betas[b]), sb = sin(betas[b]);
        for (int64_t d = 0; d < ndof; ++d) {
            const double* f = dofs + d * nf * 3;
            zc force = 0;
            for (int64_t i = 0; i < nf; ++i) {
                double x = C[i], y = C[i + nf], z = C[i + 2 * nf];
                double wbar = x * cb + y * sb;
                zc phi0 = -I * g / omega * exp(k * z) * cexp(I * k * wbar);
                zc p = I * omega * rho * phi0;
                double amp = -(f[i] * N[i] + f[i + nf] * N[i + nf] +
                               f[i + 2 * nf] * N[i + 2 * nf]);
                force += p * amp * A[i];
            }
            FK[2 * (d + b * ndof)] = creal(force);
            FK[2 * (d + b * ndof) + 1] = cimag(force);
        }
    }
}
