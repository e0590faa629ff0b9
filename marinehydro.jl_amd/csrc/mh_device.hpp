// Device-side Green-function terms for the boundary-element pair kernels.
//
// Each function restates one term of the reference (MarineHydro.jl, file:line
// given per function) for gfx950.  They are written once over a scalar type T:
// T = double is the FP64 assembly, T = Dual<P> (mh_dual.hpp) carries P
// directional derivatives for the gradient kernels.  The Rankine near-field
// integral and its gradient are evaluated in one pass over the four edges
// (the reference walks the edges twice, once in `integral` and once in
// `both_integral_and_integral_gradient`), and the wave term returns value and
// both derivatives together.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "delh_table.hpp"
#include "delhommeau.hpp"
#include "mh_dual.hpp"

namespace mh {

constexpr double kPi = 3.14159265358979323846;
constexpr double kInv4Pi = 1.0 / (4.0 * kPi);

// Green-function term mask (same bits as the C-ABI header).
enum : int {
    GF_RANKINE = 1,
    GF_RANKINE_REFLECTED = 2,
    GF_RANKINE_REFLECTED_NEG = 4,
    GF_WU = 8,
    GF_DELHOMMEAU = 16,
    GF_DELH_TABLE = 32,  // the Delhommeau term from the table (delh_table.hpp)
};
constexpr int kGfWaveMask = GF_WU | GF_DELHOMMEAU | GF_DELH_TABLE;
// Internal bit (never accepted from the C-ABI): the Rankine and image far
// field of the FP64 pair code evaluated in FP32 (rankine_panel).
constexpr int kGfFar32 = 1 << 8;

// Packed source-panel record (one per face, 192 B), read wave-uniformly by
// the pair kernels so it lands in scalar registers.
struct __attribute__((aligned(64))) PanelRec {
    double c[3];
    double n[3];
    double area;
    double radius;
    double v[4][3];
    double pad[4];
};
static_assert(sizeof(PanelRec) == 192, "PanelRec layout");

// Source panel with T-valued geometry (radius only selects a branch).
template <class T>
struct PanelT {
    T c[3];
    T n[3];
    T area;
    double radius;
    T v[4][3];
};

template <class T>
struct cplxT {
    T re, im;
};
typedef cplxT<double> cplx;

// ---------------------------------------------------------------------------
// Rankine panel integral and gradient, reference
// src/green_functions/rankine.jl:24-63 (value) and :65-123 (gradient).
// Output: I = integral of 1/|x - xi| over the panel; g = gradient w.r.t. the
// source point xi (the reference's with_respect_to_first_variable=false).
// PT: PanelRec (T = double) or PanelT<T>.
// ---------------------------------------------------------------------------
template <class T, class PT>
__device__ __forceinline__ void rankine_panel(T x0, T x1, T x2, const PT& p, T& I, T g[3], bool far32 = false) {
    const T rb0 = x0 - p.c[0], rb1 = x1 - p.c[1], rb2 = x2 - p.c[2];
    const double R = p.radius;
    if constexpr (sizeof(T) == sizeof(double)) {
        // far field in FP32 (the offsets formed in FP64, then area / r and
        // its gradient from the FP32 reciprocal square root of r^2): the same
        // branch as below, decided on r^2 > 49 R^2 instead of r > 7 R
        if (far32) {
            const double r2 = rb0 * rb0 + rb1 * rb1 + rb2 * rb2;
            if (r2 > 49.0 * R * R) {
                const float ir = rsqrtf((float)r2);
                const float a = (float)p.area;
                I = (double)(a * ir);
                const float s = (ir * ir) * (ir * a);
                g[0] = (double)((float)rb0 * s);
                g[1] = (double)((float)rb1 * s);
                g[2] = (double)((float)rb2 * s);
                return;
            }
        }
    }
    const T r = sqrt(rb0 * rb0 + rb1 * rb1 + rb2 * rb2);
    if (val(r) > 7.0 * R) {  // rankine.jl:33-34, :74-77
        const T ir = 1.0 / r;
        I = p.area * ir;
        const T s = (ir * ir * ir) * p.area;
        g[0] = rb0 * s;
        g[1] = rb1 * s;
        g[2] = rb2 * s;
        return;
    }
    const T n0 = p.n[0], n1 = p.n[1], n2 = p.n[2];
    const T GZ = rb0 * n0 + rb1 * n1 + rb2 * n2;
    const T aGZ = fabs(GZ);
    const double sg = (val(GZ) > 0.0) ? 1.0 : ((val(GZ) < 0.0) ? -1.0 : 0.0);
    const bool use_at = val(aGZ) >= 1e-4 * R;
    T RR[4], D[4][3];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        D[v][0] = x0 - p.v[v][0];
        D[v][1] = x1 - p.v[v][1];
        D[v][2] = x2 - p.v[v][2];
        RR[v] = sqrt(D[v][0] * D[v][0] + D[v][1] * D[v][1] + D[v][2] * D[v][2]);
    }
    T acc = Cst<T>::make(0.0), g0 = Cst<T>::make(0.0), g1 = Cst<T>::make(0.0),
      g2 = Cst<T>::make(0.0);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int w = (v + 1) & 3;
        const T e0 = p.v[w][0] - p.v[v][0];
        const T e1 = p.v[w][1] - p.v[v][1];
        const T e2 = p.v[w][2] - p.v[v][2];
        const T L = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        if (val(L) >= 1e-3 * R) {  // degenerate edges skipped, rankine.jl:45, :90
            const T iL = 1.0 / L;
            const T u0 = e0 * iL, u1 = e1 * iL, u2 = e2 * iL;
            const T y0 = n1 * u2 - n2 * u1;  // GYX = n x (edge / L)
            const T y1 = n2 * u0 - n0 * u2;
            const T y2 = n0 * u1 - n1 * u0;
            const T GY = D[v][0] * y0 + D[v][1] * y1 + D[v][2] * y2;
            const T s = RR[w] + RR[v];
            const T ANT = 2.0 * GY * L;
            const T DNT = s * s - L * L + 2.0 * aGZ * s;
            const T ANL = s + L;
            const T DNL = s - L;
            const T ALDEN = log(ANL / DNL);
            const T AT = use_at ? atan2(ANT, DNT) : Cst<T>::make(0.0);
            // value (rankine.jl:47-58)
            if (use_at) acc -= 2.0 * aGZ * AT;
            if (::fabs(val(GY)) > 1e-5) acc += GY * ALDEN;
            // gradient (rankine.jl:105-114)
            const T iw = 1.0 / RR[w], iv = 1.0 / RR[v];
            const T a0 = D[w][0] * iw + D[v][0] * iv;  // ANLX
            const T a1 = D[w][1] * iw + D[v][1] * iv;
            const T a2 = D[w][2] * iw + D[v][2] * iv;
            const T cL = GY * (DNL - ANL) / (ANL * DNL);
            const T den = ANT * ANT + DNT * DNT;
            const T cT = 2.0 * aGZ / den;
            const T dnt_a = 2.0 * (s + aGZ);
            const T dnt_n = (2.0 * sg) * s;
            const T at_n = (2.0 * sg) * AT;
            {
                const T ANTX = 2.0 * L * y0, DNTX = dnt_a * a0 + dnt_n * n0;
                g0 -= ALDEN * y0 - at_n * n0 + cL * a0 - cT * (ANTX * DNT - DNTX * ANT);
            }
            {
                const T ANTX = 2.0 * L * y1, DNTX = dnt_a * a1 + dnt_n * n1;
                g1 -= ALDEN * y1 - at_n * n1 + cL * a1 - cT * (ANTX * DNT - DNTX * ANT);
            }
            {
                const T ANTX = 2.0 * L * y2, DNTX = dnt_a * a2 + dnt_n * n2;
                g2 -= ALDEN * y2 - at_n * n2 + cL * a2 - cT * (ANTX * DNT - DNTX * ANT);
            }
        }
    }
    I = acc;
    g[0] = g0;
    g[1] = g1;
    g[2] = g2;
}

// ---------------------------------------------------------------------------
// Wu approximation of the wave term, reference src/green_functions/wu.jl.
// Coefficients are the published fits the reference uses (must be identical).
// Struve H0/H1 (wu.jl:295-362) take Y0/Y1 precomputed (used for x > 3).
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T struve_h0(T xx, const T& Y0) {  // wu.jl:295-327
    if (val(xx) <= 3.0) {
        const T yy = (xx / 3.0) * (xx / 3.0);
        const T h = 1.909859164 +
                    (-1.909855001 +
                     (0.687514637 +
                      (-0.126164557 + (0.013828813 + -0.000876918 * yy) * yy) * yy) *
                         yy) *
                        yy;
        return h * (xx / 3.0);
    }
    const T yy = (3.0 / xx) * (3.0 / xx);
    const T c1 = 2.0 * (0.99999906 + (4.77228920 + (3.85542044 + 0.32303607 * yy) * yy) * yy);
    const T c2 = (kPi * xx) * (1.0 + (4.88331068 + (4.28957333 + 0.52120508 * yy) * yy) * yy);
    return c1 / c2 + Y0;
}

template <class T>
__device__ __forceinline__ T struve_h1(T xx, const T& Y1) {  // wu.jl:331-362
    if (val(xx) <= 3.0) {
        const T yy = (xx / 3.0) * (xx / 3.0);
        return (1.909859286 +
                (-1.145914713 +
                 (0.294656958 +
                  (-0.042070508 + (0.003785727 + -0.000207183 * yy) * yy) * yy) *
                     yy) *
                    yy) *
               yy;
    }
    const T yy = (3.0 / xx) * (3.0 / xx);
    const T c1 = 2.0 * (1.00000004 + (3.92205313 + (2.64893033 + 0.27450895 * yy) * yy) * yy);
    const T c2 = kPi * (1.0 + (3.81095112 + (2.26216956 + 0.10885141 * yy) * yy) * yy);
    return c1 / c2 + Y1;
}

// Horner in the reference's association: (c[d-1] + c[d] t) t, (c[i] + acc) t,
// ..., c[0] + acc.  For a dual argument the value is formed the same way and
// the polynomial's derivative in t alongside (one extra FMA per step), then
// chained to t's tangents once.
template <class T>
struct DualWidth {
    static constexpr int value = 0;
};
template <int P>
struct DualWidth<Dual<P>> {
    static constexpr int value = P;
};
template <int DEG, class T>
__device__ __forceinline__ T horner_ref(const double (&c)[DEG + 1], T t) {
    if constexpr (DualWidth<T>::value > 0) {
        const double x = t.v;
        double acc = (c[DEG - 1] + c[DEG] * x) * x;
        double dacc = c[DEG - 1] + 2.0 * c[DEG] * x;
#pragma unroll
        for (int i = DEG - 2; i >= 1; --i) {
            const double u = c[i] + acc;
            dacc = dacc * x + u;
            acc = u * x;
        }
        return chain(t, c[0] + acc, dacc);
    } else {
        T acc = (c[DEG - 1] + c[DEG] * t) * t;
#pragma unroll
        for (int i = DEG - 2; i >= 1; --i) acc = (c[i] + acc) * t;
        return c[0] + acc;
    }
}

// G(h, v), dG/dh, dG/dv of the Wu wave term in reduced coordinates
// (greens wu.jl:5-13, both_greens_and_gradient_greens wu.jl:19-30).
// If WANT_GRAD is false only G is computed (the reference's `greens`).
template <bool WANT_GRAD, class T>
__device__ __forceinline__ void wu_dimless(T hh, T vv, cplxT<T>& G, cplxT<T>& dGdh,
                                           cplxT<T>& dGdv) {
    constexpr double A[10] = {1.21, -13.328, 215.896, -1763.96, 8418.94,
                              -24314.21, 42002.57, -41592.9, 21859.0, -4838.6};
    constexpr double B[10] = {0.938, 5.737, -67.92, 796.534, -4780.77,
                              17137.74, -36618.81, 44894.06, -29030.24, 7671.22};
    constexpr double Cc[8] = {1.268, -9.747, 209.653, -1397.89, 5155.67, -9844.35, 9136.4, -3272.62};
    constexpr double Dd[10] = {0.632, -40.97, 667.16, -6072.07, 31127.39,
                               -96293.05, 181856.75, -205690.43, 128170.2, -33744.6};
    const T dd = sqrt(hh * hh + vv * vv);
    const T alpha = -vv / dd;
    const T beta = hh / dd;
    const T rho = dd / (1.0 + dd);
    const T ev = exp(vv);
    const T d3 = dd * dd * dd;
    const T om = 1.0 - rho;
    const T om3 = om * om * om;
    T J0, J1, Y0, Y1;
    bessel_j01(hh, J0, J1);
    if (val(hh) > 3.0) {
        bessel_y01(hh, Y0, Y1);
    } else {
        Y0 = Cst<T>::make(0.0);
        Y1 = Cst<T>::make(0.0);
    }
    // L0 (wu.jl:282-293) with Lp (wu.jl:266-280)
    T PP = log(0.5 * (dd - vv)) + 0.5772156649 - 2.0 * (dd * dd);
    PP = ev * PP;
    PP = PP + dd * dd - vv;
    T RR = (1.0 - beta) * horner_ref<9>(A, rho);
    RR -= beta * horner_ref<9>(B, rho);
    RR -= alpha * horner_ref<7>(Cc, rho) / (1.0 + 6.0 * alpha * rho * (1.0 - rho));
    RR += beta * (1.0 - beta) * horner_ref<9>(Dd, rho);
    const T L0 = 2.0 * PP / (1.0 + d3) + 2.0 * (rho * om3 * RR);
    const double twopi = 2.0 * kPi;
    const T H0 = struve_h0(hh, Y0);
    G.re = -L0 - twopi * H0 * ev;  // -L0 - W, W = 2 pi (H0 - i J0) e^v
    G.im = twopi * J0 * ev;
    if (WANT_GRAD) {
        constexpr double dA[10] = {2.948, -24.53, 249.69, -754.85, -1187.71,
                                   16370.75, -48811.41, 68220.87, -46688.0, 12622.25};
        constexpr double dB[10] = {1.11, 2.894, -76.765, 1565.35, -11336.19,
                                   44270.15, -97014.11, 118879.26, -76209.82, 19923.28};
        constexpr double dC[6] = {14.19, -148.24, 847.8, -2318.58, 3168.35, -1590.27};
        // Ls (wu.jl:145-157) with Lsp (wu.jl:131-143)
        T PS = (beta + hh) / (dd - vv);
        PS = PS - 2.0 * beta + 2.0 * ev * dd - hh;
        T QS = exp(-dd) * (1.0 - beta);
        QS = QS * (1.0 + dd / (1.0 + d3));
        T R2 = beta * horner_ref<9>(dA, rho);
        R2 -= (1.0 - alpha) * horner_ref<9>(dB, rho);
        R2 += beta * (1.0 - beta) * rho * (1.0 - 2.0 * rho) * horner_ref<5>(dC, rho);
        const T Ls = 2.0 * PS / (1.0 + d3) - 4.0 * QS + 2.0 * (rho * om3 * R2);
        const T H1 = struve_h1(hh, Y1);
        // -Ls - Wh, Wh = 2 pi (2/pi - H1 + i J1) e^v (wu.jl:122-129)
        dGdh.re = -Ls - twopi * (2.0 / kPi - H1) * ev;
        dGdh.im = -(twopi * J1 * ev);
        dGdv.re = G.re + 2.0 / dd;  // wu.jl:27
        dGdv.im = G.im;
    }
}

// Horizontal distance as the reference computes it, including its
// perturbation at zero (abstract_greens_function.jl:91-99).
template <class T>
__device__ __forceinline__ T horiz_dist(T dx, T dy) {
    const T d2 = dx * dx + dy * dy;
    const double purt = 1e-18;
    return (val(d2) == 0.0) ? sqrt(d2 + purt * purt) - purt : sqrt(d2);
}

// ---------------------------------------------------------------------------
// Exact Guevel-Delhommeau wave term (delhommeau.hpp) in the same interface:
// double values, or Dual<2> with local tangents (d/dr, d/dz) seeded by the
// caller; dG/dr is the exact r-derivative of G and dG/dz = G + 2/hypot(r, z)
// (exact_Guevel_Delhommeau.jl:52-59), so the local second derivatives are
// two more theta-integrals and closed forms.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void delhommeau_dimless(double r, double z, cplx& G, cplx& dGdr,
                                                   cplx& dGdz) {
    double g[2], gr[2], gz[2];
    delh::delhommeau_exact<false>(r, z, g, gr, gz);
    G = {g[0], g[1]};
    dGdr = {gr[0], gr[1]};
    dGdz = {gz[0], gz[1]};
}

__device__ __forceinline__ void delhommeau_dimless(const Dual<2>& rl, const Dual<2>& zl,
                                                   cplxT<Dual<2>>& G, cplxT<Dual<2>>& dGdr,
                                                   cplxT<Dual<2>>& dGdz) {
    // rl, zl carry the local seeds (1, 0) and (0, 1)
    const double r = rl.v, z = zl.v;
    double g[2], gr[2], gz[2], grr[2], grz[2];
    delh::delhommeau_exact<true>(r, z, g, gr, gz, grr, grz);
    const double d = ::hypot(r, z), d3 = d * d * d;
    auto mk = [](double v, double dr, double dz) {
        Dual<2> o;
        o.v = v;
        o.d[0] = dr;
        o.d[1] = dz;
        return o;
    };
    // G_r = Gr, G_z = Gz;  Gr_r = Grr, Gr_z = Grz;  Gz = G + 2/d
    G.re = mk(g[0], gr[0], gz[0]);
    G.im = mk(g[1], gr[1], gz[1]);
    dGdr.re = mk(gr[0], grr[0], grz[0]);
    dGdr.im = mk(gr[1], grr[1], grz[1]);
    dGdz.re = mk(gz[0], gr[0] - 2.0 * r / d3, gz[0] - 2.0 * z / d3);
    dGdz.im = mk(gz[1], gr[1], gz[1]);
}

// ---------------------------------------------------------------------------
// Pair terms shared by the assembly and gradient kernels
// ---------------------------------------------------------------------------
// Rankine + image Rankine: value and gradients w.r.t. the source point (gxi)
// and the collocation point (gx) for one pair.
template <class T, class PT>
__device__ __forceinline__ void rankine_terms(int gf, T x0, T x1, T x2, const PT& p, T& I,
                                              T gxi[3], T gx[3]) {
    I = Cst<T>::make(0.0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        gxi[c] = Cst<T>::make(0.0);
        gx[c] = Cst<T>::make(0.0);
    }
    T v, g[3];
    if (gf & GF_RANKINE) {  // rankine.jl:24-128
        rankine_panel(x0, x1, x2, p, v, g, (gf & kGfFar32) != 0);
        I += v;
        gxi[0] += g[0];
        gxi[1] += g[1];
        gxi[2] += g[2];
        gx[0] -= g[0];
        gx[1] -= g[1];
        gx[2] -= g[2];
    }
    if (gf & (GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG)) {
        // collocation point reflected across z = 0 (rankine_reflected.jl:19-35)
        rankine_panel(x0, x1, -x2, p, v, g, (gf & kGfFar32) != 0);
        const double s = ((gf & GF_RANKINE_REFLECTED) ? 1.0 : 0.0) -
                         ((gf & GF_RANKINE_REFLECTED_NEG) ? 1.0 : 0.0);
        I += s * v;
        gxi[0] += s * g[0];
        gxi[1] += s * g[1];
        gxi[2] += s * g[2];
        // w.r.t. x: vertical reflection of minus the source gradient
        gx[0] -= s * g[0];
        gx[1] -= s * g[1];
        gx[2] += s * g[2];
    }
}

// Wave term of one pair at wavenumber k without the source area: G, dG/dr,
// dG/dz in reduced coordinates and d1, d2 = dr/dxi_1, dr/dxi_2 (zero at
// r <= 1e-6), for horizontal offset (dx, dy) of horizontal length hd.
// WAVE: 1 Wu, 2 exact Delhommeau, 3 tabulated Delhommeau (needs tab).
template <int WAVE>
__device__ __forceinline__ void wave_raw(double k, double dx, double dy, double zsum, double hd, cplx& G,
                                         cplx& Gr, cplx& Gz, double& d1, double& d2, const DelhTab& tab = DelhTab{}) {
    const double r = k * hd;
    const double z = k * zsum;
    if (WAVE == 1) wu_dimless<true>(r, z, G, Gr, Gz);
    if (WAVE == 2) delhommeau_dimless(r, z, G, Gr, Gz);
    if constexpr (WAVE == 3) {
        double g[2], gr[2], gz[2];
        delh_table_eval(tab, r, z, g, gr, gz);
        G = {g[0], g[1]};
        Gr = {gr[0], gr[1]};
        Gz = {gz[0], gz[1]};
    }
    d1 = 0.0;
    d2 = 0.0;
    if (::fabs(r) > 1e-6) {
        const double kk = k * k / r;
        d1 = kk * (-dx);
        d2 = kk * (-dy);
    }
}

// Wave term for one pair at wavenumber k, already multiplied by the source
// area (one-point rule, wu.jl:33-47): value, gradient w.r.t. the source
// point (gxi) and w.r.t. the collocation point (gx).  Reduced coordinates per
// abstract_greens_function.jl:88-155.
template <int WAVE, class T>
__device__ __forceinline__ void wave_term(T k, T dx, T dy, T zsum, T area, cplxT<T>& Sv,
                                          cplxT<T> gxi[3], cplxT<T> gx[3], const DelhTab& tab = DelhTab{}) {
    const T r = k * horiz_dist(dx, dy);
    const T z = k * zsum;
    cplxT<T> G, dGdr, dGdz;
    if (WAVE == 1) wu_dimless<true>(r, z, G, dGdr, dGdz);
    if (WAVE == 2) delhommeau_dimless(r, z, G, dGdr, dGdz);
    if constexpr (WAVE == 3) {  // values only (T = double)
        double g[2], gr[2], gz[2];
        delh_table_eval(tab, r, z, g, gr, gz);
        G = {g[0], g[1]};
        dGdr = {gr[0], gr[1]};
        dGdz = {gz[0], gz[1]};
    }
    T d1 = Cst<T>::make(0.0), d2 = Cst<T>::make(0.0);  // dr/dxi_1, dr/dxi_2
    if (::fabs(val(r)) > 1e-6) {
        const T kk = k * k / r;
        d1 = kk * (-dx);
        d2 = kk * (-dy);
    }
    const T ka = k * area;
    Sv.re = ka * G.re;
    Sv.im = ka * G.im;
    gxi[0].re = ka * (d1 * dGdr.re);
    gxi[0].im = ka * (d1 * dGdr.im);
    gxi[1].re = ka * (d2 * dGdr.re);
    gxi[1].im = ka * (d2 * dGdr.im);
    gxi[2].re = ka * (k * dGdz.re);
    gxi[2].im = ka * (k * dGdz.im);
    gx[0].re = -gxi[0].re;
    gx[0].im = -gxi[0].im;
    gx[1].re = -gxi[1].re;
    gx[1].im = -gxi[1].im;
    gx[2] = gxi[2];
}


// Wave term with P tangent directions.  G depends on the parameters only
// through the reduced coordinates (r, z), so the Wu functions run with two
// local tangents (d/dr, d/dz) and the result is chained to the P parameter
// directions: X_m = X_r dr_m + X_z dz_m.  Same values and derivatives as
// running the generic version in Dual<P>, at a cost independent of P.
template <int P>
__device__ __forceinline__ Dual<P> lift_rz(const Dual<2>& X, const Dual<P>& r, const Dual<P>& z) {
    Dual<P> o;
    o.v = X.v;
#pragma unroll
    for (int m = 0; m < P; ++m) o.d[m] = X.d[0] * r.d[m] + X.d[1] * z.d[m];
    return o;
}

template <int WAVE, int P>
__device__ __forceinline__ void wave_term(Dual<P> k, Dual<P> dx, Dual<P> dy, Dual<P> zsum,
                                          Dual<P> area, cplxT<Dual<P>>& Sv, cplxT<Dual<P>> gxi[3],
                                          cplxT<Dual<P>> gx[3]) {
    typedef Dual<P> T;
    const T r = k * horiz_dist(dx, dy);
    const T z = k * zsum;
    Dual<2> rl, zl;
    rl.v = r.v;
    rl.d[0] = 1.0;
    rl.d[1] = 0.0;
    zl.v = z.v;
    zl.d[0] = 0.0;
    zl.d[1] = 1.0;
    cplxT<Dual<2>> Gl, Grl, Gzl;
    if (WAVE == 1) wu_dimless<true>(rl, zl, Gl, Grl, Gzl);
    if (WAVE == 2) delhommeau_dimless(rl, zl, Gl, Grl, Gzl);
    cplxT<T> G, dGdr, dGdz;
    G.re = lift_rz(Gl.re, r, z);
    G.im = lift_rz(Gl.im, r, z);
    dGdr.re = lift_rz(Grl.re, r, z);
    dGdr.im = lift_rz(Grl.im, r, z);
    dGdz.re = lift_rz(Gzl.re, r, z);
    dGdz.im = lift_rz(Gzl.im, r, z);
    T d1 = Cst<T>::make(0.0), d2 = Cst<T>::make(0.0);
    if (::fabs(val(r)) > 1e-6) {
        const T kk = k * k / r;
        d1 = kk * (-dx);
        d2 = kk * (-dy);
    }
    const T ka = k * area;
    Sv.re = ka * G.re;
    Sv.im = ka * G.im;
    gxi[0].re = ka * (d1 * dGdr.re);
    gxi[0].im = ka * (d1 * dGdr.im);
    gxi[1].re = ka * (d2 * dGdr.re);
    gxi[1].im = ka * (d2 * dGdr.im);
    gxi[2].re = ka * (k * dGdz.re);
    gxi[2].im = ka * (k * dGdz.im);
    gx[0].re = -gxi[0].re;
    gx[0].im = -gxi[0].im;
    gx[1].re = -gxi[1].re;
    gx[1].im = -gxi[1].im;
    gx[2] = gxi[2];
}

}  // namespace mh
