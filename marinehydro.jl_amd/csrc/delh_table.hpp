// Tabulated Guevel-Delhommeau wave term (an approximation of the exact term
// of exact_Guevel_Delhommeau.jl:15-59, which is what it is tested against).
//
// Imaginary parts are closed forms and are evaluated directly:
//   Im G = 2 pi e^z J0(r),  Im dG/dr = -2 pi e^z J1(r),  Im dG/dz = Im G.
// The real part has a logarithmic singularity at (r, z) -> 0:
//   Re G ~ -2 log(d - z) + const,  d = hypot(r, z),
// which is subtracted analytically; the table holds the remainder
//   T(r, z) = Re G + 2 e^z log(d - z)
// and its r-derivative Tr = Re dG/dr + 2 e^z r / (d (d - z)), both bounded
// and continuous.  Since dG/dz = G + 2/d (exact_Guevel_Delhommeau.jl:56),
//   Tz  = T + 2 (1 - e^z) / d,        Trz  = Tr - 2 (1 - e^z) r / d^3,
//   Trrz = Trr - 2 (1 - e^z) (z^2 - 2 r^2) / d^5,
// so the exact values of (T, Tr, Trr) at a node give the bicubic Hermite data
// (value, both first derivatives, cross derivative) of T and of Tr.
//
// Grids (r >= 0, |z| = -z >= 0, m = max(r, |z|)):
//   level 0: r, |z| in [0, kTabR0] with spacing kTabH0, used for m > kTabR1;
//   levels l = 1 .. kTabLevels: [0, R_l]^2 with R_l = kTabR1 2^(1-l) and
//   kTabNn nodes per side, used for R_l / 2 < m <= R_l.  Every level sees a
//   point at least (kTabNn - 1) / 2 cells from the origin, where T's
//   derivatives grow like 1/d^n, so the interpolation error is the same
//   fraction of the local scale on every level.  Points beyond level 0 use
//   the exact term; points below the finest level take its nearest cells,
//   where T differs from its limit at the origin by O(m).
// Node record: T, Tr, Tz, Trz, Trr, Trrz (6 doubles).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "delhommeau.hpp"
#include "mh_dual.hpp"

namespace mh {

constexpr double kTabR0 = 32.0;   // level 0 extent in r and |z|
constexpr double kTabH0 = 0.02;   // level 0 spacing
constexpr int kTabN0 = 1601;      // level 0 nodes per side
constexpr double kTabR1 = 2.0;    // extent of level 1
constexpr int kTabNn = 129;       // nodes per side of levels 1 ..
constexpr int kTabLevels = 26;    // R_26 = 2^-24 kTabR1
constexpr int kTabRec = 6;
// Re G + 2 log(d - z) at the origin: 2 log 2 - 2 gamma (from e^z E1(z) =
// -gamma - log z + O(z) in the theta-integral), used only for nodes at d = 0
constexpr double kTabT0 = 0.23186312590795463;

struct DelhTab {
    const double* lv0;  // kTabN0 x kTabN0 records, r fastest
    const double* lvs;  // kTabLevels levels of kTabNn x kTabNn records
};

__host__ __device__ inline int64_t tab_records() {
    return (int64_t)kTabN0 * kTabN0 + (int64_t)kTabLevels * kTabNn * kTabNn;
}

// Node record from the exact term (for the table builder).
__host__ __device__ inline void tab_node(double r, double z, double* rec) {
    const double d = hypot(r, z);
    if (d == 0.0) {
        rec[0] = kTabT0;
        for (int q = 1; q < kTabRec; ++q) rec[q] = 0.0;
        return;
    }
    double G[2], Gr[2], Gz[2], Grr[2], Grz[2];
    delh::delhommeau_exact<true>(r, z, G, Gr, Gz, Grr, Grz);
    const double ez = exp(z), dz = d - z;
    const double lg = log(dz);
    const double dl = r / (d * dz);                               // d log(d - z) / dr
    const double dll = (d * (z * z - r * r) - z * z * z) / (d * d * d * dz * dz);  // d2 / dr2
    const double T = G[0] + 2.0 * ez * lg;
    const double Tr = Gr[0] + 2.0 * ez * dl;
    const double Trr = Grr[0] + 2.0 * ez * dll;
    const double d3 = d * d * d, d5 = d3 * d * d;
    rec[0] = T;
    rec[1] = Tr;
    rec[2] = T + 2.0 * (1.0 - ez) / d;
    rec[3] = Tr - 2.0 * (1.0 - ez) * r / d3;
    rec[4] = Trr;
    rec[5] = Trr - 2.0 * (1.0 - ez) * (z * z - 2.0 * r * r) / d5;
}

// Cubic Hermite basis on [0, 1]: values at 0 and 1, unit derivatives at 0 and 1.
__device__ __forceinline__ void hermite(double t, double& a0, double& a1, double& b0, double& b1) {
    const double t2 = t * t, t3 = t2 * t;
    a0 = 2.0 * t3 - 3.0 * t2 + 1.0;
    a1 = -2.0 * t3 + 3.0 * t2;
    b0 = t3 - 2.0 * t2 + t;
    b1 = t3 - t2;
}

// G, dG/dr, dG/dz (re, im) of the wave term at (r, z), z <= 0, from the table.
__device__ inline void delh_table_eval(const DelhTab& tab, double r, double z, double G[2], double Gr[2],
                                       double Gz[2]) {
    const double az = -z;
    const double m = fmax(r, az);
    const double* base;
    double h;
    int nn;
    if (m > kTabR1) {
        if (r >= kTabR0 || az >= kTabR0) {  // beyond the table: the exact term
            delh::delhommeau_exact<false>(r, z, G, Gr, Gz);
            return;
        }
        base = tab.lv0;
        h = kTabH0;
        nn = kTabN0;
    } else {
        int l = (m > 0.0) ? 1 + (int)floor(log2(kTabR1 / m)) : kTabLevels;
        if (l < 1) l = 1;
        if (l > kTabLevels) l = kTabLevels;
        h = ldexp(kTabR1, 1 - l) / (double)(kTabNn - 1);
        base = tab.lvs + (size_t)(l - 1) * kTabNn * kTabNn * kTabRec;
        nn = kTabNn;
    }
    double x = r / h, y = az / h;
    int i = (int)x, j = (int)y;
    if (i > nn - 2) i = nn - 2;
    if (j > nn - 2) j = nn - 2;
    x -= (double)i;
    y -= (double)j;
    double ax0, ax1, bx0, bx1, ay0, ay1, by0, by1;
    hermite(x, ax0, ax1, bx0, bx1);
    hermite(y, ay0, ay1, by0, by1);
    // unit-cell derivatives: d/dx = h d/dr, d/dy = -h d/dz (y runs along |z|)
    const double wv[4] = {ax0 * ay0, ax1 * ay0, ax0 * ay1, ax1 * ay1};
    const double wx[4] = {bx0 * ay0, bx1 * ay0, bx0 * ay1, bx1 * ay1};
    const double wy[4] = {ax0 * by0, ax1 * by0, ax0 * by1, ax1 * by1};
    const double wxy[4] = {bx0 * by0, bx1 * by0, bx0 * by1, bx1 * by1};
    const size_t o[4] = {((size_t)j * nn + i) * kTabRec, ((size_t)j * nn + i + 1) * kTabRec,
                         ((size_t)(j + 1) * nn + i) * kTabRec, ((size_t)(j + 1) * nn + i + 1) * kTabRec};
    double T = 0.0, Tr = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double* q = base + o[c];
        T += wv[c] * q[0] + h * (wx[c] * q[1] - wy[c] * q[2]) - h * h * wxy[c] * q[3];
        Tr += wv[c] * q[1] + h * (wx[c] * q[4] - wy[c] * q[3]) - h * h * wxy[c] * q[5];
    }
    const double d = hypot(r, z);
    const double ez = exp(z), dz = d - z;
    const double reG = T - 2.0 * ez * log(dz);
    const double reGr = Tr - 2.0 * ez * r / (d * dz);
    double J0, J1;
    bessel_j01(r, J0, J1);
    const double tpe = 6.28318530717958647692 * ez;
    G[0] = reG;
    G[1] = tpe * J0;
    Gr[0] = reGr;
    Gr[1] = -tpe * J1;
    Gz[0] = reG + 2.0 / d;
    Gz[1] = tpe * J0;
}

}  // namespace mh
