// Gradients of A, B and F with respect to P shape/frequency parameters.
//
// For one frequency, direct formulation (src/matrix_assembly.jl:125-126,
// src/waves.jl:21-37): D phi^r = S b^r and F^{i,r} = c^i . phi^r with
// c^i_p = i rho omega (-n_p . dof^i_p) a_p.  With the adjoint vector
// lambda^i = D^{-T} c^i (one transposed solve with the LU factors),
//   dF^{i,r} = (dc^i) . phi^r + (S^T lambda^i) . db^r
//            + sum_pq lambda^i_p (dS_pq b^r_q - dD_pq phi^r_q).
// Indirect (src/matrix_assembly.jl:127-131): K sigma^r = b^r, phi = S sigma,
// lambda^i = K^{-T} S^T c^i and
//   dF^{i,r} = (dc^i) . phi^r + lambda^i . db^r
//            + sum_pq (c^i_p dS_pq - lambda^i_p dK_pq) sigma^r_q.
// The derivatives are split by Green-function term.  The Rankine and image
// parts of dS and dD/dK do not depend on the frequency: one pass over all
// panel pairs (the FP64 pair code of mh_device.hpp run in Dual<P>) stores
// them as real n x n matrices per parameter, and two GEMMs with all
// frequencies' right vectors at once, then dot products with the left
// vectors, contract them.  The wave-term parts are recomputed per frequency
// and contracted on the fly with the adjoint weights, never stored.  The
// O(N) boundary-condition and force terms are handled by a third kernel.  The derivative is that of the
// code as written (the Wu fits, the branch thresholds), i.e. what automatic
// differentiation of the reference gives and what central finite
// differences of the forward path converge to.
#include <vector>

#include "mh_cx.hpp"
#include "mh_internal.hpp"

namespace mh {

constexpr int kTanFields = 19;  // dc[3], dn[3], da, dV[4][3]

// Tangent record of panel q for parameters [m0, m0+P): tan[(q*19 + f)*P + m].
template <int P>
__global__ void k_pack_tangents(int64_t nf, int64_t nv, int64_t nparam, int64_t m0,
                                const int64_t* __restrict__ F, const double* __restrict__ dV,
                                const double* __restrict__ dC, const double* __restrict__ dN,
                                const double* __restrict__ dA, double* __restrict__ tan) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nf) return;
    double* t = tan + q * kTanFields * P;
    for (int m = 0; m < P; ++m) {
        const int64_t pm = m0 + m;
        const bool ok = pm < nparam;
        for (int c = 0; c < 3; ++c) {
            t[c * P + m] = ok && dC ? dC[pm * nf * 3 + q + c * nf] : 0.0;
            t[(3 + c) * P + m] = ok && dN ? dN[pm * nf * 3 + q + c * nf] : 0.0;
        }
        t[6 * P + m] = ok && dA ? dA[pm * nf + q] : 0.0;
        for (int v = 0; v < 4; ++v) {
            int64_t iv = F[q + v * nf];
            iv = (iv < 0 || iv >= nv) ? 0 : iv;
            for (int c = 0; c < 3; ++c)
                t[(7 + 3 * v + c) * P + m] = ok && dV ? dV[pm * nv * 3 + iv + c * nv] : 0.0;
        }
    }
}

template <int P>
__device__ __forceinline__ Dual<P> tangent_of(double v, const double* t, int field) {
    Dual<P> r;
    r.v = v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = t[field * P + m];
    return r;
}

__device__ __forceinline__ double2 cmac(double2 acc, double ar, double ai, double2 b) {
    // acc + (ar + i ai) * b
    return make_double2(acc.x + ar * b.x - ai * b.y, acc.y + ar * b.y + ai * b.x);
}

// ---------------------------------------------------------------------------
// pair tangent-contraction kernel
//   part[w][blk][i][r][m] = sum_{p in tile rows, q in tile cols}
//       aS^i_p dS_pq,m sS^r_q - aD^i_p dDK_pq,m sD^r_q
// FORM 1: DK = D (source normal, gradient w.r.t. the source point);
// FORM 2: DK = K (collocation normal, gradient w.r.t. the collocation point).
// One launch covers the right-hand sides r0 .. r0 + rn - 1 (rn <= NR) of the
// nr in the part layout, and every i.
// ---------------------------------------------------------------------------
template <int WAVE, int FORM, int P, int NR>
__global__ __launch_bounds__(256) void k_tangent_pairs(
    int64_t n, const double* __restrict__ C, const double* __restrict__ N,
    const PanelRec* __restrict__ panels, const double* __restrict__ tan, int gf,
    const double* __restrict__ ks, const double* __restrict__ dks, int64_t ndof, int64_t nr, int64_t r0,
    int64_t rn, const double2* __restrict__ aS, const double2* __restrict__ aD, const double2* __restrict__ sS,
    const double2* __restrict__ sD, int64_t wstride_a, int64_t wstride_s,
    double2* __restrict__ part) {
    typedef Dual<P> D_;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = blockIdx.z;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double* ti = tan + ic * kTanFields * P;
    const D_ x0 = tangent_of<P>(C[ic], ti, 0);
    const D_ x1 = tangent_of<P>(C[ic + n], ti, 1);
    const D_ x2 = tangent_of<P>(C[ic + 2 * n], ti, 2);
    D_ ni[3];
    if (FORM == 2) {
        ni[0] = tangent_of<P>(N[ic], ti, 3);
        ni[1] = tangent_of<P>(N[ic + n], ti, 4);
        ni[2] = tangent_of<P>(N[ic + 2 * n], ti, 5);
    }
    D_ k;
    k.v = ks[w];
#pragma unroll
    for (int m = 0; m < P; ++m) k.d[m] = dks[w * P + m];
    const double2* sSw = sS + w * wstride_s + r0 * n;
    const double2* sDw = sD + w * wstride_s + r0 * n;
    double2 RS[NR][P], RD[NR][P];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int m = 0; m < P; ++m) {
            RS[r][m] = make_double2(0.0, 0.0);
            RD[r][m] = make_double2(0.0, 0.0);
        }
    const int64_t jb = (int64_t)blockIdx.y * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& pr = panels[j];
        const double* tj = tan + j * kTanFields * P;
        PanelT<D_> p;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            p.c[c] = tangent_of<P>(pr.c[c], tj, c);
            p.n[c] = tangent_of<P>(pr.n[c], tj, 3 + c);
        }
        p.area = tangent_of<P>(pr.area, tj, 6);
        p.radius = pr.radius;
        D_ IR, gxiR[3], gxR[3];
        if (gf & (GF_RANKINE | GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG)) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int c = 0; c < 3; ++c) p.v[v][c] = tangent_of<P>(pr.v[v][c], tj, 7 + 3 * v + c);
            rankine_terms(gf, x0, x1, x2, p, IR, gxiR, gxR);
        } else {
            IR = dconst<P>(0.0);
#pragma unroll
            for (int c = 0; c < 3; ++c) gxiR[c] = gxR[c] = dconst<P>(0.0);
        }
        cplxT<D_> Sv, gxi[3], gx[3];
        if (WAVE) {
            wave_term<WAVE>(k, x0 - p.c[0], x1 - p.c[1], x2 + p.c[2], p.area, Sv, gxi, gx);
        } else {
            Sv.re = Sv.im = dconst<P>(0.0);
#pragma unroll
            for (int c = 0; c < 3; ++c) gxi[c].re = gxi[c].im = gx[c].re = gx[c].im = dconst<P>(0.0);
        }
        // S = -1/(4 pi) (IR + Sv);  DK = delta/2 - 1/(4 pi) n . g
        const D_ s_re = IR + Sv.re;
        const D_& s_im = Sv.im;
        D_ k_re, k_im;
        if (FORM == 1) {
            k_re = p.n[0] * (gxiR[0] + gxi[0].re) + p.n[1] * (gxiR[1] + gxi[1].re) +
                   p.n[2] * (gxiR[2] + gxi[2].re);
            k_im = p.n[0] * gxi[0].im + p.n[1] * gxi[1].im + p.n[2] * gxi[2].im;
        } else {
            k_re = ni[0] * (gxR[0] + gx[0].re) + ni[1] * (gxR[1] + gx[1].re) +
                   ni[2] * (gxR[2] + gx[2].re);
            k_im = ni[0] * gx[0].im + ni[1] * gx[1].im + ni[2] * gx[2].im;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r < rn) {
                const double2 ss = sSw[j + r * n];
                const double2 sd = sDw[j + r * n];
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    RS[r][m] = cmac(RS[r][m], -kInv4Pi * s_re.d[m], -kInv4Pi * s_im.d[m], ss);
                    RD[r][m] = cmac(RD[r][m], -kInv4Pi * k_re.d[m], -kInv4Pi * k_im.d[m], sd);
                }
            }
        }
    }
    // contract with the left weights and reduce over the workgroup
    __shared__ double2 red[4][NR * P];
    const int64_t nblk = (int64_t)gridDim.x * gridDim.y;
    const int64_t blk = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    const double2* aSw = aS + w * wstride_a;
    const double2* aDw = aD + w * wstride_a;
    for (int64_t d = 0; d < ndof; ++d) {
        const double2 as = row_ok ? aSw[i + d * n] : make_double2(0.0, 0.0);
        const double2 ad = row_ok ? aDw[i + d * n] : make_double2(0.0, 0.0);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r >= rn) continue;
#pragma unroll
            for (int m = 0; m < P; ++m) {
                double vx = as.x * RS[r][m].x - as.y * RS[r][m].y - (ad.x * RD[r][m].x - ad.y * RD[r][m].y);
                double vy = as.x * RS[r][m].y + as.y * RS[r][m].x - (ad.x * RD[r][m].y + ad.y * RD[r][m].x);
                for (int off = 32; off > 0; off >>= 1) {
                    vx += __shfl_xor(vx, off);
                    vy += __shfl_xor(vy, off);
                }
                if (lane == 0) red[wave][r * P + m] = make_double2(vx, vy);
                __syncthreads();
                if (threadIdx.x == 0) {
                    const double2 a = red[0][r * P + m], b = red[1][r * P + m],
                                  c2 = red[2][r * P + m], e = red[3][r * P + m];
                    part[(((w * nblk + blk) * ndof + d) * nr + r0 + r) * P + m] =
                        make_double2(a.x + b.x + c2.x + e.x, a.y + b.y + c2.y + e.y);
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------------------
// wave-term-only tangent contraction, evaluated once per unordered pair.
// The wave term depends on a pair only through r = |x_p - x_q|_h and
// z = z_p + z_q (collocation points are the panel centres), both symmetric,
// so the expensive part -- G, dG/dr, dG/dz with their tangents -- is shared by
// (p, q) and (q, p).  With A = (d1 dG/dr, d2 dG/dr, k dG/dz) for the pair
// (p, q) (d1, d2 = dr/dxi), the source gradient of (q, p) is (-A1, -A2, A3),
// and with Eq = n_q . A, Ep = (-n_p1, -n_p2, n_p3) . A:
//   S_pq = k a_q G, S_qp = k a_p G
//   FORM 1: D_pq = k a_q Eq, D_qp = k a_p Ep;  FORM 2: K_pq = k a_q Ep, K_qp = k a_p Eq
// (before the -1/(4 pi) factor).  Tile pairs P < Q evaluate both orders,
// diagonal tiles only (p, q), tiles P > Q are covered by their partner.
// Contraction: (p, q) terms accumulate dS_pq sS_q per row p (weighted by aS_p
// at the end, as in k_tangent_pairs); (q, p) terms accumulate aS_q dS_qp per
// column p (weighted by sS_p at the end).  One launch covers the dofs d0 ..
// d0 + dn - 1 and the right-hand sides r0 .. r0 + rn - 1 (dn, rn <= NR) of
// the ndof x nr outputs.
// ---------------------------------------------------------------------------
template <int WAVE, int FORM, int P, int NR>
__global__ __launch_bounds__(256) void k_tangent_sym(
    int64_t n, const double* __restrict__ C, const PanelRec* __restrict__ panels, const double* __restrict__ tan,
    const double* __restrict__ ks, const double* __restrict__ dks, int64_t ndof, int64_t nr, int64_t d0,
    int64_t dn, int64_t r0, int64_t rn, const double2* __restrict__ aS, const double2* __restrict__ aD,
    const double2* __restrict__ sS, const double2* __restrict__ sD, int64_t wstride_a, int64_t wstride_s,
    double2* __restrict__ part) {
    typedef Dual<P> D_;
    typedef cplxT<D_> CD;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = blockIdx.z;
    const int64_t nblk = (int64_t)gridDim.x * gridDim.y;
    const int64_t blk = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    const int64_t nout = ndof * nr * P;
    double2* pw = part + (w * nblk + blk) * nout;
    if (blockIdx.x > blockIdx.y) {  // covered by the tile pair (y, x)
        for (int64_t o = threadIdx.x; o < dn * rn * P; o += 256) {
            const int64_t d = o / (rn * P), r = (o / P) % rn, m = o % P;
            pw[((d0 + d) * nr + r0 + r) * P + m] = make_double2(0.0, 0.0);
        }
        return;
    }
    const bool sym = blockIdx.x < blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double* ti = tan + ic * kTanFields * P;
    const D_ x0 = tangent_of<P>(C[ic], ti, 0);
    const D_ x1 = tangent_of<P>(C[ic + n], ti, 1);
    const D_ x2 = tangent_of<P>(C[ic + 2 * n], ti, 2);
    const PanelRec& pi_ = panels[ic];
    const D_ np0 = tangent_of<P>(pi_.n[0], ti, 3), np1 = tangent_of<P>(pi_.n[1], ti, 4),
             np2 = tangent_of<P>(pi_.n[2], ti, 5);
    const D_ ap = tangent_of<P>(pi_.area, ti, 6);
    D_ k;
    k.v = ks[w];
#pragma unroll
    for (int m = 0; m < P; ++m) k.d[m] = dks[w * P + m];
    const double2* sSw = sS + w * wstride_s + r0 * n;
    const double2* sDw = sD + w * wstride_s + r0 * n;
    const double2* aSw = aS + w * wstride_a + d0 * n;
    const double2* aDw = aD + w * wstride_a + d0 * n;
    // (q, p) accumulators in LDS (one slot per thread) when small enough:
    // keeps the pair loop within two waves per SIMD
    constexpr bool kLdsU = NR * P <= 8;
    __shared__ double2 sU[kLdsU ? 2 * NR * P : 1][256];
    double2 RS[NR][P], RD[NR][P], US[kLdsU ? 1 : NR][P], UD[kLdsU ? 1 : NR][P];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int m = 0; m < P; ++m) RS[r][m] = RD[r][m] = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < (kLdsU ? 1 : NR); ++r)
#pragma unroll
        for (int m = 0; m < P; ++m) US[r][m] = UD[r][m] = make_double2(0.0, 0.0);
    if (kLdsU)
        for (int e = 0; e < 2 * NR * P; ++e) sU[e][threadIdx.x] = make_double2(0.0, 0.0);
    auto cmul = [](const CD& a, const CD& b) {
        CD o;
        o.re = a.re * b.re - a.im * b.im;
        o.im = a.re * b.im + a.im * b.re;
        return o;
    };
    auto scl = [](const D_& a, const CD& b) {
        CD o;
        o.re = a * b.re;
        o.im = a * b.im;
        return o;
    };
    const int64_t jb = (int64_t)blockIdx.y * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& pr = panels[j];
        const double* tj = tan + j * kTanFields * P;
        const D_ cq0 = tangent_of<P>(pr.c[0], tj, 0), cq1 = tangent_of<P>(pr.c[1], tj, 1),
                 cq2 = tangent_of<P>(pr.c[2], tj, 2);
        const D_ nq0 = tangent_of<P>(pr.n[0], tj, 3), nq1 = tangent_of<P>(pr.n[1], tj, 4),
                 nq2 = tangent_of<P>(pr.n[2], tj, 5);
        const D_ aq = tangent_of<P>(pr.area, tj, 6);
        const D_ dx = x0 - cq0, dy = x1 - cq1, zs = x2 + cq2;
        // G, dG/dr, dG/dz with local tangents (d/dr, d/dz), chained to P
        const D_ rr = k * horiz_dist(dx, dy);
        const D_ zz = k * zs;
        Dual<2> rl, zl;
        rl.v = rr.v;
        rl.d[0] = 1.0;
        rl.d[1] = 0.0;
        zl.v = zz.v;
        zl.d[0] = 0.0;
        zl.d[1] = 1.0;
        cplxT<Dual<2>> Gl, Grl, Gzl;
        if (WAVE == 1) wu_dimless<true>(rl, zl, Gl, Grl, Gzl);
        if (WAVE == 2) delhommeau_dimless(rl, zl, Gl, Grl, Gzl);
        CD G, Gr, Gz;
        G.re = lift_rz(Gl.re, rr, zz);
        G.im = lift_rz(Gl.im, rr, zz);
        Gr.re = lift_rz(Grl.re, rr, zz);
        Gr.im = lift_rz(Grl.im, rr, zz);
        Gz.re = lift_rz(Gzl.re, rr, zz);
        Gz.im = lift_rz(Gzl.im, rr, zz);
        D_ d1 = dconst<P>(0.0), d2 = dconst<P>(0.0);
        if (::fabs(val(rr)) > 1e-6) {
            const D_ kk = k * k / rr;
            d1 = kk * (-dx);
            d2 = kk * (-dy);
        }
        const CD A1 = scl(d1, Gr), A2 = scl(d2, Gr), A3 = scl(k, Gz);
        CD Eq, Ep;
        Eq.re = nq0 * A1.re + nq1 * A2.re + nq2 * A3.re;
        Eq.im = nq0 * A1.im + nq1 * A2.im + nq2 * A3.im;
        Ep.re = np2 * A3.re - np0 * A1.re - np1 * A2.re;
        Ep.im = np2 * A3.im - np0 * A1.im - np1 * A2.im;
        const D_ kaq = k * aq;
        const CD Spq = scl(kaq, G);
        const CD Dpq = scl(kaq, FORM == 1 ? Eq : Ep);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r < rn) {
                const double2 ss = sSw[j + r * n];
                const double2 sd = sDw[j + r * n];
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    RS[r][m] = cmac(RS[r][m], -kInv4Pi * Spq.re.d[m], -kInv4Pi * Spq.im.d[m], ss);
                    RD[r][m] = cmac(RD[r][m], -kInv4Pi * Dpq.re.d[m], -kInv4Pi * Dpq.im.d[m], sd);
                }
            }
        }
        if (sym) {
            const D_ kap = k * ap;
            const CD Sqp = scl(kap, G);
            const CD Dqp = scl(kap, FORM == 1 ? Ep : Eq);
#pragma unroll
            for (int d = 0; d < NR; ++d) {
                if (d < dn) {
                    const double2 as = aSw[j + d * n];
                    const double2 ad = aDw[j + d * n];
#pragma unroll
                    for (int m = 0; m < P; ++m) {
                        if (kLdsU) {
                            double2& us = sU[d * P + m][threadIdx.x];
                            double2& ud = sU[(NR + d) * P + m][threadIdx.x];
                            us = cmac(us, -kInv4Pi * Sqp.re.d[m], -kInv4Pi * Sqp.im.d[m], as);
                            ud = cmac(ud, -kInv4Pi * Dqp.re.d[m], -kInv4Pi * Dqp.im.d[m], ad);
                        } else {
                            US[d][m] = cmac(US[d][m], -kInv4Pi * Sqp.re.d[m], -kInv4Pi * Sqp.im.d[m], as);
                            UD[d][m] = cmac(UD[d][m], -kInv4Pi * Dqp.re.d[m], -kInv4Pi * Dqp.im.d[m], ad);
                        }
                    }
                }
            }
        }
    }
    (void)cmul;
    // contract and reduce over the workgroup
    __shared__ double2 red[4][NR * P];
    for (int64_t d = 0; d < dn; ++d) {
        const double2 as = row_ok ? aSw[i + d * n] : make_double2(0.0, 0.0);
        const double2 ad = row_ok ? aDw[i + d * n] : make_double2(0.0, 0.0);
        double2 usd[P], udd[P];
#pragma unroll
        for (int m = 0; m < P; ++m) {
            usd[m] = make_double2(0.0, 0.0);
            udd[m] = make_double2(0.0, 0.0);
            if (kLdsU) {
                usd[m] = sU[d * P + m][threadIdx.x];
                udd[m] = sU[(NR + d) * P + m][threadIdx.x];
            } else {
#pragma unroll
                for (int dd = 0; dd < (kLdsU ? 1 : NR); ++dd)
                    if (dd == d) {
                        usd[m] = US[dd][m];
                        udd[m] = UD[dd][m];
                    }
            }
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (r >= rn) continue;
            const double2 ssp = row_ok ? sSw[i + r * n] : make_double2(0.0, 0.0);
            const double2 sdp = row_ok ? sDw[i + r * n] : make_double2(0.0, 0.0);
#pragma unroll
            for (int m = 0; m < P; ++m) {
                double vx = as.x * RS[r][m].x - as.y * RS[r][m].y - (ad.x * RD[r][m].x - ad.y * RD[r][m].y);
                double vy = as.x * RS[r][m].y + as.y * RS[r][m].x - (ad.x * RD[r][m].y + ad.y * RD[r][m].x);
                vx += usd[m].x * ssp.x - usd[m].y * ssp.y - (udd[m].x * sdp.x - udd[m].y * sdp.y);
                vy += usd[m].x * ssp.y + usd[m].y * ssp.x - (udd[m].x * sdp.y + udd[m].y * sdp.x);
                for (int off = 32; off > 0; off >>= 1) {
                    vx += __shfl_xor(vx, off);
                    vy += __shfl_xor(vy, off);
                }
                if (lane == 0) red[wave][r * P + m] = make_double2(vx, vy);
            }
        }
        __syncthreads();
        if (threadIdx.x < rn * P) {
            const int r = threadIdx.x / P, m = threadIdx.x % P;
            const double2 a = red[0][r * P + m], b = red[1][r * P + m], c2 = red[2][r * P + m],
                          e = red[3][r * P + m];
            pw[((d0 + d) * nr + r0 + r) * P + m] = make_double2(a.x + b.x + c2.x + e.x, a.y + b.y + c2.y + e.y);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// frequency-independent Rankine + image tangents, stored:
//   dR[(mat * P + m) * n * n + p + q * n], mat 0: dS_pq,m, mat 1: dDK_pq,m
// (the real parts of S and D/K from the Rankine terms; FORM as above)
// ---------------------------------------------------------------------------
template <int FORM, int P>
__global__ __launch_bounds__(256) void k_rankine_tangents(int64_t n, const double* __restrict__ C,
                                                          const double* __restrict__ N,
                                                          const PanelRec* __restrict__ panels,
                                                          const double* __restrict__ tan, int gf,
                                                          double* __restrict__ dR) {
    typedef Dual<P> D_;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double* ti = tan + ic * kTanFields * P;
    const D_ x0 = tangent_of<P>(C[ic], ti, 0);
    const D_ x1 = tangent_of<P>(C[ic + n], ti, 1);
    const D_ x2 = tangent_of<P>(C[ic + 2 * n], ti, 2);
    D_ ni[3];
    if (FORM == 2) {
        ni[0] = tangent_of<P>(N[ic], ti, 3);
        ni[1] = tangent_of<P>(N[ic + n], ti, 4);
        ni[2] = tangent_of<P>(N[ic + 2 * n], ti, 5);
    }
    const int64_t nn = n * n;
    const int64_t jb = (int64_t)blockIdx.y * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& pr = panels[j];
        const double* tj = tan + j * kTanFields * P;
        PanelT<D_> p;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            p.c[c] = tangent_of<P>(pr.c[c], tj, c);
            p.n[c] = tangent_of<P>(pr.n[c], tj, 3 + c);
        }
        p.area = tangent_of<P>(pr.area, tj, 6);
        p.radius = pr.radius;
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int c = 0; c < 3; ++c) p.v[v][c] = tangent_of<P>(pr.v[v][c], tj, 7 + 3 * v + c);
        D_ IR, gxiR[3], gxR[3];
        rankine_terms(gf, x0, x1, x2, p, IR, gxiR, gxR);
        const D_ kr = (FORM == 1) ? p.n[0] * gxiR[0] + p.n[1] * gxiR[1] + p.n[2] * gxiR[2]
                                  : ni[0] * gxR[0] + ni[1] * gxR[1] + ni[2] * gxR[2];
        if (row_ok) {
#pragma unroll
            for (int m = 0; m < P; ++m) {
                dR[m * nn + i + j * n] = -kInv4Pi * IR.d[m];
                dR[(P + m) * nn + i + j * n] = -kInv4Pi * kr.d[m];
            }
        }
    }
}

// Y[z] (n x ncol complex, column c at Y + z * n * ncol + c * n) =
//   dR[z] (n x n real) * B_z, z = mat * P + m; B_z column c = (w, r) =
//   (c / nr, c % nr) at Bmat + w * wstride + r * n (mat 0: B0, mat 1: B1).
// 64 x 64 tiles, 256 threads, on the FP64 matrix cores.
constexpr int kRcK = 16;
__global__ __launch_bounds__(256) void k_rc_gemm(int64_t n, int64_t ncol, int64_t nr, int P,
                                                 const double* __restrict__ dR, const double2* __restrict__ B0,
                                                 const double2* __restrict__ B1, int64_t wstride,
                                                 double2* __restrict__ Y) {
    // 64 x 64 tile per workgroup on the FP64 matrix cores: wave (wr, wc) owns
    // a 32 x 32 quarter as 2 x 2 v_mfma_f64_16x16x4f64 tiles, real and
    // imaginary parts of the complex B as two real products.  Operands are
    // swapped (the MFMA forms the transposed block) so the 16 lanes of a
    // result row hold 16 consecutive rows of Y (coalesced column stores).
    __shared__ double sA[kRcK][64];   // sA[k][row]
    __shared__ double2 sB[kRcK][64];  // sB[k][col]
    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    const int z = blockIdx.z, mat = z / P;
    const double* A = dR + (int64_t)z * n * n;
    const double2* B = mat ? B1 : B0;
    const int64_t r0 = (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
    typedef Cx<double>::acc acc_t;
    acc_t accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = (acc_t){0.0, 0.0, 0.0, 0.0};
            acci[a][b] = (acc_t){0.0, 0.0, 0.0, 0.0};
        }
    for (int64_t k0 = 0; k0 < n; k0 += kRcK) {
        // all eight loads issued before any is waited for: entries outside
        // the ranges load the first element and are replaced by zero
        double va[4], vbx[4], vby[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = tid + 256 * t;
            const int rr = q & 63, kk = q >> 6;
            const int64_t row = r0 + rr, k = k0 + kk;
            const bool ina = row < n && k < n;
            const double a = A[ina ? row + k * n : 0];
            va[t] = ina ? a : 0.0;
            const int kb = q & 15, cc = q >> 4;
            const int64_t col = c0 + cc, qq = k0 + kb;
            const bool inb = col < ncol && qq < n;
            const double2 b = B[inb ? (col / nr) * wstride + (col % nr) * n + qq : 0];
            vbx[t] = inb ? b.x : 0.0;
            vby[t] = inb ? b.y : 0.0;
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = tid + 256 * t;
            sA[q >> 6][q & 63] = va[t];
            sB[q & 15][q >> 4] = make_double2(vbx[t], vby[t]);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kRcK; kk += 4) {
            double av[2];
            double2 bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = sA[kk + (l >> 4)][32 * wr + 16 * a + (l & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = sB[kk + (l >> 4)][32 * wc + 16 * b + (l & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = Cx<double>::mfma(bv[b].x, av[a], accr[a][b]);
                    acci[a][b] = Cx<double>::mfma(bv[b].y, av[a], acci[a][b]);
                }
        }
        __syncthreads();
    }
    double2* Yz = Y + (int64_t)z * n * ncol;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = r0 + 32 * wr + 16 * a + (l & 15);
                const int64_t col = c0 + 32 * wc + 16 * b + Cx<double>::row(r, l);
                if (row < n && col < ncol) Yz[col * n + row] = make_double2(accr[a][b][r], acci[a][b][r]);
            }
}

// out[((w * ndof + d) * nr + r) * P + m] +=
//   sum_p aS^d_w,p Y[m][p, c] - aD^d_w,p Y[P + m][p, c],  c = w * nr + r
__global__ __launch_bounds__(256) void k_rc_contract(int64_t n, int64_t ndof, int64_t nr, int P, int64_t ncol,
                                                     const double2* __restrict__ aS,
                                                     const double2* __restrict__ aD, int64_t wstride_a,
                                                     const double2* __restrict__ Y, double2* __restrict__ out) {
    const int64_t o = blockIdx.x;  // ((w * ndof + d) * nr + r) * P + m
    const int m = (int)(o % P);
    const int64_t r = (o / P) % nr, d = (o / (P * nr)) % ndof, w = o / (P * nr * ndof);
    const int64_t c = w * nr + r;
    const double2* ys = Y + (int64_t)m * n * ncol + c * n;
    const double2* yd = Y + (int64_t)(P + m) * n * ncol + c * n;
    const double2* as = aS + w * wstride_a + d * n;
    const double2* ad = aD + w * wstride_a + d * n;
    double sx = 0.0, sy = 0.0;
    for (int64_t p = threadIdx.x; p < n; p += 256) {
        const double2 a = as[p], y = ys[p], b = ad[p], u = yd[p];
        sx += a.x * y.x - a.y * y.y - (b.x * u.x - b.y * u.y);
        sy += a.x * y.y + a.y * y.x - (b.x * u.y + b.y * u.x);
    }
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    __shared__ double2 red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = make_double2(sx, sy);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[o].x += red[0].x + red[1].x + red[2].x + red[3].x;
        out[o].y += red[0].y + red[1].y + red[2].y + red[3].y;
    }
}

// out[w][d][r][m] (+)= sum over blocks of part
__global__ __launch_bounds__(256) void k_reduce_partials(int64_t nw, int64_t nblk, int64_t nout,
                                                         const double2* __restrict__ part, double2* __restrict__ out,
                                                         int accumulate) {
    // one workgroup per output (w, o): the threads take the blocks in turn,
    // then the sums of the four waves are added
    const int64_t t = blockIdx.x;
    const int64_t w = t / nout, o = t % nout;
    double sx = 0.0, sy = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) {
        const double2 v = part[(w * nblk + b) * nout + o];
        sx += v.x;
        sy += v.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    __shared__ double2 s_w[4];
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = make_double2(sx, sy);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double rx = (s_w[0].x + s_w[1].x) + (s_w[2].x + s_w[3].x);
        const double ry = (s_w[0].y + s_w[1].y) + (s_w[2].y + s_w[3].y);
        if (accumulate) {
            out[t].x += rx;
            out[t].y += ry;
        } else {
            out[t] = make_double2(rx, ry);
        }
    }
}

// ---------------------------------------------------------------------------
// force weights c^i_p = i rho omega (-n_p . dof^i_p) a_p  (src/waves.jl:21-25, :34)
// ---------------------------------------------------------------------------
__global__ void k_force_weights(int64_t n, const double* __restrict__ N, const double* __restrict__ areas,
                                const double* __restrict__ omega, double rho, int64_t ndof,
                                const double* __restrict__ dofs, double2* __restrict__ cw,
                                int64_t wstride) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = blockIdx.y;
    if (p >= n) return;
    const double om = omega[w];
    for (int64_t d = 0; d < ndof; ++d) {
        const double* f = dofs + d * n * 3;
        const double amp = -(f[p] * N[p] + f[p + n] * N[p + n] + f[p + 2 * n] * N[p + 2 * n]);
        cw[w * wstride + p + d * n] = make_double2(0.0, rho * om * amp * areas[p]);
    }
}

// ---------------------------------------------------------------------------
// O(N) terms: T1[w][d][r][m] = sum_p dc^d_p,m phi^r_p + rhoW^d_p db^r_p,m
// ---------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ void dual_sincos(const Dual<P>& x, Dual<P>& s, Dual<P>& c) {
    double sv, cv;
    ::sincos(x.v, &sv, &cv);
    s = chain(x, sv, cv);
    c = chain(x, cv, -sv);
}

template <int P>
__global__ __launch_bounds__(256) void k_bc_tangent(
    int64_t n, const double* __restrict__ C, const double* __restrict__ N,
    const double* __restrict__ areas, const double* __restrict__ tan,
    const double* __restrict__ omega, const double* __restrict__ domega, double g, double rho,
    int64_t ndof, const double* __restrict__ dofs, const double* __restrict__ ddofs, int64_t nparam,
    int64_t m0, int64_t nbeta, const double* __restrict__ beta, const double2* __restrict__ phi,
    int64_t bstride, const double2* __restrict__ rhoW, int64_t wstride,
    double2* __restrict__ out /* [w][d][r][m] */) {
    typedef Dual<P> D_;
    const int64_t w = blockIdx.x;
    const int64_t nr = ndof + nbeta;
    D_ om;
    om.v = omega[w];
#pragma unroll
    for (int m = 0; m < P; ++m) om.d[m] = domega[w * P + m];  // d omega_w / d theta_(m0+m)
    const D_ k = om * om / g;
    __shared__ double2 red[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t d = 0; d < ndof; ++d)
        for (int64_t r = 0; r < nr; ++r)
            for (int m = 0; m < P; ++m) {
                double ax = 0.0, ay = 0.0;
                for (int64_t p = threadIdx.x; p < n; p += 256) {
                    const double* t = tan + p * kTanFields * P;
                    D_ nn[3], x[3];
                    for (int c = 0; c < 3; ++c) {
                        x[c] = tangent_of<P>(C[p + c * n], t, c);
                        nn[c] = tangent_of<P>(N[p + c * n], t, 3 + c);
                    }
                    const D_ a = tangent_of<P>(areas[p], t, 6);
                    // dof fields (tangents given per parameter, or zero)
                    D_ fd[3];
                    for (int c = 0; c < 3; ++c) {
                        fd[c].v = dofs[d * n * 3 + p + c * n];
                        for (int q = 0; q < P; ++q)
                            fd[c].d[q] = (ddofs && m0 + q < nparam)
                                             ? ddofs[((m0 + q) * ndof + d) * n * 3 + p + c * n]
                                             : 0.0;
                    }
                    // c^d_p = i rho omega (-n . dof^d) a : imaginary part only
                    const D_ cim = (rho * om) * (-(nn[0] * fd[0] + nn[1] * fd[1] + nn[2] * fd[2])) * a;
                    // b^r_p
                    D_ bre, bim;
                    if (r < ndof) {
                        D_ fr[3];
                        for (int c = 0; c < 3; ++c) {
                            fr[c].v = dofs[r * n * 3 + p + c * n];
                            for (int q = 0; q < P; ++q)
                                fr[c].d[q] = (ddofs && m0 + q < nparam)
                                                 ? ddofs[((m0 + q) * ndof + r) * n * 3 + p + c * n]
                                                 : 0.0;
                        }
                        const D_ nd = nn[0] * fr[0] + nn[1] * fr[1] + nn[2] * fr[2];
                        bre = dconst<P>(0.0);
                        bim = -(om * nd);
                    } else {
                        const double bb = beta[r - ndof];
                        const double cb = ::cos(bb), sb = ::sin(bb);
                        const D_ wbar = x[0] * cb + x[1] * sb;
                        const D_ cih = exp(k * x[2]);
                        const D_ amp = (g * k) / om;
                        D_ sn, cs;
                        dual_sincos(k * wbar, sn, cs);
                        const D_ pr = amp * cs, pi = amp * sn;
                        const D_ vre = pr * (cb * cih) * nn[0] + pr * (sb * cih) * nn[1] + pi * cih * nn[2];
                        const D_ vim = pi * (cb * cih) * nn[0] + pi * (sb * cih) * nn[1] - pr * cih * nn[2];
                        bre = -vre;
                        bim = -vim;
                    }
                    const double2 ph = phi[w * bstride + p + r * n];
                    const double2 rw = rhoW[w * wstride + p + d * n];
                    // dc . phi  (dc = i dcim)
                    ax += -cim.d[m] * ph.y;
                    ay += cim.d[m] * ph.x;
                    // rhoW . db
                    ax += rw.x * bre.d[m] - rw.y * bim.d[m];
                    ay += rw.x * bim.d[m] + rw.y * bre.d[m];
                }
                for (int off = 32; off > 0; off >>= 1) {
                    ax += __shfl_xor(ax, off);
                    ay += __shfl_xor(ay, off);
                }
                if (lane == 0) red[wave] = make_double2(ax, ay);
                __syncthreads();
                if (threadIdx.x == 0) {
                    out[((w * ndof + d) * nr + r) * P + m] =
                        make_double2(red[0].x + red[1].x + red[2].x + red[3].x,
                                     red[0].y + red[1].y + red[2].y + red[3].y);
                }
                __syncthreads();
            }
}

// ---------------------------------------------------------------------------
// host driver pieces
// ---------------------------------------------------------------------------
template <int WAVE, int FORM, int P, int NR>
static void launch_tp(dim3 grid, hipStream_t st, int64_t n, const double* C, const double* N,
                      const PanelRec* P_, const double* tan, int gf, const double* ks, const double* dks,
                      int64_t ndof, int64_t nr, int64_t r0, int64_t rn, const double2* aS, const double2* aD,
                      const double2* sS, const double2* sD, int64_t wa, int64_t ws, double2* part) {
    hipLaunchKernelGGL((k_tangent_pairs<WAVE, FORM, P, NR>), grid, dim3(256), 0, st, n, C, N, P_, tan, gf,
                       ks, dks, ndof, nr, r0, rn, aS, aD, sS, sD, wa, ws, part);
}

// column blocks of at most 4: cnt columns in ceil(cnt / 4) blocks of
// near-equal width; block b is [b0, b0 + bn)
static void column_block(int64_t cnt, int64_t b, int64_t* b0, int64_t* bn) {
    const int64_t nb = (cnt + 3) / 4, base = cnt / nb, extra = cnt % nb;
    *b0 = b * base + (b < extra ? b : extra);
    *bn = base + (b < extra ? 1 : 0);
}

int launch_tangent_pairs(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, const double* tan,
                         int P, int gf, int form, int64_t nw, const double* d_k, const double* d_dk,
                         int64_t ndof, int64_t nr, const double2* aS, const double2* aD,
                         const double2* sS, const double2* sD, int64_t wa, int64_t ws, double2* part,
                         int64_t* nblk_out) {
    const int64_t n = m->nfaces;
    const int64_t tiles = (n + 63) / 64;
    *nblk_out = tiles * tiles;
    dim3 grid((unsigned)tiles, (unsigned)tiles, (unsigned)nw);
    MH_REQUIRE(ndof >= 1 && ndof <= nr && nr <= 16, "gradients: 1 <= dofs <= dofs + headings <= 16");
    ProfScope ps(ctx, "tangent_pairs", 0.0, 0.0);
    // the tabulated Delhommeau term is differentiated as the exact one
    const int wave = (gf & GF_WU) ? 1 : ((gf & (GF_DELHOMMEAU | GF_DELH_TABLE)) ? 2 : 0);
    // the pair kernels hold at most 4 right-hand sides (and, wave term only,
    // 4 dofs) in registers: wider sweeps take one launch per column block,
    // each writing its own entries of the partial sums
    const int64_t nrb = (nr + 3) / 4;
    if (wave && !(gf & (GF_RANKINE | GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG))) {
        // wave term only: one evaluation per unordered pair
#define TS(WV, FM, PP, NRR)                                                                                 \
    hipLaunchKernelGGL((k_tangent_sym<WV, FM, PP, NRR>), grid, dim3(256), 0, ctx->stream, n, m->centers, panels, \
                       tan, d_k, d_dk, ndof, nr, d0, dn, r0, rn, aS, aD, sS, sD, wa, ws, part)
#define TS_NR(WV, FM, PP)           \
    do {                            \
        if (nr2)                    \
            TS(WV, FM, PP, 2);      \
        else                        \
            TS(WV, FM, PP, 4);      \
    } while (0)
#define TS_P(WV, FM)                \
    do {                            \
        if (P == 1)                 \
            TS_NR(WV, FM, 1);       \
        else if (P == 2)            \
            TS_NR(WV, FM, 2);       \
        else                        \
            TS_NR(WV, FM, 4);       \
    } while (0)
        const int64_t ndb = (ndof + 3) / 4;
        for (int64_t db = 0; db < ndb; ++db)
            for (int64_t rb = 0; rb < nrb; ++rb) {
                int64_t d0, dn, r0, rn;
                column_block(ndof, db, &d0, &dn);
                column_block(nr, rb, &r0, &rn);
                const bool nr2 = dn <= 2 && rn <= 2;
                if (wave == 1) {
                    if (form == 1) TS_P(1, 1); else TS_P(1, 2);
                } else {
                    if (form == 1) TS_P(2, 1); else TS_P(2, 2);
                }
            }
#undef TS_P
#undef TS_NR
#undef TS
        MH_HIP_CHECK(hipGetLastError());
        return MH_OK;
    }
#define TP(WV, FM, PP, NRR)                                                                        \
    launch_tp<WV, FM, PP, NRR>(grid, ctx->stream, n, m->centers, m->normals, panels, tan, gf, d_k, \
                               d_dk, ndof, nr, r0, rn, aS, aD, sS, sD, wa, ws, part)
#define TP_NR(WV, FM, PP) \
    do {                  \
        if (nr2)          \
            TP(WV, FM, PP, 2); \
        else              \
            TP(WV, FM, PP, 4); \
    } while (0)
#define TP_FORM(WV, PP)        \
    do {                       \
        if (form == 1)         \
            TP_NR(WV, 1, PP);  \
        else                   \
            TP_NR(WV, 2, PP);  \
    } while (0)
#define TP_WAVE(PP)            \
    do {                       \
        if (wave == 1)         \
            TP_FORM(1, PP);    \
        else if (wave == 2)    \
            TP_FORM(2, PP);    \
        else                   \
            TP_FORM(0, PP);    \
    } while (0)
    for (int64_t rb = 0; rb < nrb; ++rb) {
        int64_t r0, rn;
        column_block(nr, rb, &r0, &rn);
        const bool nr2 = rn <= 2;
        if (P == 1)
            TP_WAVE(1);
        else if (P == 2)
            TP_WAVE(2);
        else
            TP_WAVE(4);
    }
#undef TP_WAVE
#undef TP_FORM
#undef TP_NR
#undef TP
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_rankine_tangents(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, const double* tan, int P,
                            int gf, int form, double* dR) {
    const int64_t n = m->nfaces;
    const int64_t tiles = (n + 63) / 64;
    dim3 grid((unsigned)tiles, (unsigned)tiles);
    ProfScope ps(ctx, "rankine_tangents", 0.0, 16.0 * P * (double)n * n);
#define RT(FM, PP)                                                                                        \
    hipLaunchKernelGGL((k_rankine_tangents<FM, PP>), grid, dim3(256), 0, ctx->stream, n, m->centers, m->normals, \
                       panels, tan, gf, dR)
    if (form == 1) {
        if (P == 1) RT(1, 1); else if (P == 2) RT(1, 2); else RT(1, 4);
    } else {
        if (P == 1) RT(2, 1); else if (P == 2) RT(2, 2); else RT(2, 4);
    }
#undef RT
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_rankine_contract(Context* ctx, int64_t n, int P, int64_t nw, int64_t ndof, int64_t nr, const double* dR,
                            const double2* aS, const double2* aD, const double2* sS, const double2* sD, int64_t wa,
                            int64_t ws, double2* Y, double2* out) {
    const int64_t ncol = nw * nr;
    if (n == 0 || ncol == 0 || ndof == 0) return MH_OK;
    {
        ProfScope ps(ctx, "rc_gemm", 2.0 * 2.0 * 2.0 * P * (double)n * n * ncol, 0.0);
        dim3 grid((unsigned)((n + 63) / 64), (unsigned)((ncol + 63) / 64), (unsigned)(2 * P));
        hipLaunchKernelGGL(k_rc_gemm, grid, dim3(256), 0, ctx->stream, n, ncol, nr, P, dR, sS, sD, ws, Y);
        MH_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rc_contract, dim3((unsigned)(nw * ndof * nr * P)), dim3(256), 0, ctx->stream, n, ndof, nr,
                       P, ncol, aS, aD, wa, Y, out);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_pack_tangents(Context* ctx, const mh_mesh_t* m, int P, int64_t nparam, int64_t m0,
                         const double* dV, const double* dC, const double* dN, const double* dA,
                         double* tan) {
    const int64_t nf = m->nfaces;
    dim3 grid((unsigned)((nf + 127) / 128)), blk(128);
    if (P == 1)
        hipLaunchKernelGGL(k_pack_tangents<1>, grid, blk, 0, ctx->stream, nf, m->nvertices, nparam, m0,
                           m->faces, dV, dC, dN, dA, tan);
    else if (P == 2)
        hipLaunchKernelGGL(k_pack_tangents<2>, grid, blk, 0, ctx->stream, nf, m->nvertices, nparam, m0,
                           m->faces, dV, dC, dN, dA, tan);
    else
        hipLaunchKernelGGL(k_pack_tangents<4>, grid, blk, 0, ctx->stream, nf, m->nvertices, nparam, m0,
                           m->faces, dV, dC, dN, dA, tan);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_force_weights(Context* ctx, const mh_mesh_t* m, int64_t nw, const double* d_om, double rho,
                         int64_t ndof, const double* dofs, double2* cw, int64_t wstride) {
    const int64_t n = m->nfaces;
    hipLaunchKernelGGL(k_force_weights, dim3((unsigned)((n + 255) / 256), (unsigned)nw), dim3(256), 0,
                       ctx->stream, n, m->normals, m->areas, d_om, rho, ndof, dofs, cw, wstride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_bc_tangent(Context* ctx, const mh_mesh_t* m, int P, const double* tan, int64_t nw,
                      const double* d_om, const double* d_dom, double g, double rho, int64_t ndof,
                      const double* dofs, const double* ddofs, int64_t nparam, int64_t m0,
                      int64_t nbeta, const double* d_beta, const double2* phi, int64_t bstride,
                      const double2* rhoW, int64_t wstride, double2* out) {
    const int64_t n = m->nfaces;
    ProfScope ps(ctx, "bc_tangent", 0.0, 0.0);
    if (P == 1)
        hipLaunchKernelGGL(k_bc_tangent<1>, dim3((unsigned)nw), dim3(256), 0, ctx->stream, n, m->centers,
                           m->normals, m->areas, tan, d_om, d_dom, g, rho, ndof, dofs, ddofs, nparam, m0,
                           nbeta, d_beta, phi, bstride, rhoW, wstride, out);
    else if (P == 2)
        hipLaunchKernelGGL(k_bc_tangent<2>, dim3((unsigned)nw), dim3(256), 0, ctx->stream, n, m->centers,
                           m->normals, m->areas, tan, d_om, d_dom, g, rho, ndof, dofs, ddofs, nparam, m0,
                           nbeta, d_beta, phi, bstride, rhoW, wstride, out);
    else
        hipLaunchKernelGGL(k_bc_tangent<4>, dim3((unsigned)nw), dim3(256), 0, ctx->stream, n, m->centers,
                           m->normals, m->areas, tan, d_om, d_dom, g, rho, ndof, dofs, ddofs, nparam, m0,
                           nbeta, d_beta, phi, bstride, rhoW, wstride, out);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_reduce_partials(Context* ctx, int64_t nw, int64_t nblk, int64_t nout, const double2* part,
                           double2* out, int accumulate) {
    const int64_t tot = nw * nout;
    if (tot == 0) return MH_OK;
    MH_REQUIRE(tot <= 0x7fffffff, "gradients: too many outputs for one launch");
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)tot), dim3(256), 0, ctx->stream, nw, nblk, nout, part, out,
                       accumulate);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// Gradients of A, B and F from the contracted pair terms T1 + T2 (dF of every
// (dof d, right-hand side r) pair per parameter), the forces A, B and the
// parameter's omega tangent dw: A = Re F / omega^2, B = Im F / omega
// (src/waves.jl:36), so dA = Re dF / omega^2 - 2 A dw / omega and
// dB = Im dF / omega - B dw / omega; right-hand sides r >= ndof are the
// headings, whose dF is written as is.  One thread per (w, m, d, r).
__global__ void k_grad_finish(int64_t nw, int64_t np, int64_t m0, int P, int64_t ndof, int64_t nr,
                              int64_t nbeta, const double* __restrict__ om, const double* __restrict__ dom,
                              const double* __restrict__ A, const double* __restrict__ B,
                              const double2* __restrict__ T1, const double2* __restrict__ T2,
                              double* __restrict__ gA, double* __restrict__ gB, double* __restrict__ gF) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nw * P * ndof * nr) return;
    const int64_t r = t % nr, d = (t / nr) % ndof, m = (t / (nr * ndof)) % P, w = t / (nr * ndof * P);
    if (m0 + m >= np) return;
    const int64_t o = ((w * ndof + d) * nr + r) * P + m;
    const double fr = T1[o].x + T2[o].x, fi = T1[o].y + T2[o].y;
    if (r < ndof) {
        const double w_ = om[w], dw = dom[w * P + m];
        const int64_t ia = (w * ndof + r) * ndof + d;
        const int64_t og = ((w * np + m0 + m) * ndof + r) * ndof + d;
        if (gA) gA[og] = fr / (w_ * w_) - 2.0 * A[ia] * dw / w_;
        if (gB) gB[og] = fi / w_ - B[ia] * dw / w_;
    } else if (gF) {
        const int64_t og = ((w * np + m0 + m) * nbeta + (r - ndof)) * ndof + d;
        gF[2 * og] = fr;
        gF[2 * og + 1] = fi;
    }
}

int launch_grad_finish(Context* ctx, int64_t nw, int64_t np, int64_t m0, int P, int64_t ndof, int64_t nr,
                       int64_t nbeta, const double* d_om, const double* d_dom, const double* A, const double* B,
                       const double2* T1, const double2* T2, double* gA, double* gB, double* gF) {
    const int64_t tot = nw * P * ndof * nr;
    if (tot == 0) return MH_OK;
    hipLaunchKernelGGL(k_grad_finish, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, nw, np, m0,
                       P, ndof, nr, nbeta, d_om, d_dom, A, B, T1, T2, gA, gB, gF);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

}  // namespace mh
