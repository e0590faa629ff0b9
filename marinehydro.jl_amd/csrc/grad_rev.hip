// Reverse mode of the wave-term pair sums of the sweep gradients.
//
// The gradient of output o = (dof d, right-hand side r) of one frequency
// needs, besides O(N) terms, the pair sum (direct formulation; the
// indirect one has the same form with other weight vectors)
//   T_o = -1/(4 pi) sum_pq [ aS_p^d dS_pq sS_q^r - aD_p^d dD_pq sD_q^r ]
// (src/matrix_assembly.jl:4-42 differentiated; the weight vectors are the
// adjoint and forward solutions, as in grad.hip).  Forward mode (grad.hip,
// k_tangent_sym) carries at most 4 parameter directions per evaluation of
// the wave term, so P parameters cost ceil(P / 4) passes over all pairs.
// Here every pair is evaluated once per frequency and block of outputs:
// the wave term depends on the pair only through the local inputs
//   dx = c_p,x - c_q,x, dy = c_p,y - c_q,y, zs = c_p,z + c_q,z, and k,
// linearly on the source area and on one normal.  The Wu (or exact
// Delhommeau) function runs with local (r, z) tangents, lifted to the four
// local inputs (Dual<4>); the pair's contribution to T_o is differentiated
// with respect to them, and the cotangents are summed per panel field
// (centre 3, normal 3, area 1, per output) and per frequency (k):
//   rows p of a tile in LDS, columns q by a reduction across the wave.
// A second kernel contracts the cotangents with the tangents of any number
// of parameters (d c / d theta, d n / d theta, d a / d theta, d k / d theta),
// which is what forward mode computes, in another summation order.
#include "mh_internal.hpp"

namespace mh {

constexpr int kCotF = 7;  // cotangent fields per panel and output: c[3], n[3], area

template <int P>
__device__ __forceinline__ Dual<P> seed_local(double v, int idx) {
    Dual<P> r;
    r.v = v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = (m == idx) ? 1.0 : 0.0;
    return r;
}

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
// a * (re + i im)
__device__ __forceinline__ double2 cmulr(double2 a, double re, double im) {
    return make_double2(a.x * re - a.y * im, a.x * im + a.y * re);
}

// ---------------------------------------------------------------------------
// Pair cotangents, one workgroup per tile pair (I <= J) and frequency, 64
// rows p (lanes) x 64 columns q (4 waves x 16, wave-uniform); (I, J) with
// I < J also covers the transposed entries (q, p) with the same wave-term
// evaluation; diagonal tiles take every entry of the tile as (p, q).
// Outputs o0 .. o0 + on - 1 (on <= OB) of the ndof x nr flattened d * nr + r.
//   prow[((w tiles + J) n + p) OB 7 + ol 7 + f]  row-panel sums of pair (I, J)
//   pcol[((w tiles + I) n + q) OB 7 + ol 7 + f]  column-panel sums
//   pk[(w tiles tiles + J tiles + I) OB + ol]     k sums
// ---------------------------------------------------------------------------
template <int WAVE, int FORM, int OB>
__global__ __launch_bounds__(256) void k_wave_cot_sym(
    int64_t n, const double* __restrict__ C, const PanelRec* __restrict__ panels, const double* __restrict__ ks,
    int64_t nr, int64_t o0, int on, const double2* __restrict__ aS, const double2* __restrict__ aD,
    const double2* __restrict__ sS, const double2* __restrict__ sD, int64_t wstride_a, int64_t wstride_s,
    int64_t tiles, double2* __restrict__ prow, double2* __restrict__ pcol, double2* __restrict__ pk) {
    typedef Dual<4> D4;
    typedef cplxT<D4> C4;
    const int I = blockIdx.x, J = blockIdx.y;
    if (I > J) return;  // covered by the tile pair (J, I)
    const bool sym = I < J;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = blockIdx.z;
    const int64_t i = (int64_t)I * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double x0 = C[ic], x1 = C[ic + n], x2 = C[ic + 2 * n];
    const PanelRec& pp = panels[ic];
    const double np0 = pp.n[0], np1 = pp.n[1], np2 = pp.n[2], ap = pp.area;
    const double kv = ks[w];
    const double2* aSw = aS + w * wstride_a;
    const double2* aDw = aD + w * wstride_a;
    const double2* sSw = sS + w * wstride_s;
    const double2* sDw = sD + w * wstride_s;
    constexpr double c4 = -kInv4Pi;
    // lane-constant weights of row p for each output
    double2 aSp[OB], aDp[OB], sSp[OB], sDp[OB];
    int dof[OB], rhs[OB];
#pragma unroll
    for (int ol = 0; ol < OB; ++ol) {
        const int64_t o = o0 + (ol < on ? ol : 0);
        dof[ol] = (int)(o / nr);
        rhs[ol] = (int)(o % nr);
        const bool ok = row_ok && ol < on;
        aSp[ol] = ok ? aSw[i + dof[ol] * n] : make_double2(0.0, 0.0);
        aDp[ol] = ok ? aDw[i + dof[ol] * n] : make_double2(0.0, 0.0);
        sSp[ol] = ok ? sSw[i + rhs[ol] * n] : make_double2(0.0, 0.0);
        sDp[ol] = ok ? sDw[i + rhs[ol] * n] : make_double2(0.0, 0.0);
    }
    __shared__ double2 sRow[OB * kCotF][256];
    __shared__ double2 red[4][OB];
#pragma unroll
    for (int e = 0; e < OB * kCotF; ++e) sRow[e][threadIdx.x] = make_double2(0.0, 0.0);
    double2 ck[OB];
#pragma unroll
    for (int ol = 0; ol < OB; ++ol) ck[ol] = make_double2(0.0, 0.0);
    auto scl = [](const D4& a, const C4& b) {
        C4 o;
        o.re = a * b.re;
        o.im = a * b.im;
        return o;
    };
    const int64_t jb = (int64_t)J * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& pq = panels[j];
        const double nq0 = pq.n[0], nq1 = pq.n[1], nq2 = pq.n[2], aq = pq.area;
        // local inputs (dx, dy, zs, k)
        const D4 dx = seed_local<4>(x0 - pq.c[0], 0), dy = seed_local<4>(x1 - pq.c[1], 1);
        const D4 zs = seed_local<4>(x2 + pq.c[2], 2), kk = seed_local<4>(kv, 3);
        const D4 rr = kk * horiz_dist(dx, dy);
        const D4 zz = kk * zs;
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
        C4 G, Gr, Gz;
        G.re = lift_rz(Gl.re, rr, zz);
        G.im = lift_rz(Gl.im, rr, zz);
        Gr.re = lift_rz(Grl.re, rr, zz);
        Gr.im = lift_rz(Grl.im, rr, zz);
        Gz.re = lift_rz(Gzl.re, rr, zz);
        Gz.im = lift_rz(Gzl.im, rr, zz);
        D4 d1 = dconst<4>(0.0), d2 = dconst<4>(0.0);
        if (::fabs(rr.v) > 1e-6) {
            const D4 kq = kk * kk / rr;
            d1 = kq * (-dx);
            d2 = kq * (-dy);
        }
        const C4 A1 = scl(d1, Gr), A2 = scl(d2, Gr), A3 = scl(kk, Gz);
        // n . (A1, A2, A3) for the source normal of (p, q) [FORM 1] and the
        // collocation normal of (q, p) [FORM 2]: EQ; (-A1, -A2, A3) with n_p: EP
        C4 EQ, EP;
        EQ.re = nq0 * A1.re + nq1 * A2.re + nq2 * A3.re;
        EQ.im = nq0 * A1.im + nq1 * A2.im + nq2 * A3.im;
        EP.re = np2 * A3.re - np0 * A1.re - np1 * A2.re;
        EP.im = np2 * A3.im - np0 * A1.im - np1 * A2.im;
        const C4 KG = scl(kk, G);
        const C4 KQ = scl(kk, FORM == 1 ? EQ : EP);  // D_pq (K_pq) / a_q
        const C4 KP = scl(kk, FORM == 1 ? EP : EQ);  // D_qp (K_qp) / a_p
        // d KQ / d normal, d KP / d normal (values): k (A1, A2, A3) for n_q,
        // k (-A1, -A2, A3) for n_p
        const double2 vq[3] = {make_double2(kv * A1.re.v, kv * A1.im.v), make_double2(kv * A2.re.v, kv * A2.im.v),
                               make_double2(kv * A3.re.v, kv * A3.im.v)};
        const double2 vp[3] = {make_double2(-vq[0].x, -vq[0].y), make_double2(-vq[1].x, -vq[1].y), vq[2]};
#pragma unroll
        for (int ol = 0; ol < OB; ++ol) {
            if (ol >= on) continue;
            const double2 aSq = aSw[j + dof[ol] * n], aDq = aDw[j + dof[ol] * n];
            const double2 sSq = sSw[j + rhs[ol] * n], sDq = sDw[j + rhs[ol] * n];
            // entry (p, q): u S_pq - v D_pq with S_pq = a_q KG, D_pq = a_q KQ
            const double2 u = cmul2(aSp[ol], sSq), v = cmul2(aDp[ol], sDq);
            double2 al = make_double2(u.x * aq, u.y * aq), be = make_double2(v.x * aq, v.y * aq);
            double2 ga = make_double2(0.0, 0.0), u2 = ga, v2 = ga;
            if (sym) {  // entry (q, p): u' S_qp - v' D_qp with S_qp = a_p KG, D_qp = a_p KP
                u2 = cmul2(aSq, sSp[ol]);
                v2 = cmul2(aDq, sDp[ol]);
                al.x += u2.x * ap;
                al.y += u2.y * ap;
                ga = make_double2(v2.x * ap, v2.y * ap);
            }
            double2 Fd[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double2 t = cmulr(al, KG.re.d[m], KG.im.d[m]);
                const double2 t1 = cmulr(be, KQ.re.d[m], KQ.im.d[m]);
                const double2 t2 = cmulr(ga, KP.re.d[m], KP.im.d[m]);
                Fd[m] = make_double2(c4 * (t.x - t1.x - t2.x), c4 * (t.y - t1.y - t2.y));
            }
            // area cotangents
            const double2 gq = cmulr(u, KG.re.v, KG.im.v), hq = cmulr(v, KQ.re.v, KQ.im.v);
            const double2 daq = make_double2(c4 * (gq.x - hq.x), c4 * (gq.y - hq.y));
            double2 dap = make_double2(0.0, 0.0);
            if (sym) {
                const double2 gp = cmulr(u2, KG.re.v, KG.im.v), hp = cmulr(v2, KP.re.v, KP.im.v);
                dap = make_double2(c4 * (gp.x - hp.x), c4 * (gp.y - hp.y));
            }
            // normal cotangents: -c4 be (d KQ / dn), -c4 ga (d KP / dn)
            const double2 mbe = make_double2(-c4 * be.x, -c4 * be.y), mga = make_double2(-c4 * ga.x, -c4 * ga.y);
            double2 nrow[3], ncol[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (FORM == 1) {  // KQ: source normal n_q; KP: source normal n_p
                    ncol[c] = cmul2(mbe, vq[c]);
                    nrow[c] = cmul2(mga, vp[c]);
                } else {  // KQ: collocation normal n_p; KP: collocation normal n_q
                    nrow[c] = cmul2(mbe, vp[c]);
                    ncol[c] = cmul2(mga, vq[c]);
                }
            }
            // rows: centre p (+d/ddx, +d/ddy, +d/dzs), normal, area
            double2* sr = &sRow[ol * kCotF][0];
            if (row_ok) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    sr[c * 256 + threadIdx.x] = cadd2(sr[c * 256 + threadIdx.x], Fd[c]);
                    sr[(3 + c) * 256 + threadIdx.x] = cadd2(sr[(3 + c) * 256 + threadIdx.x], nrow[c]);
                }
                sr[6 * 256 + threadIdx.x] = cadd2(sr[6 * 256 + threadIdx.x], dap);
                ck[ol] = cadd2(ck[ol], Fd[3]);
            }
            // columns: centre q (-d/ddx, -d/ddy, +d/dzs), normal, area; summed
            // over the lanes (rows), written by lane 0
            double2 cv[kCotF] = {make_double2(-Fd[0].x, -Fd[0].y), make_double2(-Fd[1].x, -Fd[1].y), Fd[2],
                                 ncol[0], ncol[1], ncol[2], daq};
#pragma unroll
            for (int f = 0; f < kCotF; ++f) {
                double vx = row_ok ? cv[f].x : 0.0, vy = row_ok ? cv[f].y : 0.0;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    vx += __shfl_xor(vx, off);
                    vy += __shfl_xor(vy, off);
                }
                cv[f] = make_double2(vx, vy);
            }
            if (lane == 0) {
                double2* pc = pcol + (((int64_t)w * tiles + I) * n + j) * (OB * kCotF) + ol * kCotF;
#pragma unroll
                for (int f = 0; f < kCotF; ++f) pc[f] = cv[f];
            }
        }
    }
    __syncthreads();
    // the four waves share the 64 rows of the tile: wave 0 writes each row's
    // sum over the waves
    if (threadIdx.x < 64 && row_ok) {
        double2* pr = prow + (((int64_t)w * tiles + J) * n + i) * (OB * kCotF);
        for (int ol = 0; ol < on; ++ol)
#pragma unroll
            for (int f = 0; f < kCotF; ++f) {
                const int e = ol * kCotF + f;
                const double2 a = sRow[e][lane], b = sRow[e][64 + lane], c = sRow[e][128 + lane],
                              d = sRow[e][192 + lane];
                pr[e] = make_double2(a.x + b.x + c.x + d.x, a.y + b.y + c.y + d.y);
            }
    }
#pragma unroll
    for (int ol = 0; ol < OB; ++ol) {
        double vx = ck[ol].x, vy = ck[ol].y;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            vx += __shfl_xor(vx, off);
            vy += __shfl_xor(vy, off);
        }
        if (lane == 0) red[wave][ol] = make_double2(vx, vy);
    }
    __syncthreads();
    if (threadIdx.x < on) {
        const int ol = threadIdx.x;
        pk[((int64_t)w * tiles * tiles + (int64_t)J * tiles + I) * OB + ol] =
            make_double2(red[0][ol].x + red[1][ol].x + red[2][ol].x + red[3][ol].x,
                         red[0][ol].y + red[1][ol].y + red[2][ol].y + red[3][ol].y);
    }
}

// cot[((w nout + o0 + ol) n + q) 7 + f] = sum of the row slots J >= tile(q)
// and the column slots I <= tile(q); cotk[w nout + o0 + ol] = sum over the
// tile pairs I <= J
__global__ __launch_bounds__(256) void k_wave_cot_gather(int64_t n, int64_t nw, int64_t tiles, int OB, int on,
                                                        int64_t nout, int64_t o0, const double2* __restrict__ prow,
                                                        const double2* __restrict__ pcol,
                                                        const double2* __restrict__ pk, double2* __restrict__ cot,
                                                        double2* __restrict__ cotk) {
    const int64_t w = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (q, ol, f)
    const int64_t per = (int64_t)on * kCotF;
    if (t < n * per) {
        const int64_t q = t / per, e = t % per;
        const int64_t ol = e / kCotF, f = e % kCotF;
        const int64_t T = q / 64;
        double sx = 0.0, sy = 0.0;
        for (int64_t J = T; J < tiles; ++J) {
            const double2 v = prow[((w * tiles + J) * n + q) * (OB * kCotF) + ol * kCotF + f];
            sx += v.x;
            sy += v.y;
        }
        for (int64_t I = 0; I <= T; ++I) {
            const double2 v = pcol[((w * tiles + I) * n + q) * (OB * kCotF) + ol * kCotF + f];
            sx += v.x;
            sy += v.y;
        }
        cot[((w * nout + o0 + ol) * n + q) * kCotF + f] = make_double2(sx, sy);
    }
    if (blockIdx.x == 0 && threadIdx.x < on) {
        const int ol = threadIdx.x;
        double sx = 0.0, sy = 0.0;
        for (int64_t J = 0; J < tiles; ++J)
            for (int64_t I = 0; I <= J; ++I) {
                const double2 v = pk[(w * tiles * tiles + J * tiles + I) * OB + ol];
                sx += v.x;
                sy += v.y;
            }
        cotk[w * nout + o0 + ol] = make_double2(sx, sy);
    }
}

// out[((w ndof + d) nr + r) Pg + m] += sum_q sum_f cot[w][o][q][f] t_f,q(m0 + m)
// + cotk[w][o] dk[w Pg + m], one workgroup per (w, o, m); tangents from the
// device arrays of the sweep (null = zero)
__global__ __launch_bounds__(256) void k_wave_cot_contract(int64_t n, int64_t ndof, int64_t nr, int Pg,
                                                          int64_t nparam, int64_t m0, const double2* __restrict__ cot,
                                                          const double2* __restrict__ cotk,
                                                          const double* __restrict__ dC,
                                                          const double* __restrict__ dN,
                                                          const double* __restrict__ dA,
                                                          const double* __restrict__ dk, double2* __restrict__ out) {
    const int64_t nout = ndof * nr;
    const int64_t m = blockIdx.x % Pg, o = (blockIdx.x / Pg) % nout, w = blockIdx.x / (Pg * nout);
    const int64_t pm = m0 + m;
    __shared__ double2 red[4];
    double sx = 0.0, sy = 0.0;
    if (pm < nparam) {
        const double2* cw = cot + (w * nout + o) * n * kCotF;
        for (int64_t q = threadIdx.x; q < n; q += 256) {
            const double2* cq = cw + q * kCotF;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double tc = dC ? dC[pm * n * 3 + q + c * n] : 0.0;
                const double tn = dN ? dN[pm * n * 3 + q + c * n] : 0.0;
                sx += cq[c].x * tc + cq[3 + c].x * tn;
                sy += cq[c].y * tc + cq[3 + c].y * tn;
            }
            const double ta = dA ? dA[pm * n + q] : 0.0;
            sx += cq[6].x * ta;
            sy += cq[6].y * ta;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = make_double2(sx, sy);
    __syncthreads();
    if (threadIdx.x == 0) {
        double tx = red[0].x + red[1].x + red[2].x + red[3].x, ty = red[0].y + red[1].y + red[2].y + red[3].y;
        if (pm < nparam) {
            const double2 k = cotk[w * nout + o];
            const double t = dk[w * Pg + m];
            tx += k.x * t;
            ty += k.y * t;
        }
        const int64_t d = o / nr, r = o % nr;
        double2* dst = out + ((w * ndof + d) * nr + r) * Pg + m;
        *dst = make_double2(dst->x + tx, dst->y + ty);
    }
}

constexpr int kRevOB = 2;

int64_t wave_cot_ws_bytes(int64_t n, int64_t nw, int64_t ndof, int64_t nr) {
    const int64_t tiles = (n + 63) / 64;
    const int64_t part = nw * tiles * n * kRevOB * kCotF;  // double2 entries of prow and of pcol
    const int64_t pk = nw * tiles * tiles * kRevOB;
    const int64_t cot = nw * ndof * nr * n * kCotF;
    const int64_t cotk = nw * ndof * nr;
    return (int64_t)sizeof(double2) * (2 * part + pk + cot + cotk) + 8 * 256;
}

int launch_wave_cot(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, int gf, int form, int64_t nw,
                    const double* d_k, int64_t ndof, int64_t nr, const double2* aS, const double2* aD,
                    const double2* sS, const double2* sD, int64_t wa, int64_t ws, void* work, double2** cot_out,
                    double2** cotk_out) {
    const int64_t n = m->nfaces;
    const int64_t tiles = (n + 63) / 64;
    MH_REQUIRE(nw <= 65535 && tiles <= 65535, "gradients: too many frequencies or panels for one launch");
    const int wave = (gf & GF_WU) ? 1 : ((gf & (GF_DELHOMMEAU | GF_DELH_TABLE)) ? 2 : 0);
    MH_REQUIRE(wave != 0, "reverse gradients: a wave term is required");
    char* b = static_cast<char*>(work);
    auto take = [&](int64_t count) {
        double2* p = reinterpret_cast<double2*>(b);
        b += ((sizeof(double2) * (size_t)count + 255) & ~size_t(255));
        return p;
    };
    double2* prow = take(nw * tiles * n * kRevOB * kCotF);
    double2* pcol = take(nw * tiles * n * kRevOB * kCotF);
    double2* pk = take(nw * tiles * tiles * kRevOB);
    double2* cot = take(nw * ndof * nr * n * kCotF);
    double2* cotk = take(nw * ndof * nr);
    const int64_t nout = ndof * nr;
    const dim3 grid((unsigned)tiles, (unsigned)tiles, (unsigned)nw);
    ProfScope ps(ctx, "wave_cot", 0.0, 0.0);
    for (int64_t o0 = 0; o0 < nout; o0 += kRevOB) {
        const int on = (int)((nout - o0) < kRevOB ? (nout - o0) : kRevOB);
#define WC(WV, FM)                                                                                                 \
    hipLaunchKernelGGL((k_wave_cot_sym<WV, FM, kRevOB>), grid, dim3(256), 0, ctx->stream, n, m->centers, panels,   \
                       d_k, nr, o0, on, aS, aD, sS, sD, wa, ws, tiles, prow, pcol, pk)
        if (wave == 1) {
            if (form == 1) WC(1, 1); else WC(1, 2);
        } else {
            if (form == 1) WC(2, 1); else WC(2, 2);
        }
#undef WC
        MH_HIP_CHECK(hipGetLastError());
        const int64_t per = n * on * kCotF;
        hipLaunchKernelGGL(k_wave_cot_gather, dim3((unsigned)((per + 255) / 256), (unsigned)nw), dim3(256), 0,
                           ctx->stream, n, nw, tiles, kRevOB, on, nout, o0, prow, pcol, pk, cot, cotk);
        MH_HIP_CHECK(hipGetLastError());
    }
    *cot_out = cot;
    *cotk_out = cotk;
    return MH_OK;
}

int launch_wave_cot_contract(Context* ctx, int64_t n, int64_t nw, int64_t ndof, int64_t nr, int Pg, int64_t nparam,
                             int64_t m0, const double2* cot, const double2* cotk, const double* dC, const double* dN,
                             const double* dA, const double* dk, double2* out) {
    const int64_t blocks = nw * ndof * nr * Pg;
    if (blocks == 0) return MH_OK;
    MH_REQUIRE(blocks <= 0x7fffffff, "gradients: too many outputs for one launch");
    ProfScope ps(ctx, "wave_cot_contract", 0.0, 0.0);
    hipLaunchKernelGGL(k_wave_cot_contract, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, n, ndof, nr, Pg,
                       nparam, m0, cot, cotk, dC, dN, dA, dk, out);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

}  // namespace mh
