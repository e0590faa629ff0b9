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
// Delhommeau) function runs with local (r, z) tangents (r = k hd(dx, dy),
// z = k zs); per output the pair's contribution to T_o is
//   H = k al G - k^2 b1 Gz - k kh b2 Gr     (kh = k / hd)
// with complex coefficients from the weights, areas and normals (b2 linear
// in dx, dy), and its derivatives w.r.t. the local inputs follow from the
// (r, z) derivatives of G, Gr, Gz by the chain rule plus the explicit
// dependence of k, kh and b2; the cotangents are summed per panel field
// (centre 3, normal 3, area 1, per output) and per frequency (k):
//   rows p of a tile in LDS, columns q by a reduction across the wave.
// A second kernel contracts the cotangents with the tangents of any number
// of parameters (d c / d theta, d n / d theta, d a / d theta, d k / d theta),
// which is what forward mode computes, in another summation order.
#include "mh_internal.hpp"

namespace mh {

constexpr int kCotF = 7;  // cotangent fields per panel and output: c[3], n[3], area

__device__ __forceinline__ double2 cmul2(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd2(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// v of another lane of the same row of 16 by a DPP move of both halves
template <int CTRL>
__device__ __forceinline__ double dpp_row(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
// v of the last lane of the previous row (ROW_BCAST15, rows 1 and 3) or of
// lane 31 (ROW_BCAST31, rows 2 and 3); 0 in the other rows
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_bcast(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
// sum over the 64 lanes, in lane 63: within quads (DPP quad_perm [1,0,3,2]
// and [2,3,0,1]), across the quads of a row (row rotations by 4 and 8: every
// lane of a row then holds the row's sum), then the row sums carried forward
// by the row broadcasts (rows 1, 3 add the row before; rows 2, 3 add lane 31)
__device__ __forceinline__ double wave_sum63(double v) {
    v += dpp_row<0xB1>(v);
    v += dpp_row<0x4E>(v);
    v += dpp_row<0x124>(v);
    v += dpp_row<0x128>(v);
    v += dpp_bcast<0x142, 0xA>(v);
    v += dpp_bcast<0x143, 0xC>(v);
    return v;
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
    const int64_t jb = (int64_t)J * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& pq = panels[j];
        const double nq0 = pq.n[0], nq1 = pq.n[1], nq2 = pq.n[2], aq = pq.area;
        // local inputs: dx = c_p,x - c_q,x, dy, zs = c_p,z + c_q,z, and k
        const double dx = x0 - pq.c[0], dy = x1 - pq.c[1], zs = x2 + pq.c[2];
        const double dsq = dx * dx + dy * dy;
        const double hd = horiz_dist(dx, dy);
        // d hd / d(dx, dy) as the dual numbers give it: dx / sqrt(dsq), with
        // the perturbation of horiz_dist at zero (then dx = dy = 0 and so is it)
        const double sq = (dsq == 0.0) ? ::sqrt(dsq + 1e-36) : hd;
        const double hx = dx / sq, hy = dy / sq;
        const double rr = kv * hd, zz = kv * zs;
        Dual<2> rl, zl;
        rl.v = rr;
        rl.d[0] = 1.0;
        rl.d[1] = 0.0;
        zl.v = zz;
        zl.d[0] = 0.0;
        zl.d[1] = 1.0;
        cplxT<Dual<2>> Gl, Grl, Gzl;
        if (WAVE == 1) wu_dimless<true>(rl, zl, Gl, Grl, Gzl);
        if (WAVE == 2) delhommeau_dimless(rl, zl, Gl, Grl, Gzl);
        // values and (r, z) derivatives of G, dG/dr, dG/dz
        const double2 G0 = make_double2(Gl.re.v, Gl.im.v), Gr0 = make_double2(Grl.re.v, Grl.im.v),
                      Gz0 = make_double2(Gzl.re.v, Gzl.im.v);
        const double2 G_r = make_double2(Gl.re.d[0], Gl.im.d[0]), G_z = make_double2(Gl.re.d[1], Gl.im.d[1]);
        const double2 Gr_r = make_double2(Grl.re.d[0], Grl.im.d[0]), Gr_z = make_double2(Grl.re.d[1], Grl.im.d[1]);
        const double2 Gz_r = make_double2(Gzl.re.d[0], Gzl.im.d[0]), Gz_z = make_double2(Gzl.re.d[1], Gzl.im.d[1]);
        // d1 = -kh dx, d2 = -kh dy (kh = k / hd; zero near r = 0, as the
        // forward code sets them); A = (d1 Gr, d2 Gr, k Gz);
        // EQ(n) = n . A, EP(n) = n2 A3 - n0 A1 - n1 A2
        const bool far = ::fabs(rr) > 1e-6;
        const double kh = far ? kv * kv / rr : 0.0;
        const double2 A1 = make_double2(-kh * dx * Gr0.x, -kh * dx * Gr0.y);
        const double2 A2 = make_double2(-kh * dy * Gr0.x, -kh * dy * Gr0.y);
        const double2 A3 = make_double2(kv * Gz0.x, kv * Gz0.y);
        const double2 EQ = make_double2(nq0 * A1.x + nq1 * A2.x + nq2 * A3.x, nq0 * A1.y + nq1 * A2.y + nq2 * A3.y);
        const double2 EP = make_double2(np2 * A3.x - np0 * A1.x - np1 * A2.x, np2 * A3.y - np0 * A1.y - np1 * A2.y);
        // values: KG = k G, KQ = D_pq / a_q (K_pq / a_q), KP = D_qp / a_p
        const double2 KG = make_double2(kv * G0.x, kv * G0.y);
        const double2 KQ = FORM == 1 ? make_double2(kv * EQ.x, kv * EQ.y) : make_double2(kv * EP.x, kv * EP.y);
        const double2 KP = FORM == 1 ? make_double2(kv * EP.x, kv * EP.y) : make_double2(kv * EQ.x, kv * EQ.y);
        // the normals and signs of the two D (K) entries: K_beta = k E(n_beta)
        // with E(n) = n2 k Gz + s (kh (n0 dx + n1 dy)) Gr, s = -1 for EQ, +1 for EP
        const double nb0 = FORM == 1 ? nq0 : np0, nb1 = FORM == 1 ? nq1 : np1, nb2 = FORM == 1 ? nq2 : np2;
        const double ng0 = FORM == 1 ? np0 : nq0, ng1 = FORM == 1 ? np1 : nq1, ng2 = FORM == 1 ? np2 : nq2;
        const double sb = FORM == 1 ? -1.0 : 1.0, sg = -sb;
        const double rhob = nb0 * dx + nb1 * dy, rhog = ng0 * dx + ng1 * dy;
        // d (k E) / d normal (values): k (A1, A2, A3) with n_q [EQ],
        // k (-A1, -A2, A3) with n_p [EP]
        const double2 vq[3] = {make_double2(kv * A1.x, kv * A1.y), make_double2(kv * A2.x, kv * A2.y),
                               make_double2(kv * A3.x, kv * A3.y)};
        const double2 vp[3] = {make_double2(-vq[0].x, -vq[0].y), make_double2(-vq[1].x, -vq[1].y), vq[2]};
        const double k2 = kv * kv, kkh = kv * kh;
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
            // H = al KG - be KQ - ga KP = k al G - k^2 b1 Gz - k kh b2 Gr
            const double2 b1 = make_double2(be.x * nb2 + ga.x * ng2, be.y * nb2 + ga.y * ng2);
            const double2 b2 = make_double2(sb * rhob * be.x + sg * rhog * ga.x, sb * rhob * be.y + sg * rhog * ga.y);
            const double2 b2x = make_double2(sb * nb0 * be.x + sg * ng0 * ga.x, sb * nb0 * be.y + sg * ng0 * ga.y);
            const double2 b2y = make_double2(sb * nb1 * be.x + sg * ng1 * ga.x, sb * nb1 * be.y + sg * ng1 * ga.y);
            const double2 aGr = cmul2(al, G_r), aGz = cmul2(al, G_z);
            const double2 b1r = cmul2(b1, Gz_r), b1z = cmul2(b1, Gz_z);
            const double2 b2r = cmul2(b2, Gr_r), b2z = cmul2(b2, Gr_z);
            const double2 Hr = make_double2(kv * aGr.x - k2 * b1r.x - kkh * b2r.x, kv * aGr.y - k2 * b1r.y - kkh * b2r.y);
            const double2 Hz = make_double2(kv * aGz.x - k2 * b1z.x - kkh * b2z.x, kv * aGz.y - k2 * b1z.y - kkh * b2z.y);
            // explicit dependence on k (H's factors k, k^2, k kh with kh = k / hd)
            // and on dx, dy (kh and b2)
            const double2 aG = cmul2(al, G0), b1G = cmul2(b1, Gz0), b2G = cmul2(b2, Gr0);
            const double2 Hk = make_double2(aG.x - 2.0 * kv * b1G.x - 2.0 * kh * b2G.x,
                                            aG.y - 2.0 * kv * b1G.y - 2.0 * kh * b2G.y);
            const double ihd = far ? 1.0 / hd : 0.0;
            const double2 tx = make_double2(b2.x * hx * ihd - b2x.x, b2.y * hx * ihd - b2x.y);
            const double2 ty = make_double2(b2.x * hy * ihd - b2y.x, b2.y * hy * ihd - b2y.y);
            const double2 Gtx = cmul2(Gr0, tx), Gty = cmul2(Gr0, ty);
            double2 Fd[4];
            Fd[0] = make_double2(c4 * (Hr.x * kv * hx + kkh * Gtx.x), c4 * (Hr.y * kv * hx + kkh * Gtx.y));
            Fd[1] = make_double2(c4 * (Hr.x * kv * hy + kkh * Gty.x), c4 * (Hr.y * kv * hy + kkh * Gty.y));
            Fd[2] = make_double2(c4 * Hz.x * kv, c4 * Hz.y * kv);
            Fd[3] = make_double2(c4 * (Hr.x * hd + Hz.x * zs + Hk.x), c4 * (Hr.y * hd + Hz.y * zs + Hk.y));
            // area cotangents
            const double2 gq = cmul2(u, KG), hq = cmul2(v, KQ);
            const double2 daq = make_double2(c4 * (gq.x - hq.x), c4 * (gq.y - hq.y));
            double2 dap = make_double2(0.0, 0.0);
            if (sym) {
                const double2 gp = cmul2(u2, KG), hp = cmul2(v2, KP);
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
            // over the lanes (rows), written by lane 63
            double2 cv[kCotF] = {make_double2(-Fd[0].x, -Fd[0].y), make_double2(-Fd[1].x, -Fd[1].y), Fd[2],
                                 ncol[0], ncol[1], ncol[2], daq};
#pragma unroll
            for (int f = 0; f < kCotF; ++f)
                cv[f] = make_double2(wave_sum63(row_ok ? cv[f].x : 0.0), wave_sum63(row_ok ? cv[f].y : 0.0));
            if (lane == 63) {
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
