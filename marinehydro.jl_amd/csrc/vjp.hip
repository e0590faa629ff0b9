// Reverse mode of the assembly and the solve (matrix-level cotangents).
//
// Convention (Zygote / ChainRules for complex arrays): the cotangent of a
// complex entry z is zbar = dL/dRe z + i dL/dIm z, so a real input u
// receives ubar = sum Re(conj(zbar) dz/du).
//
// Assembly, reference src/matrix_assembly.jl:4-42 (S at :14, D/K at :16-26):
//   S_ij = -1/(4 pi) [IR(x_i; panel j) + k a_j G(r, z)]
//   D_ij = delta_ij/2 - 1/(4 pi) n_j . [gxiR + k a_j grad_xi G]     (direct)
//   K_ij = delta_ij/2 - 1/(4 pi) n_i . [gxR + k a_j grad_x G]       (indirect)
// with x_i = centre i (collocation point) and panel j's centre, normal,
// area and four vertices read from the mesh arrays as independent fields
// (the reference differentiates them separately, e.g. the vertex Jacobian of
// test/matrix_assembly_differentiation.jl:34-64 keeps centres and normals).
// Per pair the derivatives w.r.t. the pair's local inputs are formed in
// forward mode with unit seeds (the FP64 pair code run in Dual<P>), weighted
// by the pair's cotangents and summed per row (x_i, n_i) and per column
// (panel j's fields); nothing of size n^2 is stored.
//  - The Rankine and image parts do not depend on k: one kernel, weighted by
//    the real parts of the cotangents summed over the wavenumbers.  Seeds:
//    pass 0 = collocation point (3), panel centre (3), area; passes 1 and 2
//    = panel normal (3) and vertices (12), which enter only the near-field
//    formula (rankine.jl:36-63, :79-115), so a wave skips them when none of
//    its pairs is in the near field.  The normal also enters D (or K)
//    linearly; that part is the gradient's value, added in pass 0.
//  - The wave term (wu.jl, exact_Guevel_Delhommeau.jl) depends on the pair
//    through dx, dy (horizontal offset), z_i + z_j, the area and k: Dual<5>,
//    at the cost of the value (mh_device.hpp's wave_term chains local
//    (r, z) tangents), once per wavenumber.
// Partial sums go to per-tile buffers; a finishing kernel adds them and
// scatters the vertex cotangents of each face to the shared vertices.
#include "mh_internal.hpp"

namespace mh {

// column fields: c[3], area, n_j[3], V[12]; row fields: x[3], n_i[3]
constexpr int kColF = 19;
constexpr int kRowF = 6;

template <int P>
__device__ __forceinline__ Dual<P> seed(double v, int idx, int first) {
    Dual<P> r;
    r.v = v;
#pragma unroll
    for (int m = 0; m < P; ++m) r.d[m] = (idx == first + m) ? 1.0 : 0.0;
    return r;
}

// wave-reduce v and add lane 0's total into *dst (LDS, one wave per column)
__device__ __forceinline__ void col_add(double v, double* dst, int lane) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) *dst += v;
}

// ---------------------------------------------------------------------------
// Rankine + image parts.  Wr, Wd: n x n real weights (the real parts of the
// S and D/K cotangents summed over wavenumbers), column-major, ld n.
// Local inputs: 0-2 x, 3-5 c, 6 area, 7-9 n_j, 10-21 V[v][c].
// ---------------------------------------------------------------------------
template <int FORM, int P, int FIRST>
__device__ __forceinline__ void rankine_vjp_pass(int gf, const double x[3], const PanelRec& pr, const double ni[3],
                                                 double cS, double cD, double out[P]) {
    typedef Dual<P> D_;
    const D_ x0 = seed<P>(x[0], 0, FIRST), x1 = seed<P>(x[1], 1, FIRST), x2 = seed<P>(x[2], 2, FIRST);
    PanelT<D_> p;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        p.c[c] = seed<P>(pr.c[c], 3 + c, FIRST);
        p.n[c] = seed<P>(pr.n[c], 7 + c, FIRST);
    }
    p.area = seed<P>(pr.area, 6, FIRST);
    p.radius = pr.radius;
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int c = 0; c < 3; ++c) p.v[v][c] = seed<P>(pr.v[v][c], 10 + 3 * v + c, FIRST);
    D_ I, gxi[3], gx[3];
    rankine_terms(gf, x0, x1, x2, p, I, gxi, gx);
#pragma unroll
    for (int m = 0; m < P; ++m) {
        // the normal's linear appearance in D/K is added separately (pass 0)
        const double dk = (FORM == 1) ? pr.n[0] * gxi[0].d[m] + pr.n[1] * gxi[1].d[m] + pr.n[2] * gxi[2].d[m]
                                      : ni[0] * gx[0].d[m] + ni[1] * gx[1].d[m] + ni[2] * gx[2].d[m];
        out[m] = cS * I.d[m] + cD * dk;
    }
}

template <int FORM>
__global__ __launch_bounds__(256) void k_vjp_rankine(int64_t n, const double* __restrict__ C,
                                                     const double* __restrict__ N,
                                                     const PanelRec* __restrict__ panels, int gf,
                                                     const double* __restrict__ Wr, const double* __restrict__ Wd,
                                                     double* __restrict__ colpart, double* __restrict__ rowpart) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double x[3] = {C[ic], C[ic + n], C[ic + 2 * n]};
    double ni[3] = {0.0, 0.0, 0.0};
    if (FORM == 2) {
        ni[0] = N[ic];
        ni[1] = N[ic + n];
        ni[2] = N[ic + 2 * n];
    }
    __shared__ double col[64][kColF];
    __shared__ double rowred[4][64][kRowF];
    for (int t = threadIdx.x; t < 64 * kColF; t += 256) (&col[0][0])[t] = 0.0;
    __syncthreads();
    double racc[kRowF];
#pragma unroll
    for (int f = 0; f < kRowF; ++f) racc[f] = 0.0;
    const int64_t jb = (int64_t)blockIdx.y * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& pr = panels[j];
        const double cS = row_ok ? -kInv4Pi * Wr[i + j * n] : 0.0;
        const double cD = row_ok ? -kInv4Pi * Wd[i + j * n] : 0.0;
        // pass 0: x, c, area (+ the normal's linear part of D/K)
        double o0[7];
        rankine_vjp_pass<FORM, 7, 0>(gf, x, pr, ni, cS, cD, o0);
        racc[0] += o0[0];
        racc[1] += o0[1];
        racc[2] += o0[2];
        col_add(o0[3], &col[jj][0], lane);
        col_add(o0[4], &col[jj][1], lane);
        col_add(o0[5], &col[jj][2], lane);
        col_add(o0[6], &col[jj][3], lane);
        {
            double I, gxi[3], gx[3];
            rankine_terms(gf, x[0], x[1], x[2], pr, I, gxi, gx);
            if (FORM == 1) {
#pragma unroll
                for (int c = 0; c < 3; ++c) col_add(cD * gxi[c], &col[jj][4 + c], lane);
            } else {
#pragma unroll
                for (int c = 0; c < 3; ++c) racc[3 + c] += cD * gx[c];
            }
        }
        // passes 1, 2: normal and vertices, near field only
        const double d0 = x[0] - pr.c[0], d1 = x[1] - pr.c[1];
        const double dz = x[2] - pr.c[2], dzr = -x[2] - pr.c[2];
        const double rr = 49.0 * pr.radius * pr.radius;
        const bool near = ((gf & GF_RANKINE) && d0 * d0 + d1 * d1 + dz * dz <= rr) ||
                          ((gf & (GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG)) && d0 * d0 + d1 * d1 + dzr * dzr <= rr);
        if (__any(near && row_ok)) {
            double o1[8], o2[7];
            rankine_vjp_pass<FORM, 8, 7>(gf, x, pr, ni, cS, cD, o1);   // n_j, V0, V1[0..1]
            rankine_vjp_pass<FORM, 7, 15>(gf, x, pr, ni, cS, cD, o2);  // V1[2], V2, V3
#pragma unroll
            for (int m = 0; m < 8; ++m) col_add(o1[m], &col[jj][4 + m], lane);
#pragma unroll
            for (int m = 0; m < 7; ++m) col_add(o2[m], &col[jj][12 + m], lane);
        }
    }
#pragma unroll
    for (int f = 0; f < kRowF; ++f) rowred[wave][lane][f] = racc[f];
    __syncthreads();
    // colpart[(tile_i * n + j) * 19 + f], rowpart[(tile_j * n + i) * 6 + f]
    for (int t = threadIdx.x; t < 64 * kColF; t += 256) {
        const int jj = t / kColF, f = t % kColF;
        if (jb + jj < n) colpart[((int64_t)blockIdx.x * n + jb + jj) * kColF + f] = col[jj][f];
    }
    if (wave == 0 && row_ok) {
#pragma unroll
        for (int f = 0; f < kRowF; ++f)
            rowpart[((int64_t)blockIdx.y * n + i) * kRowF + f] =
                rowred[0][lane][f] + rowred[1][lane][f] + rowred[2][lane][f] + rowred[3][lane][f];
    }
}

// Sum over wavenumbers of the real parts of nk cotangent matrices (either
// may be null = zero): out n x n real.
__global__ void k_sum_real(int64_t n, int64_t nk, const double2* __restrict__ A, int64_t lda, int64_t sa,
                           double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int64_t i = t % n, j = t / n;
    double s = 0.0;
    if (A)
        for (int64_t w = 0; w < nk; ++w) s += A[w * sa + i + j * lda].x;
    out[t] = s;
}

// ---------------------------------------------------------------------------
// Wave term, one wavenumber per loop step.  Local seeds (Dual<5>): dx, dy,
// zsum, area, k; x0 = +dx, c0 = -dx, x1 = +dy, c1 = -dy, x2 = c2 = +zsum.
// Sb, Db: complex cotangents (nk matrices, ld, stride; either may be null).
// ---------------------------------------------------------------------------
template <int WAVE, int FORM>
__global__ __launch_bounds__(256) void k_vjp_wave(int64_t n, const double* __restrict__ C,
                                                  const double* __restrict__ N,
                                                  const PanelRec* __restrict__ panels, int64_t nk,
                                                  const double* __restrict__ ks, const double2* __restrict__ Sb,
                                                  const double2* __restrict__ Db, int64_t ld, int64_t stride,
                                                  double* __restrict__ colpart, double* __restrict__ rowpart,
                                                  double* __restrict__ kpart) {
    typedef Dual<5> D_;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double x0 = C[ic], x1 = C[ic + n], x2 = C[ic + 2 * n];
    double ni[3] = {0.0, 0.0, 0.0};
    if (FORM == 2) {
        ni[0] = N[ic];
        ni[1] = N[ic + n];
        ni[2] = N[ic + 2 * n];
    }
    constexpr int kWCol = 7;  // c[3], area, n_j[3]
    __shared__ double col[64][kWCol];
    __shared__ double rowred[4][64][kRowF];
    __shared__ double kred[4];
    for (int t = threadIdx.x; t < 64 * kWCol; t += 256) (&col[0][0])[t] = 0.0;
    __syncthreads();
    double racc[kRowF];
#pragma unroll
    for (int f = 0; f < kRowF; ++f) racc[f] = 0.0;
    const int64_t jb = (int64_t)blockIdx.y * 64;
    const int64_t blk = blockIdx.x + (int64_t)gridDim.x * blockIdx.y;
    const int64_t nblk = (int64_t)gridDim.x * gridDim.y;
    for (int64_t w = 0; w < nk; ++w) {
        double kacc = 0.0;
        const D_ k = seed<5>(ks[w], 4, 0);
        for (int jj = wave; jj < 64; jj += 4) {
            const int64_t j = jb + jj;
            if (j >= n) break;
            const PanelRec& pr = panels[j];
            const double2 sb = (row_ok && Sb) ? Sb[w * stride + i + j * ld] : make_double2(0.0, 0.0);
            const double2 db = (row_ok && Db) ? Db[w * stride + i + j * ld] : make_double2(0.0, 0.0);
            const D_ dx = seed<5>(x0 - pr.c[0], 0, 0);
            const D_ dy = seed<5>(x1 - pr.c[1], 1, 0);
            const D_ zs = seed<5>(x2 + pr.c[2], 2, 0);
            const D_ ar = seed<5>(pr.area, 3, 0);
            cplxT<D_> Sv, gxi[3], gx[3];
            wave_term<WAVE>(k, dx, dy, zs, ar, Sv, gxi, gx);
            double o[5];
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                double kre, kim;
                if (FORM == 1) {
                    kre = pr.n[0] * gxi[0].re.d[m] + pr.n[1] * gxi[1].re.d[m] + pr.n[2] * gxi[2].re.d[m];
                    kim = pr.n[0] * gxi[0].im.d[m] + pr.n[1] * gxi[1].im.d[m] + pr.n[2] * gxi[2].im.d[m];
                } else {
                    kre = ni[0] * gx[0].re.d[m] + ni[1] * gx[1].re.d[m] + ni[2] * gx[2].re.d[m];
                    kim = ni[0] * gx[0].im.d[m] + ni[1] * gx[1].im.d[m] + ni[2] * gx[2].im.d[m];
                }
                o[m] = -kInv4Pi * (sb.x * Sv.re.d[m] + sb.y * Sv.im.d[m] + db.x * kre + db.y * kim);
            }
            racc[0] += o[0];
            racc[1] += o[1];
            racc[2] += o[2];
            kacc += o[4];
            col_add(-o[0], &col[jj][0], lane);
            col_add(-o[1], &col[jj][1], lane);
            col_add(o[2], &col[jj][2], lane);
            col_add(o[3], &col[jj][3], lane);
            // the normal's linear appearance: D = n_j . g (direct), K = n_i . g
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (FORM == 1)
                    col_add(-kInv4Pi * (db.x * gxi[c].re.v + db.y * gxi[c].im.v), &col[jj][4 + c], lane);
                else
                    racc[3 + c] += -kInv4Pi * (db.x * gx[c].re.v + db.y * gx[c].im.v);
            }
        }
        // k cotangent of this wavenumber: workgroup total
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) kacc += __shfl_xor(kacc, off);
        if (lane == 0) kred[wave] = kacc;
        __syncthreads();
        if (threadIdx.x == 0) kpart[w * nblk + blk] = kred[0] + kred[1] + kred[2] + kred[3];
        __syncthreads();
    }
#pragma unroll
    for (int f = 0; f < kRowF; ++f) rowred[wave][lane][f] = racc[f];
    __syncthreads();
    for (int t = threadIdx.x; t < 64 * kWCol; t += 256) {
        const int jj = t / kWCol, f = t % kWCol;
        if (jb + jj < n) colpart[((int64_t)blockIdx.x * n + jb + jj) * kWCol + f] = col[jj][f];
    }
    if (wave == 0 && row_ok) {
#pragma unroll
        for (int f = 0; f < kRowF; ++f)
            rowpart[((int64_t)blockIdx.y * n + i) * kRowF + f] =
                rowred[0][lane][f] + rowred[1][lane][f] + rowred[2][lane][f] + rowred[3][lane][f];
    }
}

// ---------------------------------------------------------------------------
// Finish: per face q, add the tile partials of both kernels; vertices are
// shared between faces, so their cotangents are added atomically.
// Outputs (all zeroed by the caller, may be null): dV nv x 3, dC nf x 3,
// dN nf x 3, dA nf (column-major).
// ---------------------------------------------------------------------------
__global__ void k_vjp_finish(int64_t n, int64_t nv, int64_t tiles, const int64_t* __restrict__ F,
                             const double* __restrict__ colR, const double* __restrict__ rowR,
                             const double* __restrict__ colW, const double* __restrict__ rowW,
                             double* __restrict__ dV, double* __restrict__ dC, double* __restrict__ dN,
                             double* __restrict__ dA) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    double c[kColF], r[kRowF];
    for (int f = 0; f < kColF; ++f) c[f] = 0.0;
    for (int f = 0; f < kRowF; ++f) r[f] = 0.0;
    for (int64_t t = 0; t < tiles; ++t) {
        if (colR)
            for (int f = 0; f < kColF; ++f) c[f] += colR[(t * n + q) * kColF + f];
        if (colW)
            for (int f = 0; f < 7; ++f) c[f] += colW[(t * n + q) * 7 + f];
        if (rowR)
            for (int f = 0; f < kRowF; ++f) r[f] += rowR[(t * n + q) * kRowF + f];
        if (rowW)
            for (int f = 0; f < kRowF; ++f) r[f] += rowW[(t * n + q) * kRowF + f];
    }
    if (dC)
        for (int k = 0; k < 3; ++k) dC[q + k * n] = c[k] + r[k];
    if (dA) dA[q] = c[3];
    if (dN)
        for (int k = 0; k < 3; ++k) dN[q + k * n] = c[4 + k] + r[3 + k];
    if (dV)
        for (int v = 0; v < 4; ++v) {
            const int64_t iv = F[q + v * n];
            if (iv < 0 || iv >= nv) continue;  // rejected when the panels were packed
            for (int k = 0; k < 3; ++k) atomicAdd(&dV[iv + k * nv], c[7 + 3 * v + k]);
        }
}

// dk[w] = sum over workgroups of kpart[w]
__global__ void k_vjp_dk(int64_t nk, int64_t nblk, const double* __restrict__ kpart, double* __restrict__ dk) {
    const int64_t w = blockIdx.x;
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nblk; b += 256) s += kpart[w * nblk + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) dk[w] = red[0] + red[1] + red[2] + red[3];
}

int64_t assemble_vjp_ws_doubles(int64_t n, int64_t nk) {
    const int64_t tiles = (n + 63) / 64;
    return 2 * n * n + tiles * n * (kColF + 7 + 2 * kRowF) + nk * tiles * tiles;
}

int launch_assemble_vjp(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, int gf, int form, int64_t nk,
                        const double* d_k, const double2* Sb, const double2* Db, int64_t ld, int64_t stride,
                        double* ws, double* dV, double* dC, double* dN, double* dA, double* dk) {
    const int64_t n = m->nfaces;
    const int64_t tiles = (n + 63) / 64;
    MH_REQUIRE(form == 1 || form == 2, "assemble_vjp: formulation must be direct (1) or indirect (2)");
    MH_REQUIRE(tiles <= 65535, "assemble_vjp: too many panels for one launch");
    const int rank = gf & (GF_RANKINE | GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG);
    const int wavem = gf & kGfWaveMask;  // the tabulated term is differentiated as the exact one
    double* Wr = ws;
    double* Wd = Wr + n * n;
    double* colR = Wd + n * n;
    double* colW = colR + tiles * n * kColF;
    double* rowR = colW + tiles * n * 7;
    double* rowW = rowR + tiles * n * kRowF;
    double* kpart = rowW + tiles * n * kRowF;
    const dim3 grid((unsigned)tiles, (unsigned)tiles);
    if (rank) {
        ProfScope ps(ctx, "vjp_rankine", 0.0, 16.0 * (double)n * n);
        const unsigned nb = (unsigned)((n * n + 255) / 256);
        hipLaunchKernelGGL(k_sum_real, dim3(nb), dim3(256), 0, ctx->stream, n, nk, Sb, ld, stride, Wr);
        hipLaunchKernelGGL(k_sum_real, dim3(nb), dim3(256), 0, ctx->stream, n, nk, Db, ld, stride, Wd);
        if (form == 1)
            hipLaunchKernelGGL((k_vjp_rankine<1>), grid, dim3(256), 0, ctx->stream, n, m->centers, m->normals,
                               panels, rank, Wr, Wd, colR, rowR);
        else
            hipLaunchKernelGGL((k_vjp_rankine<2>), grid, dim3(256), 0, ctx->stream, n, m->centers, m->normals,
                               panels, rank, Wr, Wd, colR, rowR);
        MH_HIP_CHECK(hipGetLastError());
    }
    if (wavem) {
        ProfScope ps(ctx, "vjp_wave", 0.0, 32.0 * (double)n * n * nk);
#define VW(W, FM)                                                                                             \
    hipLaunchKernelGGL((k_vjp_wave<W, FM>), grid, dim3(256), 0, ctx->stream, n, m->centers, m->normals, panels, \
                       nk, d_k, Sb, Db, ld, stride, colW, rowW, kpart)
        if (wavem & GF_WU) {
            if (form == 1) VW(1, 1);
            else VW(1, 2);
        } else {
            if (form == 1) VW(2, 1);
            else VW(2, 2);
        }
#undef VW
        MH_HIP_CHECK(hipGetLastError());
        if (dk) {
            hipLaunchKernelGGL(k_vjp_dk, dim3((unsigned)nk), dim3(256), 0, ctx->stream, nk, tiles * tiles, kpart, dk);
            MH_HIP_CHECK(hipGetLastError());
        }
    } else if (dk) {
        MH_HIP_CHECK(hipMemsetAsync(dk, 0, sizeof(double) * nk, ctx->stream));
    }
    if (dV) MH_HIP_CHECK(hipMemsetAsync(dV, 0, sizeof(double) * 3 * m->nvertices, ctx->stream));
    hipLaunchKernelGGL(k_vjp_finish, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream, n,
                       m->nvertices, tiles, m->faces, rank ? colR : nullptr, rank ? rowR : nullptr,
                       wavem ? colW : nullptr, wavem ? rowW : nullptr, dV, dC, dN, dA);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// ---------------------------------------------------------------------------
// Solve reverse mode helpers
// ---------------------------------------------------------------------------
// Y = conj(X) (n x cols per item; may alias)
__global__ void k_conj(int64_t rows, int64_t cols, const double2* __restrict__ X, int64_t ldx, int64_t sx,
                       double2* __restrict__ Y, int64_t ldy, int64_t sy) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = blockIdx.y;
    if (t >= rows * cols) return;
    const int64_t i = t % rows, j = t / rows;
    const double2 v = X[b * sx + i + j * ldx];
    Y[b * sy + i + j * ldy] = make_double2(v.x, -v.y);
}

// M[i, j] = alpha * sum_r X[i, r] conj(Y[j, r]) (n x n per item, r < nr)
__global__ __launch_bounds__(256) void k_outer_conj(int64_t n, int64_t nr, double alpha,
                                                    const double2* __restrict__ X, const double2* __restrict__ Y,
                                                    int64_t ldb, int64_t bs, double2* __restrict__ M, int64_t ld,
                                                    int64_t stride) {
    const int64_t b = blockIdx.z;
    const int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int64_t j0 = (int64_t)blockIdx.y * 16 + (threadIdx.x >> 6) * 4;
    if (i >= n) return;
    for (int q = 0; q < 4; ++q) {
        const int64_t j = j0 + q;
        if (j >= n) break;
        double re = 0.0, im = 0.0;
        for (int64_t r = 0; r < nr; ++r) {
            const double2 x = X[b * bs + i + r * ldb];
            const double2 y = Y[b * bs + j + r * ldb];
            re += x.x * y.x + x.y * y.y;
            im += x.y * y.x - x.x * y.y;
        }
        M[b * stride + i + j * ld] = make_double2(alpha * re, alpha * im);
    }
}

int conj_batched(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const double2* X, int64_t ldx, int64_t sx,
                 double2* Y, int64_t ldy, int64_t sy) {
    if (rows * cols == 0 || batch == 0) return MH_OK;
    hipLaunchKernelGGL(k_conj, dim3((unsigned)((rows * cols + 255) / 256), (unsigned)batch), dim3(256), 0,
                       ctx->stream, rows, cols, X, ldx, sx, Y, ldy, sy);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int outer_conj_batched(Context* ctx, int64_t n, int64_t nr, int64_t batch, double alpha, const double2* X,
                       const double2* Y, int64_t ldb, int64_t bs, double2* M, int64_t ld, int64_t stride) {
    if (n == 0 || batch == 0) return MH_OK;
    MH_REQUIRE(batch <= 65535, "outer product: batch too large for one launch");
    ProfScope ps(ctx, "outer_conj", 8.0 * (double)n * n * nr * batch, 16.0 * (double)n * n * batch);
    hipLaunchKernelGGL(k_outer_conj, dim3((unsigned)((n + 63) / 64), (unsigned)((n + 15) / 16), (unsigned)batch),
                       dim3(256), 0, ctx->stream, n, nr, alpha, X, Y, ldb, bs, M, ld, stride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

}  // namespace mh
