// Fused influence-matrix assembly for gfx950.
//
// One pass over every panel pair (i = collocation row, j = source column)
// evaluates the Rankine and image-Rankine panel integrals and their gradients
// once, then the wave term for a group of wavenumbers, and writes S together
// with D (direct), K (indirect) or both.  Reference semantics:
// src/matrix_assembly.jl:4-42 (comprehension assembly; S at :14, D/K at
// :16-26), summing the terms of src/green_functions/*.jl.
//
// Launch shape: a 256-thread workgroup covers a 64 x 64 tile of (i, j);
// each wave owns 16 source columns and its 64 lanes own 64 consecutive rows.
// The source panel is wave-uniform (scalar loads of a packed 192-byte record)
// and the 64 lanes of a wave write one contiguous 1 KiB run of each
// column-major complex128 matrix.  Rankine terms do not depend on the
// wavenumber, so they are computed once per pair and reused for every
// wavenumber of the group (blockIdx.z).  With a wave term the launch is
// split: the Rankine parts once (k_rankine_parts), then the wave term once
// per unordered pair (k_assemble_sym), since it is symmetric in (i, j).
#include <mutex>

#include "mh_internal.hpp"

namespace mh {

__global__ void k_pack_panels(int64_t nf, int64_t nv, const double* __restrict__ V,
                              const int64_t* __restrict__ F, const double* __restrict__ C,
                              const double* __restrict__ N, const double* __restrict__ A,
                              const double* __restrict__ R, PanelRec* __restrict__ out,
                              int* __restrict__ flag) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nf) return;
    PanelRec p;
    for (int c = 0; c < 3; ++c) {
        p.c[c] = C[j + c * nf];
        p.n[c] = N[j + c * nf];
    }
    p.area = A[j];
    p.radius = R[j];
    for (int v = 0; v < 4; ++v) {
        int64_t iv = F[j + v * nf];  // 0-based, reference src/meshes.jl:179
        if (iv < 0 || iv >= nv) {
            atomicOr(flag, 1);
            iv = 0;
        }
        for (int c = 0; c < 3; ++c) p.v[v][c] = V[iv + c * nv];
    }
    for (int c = 0; c < 4; ++c) p.pad[c] = 0.0;
    out[j] = p;
}

constexpr int kFG = 16;  // wavenumbers per thread (blockIdx.z group)

template <int WAVE, int FORM>
__global__ __launch_bounds__(256) void k_assemble(
    int64_t n, const double* __restrict__ C, const double* __restrict__ N,
    const PanelRec* __restrict__ panels, int gf, int64_t nk, const double* __restrict__ ks,
    double2* __restrict__ S, double2* __restrict__ D, double2* __restrict__ K, int64_t ld,
    int64_t stride, float2* __restrict__ DK32, DelhTab tab) {
    // DK32 (optional): complex64 copy of D (FORM 1) or K (FORM 2), n x n per
    // wavenumber, for the FP32 factorisation of the mixed-precision path
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double x0 = C[ic], x1 = C[ic + n], x2 = C[ic + 2 * n];
    double ni0 = 0.0, ni1 = 0.0, ni2 = 0.0;
    if (FORM & 2) {
        ni0 = N[ic];
        ni1 = N[ic + n];
        ni2 = N[ic + 2 * n];
    }
    const int64_t jb = (int64_t)blockIdx.y * 64;
    const int64_t f0 = (int64_t)blockIdx.z * kFG;
    const int64_t nfg = (nk - f0) < kFG ? (nk - f0) : kFG;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& p = panels[j];
        double IR, gxiR[3], gxR[3];
        rankine_terms(gf, x0, x1, x2, p, IR, gxiR, gxR);
        const double nj0 = p.n[0], nj1 = p.n[1], nj2 = p.n[2];
        const double diag = (i == j) ? 0.5 : 0.0;
        for (int64_t f = 0; f < nfg; ++f) {
            const int64_t m = f0 + f;
            cplx Sv = {0.0, 0.0}, gxi[3] = {{0, 0}, {0, 0}, {0, 0}}, gx[3] = {{0, 0}, {0, 0}, {0, 0}};
            if (WAVE) wave_term<WAVE>(ks[m], x0 - p.c[0], x1 - p.c[1], x2 + p.c[2], p.area, Sv, gxi, gx, tab);
            if (!row_ok) continue;
            const int64_t off = m * stride + i + j * ld;
            S[off] = make_double2(-kInv4Pi * (IR + Sv.re), -kInv4Pi * Sv.im);
            if (FORM & 1) {
                const double re = nj0 * (gxiR[0] + gxi[0].re) + nj1 * (gxiR[1] + gxi[1].re) +
                                  nj2 * (gxiR[2] + gxi[2].re);
                const double im = nj0 * gxi[0].im + nj1 * gxi[1].im + nj2 * gxi[2].im;
                const double2 v = make_double2(diag - kInv4Pi * re, -kInv4Pi * im);
                D[off] = v;
                if (DK32) DK32[m * n * n + i + j * n] = make_float2((float)v.x, (float)v.y);
            }
            if (FORM & 2) {
                const double re = ni0 * (gxR[0] + gx[0].re) + ni1 * (gxR[1] + gx[1].re) +
                                  ni2 * (gxR[2] + gx[2].re);
                const double im = ni0 * gx[0].im + ni1 * gx[1].im + ni2 * gx[2].im;
                const double2 v = make_double2(diag - kInv4Pi * re, -kInv4Pi * im);
                K[off] = v;
                if (DK32 && !(FORM & 1)) DK32[m * n * n + i + j * n] = make_float2((float)v.x, (float)v.y);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Symmetric wave term.  The collocation points are the panel centres, so the
// wave term of the pair (i, j) and of (j, i) is one evaluation at the same
// reduced coordinates (horizontal distance and z_i + z_j); only the source
// area, the sign of the horizontal gradient and the normals differ.
// k_rankine_parts stores the frequency-independent Rankine parts once per
// launch (real n x n: RS = I, RD = n_j . grad_xi, RK = n_i . grad_x), and
// k_assemble_sym covers each unordered tile pair (I <= J) once: entry (i, j)
// is written straight from the lanes (rows i), entry (j, i) goes through LDS
// so that it is written down columns as well.
// ---------------------------------------------------------------------------
template <int FORM>
__global__ __launch_bounds__(256) void k_rankine_parts(int64_t n, const double* __restrict__ C,
                                                       const double* __restrict__ N,
                                                       const PanelRec* __restrict__ panels, int gf,
                                                       double* __restrict__ RS, double* __restrict__ RD,
                                                       double* __restrict__ RK) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double x0 = C[ic], x1 = C[ic + n], x2 = C[ic + 2 * n];
    double ni0 = 0.0, ni1 = 0.0, ni2 = 0.0;
    if (FORM & 2) {
        ni0 = N[ic];
        ni1 = N[ic + n];
        ni2 = N[ic + 2 * n];
    }
    const int64_t jb = (int64_t)blockIdx.y * 64;
    for (int jj = wave; jj < 64; jj += 4) {
        const int64_t j = jb + jj;
        if (j >= n) break;
        const PanelRec& p = panels[j];
        double IR, gxiR[3], gxR[3];
        rankine_terms(gf, x0, x1, x2, p, IR, gxiR, gxR);
        if (!row_ok) continue;
        const int64_t off = i + j * n;
        RS[off] = IR;
        if (FORM & 1) RD[off] = p.n[0] * gxiR[0] + p.n[1] * gxiR[1] + p.n[2] * gxiR[2];
        if (FORM & 2) RK[off] = ni0 * gxR[0] + ni1 * gxR[1] + ni2 * gxR[2];
    }
}

constexpr int kSymFG = 8;  // wavenumbers per workgroup (blockIdx.y group)
// rounds (4 source columns each, one per wave) per transposed write-out: 4
// where the kernel's time goes into its matrix writes (Wu, tabulated
// Delhommeau), 1 for the exact Delhommeau term, whose quadrature per pair
// dominates and which keeps its registers (and waves per SIMD) that way
template <int WAVE>
constexpr int sym_rounds() {
    return WAVE == 2 ? 1 : 4;
}

// Loop order: groups of RG rounds (4 RG source columns j), then the
// workgroup's wavenumbers, then the rounds of the group.  A round's lanes
// write the (i, j) entries straight down column j; its (j, i) entries are
// staged in LDS and, once the group's 4 RG rows are there, written 4 RG rows
// at a time: with RG = 4, 256-byte runs per column (128 for the FP32 copy)
// instead of the 64-byte runs of one round, so the L2 lines of the
// transposed half leave the cache whole rather than half written.  The group's horizontal distances
// and Rankine parts stay in registers across the wavenumbers.
// four per-round values in named registers, read by selection (an indexed
// register array would be placed in scratch memory)
struct Held4 {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    __device__ void set(int t, double v) {
        if (t == 0) a0 = v;
        if (t == 1) a1 = v;
        if (t == 2) a2 = v;
        if (t == 3) a3 = v;
    }
    __device__ double get(int t) const { return t == 0 ? a0 : (t == 1 ? a1 : (t == 2 ? a2 : a3)); }
};

template <int WAVE, int FORM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVE == 1 && FORM != 3 ? 3 : 1))) void k_assemble_sym(
    int64_t n, const double* __restrict__ C, const double* __restrict__ N, const PanelRec* __restrict__ panels,
    int64_t nk, const double* __restrict__ ks, const double* __restrict__ RS, const double* __restrict__ RD,
    const double* __restrict__ RK, double2* __restrict__ S, double2* __restrict__ D, double2* __restrict__ K,
    int64_t ld, int64_t stride, float2* __restrict__ DK32, DelhTab tab) {
    // tile pair (I, J), I <= J: blockIdx.x = J (J + 1) / 2 + I
    const int pidx = (int)blockIdx.x;
    int J = (int)((::sqrt(8.0 * pidx + 1.0) - 1.0) * 0.5);
    while ((J + 1) * (J + 2) / 2 <= pidx) ++J;
    while (J * (J + 1) / 2 > pidx) --J;
    const int I = pidx - J * (J + 1) / 2;
    const bool dtile = I == J;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t i = (int64_t)I * 64 + lane;
    const bool row_ok = i < n;
    const int64_t ic = row_ok ? i : n - 1;
    const double x0 = C[ic], x1 = C[ic + n], x2 = C[ic + 2 * n];
    const double ni0 = N[ic], ni1 = N[ic + n], ni2 = N[ic + 2 * n];
    const double ai = panels[ic].area;
    const int64_t f0 = (int64_t)blockIdx.y * kSymFG;
    const int nfg = (int)((nk - f0) < kSymFG ? (nk - f0) : kSymFG);
    constexpr int RG = sym_rounds<WAVE>();
    constexpr int kRows = 4 * RG;  // rows j of a group
    // wave parts of the group's (j, i) entries for one wavenumber:
    // [j - first row of the group][i - column], rows padded against bank conflicts
    __shared__ double2 sS[kRows][64 + 1];
    __shared__ double2 sD[(FORM & 1) ? kRows : 1][64 + 1];
    __shared__ double2 sK[(FORM & 2) ? kRows : 1][64 + 1];
    // write-out: thread -> entries (row j = first + (e & 15), column i = I*64 + (e >> 4)), e = tid + 256 u
    constexpr int kOut = kRows * 64 / 256;
    static_assert(RG == 1 || RG == 4, "the held values are four named registers");
    for (int tg = 0; tg < 16 / RG; ++tg) {
        Held4 hd, rs, rd, rk;
#pragma unroll
        for (int tt = 0; tt < RG; ++tt) {
            const int64_t j = (int64_t)J * 64 + 4 * (tg * RG + tt) + wave;  // wave-uniform source column
            const bool jv = j < n;
            const bool dir = row_ok && jv && (!dtile || j >= i);
            const PanelRec& pj = panels[jv ? j : n - 1];
            double vs = 0.0, vd = 0.0, vk = 0.0;
            if (dir && RS) {
                vs = RS[i + j * n];
                if (FORM & 1) vd = RD[i + j * n];
                if (FORM & 2) vk = RK[i + j * n];
            }
            hd.set(tt, horiz_dist(x0 - pj.c[0], x1 - pj.c[1]));
            rs.set(tt, vs);
            rd.set(tt, vd);
            rk.set(tt, vk);
        }
        // Rankine parts of this thread's transposed entries
        const int64_t qrow0 = (int64_t)J * 64 + kRows * tg + (threadIdx.x & (kRows - 1));
        auto qcol = [&](int u) { return (int64_t)I * 64 + (int)(threadIdx.x + 256 * u) / kRows; };
        auto qok = [&](int u) { return qrow0 < n && qcol(u) < n && (!dtile || qrow0 > qcol(u)); };
        double qs[kOut], qd[kOut], qk[kOut];
#pragma unroll
        for (int u = 0; u < kOut; ++u) {
            qs[u] = qd[u] = qk[u] = 0.0;
            if (qok(u) && RS) {
                qs[u] = RS[qrow0 + qcol(u) * n];
                if (FORM & 1) qd[u] = RD[qrow0 + qcol(u) * n];
                if (FORM & 2) qk[u] = RK[qrow0 + qcol(u) * n];
            }
        }
#pragma unroll 1
        for (int f = 0; f < nfg; ++f) {
            const int64_t m = f0 + f;
            const double k = ks[m];
#pragma unroll 1
            for (int tt = 0; tt < RG; ++tt) {
                const int r = 4 * tt + wave;  // row of the group
                const int64_t j = (int64_t)J * 64 + kRows * tg + r;
                const bool jv = j < n;
                const bool dir = row_ok && jv && (!dtile || j >= i);
                const bool trn = row_ok && jv && (!dtile || j > i);
                if (!(dir || trn)) continue;
                const PanelRec& pj = panels[j];
                const double nj0 = pj.n[0], nj1 = pj.n[1], nj2 = pj.n[2];
                const double aj = pj.area;
                const double dx = x0 - pj.c[0], dy = x1 - pj.c[1], zs = x2 + pj.c[2];
                cplx G, Gr, Gz;
                double d1, d2;
                wave_raw<WAVE>(k, dx, dy, zs, hd.get(tt), G, Gr, Gz, d1, d2, tab);
                // gradient w.r.t. the source point of (i, j), without the area:
                // (d1 Gr, d2 Gr, k Gz); of (j, i): (-d1 Gr, -d2 Gr, k Gz)
                const double kaj = k * aj, kai = k * ai;
                if (dir) {
                    const double dg = (i == j) ? 0.5 : 0.0;
                    const int64_t off = m * stride + i + j * ld;
                    S[off] = make_double2(-kInv4Pi * (rs.get(tt) + kaj * G.re), -kInv4Pi * (kaj * G.im));
                    if (FORM & 1) {
                        const double wr = kaj * (nj0 * (d1 * Gr.re) + nj1 * (d2 * Gr.re) + nj2 * (k * Gz.re));
                        const double wi = kaj * (nj0 * (d1 * Gr.im) + nj1 * (d2 * Gr.im) + nj2 * (k * Gz.im));
                        const double2 v = make_double2(dg - kInv4Pi * (rd.get(tt) + wr), -kInv4Pi * wi);
                        D[off] = v;
                        if (DK32) DK32[m * n * n + i + j * n] = make_float2((float)v.x, (float)v.y);
                    }
                    if (FORM & 2) {
                        const double wr = kaj * (ni2 * (k * Gz.re) - ni0 * (d1 * Gr.re) - ni1 * (d2 * Gr.re));
                        const double wi = kaj * (ni2 * (k * Gz.im) - ni0 * (d1 * Gr.im) - ni1 * (d2 * Gr.im));
                        const double2 v = make_double2(dg - kInv4Pi * (rk.get(tt) + wr), -kInv4Pi * wi);
                        K[off] = v;
                        if (DK32 && !(FORM & 1)) DK32[m * n * n + i + j * n] = make_float2((float)v.x, (float)v.y);
                    }
                }
                if (trn) {
                    sS[r][lane] = make_double2(kai * G.re, kai * G.im);
                    if (FORM & 1)
                        sD[r][lane] = make_double2(kai * (ni2 * (k * Gz.re) - ni0 * (d1 * Gr.re) - ni1 * (d2 * Gr.re)),
                                                   kai * (ni2 * (k * Gz.im) - ni0 * (d1 * Gr.im) - ni1 * (d2 * Gr.im)));
                    if (FORM & 2)
                        sK[r][lane] = make_double2(kai * (nj0 * (d1 * Gr.re) + nj1 * (d2 * Gr.re) + nj2 * (k * Gz.re)),
                                                   kai * (nj0 * (d1 * Gr.im) + nj1 * (d2 * Gr.im) + nj2 * (k * Gz.im)));
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kOut; ++u) {
                if (!qok(u)) continue;
                const int e = threadIdx.x + 256 * u;
                const int r = e & (kRows - 1), c = e / kRows;
                const int64_t off = m * stride + qrow0 + qcol(u) * ld;
                const double2 ws = sS[r][c];
                S[off] = make_double2(-kInv4Pi * (qs[u] + ws.x), -kInv4Pi * ws.y);
                if (FORM & 1) {
                    const double2 w = sD[r][c];
                    const double2 v = make_double2(-kInv4Pi * (qd[u] + w.x), -kInv4Pi * w.y);
                    D[off] = v;
                    if (DK32) DK32[m * n * n + qrow0 + qcol(u) * n] = make_float2((float)v.x, (float)v.y);
                }
                if (FORM & 2) {
                    const double2 w = sK[r][c];
                    const double2 v = make_double2(-kInv4Pi * (qk[u] + w.x), -kInv4Pi * w.y);
                    K[off] = v;
                    if (DK32 && !(FORM & 1))
                        DK32[m * n * n + qrow0 + qcol(u) * n] = make_float2((float)v.x, (float)v.y);
                }
            }
            __syncthreads();
        }
    }
}

__global__ void k_clear_int(int* __restrict__ p, int count) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < count) p[t] = 0;
}

int clear_ints(Context* ctx, int* p, int64_t count) {
    if (count <= 0) return MH_OK;
    MH_REQUIRE(count <= 0x7fffffff, "clear_ints: count too large");
    hipLaunchKernelGGL(k_clear_int, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, ctx->stream, p,
                       (int)count);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

__global__ void k_fill_zero16(uint4* __restrict__ p, int64_t count) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = make_uint4(0u, 0u, 0u, 0u);
}

int fill_zero(Context* ctx, void* p, size_t bytes) {
    if (bytes == 0) return MH_OK;
    MH_REQUIRE(bytes % 16 == 0 && (reinterpret_cast<uintptr_t>(p) & 15) == 0,
               "fill_zero: 16-byte aligned multiples of 16 bytes only");
    const int64_t count = (int64_t)(bytes / 16);
    const int64_t blocks = (count + 255) / 256 < 1024 ? (count + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_fill_zero16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, static_cast<uint4*>(p),
                       count);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

__global__ void k_copy_d2(double2* __restrict__ dst, int64_t ldd, const double2* __restrict__ src, int64_t lds,
                          int64_t rows, int64_t cols) {
    for (int64_t j = blockIdx.y; j < cols; j += gridDim.y)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x)
            dst[i + j * ldd] = src[i + j * lds];
}

int copy_d2(Context* ctx, double2* dst, int64_t ldd, const double2* src, int64_t lds, int64_t rows, int64_t cols) {
    if (rows <= 0 || cols <= 0) return MH_OK;
    const int64_t bx = (rows + 255) / 256 < 512 ? (rows + 255) / 256 : 512;
    const int64_t by = cols < 1024 ? cols : 1024;
    hipLaunchKernelGGL(k_copy_d2, dim3((unsigned)bx, (unsigned)by), dim3(256), 0, ctx->stream, dst, ldd, src, lds,
                       rows, cols);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

__global__ void k_or_words(const int* __restrict__ src, int64_t count, int* __restrict__ dst) {
    int v = 0;
    for (int64_t t = threadIdx.x; t < count; t += blockDim.x) v |= src[t];
    if (v) atomicOr(dst, v);
}

int or_words(Context* ctx, const int* src, int64_t count, int* dst) {
    if (count <= 0) return MH_OK;
    hipLaunchKernelGGL(k_or_words, dim3(1), dim3(256), 0, ctx->stream, src, count, dst);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

__global__ void k_ir_verdict(const double* __restrict__ nrm, int64_t count, double tol, int* __restrict__ dst) {
    int v = 0;
    for (int64_t t = threadIdx.x; t < count; t += blockDim.x) {
        const double x = nrm[t];
        if (!(x < tol)) v |= 1;  // a NaN compares false: not converged
        if (!(x <= 1e300)) v |= 2;
    }
    if (v) atomicOr(dst, v);
}

int ir_verdict(Context* ctx, const double* nrm, int64_t count, double tol, int* dst) {
    if (count <= 0) return MH_OK;
    hipLaunchKernelGGL(k_ir_verdict, dim3(1), dim3(256), 0, ctx->stream, nrm, count, tol, dst);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int launch_pack_panels(Context* ctx, const mh_mesh_t* m, PanelRec* out, int* flag_buf) {
    const int64_t nf = m->nfaces;
    if (nf == 0) return MH_OK;
    int* d_flag = flag_buf ? flag_buf : ctx->d_flag;
    if (flag_buf)  // a fill kernel, not a memset node, for callers that capture
        hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, ctx->stream, d_flag, 1);
    else
        MH_HIP_CHECK(hipMemsetAsync(d_flag, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_pack_panels, dim3((unsigned)((nf + 127) / 128)), dim3(128), 0, ctx->stream,
                       nf, m->nvertices, m->vertices, m->faces, m->centers, m->normals, m->areas,
                       m->radii, out, d_flag);
    MH_HIP_CHECK(hipGetLastError());
    if (ctx->capturing) return MH_OK;  // the flag is read after the graph's launch
    int flag = 0;
    MH_HIP_CHECK(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (flag) {
        set_error("mesh: a face references a vertex index outside [0, nvertices)");
        return MH_ERR_ARGUMENT;
    }
    return MH_OK;
}

template <int WAVE>
static void dispatch_form(int form, dim3 grid, hipStream_t st, int64_t n, const double* C,
                          const double* N, const PanelRec* P, int gf, int64_t nk,
                          const double* ks, double2* S, double2* D, double2* K, int64_t ld,
                          int64_t stride, float2* DK32, DelhTab tab = DelhTab{}) {
    switch (form) {
        case 1:
            hipLaunchKernelGGL((k_assemble<WAVE, 1>), grid, dim3(256), 0, st, n, C, N, P, gf, nk, ks,
                               S, D, K, ld, stride, DK32, tab);
            break;
        case 2:
            hipLaunchKernelGGL((k_assemble<WAVE, 2>), grid, dim3(256), 0, st, n, C, N, P, gf, nk, ks,
                               S, D, K, ld, stride, DK32, tab);
            break;
        default:
            hipLaunchKernelGGL((k_assemble<WAVE, 3>), grid, dim3(256), 0, st, n, C, N, P, gf, nk, ks,
                               S, D, K, ld, stride, DK32, tab);
            break;
    }
}

int launch_assemble(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, int gf, int form,
                    int64_t nk, const double* d_k, double2* S, double2* D, double2* K, int64_t ld,
                    int64_t stride, float2* DK32) {
    const int64_t n = m->nfaces;
    if (n == 0 || nk == 0) return MH_OK;
    if (ctx->far32) gf |= kGfFar32;  // read by the Rankine pair code only
    DelhTab tab{};
    if (gf & GF_DELH_TABLE) TRY_RC(get_delh_table(ctx, &tab));
    const int64_t tiles = (n + 63) / 64;
    const int64_t groups = (nk + kFG - 1) / kFG;
    if (tiles > 0x7fffffff || groups > 65535) {
        set_error("assemble: problem too large for one launch");
        return MH_ERR_ARGUMENT;
    }
    dim3 grid((unsigned)tiles, (unsigned)tiles, (unsigned)groups);
    const double nmat = 1.0 + ((form & 1) ? 1.0 : 0.0) + ((form & 2) ? 1.0 : 0.0);
    if (gf & kGfWaveMask) {
        // Rankine parts once, then the symmetric wave-term pass
        const bool rank = (gf & (GF_RANKINE | GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG)) != 0;
        const int nparts = 1 + ((form & 1) ? 1 : 0) + ((form & 2) ? 1 : 0);
        double *RS = nullptr, *RD = nullptr, *RK = nullptr;
        if (rank) {
            const size_t need = sizeof(double) * (size_t)n * (size_t)n * (size_t)nparts;
            if (need > ctx->rk_bytes) {
                if (ctx->rk) {
                    MH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                    MH_HIP_CHECK(hipFree(ctx->rk));
                    ctx->rk = nullptr;
                    ctx->rk_bytes = 0;
                }
                MH_HIP_CHECK(hipMalloc(&ctx->rk, need));
                ctx->rk_bytes = need;
            }
            double* b = static_cast<double*>(ctx->rk);
            RS = b;
            int q = 1;
            if (form & 1) RD = b + (size_t)(q++) * n * n;
            if (form & 2) RK = b + (size_t)(q++) * n * n;
            ProfScope pr(ctx, "rankine_parts", 0.0, 8.0 * nparts * (double)n * (double)n);
            dim3 g2((unsigned)tiles, (unsigned)tiles);
#define MH_RP(F) hipLaunchKernelGGL((k_rankine_parts<F>), g2, dim3(256), 0, ctx->stream, n, m->centers, \
                                    m->normals, panels, gf, RS, RD, RK)
            if (form == 1) MH_RP(1);
            else if (form == 2) MH_RP(2);
            else MH_RP(3);
#undef MH_RP
            MH_HIP_CHECK(hipGetLastError());
        }
        const int64_t pairs = tiles * (tiles + 1) / 2;
        const int64_t sgroups = (nk + kSymFG - 1) / kSymFG;
        if (pairs > 0x7fffffff || sgroups > 65535) {
            set_error("assemble: problem too large for one launch");
            return MH_ERR_ARGUMENT;
        }
        dim3 gs((unsigned)pairs, (unsigned)sgroups);
        ProfScope ps(ctx, "assemble", 0.0, 16.0 * nmat * (double)n * (double)n * (double)nk);
#define MH_AS(W, F) hipLaunchKernelGGL((k_assemble_sym<W, F>), gs, dim3(256), 0, ctx->stream, n, m->centers, \
                                       m->normals, panels, nk, d_k, RS, RD, RK, S, D, K, ld, stride, DK32, tab)
        if (gf & GF_WU) {
            if (form == 1) MH_AS(1, 1);
            else if (form == 2) MH_AS(1, 2);
            else MH_AS(1, 3);
        } else if (gf & GF_DELH_TABLE) {
            if (form == 1) MH_AS(3, 1);
            else if (form == 2) MH_AS(3, 2);
            else MH_AS(3, 3);
        } else {
            if (form == 1) MH_AS(2, 1);
            else if (form == 2) MH_AS(2, 2);
            else MH_AS(2, 3);
        }
#undef MH_AS
        MH_HIP_CHECK(hipGetLastError());
        return MH_OK;
    }
    ProfScope ps(ctx, "assemble", 0.0, 16.0 * nmat * (double)n * (double)n * (double)nk);
    if (gf & GF_WU)
        dispatch_form<1>(form, grid, ctx->stream, n, m->centers, m->normals, panels, gf, nk, d_k, S,
                         D, K, ld, stride, DK32);
    else if (gf & GF_DELHOMMEAU)
        dispatch_form<2>(form, grid, ctx->stream, n, m->centers, m->normals, panels, gf, nk, d_k, S,
                         D, K, ld, stride, DK32);
    else
        dispatch_form<0>(form, grid, ctx->stream, n, m->centers, m->normals, panels, gf, nk, d_k, S,
                         D, K, ld, stride, DK32);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// One-pass assembly (k_assemble: Rankine and wave terms of each pair in one
// kernel, no frequency-independent scratch matrices).  No kernel timer.
int launch_assemble_onepass(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, int gf, int form,
                            int64_t nk, const double* d_k, double2* S, double2* D, double2* K, int64_t ld,
                            int64_t stride) {
    const int64_t n = m->nfaces;
    if (n == 0 || nk == 0) return MH_OK;
    const int64_t tiles = (n + 63) / 64;
    const int64_t groups = (nk + kFG - 1) / kFG;
    if (tiles > 0x7fffffff || groups > 65535) {
        set_error("assemble: problem too large for one launch");
        return MH_ERR_ARGUMENT;
    }
    dim3 grid((unsigned)tiles, (unsigned)tiles, (unsigned)groups);
    MH_REQUIRE(!(gf & GF_DELH_TABLE), "assemble: the one-pass kernel takes the exact Delhommeau term only");
    if (gf & GF_WU)
        dispatch_form<1>(form, grid, ctx->stream, n, m->centers, m->normals, panels, gf, nk, d_k, S, D, K, ld,
                         stride, nullptr);
    else if (gf & GF_DELHOMMEAU)
        dispatch_form<2>(form, grid, ctx->stream, n, m->centers, m->normals, panels, gf, nk, d_k, S, D, K, ld,
                         stride, nullptr);
    else
        dispatch_form<0>(form, grid, ctx->stream, n, m->centers, m->normals, panels, gf, nk, d_k, S, D, K, ld,
                         stride, nullptr);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// ---------------------------------------------------------------------------
// Per-pair evaluation (greens / gradient_greens / integral / integral_gradient)
// ---------------------------------------------------------------------------
__global__ void k_gf_eval(int gf, int op, int wrt_first, int64_t np,
                          const double* __restrict__ pts, const mh_panel_t* __restrict__ pan,
                          double k, double2* __restrict__ out, DelhTab tab) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const double x0 = pts[t], x1 = pts[t + np], x2 = pts[t + 2 * np];
    const mh_panel_t& q = pan[t];
    PanelRec p;
    for (int c = 0; c < 3; ++c) {
        p.c[c] = q.center[c];
        p.n[c] = q.normal[c];
        for (int v = 0; v < 4; ++v) p.v[v][c] = q.vertices[v][c];
    }
    p.area = q.area;
    p.radius = q.radius;
    double val_re = 0.0, val_im = 0.0;
    double g_re[3] = {0, 0, 0}, g_im[3] = {0, 0, 0};
    const bool grad = (op == MH_OP_GRADIENT_GREENS || op == MH_OP_INTEGRAL_GRADIENT);
    const bool panel_int = (op == MH_OP_INTEGRAL || op == MH_OP_INTEGRAL_GRADIENT);
    // Rankine terms
    const double sref = ((gf & GF_RANKINE_REFLECTED) ? 1.0 : 0.0) -
                        ((gf & GF_RANKINE_REFLECTED_NEG) ? 1.0 : 0.0);
    if (panel_int) {
        double I, gxi[3], gx[3];
        rankine_terms(gf, x0, x1, x2, p, I, gxi, gx);
        val_re += I;
        for (int c = 0; c < 3; ++c) g_re[c] += wrt_first ? gx[c] : gxi[c];
    } else {
        // point values: rankine.jl:8-22, rankine_reflected.jl:1-14
        if (gf & GF_RANKINE) {
            const double d0 = x0 - p.c[0], d1 = x1 - p.c[1], d2 = x2 - p.c[2];
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            val_re += 1.0 / r;
            const double r3 = r * r * r;
            const double s = wrt_first ? 1.0 : -1.0;  // grad_x = (xi - x)/r^3
            g_re[0] += s * (-d0) / r3;
            g_re[1] += s * (-d1) / r3;
            g_re[2] += s * (-d2) / r3;
        }
        if (sref != 0.0) {
            // value with the collocation point reflected
            const double d0 = x0 - p.c[0], d1 = x1 - p.c[1], d2 = -x2 - p.c[2];
            const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            val_re += sref / r;
            const double r3 = r * r * r;
            if (wrt_first) {
                // gradient_greens(Rankine, e1, reflect(e2); true) = (reflect(xi) - x)/r^3
                g_re[0] += sref * (p.c[0] - x0) / r3;
                g_re[1] += sref * (p.c[1] - x1) / r3;
                g_re[2] += sref * (-p.c[2] - x2) / r3;
            } else {
                // gradient_greens(Rankine, reflect(e1), e2; false) = (reflect(x) - xi)/r^3
                g_re[0] += sref * (x0 - p.c[0]) / r3;
                g_re[1] += sref * (x1 - p.c[1]) / r3;
                g_re[2] += sref * (-x2 - p.c[2]) / r3;
            }
        }
    }
    if (gf & kGfWaveMask) {
        cplx Sv, gxi[3], gx[3];
        const double area = panel_int ? p.area : 1.0;
        if (gf & GF_WU)
            wave_term<1>(k, x0 - p.c[0], x1 - p.c[1], x2 + p.c[2], area, Sv, gxi, gx);
        else if (gf & GF_DELHOMMEAU)
            wave_term<2>(k, x0 - p.c[0], x1 - p.c[1], x2 + p.c[2], area, Sv, gxi, gx);
        else
            wave_term<3>(k, x0 - p.c[0], x1 - p.c[1], x2 + p.c[2], area, Sv, gxi, gx, tab);
        val_re += Sv.re;
        val_im += Sv.im;
        for (int c = 0; c < 3; ++c) {
            g_re[c] += wrt_first ? gx[c].re : gxi[c].re;
            g_im[c] += wrt_first ? gx[c].im : gxi[c].im;
        }
    }
    if (grad) {
        for (int c = 0; c < 3; ++c) out[t + c * np] = make_double2(g_re[c], g_im[c]);
    } else {
        out[t] = make_double2(val_re, val_im);
    }
}

int launch_gf_eval(Context* ctx, int gf, int op, int wrt_first, int64_t np, const double* pts,
                   const mh_panel_t* pan, double k, double2* out) {
    if (np == 0) return MH_OK;
    DelhTab tab{};
    if (gf & GF_DELH_TABLE) TRY_RC(get_delh_table(ctx, &tab));
    hipLaunchKernelGGL(k_gf_eval, dim3((unsigned)((np + 127) / 128)), dim3(128), 0, ctx->stream, gf,
                       op, wrt_first, np, pts, pan, k, out, tab);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

__global__ void k_wave_term(int term, int64_t n, const double* __restrict__ r,
                            const double* __restrict__ z, double2* __restrict__ out, DelhTab tab) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    cplx G, dr, dz;
    if (term == GF_WU) {
        wu_dimless<true>(r[t], z[t], G, dr, dz);
    } else if (term == GF_DELHOMMEAU) {
        delhommeau_dimless(r[t], z[t], G, dr, dz);
    } else {
        double g[2], gr[2], gz[2];
        delh_table_eval(tab, r[t], z[t], g, gr, gz);
        G = {g[0], g[1]};
        dr = {gr[0], gr[1]};
        dz = {gz[0], gz[1]};
    }
    out[t] = make_double2(G.re, G.im);
    out[t + n] = make_double2(dr.re, dr.im);
    out[t + 2 * n] = make_double2(dz.re, dz.im);
}

int launch_wave_term(Context* ctx, int term, int64_t n, const double* r, const double* z,
                     double2* out) {
    if (term != GF_WU && term != GF_DELHOMMEAU && term != GF_DELH_TABLE) {
        set_error("wave_term: term must be MH_GF_WU, MH_GF_DELHOMMEAU or MH_GF_DELHOMMEAU_TABLE");
        return MH_ERR_ARGUMENT;
    }
    if (n == 0) return MH_OK;
    DelhTab tab{};
    if (term == GF_DELH_TABLE) TRY_RC(get_delh_table(ctx, &tab));
    hipLaunchKernelGGL(k_wave_term, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream,
                       term, n, r, z, out, tab);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// ---------------------------------------------------------------------------
// Table of the Delhommeau term (delh_table.hpp): built once per device from
// the exact FP64 term on first use and kept until the process ends.
// ---------------------------------------------------------------------------
__global__ void k_build_tab(double* __restrict__ lv0, double* __restrict__ lvs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n0 = (int64_t)kTabN0 * kTabN0;
    if (t >= tab_records()) return;
    double r, z, *rec;
    if (t < n0) {
        const int64_t i = t % kTabN0, j = t / kTabN0;
        r = (double)i * kTabH0;
        z = -(double)j * kTabH0;
        rec = lv0 + t * kTabRec;
    } else {
        const int64_t u = t - n0, per = (int64_t)kTabNn * kTabNn;
        const int l = (int)(u / per) + 1;
        const int64_t v = u % per, i = v % kTabNn, j = v / kTabNn;
        const double h = ldexp(kTabR1, 1 - l) / (double)(kTabNn - 1);
        r = (double)i * h;
        z = -(double)j * h;
        rec = lvs + u * kTabRec;
    }
    tab_node(r, z, rec);
}

static std::mutex g_tab_mutex;
static double* g_tab_mem[64] = {};

int get_delh_table(Context* ctx, DelhTab* out) {
    MH_REQUIRE(ctx->device >= 0 && ctx->device < 64, "Delhommeau table: device index beyond the supported range");
    std::lock_guard<std::mutex> lock(g_tab_mutex);
    double*& mem = g_tab_mem[ctx->device];
    if (!mem) {
        const int64_t nrec = tab_records();
        double* p = nullptr;
        MH_HIP_CHECK(hipMalloc(&p, sizeof(double) * kTabRec * (size_t)nrec));
        hipLaunchKernelGGL(k_build_tab, dim3((unsigned)((nrec + 127) / 128)), dim3(128), 0, ctx->stream, p,
                           p + (size_t)kTabN0 * kTabN0 * kTabRec);
        MH_HIP_CHECK(hipGetLastError());
        MH_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // other streams may read it at once
        mem = p;
    }
    out->lv0 = mem;
    out->lvs = mem + (size_t)kTabN0 * kTabN0 * kTabRec;
    return MH_OK;
}

}  // namespace mh
