// Batched dense complex LU (partial pivoting) and triangular solves, FP64.
//
// Right-looking blocked LU, one batch item per matrix (one frequency):
//   for each block column k0 (width NB):
//     k_lu_panel      factor the (n-k0) x NB panel, pivot = max |re|+|im|
//                     (LAPACK zgetf2 / izamax semantics, first index on ties)
//     k_lu_swap_trsm  apply the NB row interchanges to every other column and
//                     solve the unit-lower L11 U12 = A12 block row
//     k_lu_gemm       trailing update A22 -= L21 U12 (complex)
// Reference: the LAPACK factorisation behind `\` / implicit_linear
// (src/matrix_assembly.jl:108-126).
#include "mh_internal.hpp"

namespace mh {

constexpr int kNB = 32;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
}
// a - b*c
__device__ __forceinline__ double2 cfms(double2 a, double2 b, double2 c) {
    return make_double2(a.x - (b.x * c.x - b.y * c.y), a.y - (b.x * c.y + b.y * c.x));
}
__device__ __forceinline__ double2 crecip(double2 a) {
    // Smith's algorithm for 1/a
    if (fabs(a.x) >= fabs(a.y)) {
        const double r = a.y / a.x, d = a.x + a.y * r;
        return make_double2(1.0 / d, -r / d);
    }
    const double r = a.x / a.y, d = a.y + a.x * r;
    return make_double2(r / d, -1.0 / d);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) { return cmul(a, crecip(b)); }

// ---------------------------------------------------------------------------
// panel factorisation: one 256-thread workgroup per batch item
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lu_panel(int64_t n, double2* __restrict__ A, int64_t ld,
                                                  int64_t stride, int64_t k0, int nb,
                                                  int64_t* __restrict__ ipiv,
                                                  int* __restrict__ info) {
    double2* M = A + (int64_t)blockIdx.x * stride;
    int64_t* piv = ipiv + (int64_t)blockIdx.x * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double s_v[4];
    __shared__ int64_t s_i[4];
    __shared__ double2 s_row[kNB];
    __shared__ int64_t s_p;
    __shared__ int s_zero;
    for (int c = 0; c < nb; ++c) {
        const int64_t col = k0 + c;
        double best = -1.0;
        int64_t bi = n;
        for (int64_t r = col + tid; r < n; r += 256) {
            const double2 a = M[r + col * ld];
            const double v = fabs(a.x) + fabs(a.y);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(best, off);
            const int64_t oi = __shfl_xor(bi, off);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            s_v[wave] = best;
            s_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double b = s_v[0];
            int64_t p = s_i[0];
            for (int w = 1; w < 4; ++w)
                if (s_v[w] > b || (s_v[w] == b && s_i[w] < p)) {
                    b = s_v[w];
                    p = s_i[w];
                }
            if (b <= 0.0) {  // exact zero pivot: record, no swap, no update
                p = col;
                s_zero = 1;
                atomicOr(info + blockIdx.x, 1);
            } else {
                s_zero = 0;
            }
            s_p = p;
            piv[col] = p;
        }
        __syncthreads();
        const int64_t p = s_p;
        if (tid < nb && p != col) {
            const int64_t cj = (k0 + tid) * ld;
            const double2 t = M[col + cj];
            M[col + cj] = M[p + cj];
            M[p + cj] = t;
        }
        __syncthreads();
        if (tid < nb) s_row[tid] = M[col + (k0 + tid) * ld];
        __syncthreads();
        if (!s_zero) {
            const double2 inv = crecip(s_row[c]);
            for (int64_t r = col + 1 + tid; r < n; r += 256) {
                const double2 l = cmul(M[r + col * ld], inv);
                M[r + col * ld] = l;
                for (int j = c + 1; j < nb; ++j) {
                    const int64_t o = r + (k0 + j) * ld;
                    M[o] = cfms(M[o], l, s_row[j]);
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// row interchanges on all non-panel columns + L11 solve on the right block row
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lu_swap_trsm(int64_t n, int64_t ncols,
                                                      double2* __restrict__ A, int64_t ld,
                                                      int64_t stride, int64_t k0, int nb,
                                                      const int64_t* __restrict__ ipiv) {
    double2* M = A + (int64_t)blockIdx.y * stride;
    const int64_t* piv = ipiv + (int64_t)blockIdx.y * n;
    __shared__ double2 sL[kNB * kNB];  // sL[c*nb + r] = L11(r, c)
    __shared__ int64_t sp[kNB];
    const int tid = threadIdx.x;
    for (int q = tid; q < nb * nb; q += 256) {
        const int r = q % nb, c = q / nb;
        sL[q] = M[(k0 + r) + (k0 + c) * ld];
    }
    if (tid < nb) sp[tid] = piv[k0 + tid];
    __syncthreads();
    const int64_t j = (int64_t)blockIdx.x * 256 + tid;
    if (j >= ncols || (j >= k0 && j < k0 + nb)) return;
    double2* col = M + j * ld;
    for (int c = 0; c < nb; ++c) {
        const int64_t p = sp[c];
        if (p != k0 + c) {
            const double2 t = col[k0 + c];
            col[k0 + c] = col[p];
            col[p] = t;
        }
    }
    if (j < k0 + nb) return;  // left of the panel: interchanges only
    double2 x[kNB];
#pragma unroll
    for (int r = 0; r < kNB; ++r)
        if (r < nb) x[r] = col[k0 + r];
#pragma unroll
    for (int c = 0; c < kNB; ++c) {
#pragma unroll
        for (int r = c + 1; r < kNB; ++r)
            if (r < nb) x[r] = cfms(x[r], sL[c * nb + r], x[c]);
    }
#pragma unroll
    for (int r = 0; r < kNB; ++r)
        if (r < nb) col[k0 + r] = x[r];
}

// ---------------------------------------------------------------------------
// trailing update C -= A B, C = M[r0:n, c0:ncols], A = M[r0:n, k0:k0+nb],
// B = M[k0:k0+nb, c0:ncols]; 64 x 64 complex tile per 256-thread workgroup,
// 4 x 4 complex outputs per thread.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lu_gemm(int64_t n, int64_t ncols, double2* __restrict__ A,
                                                 int64_t ld, int64_t stride, int64_t k0, int nb) {
    double2* M = A + (int64_t)blockIdx.z * stride;
    const int64_t r0 = k0 + nb, c0 = k0 + nb;
    const int64_t tr = r0 + (int64_t)blockIdx.x * 64;
    const int64_t tc = c0 + (int64_t)blockIdx.y * 64;
    __shared__ double2 sA[kNB][64];  // sA[k][row]
    __shared__ double2 sB[kNB][64];  // sB[k][col]
    const int tid = threadIdx.x;
    for (int q = tid; q < nb * 64; q += 256) {
        const int r = q & 63, k = q >> 6;
        const int64_t gr = tr + r;
        sA[k][r] = (gr < n) ? M[gr + (k0 + k) * ld] : make_double2(0.0, 0.0);
        const int cc = q / nb, kk = q % nb;  // B: walk down each column
        const int64_t gc = tc + cc;
        sB[kk][cc] = (gc < ncols) ? M[(k0 + kk) + gc * ld] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = make_double2(0.0, 0.0);
    for (int k = 0; k < nb; ++k) {
        double2 av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) av[a] = sA[k][tx + 16 * a];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = sB[k][ty + 16 * b];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                acc[a][b].x = fma(av[a].x, bv[b].x, acc[a][b].x);
                acc[a][b].x = fma(-av[a].y, bv[b].y, acc[a][b].x);
                acc[a][b].y = fma(av[a].x, bv[b].y, acc[a][b].y);
                acc[a][b].y = fma(av[a].y, bv[b].x, acc[a][b].y);
            }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int64_t gc = tc + ty + 16 * b;
        if (gc >= ncols) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int64_t gr = tr + tx + 16 * a;
            if (gr >= n) continue;
            const int64_t o = gr + gc * ld;
            M[o] = csub(M[o], acc[a][b]);
        }
    }
}

int lu_factor(Context* ctx, int64_t n, int64_t ncols, int64_t batch, double2* A, int64_t ld,
              int64_t stride, int64_t* ipiv, int* d_info) {
    if (n == 0 || batch == 0) return MH_OK;
    MH_REQUIRE(batch <= 65535, "lu: batch too large for one launch");
    for (int64_t k0 = 0; k0 < n; k0 += kNB) {
        const int nb = (int)((n - k0) < kNB ? (n - k0) : kNB);
        const double m_ = (double)(n - k0);
        const double rem = (double)(ncols - k0 - nb);
        {
            ProfScope ps(ctx, "lu_panel", 8.0 * m_ * nb * nb / 2.0 * batch, 32.0 * m_ * nb * batch);
        hipLaunchKernelGGL(k_lu_panel, dim3((unsigned)batch), dim3(256), 0, ctx->stream, n, A, ld,
                           stride, k0, nb, ipiv, d_info);
        }
        {
            ProfScope ps(ctx, "lu_swap_trsm", 4.0 * nb * nb * rem * batch,
                         16.0 * 2.0 * nb * (double)ncols * batch);
        hipLaunchKernelGGL(k_lu_swap_trsm, dim3((unsigned)((ncols + 255) / 256), (unsigned)batch),
                           dim3(256), 0, ctx->stream, n, ncols, A, ld, stride, k0, nb, ipiv);
        }
        const int64_t rem_r = n - k0 - nb, rem_c = ncols - k0 - nb;
        if (rem_r > 0 && rem_c > 0) {
            ProfScope ps(ctx, "lu_gemm", 8.0 * (double)rem_r * (double)rem_c * nb * batch,
                         16.0 * (2.0 * rem_r * rem_c + nb * (double)(rem_r + rem_c)) * batch);
            hipLaunchKernelGGL(k_lu_gemm,
                               dim3((unsigned)((rem_r + 63) / 64), (unsigned)((rem_c + 63) / 64),
                                    (unsigned)batch),
                               dim3(256), 0, ctx->stream, n, ncols, A, ld, stride, k0, nb);
        }
    }
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// ---------------------------------------------------------------------------
// triangular solves with the LU factors, one 256-thread workgroup per
// (batch item, right-hand side); the vector lives in LDS.
// ---------------------------------------------------------------------------
constexpr int kMaxSolveN = 4096;  // vector held in 64 KiB of LDS

__global__ __launch_bounds__(256) void k_lu_solve(int64_t n, const double2* __restrict__ A,
                                                  int64_t ld, int64_t stride,
                                                  const int64_t* __restrict__ ipiv, int trans,
                                                  double2* __restrict__ B, int64_t ldb,
                                                  int64_t bstride) {
    extern __shared__ double2 sb[];
    const int64_t item = blockIdx.y;
    const double2* M = A + item * stride;
    const int64_t* piv = ipiv + item * n;
    double2* b = B + item * bstride + (int64_t)blockIdx.x * ldb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int64_t i = tid; i < n; i += 256) sb[i] = b[i];
    __syncthreads();
    if (!trans) {
        if (tid == 0)
            for (int64_t k = 0; k < n; ++k) {
                const int64_t p = piv[k];
                if (p != k) {
                    const double2 t = sb[k];
                    sb[k] = sb[p];
                    sb[p] = t;
                }
            }
        __syncthreads();
        // L y = b (unit lower), 64-column blocks
        for (int64_t j0 = 0; j0 < n; j0 += 64) {
            const int64_t jn = (n - j0) < 64 ? (n - j0) : 64;
            if (wave == 0) {
                double2 v = (lane < jn) ? sb[j0 + lane] : make_double2(0.0, 0.0);
                for (int c = 0; c < jn; ++c) {
                    const double2 yc = make_double2(__shfl(v.x, c), __shfl(v.y, c));
                    if (lane > c && lane < jn) v = cfms(v, M[(j0 + lane) + (j0 + c) * ld], yc);
                }
                if (lane < jn) sb[j0 + lane] = v;
            }
            __syncthreads();
            for (int64_t i = j0 + jn + tid; i < n; i += 256) {
                double2 v = sb[i];
                for (int c = 0; c < jn; ++c) v = cfms(v, M[i + (j0 + c) * ld], sb[j0 + c]);
                sb[i] = v;
            }
            __syncthreads();
        }
        // U x = y (upper), blocks from the bottom
        for (int64_t j1 = n; j1 > 0; j1 -= 64) {
            const int64_t j0 = (j1 - 64) > 0 ? (j1 - 64) : 0;
            const int64_t jn = j1 - j0;
            if (wave == 0) {
                double2 v = (lane < jn) ? sb[j0 + lane] : make_double2(0.0, 0.0);
                for (int c = (int)jn - 1; c >= 0; --c) {
                    if (lane == c) v = cdiv(v, M[(j0 + c) + (j0 + c) * ld]);
                    const double2 xc = make_double2(__shfl(v.x, c), __shfl(v.y, c));
                    if (lane < c) v = cfms(v, M[(j0 + lane) + (j0 + c) * ld], xc);
                }
                if (lane < jn) sb[j0 + lane] = v;
            }
            __syncthreads();
            for (int64_t i = tid; i < j0; i += 256) {
                double2 v = sb[i];
                for (int c = 0; c < jn; ++c) v = cfms(v, M[i + (j0 + c) * ld], sb[j0 + c]);
                sb[i] = v;
            }
            __syncthreads();
        }
    } else {
        // U^T y = b (lower, non-unit): dot-product form over columns of U
        for (int64_t i0 = 0; i0 < n; i0 += 64) {
            const int64_t in = (n - i0) < 64 ? (n - i0) : 64;
            // b[i] -= sum_{j < i0} U[j, i] y[j]; one wave per row, lanes over j
            for (int64_t r = wave; r < in; r += 4) {
                const int64_t i = i0 + r;
                double sx = 0.0, sy = 0.0;
                for (int64_t j = lane; j < i0; j += 64) {
                    const double2 u = M[j + i * ld], y = sb[j];
                    sx += u.x * y.x - u.y * y.y;
                    sy += u.x * y.y + u.y * y.x;
                }
                for (int off = 32; off > 0; off >>= 1) {
                    sx += __shfl_xor(sx, off);
                    sy += __shfl_xor(sy, off);
                }
                if (lane == 0) sb[i] = make_double2(sb[i].x - sx, sb[i].y - sy);
            }
            __syncthreads();
            if (wave == 0) {
                double2 v = (lane < in) ? sb[i0 + lane] : make_double2(0.0, 0.0);
                for (int c = 0; c < in; ++c) {
                    if (lane == c) v = cdiv(v, M[(i0 + c) + (i0 + c) * ld]);
                    const double2 yc = make_double2(__shfl(v.x, c), __shfl(v.y, c));
                    if (lane > c && lane < in) v = cfms(v, M[(i0 + c) + (i0 + lane) * ld], yc);
                }
                if (lane < in) sb[i0 + lane] = v;
            }
            __syncthreads();
        }
        // L^T z = y (upper, unit): from the bottom, dot over columns of L
        for (int64_t i1 = n; i1 > 0; i1 -= 64) {
            const int64_t i0 = (i1 - 64) > 0 ? (i1 - 64) : 0;
            const int64_t in = i1 - i0;
            for (int64_t r = wave; r < in; r += 4) {
                const int64_t i = i0 + r;
                double sx = 0.0, sy = 0.0;
                for (int64_t j = i1 + lane; j < n; j += 64) {
                    const double2 l = M[j + i * ld], z = sb[j];
                    sx += l.x * z.x - l.y * z.y;
                    sy += l.x * z.y + l.y * z.x;
                }
                for (int off = 32; off > 0; off >>= 1) {
                    sx += __shfl_xor(sx, off);
                    sy += __shfl_xor(sy, off);
                }
                if (lane == 0) sb[i] = make_double2(sb[i].x - sx, sb[i].y - sy);
            }
            __syncthreads();
            if (wave == 0) {
                double2 v = (lane < in) ? sb[i0 + lane] : make_double2(0.0, 0.0);
                for (int c = (int)in - 1; c >= 0; --c) {
                    const double2 zc = make_double2(__shfl(v.x, c), __shfl(v.y, c));
                    if (lane < c) v = cfms(v, M[(i0 + c) + (i0 + lane) * ld], zc);
                }
                if (lane < in) sb[i0 + lane] = v;
            }
            __syncthreads();
        }
        if (tid == 0)
            for (int64_t k = n - 1; k >= 0; --k) {
                const int64_t p = piv[k];
                if (p != k) {
                    const double2 t = sb[k];
                    sb[k] = sb[p];
                    sb[p] = t;
                }
            }
        __syncthreads();
    }
    for (int64_t i = tid; i < n; i += 256) b[i] = sb[i];
}

int lu_solve(Context* ctx, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
             const int64_t* ipiv, int trans, int64_t nrhs, double2* B, int64_t ldb,
             int64_t bstride) {
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(n <= kMaxSolveN, "lu_solve: n larger than 4096 is not supported");
    MH_REQUIRE(batch <= 65535 && nrhs <= 65535, "lu_solve: batch or nrhs too large");
    const size_t lds = sizeof(double2) * (size_t)n;
    ProfScope ps(ctx, "lu_solve", 8.0 * (double)n * n * nrhs * batch, 16.0 * (double)n * n * nrhs * batch);
    hipLaunchKernelGGL(k_lu_solve, dim3((unsigned)nrhs, (unsigned)batch), dim3(256), lds,
                       ctx->stream, n, LU, ld, stride, ipiv, trans, B, ldb, bstride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// ---------------------------------------------------------------------------
// Y = S X for small nrhs: one thread per row, S read once, coalesced
// ---------------------------------------------------------------------------
constexpr int kMaxRhs = 16;

__global__ __launch_bounds__(256) void k_zgemv_multi(int64_t n, int64_t nrhs,
                                                     const double2* __restrict__ S, int64_t ld,
                                                     int64_t stride, const double2* __restrict__ X,
                                                     int64_t ldx, int64_t xstride,
                                                     double2* __restrict__ Y, int64_t ldy,
                                                     int64_t ystride) {
    const int64_t item = blockIdx.y;
    const double2* Sm = S + item * stride;
    const double2* Xm = X + item * xstride;
    double2* Ym = Y + item * ystride;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    __shared__ double2 sx[64][kMaxRhs];
    double2 acc[kMaxRhs];
#pragma unroll
    for (int r = 0; r < kMaxRhs; ++r) acc[r] = make_double2(0.0, 0.0);
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
        const int64_t jn = (n - j0) < 64 ? (n - j0) : 64;
        __syncthreads();
        for (int q = threadIdx.x; q < 64 * kMaxRhs; q += 256) {
            const int jj = q % 64, r = q / 64;
            sx[jj][r] = (jj < jn && r < nrhs) ? Xm[(j0 + jj) + r * ldx] : make_double2(0.0, 0.0);
        }
        __syncthreads();
        if (i < n) {
            for (int jj = 0; jj < jn; ++jj) {
                const double2 s = Sm[i + (j0 + jj) * ld];
#pragma unroll
                for (int r = 0; r < kMaxRhs; ++r) {
                    const double2 x = sx[jj][r];
                    acc[r].x = fma(s.x, x.x, acc[r].x);
                    acc[r].x = fma(-s.y, x.y, acc[r].x);
                    acc[r].y = fma(s.x, x.y, acc[r].y);
                    acc[r].y = fma(s.y, x.x, acc[r].y);
                }
            }
        }
    }
    if (i < n)
        for (int r = 0; r < nrhs; ++r) Ym[i + r * ldy] = acc[r];
}

int zgemm_nn_batched(Context* ctx, int64_t n, int64_t nrhs, int64_t batch, const double2* S,
                     int64_t ld, int64_t stride, const double2* X, int64_t ldx, int64_t xstride,
                     double2* Y, int64_t ldy, int64_t ystride) {
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(nrhs <= kMaxRhs, "matvec: at most 16 right-hand sides per call");
    ProfScope ps(ctx, "zgemv", 8.0 * (double)n * n * nrhs * batch, 16.0 * (double)n * n * batch);
    hipLaunchKernelGGL(k_zgemv_multi, dim3((unsigned)((n + 255) / 256), (unsigned)batch), dim3(256),
                       0, ctx->stream, n, nrhs, S, ld, stride, X, ldx, xstride, Y, ldy, ystride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

}  // namespace mh
