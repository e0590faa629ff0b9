// Batched triangular solves with the LU factors, pipelined across many
// workgroups per matrix.
//
// A solve with one n x n triangle and a few right-hand sides reads the whole
// triangle once; one workgroup per matrix would leave most of the chip idle
// and be bound by one CU's bandwidth.  Here every matrix is cut into slabs of
// 256 rows, one workgroup per slab (one row per thread).  The unknowns are
// produced in 32-row blocks in order.  A workgroup keeps its rows' partial
// sums in registers; it applies the blocks of earlier slabs as their
// solutions are published (a per-block flag written by the block's owner),
// several ready blocks per round, with its rows' slice of the next block
// loaded ahead of time.
// For its own blocks it multiplies the block's residual by the inverse of
// the 32 x 32 diagonal block (precomputed once per factorisation, so the
// dependency chain carries no sequential substitution), publishes the
// solution and the block's flag, and updates its remaining rows.
//
// The four triangles of A = P L U are one kernel in "logical" coordinates.
// With np = n rounded up to a multiple of 32, unknown i of the logical
// lower-triangular system sits at physical row p(i) = i (forward) or
// np-1-i (backward; rows >= n are padding with zero right-hand side), so
// logical blocks are aligned with the physical 32-row blocks, and T[i][j] is
// M[p(i), p(j)] (L, U as stored) or M[p(j), p(i)] (L^T, U^T):
//     L y = P b              forward,  L^-1 blocks, gather b through perm
//     U x = y                backward, U^-1 blocks
//     U^T y = b              forward,  U^-1 blocks transposed
//     L^T z = y, x = P^T z   backward, L^-1 blocks transposed, scatter
//
// Synchronisation: solution entries and flags are written and read with
// device-scope (agent) atomic accesses, which are coherent across the XCDs'
// L2 caches; a producer waits for its solution stores to complete before it
// stores the flag.  Workgroups take their (slab, matrix) from an atomic
// ticket in the order they start, so every workgroup a running one waits for
// has started before it.  Waits are bounded; on timeout an error word is set
// and the kernel runs to completion (the host reports the error).
// Reference: the LAPACK triangular solves behind `\` / implicit_linear
// (src/matrix_assembly.jl:108-126).
#include "mh_cx.hpp"
#include "mh_internal.hpp"

namespace mh {

constexpr int kTsRows = 256;  // rows per workgroup, one per thread
constexpr int kTsCB = 32;     // block of unknowns = diagonal block
constexpr int kTsNR = 4;      // right-hand sides per launch
constexpr int kTsOwn = kTsRows / kTsCB;
constexpr int kTsBatch = 8;   // ready blocks of earlier slabs applied per round
constexpr unsigned kSpinLimit = 1u << 22;

// Thread 0 waits until *flag == epoch (or an error is flagged), then the
// workgroup proceeds.
__device__ __forceinline__ void wait_flag(const unsigned* flag, unsigned epoch, int* err) {
    if (threadIdx.x == 0) {
        unsigned spins = 0;
        unsigned* f = const_cast<unsigned*>(flag);
        while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++spins > kSpinLimit) {
                atomicOr(err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

template <class C2>
__device__ __forceinline__ C2 load_published(const C2* p) {
    typedef decltype(C2().x) T;
    T* q = reinterpret_cast<T*>(const_cast<C2*>(p));
    return mkc<C2>(__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                   __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
template <class C2>
__device__ __forceinline__ void store_published(C2* p, C2 v) {
    typedef decltype(C2().x) T;
    T* q = reinterpret_cast<T*>(p);
    __hip_atomic_store(q, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T, bool TR, bool REV, int NR>
__global__ __launch_bounds__(kTsRows) void k_trsv(
    int64_t n, int64_t batch, const typename Cx<T>::t* __restrict__ A, int64_t ld, int64_t stride,
    const typename Cx<T>::t* __restrict__ Inv, const int* __restrict__ gperm,
    const int* __restrict__ sperm, int nrhs, const typename Cx<T>::t* in, int64_t ldi, int64_t istride,
    typename Cx<T>::t* out, int64_t ldo, int64_t ostride, int nslab, unsigned* __restrict__ flags,
    int nsubp, unsigned epoch, unsigned long long* __restrict__ ticket, int* __restrict__ err) {
    typedef typename Cx<T>::t C2;
    __shared__ unsigned long long s_tk;
    __shared__ int s_ok[kTsBatch];
    __shared__ C2 sI[2][kTsCB][kTsCB + 1];  // inverse diagonal blocks, double-buffered
    __shared__ C2 sx[kTsBatch][kTsCB][NR];  // solutions of the blocks being applied
    __shared__ C2 sr[kTsCB][NR];            // residual of the current own block
    const int tid = threadIdx.x, wave = tid >> 6;
    if (tid == 0) s_tk = atomicAdd(ticket, 1ull);
    __syncthreads();
    const int64_t tk = (int64_t)s_tk;
    if (tk >= (int64_t)nslab * batch) return;  // cannot happen with a reset ticket
    const int slab = (int)(tk % nslab);
    const int64_t item = tk / nslab;
    const int64_t np = (int64_t)nsubp * kTsCB;
    const C2* M = A + item * stride;
    const int* gp = gperm ? gperm + item * n : nullptr;
    const int* sp = sperm ? sperm + item * n : nullptr;
    const C2* bin = in + item * istride;
    C2* bout = out + item * ostride;
    unsigned* fl = flags + item * nsubp;
    const C2 zero = mkc<C2>(T(0), T(0));
    auto phys = [&](int64_t i) -> int64_t { return REV ? np - 1 - i : i; };

    const int64_t i0 = (int64_t)slab * kTsRows;
    const int64_t i = i0 + tid;
    const int64_t pi = phys(i);
    const bool valid = i < np && pi < n;
    const int cown = (int)(i0 / kTsCB);
    const int cend = (cown + kTsOwn) < nsubp ? cown + kTsOwn : nsubp;

    // inverse of own block s (physical block K, element (a, b) at
    // Inv[((item * nsubp + K) * 32 + b) * 32 + a]) into buffer s & 1
    auto load_inv = [&](int s) {
        const int c = cown + s;
        const int64_t K = REV ? nsubp - 1 - c : c;
        const C2* src = Inv + (item * nsubp + K) * kTsCB * kTsCB;
        for (int e = tid; e < kTsCB * kTsCB; e += kTsRows) sI[s & 1][e % kTsCB][e / kTsCB] = src[e];
    };
    if (cend > cown) load_inv(0);

    C2 v[NR];
    {
        const int64_t src = valid ? (gp ? (int64_t)gp[pi] : pi) : 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) v[r] = (valid && r < nrhs) ? bin[src + r * ldi] : zero;
    }

    // this row's slice of block c (logical columns 32c .. 32c+31), for rows
    // after the block, and its application with the block's solution x
    C2 m[kTsCB];
    auto prefetch = [&](int c) {
        const bool need = valid && i >= (int64_t)(c + 1) * kTsCB;
#pragma unroll
        for (int q = 0; q < kTsCB; ++q) {
            const int64_t pj = phys((int64_t)c * kTsCB + q);
            m[q] = (need && pj < n) ? (TR ? M[pj + pi * ld] : M[pi + pj * ld]) : zero;
        }
    };
    auto apply = [&](int c, C2 (*x)[NR]) {
        if (valid && i >= (int64_t)(c + 1) * kTsCB) {
#pragma unroll
            for (int q = 0; q < kTsCB; ++q)
#pragma unroll
                for (int r = 0; r < NR; ++r)
                    if (r < nrhs) v[r] = cfms(v[r], m[q], x[q][r]);
        }
    };
    prefetch(0);

    // 1. blocks of earlier slabs, applied in rounds of up to kTsBatch blocks
    //    that are ready
    int c = 0;
    while (c < cown) {
        wait_flag(fl + c, epoch, err);
        if (tid < kTsBatch) {
            bool ok = tid == 0;
            if (tid > 0 && c + tid < cown)
                ok = __hip_atomic_load(fl + c + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch;
            s_ok[tid] = ok ? 1 : 0;
        }
        __syncthreads();
        int nb = 1;
        while (nb < kTsBatch && s_ok[nb]) ++nb;
        for (int e = tid; e < nb * kTsCB * NR; e += kTsRows) {
            const int b = e / (kTsCB * NR), q = e % kTsCB, r = (e / kTsCB) % NR;
            const int64_t pj = phys((int64_t)(c + b) * kTsCB + q);
            sx[b][q][r] = (r < nrhs && pj < n) ? load_published(bout + (sp ? (int64_t)sp[pj] : pj) + r * ldo)
                                               : zero;
        }
        __syncthreads();
        apply(c, sx[0]);
        for (int b = 1; b < nb; ++b) {
            prefetch(c + b);
            apply(c + b, sx[b]);
        }
        c += nb;
        __syncthreads();
        prefetch(c);
    }

    // 2. this slab's own blocks: x = T_cc^-1 r, published, then applied to
    //    the rows after the block
    for (int s = 0; s < cend - cown; ++s) {
        const int cb = cown + s, base = s * kTsCB;
        const bool mine = tid >= base && tid < base + kTsCB;
        if (mine) {
#pragma unroll
            for (int r = 0; r < NR; ++r) sr[tid - base][r] = v[r];
        }
        if (s + 1 < cend - cown) load_inv(s + 1);
        __syncthreads();
        if (mine) {
            const int a = tid - base;
            const int pa = REV ? kTsCB - 1 - a : a;
            C2 acc[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = zero;
#pragma unroll 8
            for (int b = 0; b < kTsCB; ++b) {
                const int pb = REV ? kTsCB - 1 - b : b;
                const C2 w = TR ? sI[s & 1][pb][pa] : sI[s & 1][pa][pb];
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const C2 y = sr[b][r];
                    acc[r].x += w.x * y.x - w.y * y.y;
                    acc[r].y += w.x * y.y + w.y * y.x;
                }
            }
            const int64_t dst = sp ? (valid ? (int64_t)sp[pi] : 0) : pi;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                v[r] = acc[r];
                sx[0][a][r] = acc[r];
                if (valid && r < nrhs) store_published(bout + dst + r * ldo, acc[r]);
            }
        }
        if (wave == (base >> 6)) __builtin_amdgcn_s_waitcnt(0);  // solution stores complete
        __syncthreads();
        if (tid == 0) __hip_atomic_store(fl + cb, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        apply(cb, sx[0]);
        __syncthreads();
        if (s + 1 < cend - cown) prefetch(cb + 1);
    }
}

// Inverses of the 32 x 32 diagonal blocks of the unit-lower L and of U
// (physical blocks K = 0 .. nsubp-1; a last partial block is padded with the
// identity).  One wave per block: lane b < 32 builds column b of L^-1 by
// forward substitution, lane 32 + b column b of U^-1 by back substitution.
template <class T>
__global__ __launch_bounds__(64) void k_tri_inv(int64_t n, int64_t batch, const typename Cx<T>::t* __restrict__ A,
                                                int64_t ld, int64_t stride, int nsubp,
                                                typename Cx<T>::t* __restrict__ inv) {
    typedef typename Cx<T>::t C2;
    __shared__ C2 sB[kTsCB][kTsCB + 1];
    __shared__ C2 sY[2][kTsCB][kTsCB + 1];  // columns of L^-1 and U^-1
    const int K = blockIdx.x;
    const int64_t item = blockIdx.y;
    const int lane = threadIdx.x;
    const C2* M = A + item * stride;
    const int64_t r0 = (int64_t)K * kTsCB;
    const int jn = (int)((n - r0) < kTsCB ? (n - r0) : kTsCB);
    for (int e = lane; e < kTsCB * kTsCB; e += 64) {
        const int a = e % kTsCB, b = e / kTsCB;
        sB[a][b] = (a < jn && b < jn) ? M[(r0 + a) + (r0 + b) * ld] : mkc<C2>(T(a == b ? 1 : 0), T(0));
    }
    __syncthreads();
    const int b = lane & 31, u = lane >> 5;
    if (u == 0) {
        for (int a = 0; a < kTsCB; ++a) {  // L y = e_b, unit diagonal
            C2 acc = mkc<C2>(T(a == b ? 1 : 0), T(0));
            for (int c = 0; c < a; ++c) acc = cfms(acc, sB[a][c], sY[0][c][b]);
            sY[0][a][b] = acc;
        }
    } else {
        for (int a = kTsCB - 1; a >= 0; --a) {  // U y = e_b
            C2 acc = mkc<C2>(T(a == b ? 1 : 0), T(0));
            for (int c = a + 1; c < kTsCB; ++c) acc = cfms(acc, sB[a][c], sY[1][c][b]);
            sY[1][a][b] = cdiv(acc, sB[a][a]);
        }
    }
    C2* dst = inv + ((u ? batch : 0) + item) * (int64_t)nsubp * kTsCB * kTsCB + (int64_t)K * kTsCB * kTsCB +
              (int64_t)b * kTsCB;
    for (int a = 0; a < kTsCB; ++a) dst[a] = sY[u][a][b];
}

// net row permutation of the interchanges ipiv (0-based, applied in order):
// (P b)[i] = b[perm[i]]
__global__ __launch_bounds__(256) void k_perm_from_ipiv(int64_t n, const int64_t* __restrict__ ipiv,
                                                        int* __restrict__ perm, int* __restrict__ err) {
    __shared__ int s[kMaxSolveN];
    __shared__ int sp[kMaxSolveN];
    const int64_t item = blockIdx.x;
    const int64_t* pv = ipiv + item * n;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
        s[k] = (int)k;
        const int64_t p = pv[k];
        sp[k] = (p < k || p >= n) ? -1 : (int)p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < (int)n; ++k) {
            const int p = sp[k];
            if (p < 0) {
                atomicOr(err, 2);
                continue;
            }
            if (p != k) {
                const int t = s[k];
                s[k] = s[p];
                s[p] = t;
            }
        }
    }
    __syncthreads();
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) perm[item * n + k] = s[k];
}

__global__ void k_reset_ticket(unsigned long long* t) { *t = 0ull; }

template <class C2>
__global__ void k_copy_cols(int64_t rows, int64_t cols, int64_t batch, const C2* __restrict__ src,
                            int64_t lds, int64_t sstride, C2* __restrict__ dst, int64_t ldd, int64_t dstride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = rows * cols * batch;
    if (t >= tot) return;
    const int64_t r = t % rows, c = (t / rows) % cols, b = t / (rows * cols);
    dst[b * dstride + c * ldd + r] = src[b * sstride + c * lds + r];
}

// per-context synchronisation state: [ticket u64 | pad | flags u32 ...], the
// error word and a solution workspace for the out-of-place steps
static int tri_state(Context* c, int64_t batch, int64_t nsub, size_t ws_bytes) {
    Context::TriSlot& t = c->tri[c->tri_slot];
    const size_t need = 256 + sizeof(unsigned) * (size_t)(batch * nsub);
    if (need > t.sync_bytes) {
        if (t.sync) {
            MH_HIP_CHECK(hipStreamSynchronize(c->stream));
            MH_HIP_CHECK(hipFree(t.sync));
            t.sync = nullptr;
            t.sync_bytes = 0;
        }
        MH_HIP_CHECK(hipMalloc(&t.sync, need));
        MH_HIP_CHECK(hipMemset(t.sync, 0, need));
        t.sync_bytes = need;
    }
    if (ws_bytes > t.ws_bytes) {
        if (t.ws) {
            MH_HIP_CHECK(hipStreamSynchronize(c->stream));
            MH_HIP_CHECK(hipFree(t.ws));
            t.ws = nullptr;
            t.ws_bytes = 0;
        }
        MH_HIP_CHECK(hipMalloc(&t.ws, ws_bytes));
        t.ws_bytes = ws_bytes;
    }
    if (!c->tri_err) {
        MH_HIP_CHECK(hipMalloc(&c->tri_err, sizeof(int)));
        MH_HIP_CHECK(hipMemset(c->tri_err, 0, sizeof(int)));
    }
    return MH_OK;
}

int64_t tri_inv_elems(int64_t n, int64_t batch) {
    const int64_t nsubp = (n + kTsCB - 1) / kTsCB;
    return 2 * batch * nsubp * kTsCB * kTsCB;
}

template <class T>
static int tri_prepare_t(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* LU, int64_t ld,
                         int64_t stride, const int64_t* ipiv, int* perm, typename Cx<T>::t* inv) {
    if (n == 0 || batch == 0) return MH_OK;
    MH_REQUIRE(n <= kMaxSolveN, "lu_solve: n larger than 4096 is not supported");
    MH_REQUIRE(batch <= 65535, "lu_solve: batch too large");
    TRY_RC(tri_state(c, 1, 1, 0));
    if (perm) {
        hipLaunchKernelGGL(k_perm_from_ipiv, dim3((unsigned)batch), dim3(256), 0, c->stream, n, ipiv, perm,
                           c->tri_err);
        MH_HIP_CHECK(hipGetLastError());
    }
    const int nsubp = (int)((n + kTsCB - 1) / kTsCB);
    ProfScope ps(c, sizeof(T) == 8 ? "tri_inv" : "tri_inv_f32", 0.0, 0.0);
    hipLaunchKernelGGL(k_tri_inv<T>, dim3((unsigned)nsubp, (unsigned)batch), dim3(64), 0, c->stream, n, batch,
                       LU, ld, stride, nsubp, inv);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int tri_prepare(Context* c, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                const int64_t* ipiv, int* perm, double2* inv) {
    return tri_prepare_t<double>(c, n, batch, LU, ld, stride, ipiv, perm, inv);
}
int tri_prepare_f32(Context* c, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                    const int64_t* ipiv, int* perm, float2* inv) {
    return tri_prepare_t<float>(c, n, batch, LU, ld, stride, ipiv, perm, inv);
}

template <class T, bool TR, bool REV>
static int launch_trsv(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* A, int64_t ld,
                       int64_t stride, const typename Cx<T>::t* inv, const int* gperm, const int* sperm,
                       int nrhs, const typename Cx<T>::t* in, int64_t ldi, int64_t istride,
                       typename Cx<T>::t* out, int64_t ldo, int64_t ostride) {
    const int nsubp = (int)((n + kTsCB - 1) / kTsCB);
    const int nslab = (int)((nsubp * (int64_t)kTsCB + kTsRows - 1) / kTsRows);
    Context::TriSlot& ts = c->tri[c->tri_slot];
    unsigned long long* ticket = reinterpret_cast<unsigned long long*>(ts.sync);
    unsigned* flags = reinterpret_cast<unsigned*>(static_cast<char*>(ts.sync) + 256);
    if (++ts.epoch == 0) ts.epoch = 1;
    hipLaunchKernelGGL(k_reset_ticket, dim3(1), dim3(1), 0, c->stream, ticket);
#define MH_TRSV(NRR)                                                                                         \
    hipLaunchKernelGGL((k_trsv<T, TR, REV, NRR>), dim3((unsigned)(nslab * batch)), dim3(kTsRows), 0, c->stream, \
                       n, batch, A, ld, stride, inv, gperm, sperm, nrhs, in, ldi, istride, out, ldo, ostride,    \
                       nslab, flags, nsubp, ts.epoch, ticket, c->tri_err)
    if (nrhs <= 1)
        MH_TRSV(1);
    else if (nrhs <= 2)
        MH_TRSV(2);
    else
        MH_TRSV(4);
#undef MH_TRSV
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

template <class T>
static int tri_solve_t(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* LU, int64_t ld,
                       int64_t stride, const int* perm, const typename Cx<T>::t* inv, int trans, int part,
                       int64_t nrhs, typename Cx<T>::t* B, int64_t ldb, int64_t bstride) {
    typedef typename Cx<T>::t C2;
    if (n == 0 || batch == 0 || nrhs == 0 || part == 0) return MH_OK;
    MH_REQUIRE(n <= kMaxSolveN, "lu_solve: n larger than 4096 is not supported");
    const int64_t nsubp = (n + kTsCB - 1) / kTsCB;
    const int64_t nslab = (nsubp * kTsCB + kTsRows - 1) / kTsRows;
    MH_REQUIRE(nslab * batch <= 0x7fffffff, "lu_solve: batch too large");
    MH_REQUIRE((part & 1) == 0 || perm, "lu_solve: permutation missing");
    MH_REQUIRE(inv, "lu_solve: inverse diagonal blocks missing");
    TRY_RC(tri_state(c, batch, nsubp, sizeof(C2) * (size_t)(n * kTsNR * batch)));
    C2* W = static_cast<C2*>(c->tri[c->tri_slot].ws);
    const int64_t ws = n * kTsNR;
    const C2* invL = inv;
    const C2* invU = inv + batch * nsubp * kTsCB * kTsCB;
    const double tri = ((part & 1) ? 0.5 : 0.0) + ((part & 2) ? 0.5 : 0.0);
    ProfScope ps(c, sizeof(T) == 8 ? "lu_solve" : "lu_solve_f32", 8.0 * tri * (double)n * n * nrhs * batch,
                 sizeof(C2) * tri * (double)n * n * batch * (double)((nrhs + kTsNR - 1) / kTsNR));
    for (int64_t r0 = 0; r0 < nrhs; r0 += kTsNR) {
        const int nr = (int)((nrhs - r0) < kTsNR ? (nrhs - r0) : kTsNR);
        C2* b = B + r0 * ldb;
        const int64_t tot = n * nr * batch;
        const unsigned cgrid = (unsigned)((tot + 255) / 256);
        if (!trans) {
            if (part == 3) {
                TRY_RC((launch_trsv<T, false, false>(c, n, batch, LU, ld, stride, invL, perm, nullptr, nr, b, ldb,
                                                     bstride, W, n, ws)));
                TRY_RC((launch_trsv<T, false, true>(c, n, batch, LU, ld, stride, invU, nullptr, nullptr, nr, W, n,
                                                    ws, b, ldb, bstride)));
            } else if (part == 2) {
                TRY_RC((launch_trsv<T, false, true>(c, n, batch, LU, ld, stride, invU, nullptr, nullptr, nr, b,
                                                    ldb, bstride, b, ldb, bstride)));
            } else {
                TRY_RC((launch_trsv<T, false, false>(c, n, batch, LU, ld, stride, invL, perm, nullptr, nr, b, ldb,
                                                     bstride, W, n, ws)));
                hipLaunchKernelGGL(k_copy_cols<C2>, dim3(cgrid), dim3(256), 0, c->stream, n, (int64_t)nr, batch,
                                   (const C2*)W, n, ws, b, ldb, bstride);
                MH_HIP_CHECK(hipGetLastError());
            }
        } else {
            if (part == 3) {
                TRY_RC((launch_trsv<T, true, false>(c, n, batch, LU, ld, stride, invU, nullptr, nullptr, nr, b,
                                                    ldb, bstride, W, n, ws)));
                TRY_RC((launch_trsv<T, true, true>(c, n, batch, LU, ld, stride, invL, nullptr, perm, nr, W, n, ws,
                                                   b, ldb, bstride)));
            } else if (part == 2) {
                TRY_RC((launch_trsv<T, true, false>(c, n, batch, LU, ld, stride, invU, nullptr, nullptr, nr, b,
                                                    ldb, bstride, b, ldb, bstride)));
            } else {
                hipLaunchKernelGGL(k_copy_cols<C2>, dim3(cgrid), dim3(256), 0, c->stream, n, (int64_t)nr, batch,
                                   (const C2*)b, ldb, bstride, W, n, ws);
                MH_HIP_CHECK(hipGetLastError());
                TRY_RC((launch_trsv<T, true, true>(c, n, batch, LU, ld, stride, invL, nullptr, perm, nr, W, n, ws,
                                                   b, ldb, bstride)));
            }
        }
    }
    return MH_OK;
}

int tri_solve(Context* c, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
              const int* perm, const double2* inv, int trans, int part, int64_t nrhs, double2* B, int64_t ldb,
              int64_t bstride) {
    return tri_solve_t<double>(c, n, batch, LU, ld, stride, perm, inv, trans, part, nrhs, B, ldb, bstride);
}

int tri_solve_f32(Context* c, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                  const int* perm, const float2* inv, int trans, int part, int64_t nrhs, float2* B, int64_t ldb,
                  int64_t bstride) {
    return tri_solve_t<float>(c, n, batch, LU, ld, stride, perm, inv, trans, part, nrhs, B, ldb, bstride);
}

// Reads and clears the inter-workgroup error word (synchronises the stream).
int tri_check(Context* c) {
    if (!c->tri_err) return MH_OK;
    int h = 0;
    MH_HIP_CHECK(hipMemcpyAsync(&h, c->tri_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (h) {
        MH_HIP_CHECK(hipMemsetAsync(c->tri_err, 0, sizeof(int), c->stream));
        MH_HIP_CHECK(hipStreamSynchronize(c->stream));
        if (h & 2) {
            set_error("lu_solve: pivot index out of range");
            return MH_ERR_ARGUMENT;
        }
        set_error("lu_solve: inter-workgroup wait timed out");
        return MH_ERR_HIP;
    }
    return MH_OK;
}

}  // namespace mh
