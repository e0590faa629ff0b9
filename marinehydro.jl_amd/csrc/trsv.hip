// Batched triangular solves with the LU factors, pipelined across many
// workgroups per matrix.
//
// A solve with one n x n triangle and a few right-hand sides reads the whole
// triangle once; one workgroup per matrix would leave most of the chip idle
// and be bound by one CU's bandwidth.  Here every matrix is cut into slabs of
// 256 rows, one workgroup per slab (one row per thread).  The unknowns are
// produced in 32-row blocks in order.  A workgroup keeps its rows' partial
// sums in registers; it applies the blocks of earlier slabs as their
// solutions are published, several ready blocks per round, with its rows'
// slice of the next block in flight while the current one is applied.
// For its own blocks it multiplies the block's residual by the inverse of
// the 32 x 32 diagonal block (precomputed once per factorisation and fetched
// one block ahead, so the dependency chain carries no sequential
// substitution and no memory latency), publishes the solution and updates
// its remaining rows.
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
// For the transposed solves the factors can also be given as an explicit
// transposed copy, which turns the strided reads into coalesced ones.
//
// Publication: the output is filled with an all-ones bit pattern ("not yet
// written") before the launch; a producer writes each solution entry with
// 64-bit device-scope (agent) atomic stores, which are coherent across the
// XCDs' L2 caches, and consumers poll the entries themselves, so a block is
// usable as soon as its values land, with no flag and no wait for store
// completion on the producer's side.  A computed value with the reserved
// pattern (a NaN) is stored as the canonical NaN.  Workgroups take their
// (slab, matrix) from an atomic ticket in the order they start, so every
// workgroup a running one waits for has started before it.  Waits are
// bounded; on timeout an error word is set and the kernel runs to
// completion (the host reports the error).
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
constexpr unsigned long long kUnset = ~0ull;  // "not yet written"
constexpr int kNBPerm = 128;  // interchanges folded per k_perm_update launch (a block column)

__device__ __forceinline__ unsigned long long ld_pub(const void* p) {
    return __hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<void*>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_pub(void* p, unsigned long long b) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Loads a solution entry; true once it has been written.
__device__ __forceinline__ bool pub_load(const float2* p, float2& v) {
    const unsigned long long b = ld_pub(p);
    v = make_float2(__uint_as_float((unsigned)b), __uint_as_float((unsigned)(b >> 32)));
    return b != kUnset;
}
__device__ __forceinline__ bool pub_load(const double2* p, double2& v) {
    const unsigned long long a = ld_pub(p), b = ld_pub(reinterpret_cast<const double*>(p) + 1);
    v = make_double2(__longlong_as_double((long long)a), __longlong_as_double((long long)b));
    return a != kUnset && b != kUnset;
}
__device__ __forceinline__ void pub_store(float2* p, float2 v) {
    unsigned long long b =
        (unsigned long long)__float_as_uint(v.x) | ((unsigned long long)__float_as_uint(v.y) << 32);
    if (b == kUnset) b = 0x7fc000007fc00000ull;
    st_pub(p, b);
}
__device__ __forceinline__ void pub_store(double2* p, double2 v) {
    unsigned long long a = (unsigned long long)__double_as_longlong(v.x);
    unsigned long long b = (unsigned long long)__double_as_longlong(v.y);
    if (a == kUnset) a = 0x7ff8000000000000ull;
    if (b == kUnset) b = 0x7ff8000000000000ull;
    st_pub(p, a);
    st_pub(reinterpret_cast<double*>(p) + 1, b);
}

// A row's slice of block c (logical columns 32c .. 32c+31) for rows after
// the block, and its application with the block's solution x.  Every load
// stays inside the matrix (offsets of rows/columns outside it are replaced
// by 0 and the value by zero).
template <bool TRM, bool REV, class C2>
__device__ __forceinline__ void trsv_fetch(C2 (&m)[kTsCB], __amdgpu_buffer_rsrc_t rs, uint32_t vo, uint32_t es,
                                           int np, int n, int nsubp, int i, bool valid, int c) {
    const bool need = valid && c < nsubp && i >= (c + 1) * kTsCB;
    // physical row/column of entry q: p0 + q (forward) or p0 - q (backward);
    // entries with p >= n are padding
    const int p0 = REV ? np - 1 - c * kTsCB : c * kTsCB;
#pragma unroll
    for (int q = 0; q < kTsCB; ++q) {
        const int pj = REV ? p0 - q : p0 + q;
        const bool in = pj < n;
        const uint32_t so = (uint32_t)(in ? pj : 0) * es;
        const C2 val = buf_load(rs, vo, so, C2());
        m[q] = (need && in) ? val : mkc<C2>(0, 0);
    }
}
template <int NR, class C2>
__device__ __forceinline__ void trsv_apply(const C2 (&m)[kTsCB], C2 (&v)[NR], int nrhs, int i, bool valid, int c,
                                           const C2 (*x)[NR]) {
    if (valid && i >= (c + 1) * kTsCB) {
#pragma unroll
        for (int q = 0; q < kTsCB; ++q)
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (r < nrhs) v[r] = cfms(v[r], m[q], x[q][r]);
    }
}

// Inverse of the diagonal block of logical block cc (physical block K,
// element (a, b) at Inv[((item * nsubp + K) * 32 + b) * 32 + a]): fetched
// into registers one block ahead, staged into LDS when its turn comes.
// The parts are held as separate real arrays: an array of HIP's vector
// structs carried around the block loop stays in scratch memory.
constexpr int kTsIE = kTsCB * kTsCB / kTsRows;
template <bool REV, class T, class C2>
__device__ __forceinline__ void fetch_inv(T (&ibx)[kTsIE], T (&iby)[kTsIE], const C2* __restrict__ Inv, int64_t item,
                                          int nsubp, int cc, int tid) {
    const int64_t K = REV ? nsubp - 1 - cc : cc;
    const C2* src = Inv + (item * nsubp + K) * kTsCB * kTsCB;
#pragma unroll
    for (int e = 0; e < kTsIE; ++e) {
        const C2 v = src[tid + kTsRows * e];
        ibx[e] = v.x;
        iby[e] = v.y;
    }
}
template <class T, class C2>
__device__ __forceinline__ void stash_inv(C2 (*dst)[kTsCB + 1], const T (&ibx)[kTsIE], const T (&iby)[kTsIE], int tid) {
#pragma unroll
    for (int e = 0; e < kTsIE; ++e) {
        const int x = tid + kTsRows * e;
        dst[x % kTsCB][x / kTsCB] = mkc<C2>(ibx[e], iby[e]);
    }
}

// TR: transposed solve (inverse blocks read transposed); TRM: the matrix is
// read transposed (false when A is an explicit transposed copy).
template <class T, bool TR, bool TRM, bool REV, int NR>
__global__ __launch_bounds__(kTsRows) __attribute__((amdgpu_waves_per_eu(2))) void k_trsv(
    int64_t n, int64_t batch, const typename Cx<T>::t* __restrict__ A, int64_t ld, int64_t stride,
    const typename Cx<T>::t* __restrict__ Inv, const int* __restrict__ gperm,
    const int* __restrict__ sperm, int nrhs, const typename Cx<T>::t* in, int64_t ldi, int64_t istride,
    typename Cx<T>::t* out, int64_t ldo, int64_t ostride, int nslab, int nsubp,
    unsigned long long* __restrict__ ticket, int* __restrict__ err) {
    typedef typename Cx<T>::t C2;
    __shared__ unsigned long long s_tk;
    __shared__ unsigned s_bad[2];           // per round (parity): blocks c+b not yet complete
    __shared__ C2 sI[2][kTsCB][kTsCB + 1];  // inverse diagonal blocks, double-buffered
    __shared__ C2 sx[kTsBatch][kTsCB][NR];  // solutions of the blocks being applied
    __shared__ C2 sr[kTsCB][NR];            // residual of the current own block
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        s_tk = atomicAdd(ticket, 1ull);
        s_bad[0] = 0u;
        s_bad[1] = 0u;
    }
    __syncthreads();
    // values read from LDS are wave-uniform but not known to be so: read them
    // back as scalars, so the indices derived from them (and the scalar
    // offsets of the buffer loads) stay in scalar registers
    const unsigned long long tk_lds = s_tk;
    const int64_t tk = (int64_t)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(tk_lds >> 32)) << 32) |
                                 (unsigned)__builtin_amdgcn_readfirstlane((unsigned)tk_lds));
    if (tk >= (int64_t)nslab * batch) return;  // cannot happen with a reset ticket
    const int slab = (int)(tk % nslab);
    const int64_t item = tk / nslab;
    const int64_t np = (int64_t)nsubp * kTsCB;
    const C2* M = A + item * stride;
    // extent (n-1) * (ld+1) + 1 entries; the host checks it fits 31 bits
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<C2*>(M), (short)0, (int)(((n - 1) * (ld + 1) + 1) * (int64_t)sizeof(C2)), 0x00020000);
    const int* gp = gperm ? gperm + item * n : nullptr;
    const int* sp = sperm ? sperm + item * n : nullptr;
    const C2* bin = in + item * istride;
    C2* bout = out + item * ostride;
    auto phys = [&](int64_t i) -> int64_t { return REV ? np - 1 - i : i; };
    auto dst_of = [&](int64_t p) -> int64_t { return sp ? (int64_t)sp[p] : p; };

    const int64_t i0 = (int64_t)slab * kTsRows;
    const int64_t i = i0 + tid;
    const int64_t pi = phys(i);
    const bool valid = i < np && pi < n;
    // byte offsets into the matrix: per lane (row pi, or column pi of the
    // transposed read) and per entry of a slice (scaled by es)
    const uint32_t vo = valid ? (uint32_t)(TRM ? pi * ld : pi) * (uint32_t)sizeof(C2) : 0u;
    const uint32_t es = (uint32_t)(TRM ? 1 : ld) * (uint32_t)sizeof(C2);
    const int cown = (int)(i0 / kTsCB);
    const int cend = (cown + kTsOwn) < nsubp ? cown + kTsOwn : nsubp;

    // inverse of own block s: fetched one block ahead, staged into LDS
    // buffer s & 1 when its turn comes
    T ibx[kTsIE], iby[kTsIE];
    if (cend > cown) fetch_inv<REV>(ibx, iby, Inv, item, nsubp, cown, tid);

    C2 v[NR];
    {
        const int64_t src = valid ? (gp ? (int64_t)gp[pi] : pi) : 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) v[r] = (valid && r < nrhs) ? bin[src + r * ldi] : mkc<C2>(T(0), T(0));
    }

    // the row's slices are consumed in block order 0, 1, ..., cend-1: apply
    // block c, then fetch block c + 1's slice
    C2 ma[kTsCB];
#define MH_TRSV_STEP(C, X)                                                                      \
    do {                                                                                        \
        const int c_ = (C);                                                                     \
        trsv_apply<NR>(ma, v, nrhs, (int)i, valid, c_, (X));                                    \
        trsv_fetch<TRM, REV>(ma, rs, vo, es, (int)np, (int)n, nsubp, (int)i, valid, c_ + 1);    \
    } while (0)
    trsv_fetch<TRM, REV>(ma, rs, vo, es, (int)np, (int)n, nsubp, (int)i, valid, 0);

    // 1. blocks of earlier slabs, in rounds: wave 0 waits for block c while
    //    the other waves take the entries of blocks c+1 .. c+7 that are
    //    already written; the round applies block c and the complete ones
    //    that follow it
    int c = 0, par = 0;
    while (c < cown) {
        if (tid == 0) s_bad[par ^ 1] = 0u;
        if (wave == 0) {
            unsigned spins = 0;
            for (;;) {
                bool ok = true;
#pragma unroll 1
                for (int e = lane; e < kTsCB * NR; e += 64) {
                    const int q = e % kTsCB, r = e / kTsCB;
                    const int64_t pj = phys((int64_t)c * kTsCB + q);
                    C2 val = mkc<C2>(T(0), T(0));
                    if (r < nrhs && pj < n) ok = pub_load(bout + dst_of(pj) + r * ldo, val) && ok;
                    sx[0][q][r] = val;
                }
                if (__all(ok)) break;
                const int bad = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__any(bad != 0)) break;
                if (++spins > kSpinLimit) {
                    if (lane == 0) atomicOr(err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        } else {
            unsigned bad = 0u;
#pragma unroll 1
            for (int e = tid - 64; e < (kTsBatch - 1) * kTsCB * NR; e += kTsRows - 64) {
                const int b = 1 + e / (kTsCB * NR), q = e % kTsCB, r = (e / kTsCB) % NR;
                if (c + b >= cown) {
                    bad |= 1u << b;
                    continue;
                }
                const int64_t pj = phys((int64_t)(c + b) * kTsCB + q);
                C2 val = mkc<C2>(T(0), T(0));
                if (r < nrhs && pj < n && !pub_load(bout + dst_of(pj) + r * ldo, val)) bad |= 1u << b;
                sx[b][q][r] = val;
            }
            if (bad) atomicOr(&s_bad[par], bad);
        }
        __syncthreads();
        const int nb = __builtin_amdgcn_readfirstlane(__builtin_ctz(s_bad[par] | (1u << kTsBatch)));
        for (int b = 0; b < nb; ++b) MH_TRSV_STEP(c + b, sx[b]);
        c += nb;
        par ^= 1;
        __syncthreads();
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
        stash_inv(sI[s & 1], ibx, iby, tid);
        if (s + 1 < cend - cown) fetch_inv<REV>(ibx, iby, Inv, item, nsubp, cown + s + 1, tid);
        __syncthreads();
        if (mine) {
            const int a = tid - base;
            const int pa = REV ? kTsCB - 1 - a : a;
            C2 acc[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = mkc<C2>(T(0), T(0));
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
            const int64_t dst = valid ? dst_of(pi) : 0;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                v[r] = acc[r];
                sx[0][a][r] = acc[r];
                if (valid && r < nrhs) pub_store(bout + dst + r * ldo, acc[r]);
            }
        }
        __syncthreads();
        MH_TRSV_STEP(cb, sx[0]);
    }
#undef MH_TRSV_STEP
}

// Marks the n x cols x batch output of a solve as not yet written.
template <class C2>
__global__ void k_fill_unset(int64_t rows, int64_t cols, int64_t batch, C2* __restrict__ out, int64_t ldo,
                             int64_t ostride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t tot = rows * cols * batch;
    if (t >= tot) return;
    const int64_t r = t % rows, c = (t / rows) % cols, b = t / (rows * cols);
    unsigned long long* p = reinterpret_cast<unsigned long long*>(out + b * ostride + c * ldo + r);
#pragma unroll
    for (int w = 0; w < (int)(sizeof(C2) / 8); ++w) p[w] = kUnset;
}

// Inverses of the 32 x 32 diagonal blocks of the unit-lower L and of U
// (physical blocks K = 0 .. nsubp-1; a last partial block is padded with the
// identity).  One wave per block: lane b < 32 builds column b of L^-1 by
// forward substitution, lane 32 + b column b of U^-1 by back substitution.
template <class T>
__global__ __launch_bounds__(64) void k_tri_inv(int64_t n, int64_t batch, const typename Cx<T>::t* __restrict__ A,
                                                int64_t ld, int64_t stride, int nsubp,
                                                typename Cx<T>::t* __restrict__ inv, int K0) {
    typedef typename Cx<T>::t C2;
    __shared__ C2 sB[kTsCB][kTsCB + 1];
    __shared__ C2 sY[2][kTsCB][kTsCB + 1];  // columns of L^-1 and U^-1
    const int K = K0 + blockIdx.x;
    const int64_t item = blockIdx.y;
    const int lane = threadIdx.x;
    const C2* M = A + item * stride;
    const int64_t r0 = (int64_t)K * kTsCB;
    const int jn = (int)((n - r0) < kTsCB ? (n - r0) : kTsCB);
    {  // the diagonal block, every load issued before any is waited for
       // (outside the matrix: element 0 loaded, identity entries used)
        constexpr int kPer = kTsCB * kTsCB / 64;
        T vx[kPer], vy[kPer];
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int q = lane + 64 * e, ra = q % kTsCB, cb = q / kTsCB;
            const bool in = ra < jn && cb < jn;
            const C2 v = M[in ? (r0 + ra) + (r0 + cb) * ld : 0];
            vx[e] = in ? v.x : T(ra == cb ? 1 : 0);
            vy[e] = in ? v.y : T(0);
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int q = lane + 64 * e;
            sB[q % kTsCB][q / kTsCB] = mkc<C2>(vx[e], vy[e]);
        }
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

// The interchanges ipiv[k0 .. k0+w) folded into the net permutation perm
// (initialised to the identity when k0 = 0); they only move entries >= k0,
// which are staged in LDS.  One workgroup per item, one thread applies the
// interchanges in order.
__global__ __launch_bounds__(256) void k_perm_update(int64_t n, const int64_t* __restrict__ ipiv,
                                                     int* __restrict__ perm, int* __restrict__ err, int64_t k0,
                                                     int w) {
    __shared__ int s[kMaxSolveN];
    __shared__ int sp[kNBPerm];
    const int64_t item = blockIdx.x;
    const int64_t* pv = ipiv + item * n;
    int* pm = perm + item * n;
    for (int64_t k = k0 + threadIdx.x; k < n; k += blockDim.x) s[k - k0] = k0 == 0 ? (int)k : pm[k];
    for (int k = threadIdx.x; k < w; k += blockDim.x) {
        const int64_t p = pv[k0 + k];
        sp[k] = (p < k0 + k || p >= n) ? -1 : (int)(p - k0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < w; ++k) {
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
    for (int64_t k = k0 + threadIdx.x; k < n; k += blockDim.x) pm[k] = s[k - k0];
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

// per-context synchronisation state: the workgroup ticket, the error word
// and a solution workspace for the out-of-place steps
static int tri_state(Context* c, size_t ws_bytes) {
    Context::TriSlot& t = c->tri[c->tri_slot];
    const size_t need = 256;
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
    TRY_RC(tri_state(c, 0));
    if (perm) {
        hipLaunchKernelGGL(k_perm_from_ipiv, dim3((unsigned)batch), dim3(256), 0, c->stream, n, ipiv, perm,
                           c->tri_err);
        MH_HIP_CHECK(hipGetLastError());
    }
    const int nsubp = (int)((n + kTsCB - 1) / kTsCB);
    ProfScope ps(c, sizeof(T) == 8 ? "tri_inv" : "tri_inv_f32", 0.0, 0.0);
    hipLaunchKernelGGL(k_tri_inv<T>, dim3((unsigned)nsubp, (unsigned)batch), dim3(64), 0, c->stream, n, batch,
                       LU, ld, stride, nsubp, inv, 0);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// The part of tri_prepare that block column [k0, k0+w) of a factorisation in
// progress allows (w <= 128; k0 and w multiples of 32 except at the end):
// its interchanges folded into perm, the inverses of its diagonal blocks.
template <class T>
static int tri_prepare_part_t(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* LU, int64_t ld,
                              int64_t stride, const int64_t* ipiv, int* perm, typename Cx<T>::t* inv, int64_t k0,
                              int w) {
    if (n == 0 || batch == 0 || w <= 0) return MH_OK;
    MH_REQUIRE(n <= kMaxSolveN, "lu_solve: n larger than 4096 is not supported");
    MH_REQUIRE(batch <= 65535, "lu_solve: batch too large");
    MH_REQUIRE(w <= kNBPerm && k0 % kTsCB == 0 && (k0 + w == n || w % kTsCB == 0), "lu_solve: bad block column");
    TRY_RC(tri_state(c, 0));
    if (perm) {
        hipLaunchKernelGGL(k_perm_update, dim3((unsigned)batch), dim3(256), 0, c->stream, n, ipiv, perm, c->tri_err,
                           k0, w);
        MH_HIP_CHECK(hipGetLastError());
    }
    const int nsubp = (int)((n + kTsCB - 1) / kTsCB);
    const int K0 = (int)(k0 / kTsCB), K1 = (int)((k0 + w + kTsCB - 1) / kTsCB);
    ProfScope ps(c, sizeof(T) == 8 ? "tri_inv" : "tri_inv_f32", 0.0, 0.0);
    hipLaunchKernelGGL(k_tri_inv<T>, dim3((unsigned)(K1 - K0), (unsigned)batch), dim3(64), 0, c->stream, n, batch,
                       LU, ld, stride, nsubp, inv, K0);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}
int tri_prepare_part(Context* c, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                     const int64_t* ipiv, int* perm, double2* inv, int64_t k0, int w) {
    return tri_prepare_part_t<double>(c, n, batch, LU, ld, stride, ipiv, perm, inv, k0, w);
}
int tri_prepare_part(Context* c, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                     const int64_t* ipiv, int* perm, float2* inv, int64_t k0, int w) {
    return tri_prepare_part_t<float>(c, n, batch, LU, ld, stride, ipiv, perm, inv, k0, w);
}

int tri_prepare(Context* c, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                const int64_t* ipiv, int* perm, double2* inv) {
    return tri_prepare_t<double>(c, n, batch, LU, ld, stride, ipiv, perm, inv);
}
int tri_prepare_f32(Context* c, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                    const int64_t* ipiv, int* perm, float2* inv) {
    return tri_prepare_t<float>(c, n, batch, LU, ld, stride, ipiv, perm, inv);
}

// One triangle (see the kernel); A is the factor matrix, or its explicit
// transposed copy when TRM is false for a transposed solve.  `out` must not
// overlap `in`.
template <class T, bool TR, bool TRM, bool REV>
static int launch_trsv(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* A, int64_t ld,
                       int64_t stride, const typename Cx<T>::t* inv, const int* gperm, const int* sperm,
                       int nrhs, const typename Cx<T>::t* in, int64_t ldi, int64_t istride,
                       typename Cx<T>::t* out, int64_t ldo, int64_t ostride) {
    typedef typename Cx<T>::t C2;
    const int nsubp = (int)((n + kTsCB - 1) / kTsCB);
    const int nslab = (int)((nsubp * (int64_t)kTsCB + kTsRows - 1) / kTsRows);
    Context::TriSlot& ts = c->tri[c->tri_slot];
    unsigned long long* ticket = reinterpret_cast<unsigned long long*>(ts.sync);
    hipLaunchKernelGGL(k_reset_ticket, dim3(1), dim3(1), 0, c->stream, ticket);
    const int64_t tot = n * nrhs * batch;
    hipLaunchKernelGGL(k_fill_unset<C2>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, n,
                       (int64_t)nrhs, batch, out, ldo, ostride);
#define MH_TRSV(NRR)                                                                                              \
    hipLaunchKernelGGL((k_trsv<T, TR, TRM, REV, NRR>), dim3((unsigned)(nslab * batch)), dim3(kTsRows), 0,         \
                       c->stream, n, batch, A, ld, stride, inv, gperm, sperm, nrhs, in, ldi, istride, out, ldo,   \
                       ostride, nslab, nsubp, ticket, c->tri_err)
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

// Rows [r_lo, r_hi) of each n x n matrix, transposed into columns
// [r_lo, r_hi) of At: 64 x 64 tiles through LDS, reads and writes both run
// down columns.
template <class C2>
__global__ __launch_bounds__(256) void k_transpose(int64_t n, const C2* __restrict__ A, int64_t ld,
                                                   int64_t stride, C2* __restrict__ At, int64_t r_lo,
                                                   int64_t r_hi) {
    __shared__ C2 t[64][65];
    const int64_t item = blockIdx.z;
    const int64_t r0 = r_lo + (int64_t)blockIdx.x * 64, c0 = (int64_t)blockIdx.y * 64;
    const C2* src = A + item * stride;
    C2* dst = At + item * stride;
    const int x = threadIdx.x & 63, y = threadIdx.x >> 6;
    for (int cc = y; cc < 64; cc += 4)
        if (r0 + x < r_hi && c0 + cc < n) t[cc][x] = src[(r0 + x) + (c0 + cc) * ld];
    __syncthreads();
    // At[c0 + x, r0 + rr] = A[r0 + rr, c0 + x]
    for (int rr = y; rr < 64; rr += 4)
        if (c0 + x < n && r0 + rr < r_hi) dst[(c0 + x) + (r0 + rr) * ld] = t[x][rr];
}

template <class T>
static int tri_transpose_t(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* LU, int64_t ld,
                           int64_t stride, typename Cx<T>::t* LUt, int64_t r_lo, int64_t r_hi) {
    if (r_hi < 0) r_hi = n;
    if (n == 0 || batch == 0 || r_hi <= r_lo) return MH_OK;
    MH_REQUIRE(batch <= 65535, "lu_solve: batch too large");
    MH_REQUIRE(r_lo >= 0 && r_hi <= n, "transpose: row range outside the matrix");
    const unsigned tr = (unsigned)((r_hi - r_lo + 63) / 64), tc = (unsigned)((n + 63) / 64);
    ProfScope ps(c, "transpose", 0.0, 2.0 * sizeof(typename Cx<T>::t) * (double)(r_hi - r_lo) * n * batch);
    // The FP32 transposes run beside the LU's panel chain: 24 KB of LDS they
    // do not use (57 KB a workgroup) keeps them at two workgroups per CU
    // instead of four, leaving room for the chain (33.3-33.5 -> 33.1 ms per
    // sweep at 2,048 panels x 64 frequencies; 48 KB was slower).  The FP64
    // tile already takes 66 KB.
    const unsigned pad = sizeof(T) == 4 ? 24576u : 0u;
    hipLaunchKernelGGL(k_transpose<typename Cx<T>::t>, dim3(tr, tc, (unsigned)batch), dim3(256), pad, c->stream, n,
                       LU, ld, stride, LUt, r_lo, r_hi);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}
int tri_transpose(Context* c, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                  double2* LUt, int64_t r_lo, int64_t r_hi) {
    return tri_transpose_t<double>(c, n, batch, LU, ld, stride, LUt, r_lo, r_hi);
}
int tri_transpose_f32(Context* c, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                      float2* LUt, int64_t r_lo, int64_t r_hi) {
    return tri_transpose_t<float>(c, n, batch, LU, ld, stride, LUt, r_lo, r_hi);
}

// One triangle: as stored, or transposed (from the explicit copy LUt when
// given, else by strided reads of LU).
template <class T, bool REV>
static int one_tri(Context* c, bool trans, int64_t n, int64_t batch, const typename Cx<T>::t* LU,
                   const typename Cx<T>::t* LUt, int64_t ld, int64_t stride, const typename Cx<T>::t* inv,
                   const int* gperm, const int* sperm, int nr, const typename Cx<T>::t* in, int64_t ldi,
                   int64_t istride, typename Cx<T>::t* out, int64_t ldo, int64_t ostride) {
    if (!trans)
        return launch_trsv<T, false, false, REV>(c, n, batch, LU, ld, stride, inv, gperm, sperm, nr, in, ldi,
                                                 istride, out, ldo, ostride);
    if (LUt)
        return launch_trsv<T, true, false, REV>(c, n, batch, LUt, ld, stride, inv, gperm, sperm, nr, in, ldi,
                                                istride, out, ldo, ostride);
    return launch_trsv<T, true, true, REV>(c, n, batch, LU, ld, stride, inv, gperm, sperm, nr, in, ldi, istride,
                                           out, ldo, ostride);
}

template <class T>
static int tri_solve_t(Context* c, int64_t n, int64_t batch, const typename Cx<T>::t* LU, int64_t ld,
                       int64_t stride, const int* perm, const typename Cx<T>::t* inv, int trans, int part,
                       int64_t nrhs, typename Cx<T>::t* B, int64_t ldb, int64_t bstride,
                       const typename Cx<T>::t* LUt) {
    typedef typename Cx<T>::t C2;
    if (n == 0 || batch == 0 || nrhs == 0 || part == 0) return MH_OK;
    MH_REQUIRE(n <= kMaxSolveN, "lu_solve: n larger than 4096 is not supported");
    const int64_t nsubp = (n + kTsCB - 1) / kTsCB;
    const int64_t nslab = (nsubp * kTsCB + kTsRows - 1) / kTsRows;
    MH_REQUIRE(nslab * batch <= 0x7fffffff, "lu_solve: batch too large");
    MH_REQUIRE((part & 1) == 0 || perm, "lu_solve: permutation missing");
    MH_REQUIRE(inv, "lu_solve: inverse diagonal blocks missing");
    MH_REQUIRE(((n - 1) * (ld + 1) + 1) * (int64_t)sizeof(C2) <= 0x7fffffff,
               "lu_solve: matrix extent (leading dimension x n) above 2 GiB");
    TRY_RC(tri_state(c, sizeof(C2) * (size_t)(n * kTsNR * batch)));
    C2* W = static_cast<C2*>(c->tri[c->tri_slot].ws);
    const int64_t ws = n * kTsNR;
    const C2* invL = inv;
    const C2* invU = inv + batch * nsubp * kTsCB * kTsCB;
    const double tri = ((part & 1) ? 0.5 : 0.0) + ((part & 2) ? 0.5 : 0.0);
    ProfScope ps(c, sizeof(T) == 8 ? "lu_solve" : "lu_solve_f32", 8.0 * tri * (double)n * n * nrhs * batch,
                 sizeof(C2) * tri * (double)n * n * batch * (double)((nrhs + kTsNR - 1) / kTsNR));
    // first triangle: L (forward, gather through perm) or U^T; second: U or
    // L^T (scatter through perm); each solve writes a buffer it does not read
    const bool tr = trans != 0;
    const bool do_first = tr ? (part & 2) != 0 : (part & 1) != 0;
    const bool do_second = tr ? (part & 1) != 0 : (part & 2) != 0;
    const C2* inv_first = tr ? invU : invL;
    const C2* inv_second = tr ? invL : invU;
    const int* gp = tr ? nullptr : perm;
    const int* sp = tr ? perm : nullptr;
    for (int64_t r0 = 0; r0 < nrhs; r0 += kTsNR) {
        const int nr = (int)((nrhs - r0) < kTsNR ? (nrhs - r0) : kTsNR);
        C2* b = B + r0 * ldb;
        const int64_t tot = n * nr * batch;
        const unsigned cgrid = (unsigned)((tot + 255) / 256);
        if (do_first)
            TRY_RC((one_tri<T, false>(c, tr, n, batch, LU, LUt, ld, stride, inv_first, gp, nullptr, nr, b, ldb,
                                      bstride, W, n, ws)));
        else {
            hipLaunchKernelGGL(k_copy_cols<C2>, dim3(cgrid), dim3(256), 0, c->stream, n, (int64_t)nr, batch,
                               (const C2*)b, ldb, bstride, W, n, ws);
            MH_HIP_CHECK(hipGetLastError());
        }
        if (do_second)
            TRY_RC((one_tri<T, true>(c, tr, n, batch, LU, LUt, ld, stride, inv_second, nullptr, sp, nr, W, n, ws,
                                     b, ldb, bstride)));
        else {
            hipLaunchKernelGGL(k_copy_cols<C2>, dim3(cgrid), dim3(256), 0, c->stream, n, (int64_t)nr, batch,
                               (const C2*)W, n, ws, b, ldb, bstride);
            MH_HIP_CHECK(hipGetLastError());
        }
    }
    return MH_OK;
}

int tri_solve(Context* c, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
              const int* perm, const double2* inv, int trans, int part, int64_t nrhs, double2* B, int64_t ldb,
              int64_t bstride, const double2* LUt) {
    return tri_solve_t<double>(c, n, batch, LU, ld, stride, perm, inv, trans, part, nrhs, B, ldb, bstride, LUt);
}

int tri_solve_f32(Context* c, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                  const int* perm, const float2* inv, int trans, int part, int64_t nrhs, float2* B, int64_t ldb,
                  int64_t bstride, const float2* LUt) {
    return tri_solve_t<float>(c, n, batch, LU, ld, stride, perm, inv, trans, part, nrhs, B, ldb, bstride, LUt);
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
