// Batched dense complex LU (partial pivoting) and triangular solves, FP64 and FP32.
//
// Right-looking LU with block columns of NB = 128 (one batch item per
// frequency).  Each block column is factored recursively (left half, update
// of the right half, right half; as in LAPACK's zgetrf2) down to leaves of
// 4-16 columns, and every update is a trailing GEMM on the matrix cores:
//   k_lu_leaf      register-resident factorisation of an (n-k0) x <=16 panel,
//                  one workgroup per matrix, pivot = max |re|+|im| (izamax
//                  semantics, first index on ties)
//   k_lu_swap      row interchanges of a pivot block on a column range
//   k_lu_trsm_blk  unit-lower L11 X = B on a whole <=128-row block, column range
//   k_lu_gemm      C -= A B on arbitrary row/column/depth ranges, complex
//                  product as four real v_mfma_f64_16x16x4 products
// Extra columns beyond n (ncols > n) are carried through every step, so a
// right-hand side appended to the matrix leaves the factorisation already
// forward-substituted.  Reference: the LAPACK factorisation behind `\` /
// implicit_linear (src/matrix_assembly.jl:108-126).
#include "mh_cx.hpp"
#include "mh_internal.hpp"

namespace mh {

constexpr int kNBOuter = 128;  // block-column width of the trailing updates

// ---------------------------------------------------------------------------
// Wave reductions on DPP lane moves.  Within each row of 16 lanes the moves
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror and row_mirror
// (gfx9 DPP control codes below) pair every lane with one in the other half
// of a group of 2, 4, 8 and 16 lanes, so four steps leave the row's
// reduction in every lane of the row; the four rows are then combined from
// lanes 0, 16, 32 and 48 by readlane (a wave-uniform result).  Every lane of
// the wave must be active.
// ---------------------------------------------------------------------------
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float x) {
    return __int_as_float(dpp_i32<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = dpp_i32<CTRL>((int)b), hi = dpp_i32<CTRL>((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ int dpp_move(int x) {
    return dpp_i32<CTRL>(x);
}
__device__ __forceinline__ float read_lane(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double read_lane(double x, int l) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ int read_lane(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

// maximum (MAX) or minimum over each row of 16 lanes; x never NaN
template <bool MAX, class V>
__device__ __forceinline__ V row_reduce(V x) {
    V y = dpp_move<kDppXor1>(x);
    x = MAX ? (y > x ? y : x) : (y < x ? y : x);
    y = dpp_move<kDppXor2>(x);
    x = MAX ? (y > x ? y : x) : (y < x ? y : x);
    y = dpp_move<kDppHalfMirror>(x);
    x = MAX ? (y > x ? y : x) : (y < x ? y : x);
    y = dpp_move<kDppMirror>(x);
    x = MAX ? (y > x ? y : x) : (y < x ? y : x);
    return x;
}
template <bool MAX, class V>
__device__ __forceinline__ V wave_reduce(V x) {
    x = row_reduce<MAX>(x);
    V r = read_lane(x, 0);
#pragma unroll
    for (int l = 16; l < 64; l += 16) {
        const V y = read_lane(x, l);
        r = MAX ? (y > r ? y : r) : (y < r ? y : r);
    }
    return r;
}

// 1/p for a leaf pivot without a division: p scaled by a power of two so its
// larger component lies in [0.5, 1) (exact), 1/|p|^2 from the hardware
// reciprocal estimate refined by Newton steps (one in FP32, two in FP64),
// then scaled back (exact).  Within an ulp or two of Smith's algorithm;
// reciprocals of powers of two and of +-1 +-i stay exact.
__device__ __forceinline__ float rcp_refined(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return __builtin_fmaf(r, __builtin_fmaf(-d, r, 1.0f), r);
}
__device__ __forceinline__ double rcp_refined(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}
__device__ __forceinline__ int frexp_exponent(float x) { return __builtin_amdgcn_frexp_expf(x); }
__device__ __forceinline__ int frexp_exponent(double x) { return __builtin_amdgcn_frexp_exp(x); }
__device__ __forceinline__ float scale2(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }
__device__ __forceinline__ double scale2(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }
template <class C2>
__device__ __forceinline__ C2 pivot_recip(C2 p) {
    const int e = frexp_exponent(fmax(fabs(p.x), fabs(p.y)));
    const auto x = scale2(p.x, -e), y = scale2(p.y, -e);
    const auto r = rcp_refined(fma(x, x, y * y));
    return mkc<C2>(scale2(x * r, -e), scale2(-y * r, -e));
}

// x_c broadcast from lane L of each row of 16 lanes to the whole row (DPP
// row_newbcast); every lane of the row must be active
template <int L>
__device__ __forceinline__ float row_bcast(float x) {
    static_assert(L >= 0 && L < 16, "lane within a row of 16");
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x150 + L, 0xF, 0xF, false));
}
template <int L>
__device__ __forceinline__ double row_bcast(double x) {
    static_assert(L >= 0 && L < 16, "lane within a row of 16");
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x150 + L, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x150 + L, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// One substitution step of the leaf's optional solve: every lane of a DPP
// row holds row g of one column, x_g -= L11(g, C) x_C (L11 from LDS, zero on
// and above the diagonal).
template <class T, int C, int K>
struct LeafXStep {
    typedef typename Cx<T>::t C2;
    __device__ __forceinline__ static void run(T& xr, T& xi, const C2 (*sL)[16 + 1], int g, int w) {
        if (C < w) {  // uniform
            const T br = row_bcast<C>(xr), bi = row_bcast<C>(xi);
            const C2 l = sL[C][g];
            xr = fma(-l.x, br, fma(l.y, bi, xr));
            xi = fma(-l.x, bi, fma(-l.y, br, xi));
            LeafXStep<T, C + 1, K>::run(xr, xi, sL, g, w);
        }
    }
};
template <class T, int K>
struct LeafXStep<T, K, K> {
    typedef typename Cx<T>::t C2;
    __device__ __forceinline__ static void run(T&, T&, const C2 (*)[16 + 1], int, int) {}
};
constexpr int kLeafX = 16;  // columns the leaf's optional solve covers (one per DPP row of waves 0-3)

// ---------------------------------------------------------------------------
// leaf panel: columns [k0, k0+w), rows [k0, n), w <= kLeaf, held in
// registers.  Thread t owns local rows t + NT q, q < R.  Per column: every
// wave finds its largest |re|+|im| (first row on ties) and the lane owning
// that row publishes the row with the wave's maximum; after one barrier
// every wave picks the winner among the 16 candidates and reads its row.
// The leaf's row interchanges are then applied to the other columns of the
// enclosing block column [pc_lo, pc_hi) (at most kNBOuter wide), which
// replaces the separate swap passes of the recursion: the left columns hold
// final L values whose rows are only re-ordered, the right columns are not
// updated before the interchanges of every earlier leaf have reached them.
// ---------------------------------------------------------------------------
template <class T, int R, int kLeaf, int NT>
__global__ __launch_bounds__(NT) void k_lu_leaf(int64_t n, typename Cx<T>::t* __restrict__ A, int64_t ld,
                                                int64_t stride, int64_t k0, int w, int64_t* __restrict__ ipiv,
                                                int* __restrict__ info, int64_t pc_lo, int64_t pc_hi, int64_t x_lo,
                                                int64_t x_hi) {
    // panel-chain kernel: its waves win issue arbitration against those of
    // a trailing-update GEMM sharing the CU (the chain is the LU's critical
    // path)
    __builtin_amdgcn_s_setprio(3);
    typedef typename Cx<T>::t C2;
    static_assert(NT == 256 || NT == 512 || NT == 1024, "4, 8 or 16 waves");
    constexpr int NW = NT / 64;
    C2* M = A + (int64_t)blockIdx.x * stride;
    int64_t* piv = ipiv + (int64_t)blockIdx.x * n;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = (int)(n - k0);
    // double-buffered by column parity: one barrier per column (a wave writes
    // buffer c & 1 again only after every wave has passed the barrier of
    // column c + 1, i.e. finished reading column c)
    __shared__ T s_v[2][NW];
    __shared__ int s_i[2][NW];
    __shared__ C2 s_cand[2][NW][kLeaf];           // each wave's candidate pivot row
    __shared__ C2 s_crow[2][kLeaf];
    __shared__ int s_p[kLeaf];                    // local pivot rows
    __shared__ int s_rows[2 * kLeaf], s_src[2 * kLeaf];
    __shared__ C2 s_L[kLeaf <= 16 ? kLeaf : 16][16 + 1];  // L11 for the optional solve, s_L[c][r]
    // buffer addressing from the panel's first row M + k0: entry (k0 + r,
    // col) at byte (r + col ld) sizeof(C2) (the host checks the extent fits
    // 31 bits); rows outside the panel load from offset 0 and are zeroed
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        M + k0, (short)0, (int)(((int64_t)(m - 1) + (n - 1) * ld + 1) * (int64_t)sizeof(C2)), 0x00020000);
    C2 a[R][kLeaf];
    bool any_zero = false;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = tid + NT * q;
        const uint32_t vo = r < m ? (uint32_t)r * (uint32_t)sizeof(C2) : 0u;
#pragma unroll
        for (int c = 0; c < kLeaf; ++c) {
            a[q][c] = mkc<C2>(T(0), T(0));
            if (c < w) {  // uniform
                const C2 v = buf_load(rs, vo, (uint32_t)((k0 + c) * ld) * (uint32_t)sizeof(C2), C2());
                if (r < m) a[q][c] = v;
            }
        }
    }
    // the panel is in registers: nothing of it stays outstanding in the
    // vector memory counter, which later stores share
    __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
    for (int c = 0; c < kLeaf; ++c) {
        if (c < w) {  // uniform
        const int bf = c & 1;
        T best = T(-1);
        int bi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int r = tid + NT * q;
            if (r >= c && r < m) {
                const T v = fabs(a[q][c].x) + fabs(a[q][c].y);
                if (v > best) {
                    best = v;
                    bi = r;
                }
            }
        }
        // the wave's maximum, then the first row that attains it; the owner
        // of that row publishes it
        const T wv = wave_reduce<true>(best);
        const int wi = wave_reduce<false>(best == wv ? bi : 0x7fffffff);
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (tid + NT * q == wi) {
#pragma unroll
                for (int j = 0; j < kLeaf; ++j) s_cand[bf][wave][j] = a[q][j];
            }
        if (lane == 0) {
            s_v[bf][wave] = wv;
            s_i[bf][wave] = wi;
        }
        if (tid == c) {
#pragma unroll
            for (int j = 0; j < kLeaf; ++j) s_crow[bf][j] = a[0][j];
        }
        __syncthreads();
        // every wave reduces the wave maxima itself (first row on ties)
        const T bv = s_v[bf][lane % NW];
        const int biv = s_i[bf][lane % NW];
        const T bm = row_reduce<true>(bv);
        const int pm = row_reduce<false>(bv == bm ? biv : 0x7fffffff);
        const T b = read_lane(bm, 0);
        int p = read_lane(pm, 0);
        const bool zero = !(b > T(0));  // exact zero pivot: record, no interchange, no update
        if (zero) p = c;
        const int po = p % NT, pq = p / NT;  // owner thread and register slot of row p
        const C2* prow = s_cand[bf][po >> 6];
        // pivots and the zero-pivot flag stay on chip until the loop ends: a
        // global store here would be waited for (the vector memory counter
        // also counts stores) wherever a register still waits for its load
        if (tid == 0) s_p[c] = p;
        any_zero = any_zero || zero;
        if (p != c) {  // uniform: rows c and p exchange
            if (tid == c) {
#pragma unroll
                for (int j = 0; j < kLeaf; ++j) a[0][j] = prow[j];
            }
            if (tid == po) {
#pragma unroll
                for (int q = 0; q < R; ++q)
                    if (q == pq) {
#pragma unroll
                        for (int j = 0; j < kLeaf; ++j) a[q][j] = s_crow[bf][j];
                    }
            }
        }
        if (!zero) {
            const C2 inv = pivot_recip(prow[c]);
            // the pivot row's remaining entries, read once for all R rows
            T ux[kLeaf], uy[kLeaf];  // scalar arrays: an array of vector structs would go to scratch
#pragma unroll
            for (int j = 1; j < kLeaf; ++j)  // constant trip count: unrolled before c is known
                if (j > c) {
                    const C2 v = prow[j];
                    ux[j] = v.x;
                    uy[j] = v.y;
                }
            // rows above the pivot and outside the panel take multiplier 0, so
            // every row runs the same straight-line update (a - 0 u = a)
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int r = tid + NT * q;
                const bool act = r > c && r < m;
                const C2 lv = cmul(a[q][c], inv);
                // selects per component: a conditional on the vector structs
                // selects between addresses and puts the panel in scratch
                const T lx = act ? lv.x : T(0), ly = act ? lv.y : T(0);
                a[q][c] = mkc<C2>(act ? lv.x : a[q][c].x, act ? lv.y : a[q][c].y);
                const C2 l = mkc<C2>(lx, ly);
#pragma unroll
                for (int j = 1; j < kLeaf; ++j)
                    if (j > c) {  // a -= l u as four fused multiply-adds
                        T x = fma(-l.x, ux[j], a[q][j].x), y = fma(-l.x, uy[j], a[q][j].y);
                        a[q][j] = mkc<C2>(fma(l.y, uy[j], x), fma(-l.y, ux[j], y));
                    }
            }
        }
        }
    }
    if (x_hi > x_lo && tid < 16) {  // uniform condition; L11 strictly below the diagonal, zero elsewhere
#pragma unroll
        for (int j = 0; j < kLeaf && j < 16; ++j)
            s_L[j][tid] = mkc<C2>(j < tid && tid < w ? a[0][j].x : T(0), j < tid && tid < w ? a[0][j].y : T(0));
    }
    if (tid < w) piv[k0 + tid] = k0 + s_p[tid];  // s_p written by lane 0 of this wave
    if (tid == 0 && any_zero) atomicOr(info + blockIdx.x, 1);
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int r = tid + NT * q;
        if (r < m) {
            const uint32_t vo = (uint32_t)r * (uint32_t)sizeof(C2);
            if (w == kLeaf) {  // uniform; straight-line stores
#pragma unroll
                for (int c = 0; c < kLeaf; ++c)
                    buf_store(a[q][c], rs, vo, (uint32_t)((k0 + c) * ld) * (uint32_t)sizeof(C2));
            } else {
#pragma unroll
                for (int c = 0; c < kLeaf; ++c)
                    if (c < w) buf_store(a[q][c], rs, vo, (uint32_t)((k0 + c) * ld) * (uint32_t)sizeof(C2));
            }
        }
    }
    // interchanges on the block column's other columns: the net permutation
    // of the rows they touch (value of row s_rows[s] <- row s_src[s]; slot
    // j < w is row j, slot w + j the pivot row of column j when it lies below
    // the leaf and first appears there, else unused = -1), then every
    // (column, row) gathered before any is written
    const int nleft = (int)(k0 - pc_lo), nright = (int)(pc_hi - (k0 + w));
    if (nleft + nright <= 0) return;
    if (wave == 0) {
        // lane j < w: slot of column j's pivot row; lanes 0..2w-1: slot rows
        const int pj = lane < w ? s_p[lane] : 0;
        int first = lane;
        if (lane < w && pj >= w)
            for (int i = 0; i < lane; ++i)
                if (s_p[i] == pj) {
                    first = i;
                    break;
                }
        const int bslot = lane < w ? (pj < w ? pj : w + first) : 0;
        const int jw = (lane >= w && lane < 2 * w) ? lane - w : 0;
        const int pr = __shfl(pj, jw);  // every lane takes part in the shuffles
        const int fr = __shfl(first, jw);
        const int row = lane < w ? lane : ((lane < 2 * w && pr >= w && fr == jw) ? pr : -1);
        // labels: the original row whose value each slot holds, interchanged in order
        int lab = row;
        for (int j = 0; j < w; ++j) {
            const int b = __shfl(bslot, j);
            const int va = __shfl(lab, j), vb = __shfl(lab, b);
            if (b != j) lab = lane == j ? vb : (lane == b ? va : lab);
        }
        if (lane < 2 * kLeaf) {
            s_rows[lane] = lane < 2 * w ? row : -1;
            s_src[lane] = lane < 2 * w ? lab : -1;
        }
    }
    __syncthreads();
    const int nr = 2 * w;
    const int items = (nleft + nright) * nr;  // <= kNBOuter * 2 * kLeaf
    constexpr int kItems = (kNBOuter * 2 * kLeaf + NT - 1) / NT;
    C2 g[kItems];
    uint32_t vd[kItems];  // destination byte offsets, computed before any store
#pragma unroll
    for (int e = 0; e < kItems; ++e) {
        const int it = tid + NT * e;
        vd[e] = 0xffffffffu;  // nothing to store
        g[e] = mkc<C2>(T(0), T(0));
        if (it < items) {
            const int ci = it / nr, si = it % nr;
            const int64_t col = ci < nleft ? pc_lo + ci : k0 + w + (ci - nleft);
            const int sr = s_src[si], rr = s_rows[si];
            const uint32_t cb = (uint32_t)(col * ld) * (uint32_t)sizeof(C2);
            if (sr >= 0 && rr >= 0) {
                g[e] = buf_load(rs, (uint32_t)sr * (uint32_t)sizeof(C2) + cb, 0u, C2());
                vd[e] = (uint32_t)rr * (uint32_t)sizeof(C2) + cb;
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0);  // every gather has returned before any row is overwritten
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kItems; ++e)
        if (vd[e] != 0xffffffffu) buf_store(g[e], rs, vd[e], 0u);
    // optional: X = L11^{-1} X on rows [k0, k0+w) of the (at most kLeafX)
    // columns [x_lo, x_hi) right of the leaf, once the interchanges above have
    // reached them -- the recursion's TRSM between two leaves, without a launch
    if (kLeaf <= 16 && x_hi > x_lo) {  // uniform
        __syncthreads();  // the rows written above are visible to the workgroup
        if (tid < 16 * kLeafX) {  // waves 0-3: one column per row of 16 lanes
            const int gr = tid & 15;
            const int64_t col = x_lo + (tid >> 4);
            const bool ok = col < x_hi && gr < w;
            // per-lane byte offset (row + column; the column differs between
            // the rows of 16 lanes, so it cannot be the wave-uniform offset)
            const uint32_t vo = ok ? (uint32_t)(gr + col * ld) * (uint32_t)sizeof(C2) : 0u;
            const C2 v = buf_load(rs, vo, 0u, C2());
            T xr = ok ? v.x : T(0), xi = ok ? v.y : T(0);
            LeafXStep<T, 0, (kLeaf < 16 ? kLeaf : 16)>::run(xr, xi, s_L, gr, w);
            if (ok) buf_store(mkc<C2>(xr, xi), rs, vo, 0u);
        }
    }
}


// ---------------------------------------------------------------------------
// row interchanges of pivots [k0, k0+npiv) on columns [c_lo, c_hi)
// ---------------------------------------------------------------------------
template <class C2>
__global__ __launch_bounds__(256) void k_lu_swap(C2* __restrict__ A, int64_t ld, int64_t stride,
                                                 int64_t n, const int64_t* __restrict__ ipiv,
                                                 int64_t k0, int npiv, int64_t c_lo, int64_t c_hi) {
    C2* M = A + (int64_t)blockIdx.y * stride;
    const int64_t* piv = ipiv + (int64_t)blockIdx.y * n;
    __shared__ int64_t sp[kNBOuter];
    for (int t = threadIdx.x; t < npiv; t += 256) sp[t] = piv[k0 + t];
    __syncthreads();
    const int64_t j = c_lo + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= c_hi) return;
    C2* col = M + j * ld;
    for (int t = 0; t < npiv; ++t) {
        const int64_t p = sp[t];
        if (p != k0 + t) {
            const C2 x = col[k0 + t];
            col[k0 + t] = col[p];
            col[p] = x;
        }
    }
}

// Split operand images of the mixed LU's trailing updates (layout with
// k_split below): 16-byte units of 8 halves, hi/lo planes, and the
// per-row or per-column scaling exponent.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
constexpr int kHK = 16;                      // complex depths per chunk
constexpr int kHChunks = kNBOuter / kHK;     // chunks of a full block column
constexpr int kHUnits = kHChunks * 2 * 4;    // image units per row or column
constexpr int kHEMax = 14;                   // largest |x| of a row or column scaled into [2^13, 2^14)

__device__ __forceinline__ void split_h(float xs, _Float16& hi, _Float16& lo) {
    hi = (_Float16)xs;
    lo = (_Float16)(xs - (float)hi);  // exact in FP32
}

// exponent scaling a row's or column's largest magnitude mx into [2^13, 2^14)
__device__ __forceinline__ int split_exp(float mx) {
    if (!(mx > 0.0f)) return 0;
    int E;
    (void)frexpf(mx, &E);  // mx = m 2^E, m in [0.5, 1)
    return kHEMax - E;
}

__device__ __forceinline__ bool finite2(float2 v) {
    return ::fabsf(v.x) <= 3.402823466e38f && ::fabsf(v.y) <= 3.402823466e38f;  // false for NaN
}

// ---------------------------------------------------------------------------
// X = L11^{-1} X in one launch for a whole block column: L11 = unit lower
// block at (k0, k0) of size w <= kNBOuter, X = rows [k0, k0+w) of columns
// [c_lo, c_hi).  A workgroup solves kTbCols columns of one matrix: the 16
// lanes of a DPP row hold one column, lane g rows g + 16 i (i < 8), in
// registers.  Forward substitution one unknown at a time, fully unrolled so
// every register index is a constant: x_c is broadcast within the row
// (row_newbcast), then every lane subtracts L(r, c) x_c from its rows below
// c; L11's columns are staged in LDS kTbStage at a time (zero on and above
// the diagonal and beyond w), the next block's loads in flight while the
// current block's steps run.  One launch instead of a 32-row blocked
// TRSM + GEMM sequence (7 launches for a 128-wide block column).
// ---------------------------------------------------------------------------
constexpr int kTbCols = 16, kTbStage = 32, kTbSlots = kNBOuter / 16;
template <class T, int C>
struct TbStep {  // unknown C of the substitution
    typedef typename Cx<T>::t C2;
    __device__ __forceinline__ static void run(T (&xr)[kTbSlots], T (&xi)[kTbSlots], const C2 (*sL)[kNBOuter + 1],
                                               int g) {
        constexpr int ic = C / 16, gc = C % 16, cc = C % kTbStage;
        const T br = row_bcast<gc>(xr[ic]), bi = row_bcast<gc>(xi[ic]);
#pragma unroll
        for (int i = ic; i < kTbSlots; ++i) {
            const C2 l = sL[cc][g + 16 * i];  // zero for rows r <= C (and r >= w)
            xr[i] = fma(-l.x, br, fma(l.y, bi, xr[i]));
            xi[i] = fma(-l.x, bi, fma(-l.y, br, xi[i]));
        }
    }
};
template <class T, int C0, int N>
struct TbStage {  // unknowns C0 .. C0 + N - 1 (one staged block), uniform bound w
    typedef typename Cx<T>::t C2;
    __device__ __forceinline__ static void run(T (&xr)[kTbSlots], T (&xi)[kTbSlots], const C2 (*sL)[kNBOuter + 1],
                                               int g, int w) {
        if (C0 < w) {
            TbStep<T, C0>::run(xr, xi, sL, g);
            TbStage<T, C0 + 1, N - 1>::run(xr, xi, sL, g, w);
        }
    }
};
template <class T, int C0>
struct TbStage<T, C0, 0> {
    typedef typename Cx<T>::t C2;
    __device__ __forceinline__ static void run(T (&)[kTbSlots], T (&)[kTbSlots], const C2 (*)[kNBOuter + 1], int,
                                               int) {}
};
// L11 columns S .. S + kTbStage - 1 (rows below the diagonal, rows and
// columns below w; zero elsewhere) into registers: kTbPer entries per thread,
// every load issued before any is waited for (addresses outside the block
// clamped to its first entry, values replaced by zero)
constexpr int kTbPer = kTbStage * kNBOuter / 256;
template <class T, int S>
__device__ __forceinline__ void tb_load(T (&tx)[kTbPer], T (&ty)[kTbPer], const typename Cx<T>::t* __restrict__ M,
                                        int64_t ld, int64_t k0, int w, int tid) {
    typedef typename Cx<T>::t C2;
    const int r = tid % kNBOuter;
#pragma unroll
    for (int e = 0; e < kTbPer; ++e) {
        const int c = S + 2 * e + tid / kNBOuter;  // tid + 256 e = (c - S) kNBOuter + r
        const bool in = c < w && r < w && r > c;
        const C2 v = M[(k0 + (in ? r : 0)) + (k0 + (in ? c : 0)) * ld];
        tx[e] = in ? v.x : T(0);
        ty[e] = in ? v.y : T(0);
    }
}
template <class T, int S>
__device__ __forceinline__ void tb_stage(T (&xr)[kTbSlots], T (&xi)[kTbSlots], typename Cx<T>::t (*sL)[kNBOuter + 1],
                                         T (&tx)[kTbPer], T (&ty)[kTbPer], const typename Cx<T>::t* __restrict__ M,
                                         int64_t ld, int64_t k0, int w, int tid, int g) {
    typedef typename Cx<T>::t C2;
    if (S >= w) return;  // uniform
    __syncthreads();     // the previous block's reads of sL are done
#pragma unroll
    for (int e = 0; e < kTbPer; ++e) sL[2 * e + tid / kNBOuter][tid % kNBOuter] = mkc<C2>(tx[e], ty[e]);
    __syncthreads();
    if (S + kTbStage < w) tb_load<T, S + kTbStage>(tx, ty, M, ld, k0, w, tid);  // next block, during this one
    TbStage<T, S, kTbStage>::run(xr, xi, sL, g, w);
}
//
// IMG (FP32 only): the solved columns are the U12 operand of a split-image
// trailing update; the epilogue also writes their image units and column
// exponents (the values k_split<false> would write from them) from X staged
// in LDS, so the update needs no separate split pass over U12.
template <class T, bool IMG>
__global__ __launch_bounds__(256) void k_lu_trsm_blk(typename Cx<T>::t* __restrict__ A, int64_t ld, int64_t stride,
                                                     int64_t k0, int w, int64_t c_lo, int64_t c_hi,
                                                     u4v* __restrict__ img, int64_t img_stride, int64_t pad,
                                                     int* __restrict__ xexp, int64_t xexp_stride,
                                                     int* __restrict__ info) {
    static_assert(!IMG || sizeof(T) == 4, "split images are FP32 only");
    // panel-chain kernel: its waves win issue arbitration against those of
    // a trailing-update GEMM sharing the CU (the chain is the LU's critical
    // path)
    __builtin_amdgcn_s_setprio(3);
    typedef typename Cx<T>::t C2;
    static_assert(kNBOuter == 128 && kTbStage == 32, "four staged blocks of 32 unknowns");
    C2* M = A + (int64_t)blockIdx.y * stride;
    __shared__ C2 sL[kTbStage][kNBOuter + 1];  // sL[c - S][r] = L11(r, c)
    const int tid = threadIdx.x, g = tid & 15;
    const int64_t j = c_lo + (int64_t)blockIdx.x * kTbCols + (tid >> 4);
    const bool ok = j < c_hi;
    C2* col = M + (ok ? j : c_lo) * ld + k0;
    T xr[kTbSlots], xi[kTbSlots];
#pragma unroll
    for (int i = 0; i < kTbSlots; ++i) {
        const int r = g + 16 * i;
        const C2 v = (ok && r < w) ? col[r] : mkc<C2>(T(0), T(0));
        xr[i] = v.x;
        xi[i] = v.y;
    }
    T tx[kTbPer], ty[kTbPer];
    tb_load<T, 0>(tx, ty, M, ld, k0, w, tid);
    // every lane takes part in every step (the row broadcasts read all 16 lanes)
    tb_stage<T, 0>(xr, xi, sL, tx, ty, M, ld, k0, w, tid, g);
    tb_stage<T, 32>(xr, xi, sL, tx, ty, M, ld, k0, w, tid, g);
    tb_stage<T, 64>(xr, xi, sL, tx, ty, M, ld, k0, w, tid, g);
    tb_stage<T, 96>(xr, xi, sL, tx, ty, M, ld, k0, w, tid, g);
    if (ok) {
#pragma unroll
        for (int i = 0; i < kTbSlots; ++i) {
            const int r = g + 16 * i;
            if (r < w) col[r] = mkc<C2>(xr[i], xi[i]);
        }
    }
    if constexpr (IMG) {
        // the column's largest |re| or |im| and non-finite flag over its 16
        // lanes (rows from w on hold zeros)
        float mx = 0.0f, bad = 0.0f;
#pragma unroll
        for (int i = 0; i < kTbSlots; ++i) {
            if (!finite2(make_float2(xr[i], xi[i]))) bad = 1.0f;
            mx = fmaxf(mx, fmaxf(::fabsf(xr[i]), ::fabsf(xi[i])));
        }
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, o));
            bad = fmaxf(bad, __shfl_xor(bad, o));
        }
        __shared__ int sE[kTbCols];
        __syncthreads();  // the last stage's reads of sL are done: X goes there, sL[column][row]
        const int jl = tid >> 4;
#pragma unroll
        for (int i = 0; i < kTbSlots; ++i) sL[jl][g + 16 * i] = mkc<C2>(xr[i], xi[i]);
        if (g == 0) {
            int e = split_exp(mx);
            if (bad != 0.0f) {
                if (ok) atomicOr(info + blockIdx.y, 2);
                e = 0;
            }
            sE[jl] = e;
            if (ok) xexp[(int64_t)blockIdx.y * xexp_stride + j] = e;
        }
        __syncthreads();
        // as k_split<false>: lane (column xl, unit q) writes the 8 depths of
        // unit q of chunks c and c + 4 (wave c), hi and lo planes
        const int xl = (tid >> 2) & 15, q = tid & 3;
        const int64_t x = c_lo + (int64_t)blockIdx.x * kTbCols + xl;
        if (x < c_hi) {
            const int e = sE[xl];
            u4v* out = img + (int64_t)blockIdx.y * img_stride;
#pragma unroll
            for (int h2 = 0; h2 < 2; ++h2) {
                const int c = (tid >> 6) + 4 * h2;
                if (kHK * c < w) {
                    const int d0 = kHK * c + 8 * (q & 1);
                    h8 hi, lo;
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        const C2 z = sL[xl][d0 + t];
                        _Float16 a0, a1;
                        split_h(__builtin_amdgcn_ldexpf((q & 2) ? z.y : z.x, e), a0, a1);
                        hi[t] = a0;
                        lo[t] = a1;
                    }
                    out[((int64_t)(2 * c) * pad + x) * 4 + q] = __builtin_bit_cast(u4v, hi);
                    out[((int64_t)(2 * c + 1) * pad + x) * 4 + q] = __builtin_bit_cast(u4v, lo);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// C[r_lo:r_hi, c_lo:c_hi] -= A[r_lo:r_hi, k_lo:k_lo+kw] * A[k_lo:k_lo+kw, c_lo:c_hi]
// 64 x 64 complex tile per 256-thread workgroup; wave (wr, wc) owns a 32 x 32
// quarter as 2 x 2 MFMA tiles of 16 x 16, real and imaginary accumulators.
// v_mfma_f64_16x16x4f64 lane layout: A[i = l%16][k = l/16],
// B[k = l/16][j = l%16], D register r -> D[4r + l/16][l%16].
// ---------------------------------------------------------------------------
constexpr int kKC = 16;  // depth per LDS stage

template <class T>
__global__ __launch_bounds__(256) void k_lu_gemm(typename Cx<T>::t* __restrict__ A, int64_t ld, int64_t stride,
                                                 int64_t r_lo, int64_t r_hi, int64_t c_lo,
                                                 int64_t c_hi, int64_t k_lo, int kw) {
    // panel-chain kernel: its waves win issue arbitration against those of
    // a trailing-update GEMM sharing the CU (the chain is the LU's critical
    // path)
    __builtin_amdgcn_s_setprio(3);
    typedef typename Cx<T>::t C2;
    typedef typename Cx<T>::acc acc_t;
    C2* M = A + (int64_t)blockIdx.z * stride;
    const int64_t tr = r_lo + (int64_t)blockIdx.x * 64;
    const int64_t tc = c_lo + (int64_t)blockIdx.y * 64;
    __shared__ C2 sA[kKC][64];  // sA[k][row]
    __shared__ C2 sB[kKC][64];  // sB[k][col]
    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;
    acc_t accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = (acc_t){T(0), T(0), T(0), T(0)};
            acci[a][b] = (acc_t){T(0), T(0), T(0), T(0)};
        }
    for (int kc = 0; kc < kw; kc += kKC) {
        // every operand load issued before any is waited for: entries outside
        // the ranges load from a clamped address (the first entry of the
        // operand) and are replaced by zero; scalar arrays stay in registers
        T ax[4], ay[4], bx[4], by[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = tid + 256 * t;
            {
                const int r = q & 63, k = q >> 6;
                const int64_t gr = tr + r;
                const bool in = gr < r_hi && kc + k < kw;
                const C2 v = M[(in ? gr : r_lo) + (k_lo + (in ? kc + k : 0)) * ld];
                ax[t] = in ? v.x : T(0);
                ay[t] = in ? v.y : T(0);
            }
            {
                const int k = q & 15, c = q >> 4;
                const int64_t gc = tc + c;
                const bool in = gc < c_hi && kc + k < kw;
                const C2 v = M[(k_lo + (in ? kc + k : 0)) + (in ? gc : c_lo) * ld];
                bx[t] = in ? v.x : T(0);
                by[t] = in ? v.y : T(0);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = tid + 256 * t;
            sA[q >> 6][q & 63] = mkc<C2>(ax[t], ay[t]);
            sB[q & 15][q >> 4] = mkc<C2>(bx[t], by[t]);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKC; kk += 4) {
            C2 av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = sA[kk + (l >> 4)][32 * wr + 16 * a + (l & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = sB[kk + (l >> 4)][32 * wc + 16 * b + (l & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    // operands swapped: the MFMA computes the transposed block, so
                    // the 16 lanes of a result row hold 16 consecutive rows of C
                    // (coalesced column-major read-modify-write in the epilogue)
                    accr[a][b] = Cx<T>::mfma(bv[b].x, av[a].x, accr[a][b]);
                    accr[a][b] = Cx<T>::mfma(-bv[b].y, av[a].y, accr[a][b]);
                    acci[a][b] = Cx<T>::mfma(bv[b].y, av[a].x, acci[a][b]);
                    acci[a][b] = Cx<T>::mfma(bv[b].x, av[a].y, acci[a][b]);
                }
        }
        __syncthreads();
    }
    // C -= acc: every entry of this lane read before any is written (a store
    // between two loads would make each load wait for the store as well)
    C2 cv[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gr = tr + 32 * wr + 16 * a + (l & 15);
                const int64_t gc = tc + 32 * wc + 16 * b + Cx<T>::row(r, l);
                const bool in = gr < r_hi && gc < c_hi;  // clamped address outside: no branch
                cv[a][b][r] = M[(in ? gr : r_lo) + (in ? gc : c_lo) * ld];
            }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gr = tr + 32 * wr + 16 * a + (l & 15);
                const int64_t gc = tc + 32 * wc + 16 * b + Cx<T>::row(r, l);
                if (gr < r_hi && gc < c_hi)
                    M[gr + gc * ld] = mkc<C2>(cv[a][b][r].x - accr[a][b][r], cv[a][b][r].y - acci[a][b][r]);
            }
}

// Large trailing updates: 128 x 128 tile per 512-thread workgroup (8 waves,
// each 64 x 32 = 4 x 2 MFMA blocks), the next depth chunk loaded into
// registers while the current one is multiplied out of LDS.  Twice the
// arithmetic intensity of the 64 x 64 kernel per byte staged.
constexpr int kBT = 128;
template <class T>
__global__ __launch_bounds__(512) void k_lu_gemm_big(typename Cx<T>::t* __restrict__ A, int64_t ld, int64_t stride,
                                                     int64_t r_lo, int64_t r_hi, int64_t c_lo, int64_t c_hi,
                                                     int64_t k_lo, int kw) {
    typedef typename Cx<T>::t C2;
    typedef typename Cx<T>::acc acc_t;
    C2* M = A + (int64_t)blockIdx.z * stride;
    const int64_t tr = r_lo + (int64_t)blockIdx.x * kBT;
    const int64_t tc = c_lo + (int64_t)blockIdx.y * kBT;
    __shared__ C2 sA[kKC][kBT];      // sA[k][row]
    __shared__ C2 sB[kKC][kBT + 1];  // sB[k][col], padded against store conflicts
    const int tid = threadIdx.x, l = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wc = wave >> 1;  // 2 x 4 waves: 64 rows x 32 columns each
    // four real products per complex product (re = Ar Br - Ai Bi,
    // im = Ar Bi + Ai Br); Gauss's three-product form measured slower here
    acc_t accr[4][2], acci[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = (acc_t){T(0), T(0), T(0), T(0)};
            acci[a][b] = (acc_t){T(0), T(0), T(0), T(0)};
        }
    C2 pa[4], pb[4];
    auto gload = [&](int kc) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = tid + 512 * t;
            {
                const int r = q & (kBT - 1), k = q >> 7;
                const int64_t gr = tr + r;
                pa[t] = (gr < r_hi && kc + k < kw) ? M[gr + (k_lo + kc + k) * ld] : mkc<C2>(T(0), T(0));
            }
            {
                const int k = q & (kKC - 1), c = q >> 4;
                const int64_t gc = tc + c;
                pb[t] = (gc < c_hi && kc + k < kw) ? M[(k_lo + kc + k) + gc * ld] : mkc<C2>(T(0), T(0));
            }
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int q = tid + 512 * t;
            sA[q >> 7][q & (kBT - 1)] = pa[t];
            sB[q & (kKC - 1)][q >> 4] = pb[t];
        }
    };
    gload(0);
    lstore();
    __syncthreads();
    for (int kc = 0; kc < kw; kc += kKC) {
        const bool more = kc + kKC < kw;
        if (more) gload(kc + kKC);
#pragma unroll
        for (int kk = 0; kk < kKC; kk += 4) {
            C2 av[4], bv[2];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = sA[kk + (l >> 4)][64 * wr + 16 * a + (l & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = sB[kk + (l >> 4)][32 * wc + 16 * b + (l & 15)];
            // the two products into each accumulator are 8 MFMAs apart
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = Cx<T>::mfma(bv[b].x, av[a].x, accr[a][b]);
                    acci[a][b] = Cx<T>::mfma(bv[b].y, av[a].x, acci[a][b]);
                }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accr[a][b] = Cx<T>::mfma(-bv[b].y, av[a].y, accr[a][b]);
                    acci[a][b] = Cx<T>::mfma(bv[b].x, av[a].y, acci[a][b]);
                }
        }
        __syncthreads();
        if (more) {
            lstore();
            __syncthreads();
        }
    }
    // C -= acc, transposed blocks (operands swapped): lanes run along rows of
    // C.  Per row block, its eight entries of this lane are read before any is
    // written (a store between two loads would make each load wait for the
    // store as well).
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int64_t gr = tr + 64 * wr + 16 * a + (l & 15);
        C2 cv[2][4];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gc = tc + 32 * wc + 16 * b + Cx<T>::row(r, l);
                cv[b][r] = (gr < r_hi && gc < c_hi) ? M[gr + gc * ld] : mkc<C2>(T(0), T(0));
            }
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gc = tc + 32 * wc + 16 * b + Cx<T>::row(r, l);
                if (gr < r_hi && gc < c_hi)
                    M[gr + gc * ld] = mkc<C2>(cv[b][r].x - accr[a][b][r], cv[b][r].y - acci[a][b][r]);
            }
    }
}

// FP32 trailing updates of the mixed-precision LU on half-precision matrix
// cores, C -= L21 U12 with L21 = A[r, k_lo:k_lo+kw] and U12 = A[k_lo:k_lo+kw, c].
//
// Operand images.  Once per block column, L21 and U12 are scaled (row i of
// L21 by 2^er_i, column j of U12 by 2^ec_j, powers of two chosen so that the
// row's or column's largest |re| or |im| lands in [2^13, 2^14)) and split as
// x = hi + lo (hi = fp16(x), lo = fp16(x - hi): 22 significant bits
// together).  Each real product is then formed as hi*hi + hi*lo + lo*hi with
// FP32 accumulation and the sum scaled back by 2^-(er_i + ec_j), which is
// exact.  The dropped lo*lo term is below 2^-22 of the product, the size of
// the FP32 accumulation error over a 128-deep update; an entry more than
// 2^17 below its row's or column's maximum has a subnormal lo, an absolute
// error under 2^-39 of that maximum.  No finite entry is out of range; a
// non-finite one sets bit 1 of info[item].
//
// The images hold 16-byte units of 8 halves.  Per item, for depth chunk c
// (16 complex depths), hi/lo plane h and row (or column) x:
//   unit ((c * 2 + h) * pad + x) * 4 + q,
//   q = 0: re of depths 0-7, 1: re 8-15, 2: im 0-7, 3: im 8-15,
// so one chunk of one 128-row tile is two contiguous 8 KB blocks.  The L21
// image of block column k sits in one of two slots (k mod 2: the next
// panel's image is written while the wide update of this one still reads
// its own); the U12 image is indexed by absolute column, since the narrow
// and the wide update cover disjoint columns.
//
// The complex product folds into one real product of depth 32 per chunk:
// with the depths of a chunk laid out as [re (16) | im (16)],
//   Re C = [Ar | Ai] . [Br ; -Bi],   Im C = [Ar | Ai] . [Bi ; Br]
// (the -Bi half negated in registers), one v_mfma_f32_16x16x32_f16 per
// 16 x 16 block and split term: six per block and chunk.  Lane layout of
// that instruction: A[m = l%16][k = 8 (l/16) + e], B[k = 8 (l/16) + e]
// [n = l%16], result register r is row 4 (l/16) + r.  As in k_lu_gemm_big
// the tile's columns are the instruction's rows, so the results run along
// columns of C.
// Operand images: 16 rows of L21 (ROWS) or columns of U12 per 512-thread
// workgroup; wave c takes depth chunk c, lane (row or column x, unit q) the
// 8 depths of image unit q (re or im), so each store of a wave writes 1 KB
// of one plane contiguously.  The maximum of a row (column) is reduced over
// its four units in registers and over the chunks through LDS.  L21: rows
// [x0, x1) with x1 = npad, rows from n on written as zeros; U12: columns
// [x0, x1).
template <bool ROWS>
__global__ __launch_bounds__(512) void k_split(const float2* __restrict__ A, int64_t ld, int64_t stride, int64_t n,
                                               int64_t x0, int64_t x1, int64_t pad, int64_t k_lo, int kw,
                                               u4v* __restrict__ img, int64_t img_stride, int* __restrict__ xexp,
                                               int64_t xexp_stride, int* __restrict__ info) {
    const int64_t item = blockIdx.y;
    const int tid = threadIdx.x, c = tid >> 6, xl = (tid >> 2) & 15, q = tid & 3;
    const int64_t x = x0 + (int64_t)blockIdx.x * 16 + xl;
    const bool out_ok = x < x1, live = out_ok && (!ROWS || x < n);
    const int d0 = kHK * c + 8 * (q & 1);
    const float2* M = A + item * stride;
    float v[8];
    float mx = 0.0f, bad = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int64_t d = k_lo + d0 + e;
        float2 z = make_float2(0.f, 0.f);
        if (live && d0 + e < kw) z = ROWS ? M[x + d * ld] : M[d + x * ld];
        v[e] = (q & 2) ? z.y : z.x;
        if (!finite2(z)) bad = 1.0f;
        mx = fmaxf(mx, fmaxf(::fabsf(z.x), ::fabsf(z.y)));
    }
    mx = fmaxf(mx, __shfl_xor(mx, 1));  // units q and q ^ 1 cover the chunk's 16 depths
    bad = fmaxf(bad, __shfl_xor(bad, 1));
    __shared__ float smx[8][16], sbad[8][16];
    if (q == 0) {
        smx[c][xl] = mx;
        sbad[c][xl] = bad;
    }
    __syncthreads();
    mx = 0.0f;
    bad = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        mx = fmaxf(mx, smx[t][xl]);
        bad = fmaxf(bad, sbad[t][xl]);
    }
    if (!out_ok) return;
    int e = split_exp(mx);
    if (bad != 0.0f) {
        if (c == 0 && q == 0) atomicOr(info + item, 2);
        e = 0;
    }
    if (kHK * c < kw) {
        h8 hi, lo;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            _Float16 a0, a1;
            split_h(__builtin_amdgcn_ldexpf(v[t], e), a0, a1);
            hi[t] = a0;
            lo[t] = a1;
        }
        u4v* out = img + item * img_stride;
        out[((int64_t)(2 * c) * pad + x) * 4 + q] = __builtin_bit_cast(u4v, hi);
        out[((int64_t)(2 * c + 1) * pad + x) * 4 + q] = __builtin_bit_cast(u4v, lo);
    }
    if (c == 0 && q == 0) xexp[item * xexp_stride + x] = e;
}

constexpr int kXcd = 8;  // XCDs: workgroup w of a launch runs on XCD w % 8

// LDS slot of unit q of tile row (or column) x: the four units of a row are
// permuted by the XOR key (x >> 2) & 3.  Rows x, x + 4, x + 8, x + 12 share
// banks (64 banks of 4 bytes = four 64-byte rows); the key gives them four
// different slots, and rows x and x + 4 read at units q and q ^ 2 (the two
// halves of a wave) different slots as well.
__device__ __forceinline__ int hkey(int x) { return (x >> 2) & 3; }
__device__ __forceinline__ int hslot(int x, int q) { return x * 4 + (q ^ hkey(x)); }

typedef __attribute__((address_space(3))) void* lptr_t;

// Copies chunk c of a tile's L21 and U12 images into LDS (four 16-byte
// buffer loads per lane straight to LDS, no registers): LDS unit p = (h 128
// + x) 4 + s holds image unit q = s ^ key(x) of row (column) x, plane h.
// vo = (4 x + q) 16 is the lane's byte offset within a plane of the tile;
// planes lie npad (ncpad) rows apart, so the chunk and plane select a
// wave-uniform offset.
__device__ __forceinline__ void h_stage(__amdgpu_buffer_rsrc_t rl, __amdgpu_buffer_rsrc_t ru, int64_t npad,
                                        int64_t ncpad, int c, u4v* sl, u4v* su, uint32_t vo, int wave) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t pl = (uint32_t)((2 * c + h) * npad * 64), pu = (uint32_t)((2 * c + h) * ncpad * 64);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rl, (lptr_t)(sl + 512 * h + 64 * wave), 16, vo, pl, 0, 0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(ru, (lptr_t)(su + 512 * h + 64 * wave), 16, vo, pu, 0, 0);
    }
}

// One chunk: 48 matrix-core products per wave out of one LDS stage.
__device__ __forceinline__ void h_chunk(const u4v* sl, const u4v* su, int l, int wr, int wc, f4 (&accr)[4][2],
                                        f4 (&acci)[4][2]) {
    const int g = l >> 4;
    const unsigned neg = g >= 2 ? 0x80008000u : 0u;  // -Bi in the upper half of [Br ; -Bi]
    const u4v m = (u4v){neg, neg, neg, neg};
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int col = 32 * wc + 16 * b + (l & 15);
        const h8 b1h = __builtin_bit_cast(h8, su[hslot(col, g)] ^ m);
        const h8 b1l = __builtin_bit_cast(h8, su[512 + hslot(col, g)] ^ m);
        const h8 b2h = __builtin_bit_cast(h8, su[hslot(col, g ^ 2)]);
        const h8 b2l = __builtin_bit_cast(h8, su[512 + hslot(col, g ^ 2)]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int row = 64 * wr + 16 * a + (l & 15);
            const h8 ah = __builtin_bit_cast(h8, sl[hslot(row, g)]);
            const h8 al = __builtin_bit_cast(h8, sl[512 + hslot(row, g)]);
            accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b1h, ah, accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b2h, ah, acci[a][b], 0, 0, 0);
            accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b1l, ah, accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b2l, ah, acci[a][b], 0, 0, 0);
            accr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b1h, al, accr[a][b], 0, 0, 0);
            acci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b2h, al, acci[a][b], 0, 0, 0);
        }
        // the second column block's fragments are not read ahead into this
        // block's products (which would need 16 more VGPRs than 128 allow)
        if (b == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// Workgroup barriers of the depth loop, without the release fence of
// __syncthreads (which would wait for every load in flight to LDS, those
// into the other stage included).  The first waits for this wave's loads of
// the stage about to be read (all but the four of the next chunk, or all),
// the second for its LDS reads, which the next loads into the same stage
// must not overtake.  The waits are compiler builtins, so the compiler's
// own wait insertion knows about them; the memory clobber keeps it from
// moving LDS accesses across the barrier.
__device__ __forceinline__ void h_barrier_ready(bool next_in_flight) {
    if (next_in_flight)
        __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt(4)
    else
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    asm volatile("s_barrier" ::: "memory");
}
__device__ __forceinline__ void h_barrier_done() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    asm volatile("s_barrier" ::: "memory");
}

// Tiles of a launch: (item, column tile, row tile) order, the range
// [x per, (x + 1) per) to XCD x (workgroup w runs on XCD w % 8), which keeps
// each item's operand tiles in one XCD's L2.
__device__ __forceinline__ void h_coords(int t, int ntr, int ntc, int& item, int& ti, int& tj) {
    item = t / (ntr * ntc);
    const int rem = t - item * (ntr * ntc);
    tj = rem / ntr;
    ti = rem - tj * ntr;
}

// Arguments of the split-image update; the host checks that every index
// and byte offset within an item fits 31 bits.
// Per item: kHUnits npad (ncpad) image units and npad (ncpad) exponents.
struct HArgs {
    float2* A;
    int64_t stride;
    int ld, r_lo, r_hi, c_lo, c_hi;
    const u4v* Limg;
    const int* rexp;
    int npad;
    const u4v* Uimg;
    const int* cexp;
    int ncpad;
    int ntr, ntc, ntiles;
    int chain;  // 1: the narrow update on the panel chain (raised wave priority)
};

// C -= 2^-(er + ec) acc for one tile, one block of 16 rows (a) at a time,
// through buffer loads and stores with 32-bit offsets in the item's
// matrix; an entry outside [r_lo, r_hi) x [c_lo, c_hi) gets the offset
// 2^31, beyond the buffer's range (below 2^31 bytes): the hardware's range
// check returns zero for its load and drops its store, with no branch.  Block
// a + 1's loads are issued before block a's stores, so waiting for them
// does not wait for stores.
// C is read and written once per update: non-temporal accesses (cache
// policy bit NT), so its stream does not evict the operand images, which
// every tile of a row (column) reads again, from the XCD's L2
constexpr int kHNt = 2;
// Copies rows [0, 16) x columns [0, 32) of this wave's C block (rows tr +
// 64 wr, columns tc + 32 wc) into 4 KB of LDS (column x at 128 x bytes, 16
// rows of 8 bytes), four 16-byte buffer loads per lane straight to LDS;
// columns beyond c_hi read zeros.  Every row of the tile is inside (so no
// 16-byte load straddles the end of the matrix).
__device__ __forceinline__ void h_cstage(const HArgs& h, __amdgpu_buffer_rsrc_t rc, int tr, int tc, int wr, int wc,
                                         int lane, u4v* region) {
    const int gr = tr + 64 * wr + 2 * (lane & 7);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gc = tc + 32 * wc + 8 * i + (lane >> 3);
        const uint32_t vo = gc < h.c_hi ? ((uint32_t)gr + (uint32_t)gc * (uint32_t)h.ld) * 8u : 0x80000000u;
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rc, (lptr_t)(region + 64 * i), 16, vo, 0, 0, kHNt);
    }
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t h_crsrc(const HArgs& h, int item) {
    return __builtin_amdgcn_make_buffer_rsrc(h.A + item * h.stride, (short)0,
                                             (int)(((uint32_t)(h.c_hi - 1) * (uint32_t)h.ld + (uint32_t)h.r_hi) * 8u),
                                             0x00020000);
}

// PRE: this wave's C block 0 is in LDS at c0 (h_cstage); else loaded here.
// acc *= 2^-(er + ec) for this lane's entries (rows 16 a + l % 16, columns
// 16 b + 4 (l / 16) + r of the wave's block), once the operand fragments are
// dead, so the epilogue holds no exponents.
// The tile's exponents (rows, then columns) were copied to sE by h_estage.
__device__ __forceinline__ void h_unscale(const int* sE, int wr, int wc, int l, f4 (&accr)[4][2],
                                          f4 (&acci)[4][2]) {
    int er[4], ec[2][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) er[a] = sE[64 * wr + 16 * a + (l & 15)];
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) ec[b][r] = sE[kBT + 32 * wc + 16 * b + 4 * (l >> 4) + r];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int sh = -(er[a] + ec[b][r]);
                accr[a][b][r] = __builtin_amdgcn_ldexpf(accr[a][b][r], sh);
                acci[a][b][r] = __builtin_amdgcn_ldexpf(acci[a][b][r], sh);
            }
}

// Copies the exponents of the tile's 128 rows and 128 columns into sE (a
// 4-byte buffer load per lane straight to LDS: two by wave 0, two by wave 1),
// issued before the first operand chunk, so they have landed long before
// the depth loop ends; the loop's barriers make them visible to every wave.
__device__ __forceinline__ void h_estage(const HArgs& h, int item, int tr, int tc, int wave, int lane, int* sE) {
    if (wave > 1) return;
    const int* e = wave == 0 ? h.rexp + (int64_t)item * h.npad + tr : h.cexp + (int64_t)item * h.ncpad + tc;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<int*>(e), (short)0, 4 * kBT, 0x00020000);
#pragma unroll
    for (int i = 0; i < 2; ++i)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (lptr_t)(sE + kBT * wave + 64 * i), 4, (uint32_t)(64 * i + lane) * 4u,
                                                 0, 0, 0);
}

// The accumulators are already scaled back (h_unscale).
template <bool PRE>
__device__ __forceinline__ void h_epilogue(const HArgs& h, int item, int tr, int tc, int wr, int wc,
                                           const f4 (&accr)[4][2], const f4 (&acci)[4][2], const float2* c0) {
    // the lane index computed here, opaque to the compiler (which would
    // otherwise keep the thread index live across the depth loop for it and
    // spill it)
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    const __amdgpu_buffer_rsrc_t rc = h_crsrc(h, item);
    const int gr0 = tr + 64 * wr + (l & 15), gc0 = tc + 32 * wc + 4 * (l >> 4);
    auto off = [&](int a, int b, int r) -> uint32_t {
        const int gr = gr0 + 16 * a, gc = gc0 + 16 * b + r;
        return (gr < h.r_hi && gc < h.c_hi) ? ((uint32_t)gr + (uint32_t)gc * (uint32_t)h.ld) * 8u : 0x80000000u;
    };
    float2 cv[2][2][4];
    auto load = [&](int a, float2 (&v)[2][4]) {
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                v[b][r] = __builtin_bit_cast(float2,
                                             __builtin_amdgcn_raw_buffer_load_b64(rc, (int)off(a, b, r), 0, kHNt));
    };
    if constexpr (PRE) {  // read before any further load is issued, so the wait covers the copy only
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) cv[0][b][r] = c0[(16 * b + 4 * (l >> 4) + r) * 16 + (l & 15)];
    } else {
        load(0, cv[0]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        if (a + 1 < 4) load(a + 1, cv[(a + 1) & 1]);
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float2 v = make_float2(cv[a & 1][b][r].x - accr[a][b][r], cv[a & 1][b][r].y - acci[a][b][r]);
                __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), rc, (int)off(a, b, r), 0, kHNt);
            }
    }
}

__device__ __forceinline__ void h_rsrc(const HArgs& h, int item, int tr, int tc, __amdgpu_buffer_rsrc_t& rl,
                                       __amdgpu_buffer_rsrc_t& ru) {
    // buffer ranges end with the item's image (reads beyond return zeros)
    const int64_t ls = (int64_t)kHUnits * h.npad, us = (int64_t)kHUnits * h.ncpad;
    rl = __builtin_amdgcn_make_buffer_rsrc(const_cast<u4v*>(h.Limg + item * ls + 4 * tr), (short)0,
                                           (int)((ls - 4 * tr) * 16), 0x00020000);
    ru = __builtin_amdgcn_make_buffer_rsrc(const_cast<u4v*>(h.Uimg + item * us + 4 * tc), (short)0,
                                           (int)((us - 4 * tc) * 16), 0x00020000);
}

// One tile per workgroup: tile x per + w / 8 of XCD x = w % 8.  The depth
// loop (kNBOuter = 8 chunks of 16) runs over two LDS stages, chunk c in
// stage c mod 2.  The stages are distinct restrict parameters, so the
// compiler's own wait before reading one stage does not include the loads
// in flight into the other.
__device__ __forceinline__ void h_run(const HArgs& h, u4v* __restrict__ sS0, u4v* __restrict__ sS1,
                                      int* __restrict__ sE) {
    // stage s: L21 units [0, 1024), U12 units [1024, 2048)
    u4v* const sL0 = sS0;
    u4v* const sU0 = sS0 + 1024;
    u4v* const sL1 = sS1;
    u4v* const sU1 = sS1 + 1024;
    const int per = (h.ntiles + kXcd - 1) / kXcd, x8 = blockIdx.x % kXcd;
    const int t = x8 * per + blockIdx.x / kXcd;
    if (t >= h.ntiles || t >= (x8 + 1) * per) return;
    const int tid = threadIdx.x, l = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave & 1, wc = wave >> 1;  // 2 x 4 waves: 64 rows x 32 columns each
    // this lane's staging unit: row (column) x of the tile, image unit q
    const int x = tid >> 2, q = (tid & 3) ^ hkey(x);
    const uint32_t vo = (uint32_t)(4 * x + q) * 16u;
    int item, ti, tj;
    h_coords(t, h.ntr, h.ntc, item, ti, tj);
    const int tr = h.r_lo + ti * kBT, tc = h.c_lo + tj * kBT;
    __amdgpu_buffer_rsrc_t rl, ru;
    h_rsrc(h, item, tr, tc, rl, ru);
    h_estage(h, item, tr, tc, wave, l, sE);
    h_stage(rl, ru, h.npad, h.ncpad, 0, sL0, sU0, vo, wave);
    h_stage(rl, ru, h.npad, h.ncpad, 1, sL1, sU1, vo, wave);
    // every row of the tile inside: C block 0 is copied to LDS during the
    // last chunk (4 KB per wave in stage 0)
    const bool interior = tr + kBT <= h.r_hi;
    u4v* creg = sS0 + 256 * wave;
    f4 accr[4][2], acci[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accr[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
            acci[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
    for (int c = 0; c < kHChunks; c += 2) {
        h_barrier_ready(true);
        h_chunk(sL0, sU0, l, wr, wc, accr, acci);
        if (c + 2 < kHChunks) {
            h_barrier_done();  // every wave is done with this stage
            h_stage(rl, ru, h.npad, h.ncpad, c + 2, sL0, sU0, vo, wave);
        }
        h_barrier_ready(c + 2 < kHChunks);
        // last chunk: stage 0 is free (every wave has passed the barrier
        // after its last reads of it); each wave copies its C block 0 there
        if (c + 2 >= kHChunks && interior) h_cstage(h, h_crsrc(h, item), tr, tc, wr, wc, l, creg);
        h_chunk(sL1, sU1, l, wr, wc, accr, acci);
        if (c + 3 < kHChunks) {
            h_barrier_done();
            h_stage(rl, ru, h.npad, h.ncpad, c + 3, sL1, sU1, vo, wave);
        }
    }
    h_unscale(sE, wr, wc, l, accr, acci);
    if (interior)
        h_epilogue<true>(h, item, tr, tc, wr, wc, accr, acci, reinterpret_cast<const float2*>(creg));
    else
        h_epilogue<false>(h, item, tr, tc, wr, wc, accr, acci, nullptr);
}

// 128 x 128 tile per 512-thread workgroup (8 waves, each 64 x 32 = 4 x 2
// blocks), operand chunks double-buffered in 64 KB of LDS and at most 128
// VGPRs (two workgroups per CU), with the loads of chunk c + 1 in flight
// while chunk c is multiplied.  Depth kNBOuter (the width of every block
// column that has a trailing update).
__global__ __launch_bounds__(512, 4) void k_lu_gemm_hs(HArgs h) {
    if (h.chain) __builtin_amdgcn_s_setprio(3);  // as the other panel-chain kernels
    __shared__ u4v sS0[2048], sS1[2048];
    __shared__ int sE[2 * kBT];
    h_run(h, sS0, sS1, sE);
}

// ---------------------------------------------------------------------------
// host-side recursion
// ---------------------------------------------------------------------------
// Split operand images of the FP32 LU's trailing updates (in the context's
// grow-only image buffer): two L21 slots (block column parity) and the U12
// image, with their exponents; what each L21 slot holds (rows from r0,
// depths k_lo .. k_lo + kw) is checked by the updates that read it.
struct HImg {
    u4v* L[2] = {nullptr, nullptr};
    int* rexp[2] = {nullptr, nullptr};
    u4v* U = nullptr;
    int* cexp = nullptr;
    int64_t npad = 0, ncpad = 0;
    int64_t r0[2] = {-1, -1}, k_lo[2] = {-1, -1};
    int kw[2] = {0, 0};
    int slot = 0;  // slot of the block column the next updates use
};

// Unused LDS reserved by each workgroup of the wide trailing update of the
// mixed LU (k_lu_gemm_hs, 66,560 B of its own): 16 KB more puts it above
// half of a CU's 160 KB, so one fits per CU instead of two and half of every
// CU stays free for the panel chain's kernels beside it.  Measured at 2,048
// panels x 64 frequencies: 34.2-34.5 -> 33.0-33.5 ms per sweep (16-48 KB
// alike, more slower; reserving LDS in the wide TRSM as well did not help,
// nor reserving it only for the smaller updates of the later block columns);
// the uneven 256-body batch is the same with or without it.
constexpr unsigned kWideLdsPad = 16384;

template <class T>
struct LuPlan {
    typedef typename Cx<T>::t C2;
    Context* ctx;
    C2* A;
    int64_t n, ncols, ld, stride, batch;
    int64_t* ipiv;
    int* info;
    int64_t pc_lo, pc_hi;  // block column being factored
    HImg* img = nullptr;   // FP32 only
    bool chain = false;    // updates issued on the panel chain's stream
    bool wide = false;     // the wide trailing update (beside the panel chain)
};

// Bytes of the image buffer for `batch` matrices of n rows and ncols columns.
static size_t himg_bytes(int64_t n, int64_t ncols, int64_t batch) {
    const int64_t npad = (n + kBT - 1) / kBT * kBT, ncpad = (ncols + kBT - 1) / kBT * kBT;
    return (size_t)batch * ((size_t)(2 * npad + ncpad) * kHUnits * 16 + (size_t)(2 * npad + ncpad) * 4);
}

static int himg_setup(Context* ctx, int64_t n, int64_t ncols, int64_t batch, HImg& H) {
    const size_t need = himg_bytes(n, ncols, batch);
    if (need > ctx->himg_bytes) {
        if (ctx->himg) {  // the previous factorisations have completed on the context's stream
            MH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->himg);
            ctx->himg = nullptr;
            ctx->himg_bytes = 0;
        }
        if (hipMalloc(&ctx->himg, need) != hipSuccess) {
            ctx->himg = nullptr;
            set_error("lu: hipMalloc of " + std::to_string(need) + " bytes for the split operand images failed");
            return MH_ERR_HIP;
        }
        ctx->himg_bytes = need;
    }
    H.npad = (n + kBT - 1) / kBT * kBT;
    H.ncpad = (ncols + kBT - 1) / kBT * kBT;
    char* p = static_cast<char*>(ctx->himg);
    const size_t lb = (size_t)batch * H.npad * kHUnits * 16, ub = (size_t)batch * H.ncpad * kHUnits * 16;
    H.L[0] = reinterpret_cast<u4v*>(p);
    H.L[1] = reinterpret_cast<u4v*>(p + lb);
    H.U = reinterpret_cast<u4v*>(p + 2 * lb);
    int* e = reinterpret_cast<int*>(p + 2 * lb + ub);
    H.rexp[0] = e;
    H.rexp[1] = e + batch * H.npad;
    H.cexp = e + 2 * batch * H.npad;
    return MH_OK;
}

// L21 image of the block column whose trailing rows start at r0
static int lu_split_rows(const LuPlan<float>& P, int slot, int64_t r0, int64_t k_lo, int kw) {
    HImg& H = *P.img;
    const int64_t rows = H.npad - r0;
    if (rows <= 0) return MH_OK;
    ProfScope ps(P.ctx, "lu_split", 0.0, (8.0 * 2 + 8.0) * (double)(P.n - r0) * kw * P.batch);
    hipLaunchKernelGGL(k_split<true>, dim3((unsigned)((rows + 15) / 16), (unsigned)P.batch), dim3(512), 0,
                       P.ctx->stream, P.A, P.ld, P.stride, P.n, r0, H.npad, H.npad, k_lo, kw, H.L[slot],
                       (int64_t)kHUnits * H.npad, H.rexp[slot], H.npad, P.info);
    H.r0[slot] = r0;
    H.k_lo[slot] = k_lo;
    H.kw[slot] = kw;
    return MH_OK;
}

template <class T>
static int lu_swap(const LuPlan<T>& P, int64_t k0, int npiv, int64_t c_lo, int64_t c_hi) {
    if (c_hi <= c_lo || npiv <= 0) return MH_OK;
    const int64_t nc = c_hi - c_lo;
    ProfScope ps(P.ctx, "lu_swap", 0.0, 64.0 * npiv * (double)nc * P.batch);
    hipLaunchKernelGGL(k_lu_swap<typename Cx<T>::t>, dim3((unsigned)((nc + 255) / 256), (unsigned)P.batch), dim3(256), 0,
                       P.ctx->stream, P.A, P.ld, P.stride, P.n, P.ipiv, k0, npiv, c_lo, c_hi);
    return MH_OK;
}

// Whether the FP32 trailing update of rows [r_lo, r_hi) x columns [c_lo,
// c_hi) by kw depths runs on the split images: the split-image kernel takes
// tiles that start on multiples of kBT (the images and exponents end at the
// padded size) and full-depth blocks.
template <class T>
static bool lu_split_update(int64_t r_lo, int64_t r_hi, int64_t c_lo, int64_t c_hi, int kw) {
    return sizeof(T) == 4 && r_hi - r_lo >= kBT && c_hi - c_lo >= kBT && r_lo % kBT == 0 && c_lo % kBT == 0 &&
           kw == kNBOuter;
}

// u_ready: the TRSM that solved U12 has written its image (k_lu_trsm_blk<.., true>)
template <class T>
static int lu_gemm(const LuPlan<T>& P, int64_t r_lo, int64_t r_hi, int64_t c_lo, int64_t c_hi,
                   int64_t k_lo, int kw, bool u_ready = false) {
    if (r_hi <= r_lo || c_hi <= c_lo || kw <= 0) return MH_OK;
    const int64_t mr = r_hi - r_lo, mc = c_hi - c_lo;
    const bool big = sizeof(T) == 8 ? (mr >= kBT && mc >= kBT) : lu_split_update<T>(r_lo, r_hi, c_lo, c_hi, kw);
    MH_REQUIRE(!u_ready || (big && sizeof(T) == 4), "lu: U12 image written for an update that does not read it");
    const char* name = big ? (sizeof(T) == 8 ? "lu_gemm_big" : "lu_gemm_f16x3") : (sizeof(T) == 8 ? "lu_gemm" : "lu_gemm_f32");
    if (big) {
        const dim3 grid((unsigned)((mr + kBT - 1) / kBT), (unsigned)((mc + kBT - 1) / kBT), (unsigned)P.batch);
        if constexpr (sizeof(T) == 4) {  // FP32 LU of the mixed path: split-half matrix cores (kw = kNBOuter)
            HImg* H = P.img;
            MH_REQUIRE(H != nullptr, "lu: split operand images missing");
            const int sl = H->slot;
            MH_REQUIRE(H->r0[sl] >= 0 && r_lo >= H->r0[sl] && k_lo == H->k_lo[sl] && kw == H->kw[sl] &&
                           kw == kNBOuter && r_hi <= H->npad && c_hi <= H->ncpad,
                       "lu: trailing update does not match the split L21 image");
            if (!u_ready) {  // U12 image of these columns
                ProfScope pss(P.ctx, "lu_split", 0.0, (8.0 * 2 + 8.0) * (double)mc * kw * P.batch);
                hipLaunchKernelGGL(k_split<false>, dim3((unsigned)((mc + 15) / 16), (unsigned)P.batch),
                                   dim3(512), 0, P.ctx->stream, P.A, P.ld, P.stride, P.n, c_lo, c_hi, H->ncpad,
                                   k_lo, kw, H->U, (int64_t)kHUnits * H->ncpad, H->cexp, H->ncpad, P.info);
            }
            ProfScope ps(P.ctx, name, 8.0 * (double)mr * (double)mc * kw * P.batch,
                         2.0 * sizeof(T) * (2.0 * mr * mc) * P.batch + 8.0 * (double)kw * (mr + mc) * P.batch);
            const int64_t ntr = grid.x, ntc = grid.y, ntiles = ntr * ntc * P.batch;
            const int64_t nwg = (ntiles + kXcd - 1) / kXcd * kXcd;
            MH_REQUIRE(ntiles <= 0x7fffffff && (int64_t)kHUnits * H->npad * 16 <= 0x7fffffff &&
                           (int64_t)kHUnits * H->ncpad * 16 <= 0x7fffffff,
                       "lu: too many tiles for one launch");
            const HArgs ha{P.A,       P.stride,    (int)P.ld,   (int)r_lo,       (int)r_hi,
                           (int)c_lo, (int)c_hi,   H->L[sl],    H->rexp[sl],     (int)H->npad,
                           H->U,      H->cexp,     (int)H->ncpad, (int)ntr,      (int)ntc,
                           (int)ntiles, P.chain ? 1 : 0};
            // The wide update reserves LDS it does not use (kWideLdsPad), so
            // that one of its workgroups fits per CU instead of two: half of
            // every CU's registers and LDS stay free for the panel chain's
            // kernels beside it (the recursion's GEMMs and TRSMs, the narrow
            // leaves), which otherwise wait for a tile to retire.
            hipLaunchKernelGGL(k_lu_gemm_hs, dim3((unsigned)nwg), dim3(512), P.wide ? kWideLdsPad : 0u,
                               P.ctx->stream, ha);
        } else {
            ProfScope ps(P.ctx, name, 8.0 * (double)mr * (double)mc * kw * P.batch,
                         2.0 * sizeof(T) * (2.0 * mr * mc + (double)kw * (mr + mc)) * P.batch);
            hipLaunchKernelGGL(k_lu_gemm_big<T>, grid, dim3(512), 0, P.ctx->stream, P.A, P.ld, P.stride, r_lo,
                               r_hi, c_lo, c_hi, k_lo, kw);
        }
    } else {
        ProfScope ps(P.ctx, name, 8.0 * (double)mr * (double)mc * kw * P.batch,
                     2.0 * sizeof(T) * (2.0 * mr * mc + (double)kw * (mr + mc)) * P.batch);
        hipLaunchKernelGGL(k_lu_gemm<T>, dim3((unsigned)((mr + 63) / 64), (unsigned)((mc + 63) / 64), (unsigned)P.batch),
                           dim3(256), 0, P.ctx->stream, P.A, P.ld, P.stride, r_lo, r_hi, c_lo, c_hi, k_lo, kw);
    }
    return MH_OK;
}

// X = L[k0:k0+w, k0:k0+w]^{-1} X on columns [c_lo, c_hi), one launch
// img (FP32): also write the U12 image of the solved columns for the split-image update
template <class T>
static int lu_trsm(const LuPlan<T>& P, int64_t k0, int w, int64_t c_lo, int64_t c_hi, bool img = false) {
    if (c_hi <= c_lo) return MH_OK;
    const int64_t nc = c_hi - c_lo;
    MH_REQUIRE(w <= kNBOuter, "lu: TRSM block wider than a block column");
    const dim3 grid((unsigned)((nc + kTbCols - 1) / kTbCols), (unsigned)P.batch);
    if constexpr (sizeof(T) == 4) {
        if (img) {
            HImg* H = P.img;
            MH_REQUIRE(H != nullptr && w == kNBOuter && c_hi <= H->ncpad, "lu: U12 image does not fit the TRSM");
            ProfScope ps(P.ctx, "lu_trsm", 4.0 * w * w * (double)nc * P.batch,
                         (4.0 * sizeof(T) + 8.0) * w * (double)nc * P.batch);
            hipLaunchKernelGGL((k_lu_trsm_blk<T, true>), grid, dim3(256), 0, P.ctx->stream, P.A, P.ld, P.stride, k0, w,
                               c_lo, c_hi, H->U, (int64_t)kHUnits * H->ncpad, H->ncpad, H->cexp, H->ncpad, P.info);
            return MH_OK;
        }
    }
    MH_REQUIRE(!img, "lu: split images are FP32 only");
    ProfScope ps(P.ctx, "lu_trsm", 4.0 * w * w * (double)nc * P.batch, 4.0 * sizeof(T) * w * (double)nc * P.batch);
    hipLaunchKernelGGL((k_lu_trsm_blk<T, false>), grid, dim3(256), 0, P.ctx->stream, P.A, P.ld, P.stride, k0, w, c_lo,
                       c_hi, nullptr, (int64_t)0, (int64_t)0, nullptr, (int64_t)0, nullptr);
    return MH_OK;
}

// Leaf shape for a panel of m rows: 512 threads (two waves per SIMD, so the
// per-wave work of a column, its reductions and reciprocal, is paid by 8
// waves instead of 16), R = 2, 4 or 8 rows per thread and a width that keeps
// the panel in 128 VGPRs per thread (no spills).
constexpr int kLeafNT = 512;
template <class T>
static int leaf_width(int64_t m) {
    const int R = m <= 1024 ? 2 : (m <= 2048 ? 4 : 8);
    if (sizeof(T) == 4) return R <= 4 ? 16 : 8;
    return R == 2 ? 16 : (R == 4 ? 8 : 4);
}

template <class T>
static int lu_leaf(const LuPlan<T>& P, int64_t k0, int w, int64_t x_lo = 0, int64_t x_hi = 0) {
    const int64_t m = P.n - k0;
    ProfScope ps(P.ctx, "lu_leaf", 8.0 * (double)m * w * w / 2.0 * P.batch, 32.0 * (double)m * w * P.batch);
#define LEAF(RR, WW)                                                                                   \
    hipLaunchKernelGGL((k_lu_leaf<T, RR, WW, kLeafNT>), dim3((unsigned)P.batch), dim3(kLeafNT), 0,     \
                       P.ctx->stream, P.n, P.A, P.ld, P.stride, k0, w, P.ipiv, P.info, P.pc_lo, P.pc_hi, x_lo, x_hi)
    if (m <= 1024) {
        LEAF(2, 16);
    } else if (m <= 2048) {
        if constexpr (sizeof(T) == 4) LEAF(4, 16);
        else LEAF(4, 8);
    } else {
        if constexpr (sizeof(T) == 4) LEAF(8, 8);
        else LEAF(8, 4);
    }
#undef LEAF
    return MH_OK;
}

// recursive factorisation of the panel [k0, k0+w) (rows k0..n), touching only
// the panel's columns (zgetrf2 order)
// (each leaf applies its interchanges to the rest of the block column)
template <class T>
static int lu_panel(const LuPlan<T>& P, int64_t k0, int w) {
    int L = leaf_width<T>(P.n - P.pc_lo);
    // a narrower leaf (the autotuner's choice): the same kernels, called
    // with fewer columns each
    if (P.ctx->leaf_cap > 0 && P.ctx->leaf_cap < L) L = P.ctx->leaf_cap;
    if (w <= L) return lu_leaf(P, k0, w);
    int h = ((w / 2) + L - 1) / L * L;
    if (h >= w) h = w - L;
    if (h <= L && w - h <= kLeafX && L <= 16) {  // the left part is one leaf: it also solves the right part's top rows
        MH_REQUIRE(k0 + w <= P.pc_hi, "lu: leaf solve beyond the block column");
        TRY_RC(lu_leaf(P, k0, h, k0 + h, k0 + w));
    } else {
        TRY_RC(lu_panel(P, k0, h));
        TRY_RC(lu_trsm(P, k0, h, k0 + h, k0 + w));
    }
    TRY_RC(lu_gemm(P, k0 + h, P.n, k0 + h, k0 + w, k0, h));
    TRY_RC(lu_panel(P, k0 + h, w - h));
    return MH_OK;
}

// Right-looking blocked LU with look-ahead: the factorisation of panel k+1
// (latency-bound: pivot searches, small updates) runs on `stream` while the
// wide trailing update of panel k (the large GEMM) runs on `aux`:
//   stream: [wait B(k-2)] factor P_k | [wait B(k-1)] swap + TRSM + GEMM on
//           the next panel's columns only
//   aux:    [wait P_k factored] left row swaps of P_k, swap + TRSM + GEMM on
//           the columns beyond the next panel
// The row interchanges of P_k on the columns left of it are applied in B(k),
// after every update that reads those columns has completed.  With the
// kernel timer on, the same schedule runs on one stream (each timed kernel
// alone on the device).
static int transpose_rows(Context* c, int64_t n, int64_t batch, const double2* A, int64_t ld, int64_t stride,
                          double2* At, int64_t r_lo, int64_t r_hi) {
    return tri_transpose(c, n, batch, A, ld, stride, At, r_lo, r_hi);
}
static int transpose_rows(Context* c, int64_t n, int64_t batch, const float2* A, int64_t ld, int64_t stride,
                          float2* At, int64_t r_lo, int64_t r_hi) {
    return tri_transpose_f32(c, n, batch, A, ld, stride, At, r_lo, r_hi);
}

//
// With LUt, block row k of the factors (final once panel k, its narrow and
// its wide update are done: later interchanges only touch rows below it) is
// transposed into block column k of LUt on the side stream.
// With inv (and perm), the permutation and inverse diagonal blocks the
// triangular solves use are built per block column on the side stream, once
// its panel is factored.
template <class T>
static int lu_factor_t(Context* ctx, int64_t n, int64_t ncols, int64_t batch, typename Cx<T>::t* A,
                       int64_t ld, int64_t stride, int64_t* ipiv, int* d_info, typename Cx<T>::t* LUt, int* perm,
                       typename Cx<T>::t* inv) {
    if (n == 0 || batch == 0) return MH_OK;
    MH_REQUIRE(batch <= 65535, "lu: batch too large for one launch");
    MH_REQUIRE(n <= 4096, "lu: n larger than 4096 is not supported by the register-resident panel");
    MH_REQUIRE(sizeof(T) == 8 || (ncols * ld) * 8 <= 0x7fffffff, "lu: FP32 matrix extent above 2 GiB");
    MH_REQUIRE(n * ld * (int64_t)sizeof(typename Cx<T>::t) <= 0x7fffffff, "lu: matrix extent above 2 GiB");
    LuPlan<T> P{ctx, A, n, ncols, ld, stride, batch, ipiv, d_info, 0, 0};
    HImg H;
    if constexpr (sizeof(T) == 4) {
        TRY_RC(himg_setup(ctx, n, ncols, batch, H));
        P.img = &H;
    }
    const int64_t npan = (n + kNBOuter - 1) / kNBOuter;
    TRY_RC(ensure_aux(ctx));
    TRY_RC(ensure_events(ctx, (size_t)(3 * npan + 2)));
    hipEvent_t* evF = ctx->ev_pool.data();             // panel k factored (stream)
    hipEvent_t* evB = ctx->ev_pool.data() + npan;      // wide update k done (aux)
    hipEvent_t* evN = ctx->ev_pool.data() + 2 * npan;  // narrow update k done (stream)
    hipEvent_t ev0 = ctx->ev_pool[3 * npan];
    hipEvent_t evT = ctx->ev_pool[3 * npan + 1];       // last transposed block row (aux)
    const hipStream_t sA = ctx->stream, sB = ctx->profiling ? ctx->stream : ctx->aux;
    const hipStream_t sT = ctx->profiling ? ctx->stream : ctx->side;
    MH_HIP_CHECK(hipEventRecord(ev0, sA));  // aux starts after the work already issued
    MH_HIP_CHECK(hipStreamWaitEvent(sB, ev0, 0));
    int rc = MH_OK;
    for (int64_t k = 0; k < npan && rc == MH_OK; ++k) {
        const int64_t k0 = k * kNBOuter;
        const int w = (int)((n - k0) < kNBOuter ? (n - k0) : kNBOuter);
        const int64_t kn0 = k0 + w;
        const int wn = (int)((n - kn0) < kNBOuter ? (n - kn0) : kNBOuter);
        const int64_t cw0 = kn0 + wn;  // first column of the wide update
        if (k >= 2) MH_HIP_CHECK(hipStreamWaitEvent(sA, evB[k - 2], 0));
        P.pc_lo = k0;
        P.pc_hi = k0 + w;
        rc = lu_panel(P, k0, w);
        if constexpr (sizeof(T) == 4)  // L21 image of this block column, read by both updates
            if (rc == MH_OK && kn0 < n) rc = lu_split_rows(P, (int)(k & 1), kn0, k0, w);
        if (rc != MH_OK) break;
        MH_HIP_CHECK(hipEventRecord(evF[k], sA));
        H.slot = (int)(k & 1);
        if (inv) {  // this block column's pivots and diagonal blocks are final
            MH_HIP_CHECK(hipStreamWaitEvent(sT, evF[k], 0));
            ctx->stream = sT;
            rc = tri_prepare_part(ctx, n, batch, A, ld, stride, ipiv, perm, inv, k0, w);
            ctx->stream = sA;
            if (rc != MH_OK) break;
        }
        // wide update on aux
        ctx->stream = sB;
        MH_HIP_CHECK(hipStreamWaitEvent(sB, evF[k], 0));
        P.wide = true;
        rc = lu_swap(P, k0, w, 0, k0);
        if (rc == MH_OK) rc = lu_swap(P, k0, w, cw0, ncols);
        const bool uw = lu_split_update<T>(kn0, n, cw0, ncols, w);  // the TRSM writes U12's image
        if (rc == MH_OK) rc = lu_trsm(P, k0, w, cw0, ncols, uw);
        if (rc == MH_OK) rc = lu_gemm(P, kn0, n, cw0, ncols, k0, w, uw);
        P.wide = false;
        ctx->stream = sA;
        if (rc != MH_OK) break;
        MH_HIP_CHECK(hipEventRecord(evB[k], sB));
        // narrow update of the next panel on stream
        if (wn > 0) {
            if (k >= 1) MH_HIP_CHECK(hipStreamWaitEvent(sA, evB[k - 1], 0));
            rc = lu_swap(P, k0, w, kn0, cw0);
            const bool un = lu_split_update<T>(kn0, n, kn0, cw0, w);
            if (rc == MH_OK) rc = lu_trsm(P, k0, w, kn0, cw0, un);
            P.chain = true;
            if (rc == MH_OK) rc = lu_gemm(P, kn0, n, kn0, cw0, k0, w, un);
            P.chain = false;
        }
        if (LUt && rc == MH_OK) {  // block row k is final: transpose it on the side stream
            MH_HIP_CHECK(hipEventRecord(evN[k], sA));
            MH_HIP_CHECK(hipStreamWaitEvent(sT, evN[k], 0));
            MH_HIP_CHECK(hipStreamWaitEvent(sT, evB[k], 0));
            ctx->stream = sT;
            rc = transpose_rows(ctx, n, batch, A, ld, stride, LUt, k0, k0 + w);
            ctx->stream = sA;
        }
    }
    ctx->stream = sA;
    // stream continues after all of aux's updates (and the transposes)
    MH_HIP_CHECK(hipStreamWaitEvent(sA, evB[npan - 1], 0));
    if (LUt || inv) {
        MH_HIP_CHECK(hipEventRecord(evT, sT));
        MH_HIP_CHECK(hipStreamWaitEvent(sA, evT, 0));
    }
    if (rc != MH_OK) return rc;
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int lu_factor(Context* ctx, int64_t n, int64_t ncols, int64_t batch, double2* A, int64_t ld,
              int64_t stride, int64_t* ipiv, int* d_info, double2* LUt, int* perm, double2* inv) {
    return lu_factor_t<double>(ctx, n, ncols, batch, A, ld, stride, ipiv, d_info, LUt, perm, inv);
}

int lu_factor_f32(Context* ctx, int64_t n, int64_t ncols, int64_t batch, float2* A, int64_t ld,
                  int64_t stride, int64_t* ipiv, int* d_info, float2* LUt, int* perm, float2* inv) {
    return lu_factor_t<float>(ctx, n, ncols, batch, A, ld, stride, ipiv, d_info, LUt, perm, inv);
}

// ---------------------------------------------------------------------------
// Y = S X (trans = 0) or Y = S^T X (trans = 1) for a few right-hand sides,
// optionally Y = Z - op(S) X.  HBM-bound: S is read once per launch for up to
// 16 columns of X.
// ---------------------------------------------------------------------------
constexpr int kMaxRhs = 16;

// acc[r] += s x[r] for all NR columns (columns beyond nrhs hold zeros)
template <int NR>
__device__ __forceinline__ void zmac(double2 (&acc)[NR], double2 s, const double2* x) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        acc[r].x = fma(s.x, x[r].x, acc[r].x);
        acc[r].x = fma(-s.y, x[r].y, acc[r].x);
        acc[r].y = fma(s.x, x[r].y, acc[r].y);
        acc[r].y = fma(s.y, x[r].x, acc[r].y);
    }
}

// one thread per row of S, columns in chunks of 64 staged with X in LDS
template <int NR>
__global__ __launch_bounds__(256) void k_zgemv_nn(int64_t n, int64_t nrhs, const double2* __restrict__ S, int64_t ld,
                                                  int64_t stride, const double2* __restrict__ X, int64_t ldx,
                                                  int64_t xstride, double2* __restrict__ Y, int64_t ldy,
                                                  int64_t ystride, const double2* __restrict__ Z) {
    const int64_t item = blockIdx.y;
    const double2* Sm = S + item * stride;
    const double2* Xm = X + item * xstride;
    double2* Ym = Y + item * ystride;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int nr = (int)nrhs;
    __shared__ double2 sx[64][NR];
    double2 acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = make_double2(0.0, 0.0);
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
        const int64_t jn = (n - j0) < 64 ? (n - j0) : 64;
        // this block of X: every load issued before the barrier, none waited
        // for alone (entries outside load element 0 and count zero)
        constexpr int kPer = (64 * NR + 255) / 256;
        double tvx[kPer], tvy[kPer];
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int q = (int)threadIdx.x + 256 * e, jj = q % 64, r = q / 64;
            const bool in = q < 64 * NR && jj < jn && r < nr;
            const double2 v = Xm[in ? (j0 + jj) + r * ldx : 0];
            tvx[e] = in ? v.x : 0.0;
            tvy[e] = in ? v.y : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int q = (int)threadIdx.x + 256 * e;
            if (q < 64 * NR) sx[q % 64][q / 64] = make_double2(tvx[e], tvy[e]);
        }
        __syncthreads();
        if (i < n) {
            const double2* col = Sm + i + j0 * ld;
            if (jn == 64) {
                for (int jb = 0; jb < 64; jb += 16) {
                    double2 sv[16];  // 16 column loads in flight
#pragma unroll
                    for (int u = 0; u < 16; ++u) sv[u] = col[(jb + u) * ld];
#pragma unroll
                    for (int u = 0; u < 16; ++u) zmac<NR>(acc, sv[u], sx[jb + u]);
                }
            } else {
                for (int jj = 0; jj < jn; ++jj) zmac<NR>(acc, col[jj * ld], sx[jj]);
            }
        }
    }
    if (i < n) {
        const double2* Zm = Z ? Z + item * ystride : nullptr;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r < nr) {
                if (Zm) {
                    const double2 zv = Zm[i + r * ldy];
                    Ym[i + r * ldy] = make_double2(zv.x - acc[r].x, zv.y - acc[r].y);
                } else {
                    Ym[i + r * ldy] = acc[r];
                }
            }
    }
}

// Y[:, r] = S^T X[:, r]: one wave per output row i (a column of S), lanes
// along the column, coalesced; the column is read once for all r.
template <int NR>
__global__ __launch_bounds__(256) void k_zgemv_tn(int64_t n, int64_t nrhs, const double2* __restrict__ S, int64_t ld,
                                                  int64_t stride, const double2* __restrict__ X, int64_t ldx,
                                                  int64_t xstride, double2* __restrict__ Y, int64_t ldy,
                                                  int64_t ystride, const double2* __restrict__ Z) {
    const int64_t item = blockIdx.y;
    const double2* Sm = S + item * stride;
    const double2* Xm = X + item * xstride;
    double2* Ym = Y + item * ystride;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int nr = (int)nrhs;
    const double2* col = Sm + i * ld;
    double2 acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = make_double2(0.0, 0.0);
    int64_t j0 = 0;
    for (; j0 + 8 * 64 <= n; j0 += 8 * 64) {  // 8 loads of the column in flight
        double2 sv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) sv[u] = col[j0 + u * 64 + lane];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            double2 xv[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) {  // unused columns load column 0 (no branch) and count zero
                const double2 v = Xm[j0 + u * 64 + lane + (r < nr ? r : 0) * ldx];
                xv[r] = make_double2(r < nr ? v.x : 0.0, r < nr ? v.y : 0.0);
            }
            zmac<NR>(acc, sv[u], xv);
        }
    }
    for (int64_t j = j0 + lane; j < n; j += 64) {
        const double2 sj = col[j];
        double2 xv[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double2 v = Xm[j + (r < nr ? r : 0) * ldx];
            xv[r] = make_double2(r < nr ? v.x : 0.0, r < nr ? v.y : 0.0);
        }
        zmac<NR>(acc, sj, xv);
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (r >= nr) continue;
        double sr = acc[r].x, si = acc[r].y;
        for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off);
            si += __shfl_xor(si, off);
        }
        if (lane == 0) {
            if (Z) {
                const double2 zv = Z[item * ystride + i + r * ldy];
                Ym[i + r * ldy] = make_double2(zv.x - sr, zv.y - si);
            } else {
                Ym[i + r * ldy] = make_double2(sr, si);
            }
        }
    }
}

// Residuals of a forward and a transposed system with the same matrix, in
// one pass over it: Rf = Bf - A Xf (nrf <= NRF columns) and Ra = Ba - A^T Xa
// (nra <= NRA columns).  One thread per row of a 256-row tile, columns in
// chunks of 64 with Xf staged in LDS and 16 column loads in flight.  For
// A^T Xa every lane forms A[i, j] Xa[i] for 16 columns; recursive halving
// across the wave (lanes exchanging half of their columns at distances 32,
// 16, 8, 4, then summing at 2 and 1) leaves each column's wave sum in 4
// lanes, the 4 waves are added in LDS and the tile's partial sums are
// written to Pa[item][tile][r][j] for k_pair_finish.
constexpr int kPairRows = 256;
template <int NRF, int NRA>
__global__ __launch_bounds__(256) void k_zgemv_pair(int64_t n, const double2* __restrict__ A, int64_t ld,
                                                    int64_t stride, int nrf, const double2* __restrict__ Xf,
                                                    const double2* __restrict__ Bf, double2* __restrict__ Rf,
                                                    int64_t ldf, int64_t fstride, int nra,
                                                    const double2* __restrict__ Xa, int64_t lda, int64_t astride,
                                                    double2* __restrict__ Pa, int ntile) {
    const int64_t item = blockIdx.y;
    const int tile = blockIdx.x;
    const double2* Am = A + item * stride;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t i = (int64_t)tile * kPairRows + tid;
    const bool ok = i < n;
    __shared__ double2 sx[64][NRF];
    __shared__ double2 scol[4][NRA][64];  // per wave: column sums of the current 64-column chunk
    double2 xa[NRA];
#pragma unroll
    for (int r = 0; r < NRA; ++r)
        xa[r] = (ok && r < nra) ? Xa[item * astride + i + r * lda] : make_double2(0.0, 0.0);
    double2 acc[NRF];
#pragma unroll
    for (int r = 0; r < NRF; ++r) acc[r] = make_double2(0.0, 0.0);
    const double2* row = Am + (ok ? i : 0);
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
        const int jn = (int)((n - j0) < 64 ? (n - j0) : 64);
        __syncthreads();  // previous chunk's sx and scol consumed
        for (int q = tid; q < 64 * NRF; q += 256) {
            const int jj = q % 64, r = q / 64;
            sx[jj][r] = (jj < jn && r < nrf) ? Xf[item * fstride + (j0 + jj) + r * ldf] : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 1
        for (int jb = 0; jb < 64; jb += 16) {
            double2 av[16];  // 16 column loads in flight (zero outside the matrix)
#pragma unroll
            for (int u = 0; u < 16; ++u)
                av[u] = (ok && jb + u < jn) ? row[(j0 + jb + u) * ld] : make_double2(0.0, 0.0);
#pragma unroll
            for (int u = 0; u < 16; ++u) zmac<NRF>(acc, av[u], sx[jb + u]);
#pragma unroll
            for (int r = 0; r < NRA; ++r) {
                // p[u] = A[i, j0+jb+u] Xa[i, r], reduced over the wave's 64 rows
                double px[16], py[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    px[u] = av[u].x * xa[r].x - av[u].y * xa[r].y;
                    py[u] = av[u].x * xa[r].y + av[u].y * xa[r].x;
                }
                // distance 32: keep columns 0-7 (lanes < 32) or 8-15
                const bool h32 = (lane & 32) != 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const double sx_ = h32 ? px[u] : px[u + 8], sy_ = h32 ? py[u] : py[u + 8];
                    const double ox = __shfl_xor(sx_, 32), oy = __shfl_xor(sy_, 32);
                    px[u] = (h32 ? px[u + 8] : px[u]) + ox;
                    py[u] = (h32 ? py[u + 8] : py[u]) + oy;
                }
                const bool h16 = (lane & 16) != 0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double sx_ = h16 ? px[u] : px[u + 4], sy_ = h16 ? py[u] : py[u + 4];
                    const double ox = __shfl_xor(sx_, 16), oy = __shfl_xor(sy_, 16);
                    px[u] = (h16 ? px[u + 4] : px[u]) + ox;
                    py[u] = (h16 ? py[u + 4] : py[u]) + oy;
                }
                const bool h8 = (lane & 8) != 0;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const double sx_ = h8 ? px[u] : px[u + 2], sy_ = h8 ? py[u] : py[u + 2];
                    const double ox = __shfl_xor(sx_, 8), oy = __shfl_xor(sy_, 8);
                    px[u] = (h8 ? px[u + 2] : px[u]) + ox;
                    py[u] = (h8 ? py[u + 2] : py[u]) + oy;
                }
                const bool h4 = (lane & 4) != 0;
                {
                    const double sx_ = h4 ? px[0] : px[1], sy_ = h4 ? py[0] : py[1];
                    const double ox = __shfl_xor(sx_, 4), oy = __shfl_xor(sy_, 4);
                    px[0] = (h4 ? px[1] : px[0]) + ox;
                    py[0] = (h4 ? py[1] : py[0]) + oy;
                }
                px[0] += __shfl_xor(px[0], 2);
                py[0] += __shfl_xor(py[0], 2);
                px[0] += __shfl_xor(px[0], 1);
                py[0] += __shfl_xor(py[0], 1);
                // this lane holds column 8 b5 + 4 b4 + 2 b3 + b2 of the 16
                const int u = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
                if ((lane & 3) == 0) scol[wave][r][jb + u] = make_double2(px[0], py[0]);
            }
        }
        __syncthreads();
        if (tid < 64) {  // the tile's partial sums of this chunk's columns
#pragma unroll
            for (int r = 0; r < NRA; ++r) {
                if (r >= nra || tid >= jn) continue;
                double2 t = scol[0][r][tid];
#pragma unroll
                for (int w = 1; w < 4; ++w) {
                    t.x += scol[w][r][tid].x;
                    t.y += scol[w][r][tid].y;
                }
                Pa[((item * ntile + tile) * NRA + r) * n + j0 + tid] = t;
            }
        }
    }
    if (ok) {
#pragma unroll
        for (int r = 0; r < NRF; ++r)
            if (r < nrf) {
                const double2 b = Bf[item * fstride + i + r * ldf];
                Rf[item * fstride + i + r * ldf] = make_double2(b.x - acc[r].x, b.y - acc[r].y);
            }
    }
}

// Ra[:, r] = Ba[:, r] - sum over tiles of the partial sums
template <int NRA>
__global__ __launch_bounds__(256) void k_pair_finish(int64_t n, int nra, int ntile, const double2* __restrict__ Pa,
                                                     const double2* __restrict__ Ba, double2* __restrict__ Ra,
                                                     int64_t lda, int64_t astride) {
    const int64_t item = blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    for (int r = 0; r < nra; ++r) {
        double2 t = make_double2(0.0, 0.0);
        for (int tl = 0; tl < ntile; ++tl) {
            const double2 p = Pa[((item * ntile + tl) * NRA + r) * n + j];
            t.x += p.x;
            t.y += p.y;
        }
        const double2 b = Ba[item * astride + j + r * lda];
        Ra[item * astride + j + r * lda] = make_double2(b.x - t.x, b.y - t.y);
    }
}

int64_t zgemv_pair_ws_elems(int64_t n, int64_t batch) {
    return batch * ((n + kPairRows - 1) / kPairRows) * 4 * n;
}

int zgemv_pair(Context* ctx, int64_t n, int64_t batch, const double2* A, int64_t ld, int64_t stride, int nrf,
               const double2* Xf, const double2* Bf, double2* Rf, int64_t ldf, int64_t fstride, int nra,
               const double2* Xa, const double2* Ba, double2* Ra, int64_t lda, int64_t astride, double2* Pa) {
    if (n == 0 || batch == 0) return MH_OK;
    MH_REQUIRE(batch <= 65535, "matvec: batch too large");
    MH_REQUIRE(nrf >= 1 && nrf <= 4 && nra >= 1 && nra <= 4, "matvec pair: 1 to 4 columns per system");
    const int ntile = (int)((n + kPairRows - 1) / kPairRows);
    ProfScope ps(ctx, "zgemv_pair", 8.0 * (double)n * n * (nrf + nra) * batch, 16.0 * (double)n * n * batch);
    const dim3 grid((unsigned)ntile, (unsigned)batch), fgrid((unsigned)((n + 255) / 256), (unsigned)batch);
#define MH_PAIR(F, R)                                                                                            \
    do {                                                                                                         \
        hipLaunchKernelGGL((k_zgemv_pair<F, R>), grid, dim3(256), 0, ctx->stream, n, A, ld, stride, nrf, Xf, Bf, \
                           Rf, ldf, fstride, nra, Xa, lda, astride, Pa, ntile);                                  \
        hipLaunchKernelGGL((k_pair_finish<R>), fgrid, dim3(256), 0, ctx->stream, n, nra, ntile, Pa, Ba, Ra, lda,  \
                           astride);                                                                             \
    } while (0)
    const int f = nrf <= 1 ? 1 : (nrf <= 2 ? 2 : 4), r = nra <= 1 ? 1 : (nra <= 2 ? 2 : 4);
    if (f == 1 && r == 1) MH_PAIR(1, 1);
    else if (f == 1 && r == 2) MH_PAIR(1, 2);
    else if (f == 1) MH_PAIR(1, 4);
    else if (f == 2 && r == 1) MH_PAIR(2, 1);
    else if (f == 2 && r == 2) MH_PAIR(2, 2);
    else if (f == 2) MH_PAIR(2, 4);
    else if (r == 1) MH_PAIR(4, 1);
    else if (r == 2) MH_PAIR(4, 2);
    else MH_PAIR(4, 4);
#undef MH_PAIR
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

#define MH_ZGEMV_DISPATCH(KERNEL, GRID)                                                                  \
    do {                                                                                                 \
        if (nr <= 1)                                                                                     \
            hipLaunchKernelGGL(KERNEL<1>, GRID, dim3(256), 0, ctx->stream, n, nr, S, ld, stride, Xc, ldx, \
                               xstride, Yc, ldy, ystride, Zc);                                           \
        else if (nr <= 2)                                                                                \
            hipLaunchKernelGGL(KERNEL<2>, GRID, dim3(256), 0, ctx->stream, n, nr, S, ld, stride, Xc, ldx, \
                               xstride, Yc, ldy, ystride, Zc);                                           \
        else if (nr <= 4)                                                                                \
            hipLaunchKernelGGL(KERNEL<4>, GRID, dim3(256), 0, ctx->stream, n, nr, S, ld, stride, Xc, ldx, \
                               xstride, Yc, ldy, ystride, Zc);                                           \
        else if (nr <= 8)                                                                                \
            hipLaunchKernelGGL(KERNEL<8>, GRID, dim3(256), 0, ctx->stream, n, nr, S, ld, stride, Xc, ldx, \
                               xstride, Yc, ldy, ystride, Zc);                                           \
        else                                                                                             \
            hipLaunchKernelGGL(KERNEL<16>, GRID, dim3(256), 0, ctx->stream, n, nr, S, ld, stride, Xc,     \
                               ldx, xstride, Yc, ldy, ystride, Zc);                                      \
    } while (0)

int zgemm_nn_batched(Context* ctx, int64_t n, int64_t nrhs, int64_t batch, const double2* S,
                     int64_t ld, int64_t stride, const double2* X, int64_t ldx, int64_t xstride,
                     double2* Y, int64_t ldy, int64_t ystride, const double2* Z) {
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(batch <= 65535, "matvec: batch too large");
    ProfScope ps(ctx, "zgemv", 8.0 * (double)n * n * nrhs * batch,
                 16.0 * (double)n * n * batch * (double)((nrhs + kMaxRhs - 1) / kMaxRhs));
    for (int64_t r0 = 0; r0 < nrhs; r0 += kMaxRhs) {
        const int64_t nr = (nrhs - r0) < kMaxRhs ? (nrhs - r0) : kMaxRhs;
        const double2* Xc = X + r0 * ldx;
        double2* Yc = Y + r0 * ldy;
        const double2* Zc = Z ? Z + r0 * ldy : nullptr;
        MH_ZGEMV_DISPATCH(k_zgemv_nn, dim3((unsigned)((n + 255) / 256), (unsigned)batch));
        MH_HIP_CHECK(hipGetLastError());
    }
    return MH_OK;
}

int zgemm_tn_batched(Context* ctx, int64_t n, int64_t nrhs, int64_t batch, const double2* S,
                     int64_t ld, int64_t stride, const double2* X, int64_t ldx, int64_t xstride,
                     double2* Y, int64_t ldy, int64_t ystride, const double2* Z) {
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(batch <= 65535, "matvec: batch too large");
    ProfScope ps(ctx, "zgemv_t", 8.0 * (double)n * n * nrhs * batch,
                 16.0 * (double)n * n * batch * (double)((nrhs + kMaxRhs - 1) / kMaxRhs));
    for (int64_t r0 = 0; r0 < nrhs; r0 += kMaxRhs) {
        const int64_t nr = (nrhs - r0) < kMaxRhs ? (nrhs - r0) : kMaxRhs;
        const double2* Xc = X + r0 * ldx;
        double2* Yc = Y + r0 * ldy;
        const double2* Zc = Z ? Z + r0 * ldy : nullptr;
        MH_ZGEMV_DISPATCH(k_zgemv_tn, dim3((unsigned)((n + 3) / 4), (unsigned)batch));
        MH_HIP_CHECK(hipGetLastError());
    }
    return MH_OK;
}
#undef MH_ZGEMV_DISPATCH

// ---------------------------------------------------------------------------
// mixed precision: FP32 copies, updates and residual norms for iterative
// refinement of an FP32 LU against the FP64 matrix
// ---------------------------------------------------------------------------
__global__ void k_to_f32(int64_t rows, int64_t cols, const double2* __restrict__ src, int64_t lds,
                         int64_t sstride, float2* __restrict__ dst, int64_t ldd, int64_t dstride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t item = blockIdx.y;
    if (t >= rows * cols) return;
    const int64_t i = t % rows, j = t / rows;
    const double2 v = src[item * sstride + i + j * lds];
    dst[item * dstride + i + j * ldd] = make_float2((float)v.x, (float)v.y);
}

// x += upcast(d)
__global__ void k_add_f32(int64_t rows, int64_t cols, const float2* __restrict__ d, int64_t ldd,
                          int64_t dstride, double2* __restrict__ x, int64_t ldx, int64_t xstride) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t item = blockIdx.y;
    if (t >= rows * cols) return;
    const int64_t i = t % rows, j = t / rows;
    const float2 v = d[item * dstride + i + j * ldd];
    double2& o = x[item * xstride + i + j * ldx];
    o.x += (double)v.x;
    o.y += (double)v.y;
}

// out[item*cols + j] = max_i |r_ij|_1 / max_i |b_ij|_1   (one workgroup per column)
__global__ __launch_bounds__(256) void k_rel_residual(int64_t rows, const double2* __restrict__ r,
                                                      const double2* __restrict__ b, int64_t ld,
                                                      int64_t stride, double* __restrict__ out) {
    const int64_t item = blockIdx.y, j = blockIdx.x;
    const double2* rc = r + item * stride + j * ld;
    const double2* bc = b + item * stride + j * ld;
    // fmax drops NaN operands: a non-finite residual entry is mapped to +Inf
    // first, so a NaN or Inf solution can never look converged
    double mr = 0.0, mb = 0.0;
    for (int64_t i = threadIdx.x; i < rows; i += 256) {
        const double ar = fabs(rc[i].x) + fabs(rc[i].y);
        const double ab = fabs(bc[i].x) + fabs(bc[i].y);
        mr = fmax(mr, isfinite(ar) ? ar : INFINITY);
        mb = fmax(mb, isfinite(ab) ? ab : INFINITY);
    }
    for (int off = 32; off > 0; off >>= 1) {
        mr = fmax(mr, __shfl_xor(mr, off));
        mb = fmax(mb, __shfl_xor(mb, off));
    }
    __shared__ double s[2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s[0][wave] = mr;
        s[1][wave] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double R = fmax(fmax(s[0][0], s[0][1]), fmax(s[0][2], s[0][3]));
        const double B = fmax(fmax(s[1][0], s[1][1]), fmax(s[1][2], s[1][3]));
        out[item * gridDim.x + j] = !isfinite(B) ? INFINITY : ((B > 0.0) ? R / B : R);
    }
}

int convert_to_f32(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const double2* src,
                   int64_t lds, int64_t sstride, float2* dst, int64_t ldd, int64_t dstride) {
    const int64_t tot = rows * cols;
    if (tot == 0 || batch == 0) return MH_OK;
    ProfScope ps(ctx, "to_f32", 0.0, 24.0 * (double)tot * batch);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((tot + 255) / 256), (unsigned)batch), dim3(256), 0,
                       ctx->stream, rows, cols, src, lds, sstride, dst, ldd, dstride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int add_from_f32(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const float2* d,
                 int64_t ldd, int64_t dstride, double2* x, int64_t ldx, int64_t xstride) {
    const int64_t tot = rows * cols;
    if (tot == 0 || batch == 0) return MH_OK;
    hipLaunchKernelGGL(k_add_f32, dim3((unsigned)((tot + 255) / 256), (unsigned)batch), dim3(256), 0,
                       ctx->stream, rows, cols, d, ldd, dstride, x, ldx, xstride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

int rel_residual(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const double2* r,
                 const double2* b, int64_t ld, int64_t stride, double* out) {
    if (rows == 0 || cols == 0 || batch == 0) return MH_OK;
    hipLaunchKernelGGL(k_rel_residual, dim3((unsigned)cols, (unsigned)batch), dim3(256), 0, ctx->stream,
                       rows, r, b, ld, stride, out);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

}  // namespace mh
