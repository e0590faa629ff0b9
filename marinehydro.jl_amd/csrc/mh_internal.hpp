// Host-side internals shared by the libmhip translation units: the context,
// error reporting and the launchers each unit exports to the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mhip.h"
#include "mh_device.hpp"

namespace mh {

// Thread-local last error message (read through mh_last_error()).
void set_error(const std::string& msg);
const char* last_error();

// One timed kernel launch (kernel timer, enabled per context).
struct ProfRec {
    const char* name;
    hipEvent_t start, stop;
    double flops, bytes;  // algorithmic work of the launch
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    // growable device scratch: one bump region per call
    void* ws = nullptr;
    size_t ws_bytes = 0;
    int* d_flag = nullptr;  // device error / info flag (one int)
    // triangular solves, one state per stream that issues them (0: stream,
    // 1: aux): the workgroup ticket (u64) and an out-of-place solution
    // workspace; one shared error word
    struct TriSlot {
        void* sync = nullptr;
        size_t sync_bytes = 0;
        void* ws = nullptr;
        size_t ws_bytes = 0;
    } tri[2];
    int tri_slot = 0;
    int* tri_err = nullptr;
    // second stream for independent work (forward and adjoint solves),
    // default priority, ordered against `stream` by events
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // third stream for independent side work while aux serves the LU
    hipStream_t side = nullptr;
    hipEvent_t ev_side_fork = nullptr, ev_side_join = nullptr;
    // pinned host scratch for small device-to-host reads
    double* h_pin = nullptr;
    size_t h_pin_count = 0;
    // events for ordering work between stream and aux (LU look-ahead)
    std::vector<hipEvent_t> ev_pool;
    // frequency-independent Rankine parts of the assembly (grow-only)
    void* rk = nullptr;
    size_t rk_bytes = 0;
    // split operand images of the FP32 LU's trailing updates (grow-only)
    void* himg = nullptr;
    size_t himg_bytes = 0;
    bool profiling = false;
    std::vector<ProfRec> prof;
    // mixed precision: refinement sweeps allowed, whether a frequency chunk
    // that does not converge is solved again in FP64, and how often that
    // happened since the context was created
    int ir_max_sweeps = 12;
    bool fp64_fallback = true;
    int64_t fallbacks = 0;
    // LU leaf width cap (0: the default width for the panel height; 4 or 8)
    int leaf_cap = 0;
    // assembly: the Rankine and image far field (pairs beyond 7 panel radii)
    // in FP32 (mh_set_far_field_fp32)
    bool far32 = false;
    // gradients: how the wave-term pair sums are differentiated
    // (mh_set_gradient_mode: 0 chosen per sweep, 1 forward tangents, 2 reverse)
    int grad_mode = 0;
    // graph capture (graph.hip): while set, sweep_impl only issues device
    // work (its streams belong to a capture begun by the caller); the
    // frequencies, wavenumbers, headings and frequency tangents come from
    // the device buffers in `cap`, which the graph's owner fills before each
    // launch, and the host checks become device check words in cap->check
    bool capturing = false;
    struct CapIO* cap = nullptr;
    // refinement sweeps recorded in a captured mixed sweep (at least 2)
    int ir_graph_sweeps = 2;
    // mixed sweeps with gradients: refinement sweeps made before the
    // convergence check, which then runs on the aux stream beside the
    // gradient kernels and is read at the end of the chunk (0: the check
    // after every sweep); chunks whose deferred check failed and were
    // solved again with a check after every sweep
    int ir_deferred_sweeps = 2;
    int64_t deferred_reruns = 0;
    // sweeps: frequencies per chunk at most (0: as many as fit 32 GiB)
    int64_t chunk_limit = 0;
    // multi-body scheduler: worker contexts (each with a stream of its own)
    // and the configuration the autotuner chose (0: not tuned)
    bool owns_stream = false;
    std::vector<Context*> workers;
    int tuned_workers = 0, tuned_leaf = 0;
};

// Device inputs and check words of one captured sweep (graph.hip).  om, k:
// nomega frequencies and wavenumbers; be: the headings; dk, dom: the
// wavenumber and frequency tangents of every parameter block
// (omega_tangents' layout); check: kCheckWords words cleared at the start of
// the sweep, then [0] the mesh-index flag, [1] the OR of the LU info words,
// [2] the refinement verdict (1: a column above the tolerance, 2: a
// non-finite residual).  A nonzero word means the sweep must be run again
// without the graph, which reports or repairs what happened.
constexpr int kCheckWords = 4;
struct CapIO {
    double *om = nullptr, *k = nullptr, *be = nullptr, *dk = nullptr, *dom = nullptr;
    int* check = nullptr;
};

// Tangent directions per gradient pass for nparam parameters (1, 2 or 4).
inline int grad_block(int64_t nparam) { return nparam <= 1 ? 1 : (nparam == 2 ? 2 : 4); }
// Wavenumber and frequency tangents of every parameter block for the
// frequencies [w0, w0 + nw) of a sweep: dk and dom hold
// ceil(nparam / Pg) blocks of nw x Pg values (zero past nparam);
// dk = 2 omega / g * domega (k = omega^2 / g, src/waves.jl:30).
void omega_tangents(const double* omegas, int64_t nomega, int64_t w0, int64_t nw, const double* domega,
                    int64_t nparam, int Pg, double g, double* dk, double* dom);

// Creates a context on `device` (current afterwards) using `stream`, or a
// new non-blocking stream of default priority owned by the context when
// own_stream is set; context_free releases everything, workers included.
int context_new(int device, hipStream_t stream, bool own_stream, Context** out);
void context_free(Context* c);

// Optional gradient request of a sweep (mh_freq_sweep_grad).
struct GradArgs {
    int64_t nparam;
    const double *dV, *dC, *dN, *dA, *ddofs;  // device tangents (may be NULL)
    const double* domega;                      // host, nparam x nomega (may be NULL)
    double *dAout, *dBout, *dFout;             // device outputs
};
// One body's frequency sweep (mh_freq_sweep / mh_freq_sweep_grad after
// their argument checks), synchronous.
int sweep_impl(Context* c, const mh_mesh_t* mesh, int gf_mask, int direct, int precision, int64_t nomega,
               const double* omegas, double g, double rho, int64_t ndof, const double* dof_fields, int64_t nbeta,
               const double* betas, double* A, double* B, double* F, const GradArgs* ga);
// Argument checks of a sweep, without the context.
int sweep_args_ok(const mh_mesh_t* mesh, int gf_mask, int precision, int64_t nomega, const double* omegas,
                  double g, double rho, int64_t ndof, const double* dof_fields, int64_t nbeta,
                  const double* betas, double* A, double* B, double* F);

// Brackets one kernel launch with events on the context's stream when the
// kernel timer is on; no cost otherwise.
struct ProfScope {
    Context* c;
    int idx = -1;
    ProfScope(Context* ctx, const char* name, double flops, double bytes) : c(ctx) {
        if (!c->profiling) return;
        ProfRec r{name, nullptr, nullptr, flops, bytes};
        if (hipEventCreate(&r.start) != hipSuccess || hipEventCreate(&r.stop) != hipSuccess) return;
        (void)hipEventRecord(r.start, c->stream);
        c->prof.push_back(r);
        idx = (int)c->prof.size() - 1;
    }
    ~ProfScope() {
        if (idx >= 0) (void)hipEventRecord(c->prof[idx].stop, c->stream);
    }
};

// Grows ctx->ws to at least `bytes` (synchronous hipMalloc; only when larger).
bool ensure_workspace(Context* ctx, size_t bytes);
// Creates ctx->aux and its fork/join events on first use.
int ensure_aux(Context* ctx);
// At least `count` events (timing disabled) in ctx->ev_pool.
int ensure_events(Context* ctx, size_t count);

#define MH_HIP_CHECK(expr)                                                   \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                              \
            ::mh::set_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
            return MH_ERR_HIP;                                               \
        }                                                                    \
    } while (0)

#define TRY_RC(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != MH_OK) return rc_; \
    } while (0)

#define MH_REQUIRE(cond, msg)                                                \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ::mh::set_error(msg);                                            \
            return MH_ERR_ARGUMENT;                                          \
        }                                                                    \
    } while (0)

// ---- launchers (assemble.hip) ---------------------------------------------
// Packs the mesh into PanelRec records; a face index out of range is flagged
// in flag_buf (one int, cleared by a fill kernel first) or, when flag_buf is
// null, in ctx->d_flag, and returned as MH_ERR_ARGUMENT (synchronises).
int launch_pack_panels(Context* ctx, const mh_mesh_t* mesh, PanelRec* out, int* flag_buf = nullptr);
// p[0 .. count) = 0 by a fill kernel on ctx->stream (capturable, unlike a memset)
int clear_ints(Context* ctx, int* p, int64_t count);
// `bytes` (a multiple of 16, 16-byte aligned) of p set to zero by a fill kernel
int fill_zero(Context* ctx, void* p, size_t bytes);
// dst[i + j ldd] = src[i + j lds] for i < rows, j < cols, by a copy kernel:
// the sweeps record no memcpy nodes when captured (a 1D device-to-device
// copy recorded from hipMemcpyAsync kept zero pitches in its node and its
// graph crashed the host at launch)
int copy_d2(Context* ctx, double2* dst, int64_t ldd, const double2* src, int64_t lds, int64_t rows, int64_t cols);
// *dst |= src[0] | ... | src[count - 1]  (one workgroup)
int or_words(Context* ctx, const int* src, int64_t count, int* dst);
// *dst |= 1 if some nrm[e] is not below tol, |= 2 if some is not finite
int ir_verdict(Context* ctx, const double* nrm, int64_t count, double tol, int* dst);
// Assembly in one kernel per launch, without the frequency-independent
// scratch matrices launch_assemble keeps in the context.
int launch_assemble_onepass(Context* ctx, const mh_mesh_t* mesh, const PanelRec* panels, int gf_mask,
                            int form, int64_t nk, const double* d_k, double2* S, double2* D, double2* K,
                            int64_t ld, int64_t stride);
int launch_assemble(Context* ctx, const mh_mesh_t* mesh, const PanelRec* panels,
                    int gf_mask, int form, int64_t nk, const double* d_k, double2* S,
                    double2* D, double2* K, int64_t ld, int64_t stride, float2* DK32 = nullptr);
// Table of the Delhommeau term on ctx's device (built on first use, then
// shared by every context on the device until the process ends).
int get_delh_table(Context* ctx, DelhTab* out);
int launch_gf_eval(Context* ctx, int gf_mask, int op, int wrt_first, int64_t npairs,
                   const double* points, const mh_panel_t* panels, double k, double2* out);
int launch_wave_term(Context* ctx, int term, int64_t n, const double* r, const double* z,
                     double2* out);

// ---- launchers (solve.hip) --------------------------------------------------
// LUt (optional): the transposed factors (n x n, same ld and stride), each
// block row transposed as soon as it is final, alongside the factorisation.
// inv (optional, with perm optional): what tri_prepare produces, built block
// column by block column alongside the factorisation.
int lu_factor(Context* ctx, int64_t n, int64_t ncols, int64_t batch, double2* A, int64_t ld,
              int64_t stride, int64_t* ipiv, int* d_info, double2* LUt = nullptr, int* perm = nullptr,
              double2* inv = nullptr);
// ---- launchers (trsv.hip) ---------------------------------------------------
constexpr int kMaxSolveN = 4096;
// After an LU factorisation: perm[b*n + i], the net row permutation of the
// interchanges ipiv ((P b)[i] = b[perm[i]]; skipped when perm is null), and
// the inverses of the 32 x 32 diagonal blocks of L and U that the solves use
// (inv holds tri_inv_elems(n, batch) elements).
int64_t tri_inv_elems(int64_t n, int64_t batch);
int tri_prepare(Context* ctx, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                const int64_t* ipiv, int* perm, double2* inv);
int tri_prepare_f32(Context* ctx, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                    const int64_t* ipiv, int* perm, float2* inv);
// The same for block column [k0, k0+w) of a factorisation in progress (its
// pivots and diagonal blocks final): its interchanges folded into perm (which
// starts from the identity at k0 = 0; perm may be null) and the inverses of
// its diagonal blocks.  Block columns are processed in order.
int tri_prepare_part(Context* ctx, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                     const int64_t* ipiv, int* perm, double2* inv, int64_t k0, int w);
int tri_prepare_part(Context* ctx, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                     const int64_t* ipiv, int* perm, float2* inv, int64_t k0, int w);
// LUt = LU^T per item (same ld and stride), for transposed solves that read
// the factors row-wise; only rows [r_lo, r_hi) of LU (columns of LUt) when
// given (r_hi < 0: to n).
int tri_transpose(Context* ctx, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
                  double2* LUt, int64_t r_lo = 0, int64_t r_hi = -1);
int tri_transpose_f32(Context* ctx, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                      float2* LUt, int64_t r_lo = 0, int64_t r_hi = -1);
// Solves with the factors of A = P L U, in place on B (n x nrhs per item):
// trans 0: A X = B, 1: A^T X = B; part 1 = L (with P), 2 = U, 3 = both.
// LUt (optional): tri_transpose of LU, used by the transposed solves.
int tri_solve(Context* ctx, int64_t n, int64_t batch, const double2* LU, int64_t ld, int64_t stride,
              const int* perm, const double2* inv, int trans, int part, int64_t nrhs, double2* B,
              int64_t ldb, int64_t bstride, const double2* LUt = nullptr);
int tri_solve_f32(Context* ctx, int64_t n, int64_t batch, const float2* LU, int64_t ld, int64_t stride,
                  const int* perm, const float2* inv, int trans, int part, int64_t nrhs, float2* B,
                  int64_t ldb, int64_t bstride, const float2* LUt = nullptr);
// Reports (and clears) a failed inter-workgroup wait or a bad pivot index of
// the solves issued so far; synchronises the context stream.
int tri_check(Context* ctx);
// Reads the LU info words of a batch (synchronises): zero pivots and
// non-finite entries become MH_ERR_SINGULAR / MH_ERR_NOT_CONVERGED.
int check_info(Context* c, const int* info, int64_t batch);

// Y[b] = S[b]^T * X[b]
int zgemm_tn_batched(Context* ctx, int64_t n, int64_t nrhs, int64_t batch, const double2* S,
                     int64_t ld, int64_t stride, const double2* X, int64_t ldx, int64_t xstride,
                     double2* Y, int64_t ldy, int64_t ystride, const double2* Z = nullptr);
// Y[b] = S[b] * X[b]  (n x n times n x nrhs, complex)
// (with Z: Y[b] = Z[b] - S[b] * X[b], Z laid out like Y)
int zgemm_nn_batched(Context* ctx, int64_t n, int64_t nrhs, int64_t batch, const double2* S,
                     int64_t ld, int64_t stride, const double2* X, int64_t ldx, int64_t xstride,
                     double2* Y, int64_t ldy, int64_t ystride, const double2* Z = nullptr);
// Rf = Bf - S Xf and Ra = Ba - S^T Xa (1 to 4 columns each, n x cols per
// item with the given leading dimensions and strides) in one pass over S;
// Pa: workspace of zgemv_pair_ws_elems(n, batch) complex entries.
int64_t zgemv_pair_ws_elems(int64_t n, int64_t batch);
int zgemv_pair(Context* ctx, int64_t n, int64_t batch, const double2* S, int64_t ld, int64_t stride, int nrf,
               const double2* Xf, const double2* Bf, double2* Rf, int64_t ldf, int64_t fstride, int nra,
               const double2* Xa, const double2* Ba, double2* Ra, int64_t lda, int64_t astride, double2* Pa);
// FP32 factorisation and solves (mixed precision)
int lu_factor_f32(Context* ctx, int64_t n, int64_t ncols, int64_t batch, float2* A, int64_t ld,
                  int64_t stride, int64_t* ipiv, int* d_info, float2* LUt = nullptr, int* perm = nullptr,
                  float2* inv = nullptr);
int convert_to_f32(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const double2* src,
                   int64_t lds, int64_t sstride, float2* dst, int64_t ldd, int64_t dstride);
int add_from_f32(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const float2* d,
                 int64_t ldd, int64_t dstride, double2* x, int64_t ldx, int64_t xstride);
// out[b*cols + j] = max|r[:, j]| / max|b[:, j]| (|.| = |re| + |im|)
int rel_residual(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const double2* r,
                 const double2* b, int64_t ld, int64_t stride, double* out);

// ---- launchers (physics.hip) ------------------------------------------------
// rhs[w] columns: ndof radiation BCs then nbeta Airy BCs, for each omega.
int launch_bcs(Context* ctx, const mh_mesh_t* mesh, int64_t nomega, const double* d_omega,
               double g, int64_t ndof, const double* dof_fields, int64_t nbeta,
               const double* d_beta, double2* rhs, int64_t ldb, int64_t bstride);
int launch_forces(Context* ctx, const mh_mesh_t* mesh, int64_t nomega, const double* d_omega,
                  double rho, int64_t ndof, const double* dof_fields, int64_t nbeta,
                  const double2* phi, int64_t ldb, int64_t bstride, double* A, double* B,
                  double2* F);

// Froude-Krylov force FK (nomega x nbeta x ndof complex) and, when np > 0 and
// dFK is given, its derivatives along np parameters (nomega x np x nbeta x
// ndof); tangents as in mh_freq_sweep_grad, d_dom on the device (nomega x np).
int launch_froude_krylov(Context* ctx, const mh_mesh_t* mesh, int64_t nomega, const double* d_omega,
                         double g, double rho, int64_t ndof, const double* dofs, int64_t nbeta,
                         const double* d_beta, double2* FK, int64_t np, const double* dC,
                         const double* dN, const double* dA, const double* ddofs, const double* d_dom,
                         double2* dFK);

// ---- launchers (vjp.hip) ----------------------------------------------------
// Reverse mode of the assembly: cotangents Sb, Db (nk complex matrices, ld,
// stride; either may be null) -> dV (nv x 3), dC, dN (nf x 3), dA (nf) summed
// over the wavenumbers, and dk (nk); outputs may be null.  ws holds
// assemble_vjp_ws_doubles(n, nk) doubles.
int64_t assemble_vjp_ws_doubles(int64_t n, int64_t nk);
int launch_assemble_vjp(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, int gf, int form, int64_t nk,
                        const double* d_k, const double2* Sb, const double2* Db, int64_t ld, int64_t stride,
                        double* ws, double* dV, double* dC, double* dN, double* dA, double* dk);
// Y = conj(X), n x cols per item (X and Y may be the same buffer)
int conj_batched(Context* ctx, int64_t rows, int64_t cols, int64_t batch, const double2* X, int64_t ldx, int64_t sx,
                 double2* Y, int64_t ldy, int64_t sy);
// M = alpha X Y^H (X, Y n x nr with ldb, bs; M n x n with ld, stride)
int outer_conj_batched(Context* ctx, int64_t n, int64_t nr, int64_t batch, double alpha, const double2* X,
                       const double2* Y, int64_t ldb, int64_t bs, double2* M, int64_t ld, int64_t stride);

// ---- launchers (grad.hip) ---------------------------------------------------
int launch_pack_tangents(Context* ctx, const mh_mesh_t* m, int P, int64_t nparam, int64_t m0,
                         const double* dV, const double* dC, const double* dN, const double* dA,
                         double* tan);
int launch_tangent_pairs(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, const double* tan,
                         int P, int gf, int form, int64_t nw, const double* d_k, const double* d_dk,
                         int64_t ndof, int64_t nr, const double2* aS, const double2* aD,
                         const double2* sS, const double2* sD, int64_t wa, int64_t ws, double2* part,
                         int64_t* nblk_out);
// Frequency-independent Rankine + image tangents dR (2 P real n x n
// matrices) and their contraction with every frequency's vectors, added to
// out[w][d][r][m]; Y holds 2 P n x (nw nr) complex.
int launch_rankine_tangents(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, const double* tan, int P,
                            int gf, int form, double* dR);
int launch_rankine_contract(Context* ctx, int64_t n, int P, int64_t nw, int64_t ndof, int64_t nr, const double* dR,
                            const double2* aS, const double2* aD, const double2* sS, const double2* sD, int64_t wa,
                            int64_t ws, double2* Y, double2* out);
int launch_force_weights(Context* ctx, const mh_mesh_t* m, int64_t nw, const double* d_om, double rho,
                         int64_t ndof, const double* dofs, double2* cw, int64_t wstride);
int launch_bc_tangent(Context* ctx, const mh_mesh_t* m, int P, const double* tan, int64_t nw,
                      const double* d_om, const double* d_dom, double g, double rho, int64_t ndof,
                      const double* dofs, const double* ddofs, int64_t nparam, int64_t m0,
                      int64_t nbeta, const double* d_beta, const double2* phi, int64_t bstride,
                      const double2* rhoW, int64_t wstride, double2* out);
int launch_reduce_partials(Context* ctx, int64_t nw, int64_t nblk, int64_t nout, const double2* part,
                           double2* out, int accumulate);
// ---- reverse mode of the wave-term pair sums (grad_rev.hip) ------------------
// Bytes of the work buffer of launch_wave_cot for n panels, nw frequencies
// and ndof x nr outputs.
int64_t wave_cot_ws_bytes(int64_t n, int64_t nw, int64_t ndof, int64_t nr);
// Per-panel cotangents (centre, normal, area; n x 7 complex per frequency and
// output) and per-frequency k cotangents of the wave-term pair sums, weights
// as launch_tangent_pairs; *cot_out, *cotk_out point into `work`.
int launch_wave_cot(Context* ctx, const mh_mesh_t* m, const PanelRec* panels, int gf, int form, int64_t nw,
                    const double* d_k, int64_t ndof, int64_t nr, const double2* aS, const double2* aD,
                    const double2* sS, const double2* sD, int64_t wa, int64_t ws, void* work, double2** cot_out,
                    double2** cotk_out);
// out[w][d][r][m] += the cotangents contracted with the tangents of
// parameters m0 .. m0 + Pg - 1 (dC, dN: nparam x n x 3 column-major per
// parameter, dA: nparam x n; null = zero) and dk (nw x Pg).
int launch_wave_cot_contract(Context* ctx, int64_t n, int64_t nw, int64_t ndof, int64_t nr, int Pg, int64_t nparam,
                             int64_t m0, const double2* cot, const double2* cotk, const double* dC, const double* dN,
                             const double* dA, const double* dk, double2* out);
// dA, dB (nw x nparam x ndof x ndof) and dF (complex, nw x nparam x nbeta x
// ndof) of parameters m0 .. m0+P-1 from the pair terms T1 + T2, the forces A,
// B of the same frequencies and the omega tangents dom (nw x P); null
// outputs are skipped.
int launch_grad_finish(Context* ctx, int64_t nw, int64_t np, int64_t m0, int P, int64_t ndof, int64_t nr,
                       int64_t nbeta, const double* d_om, const double* d_dom, const double* A, const double* B,
                       const double2* T1, const double2* T2, double* gA, double* gB, double* gF);

}  // namespace mh
