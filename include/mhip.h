/*
 * mhip.h - C ABI of libmhip.so, the gfx950 (MI355X) boundary-element hot path
 * of MarineHydro.jl: Green-function panel integrals, fused S/D/K influence
 * matrix assembly, batched dense complex LU solve, boundary conditions and
 * forces.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer unless the function name ends in
 *    `_host`.  Short parameter lists (wavenumbers, frequencies, headings) are
 *    host arrays, as noted.
 *  - Matrices are column-major complex128 (interleaved re, im), exactly the
 *    memory layout of a Julia Matrix{ComplexF64}; (i, j) of batch item b sits
 *    at complex offset b*stride + i + j*ld.  Strides and leading dimensions
 *    are counted in complex elements.
 *  - Mesh arrays use the reference's Julia layout (src/meshes.jl:12-37):
 *    vertices nv x 3, faces nf x 4 (0-based, src/meshes.jl:179), centers and
 *    normals nf x 3, all column-major; areas and radii nf.
 *  - Functions return MH_OK (0) or an error code; mh_last_error() returns a
 *    message for the calling thread.  No exception crosses the ABI.
 *  - All work of one context is ordered on that context's HIP stream; the
 *    call returns after enqueueing unless the function says it synchronises.
 *    Contexts are not thread-safe; use one per host thread.
 *
 * Reference paths below are relative to the MarineHydro.jl repository root.
 */
#ifndef MHIP_H
#define MHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHIP_VERSION 100

enum mh_status {
    MH_OK = 0,
    MH_ERR_ARGUMENT = 1,        /* bad size, pointer or option            */
    MH_ERR_HIP = 2,             /* HIP runtime error                      */
    MH_ERR_SINGULAR = 3,        /* exact zero pivot in the LU             */
    MH_ERR_NOT_SUPPORTED = 4,   /* option not supported by the library   */
    MH_ERR_NO_DEVICE = 5,       /* no usable gfx950 device                */
    MH_ERR_NOT_CONVERGED = 6    /* iterative refinement did not converge  */
};

/* Green-function terms, summed like the reference's tuple of GreensFunction
 * objects (src/green_functions/abstract_greens_function.jl:18-83).       */
enum mh_gf {
    MH_GF_RANKINE = 1,               /* Rankine(), rankine.jl               */
    MH_GF_RANKINE_REFLECTED = 2,     /* RankineReflected()                  */
    MH_GF_RANKINE_REFLECTED_NEG = 4, /* RankineReflectedNegative()          */
    MH_GF_WU = 8,                    /* GFWu(), wu.jl                       */
    MH_GF_DELHOMMEAU = 16,           /* ExactGuevelDelhommeau()             */
    /* The same term from a table of the exact one (bicubic Hermite over
     * nested (r, z) grids, log singularity subtracted, exact beyond
     * r, |z| > 32): no reference counterpart; within 1e-9 of the exact term
     * and about as fast as GFWu.  Derivative entry points differentiate the
     * exact term in its place.  The mixed-precision sweep substitutes it for
     * MH_GF_DELHOMMEAU in the assembly. */
    MH_GF_DELHOMMEAU_TABLE = 32
};
/* The tuple used by calculate_radiation_forces / DiffractionForce
 * (src/matrix_assembly.jl:100-102, src/waves.jl:41-45). */
#define MH_GF_DEFAULT (MH_GF_RANKINE | MH_GF_RANKINE_REFLECTED | MH_GF_WU)

/* Formulation: the reference's `direct` keyword (src/matrix_assembly.jl:4). */
enum mh_formulation { MH_DIRECT = 1, MH_INDIRECT = 2, MH_BOTH = 3 };

/* Per-pair operation for mh_gf_eval, one per reference generic function
 * (src/green_functions/abstract_greens_function.jl:16-83). */
enum mh_gf_op {
    MH_OP_GREENS = 0,             /* greens(gf, e1, e2, k)                 */
    MH_OP_GRADIENT_GREENS = 1,    /* gradient_greens(...; wrt_first)       */
    MH_OP_INTEGRAL = 2,           /* integral(gf, e1, e2, k)               */
    MH_OP_INTEGRAL_GRADIENT = 3   /* integral_gradient(...; wrt_first)     */
};

/* Arithmetic of the dense solve. */
enum mh_precision {
    MH_FP64 = 0,    /* FP64 throughout                                       */
    MH_MIXED = 1    /* FP32 factorisation + FP64 iterative refinement         */
};

typedef struct mh_context* mh_context_t;

/* One source panel, the fields the reference reads through its element
 * accessors (src/meshes.jl:157-205): center, vertices (row v = vertex v),
 * normal, area, radius.  160 bytes. */
typedef struct {
    double center[3];
    double vertices[4][3];
    double normal[3];
    double area;
    double radius;
} mh_panel_t;

/* Device-resident mesh, field for field the reference's Mesh
 * (src/meshes.jl:12-37). */
typedef struct {
    int64_t nfaces;
    int64_t nvertices;
    const double* vertices; /* nvertices x 3, column-major           */
    const int64_t* faces;   /* nfaces x 4, column-major, 0-based     */
    const double* centers;  /* nfaces x 3, column-major              */
    const double* normals;  /* nfaces x 3, column-major              */
    const double* areas;    /* nfaces                                */
    const double* radii;    /* nfaces                                */
} mh_mesh_t;

/* ---- context ---------------------------------------------------------- */
int mh_version(void);
const char* mh_last_error(void);
/* stream: a hipStream_t, or NULL for the null stream.  No reference
 * counterpart (the reference runs on the host). */
int mh_context_create(int device, void* stream, mh_context_t* out);
int mh_context_destroy(mh_context_t ctx);
int mh_synchronize(mh_context_t ctx);

/* Kernel timer (no reference counterpart; the reference is profiled with
 * BenchmarkTools scripts under benchmark/).  When enabled, every kernel the context
 * launches is bracketed by HIP events on the context's stream.
 * mh_profile_report synchronises, writes a JSON object
 * {"kernel": {"launches": n, "ms": t, "flops": f, "bytes": b}, ...} with the
 * algorithmic flops and bytes of those launches into buf (NUL-terminated,
 * at most cap bytes) and clears the records. */
int mh_profile_enable(mh_context_t ctx, int on);
int mh_profile_report(mh_context_t ctx, char* buf, int64_t cap);

/* solve(D, S, bc; direct) (src/matrix_assembly.jl:105-133) in mixed
 * precision: an FP32 copy of D (K) is factored (trailing updates on split
 * fp16 operands with a power-of-two scale per row and column, so no finite
 * magnitude is out of range) and phi refined in FP64 against D (K) to a
 * 1e-10 relative residual.  Arguments as mh_solve; phi must not alias bc.
 * D (K) is left intact unless the refinement fails and the context allows
 * the FP64 fallback (mh_set_refinement), which solves in place as mh_solve. */
int mh_solve_mixed(mh_context_t ctx, int64_t n, int64_t batch, int direct, double* DK, const double* S,
                   int64_t ld, int64_t stride, const double* bc, int64_t nrhs, double* phi, int64_t ldb,
                   int64_t bstride);

/* Mixed-precision control (no reference counterpart; the reference solves
 * in FP64 only).  max_sweeps: iterative-refinement sweeps allowed per
 * frequency chunk (default 12, range 1-100); fp64_fallback (default 1): a
 * chunk whose refinement does not reach the 1e-10 relative residual or
 * whose FP32 factorisation meets an exact zero pivot is solved again in
 * FP64 in the same call (the FP32 factorisation scales every row and column
 * of its matrix-core operands by a power of two, so no finite magnitude is
 * out of its range); with 0 the call returns
 * MH_ERR_NOT_CONVERGED / MH_ERR_SINGULAR instead.  A NaN or Inf residual is
 * always returned as MH_ERR_NOT_CONVERGED.  mh_fallback_count: chunks solved
 * again in FP64 since the context was created. */
int mh_set_refinement(mh_context_t ctx, int max_sweeps, int fp64_fallback);
/* Far-field panel pairs in FP32 (no reference counterpart; the reference
 * evaluates every pair in FP64): with on != 0, every assembly on ctx
 * (mh_assemble, the sweeps) evaluates the Rankine and image terms of pairs
 * beyond 7 panel radii (the far-field branch of src/green_functions/
 * rankine.jl:33-34, :74-77: area / r and its gradient) in FP32, from offsets
 * formed in FP64; the branch is decided on r^2 > 49 R^2.  Nearer pairs, the
 * wave terms and everything after the assembly stay as they are.  Such an
 * entry is within ~2e-7 of its FP64 value relative to its magnitude; the
 * refinement of a mixed sweep then converges to the solution of these
 * matrices, so the results differ from the FP64 path by that order instead
 * of ~1e-13.  Default off; the gradient kernels always use FP64. */
int mh_set_far_field_fp32(mh_context_t ctx, int on);
/* How mh_freq_sweep_grad differentiates the wave-term pair sums (no
 * reference counterpart; the reference takes Zygote's reverse mode through
 * the whole path): 1 = forward tangents, one pass over the panel pairs per
 * block of at most 4 parameters (and 4 dofs x 4 right-hand sides); 2 =
 * reverse, one pass per block of 2 outputs (dof x right-hand side) that
 * leaves cotangents of every panel's centre, normal and area and of k,
 * contracted afterwards with any number of parameter tangents; 0 (default)
 * = whichever costs less, a reverse pass counted as 1.6 forward passes of
 * up to 2 parameters and 0.8 of one of 3-4.  Both give the same
 * derivatives up to rounding. */
int mh_set_gradient_mode(mh_context_t ctx, int mode);
int mh_fallback_count(mh_context_t ctx, int64_t* count);
/* Deferred refinement check (no reference counterpart; the reference solves
 * in FP64): a mixed-precision sweep with gradients makes `sweeps` refinement
 * sweeps (default 2; at most the mh_set_refinement limit) and checks the
 * last one's residuals on the device,
 * beside the gradient kernels, reading the verdict at the end of each
 * frequency chunk; a chunk that missed the 1e-10 tolerance (or whose FP32
 * factors had a zero pivot or overflowed) is solved again with the check
 * after every sweep of mh_set_refinement, so the results and errors are
 * those of the undeferred sweep.  sweeps = 0: no deferral.
 * mh_deferred_rerun_count: chunks solved again that way since the context
 * was created. */
int mh_set_deferred_check(mh_context_t ctx, int sweeps);
int mh_deferred_rerun_count(mh_context_t ctx, int64_t* count);
/* Frequencies per chunk of the sweeps at most (no reference counterpart;
 * the reference solves one frequency at a time): the sweeps assemble,
 * factor and solve their frequencies in chunks that keep the matrices under
 * 32 GiB; max_frequencies > 0 caps a chunk further (for callers short of
 * device memory), 0 restores the default. */
int mh_set_sweep_chunk(mh_context_t ctx, int64_t max_frequencies);

/* ---- Green functions ---------------------------------------------------
 * Batched per-pair evaluation of the sum of the selected terms.
 * Replaces greens / gradient_greens / integral / integral_gradient
 * (src/green_functions/abstract_greens_function.jl:16-83; terms in
 * rankine.jl:8-128, rankine_reflected.jl:1-44,
 * rankine_reflected_negative.jl:1-17, wu.jl:5-47,
 * exact_Guevel_Delhommeau.jl:28-75).
 *   points : npairs x 3, column-major (centre of element_1)
 *   panels : npairs records (element_2)
 *   out    : complex128, npairs values (ops 0, 2) or npairs x 3 column-major
 *            complex 3-vectors (ops 1, 3). */
int mh_gf_eval(mh_context_t ctx, int gf_mask, int op, int wrt_first,
               int64_t npairs, const double* points, const mh_panel_t* panels,
               double wavenumber, double* out);

/* Dimensionless wave terms G(r, z), dG/dr, dG/dz for npoints (r, z) pairs:
 * term = MH_GF_WU (wu.jl:5-30) or MH_GF_DELHOMMEAU
 * (exact_Guevel_Delhommeau.jl:15-59).  out: npoints x 3 complex,
 * column-major. */
int mh_wave_term_eval(mh_context_t ctx, int term, int64_t npoints,
                      const double* r, const double* z, double* out);

/* ---- assembly ----------------------------------------------------------
 * assemble_matrices(green_functions, mesh, k; direct) for nk wavenumbers in
 * one launch (src/matrix_assembly.jl:4-42, :97; the Wu tuple of :100-102).
 * S_ij = -1/(4 pi) sum_gf integral; D_ij = delta_ij/2 - 1/(4 pi) n_j . grad_xi;
 * K_ij = delta_ij/2 - 1/(4 pi) n_i . grad_x.  MH_BOTH writes D and K.
 * wavenumbers: HOST array of nk values.  S, D, K: nk matrices each (D or K
 * may be NULL when not requested). */
int mh_assemble(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask,
                int formulation, int64_t nk, const double* wavenumbers,
                double* S, double* D, double* K, int64_t ld, int64_t stride);

/* mh_assemble without any allocation in the call.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The panel records, wavenumbers and error flag live in one
 * fixed device array of the library (about 0.8 MB per device, held until
 * the process ends) instead of the context workspace; the one-pass pair
 * kernel needs no other scratch.  Same arguments and results as
 * mh_assemble; at most 4096 panels and 256 wavenumbers per call (larger
 * sizes return MH_ERR_ARGUMENT).  Calls on one device share the array:
 * calls that overlap on a device (two threads or unordered streams) race.
 * Synchronises (reads the mesh-index flag). */
int mh_assemble_noalloc(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask,
                        int formulation, int64_t nk, const double* wavenumbers,
                        double* S, double* D, double* K, int64_t ld, int64_t stride);

/* ---- dense solve -------------------------------------------------------
 * Batched LU with partial pivoting (zgetrf semantics; pivot = max |re|+|im|)
 * in place; ipiv: n x batch, 0-based (pivot row of step k of item b at
 * ipiv[b*n + k]).  Returns MH_ERR_SINGULAR if any exact zero pivot occurs
 * (synchronises to read the flag).  Replaces the LAPACK factorisation behind
 * `\` / ImplicitAD.implicit_linear (src/matrix_assembly.jl:108-122). */
int mh_lu_factor(mh_context_t ctx, int64_t n, int64_t batch, double* A,
                 int64_t ld, int64_t stride, int64_t* ipiv);
/* Solve with the factors: trans=0 A X = B, trans=1 A^T X = B (the adjoint
 * of the complex-linear solve).  B: n x nrhs per item, in place. */
int mh_lu_solve(mh_context_t ctx, int64_t n, int64_t batch, const double* LU,
                int64_t ld, int64_t stride, const int64_t* ipiv, int trans,
                int64_t nrhs, double* B, int64_t ldb, int64_t bstride);
/* solve(D, S, bc; direct) (src/matrix_assembly.jl:105-133) for a batch:
 * direct: phi = D \ (S bc); indirect: phi = S (K \ bc).  DK is overwritten
 * by its LU factors.  bc, phi: n x nrhs per item. */
int mh_solve(mh_context_t ctx, int64_t n, int64_t batch, int direct,
             double* DK, const double* S, int64_t ld, int64_t stride,
             const double* bc, int64_t nrhs, double* phi, int64_t ldb,
             int64_t bstride);

/* ---- reverse mode (matrix-level cotangents) ----------------------------
 * Cotangent convention of Zygote / ChainRules for complex arrays: zbar =
 * dL/dRe z + i dL/dIm z, so a real input u receives sum Re(conj(zbar) dz/du).
 *
 * Pullback of assemble_matrices(gfs, mesh, k; direct) (src/matrix_assembly.jl:
 * 4-42, what Zygote differentiates in test/matrix_assembly_differentiation.jl:
 * 34-64) for nk wavenumbers: Sbar and DKbar (nk matrices each, ld, stride;
 * either may be NULL = zero) are the cotangents of S and of D (formulation
 * MH_DIRECT) or K (MH_INDIRECT).  Outputs (device, each may be NULL):
 * dvertices nv x 3, dcenters nf x 3, dnormals nf x 3 (column-major), dareas
 * nf, summed over the wavenumbers, and dwavenumbers (nk).  The mesh fields
 * are independent inputs, as in the reference's Mesh: centres enter as the
 * collocation points and as the source-panel centres, vertices only through
 * the near-field Rankine formula.  Radii select branches and get no
 * cotangent.  The pair terms are recomputed, never stored.
 * wavenumbers: HOST array.  Synchronises. */
int mh_assemble_vjp(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask,
                    int formulation, int64_t nk, const double* wavenumbers,
                    const double* Sbar, const double* DKbar, int64_t ld,
                    int64_t stride, double* dvertices, double* dcenters,
                    double* dnormals, double* dareas, double* dwavenumbers);
/* Pullback of solve(D, S, bc; direct) (src/matrix_assembly.jl:105-133; the
 * implicit_linear rrule of ImplicitAD at :120): LU, ipiv are the factors of
 * D (direct) or K (indirect) from mh_lu_factor; phibar is the cotangent of
 * phi (n x nrhs per item, ldb, bstride, like bc).  Outputs (device, each may
 * be NULL): DKbar and Sbar (n x n per item, ld, stride), bcbar (like bc):
 *   direct  (y = S bc, phi = D \ y): ybar = D^-H phibar, Dbar = -ybar phi^H,
 *            Sbar = ybar bc^H, bcbar = S^H ybar;
 *   indirect (sigma = K \ bc, phi = S sigma): Sbar = phibar sigma^H,
 *            bcbar = K^-H S^H phibar, Kbar = -bcbar sigma^H.
 * One transposed solve on the given factors (plus one forward solve that
 * recomputes phi or sigma).  Synchronises. */
int mh_solve_vjp(mh_context_t ctx, int64_t n, int64_t batch, int direct,
                 const double* LU, const int64_t* ipiv, const double* S,
                 int64_t ld, int64_t stride, const double* bc, int64_t nrhs,
                 const double* phibar, int64_t ldb, int64_t bstride,
                 double* DKbar, double* Sbar, double* bcbar);

/* ---- one frequency sweep, end to end -----------------------------------
 * For every omega: k = omega^2/g; assemble; radiation BCs for each dof
 * field (radiation_bc, src/waves.jl:89-102) and Airy diffraction BCs for
 * each heading (AiryBC, src/waves.jl:121-142); one LU and all right-hand
 * sides; pressure and integrate_pressure (src/waves.jl:21-37, :55-59,
 * :72-85).  Outputs (device):
 *   A, B : nomega x ndof x ndof, A[w][j][i] = added mass on influenced dof
 *          i due to radiating dof j, = Re F / omega^2 (waves.jl:36);
 *          B = Im F / omega.
 *   F    : nomega x nbeta x ndof complex128, diffraction force on dof i for
 *          heading b (DiffractionForce, src/waves.jl:40-52).
 * dof_fields: device, ndof fields of nfaces x 3 (column-major) each.
 * omegas, betas: host arrays.  precision: MH_FP64 or MH_MIXED. */
int mh_freq_sweep(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask,
                  int direct, int precision, int64_t nomega,
                  const double* omegas, double g, double rho, int64_t ndof,
                  const double* dof_fields, int64_t nbeta, const double* betas,
                  double* A, double* B, double* F);

/* ---- gradients ----------------------------------------------------------
 * mh_freq_sweep plus the derivatives of A, B and F along nparam parameter
 * directions theta_m (dimensions, positions, frequency), for the Zygote /
 * ForwardDiff uses of the reference (test/outputs_differentiation.jl:59-153,
 * paper/MeshGradients_singlebody.jl, paper/meshGradients_pair.jl).  The
 * linear solve is differentiated by its adjoint (one transposed solve with
 * the LU factors per influenced dof, as ImplicitAD.implicit_linear does for
 * the reference, src/matrix_assembly.jl:120); the influence matrices'
 * derivatives are recomputed pair by pair and contracted on the fly, never
 * stored.  Tangents (device, may be NULL = zero):
 *   dvertices  : nparam x (nvertices x 3)  d vertices / d theta_m
 *   dcenters   : nparam x (nfaces x 3)     d centers  / d theta_m
 *   dnormals   : nparam x (nfaces x 3)
 *   dareas     : nparam x nfaces
 *   ddof_fields: nparam x ndof x (nfaces x 3)
 *   domegas    : HOST, nparam x nomega     d omega_w / d theta_m
 * Radii only select the near/far branch and carry no derivative.
 * Outputs (device): A, B, F as mh_freq_sweep; dA, dB: nomega x nparam x
 * ndof x ndof (same inner layout as A); dF: nomega x nparam x nbeta x ndof
 * complex128.  Synchronises.  Limits: ndof >= 1, ndof + nbeta <= 16. */
int mh_freq_sweep_grad(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask,
                       int direct, int precision, int64_t nomega,
                       const double* omegas, double g, double rho, int64_t ndof,
                       const double* dof_fields, int64_t nbeta,
                       const double* betas, int64_t nparam,
                       const double* dvertices, const double* dcenters,
                       const double* dnormals, const double* dareas,
                       const double* ddof_fields, const double* domegas,
                       double* A, double* B, double* F, double* dA, double* dB,
                       double* dF);

/* ---- Froude-Krylov and excitation forces ---------------------------------
 * FroudeKrylovForce (src/waves.jl:62-66, :150-164): the incident-wave
 * pressure i omega rho phi_Airy at the panel centres (airy_waves_pressure,
 * airy_waves_potential, :112-119, :150-154) integrated by integrate_pressure
 * (:21-25), for every frequency, heading and influenced dof.  The excitation
 * force of assemble_hydrodynamic_coefficients
 * (src/problems_and_results.jl:134-136) is the sweep's F plus FK.
 * omegas, betas: HOST arrays.  FK (device): nomega x nbeta x ndof complex128,
 * the layout of mh_freq_sweep's F.  Synchronises. */
int mh_froude_krylov(mh_context_t ctx, const mh_mesh_t* mesh, int64_t nomega,
                     const double* omegas, double g, double rho, int64_t ndof,
                     const double* dof_fields, int64_t nbeta, const double* betas,
                     double* FK);
/* The same plus dFK (device, nomega x nparam x nbeta x ndof complex128), the
 * derivatives along nparam directions; tangents as mh_freq_sweep_grad
 * (device dcenters, dnormals, dareas, ddof_fields, host domegas; each may be
 * NULL = zero).  The reference obtains them by Zygote through the same
 * functions (test/outputs_differentiation.jl).  Synchronises. */
int mh_froude_krylov_grad(mh_context_t ctx, const mh_mesh_t* mesh, int64_t nomega,
                          const double* omegas, double g, double rho, int64_t ndof,
                          const double* dof_fields, int64_t nbeta, const double* betas,
                          int64_t nparam, const double* dcenters, const double* dnormals,
                          const double* dareas, const double* ddof_fields,
                          const double* domegas, double* FK, double* dFK);

/* ---- uneven multi-body batches -----------------------------------------
 * One body's sweep, the arguments of mh_freq_sweep_grad (nparam = 0: no
 * gradients, tangent and gradient pointers unused).  omegas, betas and
 * domegas are HOST arrays; the mesh and every other array are on the
 * device. */
typedef struct {
    const mh_mesh_t* mesh;
    int gf_mask;
    int direct;
    int64_t nomega;
    const double* omegas;
    int64_t ndof;
    const double* dof_fields;
    int64_t nbeta;
    const double* betas;
    int64_t nparam;
    const double* dvertices;
    const double* dcenters;
    const double* dnormals;
    const double* dareas;
    const double* ddof_fields;
    const double* domegas;
    double *A, *B, *F, *dA, *dB, *dF;
} mh_sweep_item_t;

/* The reference's loop over problems (solve_all_problems, src/solve.jl:17-21;
 * problems_from_data, src/problems_and_results.jl:70-100) for many bodies of
 * uneven panel counts: a work list ordered longest first, taken by nworkers
 * worker contexts (each with its own stream and workspace, created on first
 * use and kept by ctx) so that bodies overlap on the GPU.  nworkers = 0 uses
 * the autotuned configuration (or 2 before tuning).  Outputs per item as
 * mh_freq_sweep_grad.  Synchronises; on failure the message names the item.
 * Limits per item as mh_freq_sweep_grad when nparam > 0, else as mh_freq_sweep. */
int mh_multi_sweep(mh_context_t ctx, int64_t nitems, const mh_sweep_item_t* items,
                   int precision, double g, double rho, int nworkers);
/* Times the given items (a sample of the workload; their outputs are
 * overwritten) under each candidate worker count (1, 2, 3, 4, 6) and LU leaf
 * width (the default for the panel height, or 8 columns), keeps the fastest
 * for the context and returns it (each output may be NULL). */
int mh_multi_autotune(mh_context_t ctx, int64_t nitems, const mh_sweep_item_t* items,
                      int precision, double g, double rho, int* best_workers,
                      int* best_leaf, double* best_ms);

/* ---- captured sweeps ----------------------------------------------------
 * No reference counterpart (the reference re-runs its loop, src/waves.jl:
 * 29-52, on every call): the FP64 forward sweep of mh_freq_sweep (precision
 * MH_FP64) for one body, recorded once as a HIP graph on a private context
 * of ctx's device and replayed with new frequencies and headings, for
 * design loops and serving that run one body (or one mesh shape) many
 * times.  The mesh arrays, dof_fields and the outputs A, B, F are the device
 * arrays given to mh_sweep_graph_create: their contents may change between
 * launches (a body reshaped in place), their addresses and sizes may not,
 * and they must outlive the graph.  Creation runs the sweep once with
 * placeholder frequencies (1 rad/s; the outputs are overwritten by every
 * launch).  Frequencies in one chunk only (matrices under 32 GiB).
 * mh_sweep_graph_launch is synchronous, starts after the work already
 * issued on ctx's stream, and reports mesh-index, pivot and solve errors as
 * mh_freq_sweep does; launches of one graph must not overlap. */
typedef struct mh_sweep_graph* mh_sweep_graph_t;
int mh_sweep_graph_create(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int direct, int64_t nomega,
                          int64_t ndof, const double* dof_fields, int64_t nbeta, double g, double rho, double* A,
                          double* B, double* F, mh_sweep_graph_t* out);
int mh_sweep_graph_launch(mh_context_t ctx, mh_sweep_graph_t graph, const double* omegas, const double* betas);
int mh_sweep_graph_destroy(mh_sweep_graph_t graph);

/* Captured batches: the items of mh_multi_sweep recorded once as one HIP
 * graph and replayed with one hipGraphLaunch.  This build captures one FP64
 * forward item on one worker (precision MH_MIXED, nparam > 0, several items
 * or nworkers > 1 return MH_ERR_NOT_SUPPORTED: such captures crashed the
 * host process and the cause was not found); it
 * replaces the reference's loop over problems (src/solve.jl:17-21) when the
 * same bodies are solved again and again.  The items are assigned longest
 * first to nworkers private worker contexts (0: the context's autotuned
 * worker count, else 2; at most the item count), each body to the worker
 * with the least work so far.  Creation runs every item once (its outputs
 * are valid results for the given frequencies) and then records the same
 * sequence.  A launch takes the items again: their host arrays (omegas,
 * betas, domegas) may change, their sizes, options and device arrays may
 * not (the mesh and tangent contents may); it is synchronous, starts after
 * the work already issued on ctx's stream, and launches of one graph must
 * not overlap.  An item whose device checks fail (a mesh index out of
 * range, a zero pivot) runs again without the graph, which reports its
 * error as mh_multi_sweep does.  mh_batch_graph_info: recorded nodes, items
 * run again so far, workers (each output may be NULL). */
typedef struct mh_batch_graph* mh_batch_graph_t;
int mh_batch_graph_create(mh_context_t ctx, int64_t nitems, const mh_sweep_item_t* items, int precision,
                          double g, double rho, int nworkers, mh_batch_graph_t* out);
int mh_batch_graph_launch(mh_context_t ctx, mh_batch_graph_t graph, int64_t nitems, const mh_sweep_item_t* items);
int mh_batch_graph_info(mh_batch_graph_t graph, int64_t* nodes, int64_t* reruns, int* nworkers);
/* The recorded graph written to `path` as Graphviz text (every node with its
 * parameters), for inspection. */
int mh_batch_graph_dot(mh_batch_graph_t graph, const char* path);
int mh_batch_graph_destroy(mh_batch_graph_t graph);

/* ---- host-array entry points (one call, synchronous) -------------------
 * calculate_radiation_forces(mesh, dof, omega) -> [A, B]
 * (src/waves.jl:29-37): Wu tuple, direct formulation, uniform dof 3-vector.
 * Mesh arrays are host arrays in the reference's layout.  out[0] = A,
 * out[1] = B. */
int mh_calculate_radiation_forces_host(
    int64_t nfaces, int64_t nvertices, const double* vertices,
    const int64_t* faces, const double* centers, const double* normals,
    const double* areas, const double* radii, const double* dof3,
    double omega, double g, double rho, double* out);
/* DiffractionForce(mesh, omega, dof) -> ComplexF64 (src/waves.jl:40-52),
 * heading 0.  out[0] = Re F, out[1] = Im F. */
int mh_diffraction_force_host(
    int64_t nfaces, int64_t nvertices, const double* vertices,
    const int64_t* faces, const double* centers, const double* normals,
    const double* areas, const double* radii, const double* dof3,
    double omega, double g, double rho, double* out);

#ifdef __cplusplus
}
#endif
#endif /* MHIP_H */
