/*
 * FP64 CPU oracle for the MarineHydro boundary-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (MarineHydro.jl, Julia) that the HIP library is tested against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (libmhip.so) never calls into it.
 *
 * Conventions (same as the reference's Julia arrays):
 *   - 3-vectors of one panel are passed as double[3];
 *   - mesh arrays are column-major (Julia layout): centers is nf x 3, i.e.
 *     cx[0..nf) then cy[0..nf) then cz[0..nf); vertices is nv x 3;
 *     faces is nf x 4, 0-based vertex indices (reference src/meshes.jl:179);
 *   - complex matrices are column-major complex128 (interleaved re, im),
 *     entry (i, j) at [i + j*n].
 *
 * Green-function selection is a bit mask mirroring the tuple of
 * GreensFunction objects the reference sums over
 * (reference src/green_functions/abstract_greens_function.jl:18-83).
 */
#ifndef MH_ORACLE_H
#define MH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_GF_RANKINE = 1,              /* rankine.jl                    */
    ORC_GF_RANKINE_REFLECTED = 2,    /* rankine_reflected.jl          */
    ORC_GF_RANKINE_REFLECTED_NEG = 4,/* rankine_reflected_negative.jl */
    ORC_GF_WU = 8,                   /* wu.jl                         */
    ORC_GF_DELHOMMEAU = 16           /* exact_Guevel_Delhommeau.jl    */
};

/* One panel: centre, 4 vertices (row v = vertex v), normal, area, radius. */
typedef struct {
    double center[3];
    double vertices[4][3];
    double normal[3];
    double area;
    double radius;
} orc_panel;

/* ---- point Green functions (greens / gradient_greens) ------------------ */
/* out: complex value (2 doubles).  gout: complex 3-vector (6 doubles:
 * re0,im0,re1,im1,re2,im2). */
void orc_greens(int gf_mask, const double x[3], const double xi[3], double k,
                double out[2]);
void orc_gradient_greens(int gf_mask, const double x[3], const double xi[3],
                         double k, int wrt_first, double gout[6]);

/* ---- panel integrals (integral / integral_gradient) --------------------- */
void orc_integral(int gf_mask, const double x[3], const orc_panel* p, double k,
                  double out[2]);
void orc_integral_gradient(int gf_mask, const double x[3], const orc_panel* p,
                           double k, int wrt_first, double gout[6]);

/* Single terms, exposed for golden-vector tests. */
double orc_rankine_integral(const double x[3], const orc_panel* p);
void orc_rankine_integral_gradient(const double x[3], const orc_panel* p,
                                   int wrt_first, double g[3]);
/* Dimensionless wave terms: G(r,z), dG/dr, dG/dz (complex, 2 doubles each). */
void orc_wu_dimless(double h, double v, double G[2], double dGdh[2],
                    double dGdv[2]);
void orc_delhommeau_dimless(double r, double z, double G[2], double dGdr[2],
                            double dGdz[2]);
/* e^zeta * E1(zeta) in long double precision, returned as 2 doubles. */
void orc_exp_e1(double zr, double zi, double out[2]);
/* Wu helper functions, for tests. */
double orc_struve_h0(double x);
double orc_struve_h1(double x);

/* ---- assembly (assemble_matrices_comprehension) ------------------------- */
/* S and DK are nf x nf column-major complex128.  direct=1: DK = D (normal of
 * the source panel, gradient w.r.t. the source point); direct=0: DK = K
 * (normal of the collocation panel, gradient w.r.t. the collocation point).
 * nthreads <= 0 uses the OpenMP default. */
void orc_assemble(int gf_mask, int64_t nf, const double* vertices,
                  int64_t nv, const int64_t* faces, const double* centers,
                  const double* normals, const double* areas,
                  const double* radii, double k, int direct, double* S,
                  double* DK, int nthreads);

/* ---- dense complex solve ------------------------------------------------ */
/* In-place LU with partial pivoting (LAPACK zgetrf semantics, pivot chosen by
 * max |re|+|im|).  ipiv is 0-based.  Returns 0 or the 1-based index of the
 * first zero pivot. */
int orc_zgetrf(int64_t n, double* A, int64_t* ipiv);
/* Solve A X = B (trans=0) or A^T X = B (trans=1) with the factors above.
 * B is n x nrhs column-major complex, overwritten by X. */
void orc_zgetrs(int64_t n, const double* LU, const int64_t* ipiv, int trans,
                int64_t nrhs, double* B);
/* reference src/matrix_assembly.jl:105-133: phi for nrhs right-hand sides.
 * DK is destroyed. */
int orc_solve(int64_t n, double* DK, const double* S, const double* bc,
              int64_t nrhs, int direct, double* phi);

/* ---- one frequency, end to end ----------------------------------------- */
/* Radiation (per dof, rows of dofs are 3-vectors applied to every panel) and
 * diffraction (per heading) for one omega.  Outputs:
 *   AB[2*ndof*ndof]: added mass A[i,j] then damping B[i,j] (radiating dof j,
 *                    influenced dof i), column-major over (i,j);
 *   F[2*ndof*nbeta]: complex diffraction force for influenced dof i and
 *                    heading b at [i + b*ndof].
 * dofs: ndof x 3 row-major (each row one rigid translation direction) or,
 * if dof_fields != NULL, ndof fields of nf x 3 column-major each. */
int orc_freq_solve(int gf_mask, int direct, int64_t nf, const double* vertices,
                   int64_t nv, const int64_t* faces, const double* centers,
                   const double* normals, const double* areas,
                   const double* radii, double omega, double g, double rho,
                   int64_t ndof, const double* dof_fields, int64_t nbeta,
                   const double* betas, double* AB, double* F, int nthreads);

/* ---- Froude-Krylov force (src/waves.jl:62-66, :150-164) ---------------- */
/* FK[2*(d + b*ndof)]: complex force on influenced dof d, heading b.  dofs:
 * ndof fields of nf x 3 column-major each. */
void orc_froude_krylov(int64_t nf, const double* C, const double* N,
                       const double* A, double omega, double g, double rho,
                       int64_t ndof, const double* dofs, int64_t nbeta,
                       const double* betas, double* FK);

#ifdef __cplusplus
}
#endif
#endif
