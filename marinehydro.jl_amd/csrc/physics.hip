// Boundary conditions and force integration for a batch of frequencies.
//
// radiation_bc  : bc = -i omega n . dof            (src/waves.jl:89-102)
// AiryBC        : bc = -(v_airy . n), infinite depth (src/waves.jl:121-142)
// pressure      : p = i rho omega phi               (src/waves.jl:34, :56)
// integrate_pressure: F = sum p (-n . dof) area     (src/waves.jl:21-25, :72-85)
// A = Re F / omega^2, B = Im F / omega              (src/waves.jl:36)
#include "mh_internal.hpp"

namespace mh {

__global__ void k_bcs(int64_t n, const double* __restrict__ C, const double* __restrict__ N,
                      const double* __restrict__ omega, double g, int64_t ndof,
                      const double* __restrict__ dofs, int64_t nbeta,
                      const double* __restrict__ beta, double2* __restrict__ rhs, int64_t ldb,
                      int64_t bstride) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t w = blockIdx.y;
    if (i >= n) return;
    const double om = omega[w];
    const double k = om * om / g;  // waves.jl:30, :123
    const double n0 = N[i], n1 = N[i + n], n2 = N[i + 2 * n];
    double2* b = rhs + w * bstride;
    for (int64_t d = 0; d < ndof; ++d) {
        const double* f = dofs + d * n * 3;
        const double nd = n0 * f[i] + n1 * f[i + n] + n2 * f[i + 2 * n];
        b[i + d * ldb] = make_double2(0.0, -om * nd);  // -1im * omega * (n . dof)
    }
    const double x = C[i], y = C[i + n], z = C[i + 2 * n];
    for (int64_t q = 0; q < nbeta; ++q) {
        const double cb = cos(beta[q]), sb = sin(beta[q]);
        const double wbar = x * cb + y * sb;
        const double cih = exp(k * z);
        const double amp = g * k / om;
        double sn, cs;
        sincos(k * wbar, &sn, &cs);
        const double pr = amp * cs, pi = amp * sn;  // (g k / omega) e^{i k wbar}
        // v = pre * (cb cih, sb cih, -i cih); bc = -(v . n)
        const double vn_re = pr * (cb * cih) * n0 + pr * (sb * cih) * n1 + pi * cih * n2;
        const double vn_im = pi * (cb * cih) * n0 + pi * (sb * cih) * n1 - pr * cih * n2;
        b[i + (ndof + q) * ldb] = make_double2(-vn_re, -vn_im);
    }
}

int launch_bcs(Context* ctx, const mh_mesh_t* m, int64_t nomega, const double* d_omega, double g,
               int64_t ndof, const double* dofs, int64_t nbeta, const double* d_beta,
               double2* rhs, int64_t ldb, int64_t bstride) {
    const int64_t n = m->nfaces;
    if (n == 0 || nomega == 0) return MH_OK;
    MH_REQUIRE(nomega <= 65535, "bcs: too many frequencies for one launch");
    ProfScope ps(ctx, "bcs", 0.0, 16.0 * (double)n * (ndof + nbeta) * nomega);
    hipLaunchKernelGGL(k_bcs, dim3((unsigned)((n + 255) / 256), (unsigned)nomega), dim3(256), 0,
                       ctx->stream, n, m->centers, m->normals, d_omega, g, ndof, dofs, nbeta,
                       d_beta, rhs, ldb, bstride);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

// One workgroup per (omega, rhs column): F_d = sum_i i rho omega phi_i
// (-n_i . dof_d,i) area_i for every influenced dof d.
constexpr int kMaxDof = 16;

__global__ __launch_bounds__(256) void k_forces(int64_t n, const double* __restrict__ N,
                                                const double* __restrict__ areas,
                                                const double* __restrict__ omega, double rho,
                                                int64_t ndof, const double* __restrict__ dofs,
                                                int64_t nbeta, const double2* __restrict__ phi,
                                                int64_t ldb, int64_t bstride,
                                                double* __restrict__ A, double* __restrict__ B,
                                                double2* __restrict__ F) {
    const int64_t w = blockIdx.x, r = blockIdx.y;
    const double om = omega[w];
    const double2* ph = phi + w * bstride + r * ldb;
    double sr[kMaxDof], si[kMaxDof];
    for (int d = 0; d < kMaxDof; ++d) sr[d] = si[d] = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double2 p = ph[i];
        // pressure = 1im * rho * omega * phi
        const double pr = -rho * om * p.y, pim = rho * om * p.x;
        const double n0 = N[i], n1 = N[i + n], n2 = N[i + 2 * n];
        const double a = areas[i];
        for (int64_t d = 0; d < ndof; ++d) {
            const double* f = dofs + d * n * 3;
            const double amp = -(f[i] * n0 + f[i + n] * n1 + f[i + 2 * n] * n2);
            sr[d] += pr * amp * a;
            si[d] += pim * amp * a;
        }
    }
    __shared__ double red[2][kMaxDof][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t d = 0; d < ndof; ++d) {
        double x = sr[d], y = si[d];
        for (int off = 32; off > 0; off >>= 1) {
            x += __shfl_xor(x, off);
            y += __shfl_xor(y, off);
        }
        if (lane == 0) {
            red[0][d][wave] = x;
            red[1][d][wave] = y;
        }
    }
    __syncthreads();
    if (threadIdx.x < ndof) {
        const int d = threadIdx.x;
        const double fr = red[0][d][0] + red[0][d][1] + red[0][d][2] + red[0][d][3];
        const double fi = red[1][d][0] + red[1][d][1] + red[1][d][2] + red[1][d][3];
        if (r < ndof) {
            const int64_t o = (w * ndof + r) * ndof + d;
            A[o] = fr / (om * om);
            B[o] = fi / om;
        } else {
            F[(w * nbeta + (r - ndof)) * ndof + d] = make_double2(fr, fi);
        }
    }
}

int launch_forces(Context* ctx, const mh_mesh_t* m, int64_t nomega, const double* d_omega,
                  double rho, int64_t ndof, const double* dofs, int64_t nbeta, const double2* phi,
                  int64_t ldb, int64_t bstride, double* A, double* B, double2* F) {
    const int64_t n = m->nfaces;
    const int64_t nr = ndof + nbeta;
    if (nomega == 0 || nr == 0) return MH_OK;
    MH_REQUIRE(ndof <= kMaxDof, "forces: at most 16 dofs");
    MH_REQUIRE(nomega <= 0x7fffffff && nr <= 65535, "forces: too many frequencies or columns");
    ProfScope ps(ctx, "forces", 0.0, 16.0 * (double)n * nr * nomega);
    hipLaunchKernelGGL(k_forces, dim3((unsigned)nomega, (unsigned)nr), dim3(256), 0, ctx->stream, n,
                       m->normals, m->areas, d_omega, rho, ndof, dofs, nbeta, phi, ldb, bstride, A,
                       B, F);
    MH_HIP_CHECK(hipGetLastError());
    return MH_OK;
}

}  // namespace mh
