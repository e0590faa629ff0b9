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

// ---------------------------------------------------------------------------
// Froude-Krylov force (FroudeKrylovForce, src/waves.jl:62-66, :150-164): the
// incident-wave pressure p = i omega rho phi_Airy, phi_Airy = -i g / omega
// e^{k z} e^{i k (x cos b + y sin b)} (airy_waves_potential, :112-119), at the
// panel centres, integrated like the scattered pressure (integrate_pressure,
// :21-25).  One workgroup per (omega, heading); FK[(w*nbeta + b)*ndof + d].
// With tangents (grid z = parameter m): the derivative along parameter m of
// the same sum, from the tangents of the centres, normals, areas, dof fields
// and omega (k = omega^2 / g): dFK[((w*np + m)*nbeta + b)*ndof + d].
// ---------------------------------------------------------------------------
template <bool TAN>
__global__ __launch_bounds__(256) void k_froude_krylov(
    int64_t n, const double* __restrict__ C, const double* __restrict__ N,
    const double* __restrict__ areas, const double* __restrict__ omega, double g, double rho,
    int64_t ndof, const double* __restrict__ dofs, int64_t nbeta, const double* __restrict__ beta,
    int64_t np, const double* __restrict__ dC, const double* __restrict__ dN,
    const double* __restrict__ dA, const double* __restrict__ ddofs, const double* __restrict__ dom,
    double2* __restrict__ out) {
    const int64_t w = blockIdx.x, b = blockIdx.y, m = blockIdx.z;
    const double om = omega[w];
    const double k = om * om / g;
    const double cb = cos(beta[b]), sb = sin(beta[b]);
    const double omd = TAN ? dom[w * np + m] : 0.0;  // d omega / d theta_m
    const double kd = 2.0 * om * omd / g;
    double sr[kMaxDof], si[kMaxDof];
    for (int d = 0; d < kMaxDof; ++d) sr[d] = si[d] = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double x = C[i], y = C[i + n], z = C[i + 2 * n];
        const double wbar = x * cb + y * sb;
        const double cih = exp(k * z);
        double sn, cs;
        sincos(k * wbar, &sn, &cs);
        // phi = (-i g / omega) cih (cs + i sn);  p = i omega rho phi
        const double f = g / om * cih;
        const double phr = f * sn, phi_ = -f * cs;
        const double pr = -om * rho * phi_, pim = om * rho * phr;
        const double n0 = N[i], n1 = N[i + n], n2 = N[i + 2 * n];
        const double a = areas[i];
        // tangent of p: p' = p (k' z + k z' + i (k' wbar + k wbar')), plus the
        // explicit omega of i omega rho (-i g / omega), which cancels
        double dpr = 0.0, dpi = 0.0, dn0 = 0.0, dn1 = 0.0, dn2 = 0.0, da = 0.0;
        if (TAN) {
            const double dx = dC ? dC[m * n * 3 + i] : 0.0;
            const double dy = dC ? dC[m * n * 3 + i + n] : 0.0;
            const double dz = dC ? dC[m * n * 3 + i + 2 * n] : 0.0;
            const double re = kd * z + k * dz;
            const double im = kd * wbar + k * (dx * cb + dy * sb);
            dpr = pr * re - pim * im;
            dpi = pr * im + pim * re;
            if (dN) {
                dn0 = dN[m * n * 3 + i];
                dn1 = dN[m * n * 3 + i + n];
                dn2 = dN[m * n * 3 + i + 2 * n];
            }
            da = dA ? dA[m * n + i] : 0.0;
        }
        for (int64_t d = 0; d < ndof; ++d) {
            const double* fd = dofs + d * n * 3;
            const double amp = -(fd[i] * n0 + fd[i + n] * n1 + fd[i + 2 * n] * n2);
            if (!TAN) {
                sr[d] += pr * amp * a;
                si[d] += pim * amp * a;
            } else {
                double damp = -(fd[i] * dn0 + fd[i + n] * dn1 + fd[i + 2 * n] * dn2);
                if (ddofs) {
                    const double* g3 = ddofs + (m * ndof + d) * n * 3;
                    damp -= g3[i] * n0 + g3[i + n] * n1 + g3[i + 2 * n] * n2;
                }
                sr[d] += dpr * amp * a + pr * (damp * a + amp * da);
                si[d] += dpi * amp * a + pim * (damp * a + amp * da);
            }
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
        const int64_t o = TAN ? (((w * np + m) * nbeta + b) * ndof + d) : ((w * nbeta + b) * ndof + d);
        out[o] = make_double2(fr, fi);
    }
}

int launch_froude_krylov(Context* ctx, const mh_mesh_t* m, int64_t nomega, const double* d_omega,
                         double g, double rho, int64_t ndof, const double* dofs, int64_t nbeta,
                         const double* d_beta, double2* FK, int64_t np, const double* dC,
                         const double* dN, const double* dA, const double* ddofs, const double* d_dom,
                         double2* dFK) {
    const int64_t n = m->nfaces;
    if (nomega == 0 || nbeta == 0 || ndof == 0) return MH_OK;
    MH_REQUIRE(ndof <= kMaxDof, "froude_krylov: at most 16 dofs");
    MH_REQUIRE(nomega <= 0x7fffffff && nbeta <= 65535 && np <= 65535,
               "froude_krylov: too many frequencies, headings or parameters");
    {
        ProfScope ps(ctx, "froude_krylov", 0.0, 8.0 * (double)n * (7 + 3 * ndof) * nomega * nbeta);
        hipLaunchKernelGGL(k_froude_krylov<false>, dim3((unsigned)nomega, (unsigned)nbeta, 1), dim3(256), 0,
                           ctx->stream, n, m->centers, m->normals, m->areas, d_omega, g, rho, ndof, dofs,
                           nbeta, d_beta, (int64_t)1, nullptr, nullptr, nullptr, nullptr, nullptr, FK);
        MH_HIP_CHECK(hipGetLastError());
    }
    if (np > 0 && dFK) {
        ProfScope ps(ctx, "froude_krylov_tangent", 0.0, 8.0 * (double)n * (14 + 6 * ndof) * nomega * nbeta * np);
        hipLaunchKernelGGL(k_froude_krylov<true>, dim3((unsigned)nomega, (unsigned)nbeta, (unsigned)np),
                           dim3(256), 0, ctx->stream, n, m->centers, m->normals, m->areas, d_omega, g, rho,
                           ndof, dofs, nbeta, d_beta, np, dC, dN, dA, ddofs, d_dom, dFK);
        MH_HIP_CHECK(hipGetLastError());
    }
    return MH_OK;
}

}  // namespace mh
