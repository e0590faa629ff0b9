// extern "C" entry points of libmhip.so (declared in include/mhip.h).
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include <algorithm>

#include "mh_internal.hpp"

namespace mh {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

bool ensure_workspace(Context* ctx, size_t bytes) {
    if (bytes <= ctx->ws_bytes) return true;
    if (ctx->ws) {
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return false;
        (void)hipFree(ctx->ws);
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
    }
    if (hipMalloc(&ctx->ws, bytes) != hipSuccess) {
        set_error("workspace: hipMalloc of " + std::to_string(bytes) + " bytes failed");
        ctx->ws = nullptr;
        return false;
    }
    ctx->ws_bytes = bytes;
    return true;
}

// Bump allocator over the context workspace (256-byte aligned slices).
struct Bump {
    size_t off = 0;
    size_t take(size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    }
};

static int check_ctx(mh_context_t c) {
    if (!c) {
        set_error("null context");
        return MH_ERR_ARGUMENT;
    }
    hipError_t e = hipSetDevice(reinterpret_cast<Context*>(c)->device);
    if (e != hipSuccess) {
        set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
        return MH_ERR_HIP;
    }
    return MH_OK;
}

static int check_mesh(const mh_mesh_t* m) {
    MH_REQUIRE(m != nullptr, "mesh: null");
    MH_REQUIRE(m->nfaces >= 0 && m->nvertices >= 0, "mesh: negative size");
    if (m->nfaces > 0)
        MH_REQUIRE(m->vertices && m->faces && m->centers && m->normals && m->areas && m->radii,
                   "mesh: null array");
    return MH_OK;
}

// A sum of Green-function terms with at most one wave term.
static int check_gf(int gf, const char* who) {
    if (gf <= 0 || gf >= 64) {
        set_error(std::string(who) + ": bad Green-function mask");
        return MH_ERR_ARGUMENT;
    }
    const int w = gf & (MH_GF_WU | MH_GF_DELHOMMEAU | MH_GF_DELHOMMEAU_TABLE);
    if (w & (w - 1)) {
        set_error(std::string(who) + ": choose one wave term");
        return MH_ERR_ARGUMENT;
    }
    return MH_OK;
}

int ensure_aux(Context* c) {
    if (c->aux) return MH_OK;
    MH_HIP_CHECK(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    MH_HIP_CHECK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    MH_HIP_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    MH_HIP_CHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    MH_HIP_CHECK(hipEventCreateWithFlags(&c->ev_side_fork, hipEventDisableTiming));
    MH_HIP_CHECK(hipEventCreateWithFlags(&c->ev_side_join, hipEventDisableTiming));
    return MH_OK;
}
int ensure_events(Context* c, size_t count) {
    while (c->ev_pool.size() < count) {
        hipEvent_t e;
        MH_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_pool.push_back(e);
    }
    return MH_OK;
}
}  // namespace mh

using namespace mh;

#define CTX(c) reinterpret_cast<mh::Context*>(c)
#define TRY(expr)                  \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != MH_OK) return rc_; \
    } while (0)

extern "C" {

int mh_version(void) { return MHIP_VERSION; }
const char* mh_last_error(void) { return mh::last_error(); }

int mh_context_create(int device, void* stream, mh_context_t* out) {
    if (!out) {
        set_error("mh_context_create: null output pointer");
        return MH_ERR_ARGUMENT;
    }
    *out = nullptr;
    Context* c = nullptr;
    TRY(context_new(device, reinterpret_cast<hipStream_t>(stream), false, &c));
    *out = reinterpret_cast<mh_context_t>(c);
    return MH_OK;
}

int mh_context_destroy(mh_context_t ctx) {
    if (!ctx) return MH_OK;
    context_free(CTX(ctx));
    return MH_OK;
}
}  // extern "C"

namespace mh {
int context_new(int device, hipStream_t stream, bool own_stream, Context** out) {
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device");
        return MH_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        set_error("device index out of range");
        return MH_ERR_ARGUMENT;
    }
    MH_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    MH_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("libmhip is built for gfx950; device is ") + prop.gcnArchName);
        return MH_ERR_NO_DEVICE;
    }
    Context* c = new Context();
    c->device = device;
    c->stream = stream;
    if (own_stream) {
        // default priority (no priority argument anywhere in the library)
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            set_error("context: hipStreamCreateWithFlags failed");
            return MH_ERR_HIP;
        }
        c->owns_stream = true;
    }
    if (hipMalloc(&c->d_flag, 256) != hipSuccess) {
        if (c->owns_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        set_error("mh_context_create: hipMalloc failed");
        return MH_ERR_HIP;
    }
    *out = c;
    return MH_OK;
}

void context_free(Context* c) {
    for (Context* w : c->workers) context_free(w);
    c->workers.clear();
    (void)hipSetDevice(c->device);
    // every stream of the context drains before any buffer it may still be
    // reading or writing is released
    (void)hipStreamSynchronize(c->stream);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (ProfRec& r : c->prof) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    c->prof.clear();
    if (c->ws) (void)hipFree(c->ws);
    if (c->d_flag) (void)hipFree(c->d_flag);
    for (auto& t : c->tri) {
        if (t.sync) (void)hipFree(t.sync);
        if (t.ws) (void)hipFree(t.ws);
    }
    if (c->tri_err) (void)hipFree(c->tri_err);
    if (c->rk) (void)hipFree(c->rk);
    if (c->himg) (void)hipFree(c->himg);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->ev_side_fork) (void)hipEventDestroy(c->ev_side_fork);
    if (c->ev_side_join) (void)hipEventDestroy(c->ev_side_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->owns_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}
int check_info(Context* c, const int* info, int64_t batch) {
    std::vector<int> h(batch);
    MH_HIP_CHECK(hipMemcpyAsync(h.data(), info, sizeof(int) * batch, hipMemcpyDeviceToHost, c->stream));
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t b = 0; b < batch; ++b) {
        if (h[b] & 1) {
            set_error("solve: exact zero pivot in batch item " + std::to_string(b));
            return MH_ERR_SINGULAR;
        }
        if (h[b] & 2) {
            set_error("mixed precision: non-finite entries in the FP32 factorisation of batch item " +
                      std::to_string(b));
            return MH_ERR_NOT_CONVERGED;
        }
    }
    return MH_OK;
}

}  // namespace mh

extern "C" {

int mh_synchronize(mh_context_t ctx) {
    TRY(check_ctx(ctx));
    MH_HIP_CHECK(hipStreamSynchronize(CTX(ctx)->stream));
    return MH_OK;
}

int mh_profile_enable(mh_context_t ctx, int on) {
    TRY(check_ctx(ctx));
    CTX(ctx)->profiling = on != 0;
    return MH_OK;
}

int mh_profile_report(mh_context_t ctx, char* buf, int64_t cap) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(buf && cap > 0, "profile_report: null buffer");
    Context* c = CTX(ctx);
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    struct Agg {
        std::string name;
        int64_t n = 0;
        double ms = 0, flops = 0, bytes = 0;
    };
    std::vector<Agg> agg;
    for (ProfRec& r : c->prof) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, r.start, r.stop);
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
        Agg* a = nullptr;
        for (Agg& x : agg)
            if (x.name == r.name) a = &x;
        if (!a) {
            agg.push_back(Agg());
            a = &agg.back();
            a->name = r.name;
        }
        a->n += 1;
        a->ms += ms;
        a->flops += r.flops;
        a->bytes += r.bytes;
    }
    c->prof.clear();
    std::string js = "{";
    for (size_t q = 0; q < agg.size(); ++q) {
        char tmp[256];
        snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"launches\": %lld, \"ms\": %.6f, \"flops\": %.6e, \"bytes\": %.6e}",
                 q ? ", " : "", agg[q].name.c_str(), (long long)agg[q].n, agg[q].ms, agg[q].flops,
                 agg[q].bytes);
        js += tmp;
    }
    js += "}";
    if ((int64_t)js.size() + 1 > cap) {
        set_error("profile_report: buffer too small");
        return MH_ERR_ARGUMENT;
    }
    memcpy(buf, js.c_str(), js.size() + 1);
    return MH_OK;
}

int mh_gf_eval(mh_context_t ctx, int gf_mask, int op, int wrt_first, int64_t npairs,
               const double* points, const mh_panel_t* panels, double k, double* out) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(npairs >= 0, "gf_eval: negative npairs");
    MH_REQUIRE(op >= 0 && op <= 3, "gf_eval: unknown op");
    TRY(check_gf(gf_mask, "gf_eval"));
    if (npairs == 0) return MH_OK;
    MH_REQUIRE(points && panels && out, "gf_eval: null array");
    return launch_gf_eval(CTX(ctx), gf_mask, op, wrt_first, npairs, points, panels, k,
                          reinterpret_cast<double2*>(out));
}

int mh_wave_term_eval(mh_context_t ctx, int term, int64_t npoints, const double* r,
                      const double* z, double* out) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(npoints >= 0, "wave_term: negative size");
    if (npoints == 0) return MH_OK;
    MH_REQUIRE(r && z && out, "wave_term: null array");
    return launch_wave_term(CTX(ctx), term, npoints, r, z, reinterpret_cast<double2*>(out));
}

static int assemble_impl(Context* c, const mh_mesh_t* mesh, int gf, int form, int64_t nk,
                         const double* ks_host, double* S, double* D, double* K, int64_t ld,
                         int64_t stride, Bump& bump, size_t base) {
    const int64_t n = mesh->nfaces;
    const size_t o_p = bump.take(sizeof(PanelRec) * (size_t)n);
    const size_t o_k = bump.take(sizeof(double) * (size_t)nk);
    if (!ensure_workspace(c, base + bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws) + base;
    PanelRec* P = reinterpret_cast<PanelRec*>(ws + o_p);
    double* dk = reinterpret_cast<double*>(ws + o_k);
    MH_HIP_CHECK(hipMemcpyAsync(dk, ks_host, sizeof(double) * nk, hipMemcpyHostToDevice, c->stream));
    TRY(launch_pack_panels(c, mesh, P));
    return launch_assemble(c, mesh, P, gf, form, nk, dk, reinterpret_cast<double2*>(S),
                           reinterpret_cast<double2*>(D), reinterpret_cast<double2*>(K), ld, stride);
}

int mh_assemble(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int formulation, int64_t nk,
                const double* wavenumbers, double* S, double* D, double* K, int64_t ld,
                int64_t stride) {
    TRY(check_ctx(ctx));
    TRY(check_mesh(mesh));
    TRY(check_gf(gf_mask, "assemble"));
    MH_REQUIRE(formulation >= 1 && formulation <= 3, "assemble: formulation must be 1, 2 or 3");
    MH_REQUIRE(nk >= 0, "assemble: negative nk");
    const int64_t n = mesh->nfaces;
    if (n == 0 || nk == 0) return MH_OK;
    MH_REQUIRE(wavenumbers && S, "assemble: null array");
    MH_REQUIRE(!(formulation & 1) || D, "assemble: D required for the direct formulation");
    MH_REQUIRE(!(formulation & 2) || K, "assemble: K required for the indirect formulation");
    MH_REQUIRE(ld >= n && (nk == 1 || stride >= ld * n), "assemble: ld/stride too small");
    Bump bump;
    return assemble_impl(CTX(ctx), mesh, gf_mask, formulation, nk, wavenumbers, S, D, K, ld,
                         stride, bump, 0);
}

int mh_lu_factor(mh_context_t ctx, int64_t n, int64_t batch, double* A, int64_t ld, int64_t stride,
                 int64_t* ipiv) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(n >= 0 && batch >= 0, "lu_factor: negative size");
    if (n == 0 || batch == 0) return MH_OK;
    MH_REQUIRE(A && ipiv, "lu_factor: null array");
    MH_REQUIRE(ld >= n && (batch == 1 || stride >= ld * n), "lu_factor: ld/stride too small");
    Context* c = CTX(ctx);
    Bump bump;
    const size_t o_info = bump.take(sizeof(int) * (size_t)batch);
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    int* info = reinterpret_cast<int*>(static_cast<char*>(c->ws) + o_info);
    MH_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int) * batch, c->stream));
    TRY(lu_factor(c, n, n, batch, reinterpret_cast<double2*>(A), ld, stride, ipiv, info));
    std::vector<int> h(batch);
    MH_HIP_CHECK(hipMemcpyAsync(h.data(), info, sizeof(int) * batch, hipMemcpyDeviceToHost, c->stream));
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int64_t b = 0; b < batch; ++b)
        if (h[b]) {
            set_error("lu_factor: exact zero pivot in batch item " + std::to_string(b));
            return MH_ERR_SINGULAR;
        }
    return MH_OK;
}

int mh_lu_solve(mh_context_t ctx, int64_t n, int64_t batch, const double* LU, int64_t ld,
                int64_t stride, const int64_t* ipiv, int trans, int64_t nrhs, double* B, int64_t ldb,
                int64_t bstride) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(n >= 0 && batch >= 0 && nrhs >= 0, "lu_solve: negative size");
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(LU && ipiv && B, "lu_solve: null array");
    MH_REQUIRE(ld >= n && ldb >= n, "lu_solve: leading dimension too small");
    Context* c = CTX(ctx);
    Bump bump;
    const size_t o_perm = bump.take(sizeof(int) * (size_t)(n * batch));
    const size_t o_inv = bump.take(sizeof(double2) * (size_t)tri_inv_elems(n, batch));
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    int* perm = reinterpret_cast<int*>(static_cast<char*>(c->ws) + o_perm);
    double2* inv = reinterpret_cast<double2*>(static_cast<char*>(c->ws) + o_inv);
    const double2* LU2 = reinterpret_cast<const double2*>(LU);
    TRY(tri_prepare(c, n, batch, LU2, ld, stride, ipiv, perm, inv));
    TRY(tri_solve(c, n, batch, LU2, ld, stride, perm, inv, trans ? 1 : 0, 3, nrhs,
                  reinterpret_cast<double2*>(B), ldb, bstride));
    return tri_check(c);
}

// shared by mh_solve and mh_freq_sweep: LU + solve for `batch` systems whose
// right-hand sides are already in rhs (overwritten by phi)
static int solve_impl(Context* c, int64_t n, int64_t batch, int direct, double2* DK,
                      const double2* S, int64_t ld, int64_t stride, double2* rhs, int64_t nrhs,
                      int64_t ldb, int64_t bstride, double2* tmp, int64_t* ipiv, int* perm, double2* inv,
                      int* info) {
    MH_HIP_CHECK(hipMemsetAsync(info, 0, sizeof(int) * batch, c->stream));
    TRY(lu_factor(c, n, n, batch, DK, ld, stride, ipiv, info));
    TRY(tri_prepare(c, n, batch, DK, ld, stride, ipiv, perm, inv));
    if (direct) {  // phi = D \ (S bc)
        TRY(zgemm_nn_batched(c, n, nrhs, batch, S, ld, stride, rhs, ldb, bstride, tmp, ldb, bstride));
        TRY(tri_solve(c, n, batch, DK, ld, stride, perm, inv, 0, 3, nrhs, tmp, ldb, bstride));
        MH_HIP_CHECK(hipMemcpyAsync(rhs, tmp, sizeof(double2) * (size_t)(bstride * batch),
                                    hipMemcpyDeviceToDevice, c->stream));
    } else {  // sigma = K \ bc; phi = S sigma
        TRY(tri_solve(c, n, batch, DK, ld, stride, perm, inv, 0, 3, nrhs, rhs, ldb, bstride));
        TRY(zgemm_nn_batched(c, n, nrhs, batch, S, ld, stride, rhs, ldb, bstride, tmp, ldb, bstride));
        MH_HIP_CHECK(hipMemcpyAsync(rhs, tmp, sizeof(double2) * (size_t)(bstride * batch),
                                    hipMemcpyDeviceToDevice, c->stream));
    }
    return MH_OK;
}


int mh_solve(mh_context_t ctx, int64_t n, int64_t batch, int direct, double* DK, const double* S,
             int64_t ld, int64_t stride, const double* bc, int64_t nrhs, double* phi, int64_t ldb,
             int64_t bstride) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(n >= 0 && batch >= 0 && nrhs >= 0, "solve: negative size");
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(DK && S && bc && phi, "solve: null array");
    MH_REQUIRE(ld >= n && ldb >= n, "solve: leading dimension too small");
    MH_REQUIRE(batch == 1 || (stride >= ld * n && bstride >= ldb * nrhs), "solve: stride too small");
    Context* c = CTX(ctx);
    const int64_t bs = (batch == 1) ? ldb * nrhs : bstride;
    Bump bump;
    const size_t o_tmp = bump.take(sizeof(double2) * (size_t)(bs * batch));
    const size_t o_piv = bump.take(sizeof(int64_t) * (size_t)(n * batch));
    const size_t o_info = bump.take(sizeof(int) * (size_t)batch);
    const size_t o_perm = bump.take(sizeof(int) * (size_t)(n * batch));
    const size_t o_inv = bump.take(sizeof(double2) * (size_t)tri_inv_elems(n, batch));
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws);
    if (phi != bc)
        MH_HIP_CHECK(hipMemcpyAsync(phi, bc, sizeof(double2) * (size_t)(bs * batch),
                                    hipMemcpyDeviceToDevice, c->stream));
    TRY(solve_impl(c, n, batch, direct, reinterpret_cast<double2*>(DK),
                   reinterpret_cast<const double2*>(S), ld, stride, reinterpret_cast<double2*>(phi),
                   nrhs, ldb, bs, reinterpret_cast<double2*>(ws + o_tmp),
                   reinterpret_cast<int64_t*>(ws + o_piv), reinterpret_cast<int*>(ws + o_perm),
                   reinterpret_cast<double2*>(ws + o_inv), reinterpret_cast<int*>(ws + o_info)));
    TRY(tri_check(c));
    return check_info(c, reinterpret_cast<int*>(ws + o_info), batch);
}

// ---- second stream: independent work (forward and adjoint solves) runs
// on ctx->aux, ordered against ctx->stream by events.  With the kernel timer
// on, everything stays on ctx->stream so each timed kernel runs alone. ------
// aux waits for everything issued on stream so far
static int fork_aux(Context* c) {
    if (c->profiling) return MH_OK;
    TRY(ensure_aux(c));
    MH_HIP_CHECK(hipEventRecord(c->ev_fork, c->stream));
    MH_HIP_CHECK(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    return MH_OK;
}
// stream waits for everything issued on aux so far
static int join_aux(Context* c) {
    if (c->profiling) return MH_OK;
    MH_HIP_CHECK(hipEventRecord(c->ev_join, c->aux));
    MH_HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return MH_OK;
}
// the side stream starts after the work issued on stream so far
static int fork_side(Context* c) {
    if (c->profiling) return MH_OK;
    TRY(ensure_aux(c));
    MH_HIP_CHECK(hipEventRecord(c->ev_side_fork, c->stream));
    MH_HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_side_fork, 0));
    return MH_OK;
}
// stream waits for everything issued on side so far
static int join_side(Context* c) {
    if (c->profiling) return MH_OK;
    MH_HIP_CHECK(hipEventRecord(c->ev_side_join, c->side));
    MH_HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_side_join, 0));
    return MH_OK;
}
// issue launches on side within a scope (no triangular solves there)
struct OnSide {
    Context* c;
    hipStream_t saved;
    explicit OnSide(Context* ctx) : c(ctx), saved(ctx->stream) {
        if (!c->profiling) c->stream = c->side;
    }
    ~OnSide() { c->stream = saved; }
};
// issue launches on aux (triangular-solve state slot 1) within a scope
struct OnAux {
    Context* c;
    hipStream_t saved;
    int saved_slot;
    OnAux(Context* ctx, bool on) : c(ctx), saved(ctx->stream), saved_slot(ctx->tri_slot) {
        if (on && !c->profiling) {
            c->stream = c->aux;
            c->tri_slot = 1;
        }
    }
    ~OnAux() {
        c->stream = saved;
        c->tri_slot = saved_slot;
    }
};
static int ensure_pinned(Context* c, size_t count) {
    if (count <= c->h_pin_count) return MH_OK;
    if (c->h_pin) {
        MH_HIP_CHECK(hipStreamSynchronize(c->stream));
        if (c->aux) MH_HIP_CHECK(hipStreamSynchronize(c->aux));
        MH_HIP_CHECK(hipHostFree(c->h_pin));
        c->h_pin = nullptr;
        c->h_pin_count = 0;
    }
    MH_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_pin), sizeof(double) * count));
    c->h_pin_count = count;
    return MH_OK;
}

// One system op(A) X = Bm (op = A or A^T) for `batch` items, refined with the
// FP32 LU factors of A against the FP64 A.  Bm, X, R share one layout (ldb,
// bstride); R32 is n x nrhs per item; nrm holds nrhs x batch residual norms.
struct IrSystem {
    int trans;
    const double2* Bm;
    int64_t nrhs, ldb, bstride;
    double2* X;
    double2* R;
    float2* R32;
    double* nrm;
};

// Iterative refinement of up to two systems with the same FP32 factors, the
// second on the aux stream so both proceed together: X <- X + LU32^{-1}
// (Bm - op(A) X) until every column's relative residual max|r| / max|b| is
// below kIrTol (at most 12 sweeps).  Each sweep gains about the FP32 LU's
// accuracy (~1e-6 relative for these boundary-element systems), so two
// sweeps typically reach 1e-11 and the results agree with the FP64 path to
// ~1e-10, well inside the 1e-6 the mixed-precision mode is specified for.
constexpr double kIrTol = 1e-10;
// LUt32 (optional): the transposed factors (built during the LU); transposed
// systems solve from it with coalesced reads.
// defer_word (not under capture): a fixed c->ir_deferred_sweeps sweeps (at
// most c->ir_max_sweeps); the
// last sweep's residuals and their convergence verdict (bits set in
// *defer_word, which the caller cleared) are formed on the aux stream, after
// the corrections, and the call returns without waiting for them: the
// caller's stream goes on with what only reads the solutions, and the
// caller reads the verdict later.
static int ir_solve(Context* c, int64_t n, int64_t batch, const double2* A64, int64_t ld, int64_t stride,
                    const float2* LU32, int64_t stride32, const int* perm, const float2* inv32, IrSystem* sys,
                    int nsys, float2* LUt32, double2* Pa, bool* nonfinite, int* defer_word = nullptr) {
    *nonfinite = false;
    // a forward and a transposed system form their residuals together, in
    // one pass over A (Pa: workspace of zgemv_pair)
    const bool pair = nsys == 2 && Pa && sys[0].trans == 0 && sys[1].trans == 1 && sys[0].nrhs <= 4 &&
                      sys[1].nrhs <= 4;
    // captured: a fixed number of sweeps, the last one's residual norms
    // turned into a device verdict (cap->check[2]) instead of a host check
    const bool cap = c->capturing;
    MH_REQUIRE(!cap || (c->cap && c->ir_graph_sweeps >= 2), "refinement: capture without check words");
    size_t hcount = 0;
    for (int k = 0; k < nsys; ++k) hcount += (size_t)(batch * sys[k].nrhs);
    if (!cap) TRY(ensure_pinned(c, hcount));
    if (nsys > 1) TRY(fork_aux(c));
    bool done[2] = {false, false};
    for (int k = 0; k < nsys; ++k) {
        OnAux on(c, k == 1);
        const IrSystem& q = sys[k];
        TRY(fill_zero(c, q.X, sizeof(double2) * (size_t)(q.bstride * batch)));
        TRY(copy_d2(c, q.R, 0, q.Bm, 0, q.bstride * batch, 1));
    }
    const bool defer = !cap && defer_word != nullptr;
    MH_REQUIRE(!defer || c->ir_deferred_sweeps >= 1, "refinement: deferred check without sweeps");
    const int max_sweeps =
        cap ? c->ir_graph_sweeps : (defer ? std::min(c->ir_deferred_sweeps, c->ir_max_sweeps) : c->ir_max_sweeps);
    bool deferred = false;
    for (int it = 0; it < max_sweeps; ++it) {
        if (defer && it == max_sweeps - 1) {  // the last correction; its residuals below
            deferred = true;
            for (int k = 0; k < nsys; ++k) {
                OnAux on(c, k == 1);
                const IrSystem& q = sys[k];
                const int64_t s32 = n * q.nrhs;
                TRY(convert_to_f32(c, n, q.nrhs, batch, q.R, q.ldb, q.bstride, q.R32, n, s32));
                TRY(tri_solve_f32(c, n, batch, LU32, n, stride32, perm, inv32, q.trans, 3, q.nrhs, q.R32, n, s32,
                                  q.trans ? LUt32 : nullptr));
                TRY(add_from_f32(c, n, q.nrhs, batch, q.R32, n, s32, q.X, q.ldb, q.bstride));
            }
            break;
        }
        const bool both = pair && !done[0] && !done[1];
        // residual norms: every sweep after the first (a captured sweep only
        // at its last one)
        const bool norms = cap ? it == max_sweeps - 1 : it > 0;
        size_t off = 0;
        for (int k = 0; k < nsys; ++k) {
            const IrSystem& q = sys[k];
            const size_t cnt = (size_t)(batch * q.nrhs);
            if (!done[k]) {
                OnAux on(c, k == 1);
                const int64_t s32 = n * q.nrhs;
                TRY(convert_to_f32(c, n, q.nrhs, batch, q.R, q.ldb, q.bstride, q.R32, n, s32));
                TRY(tri_solve_f32(c, n, batch, LU32, n, stride32, perm, inv32, q.trans, 3, q.nrhs, q.R32, n, s32,
                                  q.trans ? LUt32 : nullptr));
                TRY(add_from_f32(c, n, q.nrhs, batch, q.R32, n, s32, q.X, q.ldb, q.bstride));
                if (both) {
                    off += cnt;
                    continue;
                }
                if (q.trans)
                    TRY(zgemm_tn_batched(c, n, q.nrhs, batch, A64, ld, stride, q.X, q.ldb, q.bstride, q.R, q.ldb,
                                         q.bstride, q.Bm));
                else
                    TRY(zgemm_nn_batched(c, n, q.nrhs, batch, A64, ld, stride, q.X, q.ldb, q.bstride, q.R, q.ldb,
                                         q.bstride, q.Bm));
                if (norms) {
                    TRY(rel_residual(c, n, q.nrhs, batch, q.R, q.Bm, q.ldb, q.bstride, q.nrm));
                    if (!cap)
                        MH_HIP_CHECK(hipMemcpyAsync(c->h_pin + off, q.nrm, sizeof(double) * cnt,
                                                    hipMemcpyDeviceToHost, c->stream));
                }
            }
            off += cnt;
        }
        if (both) {  // both residuals in one pass over A, on stream
            TRY(join_aux(c));
            const IrSystem &f = sys[0], &a = sys[1];
            TRY(zgemv_pair(c, n, batch, A64, ld, stride, (int)f.nrhs, f.X, f.Bm, f.R, f.ldb, f.bstride, (int)a.nrhs,
                           a.X, a.Bm, a.R, a.ldb, a.bstride, Pa));
            if (norms) {
                size_t o2 = 0;
                for (int k = 0; k < nsys; ++k) {
                    const IrSystem& q = sys[k];
                    const size_t cnt = (size_t)(batch * q.nrhs);
                    TRY(rel_residual(c, n, q.nrhs, batch, q.R, q.Bm, q.ldb, q.bstride, q.nrm));
                    if (!cap)
                        MH_HIP_CHECK(hipMemcpyAsync(c->h_pin + o2, q.nrm, sizeof(double) * cnt,
                                                    hipMemcpyDeviceToHost, c->stream));
                    o2 += cnt;
                }
            }
            TRY(fork_aux(c));  // the next solve on aux starts after the residuals
        }
        if (cap) {
            if (it < max_sweeps - 1) continue;
            if (nsys > 1) TRY(join_aux(c));
            for (int k = 0; k < nsys; ++k)
                TRY(ir_verdict(c, sys[k].nrm, batch * sys[k].nrhs, kIrTol, c->cap->check + 2));
            return MH_OK;
        }
        // one FP32 solve leaves a residual far above the tolerance: the first
        // correction is always made, without a host round trip for the check
        // (unless it is the only sweep allowed)
        if (it == 0 && max_sweeps > 1) continue;
        if (it == 0) {  // residuals of the single correction, for the check below
            if (nsys > 1) TRY(join_aux(c));
            size_t o2 = 0;
            for (int k = 0; k < nsys; ++k) {
                const IrSystem& q = sys[k];
                const size_t cnt = (size_t)(batch * q.nrhs);
                TRY(rel_residual(c, n, q.nrhs, batch, q.R, q.Bm, q.ldb, q.bstride, q.nrm));
                MH_HIP_CHECK(hipMemcpyAsync(c->h_pin + o2, q.nrm, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                                            c->stream));
                o2 += cnt;
            }
        }
        MH_HIP_CHECK(hipStreamSynchronize(c->stream));
        if (nsys > 1 && !c->profiling) MH_HIP_CHECK(hipStreamSynchronize(c->aux));
        off = 0;
        bool all = true;
        for (int k = 0; k < nsys; ++k) {
            const size_t cnt = (size_t)(batch * sys[k].nrhs);
            if (!done[k]) {
                // every column below the tolerance; a NaN compares false and
                // keeps the system unconverged
                bool ok = true;
                for (size_t e = 0; e < cnt; ++e) {
                    ok = ok && (c->h_pin[off + e] < kIrTol);
                    if (!(c->h_pin[off + e] <= 1e300)) *nonfinite = true;
                }
                done[k] = ok;
            }
            all = all && done[k];
            off += cnt;
        }
        if (all) return MH_OK;
        if (*nonfinite) break;  // NaN or Inf: more sweeps cannot help
    }
    if (deferred) {
        // aux starts after every correction (those of system 0 on stream)
        if (nsys > 1) TRY(join_aux(c));
        TRY(fork_aux(c));
        OnAux on(c, true);
        if (pair) {
            const IrSystem &f = sys[0], &a = sys[1];
            TRY(zgemv_pair(c, n, batch, A64, ld, stride, (int)f.nrhs, f.X, f.Bm, f.R, f.ldb, f.bstride, (int)a.nrhs,
                           a.X, a.Bm, a.R, a.ldb, a.bstride, Pa));
        } else {
            for (int k = 0; k < nsys; ++k) {
                const IrSystem& q = sys[k];
                if (q.trans)
                    TRY(zgemm_tn_batched(c, n, q.nrhs, batch, A64, ld, stride, q.X, q.ldb, q.bstride, q.R, q.ldb,
                                         q.bstride, q.Bm));
                else
                    TRY(zgemm_nn_batched(c, n, q.nrhs, batch, A64, ld, stride, q.X, q.ldb, q.bstride, q.R, q.ldb,
                                         q.bstride, q.Bm));
            }
        }
        for (int k = 0; k < nsys; ++k) {
            const IrSystem& q = sys[k];
            TRY(rel_residual(c, n, q.nrhs, batch, q.R, q.Bm, q.ldb, q.bstride, q.nrm));
            TRY(ir_verdict(c, q.nrm, batch * q.nrhs, kIrTol, defer_word));
        }
        return MH_OK;
    }
    set_error(*nonfinite ? "mixed precision: non-finite residual (NaN or Inf in the system or its solution)"
                         : "mixed precision: iterative refinement did not reach a 1e-10 relative residual");
    return MH_ERR_NOT_CONVERGED;
}

// solve(D, S, bc; direct) in mixed precision: an FP32 copy of D (K) factored
// (split-image trailing updates), the system refined in FP64 against D (K)
// to a 1e-10 relative residual.  D (K) is left intact unless the refinement
// fails and the context allows the FP64 fallback, which factors it in place
// as mh_solve does.
int mh_solve_mixed(mh_context_t ctx, int64_t n, int64_t batch, int direct, double* DK, const double* S,
                   int64_t ld, int64_t stride, const double* bc, int64_t nrhs, double* phi, int64_t ldb,
                   int64_t bstride) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(n >= 0 && batch >= 0 && nrhs >= 0, "solve_mixed: negative size");
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(DK && S && bc && phi, "solve_mixed: null array");
    MH_REQUIRE(phi != bc, "solve_mixed: phi must not alias bc");
    MH_REQUIRE(ld >= n && ldb >= n, "solve_mixed: leading dimension too small");
    MH_REQUIRE(batch == 1 || (stride >= ld * n && bstride >= ldb * nrhs), "solve_mixed: stride too small");
    MH_REQUIRE(n <= 4096, "solve_mixed: at most 4096 unknowns");
    Context* c = CTX(ctx);
    const int64_t bs = (batch == 1) ? ldb * nrhs : bstride;
    Bump bump;
    const size_t o_D32 = bump.take(sizeof(float2) * (size_t)n * (size_t)n * (size_t)batch);
    const size_t o_piv = bump.take(sizeof(int64_t) * (size_t)(n * batch));
    const size_t o_info = bump.take(sizeof(int) * (size_t)batch);
    const size_t o_perm = bump.take(sizeof(int) * (size_t)(n * batch));
    const size_t o_inv = bump.take(sizeof(float2) * (size_t)tri_inv_elems(n, batch));
    const size_t o_y = bump.take(sizeof(double2) * (size_t)(bs * batch));
    const size_t o_R = bump.take(sizeof(double2) * (size_t)(bs * batch));
    const size_t o_R32 = bump.take(sizeof(float2) * (size_t)(n * nrhs * batch));
    const size_t o_nrm = bump.take(sizeof(double) * (size_t)(nrhs * batch));
    const size_t o_tmp = bump.take(sizeof(double2) * (size_t)(bs * batch));
    const size_t o_inv64 = bump.take(sizeof(double2) * (size_t)tri_inv_elems(n, batch));  // FP64 fallback
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws);
    float2* D32 = reinterpret_cast<float2*>(ws + o_D32);
    int* info = reinterpret_cast<int*>(ws + o_info);
    const double2* dS = reinterpret_cast<const double2*>(S);
    const double2* dbc = reinterpret_cast<const double2*>(bc);
    double2* dphi = reinterpret_cast<double2*>(phi);
    double2* ybuf = reinterpret_cast<double2*>(ws + o_y);
    TRY(convert_to_f32(c, n, n, batch, reinterpret_cast<const double2*>(DK), ld, stride, D32, n, n * n));
    TRY(clear_ints(c, info, batch));
    TRY(lu_factor_f32(c, n, n, batch, D32, n, n * n, reinterpret_cast<int64_t*>(ws + o_piv), info, nullptr,
                      reinterpret_cast<int*>(ws + o_perm), reinterpret_cast<float2*>(ws + o_inv)));
    IrSystem sys[1];
    if (direct) {  // phi = D \ (S bc)   (src/matrix_assembly.jl:125-126)
        TRY(zgemm_nn_batched(c, n, nrhs, batch, dS, ld, stride, dbc, ldb, bs, ybuf, ldb, bs));
        sys[0] = {0, ybuf, nrhs, ldb, bs, dphi, reinterpret_cast<double2*>(ws + o_R),
                  reinterpret_cast<float2*>(ws + o_R32), reinterpret_cast<double*>(ws + o_nrm)};
    } else {  // sigma = K \ bc; phi = S sigma   (:127-131)
        sys[0] = {0, dbc, nrhs, ldb, bs, ybuf, reinterpret_cast<double2*>(ws + o_R),
                  reinterpret_cast<float2*>(ws + o_R32), reinterpret_cast<double*>(ws + o_nrm)};
    }
    bool nonfinite = false;
    int rc = ir_solve(c, n, batch, reinterpret_cast<const double2*>(DK), ld, stride, D32, n * n,
                      reinterpret_cast<int*>(ws + o_perm), reinterpret_cast<float2*>(ws + o_inv), sys, 1, nullptr,
                      nullptr, &nonfinite);
    if (rc == MH_ERR_NOT_CONVERGED) {
        const int rc_info = check_info(c, info, batch);  // a more specific cause
        if (rc_info != MH_OK) rc = rc_info;
    }
    if (rc == MH_ERR_NOT_CONVERGED || rc == MH_ERR_SINGULAR) {
        if (nonfinite || !c->fp64_fallback) return rc;
        c->fallbacks += 1;  // FP64 in place, as mh_solve
        MH_HIP_CHECK(hipMemcpyAsync(dphi, dbc, sizeof(double2) * (size_t)(bs * batch), hipMemcpyDeviceToDevice,
                                    c->stream));
        TRY(solve_impl(c, n, batch, direct, reinterpret_cast<double2*>(DK), dS, ld, stride, dphi, nrhs, ldb, bs,
                       reinterpret_cast<double2*>(ws + o_tmp), reinterpret_cast<int64_t*>(ws + o_piv),
                       reinterpret_cast<int*>(ws + o_perm), reinterpret_cast<double2*>(ws + o_inv64), info));
        TRY(tri_check(c));
        return check_info(c, info, batch);
    }
    TRY(rc);
    if (!direct) TRY(zgemm_nn_batched(c, n, nrhs, batch, dS, ld, stride, ybuf, ldb, bs, dphi, ldb, bs));
    TRY(tri_check(c));
    return MH_OK;
}

}  // extern "C"

namespace mh {
int sweep_impl(Context* c, const mh_mesh_t* mesh, int gf_mask, int direct, int precision,
                      int64_t nomega,
                      const double* omegas, double g, double rho, int64_t ndof,
                      const double* dof_fields, int64_t nbeta, const double* betas, double* A,
                      double* B, double* F, const GradArgs* ga) {
    const int64_t n = mesh->nfaces;
    const int64_t nr = ndof + nbeta;
    // Per frequency: S and [D | S bc] (direct) or [K | bc] (indirect), n x
    // (n + nr) column-major; the right-hand sides ride along as extra columns
    // of the factorisation, which leaves them forward-substituted.  Frequencies
    // are processed in chunks that keep the matrices under 32 GiB.
    const int64_t ncols = n + nr;
    const int64_t ld = n, stride = n * ncols, ldb = n, bstride = n * nr;
    const size_t mat_aug = sizeof(double2) * (size_t)n * (size_t)ncols;
    int64_t chunk = (int64_t)((size_t(32) << 30) / (2 * mat_aug));
    if (c->chunk_limit > 0 && chunk > c->chunk_limit) chunk = c->chunk_limit;
    if (chunk < 1) chunk = 1;
    if (chunk > nomega) chunk = nomega;
    // tangent directions per gradient pass
    const int Pg = !ga ? 1 : grad_block(ga->nparam);
    const int64_t nblocks = ga ? (ga->nparam + Pg - 1) / Pg : 0;
    const int64_t wa = n * (ndof > 0 ? ndof : 1);  // per-frequency stride of n x ndof blocks
    Bump bump;
    const size_t o_S = bump.take(mat_aug * chunk);
    const size_t o_D = bump.take(mat_aug * chunk);
    const size_t o_rhs = bump.take(sizeof(double2) * (size_t)(bstride * chunk));
    const size_t o_phi = bump.take(sizeof(double2) * (size_t)(bstride * chunk));
    const size_t o_piv = bump.take(sizeof(int64_t) * (size_t)(n * chunk));
    const size_t o_info = bump.take(sizeof(int) * (size_t)chunk);
    const size_t o_perm = bump.take(sizeof(int) * (size_t)(n * chunk));
    const size_t o_inv = bump.take(sizeof(double2) * (size_t)tri_inv_elems(n, chunk));
    const size_t o_om = bump.take(sizeof(double) * (size_t)nomega);
    const size_t o_k = bump.take(sizeof(double) * (size_t)nomega);
    const size_t o_be = bump.take(sizeof(double) * (size_t)(nbeta > 0 ? nbeta : 1));
    const size_t o_P = bump.take(sizeof(PanelRec) * (size_t)n);
    const bool mixed = precision == MH_MIXED;
    const int64_t ncol_ir = nr > ndof ? nr : ndof;  // columns of the refinement buffers
    size_t o_D32 = 0, o_y = 0, o_R = 0, o_R32 = 0, o_nrm = 0, o_R2 = 0, o_R32b = 0, o_nrm2 = 0;
    // transposed factors for the adjoint solves (FP32 copy, or FP64 in the
    // layout of the augmented matrices)
    const size_t o_Dt = (ga && ndof) ? bump.take(mixed ? sizeof(float2) * (size_t)n * (size_t)n * chunk
                                                       : mat_aug * chunk)
                                     : 0;
    // partial sums of the paired residuals of the forward and adjoint systems
    const size_t o_pa = (mixed && ga && ndof) ? bump.take(sizeof(double2) * (size_t)zgemv_pair_ws_elems(n, chunk)) : 0;
    // verdict word of a deferred refinement check
    const size_t o_dv = (mixed && ga && ndof) ? bump.take(sizeof(int)) : 0;
    if (mixed) {
        o_D32 = bump.take(sizeof(float2) * (size_t)n * (size_t)n * chunk);
        o_y = bump.take(sizeof(double2) * (size_t)(n * ncol_ir * chunk));
        o_R = bump.take(sizeof(double2) * (size_t)(n * ncol_ir * chunk));
        o_R32 = bump.take(sizeof(float2) * (size_t)(n * ncol_ir * chunk));
        o_nrm = bump.take(sizeof(double) * (size_t)(ncol_ir * chunk));
        o_R2 = bump.take(sizeof(double2) * (size_t)(n * ncol_ir * chunk));
        o_R32b = bump.take(sizeof(float2) * (size_t)(n * ncol_ir * chunk));
        o_nrm2 = bump.take(sizeof(double) * (size_t)(ncol_ir * chunk));
    }
    size_t o_cw = 0, o_lam = 0, o_mu = 0, o_tan = 0, o_dk = 0, o_dom = 0, o_part = 0, o_T1 = 0, o_T2 = 0;
    size_t o_dR = 0, o_Y = 0;
    const int rank_mask = gf_mask & (GF_RANKINE | GF_RANKINE_REFLECTED | GF_RANKINE_REFLECTED_NEG);
    const int wave_mask = gf_mask & kGfWaveMask;
    // the mixed-precision assembly takes the Delhommeau term from its table
    // (within 1e-9 of the exact term); gradients differentiate the exact term
    const int asm_mask = (mixed && (gf_mask & GF_DELHOMMEAU)) ? ((gf_mask & ~GF_DELHOMMEAU) | GF_DELH_TABLE)
                                                              : gf_mask;
    const int64_t tiles = (n + 63) / 64;
    const int64_t nout = (ndof > 0 ? ndof : 1) * nr * Pg;
    // reverse mode of the wave-term pair sums when its passes (one per block
    // of 2 outputs) cost less than the forward tangents' (one per block of
    // up to 4 parameters, times dof and right-hand-side blocks of 4); in
    // units of a 2-parameter forward pass (measured at 2,048 panels x 64
    // frequencies: 9.7 ms), a 4-parameter forward pass costs 2 (19.6 ms) and
    // a reverse pass 1.6 (15.5 ms)
    bool rev = false;
    if (ga && ndof && wave_mask) {
        const int64_t fwd10 = nblocks * ((ndof + 3) / 4) * ((nr + 3) / 4) * (Pg == 4 ? 20 : 10);
        const int64_t rev10 = (ndof * nr + 1) / 2 * 16;
        rev = c->grad_mode == 2 || (c->grad_mode == 0 && rev10 < fwd10);
    }
    size_t o_rev = 0;
    if (ga) {
        if (rev) o_rev = bump.take((size_t)wave_cot_ws_bytes(n, chunk, ndof, nr));
        o_cw = bump.take(sizeof(double2) * (size_t)(wa * chunk));
        o_lam = bump.take(sizeof(double2) * (size_t)(wa * chunk));
        o_mu = bump.take(sizeof(double2) * (size_t)(wa * chunk));
        o_tan = bump.take(sizeof(double) * (size_t)(n * 19 * Pg));
        o_dk = bump.take(sizeof(double) * (size_t)(nblocks * chunk * Pg));
        o_dom = bump.take(sizeof(double) * (size_t)(nblocks * chunk * Pg));
        if (!rev) o_part = bump.take(sizeof(double2) * (size_t)(chunk * tiles * tiles * nout));
        o_T1 = bump.take(sizeof(double2) * (size_t)(chunk * nout));
        o_T2 = bump.take(sizeof(double2) * (size_t)(chunk * nout));
        if (rank_mask) {
            o_dR = bump.take(sizeof(double) * (size_t)(2 * Pg) * (size_t)n * (size_t)n);
            o_Y = bump.take(sizeof(double2) * (size_t)(2 * Pg) * (size_t)n * (size_t)(chunk * nr));
        }
    }
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws);
    double2* dS = reinterpret_cast<double2*>(ws + o_S);
    double2* dD = reinterpret_cast<double2*>(ws + o_D);
    double2* rhs = reinterpret_cast<double2*>(ws + o_rhs);
    double2* phi = reinterpret_cast<double2*>(ws + o_phi);
    int64_t* piv = reinterpret_cast<int64_t*>(ws + o_piv);
    int* info = reinterpret_cast<int*>(ws + o_info);
    int* perm = reinterpret_cast<int*>(ws + o_perm);
    double2* inv = reinterpret_cast<double2*>(ws + o_inv);
    float2* inv32 = reinterpret_cast<float2*>(ws + o_inv);
    double* d_om = reinterpret_cast<double*>(ws + o_om);
    double* d_k = reinterpret_cast<double*>(ws + o_k);
    double* d_be = reinterpret_cast<double*>(ws + o_be);
    PanelRec* P = reinterpret_cast<PanelRec*>(ws + o_P);
    // captured (graph.hip): one chunk; frequencies, wavenumbers, headings and
    // frequency tangents from the graph's device buffers, filled before
    // each launch; the check words cleared here and set by the kernels
    const bool cap = c->capturing;
    // host staging of the copies below; it lives until the sweep's final
    // synchronisation (a copy from pageable memory may still read it later)
    std::vector<double> hk, hdk, hdom;
    if (cap) {
        MH_REQUIRE(c->cap && chunk >= nomega, "sweep graph: sweeps whose matrices fit one chunk only");
        d_om = c->cap->om;
        d_k = c->cap->k;
        d_be = c->cap->be;
        TRY(clear_ints(c, c->cap->check, kCheckWords));
        TRY(launch_pack_panels(c, mesh, P, c->cap->check));
    } else {
        hk.resize(nomega);
        for (int64_t w = 0; w < nomega; ++w) hk[w] = omegas[w] * omegas[w] / g;  // waves.jl:30
        MH_HIP_CHECK(hipMemcpyAsync(d_om, omegas, sizeof(double) * nomega, hipMemcpyHostToDevice, c->stream));
        MH_HIP_CHECK(hipMemcpyAsync(d_k, hk.data(), sizeof(double) * nomega, hipMemcpyHostToDevice, c->stream));
        if (nbeta)
            MH_HIP_CHECK(hipMemcpyAsync(d_be, betas, sizeof(double) * nbeta, hipMemcpyHostToDevice, c->stream));
        TRY(launch_pack_panels(c, mesh, P));
    }
    const int form = direct ? MH_DIRECT : MH_INDIRECT;
    double2* aug = dD + n * ld;  // first right-hand-side column of item 0
    for (int64_t w0 = 0; w0 < nomega; w0 += chunk) {
        // mixed with gradients: the refinement's convergence check runs beside
        // the gradient kernels and is read at the end of the chunk
        bool defer = mixed && !cap && ga && ndof && c->ir_deferred_sweeps > 0;
        int* dword = o_dv ? reinterpret_cast<int*>(ws + o_dv) : nullptr;
    chunk_start:
        const int64_t nw = (nomega - w0) < chunk ? (nomega - w0) : chunk;
        TRY(launch_assemble(c, mesh, P, asm_mask, form, nw, d_k + w0, dS, direct ? dD : nullptr,
                            direct ? nullptr : dD, ld, stride,
                            mixed ? reinterpret_cast<float2*>(ws + o_D32) : nullptr));
        TRY(launch_bcs(c, mesh, nw, d_om + w0, g, ndof, dof_fields, nbeta, d_be, rhs, ldb, bstride));
        float2* D32 = reinterpret_cast<float2*>(ws + o_D32);
        double2* ybuf = reinterpret_cast<double2*>(ws + o_y);
        double2* Rb = reinterpret_cast<double2*>(ws + o_R);
        float2* R32 = reinterpret_cast<float2*>(ws + o_R32);
        double* d_nrm = reinterpret_cast<double*>(ws + o_nrm);
        // the frequency-independent Rankine parts of the gradient (first
        // parameter block) depend only on the mesh: side stream, while the LU runs
        const bool early_rk = ga && ndof && rank_mask;
        if (early_rk) {
            TRY(fork_side(c));
            OnSide on(c);
            double* tan = reinterpret_cast<double*>(ws + o_tan);
            TRY(launch_pack_tangents(c, mesh, Pg, ga->nparam, 0, ga->dV, ga->dC, ga->dN, ga->dA, tan));
            TRY(launch_rankine_tangents(c, mesh, P, tan, Pg, rank_mask, form, reinterpret_cast<double*>(ws + o_dR)));
        }
        // FP64 when the mixed attempt of this chunk falls back
        bool use_mixed = mixed;
        if (mixed) {
            // FP32 LU of D (or K) + FP64 iterative refinement; D/K stays FP64.
            // The forward right-hand side S bc (direct) is formed on the side
            // stream while the LU runs.
            if (direct) {
                TRY(fork_side(c));
                OnSide on(c);
                TRY(zgemm_nn_batched(c, n, nr, nw, dS, ld, stride, rhs, ldb, bstride, ybuf, ldb, bstride));
            }
            TRY(clear_ints(c, info, nw));
            if (defer) TRY(clear_ints(c, dword, 1));
            // with an adjoint system, the transposed factors are built during the LU
            float2* Dt32 = (ga && ndof) ? reinterpret_cast<float2*>(ws + o_Dt) : nullptr;
            TRY(lu_factor_f32(c, n, n, nw, D32, n, n * n, piv, info, Dt32, perm, inv32));  // + solve data
            // forward system: phi = D \ (S bc), or sigma = K \ bc (kept in ybuf)
            IrSystem sys[2];
            if (direct) {
                TRY(join_side(c));
                sys[0] = {0, ybuf, nr, ldb, bstride, phi, Rb, R32, d_nrm};
            } else {
                sys[0] = {0, rhs, nr, ldb, bstride, ybuf, Rb, R32, d_nrm};
            }
            int nsys = 1;
            if (ga && ndof) {  // adjoint system, refined alongside: lambda = D^-T c or K^-T S^T c
                double2* cw = reinterpret_cast<double2*>(ws + o_cw);
                double2* lam = reinterpret_cast<double2*>(ws + o_lam);
                double2* mu = reinterpret_cast<double2*>(ws + o_mu);
                double2* R2 = reinterpret_cast<double2*>(ws + o_R2);
                float2* R32b = reinterpret_cast<float2*>(ws + o_R32b);
                double* d_nrm2 = reinterpret_cast<double*>(ws + o_nrm2);
                TRY(launch_force_weights(c, mesh, nw, d_om + w0, rho, ndof, dof_fields, cw, wa));
                if (direct) {
                    sys[1] = {1, cw, ndof, n, wa, lam, R2, R32b, d_nrm2};
                } else {
                    TRY(zgemm_tn_batched(c, n, ndof, nw, dS, ld, stride, cw, n, wa, mu, n, wa));
                    sys[1] = {1, mu, ndof, n, wa, lam, R2, R32b, d_nrm2};
                }
                nsys = 2;
            }
            bool nonfinite = false;
            int rc = ir_solve(c, n, nw, dD, ld, stride, D32, n * n, perm, inv32, sys, nsys,
                              nsys > 1 ? reinterpret_cast<float2*>(ws + o_Dt) : nullptr,
                              o_pa ? reinterpret_cast<double2*>(ws + o_pa) : nullptr, &nonfinite,
                              defer ? dword : nullptr);
            if (rc == MH_ERR_NOT_CONVERGED) {
                const int rc_info = check_info(c, info, nw);  // a more specific cause
                if (rc_info != MH_OK) rc = rc_info;
            }
            if (rc == MH_ERR_NOT_CONVERGED || rc == MH_ERR_SINGULAR) {
                // NaN or Inf is an input error; otherwise (refinement too
                // slow, an FP32 zero pivot, entries beyond the split range)
                // the chunk is solved again in FP64 from the FP64 D/K,
                // which the mixed attempt left intact
                if (nonfinite || !c->fp64_fallback) return rc;
                if (nsys > 1) TRY(join_aux(c));
                use_mixed = false;
                c->fallbacks += 1;
            } else {
                TRY(rc);
                // (deferred: the stream already waited for the corrections on
                // aux, which now forms the residuals)
                if (nsys > 1 && !defer) TRY(join_aux(c));
                if (!direct)
                    TRY(zgemm_nn_batched(c, n, nr, nw, dS, ld, stride, ybuf, ldb, bstride, phi, ldb, bstride));
            }
        }
        if (!use_mixed) {
            if (direct) {  // phi = D \ (S bc)   (src/matrix_assembly.jl:125-126)
                TRY(zgemm_nn_batched(c, n, nr, nw, dS, ld, stride, rhs, ldb, bstride, aug, ld, stride));
            } else {  // sigma = K \ bc; phi = S sigma   (:127-131)
                TRY(copy_d2(c, aug, stride, rhs, bstride, bstride, nw));
            }
            TRY(clear_ints(c, info, nw));  // a fill kernel: capturable
            // the transposed FP64 factors only when the workspace was laid out
            // for the FP64 path (a fallback solves the adjoint from the factors)
            double2* dDt = (ga && ndof && !mixed) ? reinterpret_cast<double2*>(ws + o_Dt) : nullptr;
            TRY(lu_factor(c, n, ncols, nw, dD, ld, stride, piv, info, dDt, (ga && ndof) ? perm : nullptr,
                          inv));  // + transposed factors and solve data
            if (ga && ndof) {  // adjoint solve on aux, alongside the back substitution
                double2* cw = reinterpret_cast<double2*>(ws + o_cw);
                double2* lam = reinterpret_cast<double2*>(ws + o_lam);
                TRY(launch_force_weights(c, mesh, nw, d_om + w0, rho, ndof, dof_fields, cw, wa));
                TRY(fork_aux(c));
                OnAux on(c, true);
                if (direct)  // lambda = D^-T c
                    TRY(copy_d2(c, lam, 0, cw, 0, wa * nw, 1));
                else  // lambda = K^-T S^T c
                    TRY(zgemm_tn_batched(c, n, ndof, nw, dS, ld, stride, cw, n, wa, lam, n, wa));
                TRY(tri_solve(c, n, nw, dD, ld, stride, perm, inv, 1, 3, ndof, lam, n, wa, dDt));
            }
            TRY(tri_solve(c, n, nw, dD, ld, stride, nullptr, inv, 0, 2, nr, aug, ld, stride));
            if (ga && ndof) TRY(join_aux(c));
            if (direct) {
                TRY(copy_d2(c, phi, bstride, aug, stride, bstride, nw));
            } else {
                TRY(zgemm_nn_batched(c, n, nr, nw, dS, ld, stride, aug, ld, stride, phi, ldb, bstride));
            }
        }
        double* Aw = A ? A + w0 * ndof * ndof : nullptr;
        double* Bw = B ? B + w0 * ndof * ndof : nullptr;
        double2* Fw = F ? reinterpret_cast<double2*>(F) + w0 * nbeta * ndof : nullptr;
        TRY(launch_forces(c, mesh, nw, d_om + w0, rho, ndof, dof_fields, nbeta, phi, ldb, bstride, Aw,
                          Bw, Fw));
        // solve failures (a wait that timed out, a zero pivot) are reported
        // at the end of the chunk, after the gradient kernels: one host
        // synchronisation per chunk
        if (!ga || ndof == 0) {
            if (cap) {  // checked after each launch of the graph
                TRY(or_words(c, info, nw, c->cap->check + 1));
                continue;
            }
            TRY(tri_check(c));
            TRY(check_info(c, info, nw));
            continue;
        }
        if (early_rk) TRY(join_side(c));
        // ---- gradients (grad.hip) ----
        double2* cw = reinterpret_cast<double2*>(ws + o_cw);
        double2* lam = reinterpret_cast<double2*>(ws + o_lam);
        double2* mu = reinterpret_cast<double2*>(ws + o_mu);
        double* tan = reinterpret_cast<double*>(ws + o_tan);
        double* d_dk = reinterpret_cast<double*>(ws + o_dk);
        double* d_dom = reinterpret_cast<double*>(ws + o_dom);
        double2* part = reinterpret_cast<double2*>(ws + o_part);
        double2* T1 = reinterpret_cast<double2*>(ws + o_T1);
        double2* T2 = reinterpret_cast<double2*>(ws + o_T2);
        // lambda (and, indirect mixed, mu = S^T c) were computed with the solves
        const double2 *aS, *aDw, *sS, *sD, *rhoW;
        int64_t wss;
        if (direct) {  // lambda = D^-T c; mu = S^T lambda (on aux, alongside the pair kernels)
            TRY(fork_aux(c));
            {
                OnAux on(c, true);
                TRY(zgemm_tn_batched(c, n, ndof, nw, dS, ld, stride, lam, n, wa, mu, n, wa));
            }
            aS = lam; aDw = lam; sS = rhs; sD = phi; wss = bstride; rhoW = mu;
        } else if (use_mixed) {  // lambda = K^-T S^T c
            aS = cw; aDw = lam; sS = ybuf; sD = ybuf; wss = bstride; rhoW = lam;
        } else {
            aS = cw; aDw = lam; sS = aug; sD = aug; wss = stride; rhoW = lam;
        }
        const int64_t np = ga->nparam;
        // omega and wavenumber tangents of every parameter block (nblocks x
        // nw x Pg each), copied once per chunk (a captured sweep reads them
        // from the graph's buffers)
        if (!cap) {
            hdk.resize(nblocks * nw * Pg);
            hdom.resize(nblocks * nw * Pg);
            omega_tangents(omegas, nomega, w0, nw, ga->domega, np, Pg, g, hdk.data(), hdom.data());
            MH_HIP_CHECK(hipMemcpyAsync(d_dk, hdk.data(), sizeof(double) * hdk.size(), hipMemcpyHostToDevice,
                                        c->stream));
            MH_HIP_CHECK(hipMemcpyAsync(d_dom, hdom.data(), sizeof(double) * hdom.size(), hipMemcpyHostToDevice,
                                        c->stream));
        } else {
            d_dk = c->cap->dk;
            d_dom = c->cap->dom;
        }
        // reverse mode: the wave-term cotangents once for every parameter
        double2 *cot = nullptr, *cotk = nullptr;
        if (rev)
            TRY(launch_wave_cot(c, mesh, P, wave_mask, form, nw, d_k + w0, ndof, nr, aS, aDw, sS, sD, wa, wss,
                                ws + o_rev, &cot, &cotk));
        for (int64_t m0 = 0; m0 < np; m0 += Pg) {
            const double* bdk = d_dk + (m0 / Pg) * nw * Pg;
            const double* bdom = d_dom + (m0 / Pg) * nw * Pg;
            const bool done = early_rk && m0 == 0;  // tangents and Rankine parts already there
            if (!done) TRY(launch_pack_tangents(c, mesh, Pg, np, m0, ga->dV, ga->dC, ga->dN, ga->dA, tan));
            // the Rankine parts are contracted into T2 on the side stream while
            // the wave-term pair kernel runs; its partial sums are added after
            TRY(fill_zero(c, T2, sizeof(double2) * (size_t)(nw * ndof * nr * Pg)));
            if (rank_mask) {
                TRY(fork_side(c));
                OnSide on(c);
                double* dR = reinterpret_cast<double*>(ws + o_dR);
                double2* Yb = reinterpret_cast<double2*>(ws + o_Y);
                if (!done) TRY(launch_rankine_tangents(c, mesh, P, tan, Pg, rank_mask, form, dR));
                TRY(launch_rankine_contract(c, n, Pg, nw, ndof, nr, dR, aS, aDw, sS, sD, wa, wss, Yb, T2));
            }
            int64_t nblk = 0;
            if (wave_mask && !rev)
                TRY(launch_tangent_pairs(c, mesh, P, tan, Pg, wave_mask, form, nw, d_k + w0, bdk, ndof, nr, aS,
                                         aDw, sS, sD, wa, wss, part, &nblk));
            if (rank_mask) TRY(join_side(c));
            if (rev)
                TRY(launch_wave_cot_contract(c, n, nw, ndof, nr, Pg, np, m0, cot, cotk, ga->dC, ga->dN, ga->dA,
                                             bdk, T2));
            else
                TRY(launch_reduce_partials(c, nw, nblk, ndof * nr * Pg, part, T2, 1));
            if (direct) TRY(join_aux(c));
            TRY(launch_bc_tangent(c, mesh, Pg, tan, nw, d_om + w0, bdom, g, rho, ndof, dof_fields,
                                  ga->ddofs, np, m0, nbeta, d_be, phi, bstride, rhoW, wa, T1));
            TRY(launch_grad_finish(c, nw, np, m0, Pg, ndof, nr, nbeta, d_om + w0, bdom, Aw, Bw, T1, T2,
                                   ga->dAout ? ga->dAout + w0 * np * ndof * ndof : nullptr,
                                   ga->dBout ? ga->dBout + w0 * np * ndof * ndof : nullptr,
                                   (ga->dFout && nbeta) ? ga->dFout + 2 * w0 * np * nbeta * ndof : nullptr));
        }
        if (cap) {  // checked after each launch of the graph
            TRY(or_words(c, info, nw, c->cap->check + 1));
            continue;
        }
        if (defer) {
            // the deferred refinement verdict and the FP32 factorisation's
            // info words: a chunk that missed the tolerance (or whose FP32
            // factors had a zero pivot or overflowed) is solved again with a
            // check after every sweep, which refines further, falls back to
            // FP64 or reports the error as without the deferral
            TRY(join_aux(c));
            int hv = 0;
            std::vector<int> hinfo((size_t)nw);
            MH_HIP_CHECK(hipMemcpyAsync(&hv, dword, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            MH_HIP_CHECK(hipMemcpyAsync(hinfo.data(), info, sizeof(int) * nw, hipMemcpyDeviceToHost, c->stream));
            MH_HIP_CHECK(hipStreamSynchronize(c->stream));
            bool bad = hv != 0;
            for (int v : hinfo) bad = bad || v != 0;
            if (bad) {
                c->deferred_reruns += 1;
                defer = false;
                goto chunk_start;
            }
        }
        TRY(tri_check(c));
        TRY(check_info(c, info, nw));
    }
    return MH_OK;
}

void omega_tangents(const double* omegas, int64_t nomega, int64_t w0, int64_t nw, const double* domega,
                    int64_t nparam, int Pg, double g, double* dk, double* dom) {
    for (int64_t m0 = 0; m0 < nparam; m0 += Pg) {
        double* bdk = dk + (m0 / Pg) * nw * Pg;
        double* bdom = dom + (m0 / Pg) * nw * Pg;
        for (int64_t w = 0; w < nw; ++w)
            for (int m = 0; m < Pg; ++m) {
                const double dw = (domega && m0 + m < nparam) ? domega[(m0 + m) * nomega + w0 + w] : 0.0;
                bdom[w * Pg + m] = dw;
                bdk[w * Pg + m] = 2.0 * omegas[w0 + w] / g * dw;
            }
    }
}

int sweep_args_ok(const mh_mesh_t* mesh, int gf_mask, int precision, int64_t nomega, const double* omegas,
                  double g, double rho, int64_t ndof, const double* dof_fields, int64_t nbeta,
                  const double* betas, double* A, double* B, double* F) {
    TRY(check_mesh(mesh));
    TRY(check_gf(gf_mask, "freq_sweep"));
    MH_REQUIRE(nomega >= 0 && ndof >= 0 && nbeta >= 0, "freq_sweep: negative size");
    MH_REQUIRE(g > 0.0 && rho > 0.0, "freq_sweep: g and rho must be positive");
    MH_REQUIRE(precision == MH_FP64 || precision == MH_MIXED, "freq_sweep: unknown precision");
    const int64_t n = mesh->nfaces;
    const int64_t nr = ndof + nbeta;
    if (n == 0 || nomega == 0 || nr == 0) return MH_OK;
    MH_REQUIRE(n <= 4096, "freq_sweep: at most 4096 panels per body");
    MH_REQUIRE(nr <= 16, "freq_sweep: at most 16 dofs + headings");
    MH_REQUIRE(omegas, "freq_sweep: null omegas");
    MH_REQUIRE(ndof == 0 || (dof_fields && A && B), "freq_sweep: null dof/output array");
    MH_REQUIRE(nbeta == 0 || (betas && F), "freq_sweep: null heading/output array");
    for (int64_t w = 0; w < nomega; ++w)
        MH_REQUIRE(omegas[w] > 0.0, "freq_sweep: frequencies must be positive");
    return MH_OK;
}

}  // namespace mh

extern "C" {

static int sweep_checks(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int precision,
                        int64_t nomega, const double* omegas, double g, double rho, int64_t ndof,
                        const double* dof_fields, int64_t nbeta, const double* betas, double* A,
                        double* B, double* F) {
    TRY(check_ctx(ctx));
    return sweep_args_ok(mesh, gf_mask, precision, nomega, omegas, g, rho, ndof, dof_fields, nbeta, betas, A, B,
                         F);
}

int mh_freq_sweep(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int direct, int precision,
                  int64_t nomega, const double* omegas, double g, double rho, int64_t ndof,
                  const double* dof_fields, int64_t nbeta, const double* betas, double* A,
                  double* B, double* F) {
    TRY(sweep_checks(ctx, mesh, gf_mask, precision, nomega, omegas, g, rho, ndof, dof_fields, nbeta,
                     betas, A, B, F));
    if (mesh->nfaces == 0 || nomega == 0 || ndof + nbeta == 0) return MH_OK;
    return sweep_impl(CTX(ctx), mesh, gf_mask, direct, precision, nomega, omegas, g, rho, ndof,
                      dof_fields, nbeta, betas, A, B, F, nullptr);
}

int mh_freq_sweep_grad(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int direct,
                       int precision, int64_t nomega, const double* omegas, double g, double rho,
                       int64_t ndof, const double* dof_fields, int64_t nbeta, const double* betas,
                       int64_t nparam, const double* dvertices, const double* dcenters,
                       const double* dnormals, const double* dareas, const double* ddof_fields,
                       const double* domegas, double* A, double* B, double* F, double* dA,
                       double* dB, double* dF) {
    TRY(sweep_checks(ctx, mesh, gf_mask, precision, nomega, omegas, g, rho, ndof, dof_fields, nbeta,
                     betas, A, B, F));
    MH_REQUIRE(nparam >= 0, "freq_sweep_grad: negative nparam");
    if (mesh->nfaces == 0 || nomega == 0 || ndof + nbeta == 0) return MH_OK;
    MH_REQUIRE(ndof >= 1, "freq_sweep_grad: at least one dof (the influenced force)");
    MH_REQUIRE(nparam == 0 || (dA && dB && (nbeta == 0 || dF)), "freq_sweep_grad: null gradient output");
    GradArgs ga{nparam, dvertices, dcenters, dnormals, dareas, ddof_fields, domegas, dA, dB, dF};
    return sweep_impl(CTX(ctx), mesh, gf_mask, direct, precision, nomega, omegas, g, rho, ndof,
                      dof_fields, nbeta, betas, A, B, F, nparam > 0 ? &ga : nullptr);
}

// ---- reverse mode: assembly and solve ---------------------------------------
int mh_assemble_vjp(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int formulation, int64_t nk,
                    const double* wavenumbers, const double* Sbar, const double* DKbar, int64_t ld,
                    int64_t stride, double* dvertices, double* dcenters, double* dnormals, double* dareas,
                    double* dwavenumbers) {
    TRY(check_ctx(ctx));
    TRY(check_mesh(mesh));
    TRY(check_gf(gf_mask, "assemble_vjp"));
    MH_REQUIRE(formulation == MH_DIRECT || formulation == MH_INDIRECT,
               "assemble_vjp: formulation must be MH_DIRECT or MH_INDIRECT");
    MH_REQUIRE(nk >= 0, "assemble_vjp: negative nk");
    const int64_t n = mesh->nfaces;
    if (n == 0 || nk == 0) return MH_OK;
    MH_REQUIRE(wavenumbers, "assemble_vjp: null wavenumbers");
    MH_REQUIRE(ld >= n && (nk == 1 || stride >= ld * n), "assemble_vjp: ld/stride too small");
    Context* c = CTX(ctx);
    Bump bump;
    const size_t o_p = bump.take(sizeof(PanelRec) * (size_t)n);
    const size_t o_k = bump.take(sizeof(double) * (size_t)nk);
    const size_t o_w = bump.take(sizeof(double) * (size_t)assemble_vjp_ws_doubles(n, nk));
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws);
    PanelRec* P = reinterpret_cast<PanelRec*>(ws + o_p);
    double* dk = reinterpret_cast<double*>(ws + o_k);
    MH_HIP_CHECK(hipMemcpyAsync(dk, wavenumbers, sizeof(double) * nk, hipMemcpyHostToDevice, c->stream));
    TRY(launch_pack_panels(c, mesh, P));
    TRY(launch_assemble_vjp(c, mesh, P, gf_mask, formulation, nk, dk, reinterpret_cast<const double2*>(Sbar),
                            reinterpret_cast<const double2*>(DKbar), ld, stride,
                            reinterpret_cast<double*>(ws + o_w), dvertices, dcenters, dnormals, dareas,
                            dwavenumbers));
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MH_OK;
}

int mh_solve_vjp(mh_context_t ctx, int64_t n, int64_t batch, int direct, const double* LU, const int64_t* ipiv,
                 const double* S, int64_t ld, int64_t stride, const double* bc, int64_t nrhs, const double* phibar,
                 int64_t ldb, int64_t bstride, double* DKbar, double* Sbar, double* bcbar) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(n >= 0 && batch >= 0 && nrhs >= 0, "solve_vjp: negative size");
    if (n == 0 || batch == 0 || nrhs == 0) return MH_OK;
    MH_REQUIRE(LU && ipiv && S && bc && phibar, "solve_vjp: null array");
    MH_REQUIRE(ld >= n && ldb >= n, "solve_vjp: leading dimension too small");
    MH_REQUIRE(batch == 1 || (stride >= ld * n && bstride >= ldb * nrhs), "solve_vjp: stride too small");
    MH_REQUIRE(n <= kMaxSolveN, "solve_vjp: at most 4096 unknowns");
    Context* c = CTX(ctx);
    const int64_t bs = n * nrhs;  // internal n x nrhs blocks
    Bump bump;
    const size_t o_perm = bump.take(sizeof(int) * (size_t)(n * batch));
    const size_t o_inv = bump.take(sizeof(double2) * (size_t)tri_inv_elems(n, batch));
    const size_t o_a = bump.take(sizeof(double2) * (size_t)(bs * batch));
    const size_t o_b = bump.take(sizeof(double2) * (size_t)(bs * batch));
    const size_t o_c = bump.take(sizeof(double2) * (size_t)(bs * batch));
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws);
    int* perm = reinterpret_cast<int*>(ws + o_perm);
    double2* inv = reinterpret_cast<double2*>(ws + o_inv);
    double2* ta = reinterpret_cast<double2*>(ws + o_a);
    double2* tb = reinterpret_cast<double2*>(ws + o_b);
    double2* tc = reinterpret_cast<double2*>(ws + o_c);
    const double2* LU2 = reinterpret_cast<const double2*>(LU);
    const double2* S2 = reinterpret_cast<const double2*>(S);
    const double2* b2 = reinterpret_cast<const double2*>(bc);
    const double2* pb = reinterpret_cast<const double2*>(phibar);
    TRY(tri_prepare(c, n, batch, LU2, ld, stride, ipiv, perm, inv));
    // ta <- conj(phibar)
    TRY(conj_batched(c, n, nrhs, batch, pb, ldb, bstride, ta, n, bs));
    if (direct) {
        // y = S bc, phi = D \ y:  ybar = D^-H phibar, Dbar = -ybar phi^H,
        // Sbar = ybar bc^H, bcbar = S^H ybar
        TRY(zgemm_nn_batched(c, n, nrhs, batch, S2, ld, stride, b2, ldb, bstride, tb, n, bs));
        TRY(tri_solve(c, n, batch, LU2, ld, stride, perm, inv, 0, 3, nrhs, tb, n, bs));      // tb = phi
        TRY(tri_solve(c, n, batch, LU2, ld, stride, perm, inv, 1, 3, nrhs, ta, n, bs));      // ta = conj(ybar)
        TRY(conj_batched(c, n, nrhs, batch, ta, n, bs, tc, n, bs));                          // tc = ybar
        if (DKbar)
            TRY(outer_conj_batched(c, n, nrhs, batch, -1.0, tc, tb, n, bs, reinterpret_cast<double2*>(DKbar), ld,
                                   stride));
        if (Sbar) {
            // bc may have its own layout: copy into tb (phi is no longer needed)
            TRY(conj_batched(c, n, nrhs, batch, b2, ldb, bstride, tb, n, bs));
            TRY(conj_batched(c, n, nrhs, batch, tb, n, bs, tb, n, bs));
            TRY(outer_conj_batched(c, n, nrhs, batch, 1.0, tc, tb, n, bs, reinterpret_cast<double2*>(Sbar), ld,
                                   stride));
        }
        if (bcbar) {  // S^H ybar = conj(S^T conj(ybar))
            TRY(zgemm_tn_batched(c, n, nrhs, batch, S2, ld, stride, ta, n, bs, tb, n, bs));
            TRY(conj_batched(c, n, nrhs, batch, tb, n, bs, reinterpret_cast<double2*>(bcbar), ldb, bstride));
        }
    } else {
        // sigma = K \ bc, phi = S sigma:  Sbar = phibar sigma^H,
        // sigmabar = S^H phibar, bcbar = K^-H sigmabar, Kbar = -bcbar sigma^H
        TRY(conj_batched(c, n, nrhs, batch, b2, ldb, bstride, tb, n, bs));
        TRY(conj_batched(c, n, nrhs, batch, tb, n, bs, tb, n, bs));                          // tb = bc
        TRY(tri_solve(c, n, batch, LU2, ld, stride, perm, inv, 0, 3, nrhs, tb, n, bs));      // tb = sigma
        if (Sbar) {
            TRY(conj_batched(c, n, nrhs, batch, ta, n, bs, tc, n, bs));                      // tc = phibar
            TRY(outer_conj_batched(c, n, nrhs, batch, 1.0, tc, tb, n, bs, reinterpret_cast<double2*>(Sbar), ld,
                                   stride));
        }
        // conj(sigmabar) = S^T conj(phibar);  K^T t = conj(sigmabar) gives t = conj(bcbar)
        TRY(zgemm_tn_batched(c, n, nrhs, batch, S2, ld, stride, ta, n, bs, tc, n, bs));
        TRY(tri_solve(c, n, batch, LU2, ld, stride, perm, inv, 1, 3, nrhs, tc, n, bs));
        TRY(conj_batched(c, n, nrhs, batch, tc, n, bs, ta, n, bs));                          // ta = bcbar
        if (bcbar) TRY(conj_batched(c, n, nrhs, batch, tc, n, bs, reinterpret_cast<double2*>(bcbar), ldb, bstride));
        if (DKbar)
            TRY(outer_conj_batched(c, n, nrhs, batch, -1.0, ta, tb, n, bs, reinterpret_cast<double2*>(DKbar), ld,
                                   stride));
    }
    return tri_check(c);
}

int mh_set_refinement(mh_context_t ctx, int max_sweeps, int fp64_fallback) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(max_sweeps >= 1 && max_sweeps <= 100, "set_refinement: max_sweeps must be in [1, 100]");
    CTX(ctx)->ir_max_sweeps = max_sweeps;
    CTX(ctx)->fp64_fallback = fp64_fallback != 0;
    return MH_OK;
}

int mh_set_gradient_mode(mh_context_t ctx, int mode) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(mode >= 0 && mode <= 2, "set_gradient_mode: mode must be 0 (chosen per sweep), 1 (forward) or 2 (reverse)");
    CTX(ctx)->grad_mode = mode;
    return MH_OK;
}

int mh_set_far_field_fp32(mh_context_t ctx, int on) {
    TRY(check_ctx(ctx));
    CTX(ctx)->far32 = on != 0;
    return MH_OK;
}

int mh_set_deferred_check(mh_context_t ctx, int sweeps) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(sweeps >= 0 && sweeps <= 12, "set_deferred_check: sweeps must be in [0, 12]");
    CTX(ctx)->ir_deferred_sweeps = sweeps;
    return MH_OK;
}

int mh_set_sweep_chunk(mh_context_t ctx, int64_t max_frequencies) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(max_frequencies >= 0, "set_sweep_chunk: max_frequencies must be >= 0");
    CTX(ctx)->chunk_limit = max_frequencies;
    return MH_OK;
}

int mh_deferred_rerun_count(mh_context_t ctx, int64_t* count) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(count != nullptr, "deferred_rerun_count: null output");
    *count = CTX(ctx)->deferred_reruns;
    return MH_OK;
}

int mh_fallback_count(mh_context_t ctx, int64_t* count) {
    TRY(check_ctx(ctx));
    MH_REQUIRE(count != nullptr, "fallback_count: null output");
    *count = CTX(ctx)->fallbacks;
    return MH_OK;
}

// ---- Froude-Krylov force ----------------------------------------------------
static int fk_impl(Context* c, const mh_mesh_t* mesh, int64_t nomega, const double* omegas, double g,
                   double rho, int64_t ndof, const double* dofs, int64_t nbeta, const double* betas,
                   double* FK, int64_t np, const double* dC, const double* dN, const double* dA,
                   const double* ddofs, const double* domegas, double* dFK) {
    MH_REQUIRE(nomega >= 0 && ndof >= 0 && nbeta >= 0 && np >= 0, "froude_krylov: negative size");
    MH_REQUIRE(g > 0.0 && rho > 0.0, "froude_krylov: g and rho must be positive");
    if (mesh->nfaces == 0 || nomega == 0 || ndof == 0 || nbeta == 0) return MH_OK;
    MH_REQUIRE(omegas && betas && dofs && FK, "froude_krylov: null array");
    MH_REQUIRE(ndof <= 16, "froude_krylov: at most 16 dofs");
    for (int64_t w = 0; w < nomega; ++w)
        MH_REQUIRE(omegas[w] > 0.0, "froude_krylov: frequencies must be positive");
    MH_REQUIRE(np == 0 || dFK, "froude_krylov: null gradient output");
    Bump bump;
    const size_t o_om = bump.take(sizeof(double) * (size_t)nomega);
    const size_t o_be = bump.take(sizeof(double) * (size_t)nbeta);
    const size_t o_dom = bump.take(sizeof(double) * (size_t)(nomega * (np > 0 ? np : 1)));
    if (!ensure_workspace(c, bump.off)) return MH_ERR_HIP;
    char* ws = static_cast<char*>(c->ws);
    double* d_om = reinterpret_cast<double*>(ws + o_om);
    double* d_be = reinterpret_cast<double*>(ws + o_be);
    double* d_dom = reinterpret_cast<double*>(ws + o_dom);
    // host tangents of omega (np x nomega) -> device (nomega x np); the
    // pageable copies complete before the call returns
    std::vector<double> hdom((size_t)(nomega * (np > 0 ? np : 1)), 0.0);
    if (domegas)
        for (int64_t m = 0; m < np; ++m)
            for (int64_t w = 0; w < nomega; ++w) hdom[w * np + m] = domegas[m * nomega + w];
    MH_HIP_CHECK(hipMemcpyAsync(d_om, omegas, sizeof(double) * nomega, hipMemcpyHostToDevice, c->stream));
    MH_HIP_CHECK(hipMemcpyAsync(d_be, betas, sizeof(double) * nbeta, hipMemcpyHostToDevice, c->stream));
    if (np > 0)
        MH_HIP_CHECK(hipMemcpyAsync(d_dom, hdom.data(), sizeof(double) * hdom.size(), hipMemcpyHostToDevice,
                                    c->stream));
    TRY(launch_froude_krylov(c, mesh, nomega, d_om, g, rho, ndof, dofs, nbeta, d_be,
                             reinterpret_cast<double2*>(FK), np, dC, dN, dA, ddofs, d_dom,
                             reinterpret_cast<double2*>(dFK)));
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MH_OK;
}

int mh_froude_krylov(mh_context_t ctx, const mh_mesh_t* mesh, int64_t nomega, const double* omegas,
                     double g, double rho, int64_t ndof, const double* dof_fields, int64_t nbeta,
                     const double* betas, double* FK) {
    TRY(check_ctx(ctx));
    TRY(check_mesh(mesh));
    return fk_impl(CTX(ctx), mesh, nomega, omegas, g, rho, ndof, dof_fields, nbeta, betas, FK, 0, nullptr,
                   nullptr, nullptr, nullptr, nullptr, nullptr);
}

int mh_froude_krylov_grad(mh_context_t ctx, const mh_mesh_t* mesh, int64_t nomega, const double* omegas,
                          double g, double rho, int64_t ndof, const double* dof_fields, int64_t nbeta,
                          const double* betas, int64_t nparam, const double* dcenters,
                          const double* dnormals, const double* dareas, const double* ddof_fields,
                          const double* domegas, double* FK, double* dFK) {
    TRY(check_ctx(ctx));
    TRY(check_mesh(mesh));
    return fk_impl(CTX(ctx), mesh, nomega, omegas, g, rho, ndof, dof_fields, nbeta, betas, FK, nparam,
                   dcenters, dnormals, dareas, ddof_fields, domegas, dFK);
}

// ---- host-array entry points -----------------------------------------------
static int host_single(int64_t nf, int64_t nv, const double* V, const int64_t* Fc,
                       const double* C, const double* N, const double* Ar, const double* R,
                       const double* dof3, double omega, double g, double rho, int want_diff,
                       double* out) {
    MH_REQUIRE(nf > 0 && nv > 0, "host entry: empty mesh");
    MH_REQUIRE(V && Fc && C && N && Ar && R && dof3 && out, "host entry: null array");
    mh_context_t ctx;
    TRY(mh_context_create(0, nullptr, &ctx));
    std::vector<void*> bufs;
    auto dev = [&](const void* h, size_t bytes) -> void* {
        void* d = nullptr;
        if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
        bufs.push_back(d);
        if (h && hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return d;
    };
    int rc = MH_OK;
    std::vector<double> field(nf * 3);
    for (int c = 0; c < 3; ++c)
        for (int64_t i = 0; i < nf; ++i) field[i + c * nf] = dof3[c];
    mh_mesh_t m;
    m.nfaces = nf;
    m.nvertices = nv;
    m.vertices = (const double*)dev(V, sizeof(double) * nv * 3);
    m.faces = (const int64_t*)dev(Fc, sizeof(int64_t) * nf * 4);
    m.centers = (const double*)dev(C, sizeof(double) * nf * 3);
    m.normals = (const double*)dev(N, sizeof(double) * nf * 3);
    m.areas = (const double*)dev(Ar, sizeof(double) * nf);
    m.radii = (const double*)dev(R, sizeof(double) * nf);
    double* dfield = (double*)dev(field.data(), sizeof(double) * nf * 3);
    double* dA = (double*)dev(nullptr, sizeof(double));
    double* dB = (double*)dev(nullptr, sizeof(double));
    double* dF = (double*)dev(nullptr, sizeof(double) * 2);
    if (!m.vertices || !m.faces || !m.centers || !m.normals || !m.areas || !m.radii || !dfield ||
        !dA || !dB || !dF) {
        set_error("host entry: device allocation or copy failed");
        rc = MH_ERR_HIP;
    }
    const double beta = 0.0;
    if (rc == MH_OK)
        rc = mh_freq_sweep(ctx, &m, MH_GF_DEFAULT, 1, MH_FP64, 1, &omega, g, rho, 1, dfield,
                           want_diff ? 1 : 0, &beta, dA, dB, want_diff ? dF : nullptr);
    if (rc == MH_OK) {
        if (want_diff) {
            if (hipMemcpy(out, dF, sizeof(double) * 2, hipMemcpyDeviceToHost) != hipSuccess) rc = MH_ERR_HIP;
        } else {
            if (hipMemcpy(out, dA, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(out + 1, dB, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
                rc = MH_ERR_HIP;
        }
        if (rc != MH_OK) set_error("host entry: copy back failed");
    }
    for (void* p : bufs) (void)hipFree(p);
    mh_context_destroy(ctx);
    return rc;
}

int mh_calculate_radiation_forces_host(int64_t nfaces, int64_t nvertices, const double* vertices,
                                       const int64_t* faces, const double* centers,
                                       const double* normals, const double* areas,
                                       const double* radii, const double* dof3, double omega,
                                       double g, double rho, double* out) {
    MH_REQUIRE(omega > 0.0, "calculate_radiation_forces: omega must be positive");
    return host_single(nfaces, nvertices, vertices, faces, centers, normals, areas, radii, dof3,
                       omega, g, rho, 0, out);
}

int mh_diffraction_force_host(int64_t nfaces, int64_t nvertices, const double* vertices,
                              const int64_t* faces, const double* centers, const double* normals,
                              const double* areas, const double* radii, const double* dof3,
                              double omega, double g, double rho, double* out) {
    MH_REQUIRE(omega > 0.0, "DiffractionForce: omega must be positive");
    return host_single(nfaces, nvertices, vertices, faces, centers, normals, areas, radii, dof3,
                       omega, g, rho, 1, out);
}

}  // extern "C"
