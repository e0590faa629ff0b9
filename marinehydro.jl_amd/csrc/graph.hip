// Captured FP64 frequency sweeps: one body's forward sweep (mh_freq_sweep
// with MH_FP64) recorded once as a HIP graph and replayed for new
// frequencies and headings, for callers that run the same body, or bodies
// of the same shape, many times (design loops, serving).
//
// The graph lives on a private context (own streams, workspace and
// buffers): the addresses it recorded never move, whatever else runs on
// the caller's context.  A sizing sweep with placeholder frequencies first
// allocates everything the capture will record (workspace, Rankine parts,
// solve state, events, the look-ahead streams); the capture then records
// the same sequence from the panel packing to the forces, with the host
// steps left out: the frequencies, wavenumbers and headings are copied into
// the recorded buffers before each launch, and the mesh-index flag, the
// solve error word and the pivot info words are read after it.
#include <cmath>

#include "mh_internal.hpp"

namespace mh {

struct SweepGraph {
    Context* c = nullptr;  // private context
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev = nullptr;  // orders a launch after the caller's stream
    int64_t nomega = 0, nbeta = 0;
    double g = 9.81;
    double* pin = nullptr;  // pinned staging: omegas, wavenumbers, headings; then the check words
    double *d_om = nullptr, *d_k = nullptr, *d_be = nullptr;
    int* info = nullptr;
};

static void graph_free(SweepGraph* sg) {
    if (!sg) return;
    if (sg->c) {
        (void)hipSetDevice(sg->c->device);
        (void)hipStreamSynchronize(sg->c->stream);
        if (sg->exec) (void)hipGraphExecDestroy(sg->exec);
        if (sg->c->cap_graph) (void)hipGraphDestroy(sg->c->cap_graph);
        sg->c->cap_graph = nullptr;
    }
    if (sg->ev) (void)hipEventDestroy(sg->ev);
    if (sg->pin) (void)hipHostFree(sg->pin);
    if (sg->c) context_free(sg->c);
    delete sg;
}

// the private context's stream waits for the work already issued on the
// caller's stream (mesh or dof updates)
static int after_caller(SweepGraph* sg, Context* caller) {
    MH_HIP_CHECK(hipEventRecord(sg->ev, caller->stream));
    MH_HIP_CHECK(hipStreamWaitEvent(sg->c->stream, sg->ev, 0));
    return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_sweep_graph_create(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int direct, int64_t nomega,
                          int64_t ndof, const double* dof_fields, int64_t nbeta, double g, double rho, double* A,
                          double* B, double* F, mh_sweep_graph_t* out) {
    MH_REQUIRE(ctx != nullptr && out != nullptr, "sweep graph: null context or output handle");
    *out = nullptr;
    Context* cc = reinterpret_cast<Context*>(ctx);
    MH_HIP_CHECK(hipSetDevice(cc->device));
    MH_REQUIRE(mesh != nullptr && nomega > 0, "sweep graph: null mesh or no frequencies");
    // placeholders for the argument checks and the sizing sweep
    std::vector<double> om((size_t)nomega, 1.0), be((size_t)(nbeta > 0 ? nbeta : 1), 0.0);
    TRY_RC(sweep_args_ok(mesh, gf_mask, MH_FP64, nomega, om.data(), g, rho, ndof, dof_fields, nbeta, be.data(), A,
                         B, F));
    MH_REQUIRE(mesh->nfaces > 0 && ndof + nbeta > 0, "sweep graph: empty sweep");
    SweepGraph* sg = new SweepGraph();
    sg->nomega = nomega;
    sg->nbeta = nbeta;
    sg->g = g;
    int rc = context_new(cc->device, nullptr, true, &sg->c);
    if (rc != MH_OK) {
        delete sg;
        return rc;
    }
    if (hipEventCreateWithFlags(&sg->ev, hipEventDisableTiming) != hipSuccess) {
        set_error("sweep graph: hipEventCreate failed");
        rc = MH_ERR_HIP;
    }
    if (rc == MH_OK) rc = after_caller(sg, cc);
    // sizing sweep, then the capture of the same sequence
    if (rc == MH_OK)
        rc = sweep_impl(sg->c, mesh, gf_mask, direct, MH_FP64, nomega, om.data(), g, rho, ndof, dof_fields, nbeta,
                        be.data(), A, B, F, nullptr);
    if (rc == MH_OK) {
        sg->c->capturing = true;
        rc = sweep_impl(sg->c, mesh, gf_mask, direct, MH_FP64, nomega, om.data(), g, rho, ndof, dof_fields, nbeta,
                        be.data(), A, B, F, nullptr);
        sg->c->capturing = false;
    }
    if (rc == MH_OK && !sg->c->cap_graph) {
        set_error("sweep graph: the capture recorded no graph");
        rc = MH_ERR_HIP;
    }
    if (rc == MH_OK && hipGraphInstantiate(&sg->exec, sg->c->cap_graph, nullptr, nullptr, 0) != hipSuccess) {
        set_error("sweep graph: hipGraphInstantiate failed");
        rc = MH_ERR_HIP;
    }
    if (rc == MH_OK &&
        hipHostMalloc(reinterpret_cast<void**>(&sg->pin),
                      sizeof(double) * (size_t)(2 * nomega + nbeta + 1) + sizeof(int) * (size_t)(nomega + 2)) !=
            hipSuccess) {
        set_error("sweep graph: hipHostMalloc failed");
        rc = MH_ERR_HIP;
    }
    if (rc != MH_OK) {
        graph_free(sg);
        return rc;
    }
    sg->d_om = sg->c->cap_om;
    sg->d_k = sg->c->cap_k;
    sg->d_be = sg->c->cap_be;
    sg->info = sg->c->cap_info;
    *out = reinterpret_cast<mh_sweep_graph_t>(sg);
    return MH_OK;
}

int mh_sweep_graph_launch(mh_context_t ctx, mh_sweep_graph_t graph, const double* omegas, const double* betas) {
    MH_REQUIRE(ctx != nullptr && graph != nullptr, "sweep graph: null context or graph");
    SweepGraph* sg = reinterpret_cast<SweepGraph*>(graph);
    Context* c = sg->c;
    Context* cc = reinterpret_cast<Context*>(ctx);
    MH_REQUIRE(cc->device == c->device, "sweep graph: context on another device");
    MH_HIP_CHECK(hipSetDevice(c->device));
    MH_REQUIRE(omegas != nullptr && (sg->nbeta == 0 || betas != nullptr), "sweep graph: null omegas or betas");
    for (int64_t w = 0; w < sg->nomega; ++w)
        MH_REQUIRE(omegas[w] > 0.0, "sweep graph: frequencies must be positive");
    // the previous launch has completed (every launch synchronises), so the
    // staging buffer is free
    double* p = sg->pin;
    for (int64_t w = 0; w < sg->nomega; ++w) {
        p[w] = omegas[w];
        p[sg->nomega + w] = omegas[w] * omegas[w] / sg->g;  // waves.jl:30
    }
    for (int64_t b = 0; b < sg->nbeta; ++b) p[2 * sg->nomega + b] = betas[b];
    TRY_RC(after_caller(sg, cc));
    MH_HIP_CHECK(hipMemcpyAsync(sg->d_om, p, sizeof(double) * sg->nomega, hipMemcpyHostToDevice, c->stream));
    MH_HIP_CHECK(
        hipMemcpyAsync(sg->d_k, p + sg->nomega, sizeof(double) * sg->nomega, hipMemcpyHostToDevice, c->stream));
    if (sg->nbeta)
        MH_HIP_CHECK(hipMemcpyAsync(sg->d_be, p + 2 * sg->nomega, sizeof(double) * sg->nbeta, hipMemcpyHostToDevice,
                                    c->stream));
    MH_HIP_CHECK(hipGraphLaunch(sg->exec, c->stream));
    // the checks the capture left out, read back together: the mesh-index
    // flag, the solve error word and the pivot info words
    int* words = reinterpret_cast<int*>(p + 2 * sg->nomega + sg->nbeta + 1);
    MH_HIP_CHECK(hipMemcpyAsync(words, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    words[1] = 0;
    if (c->tri_err)
        MH_HIP_CHECK(hipMemcpyAsync(words + 1, c->tri_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MH_HIP_CHECK(
        hipMemcpyAsync(words + 2, sg->info, sizeof(int) * sg->nomega, hipMemcpyDeviceToHost, c->stream));
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (words[0]) {
        set_error("mesh: a face references a vertex index outside [0, nvertices)");
        return MH_ERR_ARGUMENT;
    }
    if (words[1]) return tri_check(c);  // reports and clears the error word
    for (int64_t w = 0; w < sg->nomega; ++w)
        if (words[2 + w]) return check_info(c, sg->info, sg->nomega);  // reports the cause
    return MH_OK;
}

int mh_sweep_graph_destroy(mh_sweep_graph_t graph) {
    graph_free(reinterpret_cast<SweepGraph*>(graph));
    return MH_OK;
}

}  // extern "C"
