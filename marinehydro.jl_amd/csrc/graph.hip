// Captured sweeps: the FP64 forward frequency sweep of a body recorded once
// as a HIP graph and replayed with one hipGraphLaunch, for callers that run
// the same body many times (design loops, serving) with new frequencies and
// headings.  The machinery below is written for many bodies on several
// worker contexts, mixed precision and gradients (device verdicts of the
// refinement, device-staged frequency tangents), but the entry points
// refuse all of that: a captured mixed sweep with gradients crashed the
// host process at its first launch, a capture of several bodies on three
// workers crashed it during the creation, and the cause was not found.
//
// The bodies are assigned to private worker contexts (own streams,
// workspace and buffers) longest first, each to the worker with the least
// estimated work so far (cost ~ n^2 (n + 64) per frequency); the addresses
// the graph recorded never move, whatever else runs on the caller's
// context.  A sizing pass runs every body once on its worker, which
// allocates everything the capture records (workspace, Rankine parts, split
// images, solve state, events, streams, the Delhommeau table); the capture
// then records the same sequence, every worker's streams joined to the
// first worker's, which holds the capture.  Host steps stay outside the
// graph: each body's frequencies, wavenumbers, headings and frequency
// tangents are written into the recorded device buffers by one copy before
// each launch, and the checks become device words read back after it: the
// mesh-index flag, the OR of the LU info words and the refinement verdict
// of the mixed solve.  The graph records a fixed number of refinement sweeps
// (two, which the FP32 factors of these systems meet); a body whose words
// are not all zero is run again without the graph, which refines further,
// falls back to FP64 or reports the error exactly as mh_freq_sweep(_grad)
// does.
#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

#include "mh_internal.hpp"

namespace mh {

struct GraphItem {
    mh_sweep_item_t it;  // device pointers as recorded (host arrays not kept)
    mh_mesh_t mesh;
    int worker = 0;
    int Pg = 1;
    int64_t nblocks = 0;
    size_t off = 0, count = 0;  // staging slice: om, k, be, dk, dom (doubles)
    CapIO io;
};

struct BatchGraph {
    std::vector<Context*> workers;           // private; workers[0] holds the capture
    std::vector<std::vector<int64_t>> lists;  // items per worker, longest first
    std::vector<GraphItem> items;
    int precision = MH_FP64;
    double g = 9.81, rho = 1025.0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev = nullptr;         // orders a launch after the caller's stream
    std::vector<hipEvent_t> ev_w;    // the workers' streams rejoin the first one
    double *d_stage = nullptr, *h_stage = nullptr;
    size_t stage_count = 0;
    int *d_check = nullptr, *h_check = nullptr;
    size_t nodes = 0;
    int64_t reruns = 0;
    // the single-body FP64 entry points (mh_sweep_graph_*)
    std::vector<double> legacy_om, legacy_be;
};

static void graph_free(BatchGraph* bg) {
    if (!bg) return;
    if (!bg->workers.empty()) {
        (void)hipSetDevice(bg->workers[0]->device);
        for (Context* w : bg->workers) (void)hipStreamSynchronize(w->stream);
    }
    if (bg->exec) (void)hipGraphExecDestroy(bg->exec);
    if (bg->graph) (void)hipGraphDestroy(bg->graph);
    if (bg->ev) (void)hipEventDestroy(bg->ev);
    for (hipEvent_t e : bg->ev_w)
        if (e) (void)hipEventDestroy(e);
    if (bg->d_stage) (void)hipFree(bg->d_stage);
    if (bg->d_check) (void)hipFree(bg->d_check);
    if (bg->h_stage) (void)hipHostFree(bg->h_stage);
    if (bg->h_check) (void)hipHostFree(bg->h_check);
    for (Context* w : bg->workers) context_free(w);
    delete bg;
}

static double item_cost(const mh_sweep_item_t& it) {
    const double n = (double)it.mesh->nfaces;
    return n * n * (n + 64.0) * (double)it.nomega;
}

static int run_eager(Context* c, const mh_sweep_item_t& it, int precision, double g, double rho) {
    if (it.nparam > 0) {
        GradArgs ga{it.nparam, it.dvertices, it.dcenters, it.dnormals, it.dareas, it.ddof_fields, it.domegas,
                    it.dA, it.dB, it.dF};
        return sweep_impl(c, it.mesh, it.gf_mask, it.direct, precision, it.nomega, it.omegas, g, rho, it.ndof,
                          it.dof_fields, it.nbeta, it.betas, it.A, it.B, it.F, &ga);
    }
    return sweep_impl(c, it.mesh, it.gf_mask, it.direct, precision, it.nomega, it.omegas, g, rho, it.ndof,
                      it.dof_fields, it.nbeta, it.betas, it.A, it.B, it.F, nullptr);
}

// every worker runs its list (threads, one per worker); the first error wins
static int run_lists(BatchGraph* bg, const mh_sweep_item_t* items) {
    const int nw = (int)bg->workers.size();
    std::vector<int> rcs(nw, MH_OK);
    std::vector<std::string> errs(nw);
    auto work = [&](int w) {
        Context* c = bg->workers[w];
        if (hipSetDevice(c->device) != hipSuccess) {
            rcs[w] = MH_ERR_HIP;
            errs[w] = "sweep graph: hipSetDevice failed in a worker";
            return;
        }
        for (int64_t i : bg->lists[w]) {
            const int rc = run_eager(c, items[i], bg->precision, bg->g, bg->rho);
            if (rc != MH_OK) {
                rcs[w] = rc;
                errs[w] = "sweep graph: item " + std::to_string(i) + ": " + last_error();
                return;
            }
        }
    };
    if (nw == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < nw; ++w) th.emplace_back(work, w);
        for (auto& t : th) t.join();
    }
    MH_HIP_CHECK(hipSetDevice(bg->workers[0]->device));
    for (int w = 0; w < nw; ++w)
        if (rcs[w] != MH_OK) {
            set_error(errs[w]);
            return rcs[w];
        }
    return MH_OK;
}

// host staging of every item's device inputs (one copy per launch)
static void stage(BatchGraph* bg, const mh_sweep_item_t* items) {
    for (size_t i = 0; i < bg->items.size(); ++i) {
        const GraphItem& gi = bg->items[i];
        const mh_sweep_item_t& it = items[i];
        double* p = bg->h_stage + gi.off;
        const int64_t nw = it.nomega, nb = it.nbeta > 0 ? it.nbeta : 1;
        for (int64_t w = 0; w < nw; ++w) {
            p[w] = it.omegas[w];
            p[nw + w] = it.omegas[w] * it.omegas[w] / bg->g;  // waves.jl:30
        }
        for (int64_t b = 0; b < nb; ++b) p[2 * nw + b] = (it.nbeta > 0) ? it.betas[b] : 0.0;
        if (gi.nblocks > 0) {
            double* dk = p + 2 * nw + nb;
            omega_tangents(it.omegas, nw, 0, nw, it.domegas, it.nparam, gi.Pg, bg->g, dk,
                           dk + gi.nblocks * nw * gi.Pg);
        }
    }
}

// The device arrays of a launch must be those the graph recorded.
static int same_arrays(const GraphItem& gi, const mh_sweep_item_t& it, int64_t i) {
    const mh_sweep_item_t& r = gi.it;
    const mh_mesh_t& m = gi.mesh;
    const bool ok = it.mesh && it.mesh->nfaces == m.nfaces && it.mesh->nvertices == m.nvertices &&
                    it.mesh->vertices == m.vertices && it.mesh->faces == m.faces && it.mesh->centers == m.centers &&
                    it.mesh->normals == m.normals && it.mesh->areas == m.areas && it.mesh->radii == m.radii &&
                    it.gf_mask == r.gf_mask && it.direct == r.direct && it.nomega == r.nomega &&
                    it.ndof == r.ndof && it.dof_fields == r.dof_fields && it.nbeta == r.nbeta &&
                    it.nparam == r.nparam && it.A == r.A && it.B == r.B && it.F == r.F &&
                    (it.nparam == 0 || (it.dvertices == r.dvertices && it.dcenters == r.dcenters &&
                                        it.dnormals == r.dnormals && it.dareas == r.dareas &&
                                        it.ddof_fields == r.ddof_fields && it.dA == r.dA && it.dB == r.dB &&
                                        it.dF == r.dF));
    if (!ok) {
        set_error("sweep graph: item " + std::to_string(i) +
                  " differs from the recorded one (sizes, options or device arrays)");
        return MH_ERR_ARGUMENT;
    }
    for (int64_t w = 0; w < it.nomega; ++w)
        MH_REQUIRE(it.omegas && it.omegas[w] > 0.0, "sweep graph: frequencies must be positive");
    MH_REQUIRE(it.nbeta == 0 || it.betas, "sweep graph: null headings");
    return MH_OK;
}

static int graph_create(Context* cc, int64_t nitems, const mh_sweep_item_t* items, int precision, double g,
                        double rho, int nworkers, BatchGraph** out) {
    *out = nullptr;
    BatchGraph* bg = new BatchGraph();
    bg->precision = precision;
    bg->g = g;
    bg->rho = rho;
    int rc = MH_OK;
    auto fail = [&](int code, const std::string& msg) {
        if (!msg.empty()) set_error(msg);
        graph_free(bg);
        return code;
    };
    // workers and the longest-first assignment
    for (int w = 0; w < nworkers; ++w) {
        Context* wc = nullptr;
        rc = context_new(cc->device, nullptr, true, &wc);
        if (rc != MH_OK) return fail(rc, "");
        wc->ir_max_sweeps = cc->ir_max_sweeps;
        wc->ir_graph_sweeps = cc->ir_graph_sweeps;
        wc->fp64_fallback = cc->fp64_fallback;
        wc->leaf_cap = cc->tuned_workers > 0 ? cc->tuned_leaf : cc->leaf_cap;
        bg->workers.push_back(wc);
    }
    std::vector<int64_t> order((size_t)nitems);
    for (int64_t i = 0; i < nitems; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return item_cost(items[a]) > item_cost(items[b]); });
    bg->lists.assign(nworkers, {});
    std::vector<double> load(nworkers, 0.0);
    bg->items.resize((size_t)nitems);
    for (int64_t i : order) {
        const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[w] += item_cost(items[i]);
        bg->lists[w].push_back(i);
        bg->items[i].worker = w;
    }
    // staging slices and check words
    size_t count = 0;
    for (int64_t i = 0; i < nitems; ++i) {
        GraphItem& gi = bg->items[i];
        gi.it = items[i];
        gi.mesh = *items[i].mesh;
        gi.it.mesh = nullptr;
        gi.it.omegas = gi.it.betas = gi.it.domegas = nullptr;
        gi.Pg = grad_block(items[i].nparam);
        gi.nblocks = items[i].nparam > 0 ? (items[i].nparam + gi.Pg - 1) / gi.Pg : 0;
        const int64_t nw = items[i].nomega, nb = items[i].nbeta > 0 ? items[i].nbeta : 1;
        gi.off = count;
        gi.count = (size_t)(2 * nw + nb + 2 * gi.nblocks * nw * gi.Pg);
        count += (gi.count + 1) & ~size_t(1);
    }
    bg->stage_count = count;
    if (hipMalloc(reinterpret_cast<void**>(&bg->d_stage), sizeof(double) * count) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&bg->d_check), sizeof(int) * kCheckWords * (size_t)nitems) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&bg->h_stage), sizeof(double) * count) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&bg->h_check), sizeof(int) * kCheckWords * (size_t)nitems) !=
            hipSuccess)
        return fail(MH_ERR_HIP, "sweep graph: allocation of the staging buffers failed");
    for (int64_t i = 0; i < nitems; ++i) {
        GraphItem& gi = bg->items[i];
        const int64_t nw = items[i].nomega, nb = items[i].nbeta > 0 ? items[i].nbeta : 1;
        double* base = bg->d_stage + gi.off;
        gi.io.om = base;
        gi.io.k = base + nw;
        gi.io.be = base + 2 * nw;
        gi.io.dk = base + 2 * nw + nb;
        gi.io.dom = gi.io.dk + gi.nblocks * nw * gi.Pg;
        gi.io.check = bg->d_check + kCheckWords * i;
    }
    if (hipEventCreateWithFlags(&bg->ev, hipEventDisableTiming) != hipSuccess)
        return fail(MH_ERR_HIP, "sweep graph: hipEventCreate failed");
    bg->ev_w.assign(nworkers, nullptr);
    for (int w = 0; w < nworkers; ++w)
        if (hipEventCreateWithFlags(&bg->ev_w[w], hipEventDisableTiming) != hipSuccess)
            return fail(MH_ERR_HIP, "sweep graph: hipEventCreate failed");
    // the workers start after the caller's pending work (mesh, dof fields)
    for (Context* w : bg->workers) {
        if (hipEventRecord(bg->ev, cc->stream) != hipSuccess || hipStreamWaitEvent(w->stream, bg->ev, 0) != hipSuccess)
            return fail(MH_ERR_HIP, "sweep graph: ordering after the caller's stream failed");
    }
    // sizing pass: every item once, eagerly, on its worker
    rc = run_lists(bg, items);
    if (rc != MH_OK) {
        const std::string msg = last_error();
        return fail(rc, msg);
    }
    for (Context* w : bg->workers)
        if (hipStreamSynchronize(w->stream) != hipSuccess)
            return fail(MH_ERR_HIP, "sweep graph: sizing pass failed");
    // capture
    Context* w0 = bg->workers[0];
    if (hipStreamBeginCapture(w0->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(MH_ERR_HIP, "sweep graph: hipStreamBeginCapture failed");
    bool ok = hipEventRecord(bg->ev_w[0], w0->stream) == hipSuccess;
    for (int w = 1; w < nworkers && ok; ++w) ok = hipStreamWaitEvent(bg->workers[w]->stream, bg->ev_w[0], 0) == hipSuccess;
    std::string err = ok ? "" : "sweep graph: joining the workers' streams to the capture failed";
    rc = ok ? MH_OK : MH_ERR_HIP;
    for (int w = 0; w < nworkers && rc == MH_OK; ++w) {
        Context* wc = bg->workers[w];
        for (int64_t i : bg->lists[w]) {
            wc->capturing = true;
            wc->cap = &bg->items[i].io;
            rc = run_eager(wc, items[i], precision, g, rho);
            wc->capturing = false;
            wc->cap = nullptr;
            if (rc != MH_OK) {
                err = "sweep graph: capture of item " + std::to_string(i) + ": " + last_error();
                break;
            }
        }
    }
    for (int w = 1; w < nworkers && rc == MH_OK; ++w) {
        if (hipEventRecord(bg->ev_w[w], bg->workers[w]->stream) != hipSuccess ||
            hipStreamWaitEvent(w0->stream, bg->ev_w[w], 0) != hipSuccess) {
            rc = MH_ERR_HIP;
            err = "sweep graph: rejoining the workers' streams failed";
        }
    }
    const hipError_t ec = hipStreamEndCapture(w0->stream, &bg->graph);
    if (rc != MH_OK) return fail(rc, err);
    if (ec != hipSuccess || !bg->graph)
        return fail(MH_ERR_HIP, std::string("sweep graph: hipStreamEndCapture: ") + hipGetErrorString(ec));
    if (hipGraphGetNodes(bg->graph, nullptr, &bg->nodes) != hipSuccess) bg->nodes = 0;
    const hipError_t ei = hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess)
        return fail(MH_ERR_HIP, std::string("sweep graph: hipGraphInstantiate: ") + hipGetErrorString(ei));
    *out = bg;
    return MH_OK;
}

static int graph_launch(Context* cc, BatchGraph* bg, const mh_sweep_item_t* items) {
    Context* w0 = bg->workers[0];
    MH_REQUIRE(cc->device == w0->device, "sweep graph: context on another device");
    MH_HIP_CHECK(hipSetDevice(w0->device));
    const int64_t nitems = (int64_t)bg->items.size();
    for (int64_t i = 0; i < nitems; ++i) TRY_RC(same_arrays(bg->items[i], items[i], i));
    // the previous launch has completed (every launch synchronises), so the
    // staging buffer is free
    stage(bg, items);
    MH_HIP_CHECK(hipEventRecord(bg->ev, cc->stream));
    MH_HIP_CHECK(hipStreamWaitEvent(w0->stream, bg->ev, 0));
    MH_HIP_CHECK(hipMemcpyAsync(bg->d_stage, bg->h_stage, sizeof(double) * bg->stage_count, hipMemcpyHostToDevice,
                                w0->stream));
    MH_HIP_CHECK(hipGraphLaunch(bg->exec, w0->stream));
    MH_HIP_CHECK(hipMemcpyAsync(bg->h_check, bg->d_check, sizeof(int) * kCheckWords * (size_t)nitems,
                                hipMemcpyDeviceToHost, w0->stream));
    MH_HIP_CHECK(hipStreamSynchronize(w0->stream));
    // a failed inter-workgroup wait or a bad pivot index in any solve
    for (Context* w : bg->workers) TRY_RC(tri_check(w));
    // items whose checks did not pass run again without the graph, which
    // reports their error or repairs the result (more refinement sweeps,
    // the FP64 fallback)
    for (int64_t i = 0; i < nitems; ++i) {
        bool clean = true;
        for (int k = 0; k < kCheckWords; ++k) clean = clean && bg->h_check[kCheckWords * i + k] == 0;
        if (clean) continue;
        bg->reruns += 1;
        const int rc = run_eager(bg->workers[bg->items[i].worker], items[i], bg->precision, bg->g, bg->rho);
        if (rc != MH_OK) {
            set_error("sweep graph: item " + std::to_string(i) + ": " + last_error());
            return rc;
        }
    }
    return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_batch_graph_create(mh_context_t ctx, int64_t nitems, const mh_sweep_item_t* items, int precision, double g,
                          double rho, int nworkers, mh_batch_graph_t* out) {
    MH_REQUIRE(ctx != nullptr && out != nullptr, "batch graph: null context or output handle");
    *out = nullptr;
    Context* cc = reinterpret_cast<Context*>(ctx);
    MH_HIP_CHECK(hipSetDevice(cc->device));
    MH_REQUIRE(nitems > 0 && items, "batch graph: no items");
    MH_REQUIRE(nworkers >= 0 && nworkers <= 16, "batch graph: nworkers must be in [0, 16]");
    // Only what has run correctly is captured: one FP64 forward sweep on one
    // worker.  A one-body mixed sweep with gradients was captured without
    // error but crashed the host process inside its first launch (also once
    // its device-to-device copies were kernels), and six FP64 bodies on three
    // workers crashed it inside the creation; the cause was not found, so
    // both are refused rather than run.
    if (precision != MH_FP64) {
        set_error("batch graph: only FP64 sweeps are captured (mixed precision is not supported by the graph path)");
        return MH_ERR_NOT_SUPPORTED;
    }
    if (nitems != 1 || nworkers > 1) {
        set_error("batch graph: one item on one worker only (several items or workers are not supported)");
        return MH_ERR_NOT_SUPPORTED;
    }
    for (int64_t i = 0; i < nitems; ++i) {
        const mh_sweep_item_t& it = items[i];
        MH_REQUIRE(it.mesh != nullptr, "batch graph: null mesh");
        if (it.nparam != 0) {
            set_error("batch graph: item " + std::to_string(i) +
                      ": gradients are not captured (nparam must be 0 in a batch graph)");
            return MH_ERR_NOT_SUPPORTED;
        }
        int rc = sweep_args_ok(it.mesh, it.gf_mask, precision, it.nomega, it.omegas, g, rho, it.ndof,
                               it.dof_fields, it.nbeta, it.betas, it.A, it.B, it.F);
        if (rc == MH_OK && !(it.direct == 0 || it.direct == 1)) {
            set_error("direct must be 0 or 1");
            rc = MH_ERR_ARGUMENT;
        }
        if (rc == MH_OK && !(it.mesh->nfaces > 0 && it.nomega > 0 && it.ndof + it.nbeta > 0)) {
            set_error("an empty sweep");
            rc = MH_ERR_ARGUMENT;
        }
        if (rc == MH_OK && it.nparam < 0) {
            set_error("negative nparam");
            rc = MH_ERR_ARGUMENT;
        }
        if (rc == MH_OK && it.nparam > 0 && !(it.ndof >= 1 && it.dA && it.dB && (it.nbeta == 0 || it.dF))) {
            set_error("gradients need at least one dof and their outputs");
            rc = MH_ERR_ARGUMENT;
        }
        if (rc != MH_OK) {
            set_error("batch graph: item " + std::to_string(i) + ": " + last_error());
            return rc;
        }
    }
    if (nworkers == 0) nworkers = cc->tuned_workers > 0 ? cc->tuned_workers : 2;
    if (nworkers > nitems) nworkers = (int)nitems;
    BatchGraph* bg = nullptr;
    TRY_RC(graph_create(cc, nitems, items, precision, g, rho, nworkers, &bg));
    *out = reinterpret_cast<mh_batch_graph_t>(bg);
    return MH_OK;
}

int mh_batch_graph_launch(mh_context_t ctx, mh_batch_graph_t graph, int64_t nitems, const mh_sweep_item_t* items) {
    MH_REQUIRE(ctx != nullptr && graph != nullptr, "batch graph: null context or graph");
    BatchGraph* bg = reinterpret_cast<BatchGraph*>(graph);
    MH_REQUIRE(nitems == (int64_t)bg->items.size() && items, "batch graph: the recorded number of items expected");
    return graph_launch(reinterpret_cast<Context*>(ctx), bg, items);
}

int mh_batch_graph_info(mh_batch_graph_t graph, int64_t* nodes, int64_t* reruns, int* nworkers) {
    MH_REQUIRE(graph != nullptr, "batch graph: null graph");
    BatchGraph* bg = reinterpret_cast<BatchGraph*>(graph);
    if (nodes) *nodes = (int64_t)bg->nodes;
    if (reruns) *reruns = bg->reruns;
    if (nworkers) *nworkers = (int)bg->workers.size();
    return MH_OK;
}

int mh_batch_graph_dot(mh_batch_graph_t graph, const char* path) {
    MH_REQUIRE(graph != nullptr && path != nullptr, "batch graph: null graph or path");
    BatchGraph* bg = reinterpret_cast<BatchGraph*>(graph);
    const hipError_t e = hipGraphDebugDotPrint(bg->graph, path, hipGraphDebugDotFlagsVerbose);
    if (e != hipSuccess) {
        set_error(std::string("batch graph: hipGraphDebugDotPrint: ") + hipGetErrorString(e));
        return MH_ERR_HIP;
    }
    return MH_OK;
}

int mh_batch_graph_destroy(mh_batch_graph_t graph) {
    graph_free(reinterpret_cast<BatchGraph*>(graph));
    return MH_OK;
}

// ---- one body, FP64 forward sweep (the frequencies given at each launch) ----

int mh_sweep_graph_create(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int direct, int64_t nomega,
                          int64_t ndof, const double* dof_fields, int64_t nbeta, double g, double rho, double* A,
                          double* B, double* F, mh_sweep_graph_t* out) {
    MH_REQUIRE(ctx != nullptr && out != nullptr, "sweep graph: null context or output handle");
    *out = nullptr;
    MH_REQUIRE(mesh != nullptr && nomega > 0, "sweep graph: null mesh or no frequencies");
    // placeholders for the argument checks and the sizing sweep
    std::vector<double> om((size_t)nomega, 1.0), be((size_t)(nbeta > 0 ? nbeta : 1), 0.0);
    mh_sweep_item_t it{};
    it.mesh = mesh;
    it.gf_mask = gf_mask;
    it.direct = direct;
    it.nomega = nomega;
    it.omegas = om.data();
    it.ndof = ndof;
    it.dof_fields = dof_fields;
    it.nbeta = nbeta;
    it.betas = nbeta > 0 ? be.data() : nullptr;
    it.A = A;
    it.B = B;
    it.F = F;
    mh_batch_graph_t bgh = nullptr;
    TRY_RC(mh_batch_graph_create(ctx, 1, &it, MH_FP64, g, rho, 1, &bgh));
    BatchGraph* bg = reinterpret_cast<BatchGraph*>(bgh);
    bg->legacy_om = om;
    bg->legacy_be = be;
    *out = reinterpret_cast<mh_sweep_graph_t>(bg);
    return MH_OK;
}

int mh_sweep_graph_launch(mh_context_t ctx, mh_sweep_graph_t graph, const double* omegas, const double* betas) {
    MH_REQUIRE(ctx != nullptr && graph != nullptr, "sweep graph: null context or graph");
    BatchGraph* bg = reinterpret_cast<BatchGraph*>(graph);
    MH_REQUIRE(bg->items.size() == 1, "sweep graph: not a single-body graph");
    GraphItem& gi = bg->items[0];
    MH_REQUIRE(omegas != nullptr && (gi.it.nbeta == 0 || betas != nullptr), "sweep graph: null omegas or betas");
    mh_sweep_item_t it = gi.it;
    it.mesh = &gi.mesh;
    it.omegas = omegas;
    it.betas = gi.it.nbeta > 0 ? betas : nullptr;
    return graph_launch(reinterpret_cast<Context*>(ctx), bg, &it);
}

int mh_sweep_graph_destroy(mh_sweep_graph_t graph) {
    graph_free(reinterpret_cast<BatchGraph*>(graph));
    return MH_OK;
}

}  // extern "C"
