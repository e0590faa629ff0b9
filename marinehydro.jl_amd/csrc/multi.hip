// Uneven multi-body batches: many (body, frequency list) sweeps of different
// panel counts in one call.
//
// The reference solves its problems one after another (src/solve.jl:17-21,
// src/problems_and_results.jl:70-100).  Here the bodies form a work list
// ordered longest first (cost ~ n^3 per frequency); a pool of worker
// contexts, each with a stream of its own and its own workspace, takes
// bodies from the list with an atomic counter, so small bodies run beside
// the large ones instead of leaving the GPU idle between launches, and the
// last bodies to finish are the cheapest.  Every body is one batched sweep
// over its frequencies (assembly, LU, solves, forces and, when asked, the
// adjoint gradients).
//
// The autotuner times the list (or a sample of it) under each candidate
// configuration -- worker count and LU leaf width (the column width of the
// register-resident panel factorisation) -- and keeps the fastest for the
// context.
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mh_internal.hpp"

namespace mh {

static int run_item(Context* c, const mh_sweep_item_t& it, int precision, double g, double rho) {
    if (it.nparam > 0) {
        GradArgs ga{it.nparam, it.dvertices, it.dcenters, it.dnormals, it.dareas, it.ddof_fields, it.domegas,
                    it.dA, it.dB, it.dF};
        return sweep_impl(c, it.mesh, it.gf_mask, it.direct, precision, it.nomega, it.omegas, g, rho, it.ndof,
                          it.dof_fields, it.nbeta, it.betas, it.A, it.B, it.F, &ga);
    }
    return sweep_impl(c, it.mesh, it.gf_mask, it.direct, precision, it.nomega, it.omegas, g, rho, it.ndof,
                      it.dof_fields, it.nbeta, it.betas, it.A, it.B, it.F, nullptr);
}

static int item_ok(const mh_sweep_item_t& it, int precision, double g, double rho) {
    MH_REQUIRE(it.mesh != nullptr, "multi_sweep: null mesh");
    int rc = sweep_args_ok(it.mesh, it.gf_mask, precision, it.nomega, it.omegas, g, rho, it.ndof, it.dof_fields,
                           it.nbeta, it.betas, it.A, it.B, it.F);
    if (rc != MH_OK) return rc;
    MH_REQUIRE(it.direct == 0 || it.direct == 1, "multi_sweep: direct must be 0 or 1");
    MH_REQUIRE(it.nparam >= 0, "multi_sweep: negative nparam");
    if (it.nparam > 0 && it.mesh->nfaces > 0 && it.nomega > 0) {
        MH_REQUIRE(it.ndof >= 1, "multi_sweep: gradients need at least one dof");
        MH_REQUIRE(it.dA && it.dB && (it.nbeta == 0 || it.dF), "multi_sweep: null gradient output");
    }
    return MH_OK;
}

static int ensure_workers(Context* c, int nworkers) {
    while ((int)c->workers.size() < nworkers) {
        Context* w = nullptr;
        TRY_RC(context_new(c->device, nullptr, true, &w));
        c->workers.push_back(w);
    }
    for (Context* w : c->workers) {
        w->ir_max_sweeps = c->ir_max_sweeps;
        w->fp64_fallback = c->fp64_fallback;
        w->leaf_cap = c->leaf_cap;
    }
    return MH_OK;
}

static int multi_run(Context* c, int64_t nitems, const mh_sweep_item_t* items, int precision, double g,
                     double rho, int nworkers) {
    TRY_RC(ensure_workers(c, nworkers));
    // the caller's stream may still be producing the inputs
    MH_HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<int64_t> order((size_t)nitems);
    for (int64_t i = 0; i < nitems; ++i) order[i] = i;
    auto cost = [&](int64_t i) {
        const double n = (double)items[i].mesh->nfaces;
        return n * n * (n + 64.0) * (double)items[i].nomega;
    };
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cost(a) > cost(b); });
    std::atomic<int64_t> next(0);
    std::atomic<bool> stop(false);
    std::mutex mu;
    std::string first_err;
    int first_rc = MH_OK;
    std::vector<int64_t> fb0(nworkers);
    for (int w = 0; w < nworkers; ++w) fb0[w] = c->workers[w]->fallbacks;
    auto worker = [&](int w) {
        Context* wc = c->workers[w];
        if (hipSetDevice(c->device) != hipSuccess) {
            std::lock_guard<std::mutex> lock(mu);
            if (first_rc == MH_OK) {
                first_rc = MH_ERR_HIP;
                first_err = "multi_sweep: hipSetDevice failed in a worker";
            }
            stop = true;
            return;
        }
        while (!stop) {
            const int64_t k = next.fetch_add(1);
            if (k >= nitems) break;
            const int64_t i = order[k];
            const int rc = run_item(wc, items[i], precision, g, rho);
            if (rc != MH_OK) {
                std::lock_guard<std::mutex> lock(mu);
                if (first_rc == MH_OK) {
                    first_rc = rc;
                    first_err = "multi_sweep: item " + std::to_string(i) + ": " + last_error();
                }
                stop = true;
            }
        }
    };
    if (nworkers == 1) {
        worker(0);
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < nworkers; ++w) th.emplace_back(worker, w);
        for (auto& t : th) t.join();
    }
    for (int w = 0; w < nworkers; ++w) c->fallbacks += c->workers[w]->fallbacks - fb0[w];
    MH_HIP_CHECK(hipSetDevice(c->device));
    if (first_rc != MH_OK) {
        set_error(first_err);
        return first_rc;
    }
    return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_multi_sweep(mh_context_t ctx, int64_t nitems, const mh_sweep_item_t* items, int precision, double g,
                   double rho, int nworkers) {
    MH_REQUIRE(ctx != nullptr, "null context");
    Context* c = reinterpret_cast<Context*>(ctx);
    MH_HIP_CHECK(hipSetDevice(c->device));
    MH_REQUIRE(nitems >= 0, "multi_sweep: negative item count");
    MH_REQUIRE(nitems == 0 || items, "multi_sweep: null item array");
    MH_REQUIRE(nworkers >= 0 && nworkers <= 16, "multi_sweep: nworkers must be in [0, 16]");
    for (int64_t i = 0; i < nitems; ++i) {
        const int rc = item_ok(items[i], precision, g, rho);
        if (rc != MH_OK) {
            set_error("multi_sweep: item " + std::to_string(i) + ": " + last_error());
            return rc;
        }
    }
    if (nitems == 0) return MH_OK;
    if (nworkers == 0) {
        nworkers = c->tuned_workers > 0 ? c->tuned_workers : 2;
        if (c->tuned_workers > 0) c->leaf_cap = c->tuned_leaf;
    }
    if (nworkers > nitems) nworkers = (int)nitems;
    return multi_run(c, nitems, items, precision, g, rho, nworkers);
}

int mh_multi_autotune(mh_context_t ctx, int64_t nitems, const mh_sweep_item_t* items, int precision, double g,
                      double rho, int* best_workers, int* best_leaf, double* best_ms) {
    MH_REQUIRE(ctx != nullptr, "null context");
    Context* c = reinterpret_cast<Context*>(ctx);
    MH_HIP_CHECK(hipSetDevice(c->device));
    MH_REQUIRE(nitems > 0 && items, "multi_autotune: empty sample");
    for (int64_t i = 0; i < nitems; ++i) {
        const int rc = item_ok(items[i], precision, g, rho);
        if (rc != MH_OK) {
            set_error("multi_autotune: item " + std::to_string(i) + ": " + last_error());
            return rc;
        }
    }
    const int cand_workers[] = {1, 2, 3, 4, 6};
    const int cand_leaf[] = {0, 8};
    const int saved_leaf = c->leaf_cap;
    // first-touch costs (workspace growth in every worker) outside the timings
    c->leaf_cap = 0;
    TRY_RC(multi_run(c, nitems, items, precision, g, rho, (int)std::min<int64_t>(6, nitems)));
    double best = 1e300;
    int bw = 1, bl = 0;
    for (int lf : cand_leaf) {
        for (int nw : cand_workers) {
            if (nw > nitems) continue;
            c->leaf_cap = lf;
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = multi_run(c, nitems, items, precision, g, rho, nw);
            const double ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (rc != MH_OK) {
                c->leaf_cap = saved_leaf;
                return rc;
            }
            if (ms < best) {
                best = ms;
                bw = nw;
                bl = lf;
            }
        }
    }
    c->tuned_workers = bw;
    c->tuned_leaf = bl;
    c->leaf_cap = bl;
    if (best_workers) *best_workers = bw;
    if (best_leaf) *best_leaf = bl;
    if (best_ms) *best_ms = best;
    return MH_OK;
}

}  // extern "C"
