// mh_assemble_noalloc: influence-matrix assembly that allocates nothing.
//
// mh_assemble stages the packed panel records and the wavenumbers in the
// context workspace (grown with hipMalloc on demand) and keeps the
// frequency-independent Rankine parts in a context buffer of 24 n^2 bytes.
// This variant stages the panel records, the wavenumbers and the mesh-index
// error flag in one file-scope device array of fixed size instead, and runs
// the one-pass assembly kernel, which needs no scratch matrices.  The caller
// still passes the mesh, S and D/K.  Each device that runs it holds its own
// copy of the array until the process ends; every call on a device shares
// it, so calls that overlap on one device (two host threads, or two streams
// without ordering between them) race on it.
#include <mutex>

#include "mh_internal.hpp"

namespace mh {

constexpr int64_t kNoAllocMaxPanels = 4096;
constexpr int64_t kNoAllocMaxWavenumbers = 256;

// layout of the array: [flag: 256 B][wavenumbers][panel records]
constexpr size_t kNaFlagOff = 0;
constexpr size_t kNaKOff = 256;
constexpr size_t kNaPanelOff = kNaKOff + sizeof(double) * kNoAllocMaxWavenumbers;
constexpr size_t kNaBytes = kNaPanelOff + sizeof(PanelRec) * kNoAllocMaxPanels;
static_assert(kNaPanelOff % 256 == 0, "panel records stay 256-byte aligned");

static __device__ __align__(256) char scratch[kNaBytes];

// Device address of `scratch` per device (the host-side name is only a
// stand-in for the symbol and must never reach a kernel or a copy).
constexpr int kMaxDevices = 64;
static std::mutex g_na_mutex;
static char* g_na_ptr[kMaxDevices] = {};

static int scratch_on_device(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("assemble_noalloc: device index beyond the supported range");
        return MH_ERR_ARGUMENT;
    }
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[device]) {
        // the caller has made `device` current (check of the context)
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("assemble_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return MH_ERR_HIP;
        }
        g_na_ptr[device] = static_cast<char*>(p);
    }
    *out = g_na_ptr[device];
    return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" int mh_assemble_noalloc(mh_context_t ctx, const mh_mesh_t* mesh, int gf_mask, int formulation,
                                   int64_t nk, const double* wavenumbers, double* S, double* D, double* K,
                                   int64_t ld, int64_t stride) {
    MH_REQUIRE(ctx != nullptr, "null context");
    Context* c = reinterpret_cast<Context*>(ctx);
    MH_HIP_CHECK(hipSetDevice(c->device));
    MH_REQUIRE(mesh != nullptr, "mesh: null");
    MH_REQUIRE(mesh->nfaces >= 0 && mesh->nvertices >= 0, "mesh: negative size");
    if (mesh->nfaces > 0)
        MH_REQUIRE(mesh->vertices && mesh->faces && mesh->centers && mesh->normals && mesh->areas &&
                       mesh->radii,
                   "mesh: null array");
    MH_REQUIRE(gf_mask > 0 && gf_mask < 32, "assemble_noalloc: bad Green-function mask (the tabulated "
                                            "Delhommeau term allocates its table and is not available here)");
    MH_REQUIRE(!((gf_mask & MH_GF_WU) && (gf_mask & MH_GF_DELHOMMEAU)), "assemble_noalloc: choose one wave term");
    MH_REQUIRE(formulation >= 1 && formulation <= 3, "assemble_noalloc: formulation must be 1, 2 or 3");
    MH_REQUIRE(nk >= 0, "assemble_noalloc: negative nk");
    const int64_t n = mesh->nfaces;
    if (n == 0 || nk == 0) return MH_OK;
    if (n > kNoAllocMaxPanels) {
        set_error("assemble_noalloc: " + std::to_string(n) + " panels exceed the fixed scratch limit of " +
                  std::to_string(kNoAllocMaxPanels) + "; use mh_assemble");
        return MH_ERR_ARGUMENT;
    }
    if (nk > kNoAllocMaxWavenumbers) {
        set_error("assemble_noalloc: " + std::to_string(nk) + " wavenumbers exceed the fixed scratch limit of " +
                  std::to_string(kNoAllocMaxWavenumbers) + " per call; split the call or use mh_assemble");
        return MH_ERR_ARGUMENT;
    }
    MH_REQUIRE(wavenumbers && S, "assemble_noalloc: null array");
    MH_REQUIRE(!(formulation & 1) || D, "assemble_noalloc: D required for the direct formulation");
    MH_REQUIRE(!(formulation & 2) || K, "assemble_noalloc: K required for the indirect formulation");
    MH_REQUIRE(ld >= n && (nk == 1 || stride >= ld * n), "assemble_noalloc: ld/stride too small");
    char* base = nullptr;
    TRY_RC(scratch_on_device(c->device, &base));
    int* flag = reinterpret_cast<int*>(base + kNaFlagOff);
    double* d_k = reinterpret_cast<double*>(base + kNaKOff);
    PanelRec* panels = reinterpret_cast<PanelRec*>(base + kNaPanelOff);
    // the wavenumbers overwrite the previous call's: stream order keeps that
    // call's kernels ahead of this copy
    MH_HIP_CHECK(hipMemcpyAsync(d_k, wavenumbers, sizeof(double) * nk, hipMemcpyHostToDevice, c->stream));
    TRY_RC(launch_pack_panels(c, mesh, panels, flag));
    return launch_assemble_onepass(c, mesh, panels, gf_mask, formulation, nk, d_k, reinterpret_cast<double2*>(S),
                                   reinterpret_cast<double2*>(D), reinterpret_cast<double2*>(K), ld, stride);
}
