// wl_cg_solve_noalloc: the single-rhs FP64 CG of wl_cg_solve (conjugate_gradient.h:46-104)
// with all of its device work memory -- the Krylov fields r, p, Dp, DDp, the odd temporary of
// M, the reduction partials, the CG state and the grid barrier -- in one fixed device array,
// `scratch`, instead of the context's work fields.  The call allocates nothing (no hipMalloc,
// hipFree or allocating library routine), for callers that cannot allocate device memory
// (tight solve loops, harnesses that forbid allocation during a run).
//
// The array is sized for lattices of up to kNaMaxStride sites per parity (16^4, or any
// lattice whose parity block rounds up to at most 2^15 sites): 5 FP64 fields of 6 MiB plus
// 8 KiB of partials and state, 30 MiB in all.  Every device the library runs on holds one copy
// of it, from the first wl_create on that device to the end of the process.  Calls on one
// device share that copy: they are serialised by a per-device lock within a process, and the
// copy keeps the last call's contents (the CG state and the barrier are re-initialised by every
// call; the fields are fully written before they are read).
#include <cmath>
#include <mutex>

#include "wl_host.h"

namespace wl {

namespace {

constexpr int kNaMaxStride = 1 << 15;                          // block stride, sites
constexpr size_t kNaField = (size_t)12 * kNaMaxStride * 16;    // one FP64 field, one rhs
constexpr size_t kNaPartials = 1024;                           // doubles
constexpr size_t kNaSmall = 256;                               // CG state, barrier, reduction out
constexpr size_t kNaBytes = 5 * kNaField + kNaPartials * sizeof(double) + 3 * kNaSmall;
static_assert(sizeof(CgState) <= kNaSmall, "CG state does not fit its workspace slot");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kNaMaxDevices = 64;
std::mutex g_setup_mu;
void *g_ptr[kNaMaxDevices] = {};          // device address of `scratch`, per device
std::mutex g_call_mu[kNaMaxDevices];      // one call at a time per device

}  // namespace

// Looks up the device address of `scratch` on the current device (once per device): wl_create
// calls it with the context's device current.  Host code never passes `scratch` itself.
int noalloc_setup(int device) {
    if (device < 0 || device >= kNaMaxDevices) return WL_OK;  // the entry point refuses such devices
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (g_ptr[device]) return WL_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(WL_E_DEVICE, "hipGetSymbolAddress of the fixed CG workspace on device %d: %s", device,
                    hipGetErrorString(e));
    g_ptr[device] = p;
    return WL_OK;
}

}  // namespace wl

using namespace wl;

extern "C" int wl_cg_solve_noalloc(wl_ctx *c, wl_field *x, const wl_field *b, double kappa, double accuracy,
                                   int maxiters, wl_cg_stats *stats) {
    if (!c || !x || !b) return fail(WL_E_ARG, "null argument");
    if (x->ctx != c || b->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (x->parity != WL_EVEN || b->parity != WL_EVEN) return fail(WL_E_ARG, "CG acts on even fields");
    if (x == b) return fail(WL_E_ARG, "x and b must be different fields");
    if (x->prec != WL_F64 || b->prec != WL_F64) return fail(WL_E_ARG, "x and b must be FP64 fields");
    if (x->nrhs != 1 || b->nrhs != 1) return fail(WL_E_ARG, "wl_cg_solve_noalloc solves one right-hand side");
    if (!(accuracy > 0)) return fail(WL_E_ARG, "accuracy must be > 0");
    if (!std::isfinite(kappa)) return fail(WL_E_ARG, "kappa must be finite");
    if (c->g.stride > kNaMaxStride)
        return fail(WL_E_ARG,
                    "wl_cg_solve_noalloc: %d sites per parity (block stride %d) exceed the fixed workspace "
                    "(block stride at most %d, e.g. 16^4)",
                    c->g.vh, c->g.stride, kNaMaxStride);
    if (cg_partials_needed(c) > kNaPartials)
        return fail(WL_E_ARG, "wl_cg_solve_noalloc: %zu partial sums exceed the fixed workspace's %zu",
                    cg_partials_needed(c), kNaPartials);
    if (c->device < 0 || c->device >= kNaMaxDevices || !g_ptr[c->device])
        return fail(WL_E_STATE, "the fixed CG workspace is not set up on device %d", c->device);
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->gauge_ok) return fail(WL_E_STATE, "no gauge field uploaded (wl_gauge_upload)");
    if (link_format(c, WL_F64) == 6 && !c->rows64_ok)
        return fail(WL_E_STATE,
                    "two-row links are on but their FP64 copy is not prepared: make one wl_meo_apply or "
                    "wl_cg_solve call first (wl_cg_solve_noalloc does not allocate it)");
    DeviceGuard dg(c->device);
    std::lock_guard<std::mutex> dlk(g_call_mu[c->device]);
    unsigned char *base = (unsigned char *)g_ptr[c->device];
    CgWork w;
    w.r = base;
    w.p = base + kNaField;
    w.Dp = base + 2 * kNaField;
    w.DDp = base + 3 * kNaField;
    w.t = base + 4 * kNaField;
    w.partials = (double *)(base + 5 * kNaField);
    w.partials_cap = kNaPartials;
    w.st = (CgState *)(base + 5 * kNaField + kNaPartials * sizeof(double));
    w.bar = base + 5 * kNaField + kNaPartials * sizeof(double) + kNaSmall;
    w.red_out = (double *)(base + 5 * kNaField + kNaPartials * sizeof(double) + 2 * kNaSmall);
    return cg_f64_run(c, w, x->data, x->bytes, b->data, kappa, accuracy, maxiters, stats);
}
