// C ABI of libhila_wilson.so: contexts, device fields, the even-odd Wilson operator and CG.
// Declarations and conventions: include/wilson_mi355x.h.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "wl_internal.h"

using namespace wl;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(WL_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

// Make the context's device current for the calling thread.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

int ensure_staging(wl_ctx *c, size_t bytes) {
    if (c->staging_bytes >= bytes) return WL_OK;
    if (c->staging) {
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->staging));
        c->staging = nullptr;
        c->staging_bytes = 0;
    }
    if (hipMalloc(&c->staging, bytes) != hipSuccess)
        return fail(WL_E_OOM, "staging buffer of %zu bytes", bytes);
    c->staging_bytes = bytes;
    return WL_OK;
}

int field_new(wl_ctx *c, int parity, int prec, int nrhs, wl_field **out) {
    wl_field *f = new (std::nothrow) wl_field;
    if (!f) return fail(WL_E_OOM, "host allocation");
    f->ctx = c;
    f->parity = parity;
    f->prec = prec;
    f->nrhs = nrhs;
    f->rhs_elems = (size_t)12 * c->g.stride;
    f->bytes = f->rhs_elems * nrhs * prec_bytes(prec);
    if (f->rhs_elems * (size_t)nrhs >= ((size_t)1 << 31)) {
        delete f;
        return fail(WL_E_ARG, "%d right-hand sides exceed 2^31 elements per field at this volume",
                    nrhs);
    }
    if (hipMalloc(&f->data, f->bytes) != hipSuccess) {
        const size_t want = f->bytes;
        delete f;
        return fail(WL_E_OOM, "device field of %zu bytes", want);
    }
    if (hipMemsetAsync(f->data, 0, f->bytes, c->stream) != hipSuccess) {
        (void)hipFree(f->data);
        delete f;
        return fail(WL_E_DEVICE, "hipMemsetAsync of a new field");
    }
    *out = f;
    return WL_OK;
}

void field_free(wl_field *f) {
    if (!f) return;
    if (f->data) (void)hipFree(f->data);
    delete f;
}

int ensure_tmp(wl_ctx *c, int prec, wl_field **slot, int parity) {
    if (*slot) return WL_OK;
    return field_new(c, parity, prec, 1, slot);
}

RedArgs red_args(wl_ctx *c) {
    RedArgs ra;
    ra.partials = c->partials;
    ra.st = c->st;
    ra.out = c->scratch_d;
    return ra;
}

const void *gauge_for(wl_ctx *c, int prec) { return prec == WL_F64 ? c->gauge64 : c->gauge32; }

int ensure_gauge(wl_ctx *c, int prec) {
    if (!c->gauge_ok) return fail(WL_E_STATE, "no gauge field uploaded (wl_gauge_upload)");
    if (prec == WL_F64) return WL_OK;
    if (prec == WL_F32) {
        if (!c->gauge32) {
            if (hipMalloc(&c->gauge32, gauge_elems(c->g) * 8) != hipSuccess)
                return fail(WL_E_OOM, "FP32 gauge copy");
        }
        if (!c->gauge32_ok) {
            HIPCHK(launch_gauge_convert(c->g, WL_F32, c->gauge64, c->gauge32, c->stream));
            c->gauge32_ok = true;
        }
        return WL_OK;
    }
    return fail(WL_E_NOTIMPL, "precision %d", prec);
}

// out_par = a * H_sign in (+ xin), one rhs
int dslash_one(wl_ctx *c, int prec, void *out, int out_par, const void *in, const void *xin,
               double a, int sign, int epi, const int *done, int fin) {
    DslashLaunch L{};
    L.out = out;
    L.in = in;
    L.xin = xin;
    L.gauge = gauge_for(c, prec);
    L.a = a;
    L.out_par = out_par;
    L.sign = sign;
    L.epi = epi;
    L.done = done;
    L.red = red_args(c);
    L.fin = fin;
    HIPCHK(launch_dslash(prec, c->g, L, c->stream));
    return WL_OK;
}

// M_s in -> out for even FP64/FP32 fields (one rhs), through the odd temporary tmp:
// tmp_o = -kappa H_oe in  (Dirac_Wilson_hop_set(..., ODD, sign), wilson.h:275,285)
// out_e = in + kappa H_eo tmp_o  (Dirac_Wilson_hop(-kappa, out, out, EVEN, sign), :277,287)
int meo_one(wl_ctx *c, int prec, void *out, const void *in, void *tmp, double kappa, int sign,
            int epi2, const int *done, int fin2) {
    int rc = dslash_one(c, prec, tmp, 1, in, nullptr, -kappa, sign, EPI_SET, done, FIN_NONE);
    if (rc) return rc;
    return dslash_one(c, prec, out, 0, tmp, in, kappa, sign, epi2, done, fin2);
}

int check_sign(int sign) { return (sign == 1 || sign == -1) ? WL_OK : fail(WL_E_ARG, "sign must be +-1"); }

int ensure_partials(wl_ctx *c, size_t n) {
    if (c->partials_cap >= n) return WL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipFree(c->partials));
    c->partials = nullptr;
    c->partials_cap = 0;
    if (hipMalloc(&c->partials, n * sizeof(double)) != hipSuccess)
        return fail(WL_E_OOM, "reduction buffer of %zu doubles", n);
    c->partials_cap = n;
    return WL_OK;
}

int ensure_st_mr(wl_ctx *c, int N) {
    if (c->st_mr_cap >= N) return WL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->st_mr) (void)hipFree(c->st_mr);
    if (c->st_mr_host) (void)hipHostFree(c->st_mr_host);
    c->st_mr = nullptr;
    c->st_mr_host = nullptr;
    c->st_mr_cap = 0;
    if (hipMalloc(&c->st_mr, N * sizeof(CgState)) != hipSuccess ||
        hipHostMalloc(&c->st_mr_host, N * sizeof(CgState), hipHostMallocDefault) != hipSuccess)
        return fail(WL_E_OOM, "CG states for %d right-hand sides", N);
    c->st_mr_cap = N;
    return WL_OK;
}

// Work field `slot` of the given precision, parity and rhs count (created on first use, kept
// resident with the context).
enum WorkSlot { W_R = 0, W_P, W_DP, W_T, W_DP64, W_DDP64, W_T64, W_X };
int work_field(wl_ctx *c, int slot, int prec, int parity, int nrhs, wl_field **out) {
    const auto key = std::make_tuple(slot, prec, parity, nrhs);
    auto it = c->work.find(key);
    if (it != c->work.end()) {
        *out = it->second;
        return WL_OK;
    }
    int rc = field_new(c, parity, prec, nrhs, out);
    if (rc == WL_OK) c->work[key] = *out;
    return rc;
}

// Right-hand sides per lane of the multi-rhs kernel (register blocking).
int rhs_per_lane(const wl_ctx *c, int prec, int N) {
    const int cap = prec == WL_F64 ? 2 : 4;
    const int nr = c->rhs_block > 0 ? c->rhs_block : cap;
    return std::max(1, std::min(nr, std::min(N, cap)));
}

// Right-hand sides per launch (a multiple of the per-lane count): each launch is one pass over
// the lattice, so a smaller group keeps the neighbour-spinor footprint of a segment small
// enough for the XCD L2, at one extra link read per pass.
int rhs_per_launch(const wl_ctx *c, int nr, int N) {
    int nl = c->rhs_launch > 0 ? c->rhs_launch : N;
    nl = std::max(nr, nl / nr * nr);
    return std::min(nl, N);
}

// out = a H_sign in (+ xin, fused CG update, ...) for all N right-hand sides, in passes of
// rhs_per_launch right-hand sides; per-rhs partial sums land in c->partials rows,
// mr_segments apart.
int dslash_mr_all(wl_ctx *c, int prec, void *out, int out_par, const void *in, const void *xin,
                  double a, int sign, int epi, int N, const int *all_done, CgState *st, void *x,
                  const void *p, void *r, int *pstride_out) {
    const int pstride = mr_segments(c->g);
    if (epi == EPI_AXPY_NORM || epi == EPI_CGUPD) {
        int rc = ensure_partials(c, (size_t)N * pstride);
        if (rc) return rc;
    }
    const int nr = rhs_per_lane(c, prec, N);
    const int nl = rhs_per_launch(c, nr, N);
    for (int j0 = 0; j0 < N; j0 += nl) {
        MrLaunch L{};
        L.out = out;
        L.in = in;
        L.xin = xin;
        L.gauge = gauge_for(c, prec);
        L.a = a;
        L.out_par = out_par;
        L.sign = sign;
        L.epi = epi;
        L.j0 = j0;
        L.NB = std::min(nl, N - j0);
        L.NR = nr;
        L.all_done = all_done;
        L.st = st;
        L.x = x;
        L.p = p;
        L.r = r;
        L.partials = c->partials;
        HIPCHK(launch_dslash_mr(prec, c->g, L, c->stream));
    }
    if (pstride_out) *pstride_out = pstride;
    return WL_OK;
}

// M_s for N right-hand sides (even fields), through the odd work field t.
int meo_mr(wl_ctx *c, int prec, void *out, const void *in, void *t, double kappa, int sign,
           int epi2, int N, const int *all_done, CgState *st, int *pstride_out) {
    int rc = dslash_mr_all(c, prec, t, 1, in, nullptr, -kappa, sign, EPI_SET, N, all_done, st,
                           nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    return dslash_mr_all(c, prec, out, 0, t, in, kappa, sign, epi2, N, all_done, st, nullptr,
                         nullptr, nullptr, pstride_out);
}

// Segments (64 checkerboard sites) per time slice, or 0 when a slice is not a whole number of
// segments (then only the memory-order traversal is available).
int segments_per_slice(const Geom &g) {
    const int64_t slice = (int64_t)g.L[0] * g.L[1] * g.L[2] / 2;
    return (slice % 64 == 0) ? (int)(slice / 64) : 0;
}

// Default traversal chunk, from measurements on MI355X: one chunk per XCD (a chunk is a
// slab of z-planes, swept over all t) when the slice splits 8 ways, else the largest divisor
// of the segments per slice up to a quarter slice.  Lattices below 2048 segments per parity
// are latency-bound and keep memory order.
int auto_chunk(const Geom &g) {
    const int sps = segments_per_slice(g);
    if (sps == 0 || g.vh / 64 < 2048) return 0;
    if (sps % 8 == 0) return sps / 8;
    for (int q = sps / 4; q >= 1; --q)
        if (sps % q == 0) return q;
    return 0;
}

int set_chunk(Geom &g, int qc) {
    const int sps = segments_per_slice(g);
    if (qc < 0) qc = auto_chunk(g);
    if (qc > 0 && (sps == 0 || sps % qc != 0))
        return fail(WL_E_ARG, "traversal chunk %d must divide the %d segments per time slice", qc, sps);
    g.qc = qc;
    g.sps = sps;
    g.ns = g.vh / 64;
    return WL_OK;
}

}  // namespace

// =================================================================== library ============

extern "C" const char *wl_version(void) { return "hila_wilson 0.1 (gfx950)"; }
extern "C" const char *wl_last_error(void) { return g_err; }

extern "C" int wl_create(const int dims[4], int antiperiodic_t, int device, wl_ctx **out) {
    if (!dims || !out) return fail(WL_E_ARG, "null argument");
    int64_t V = 1;
    for (int i = 0; i < 4; ++i) {
        if (dims[i] < 2 || (dims[i] & 1))
            return fail(WL_E_ARG, "lattice size %d in direction %d: must be even and >= 2", dims[i], i);
        V *= dims[i];
    }
    if (V / 2 > (int64_t)1 << 28) return fail(WL_E_ARG, "lattice too large for 32-bit site indexing");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(WL_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(WL_E_ARG, "device %d of %d", device, ndev);
    wl_ctx *c = new (std::nothrow) wl_ctx;
    if (!c) return fail(WL_E_OOM, "host allocation");
    c->device = device;
    c->V = V;
    c->antiperiodic = antiperiodic_t ? 1 : 0;
    for (int i = 0; i < 4; ++i) c->g.L[i] = dims[i];
    c->g.vh = (int)(V / 2);
    c->g.stride = (int)((V / 2 + kBlock - 1) / kBlock * kBlock);
    c->g.antiperiodic = c->antiperiodic;
    c->g.blk = dslash_block_auto(c->g.vh);
    (void)set_chunk(c->g, -1);
    DeviceGuard dg(device);
    int rc = WL_OK;
    // one partial-sum slot per workgroup of the largest grid (64-lane workgroups)
    const int nb = (c->g.vh + 63) / 64;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->partials, sizeof(double) * (nb + 64)) != hipSuccess ||
        hipMalloc(&c->flags_d, 64) != hipSuccess ||
        hipHostMalloc(&c->flags_h, 64, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&c->st, sizeof(CgState)) != hipSuccess ||
        hipMalloc(&c->scratch_d, 256) != hipSuccess ||
        hipHostMalloc(&c->st_host, sizeof(CgState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&c->scratch_h, 256, hipHostMallocDefault) != hipSuccess) {
        rc = fail(WL_E_DEVICE, "context allocation failed");
    } else if (hipMemsetAsync(c->st, 0, sizeof(CgState), c->stream) != hipSuccess ||
               hipMemsetAsync(c->flags_d, 0, 64, c->stream) != hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = fail(WL_E_DEVICE, "context initialisation failed");
    }
    if (rc) {
        wl_destroy(c);
        return rc;
    }
    c->partials_cap = nb + 64;
    *out = c;
    return WL_OK;
}

extern "C" int wl_destroy(wl_ctx *c) {
    if (!c) return WL_OK;
    DeviceGuard dg(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (wl_field *f : {c->tmp_odd64, c->tmp_odd32, c->cg_r, c->cg_p, c->cg_Dp, c->cg_DDp,
                        c->mx_r, c->mx_p, c->mx_Dp})
        field_free(f);
    for (auto &kv : c->work) field_free(kv.second);
    c->work.clear();
    if (c->st_mr) (void)hipFree(c->st_mr);
    if (c->st_mr_host) (void)hipHostFree(c->st_mr_host);
    if (c->flags_d) (void)hipFree(c->flags_d);
    if (c->flags_h) (void)hipHostFree(c->flags_h);
    // fields the caller still holds: release their device memory and orphan the handles
    for (wl_field *f : c->user_fields) {
        if (f->data) (void)hipFree(f->data);
        f->data = nullptr;
        f->ctx = nullptr;
    }
    c->user_fields.clear();
    for (void *p : {(void *)c->gauge64, c->gauge32, c->staging, (void *)c->partials,
                    (void *)c->st, (void *)c->scratch_d})
        if (p) (void)hipFree(p);
    if (c->st_host) (void)hipHostFree(c->st_host);
    if (c->scratch_h) (void)hipHostFree(c->scratch_h);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return WL_OK;
}

extern "C" int wl_sync(wl_ctx *c) {
    if (!c) return fail(WL_E_ARG, "null context");
    DeviceGuard dg(c->device);
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}

extern "C" int64_t wl_volume(const wl_ctx *c) { return c ? c->V : 0; }

extern "C" int wl_gauge_upload(wl_ctx *c, const double *const U[4]) {
    if (!c || !U) return fail(WL_E_ARG, "null argument");
    for (int mu = 0; mu < 4; ++mu)
        if (!U[mu]) return fail(WL_E_ARG, "null gauge direction %d", mu);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    if (!c->gauge64) {
        if (hipMalloc(&c->gauge64, gauge_elems(c->g) * 16) != hipSuccess)
            return fail(WL_E_OOM, "gauge field of %zu bytes", gauge_elems(c->g) * 16);
        HIPCHK(hipMemsetAsync(c->gauge64, 0, gauge_elems(c->g) * 16, c->stream));
    }
    const size_t bytes = (size_t)c->V * 18 * sizeof(double);
    int rc = ensure_staging(c, bytes);
    if (rc) return rc;
    for (int mu = 0; mu < 4; ++mu) {
        HIPCHK(hipMemcpyAsync(c->staging, U[mu], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_gauge_to_soa(c->g, mu, (const double *)c->staging, c->gauge64, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->gauge_ok = true;
    c->gauge32_ok = false;
    return WL_OK;
}

// ==================================================================== fields ============

extern "C" int wl_field_create(wl_ctx *c, int parity, int prec, int nrhs, wl_field **out) {
    if (!c || !out) return fail(WL_E_ARG, "null argument");
    if (parity != WL_EVEN && parity != WL_ODD) return fail(WL_E_ARG, "parity must be WL_EVEN or WL_ODD");
    if (prec != WL_F64 && prec != WL_F32) return fail(WL_E_NOTIMPL, "field precision %d", prec);
    if (nrhs < 1 || nrhs > 64) return fail(WL_E_ARG, "nrhs %d out of range 1..64", nrhs);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc = field_new(c, parity, prec, nrhs, out);
    if (rc == WL_OK) c->user_fields.insert(*out);
    return rc;
}

extern "C" int wl_field_destroy(wl_field *f) {
    if (!f) return WL_OK;
    wl_ctx *c = f->ctx;
    if (!c) {  // orphaned by wl_destroy: only the handle is left
        delete f;
        return WL_OK;
    }
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->user_fields.erase(f);
    field_free(f);
    return WL_OK;
}

extern "C" int wl_field_upload(wl_field *f, int rhs, const double *host) {
    if (!f || !host) return fail(WL_E_ARG, "null argument");
    if (!f->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    if (rhs < 0 || rhs >= f->nrhs) return fail(WL_E_ARG, "rhs %d of %d", rhs, f->nrhs);
    wl_ctx *c = f->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    const size_t bytes = (size_t)c->g.vh * 24 * sizeof(double);
    int rc = ensure_staging(c, bytes);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_aos_to_soa(f->prec, c->g, (const double *)c->staging, f->data, f->nrhs, rhs, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}

extern "C" int wl_field_download(const wl_field *f, int rhs, double *host) {
    if (!f || !host) return fail(WL_E_ARG, "null argument");
    if (!f->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    if (rhs < 0 || rhs >= f->nrhs) return fail(WL_E_ARG, "rhs %d of %d", rhs, f->nrhs);
    wl_ctx *c = f->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    const size_t bytes = (size_t)c->g.vh * 24 * sizeof(double);
    int rc = ensure_staging(c, bytes);
    if (rc) return rc;
    HIPCHK(launch_soa_to_aos(f->prec, c->g, f->data, (double *)c->staging, f->nrhs, rhs, c->stream));
    HIPCHK(hipMemcpyAsync(host, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}

extern "C" int wl_field_zero(wl_field *f) {
    if (!f) return fail(WL_E_ARG, "null argument");
    if (!f->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    DeviceGuard dg(f->ctx->device);
    HIPCHK(hipMemsetAsync(f->data, 0, f->bytes, f->ctx->stream));
    return WL_OK;
}

extern "C" int wl_field_norm2(const wl_field *f, int rhs, double *out) {
    if (!f || !out) return fail(WL_E_ARG, "null argument");
    if (!f->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    if (rhs < 0 || rhs >= f->nrhs) return fail(WL_E_ARG, "rhs %d of %d", rhs, f->nrhs);
    wl_ctx *c = f->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    HIPCHK(launch_norm2(f->prec, c->g, f->data, f->nrhs, rhs, red_args(c), FIN_STORE, nullptr, c->stream));
    HIPCHK(hipMemcpyAsync(c->scratch_h, c->scratch_d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *out = c->scratch_h[0];
    return WL_OK;
}

// ================================================================== operator ============

extern "C" int wl_dslash_eo(wl_ctx *c, wl_field *out, const wl_field *in, double kappa, int sign) {
    if (!c || !out || !in) return fail(WL_E_ARG, "null argument");
    if (check_sign(sign)) return WL_E_ARG;
    if (out->ctx != c || in->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (out->parity == in->parity) return fail(WL_E_ARG, "out and in must have opposite parity");
    if (out->prec != in->prec || out->nrhs != in->nrhs)
        return fail(WL_E_ARG, "out and in differ in precision or nrhs");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc = ensure_gauge(c, in->prec);
    if (rc) return rc;
    const int out_par = out->parity == WL_EVEN ? 0 : 1;
    if (in->nrhs == 1)
        return dslash_one(c, in->prec, out->data, out_par, in->data, nullptr, -kappa, sign, EPI_SET,
                          nullptr, FIN_NONE);
    return dslash_mr_all(c, in->prec, out->data, out_par, in->data, nullptr, -kappa, sign, EPI_SET,
                         in->nrhs, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int wl_meo_apply(wl_ctx *c, wl_field *out, const wl_field *in, double kappa, int sign) {
    if (!c || !out || !in) return fail(WL_E_ARG, "null argument");
    if (check_sign(sign)) return WL_E_ARG;
    if (out->ctx != c || in->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (out->parity != WL_EVEN || in->parity != WL_EVEN) return fail(WL_E_ARG, "M acts on even fields");
    if (out == in) return fail(WL_E_ARG, "out and in must be different fields");
    if (out->prec != in->prec || out->nrhs != in->nrhs)
        return fail(WL_E_ARG, "out and in differ in precision or nrhs");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc = ensure_gauge(c, in->prec);
    if (rc) return rc;
    if (in->nrhs == 1) {
        wl_field **tmp = in->prec == WL_F64 ? &c->tmp_odd64 : &c->tmp_odd32;
        rc = ensure_tmp(c, in->prec, tmp, WL_ODD);
        if (rc) return rc;
        return meo_one(c, in->prec, out->data, in->data, (*tmp)->data, kappa, sign, EPI_AXPY,
                       nullptr, FIN_NONE);
    }
    wl_field *t;
    if ((rc = work_field(c, W_T, in->prec, WL_ODD, in->nrhs, &t))) return rc;
    return meo_mr(c, in->prec, out->data, in->data, t->data, kappa, sign, EPI_AXPY, in->nrhs,
                  nullptr, nullptr, nullptr);
}

// =================================================================== CG =================

namespace {

// Iterations enqueued between two checks of the device `done` flag.
constexpr int kChunk = 8;

int cg_f64(wl_ctx *c, void *x, const void *b, double kappa, double accuracy, int maxiters,
           wl_cg_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    for (wl_field **w : {&c->cg_r, &c->cg_p, &c->cg_Dp, &c->cg_DDp})
        if ((rc = ensure_tmp(c, WL_F64, w, WL_EVEN))) return rc;
    if ((rc = ensure_tmp(c, WL_F64, &c->tmp_odd64, WL_ODD))) return rc;
    void *r = c->cg_r->data, *p = c->cg_p->data, *Dp = c->cg_Dp->data, *DDp = c->cg_DDp->data;
    void *t = c->tmp_odd64->data;
    const RedArgs ra = red_args(c);
    const int *done = &c->st->done;

    CgState init{};
    init.acc2 = accuracy * accuracy;
    init.maxiters = maxiters;
    *c->st_host = init;
    HIPCHK(hipMemcpyAsync(c->st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    // source_norm and target_rr (conjugate_gradient.h:61-63)
    HIPCHK(launch_norm2(WL_F64, c->g, b, 1, 0, ra, FIN_SOURCE, nullptr, c->stream));
    // r = b - M^dag M x ; p = r ; rr = |r|^2  (:65-73)
    if ((rc = meo_one(c, WL_F64, Dp, x, t, kappa, +1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    if ((rc = meo_one(c, WL_F64, DDp, Dp, t, kappa, -1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    HIPCHK(launch_cg_init(c->g, b, DDp, r, p, ra, c->stream));
    HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->st_host->source_norm == 0.0) {
        // b = 0: the solution is x = 0 (HILA would divide 0 by 0 here)
        HIPCHK(hipMemsetAsync(x, 0, c->cg_r->bytes, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->converged = 1;
        }
        return WL_OK;
    }
    if (maxiters <= 0) {
        c->st_host->done = 2;
        HIPCHK(hipMemcpyAsync(&c->st->done, &c->st_host->done, sizeof(int), hipMemcpyHostToDevice, c->stream));
    }

    // main loop (:75-96), kChunk iterations per enqueue; the kernels of an iteration after
    // convergence return at once.
    int enqueued = 0;
    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    int done_h = c->st_host->done;
    while (!done_h && enqueued < maxiters) {
        const int n = std::min(kChunk, maxiters - enqueued);
        for (int it = 0; it < n; ++it) {
            // Dp = M p, pDp = |Dp|^2, alpha = rr / pDp
            if ((rc = meo_one(c, WL_F64, Dp, p, t, kappa, +1, EPI_AXPY_NORM, done, FIN_ALPHA)))
                return rc;
            // DDp = M^dag Dp fused with x += alpha p, r -= alpha DDp, rrnew = |r|^2, test, beta
            if ((rc = dslash_one(c, WL_F64, t, 1, Dp, nullptr, -kappa, -1, EPI_SET, done, FIN_NONE)))
                return rc;
            DslashLaunch L{};
            L.out = nullptr;
            L.in = t;
            L.xin = Dp;
            L.gauge = c->gauge64;
            L.a = kappa;
            L.out_par = 0;
            L.sign = -1;
            L.epi = EPI_CGUPD;
            L.done = done;
            L.x = x;
            L.p = p;
            L.r = r;
            L.red = ra;
            L.fin = FIN_RRNEW;
            HIPCHK(launch_dslash(WL_F64, c->g, L, c->stream));
            // p = beta p + r
            HIPCHK(launch_cg_xpay(WL_F64, c->g, p, r, c->st, c->stream));
        }
        enqueued += n;
        HIPCHK(hipMemcpyAsync(&c->st_host->done, &c->st->done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(ev, c->stream));
        HIPCHK(hipEventSynchronize(ev));
        done_h = c->st_host->done;
    }
    (void)hipEventDestroy(ev);

    // true residual |b - M^dag M x|^2 recomputed in FP64
    if ((rc = meo_one(c, WL_F64, Dp, x, t, kappa, +1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    if ((rc = meo_one(c, WL_F64, DDp, Dp, t, kappa, -1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    HIPCHK(launch_diff_norm2(c->g, b, DDp, ra, FIN_TRUE, c->stream));
    HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const CgState &S = *c->st_host;
    const bool conv = S.done == 1;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = S.iter;
        stats->converged = conv ? 1 : 0;
        stats->rr_rel = S.rrnew / S.source_norm;
        stats->true_resid = std::sqrt(S.true_rr / S.source_norm);
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (!conv) return fail(WL_E_NOCONV, "CG did not converge in %d iterations (rr_rel %.3e)", S.iter,
                           S.rrnew / S.source_norm);
    return WL_OK;
}

// Mixed-precision CG on M^dag M with reliable updates (FP64 solution, FP64 reductions, inner
// operator and Krylov vectors r, p, Dp in `prec`).  The recurrences are those of HILA's CG
// (conjugate_gradient.h:75-95) run in the inner precision; the finish kernel of each iteration
// asks for a reliable update when ||r|| has dropped by `delta` since the last one, after
// `update_every` iterations, or when the iterated residual meets the target.  An update
// recomputes r = b - M^dag M x with the FP64 operator, restarts rr and beta from it and
// declares convergence only on that true residual.
int cg_mixed(wl_ctx *c, void *x, const void *b, double kappa, double accuracy, int maxiters,
             int prec, double delta, int update_every, wl_cg_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if ((rc = ensure_gauge(c, prec))) return rc;
    for (wl_field **w : {&c->cg_Dp, &c->cg_DDp})
        if ((rc = ensure_tmp(c, WL_F64, w, WL_EVEN))) return rc;
    if ((rc = ensure_tmp(c, WL_F64, &c->tmp_odd64, WL_ODD))) return rc;
    for (wl_field **w : {&c->mx_r, &c->mx_p, &c->mx_Dp}) {
        if (*w && (*w)->prec != prec) {
            field_free(*w);
            *w = nullptr;
        }
        if ((rc = ensure_tmp(c, prec, w, WL_EVEN))) return rc;
    }
    // odd scratch of the inner operator (shared with the FP64 update when prec is FP64: the
    // stream orders their uses)
    wl_field **todd = prec == WL_F64 ? &c->tmp_odd64 : &c->tmp_odd32;
    if ((rc = ensure_tmp(c, prec, todd, WL_ODD))) return rc;
    void *r = c->mx_r->data, *p = c->mx_p->data, *Dp = c->mx_Dp->data, *t = (*todd)->data;
    void *Dp64 = c->cg_Dp->data, *DDp64 = c->cg_DDp->data, *t64 = c->tmp_odd64->data;
    const RedArgs ra = red_args(c);
    const int *done = &c->st->done;

    CgState init{};
    init.acc2 = accuracy * accuracy;
    init.maxiters = maxiters;
    init.delta2 = delta * delta;
    init.update_every = update_every;
    *c->st_host = init;
    HIPCHK(hipMemcpyAsync(c->st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_norm2(WL_F64, c->g, b, 1, 0, ra, FIN_SOURCE, nullptr, c->stream));

    // reliable update: r = b - M^dag M x in FP64, r_lo = r, restart (or converged)
    auto reliable_update = [&]() -> int {
        int e;
        if ((e = meo_one(c, WL_F64, Dp64, x, t64, kappa, +1, EPI_AXPY, nullptr, FIN_NONE))) return e;
        if ((e = meo_one(c, WL_F64, DDp64, Dp64, t64, kappa, -1, EPI_AXPY, nullptr, FIN_NONE))) return e;
        HIPCHK(launch_residual_update(prec, c->g, b, DDp64, r, ra, c->stream));
        return WL_OK;
    };
    if ((rc = reliable_update())) return rc;
    HIPCHK(launch_copy(prec, c->g, r, p, 1, c->stream));  // p = r
    HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->st_host->source_norm == 0.0) {
        HIPCHK(hipMemsetAsync(x, 0, c->cg_Dp->bytes, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->converged = 1;
        }
        return WL_OK;
    }
    int done_h = c->st_host->done;
    if (maxiters <= 0 && done_h == 0) done_h = 2;

    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    while (done_h == 0 || done_h == 3 || done_h == 4) {
        if (done_h == 3 || done_h == 4) {
            if ((rc = reliable_update())) return rc;
            HIPCHK(launch_cg_xpay(prec, c->g, p, r, c->st, c->stream));  // skipped if converged
        }
        for (int it = 0; it < kChunk; ++it) {
            if ((rc = meo_one(c, prec, Dp, p, t, kappa, +1, EPI_AXPY_NORM, done, FIN_ALPHA))) return rc;
            if ((rc = dslash_one(c, prec, t, 1, Dp, nullptr, -kappa, -1, EPI_SET, done, FIN_NONE)))
                return rc;
            DslashLaunch L{};
            L.in = t;
            L.xin = Dp;
            L.gauge = gauge_for(c, prec);
            L.a = kappa;
            L.out_par = 0;
            L.sign = -1;
            L.epi = EPI_CGUPD;
            L.done = done;
            L.x = x;
            L.p = p;
            L.r = r;
            L.red = ra;
            L.fin = FIN_RRNEW_MIXED;
            HIPCHK(launch_dslash(prec, c->g, L, c->stream));
            HIPCHK(launch_cg_xpay(prec, c->g, p, r, c->st, c->stream));
        }
        HIPCHK(hipMemcpyAsync(&c->st_host->done, &c->st->done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(ev, c->stream));
        HIPCHK(hipEventSynchronize(ev));
        done_h = c->st_host->done;
    }
    (void)hipEventDestroy(ev);
    if (done_h == 2) {  // maxiters: report the true residual of the final x
        if ((rc = reliable_update())) return rc;
    }
    HIPCHK(hipMemcpyAsync(c->st_host, c->st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const CgState &S = *c->st_host;
    const bool conv = S.done == 1;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = S.iter;
        stats->converged = conv ? 1 : 0;
        stats->rr_rel = S.rrnew / S.source_norm;
        stats->true_resid = std::sqrt(S.true_rr / S.source_norm);
        stats->reliable_updates = S.nreliable - 1;  // the initial residual is not an update
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (!conv)
        return fail(WL_E_NOCONV, "mixed CG did not converge in %d iterations (true residual %.3e)",
                    S.iter, std::sqrt(S.true_rr / S.source_norm));
    return WL_OK;
}

// Batched CG for N right-hand sides that share every operator application: one CgState per
// rhs, each with HILA's recurrence (conjugate_gradient.h:75-95).  mixed = false: FP64
// throughout, each rhs stops at its own iterated criterion exactly as HILA's CG.  mixed = true:
// inner operator and Krylov vectors in `prec`, FP64 solution and reductions, reliable updates
// (recomputed FP64 residual) taken by all running rhs together whenever one of them asks
// (residual drop by delta, update_every iterations, or iterated convergence), convergence on
// the FP64 true residual.
int cg_batched(wl_ctx *c, void *x, const void *b, int N, double kappa, double accuracy,
               int maxiters, int prec, bool mixed, double delta, int update_every,
               wl_cg_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if ((rc = ensure_gauge(c, prec))) return rc;
    if ((rc = ensure_st_mr(c, N))) return rc;
    wl_field *R, *P, *DP, *T, *DP64, *DDP64, *T64;
    if ((rc = work_field(c, W_R, prec, WL_EVEN, N, &R)) || (rc = work_field(c, W_P, prec, WL_EVEN, N, &P)) ||
        (rc = work_field(c, W_DP, prec, WL_EVEN, N, &DP)) || (rc = work_field(c, W_T, prec, WL_ODD, N, &T)) ||
        (rc = work_field(c, W_DP64, WL_F64, WL_EVEN, N, &DP64)) ||
        (rc = work_field(c, W_DDP64, WL_F64, WL_EVEN, N, &DDP64)) ||
        (rc = work_field(c, W_T64, WL_F64, WL_ODD, N, &T64)))
        return rc;
    CgState *st = c->st_mr;
    const int *all_done = c->flags_d;
    for (int j = 0; j < N; ++j) {
        CgState s0{};
        s0.acc2 = accuracy * accuracy;
        s0.maxiters = maxiters;
        s0.delta2 = delta * delta;
        s0.update_every = update_every;
        c->st_mr_host[j] = s0;
    }
    HIPCHK(hipMemcpyAsync(st, c->st_mr_host, N * sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->flags_d, 0, 64, c->stream));
    const int pplain = nblocks(c->g);
    if ((rc = ensure_partials(c, (size_t)N * pplain))) return rc;
    HIPCHK(launch_norm2_mr(WL_F64, c->g, b, N, c->partials, c->stream));
    HIPCHK(launch_finish_mr(c->partials, pplain, pplain, N, st, nullptr, FIN_SOURCE, nullptr, c->stream));

    // r = b - M^dag M x, p = r
    if ((rc = meo_mr(c, WL_F64, DP64->data, x, T64->data, kappa, +1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, DDP64->data, DP64->data, T64->data, kappa, -1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
        return rc;
    if (!mixed) {
        HIPCHK(launch_cg_init_mr(c->g, b, DDP64->data, R->data, P->data, N, c->partials, c->stream));
        HIPCHK(launch_finish_mr(c->partials, pplain, pplain, N, st, nullptr, FIN_RR, nullptr, c->stream));
    } else {
        HIPCHK(launch_residual_update_mr(prec, c->g, b, DDP64->data, R->data, st, 1, N, c->partials, c->stream));
        HIPCHK(launch_finish_mr(c->partials, pplain, pplain, N, st, nullptr, FIN_UPDATE_FIRST, nullptr, c->stream));
        HIPCHK(launch_copy(prec, c->g, R->data, P->data, N, c->stream));
    }
    HIPCHK(launch_batch_sync(st, N, 0, c->flags_d, c->stream));
    if (maxiters <= 0) {
        for (int j = 0; j < N; ++j) c->st_mr_host[j].done = 2;
    }

    hipEvent_t ev;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    int enqueued = 0;
    const int cap = maxiters + 2 * kChunk + 16;
    while (maxiters > 0) {
        for (int it = 0; it < kChunk; ++it) {
            int ps = 0;
            if ((rc = meo_mr(c, prec, DP->data, P->data, T->data, kappa, +1, EPI_AXPY_NORM, N, all_done, st, &ps)))
                return rc;
            HIPCHK(launch_finish_mr(c->partials, ps, ps, N, st, nullptr, FIN_ALPHA, all_done, c->stream));
            if ((rc = dslash_mr_all(c, prec, T->data, 1, DP->data, nullptr, -kappa, -1, EPI_SET, N, all_done, st,
                                    nullptr, nullptr, nullptr, nullptr)))
                return rc;
            if ((rc = dslash_mr_all(c, prec, nullptr, 0, T->data, DP->data, kappa, -1, EPI_CGUPD, N, all_done, st,
                                    x, P->data, R->data, &ps)))
                return rc;
            HIPCHK(launch_finish_mr(c->partials, ps, ps, N, st, nullptr, mixed ? FIN_RRNEW_MIXED2 : FIN_RRNEW,
                                    all_done, c->stream));
            HIPCHK(launch_batch_sync(st, N, mixed ? 1 : 0, c->flags_d, c->stream));
            HIPCHK(launch_xpay_mr(prec, c->g, P->data, R->data, st, all_done, N, c->stream));
        }
        enqueued += kChunk;
        HIPCHK(hipMemcpyAsync(c->flags_h, c->flags_d, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipEventRecord(ev, c->stream));
        HIPCHK(hipEventSynchronize(ev));
        if (c->flags_h[0]) break;
        if (mixed && c->flags_h[1]) {  // reliable update of every running rhs
            if ((rc = meo_mr(c, WL_F64, DP64->data, x, T64->data, kappa, +1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
                return rc;
            if ((rc = meo_mr(c, WL_F64, DDP64->data, DP64->data, T64->data, kappa, -1, EPI_AXPY, N, nullptr, nullptr,
                             nullptr)))
                return rc;
            HIPCHK(launch_residual_update_mr(prec, c->g, b, DDP64->data, R->data, st, 0, N, c->partials,
                                             c->stream));
            HIPCHK(launch_finish_mr(c->partials, pplain, pplain, N, st, nullptr, FIN_UPDATE, nullptr, c->stream));
            HIPCHK(launch_xpay_mr(prec, c->g, P->data, R->data, st, nullptr, N, c->stream));
            HIPCHK(launch_batch_sync(st, N, 0, c->flags_d, c->stream));
        }
        if (enqueued > cap) break;  // every rhs stops at maxiters; this is a guard only
    }
    (void)hipEventDestroy(ev);

    // FP64 true residual of every rhs
    if ((rc = meo_mr(c, WL_F64, DP64->data, x, T64->data, kappa, +1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, DDP64->data, DP64->data, T64->data, kappa, -1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
        return rc;
    HIPCHK(launch_residual_update_mr(WL_F64, c->g, b, DDP64->data, DP64->data, st, 1, N, c->partials,
                                     c->stream));
    HIPCHK(launch_finish_mr(c->partials, pplain, pplain, N, st, nullptr, FIN_TRUE, nullptr, c->stream));
    HIPCHK(hipMemcpyAsync(c->st_mr_host, st, N * sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int worst = WL_OK;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int j = 0; j < N; ++j) {
        const CgState &S = c->st_mr_host[j];
        if (S.source_norm == 0.0) {  // b_j = 0: x_j = 0
            HIPCHK(launch_zero_rhs(c->g, x, j, c->stream));
        }
        const bool conv = S.done == 1;
        if (!conv) worst = WL_E_NOCONV;
        if (stats) {
            wl_cg_stats &o = stats[j];
            std::memset(&o, 0, sizeof(o));
            o.iterations = S.iter;
            o.converged = conv ? 1 : 0;
            o.rr_rel = S.source_norm > 0 ? S.rrnew / S.source_norm : 0.0;
            o.true_resid = S.source_norm > 0 ? std::sqrt(S.true_rr / S.source_norm) : 0.0;
            o.reliable_updates = mixed ? std::max(0, S.nreliable - 1) : 0;
            o.ms = ms;
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (worst) return fail(WL_E_NOCONV, "batched CG: not every right-hand side converged");
    return WL_OK;
}

}  // namespace

extern "C" int wl_cg_solve(wl_ctx *c, wl_field *x, const wl_field *b, double kappa, double accuracy,
                           int maxiters, wl_cg_stats *stats) {
    if (!c || !x || !b) return fail(WL_E_ARG, "null argument");
    if (x->ctx != c || b->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (x->parity != WL_EVEN || b->parity != WL_EVEN) return fail(WL_E_ARG, "CG acts on even fields");
    if (x == b) return fail(WL_E_ARG, "x and b must be different fields");
    if (x->prec != WL_F64 || b->prec != WL_F64) return fail(WL_E_ARG, "wl_cg_solve takes FP64 fields");
    if (x->nrhs != b->nrhs) return fail(WL_E_ARG, "x and b differ in nrhs");
    if (!(accuracy > 0)) return fail(WL_E_ARG, "accuracy must be > 0");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc = ensure_gauge(c, WL_F64);
    if (rc) return rc;
    if (b->nrhs == 1) return cg_f64(c, x->data, b->data, kappa, accuracy, maxiters, stats);
    return cg_batched(c, x->data, b->data, b->nrhs, kappa, accuracy, maxiters, WL_F64, false, 0.1,
                      1 << 30, stats);
}

// ================================================================ diagnostics ===========

extern "C" int wl_time_op(wl_ctx *c, int what, wl_field *out, const wl_field *in, double kappa,
                          int sign, int warmup, int iters, double *ms_per_launch) {
    if (!c || !out || !in || !ms_per_launch || iters < 1 || warmup < 0)
        return fail(WL_E_ARG, "bad argument");
    auto once = [&]() -> int {
        return what == 0 ? wl_dslash_eo(c, out, in, kappa, sign) : wl_meo_apply(c, out, in, kappa, sign);
    };
    if (what != 0 && what != 1) return fail(WL_E_ARG, "what must be 0 (hopping term) or 1 (M)");
    for (int i = 0; i < warmup; ++i)
        if (int rc = once()) return rc;
    DeviceGuard dg(c->device);
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i)
        if (int rc = once()) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_launch = (double)ms / iters;
    return WL_OK;
}

extern "C" int wl_field_copy(wl_field *dst, const wl_field *src) {
    if (!dst || !src) return fail(WL_E_ARG, "null argument");
    if (!dst->ctx || !src->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    if (dst->ctx != src->ctx) return fail(WL_E_ARG, "fields belong to different contexts");
    if (dst->parity != src->parity || dst->prec != src->prec || dst->nrhs != src->nrhs)
        return fail(WL_E_ARG, "fields differ in parity, precision or nrhs");
    if (dst == src) return WL_OK;
    wl_ctx *c = dst->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    HIPCHK(launch_copy(src->prec, c->g, src->data, dst->data, src->nrhs, c->stream));
    return WL_OK;
}

extern "C" int wl_set_option(wl_ctx *c, int option, int value) {
    if (!c) return fail(WL_E_ARG, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    switch (option) {
    case WL_OPT_TRAVERSAL_CHUNK: return set_chunk(c->g, value);
    case WL_OPT_BLOCK:
        if (value == 0) value = dslash_block_auto(c->g.vh);
        if (value != 64 && value != 128 && value != 256)
            return fail(WL_E_ARG, "workgroup size must be 64, 128 or 256");
        c->g.blk = value;
        return WL_OK;
    case WL_OPT_RHS_BLOCK:
        if (value < 0 || value > 4) return fail(WL_E_ARG, "rhs per lane must be 0 (automatic) .. 4");
        c->rhs_block = value;
        return WL_OK;
    case WL_OPT_RHS_LAUNCH:
        if (value < 0 || value > 64) return fail(WL_E_ARG, "rhs per launch must be 0 (all) .. 64");
        c->rhs_launch = value;
        return WL_OK;
    default: return fail(WL_E_ARG, "unknown option %d", option);
    }
}

extern "C" int wl_get_option(const wl_ctx *c, int option, int *value) {
    if (!c || !value) return fail(WL_E_ARG, "null argument");
    switch (option) {
    case WL_OPT_TRAVERSAL_CHUNK: *value = c->g.qc; return WL_OK;
    case WL_OPT_BLOCK: *value = c->g.blk; return WL_OK;
    case WL_OPT_RHS_BLOCK: *value = c->rhs_block; return WL_OK;
    case WL_OPT_RHS_LAUNCH: *value = c->rhs_launch; return WL_OK;
    default: return fail(WL_E_ARG, "unknown option %d", option);
    }
}

extern "C" int wl_cg_solve_mixed(wl_ctx *c, wl_field *x, const wl_field *b, double kappa,
                                 double accuracy, int maxiters, int inner_prec, double delta,
                                 int update_every, wl_cg_stats *stats) {
    if (!c || !x || !b) return fail(WL_E_ARG, "null argument");
    if (x->ctx != c || b->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (x->parity != WL_EVEN || b->parity != WL_EVEN) return fail(WL_E_ARG, "CG acts on even fields");
    if (x == b) return fail(WL_E_ARG, "x and b must be different fields");
    if (x->prec != WL_F64 || b->prec != WL_F64) return fail(WL_E_ARG, "x and b must be FP64 fields");
    if (x->nrhs != b->nrhs) return fail(WL_E_ARG, "x and b differ in nrhs");
    if (inner_prec != WL_F32 && inner_prec != WL_F64)
        return fail(WL_E_NOTIMPL, "inner precision %d", inner_prec);
    if (!(accuracy > 0) || !(delta > 0 && delta < 1) || update_every < 1)
        return fail(WL_E_ARG, "need accuracy > 0, 0 < delta < 1, update_every >= 1");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc = ensure_gauge(c, WL_F64);
    if (rc) return rc;
    if (b->nrhs == 1)
        return cg_mixed(c, x->data, b->data, kappa, accuracy, maxiters, inner_prec, delta,
                        update_every, stats);
    return cg_batched(c, x->data, b->data, b->nrhs, kappa, accuracy, maxiters, inner_prec, true,
                      delta, update_every, stats);
}
