// C ABI of libhila_wilson.so: contexts, device fields and the even-odd Wilson operator, plus
// the host-side helpers the solvers (wl_cg.hip) and the request server (wl_serve.hip) share.
// Declarations and conventions: include/wilson_mi355x.h.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "wl_host.h"
#include "wl_store.h"

namespace wl {

namespace {
thread_local char g_err[512] = "";
}  // namespace

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *last_error() { return g_err; }

int dev_alloc(wl_ctx *c, void **p, size_t bytes, const char *what) {
    *p = nullptr;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return fail(WL_E_OOM, "%s: device allocation of %zu bytes", what, bytes);
    }
    c->dev_bytes += bytes;
    return WL_OK;
}

void dev_free(wl_ctx *c, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    c->dev_bytes -= bytes;
}

int ensure_staging(wl_ctx *c, size_t bytes) {
    if (c->staging_bytes >= bytes) return WL_OK;
    if (c->staging) {
        HIPCHK(hipStreamSynchronize(c->stream));
        dev_free(c, c->staging, c->staging_bytes);
        c->staging = nullptr;
        c->staging_bytes = 0;
    }
    int rc = dev_alloc(c, &c->staging, bytes, "staging buffer");
    if (rc == WL_OK) c->staging_bytes = bytes;
    return rc;
}

int field_new(wl_ctx *c, int parity, int prec, int nrhs, wl_field **out) {
    wl_field *f = new (std::nothrow) wl_field;
    if (!f) return fail(WL_E_OOM, "host allocation");
    f->ctx = c;
    f->parity = parity;
    f->prec = prec;
    f->nrhs = nrhs;
    f->rhs_elems = (size_t)planes_of(prec) * c->g.stride;
    f->bytes = f->rhs_elems * nrhs * elem_bytes(prec);
    if (f->rhs_elems * (size_t)nrhs >= ((size_t)1 << 31)) {
        delete f;
        return fail(WL_E_ARG, "%d right-hand sides exceed 2^31 elements per field at this volume",
                    nrhs);
    }
    if (int rc = dev_alloc(c, &f->data, f->bytes, "device field")) {
        delete f;
        return rc;
    }
    if (hipMemsetAsync(f->data, 0, f->bytes, c->stream) != hipSuccess) {
        dev_free(c, f->data, f->bytes);
        delete f;
        return fail(WL_E_DEVICE, "hipMemsetAsync of a new field");
    }
    *out = f;
    return WL_OK;
}

void field_free(wl_field *f) {
    if (!f) return;
    if (f->data) {
        if (f->ctx)
            dev_free(f->ctx, f->data, f->bytes);
        else
            (void)hipFree(f->data);
    }
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

RedArgs red_args(const CgWork &w) {
    RedArgs ra;
    ra.partials = w.partials;
    ra.st = w.st;
    ra.out = w.red_out;
    return ra;
}

// Stored entries per link of the links the hopping kernels of precision prec read: 6 for
// two-row links (WL_OPT_LINK_ROWS = 2, FP64 and FP32), else 9.
int link_format(const wl_ctx *c, int prec) { return c->link_rows == 2 && prec != WL_F16 ? 6 : 9; }

const void *gauge_for(wl_ctx *c, int prec) {
    if (link_format(c, prec) == 6) return prec == WL_F64 ? c->rows64 : c->rows32;
    return prec == WL_F64 ? c->gauge64 : (prec == WL_F32 ? c->gauge32 : c->gauge16);
}

// The largest |re|, |im| of U[2][j] - conj(row 0 x row 1)_j that two-row links accept: the
// rebuilt third row then agrees with the stored one to rounding, and so do the operator's
// results.
constexpr double kRowsTol = 1e-13;

// Checks the uploaded FP64 links for two-row storage (once per upload) and builds the two-row
// copy of precision prec (FP64 or FP32).
int ensure_rows(wl_ctx *c, int prec) {
    if (c->rows_check == 0) {
        const int nb = two_rows_blocks(c->g);
        if (int rc = ensure_partials(c, (size_t)nb)) return rc;
        HIPCHK(launch_two_rows(c->g, WL_F64, c->gauge64, nullptr, c->partials, c->stream));
        std::vector<double> dev((size_t)nb);
        HIPCHK(hipMemcpyAsync(dev.data(), c->partials, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        double m = 0;
        for (double d : dev) m = std::max(m, d);
        c->rows_dev = m;
        c->rows_check = m <= kRowsTol ? 1 : -1;
    }
    if (c->rows_check < 0)
        return fail(WL_E_ARG,
                    "two-row links (WL_OPT_LINK_ROWS = 2) need special unitary links: a stored third row differs "
                    "from conj(row 0 x row 1) by %.3g (> %.0e)",
                    c->rows_dev, kRowsTol);
    void **rows = prec == WL_F64 ? &c->rows64 : &c->rows32;
    bool *ok = prec == WL_F64 ? &c->rows64_ok : &c->rows32_ok;
    if (!*rows) {
        if (int rc = dev_alloc(c, rows, rows_elems(c->g) * (prec == WL_F64 ? 16 : 8), "two-row link copy")) return rc;
    }
    if (!*ok) {
        HIPCHK(launch_two_rows(c->g, prec, c->gauge64, *rows, nullptr, c->stream));
        *ok = true;
    }
    return WL_OK;
}

int ensure_gauge(wl_ctx *c, int prec) {
    if (!c->gauge_ok) return fail(WL_E_STATE, "no gauge field uploaded (wl_gauge_upload)");
    if (link_format(c, prec) == 6) return ensure_rows(c, prec);
    if (prec == WL_F64) return WL_OK;
    if (prec == WL_F32) {
        if (!c->gauge32) {
            if (int rc = dev_alloc(c, &c->gauge32, gauge_elems(c->g) * 8, "FP32 gauge copy")) return rc;
        }
        if (!c->gauge32_ok) {
            HIPCHK(launch_gauge_convert(c->g, WL_F32, c->gauge64, c->gauge32, c->stream));
            c->gauge32_ok = true;
        }
        return WL_OK;
    }
    if (prec == WL_F16) {
        if (!c->gauge16) {
            if (int rc = dev_alloc(c, &c->gauge16, gauge_elems(c->g) * 4, "FP16 gauge copy")) return rc;
        }
        if (!c->gauge16_ok) {
            HIPCHK(launch_gauge_convert(c->g, WL_F16, c->gauge64, c->gauge16, c->stream));
            c->gauge16_ok = true;
        }
        return WL_OK;
    }
    return fail(WL_E_NOTIMPL, "precision %d", prec);
}

// out_par = a * H_sign in (+ xin), one rhs
int dslash_one(wl_ctx *c, int prec, void *out, int out_par, const void *in, const void *xin,
               double a, int sign, int epi, const int *done, int fin) {
    return dslash_one(c, red_args(c), prec, out, out_par, in, xin, a, sign, epi, done, fin);
}

int dslash_one(wl_ctx *c, const RedArgs &ra, int prec, void *out, int out_par, const void *in, const void *xin,
               double a, int sign, int epi, const int *done, int fin) {
    DslashLaunch L{};
    L.out = out;
    L.in = in;
    L.xin = xin;
    L.gauge = gauge_for(c, prec);
    L.lf = link_format(c, prec);
    if (!L.gauge) return fail(WL_E_STATE, "links of precision %d not prepared (ensure_gauge)", prec);
    L.a = a;
    L.b = 1.0 + c->op_shift;
    L.out_par = out_par;
    L.sign = sign;
    L.epi = epi;
    L.done = done;
    L.red = ra;
    L.fin = fin;
    HIPCHK(launch_dslash(prec, c->g, L, c->stream));
    return WL_OK;
}

// M_s in -> out for even FP64/FP32 fields (one rhs), through the odd temporary tmp:
// tmp_o = -kappa H_oe in  (Dirac_Wilson_hop_set(..., ODD, sign), wilson.h:275,285)
// out_e = in + kappa H_eo tmp_o  (Dirac_Wilson_hop(-kappa, out, out, EVEN, sign), :277,287)
int meo_one(wl_ctx *c, const RedArgs &ra, int prec, void *out, const void *in, void *tmp, double kappa,
            int sign, int epi2, const int *done, int fin2) {
    int rc = dslash_one(c, ra, prec, tmp, 1, in, nullptr, -kappa, sign, EPI_SET, done, FIN_NONE);
    if (rc) return rc;
    return dslash_one(c, ra, prec, out, 0, tmp, in, kappa, sign, epi2, done, fin2);
}

int meo_one(wl_ctx *c, int prec, void *out, const void *in, void *tmp, double kappa, int sign,
            int epi2, const int *done, int fin2) {
    return meo_one(c, red_args(c), prec, out, in, tmp, kappa, sign, epi2, done, fin2);
}

int check_sign(int sign) { return (sign == 1 || sign == -1) ? WL_OK : fail(WL_E_ARG, "sign must be +-1"); }

int ensure_partials(wl_ctx *c, size_t n) {
    if (c->partials_cap >= n) return WL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_free(c, c->partials, c->partials_cap * sizeof(double));
    c->partials = nullptr;
    c->partials_cap = 0;
    if (int rc = dev_alloc(c, (void **)&c->partials, n * sizeof(double), "reduction buffer")) return rc;
    c->partials_cap = n;
    return WL_OK;
}

int ensure_st_mr(wl_ctx *c, int N) {
    if (c->st_mr_cap >= N) return WL_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    dev_free(c, c->st_mr, c->st_mr_cap * sizeof(CgState));
    if (c->st_mr_host) (void)hipHostFree(c->st_mr_host);
    c->st_mr = nullptr;
    c->st_mr_host = nullptr;
    c->st_mr_cap = 0;
    if (int rc = dev_alloc(c, (void **)&c->st_mr, N * sizeof(CgState), "CG states")) return rc;
    if (hipHostMalloc(&c->st_mr_host, N * sizeof(CgState), hipHostMallocDefault) != hipSuccess) {
        c->st_mr_host = nullptr;
        return fail(WL_E_OOM, "CG states for %d right-hand sides", N);
    }
    c->st_mr_cap = N;
    return WL_OK;
}

int work_field(wl_ctx *c, int slot, int prec, int parity, int nrhs, wl_field **out) {
    const auto key = std::make_tuple(slot, prec, parity);
    auto it = c->work.find(key);
    if (it != c->work.end()) {
        if (it->second->nrhs >= nrhs) {
            *out = it->second;
            return WL_OK;
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        field_free(it->second);
        c->work.erase(it);
    }
    int rc = field_new(c, parity, prec, nrhs, out);
    if (rc == WL_OK) c->work[key] = *out;
    return rc;
}

// Right-hand sides per lane of the multi-rhs kernel (register blocking): at most 2 in FP64
// (3-4 spill) and 4 with FP32 arithmetic.  Defaults, from wl_autotune over {16^4, 24^3x48,
// 32^4} x 1-12 rhs on MI355X: FP16 one (the packed arithmetic leaves registers for occupancy,
// which the gathers need more than link reuse) up to 3 x 2^20 site-rhs per launch, two beyond;
// FP32 two on lattices of 2^17 sites per parity or more from 4 rhs (from 2 at 2^19 or more),
// else one; FP64 two, except one for 2-3 rhs below 2^17 sites per parity.
int rhs_per_lane(const wl_ctx *c, int prec, int N) {
    const int cap = prec == WL_F64 ? 2 : 4;
    int nr = c->rhs_block;
    if (nr <= 0) {
        const int vh = c->g.vh;
        if (prec == WL_F16)
            nr = (int64_t)N * vh >= 3 * (int64_t(1) << 20) ? 2 : 1;
        else if (prec == WL_F32)
            nr = (N >= 4 && vh >= (1 << 17)) || (N == 2 && vh >= (1 << 19)) ? 2 : 1;
        else
            nr = N <= 3 && vh < (1 << 17) ? 1 : 2;
    }
    return std::max(1, std::min(nr, std::min(N, cap)));
}

// Right-hand sides per launch (a multiple of the per-lane count, or all of them): each launch
// is one pass over the lattice, so a smaller group keeps the neighbour-spinor footprint of a
// segment small enough for the XCD L2, at one extra link read per pass.
int rhs_per_launch_of(int want, int nr, int N) {
    if (want >= N) return N;  // one pass; its last chunk may be partly filled
    const int nl = std::max(nr, want / nr * nr);
    return std::min(nl, N);
}

int rhs_per_launch(const wl_ctx *c, int nr, int N) {
    return rhs_per_launch_of(c->rhs_launch > 0 ? c->rhs_launch : N, nr, N);
}

// Geometry of one multi-rhs hopping launch of nb rhs, nr per lane.
// Occupancy cap on large launches: FP64 launches of four or more rhs, or of two or more on
// lattices of 2^17 sites per parity or more, hold every workgroup to 32 KB of LDS per 64
// lanes (at most five 64-lane workgroups, 320 sites, per CU); FP32 launches of four or more
// rhs at three or four per lane, on lattices of 2^17 sites per parity or more, to 24 KB (six).
// Capped launches use 64-lane workgroups unless the caller chose a size.  At full occupancy
// the neighbour spinors of the sites in flight overflow the 4 MB L2 of an XCD and are fetched
// from HBM several times per hop.  FP16, and FP32 at one or two per lane, do not gain.
// Applies only while the caller has set no LDS reserve and the links are not staged in LDS.
// Uncapped FP32 and FP16 launches of two or more rhs default to 256-lane workgroups.
Geom mr_geom(const wl_ctx *c, int prec, int nb, int nr, bool lds_links) {
    Geom g = c->g;
    int per64 = 0;  // LDS bytes per 64 lanes
    if (c->lds_auto && !lds_links) {
        if (prec == WL_F64 && (nb >= 4 || (nb >= 2 && g.vh >= (1 << 17)))) per64 = 32768;
        if (prec == WL_F32 && nr >= 3 && nb >= 4 && g.vh >= (1 << 17)) per64 = 24576;
    }
    if (per64) {
        if (c->blk_auto) g.blk = 64;
        g.lds = std::min(65536, per64 * (g.blk / 64));
    } else if (c->blk_auto && prec != WL_F64 && nb >= 2) {
        g.blk = 256;
    }
    return g;
}

int dslash_mr_all(wl_ctx *c, int prec, void *out, int out_par, const void *in, const void *xin,
                  double a, int sign, int epi, int N, const int *all_done, CgState *st, void *x,
                  const void *p, void *r, int *pstride_out) {
    const int pstride = mr_segments(c->g);
    if (epi == EPI_AXPY_NORM || epi == EPI_CGUPD) {
        int rc = ensure_partials(c, (size_t)N * pstride);
        if (rc) return rc;
    }
    if (!gauge_for(c, prec)) return fail(WL_E_STATE, "links of precision %d not prepared (ensure_gauge)", prec);
    const int nr = rhs_per_lane(c, prec, N);
    const int nl = rhs_per_launch(c, nr, N);
    for (int j0 = 0; j0 < N; j0 += nl) {
        MrLaunch L{};
        L.out = out;
        L.in = in;
        L.xin = xin;
        L.gauge = gauge_for(c, prec);
        L.lf = link_format(c, prec);
        L.a = a;
        L.b = 1.0 + c->op_shift;
        L.out_par = out_par;
        L.sign = sign;
        L.epi = epi;
        L.j0 = j0;
        L.NB = std::min(nl, N - j0);
        L.NR = nr;
        // one rhs chunk per lane: nobody reads the link again, stream it; several chunks: the
        // sibling waves of a segment re-read it from the caches
        L.link_nt = c->link_load == 2 || (c->link_load == 0 && L.NB <= nr);
        L.lds_links = c->link_load == 3;
        L.pipe = c->pipeline == 0;
        L.all_done = all_done;
        L.st = st;
        L.x = x;
        L.p = p;
        L.r = r;
        L.partials = c->partials;
        HIPCHK(launch_dslash_mr(prec, mr_geom(c, prec, L.NB, L.NR, L.lds_links), L, c->stream));
    }
    if (pstride_out) *pstride_out = pstride;
    return WL_OK;
}

int meo_mr(wl_ctx *c, int prec, void *out, const void *in, void *t, double kappa, int sign,
           int epi2, int N, const int *all_done, CgState *st, int *pstride_out) {
    int rc = dslash_mr_all(c, prec, t, 1, in, nullptr, -kappa, sign, EPI_SET, N, all_done, st,
                           nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    return dslash_mr_all(c, prec, out, 0, t, in, kappa, sign, epi2, N, all_done, st, nullptr,
                         nullptr, nullptr, pstride_out);
}

int field_upload_async(wl_ctx *c, void *data, int prec, int N, int j, const double *host) {
    const size_t bytes = (size_t)c->g.vh * 24 * sizeof(double);
    if (int rc = ensure_staging(c, bytes)) return rc;
    HIPCHK(hipMemcpyAsync(c->staging, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_aos_to_soa(prec, c->g, (const double *)c->staging, data, N, j, c->stream));
    return WL_OK;
}

int field_download_async(wl_ctx *c, const void *data, int prec, int N, int j, double *host) {
    const size_t bytes = (size_t)c->g.vh * 24 * sizeof(double);
    if (int rc = ensure_staging(c, bytes)) return rc;
    HIPCHK(launch_soa_to_aos(prec, c->g, data, (double *)c->staging, N, j, c->stream));
    HIPCHK(hipMemcpyAsync(host, c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    return WL_OK;
}

int gauge_upload_async(wl_ctx *c, const double *const U[4]) {
    if (!c->gauge64) {
        if (int rc = dev_alloc(c, &c->gauge64, gauge_elems(c->g) * 16, "gauge field")) return rc;
        HIPCHK(hipMemsetAsync(c->gauge64, 0, gauge_elems(c->g) * 16, c->stream));
    }
    const size_t bytes = (size_t)c->V * 18 * sizeof(double);
    if (int rc = ensure_staging(c, bytes)) return rc;
    for (int mu = 0; mu < 4; ++mu) {
        HIPCHK(hipMemcpyAsync(c->staging, U[mu], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_gauge_to_soa(c->g, mu, (const double *)c->staging, c->gauge64, c->stream));
    }
    c->gauge_ok = true;
    c->gauge32_ok = false;
    c->gauge16_ok = false;
    links_changed(c);
    return WL_OK;
}

namespace {

// Default block padding: 128 sites when the unpadded stride is a multiple of 2^17 sites
// (lattices such as 32^4), else none.  Unpadded, the rhs blocks of a 32^4 FP64 field sit
// 96 MiB = 3 x 2^25 B apart.  Measured on MI355X with the site-tiled layout, interleaved: the
// 12-rhs hop 0.7-1.3% faster with the pad (FP64 81.4 -> 80.8 us, FP16 21.45 -> 21.18 us per
// rhs), the single-rhs hop within 0.3%.  Other lattices (24^3x48, 48^3x96) timed within 1% for
// every pad from 0 to 256 sites.
int default_plane_pad(int base) { return base % (1 << 17) == 0 ? 128 : 0; }

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

// Block stride S in sites (a field's rhs block holds planes x S elements, site-tiled): the
// parity's sites rounded up to a workgroup, plus `pad` sites (pad < 0: default_plane_pad).
// Always a multiple of the 64-site tile.
int plane_stride(int vh, int pad) {
    const int base = (vh + kBlock - 1) / kBlock * kBlock;
    return base + (pad < 0 ? default_plane_pad(base) : pad);
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
}  // namespace wl

using namespace wl;

// =================================================================== library ============

extern "C" const char *wl_version(void) { return "hila_wilson 0.2 (gfx950)"; }
extern "C" const char *wl_last_error(void) { return last_error(); }

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
    for (int i = 0; i < 3; ++i) c->g.dm[i] = div_multiplier((unsigned)dims[i]);
    c->g.vh = (int)(V / 2);
    c->g.stride = plane_stride(c->g.vh, -1);
    c->g.antiperiodic = c->antiperiodic;
    c->g.blk = dslash_block_auto(c->g.vh);
    c->g.dir = dir_auto(c->g.vh);
    (void)set_chunk(c->g, -1);
    DeviceGuard dg(device);
    int rc = WL_OK;
    // one partial-sum slot per workgroup of the largest grid (64-lane workgroups)
    const int nb = (c->g.vh + 63) / 64;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        c->stream = nullptr;
        rc = fail(WL_E_DEVICE, "stream creation failed");
    } else if ((rc = dev_alloc(c, (void **)&c->partials, sizeof(double) * (nb + 64), "context")) ||
               (rc = dev_alloc(c, (void **)&c->flags_d, 64, "context")) ||
               (rc = dev_alloc(c, (void **)&c->st, sizeof(CgState), "context")) ||
               (rc = dev_alloc(c, (void **)&c->scratch_d, 256, "context"))) {
    } else if (hipHostMalloc(&c->flags_h, 64, hipHostMallocDefault) != hipSuccess ||
               hipHostMalloc(&c->st_host, sizeof(CgState), hipHostMallocDefault) != hipSuccess ||
               hipHostMalloc(&c->scratch_h, 256, hipHostMallocDefault) != hipSuccess) {
        rc = fail(WL_E_OOM, "pinned host allocation failed");
    } else if (hipMemsetAsync(c->st, 0, sizeof(CgState), c->stream) != hipSuccess ||
               hipMemsetAsync(c->flags_d, 0, 64, c->stream) != hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = fail(WL_E_DEVICE, "context initialisation failed");
    } else {
        rc = noalloc_setup(device);
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
    if (c->st_mr_host) (void)hipHostFree(c->st_mr_host);
    if (c->flags_h) (void)hipHostFree(c->flags_h);
    // fields the caller still holds: release their device memory and orphan the handles
    for (wl_field *f : c->user_fields) {
        if (f->data) (void)hipFree(f->data);
        f->data = nullptr;
        f->ctx = nullptr;
    }
    c->user_fields.clear();
    for (void *p : {(void *)c->st_mr, (void *)c->flags_d, (void *)c->gauge64, c->gauge32, c->gauge16, c->rows64,
                    c->rows32, c->staging, (void *)c->partials, (void *)c->st, (void *)c->scratch_d, c->force64,
                    c->mom64, c->grid_bar})
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
    if (int rc = gauge_upload_async(c, U)) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}

// ==================================================================== fields ============

extern "C" int wl_field_create(wl_ctx *c, int parity, int prec, int nrhs, wl_field **out) {
    if (!c || !out) return fail(WL_E_ARG, "null argument");
    if (parity != WL_EVEN && parity != WL_ODD) return fail(WL_E_ARG, "parity must be WL_EVEN or WL_ODD");
    if (prec != WL_F64 && prec != WL_F32 && prec != WL_F16) return fail(WL_E_NOTIMPL, "field precision %d", prec);
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
    if (int rc = field_upload_async(c, f->data, f->prec, f->nrhs, rhs, host)) return rc;
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
    if (int rc = field_download_async(c, f->data, f->prec, f->nrhs, rhs, host)) return rc;
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

extern "C" int wl_field_copy(wl_field *dst, const wl_field *src) {
    if (!dst || !src) return fail(WL_E_ARG, "null argument");
    if (!dst->ctx || !src->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    if (dst->ctx != src->ctx) return fail(WL_E_ARG, "fields belong to different contexts");
    if (dst->parity != src->parity || dst->nrhs != src->nrhs)
        return fail(WL_E_ARG, "fields differ in parity or nrhs");
    if (dst == src) return WL_OK;
    wl_ctx *c = dst->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    if (dst->prec == src->prec)
        HIPCHK(launch_copy(src->prec, c->g, src->data, dst->data, src->nrhs, c->stream));
    else
        HIPCHK(launch_convert(src->prec, dst->prec, c->g, src->data, dst->data, src->nrhs, c->stream));
    return WL_OK;
}

extern "C" int wl_field_copy_rhs(wl_field *dst, int jd, const wl_field *src, int js) {
    if (!dst || !src) return fail(WL_E_ARG, "null argument");
    if (!dst->ctx || !src->ctx) return fail(WL_E_STATE, "the field's context was destroyed");
    if (dst->ctx != src->ctx) return fail(WL_E_ARG, "fields belong to different contexts");
    if (dst->parity != src->parity || dst->prec != src->prec)
        return fail(WL_E_ARG, "fields differ in parity or precision");
    if (jd < 0 || jd >= dst->nrhs || js < 0 || js >= src->nrhs)
        return fail(WL_E_ARG, "rhs %d of %d <- rhs %d of %d out of range", jd, dst->nrhs, js, src->nrhs);
    if (dst == src && jd == js) return WL_OK;
    wl_ctx *c = dst->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    const size_t bytes = dst->rhs_elems * prec_bytes(dst->prec);  // one rhs is one contiguous block
    HIPCHK(hipMemcpyAsync((char *)dst->data + (size_t)jd * bytes, (const char *)src->data + (size_t)js * bytes, bytes,
                          hipMemcpyDeviceToDevice, c->stream));
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
    if (in->nrhs == 1 && in->prec != WL_F16)  // FP16 fields: the multi-rhs kernel, any count
        return dslash_one(c, in->prec, out->data, out_par, in->data, nullptr, -kappa, sign, EPI_SET,
                          nullptr, FIN_NONE);
    return dslash_mr_all(c, in->prec, out->data, out_par, in->data, nullptr, -kappa, sign, EPI_SET,
                         in->nrhs, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

namespace {

// M_sign + mh on every rhs of in (Dirac_Wilson_evenodd::apply/dagger, wilson.h:272-289, with
// Hasenbusch_operator's shift, Hasenbusch.h:40-48)
int meo_apply(wl_ctx *c, wl_field *out, const wl_field *in, double kappa, double mh, int sign) {
    if (!c || !out || !in) return fail(WL_E_ARG, "null argument");
    if (check_sign(sign)) return WL_E_ARG;
    if (out->ctx != c || in->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (out->parity != WL_EVEN || in->parity != WL_EVEN) return fail(WL_E_ARG, "M acts on even fields");
    if (out == in) return fail(WL_E_ARG, "out and in must be different fields");
    if (out->prec != in->prec || out->nrhs != in->nrhs)
        return fail(WL_E_ARG, "out and in differ in precision or nrhs");
    if (!std::isfinite(mh)) return fail(WL_E_ARG, "the shift must be finite");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    OpShift os(c, mh);
    int rc = ensure_gauge(c, in->prec);
    if (rc) return rc;
    if (in->nrhs == 1 && in->prec != WL_F16) {
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

}  // namespace

extern "C" int wl_meo_apply(wl_ctx *c, wl_field *out, const wl_field *in, double kappa, int sign) {
    return meo_apply(c, out, in, kappa, 0.0, sign);
}

extern "C" int wl_hasenbusch_apply(wl_ctx *c, wl_field *out, const wl_field *in, double kappa, double mh,
                                   int sign) {
    return meo_apply(c, out, in, kappa, mh, sign);
}

// ================================================================ diagnostics ===========

extern "C" int wl_time_op(wl_ctx *c, int what, wl_field *out, const wl_field *in, double kappa,
                          int sign, int warmup, int iters, double *ms_per_launch) {
    if (!c || !out || !in || !ms_per_launch || iters < 1 || warmup < 0)
        return fail(WL_E_ARG, "bad argument");
    auto once = [&]() -> int {
        if (what == 2) return wl_field_copy(out, in);
        return what == 0 ? wl_dslash_eo(c, out, in, kappa, sign) : wl_meo_apply(c, out, in, kappa, sign);
    };
    if (what < 0 || what > 2) return fail(WL_E_ARG, "what must be 0 (hopping term), 1 (M) or 2 (copy)");
    if (what == 2 && (out->ctx != c || in->ctx != c)) return fail(WL_E_ARG, "fields of another context");
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

extern "C" int wl_set_option(wl_ctx *c, int option, int value) {
    if (!c) return fail(WL_E_ARG, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    switch (option) {
    case WL_OPT_TRAVERSAL_CHUNK: return set_chunk(c->g, value);
    case WL_OPT_BLOCK:
        if (value != 0 && value != 64 && value != 128 && value != 256)
            return fail(WL_E_ARG, "workgroup size must be 64, 128 or 256");
        c->blk_auto = value == 0;
        c->g.blk = value == 0 ? dslash_block_auto(c->g.vh) : value;
        return WL_OK;
    case WL_OPT_RHS_BLOCK:
        if (value < 0 || value > 4) return fail(WL_E_ARG, "rhs per lane must be 0 (automatic) .. 4");
        c->rhs_block = value;
        return WL_OK;
    case WL_OPT_RHS_LAUNCH:
        if (value < 0 || value > 64) return fail(WL_E_ARG, "rhs per launch must be 0 (all) .. 64");
        c->rhs_launch = value;
        return WL_OK;
    case WL_OPT_LINK_LOAD:
        if (value < 0 || value > 3) return fail(WL_E_ARG, "link load mode must be 0..3");
        c->link_load = value;
        return WL_OK;
    case WL_OPT_LDS_RESERVE:
        if (value < -1 || value > 65536)
            return fail(WL_E_ARG, "LDS reserve must be -1 (none), 0 (automatic) or 1..65536 bytes");
        set_lds_reserve(c, value);
        return WL_OK;
    case WL_OPT_GRAPHS:
        if (value != 0 && value != 1) return fail(WL_E_ARG, "graphs option must be 0 or 1");
        c->use_graphs = value == 1;
        return WL_OK;
    case WL_OPT_BLOCK_KRYLOV:
        if (value != 0 && value != 1) return fail(WL_E_ARG, "block Krylov option must be 0 or 1");
        c->block_krylov = value == 1;
        return WL_OK;
    case WL_OPT_SMALL_HOP:
        if (value < -1 || value > 1)
            return fail(WL_E_ARG, "small-lattice hop option must be -1 (never), 0 (automatic) or 1 (always)");
        c->small_hop = value;
        c->g.dir = value == 0 ? dir_auto(c->g.vh) : (value > 0 ? 1 : 0);
        return WL_OK;
    case WL_OPT_CG_PERSISTENT:
        if (value < -1 || value > 1)
            return fail(WL_E_ARG, "persistent CG option must be -1 (never), 0 (automatic) or 1 (always)");
        c->cg_persistent = value;
        return WL_OK;
    case WL_OPT_PIPELINE:
        if (value != 0 && value != 1) return fail(WL_E_ARG, "pipeline option must be 0 or 1");
        c->pipeline = value;
        return WL_OK;
    case WL_OPT_LINK_ROWS: {
        if (value != 2 && value != 3) return fail(WL_E_ARG, "link rows must be 3 (full links) or 2 (two-row links)");
        const int prev = c->link_rows;
        c->link_rows = value;
        if (value == 2 && c->gauge_ok) {  // check the links now, not at the next operator call
            DeviceGuard dg(c->device);
            if (int rc = ensure_rows(c, WL_F64)) {
                c->link_rows = prev;
                return rc;
            }
        }
        return WL_OK;
    }
    case WL_OPT_PLANE_PAD: {
        if (value != -1 && (value < 0 || value > 65536 || value % 64 != 0))
            return fail(WL_E_ARG, "plane padding must be -1 or a multiple of 64 in 0..65536");
        const bool empty = !c->gauge64 && c->user_fields.empty() && c->work.empty() && !c->tmp_odd64 &&
                           !c->tmp_odd32 && !c->cg_r && !c->cg_p && !c->cg_Dp && !c->cg_DDp &&
                           !c->mx_r && !c->mx_p && !c->mx_Dp;
        if (!empty) return fail(WL_E_STATE, "plane padding must be set before any field or gauge exists");
        c->g.stride = plane_stride(c->g.vh, value);
        return WL_OK;
    }
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
    case WL_OPT_PLANE_PAD: *value = c->g.stride - plane_stride(c->g.vh, 0); return WL_OK;
    case WL_OPT_LINK_LOAD: *value = c->link_load; return WL_OK;
    case WL_OPT_LDS_RESERVE: *value = lds_reserve_of(c); return WL_OK;
    case WL_OPT_GRAPHS: *value = c->use_graphs ? 1 : 0; return WL_OK;
    case WL_OPT_PIPELINE: *value = c->pipeline; return WL_OK;
    case WL_OPT_LINK_ROWS: *value = c->link_rows; return WL_OK;
    case WL_OPT_CG_PERSISTENT: *value = c->cg_persistent; return WL_OK;
    case WL_OPT_SMALL_HOP: *value = c->small_hop; return WL_OK;
    case WL_OPT_BLOCK_KRYLOV: *value = c->block_krylov ? 1 : 0; return WL_OK;
    default: return fail(WL_E_ARG, "unknown option %d", option);
    }
}

// ================================================================== autotuner ===========

namespace {

// A tuning holds for one lattice size, precision and rhs count, and for the link format and
// link-load and pipelining modes it was timed with (they change the kernels' speed).
struct TuneKey {
    int L[4], prec, nrhs, link_format, link_load, pipeline;
    bool operator<(const TuneKey &o) const {
        return std::tie(L[0], L[1], L[2], L[3], prec, nrhs, link_format, link_load, pipeline) <
               std::tie(o.L[0], o.L[1], o.L[2], o.L[3], o.prec, o.nrhs, o.link_format, o.link_load, o.pipeline);
    }
};
TuneKey tune_key(const wl_ctx *c, int prec, int nrhs) {
    return TuneKey{{c->g.L[0], c->g.L[1], c->g.L[2], c->g.L[3]}, prec, nrhs, link_format(c, prec), c->link_load,
                   c->pipeline};
}
std::mutex g_tune_mu;
std::map<TuneKey, wl_tune_result> g_tuned;

// Average device time of `iters` hopping launches over all N rhs of `in` (after 2 warm-ups).
int time_hop_mr(wl_ctx *c, int prec, void *out, const void *in, int N, int iters, hipEvent_t e0, hipEvent_t e1,
                double *us) {
    for (int i = 0; i < 2 + iters; ++i) {
        if (i == 2) HIPCHK(hipEventRecord(e0, c->stream));
        if (int rc = dslash_mr_all(c, prec, out, 1, in, nullptr, -0.125, 1, EPI_SET, N, nullptr, nullptr, nullptr,
                                   nullptr, nullptr, nullptr))
            return rc;
    }
    HIPCHK(hipEventRecord(e1, c->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    *us = 1e3 * ms / iters;
    return WL_OK;
}

}  // namespace

namespace wl {
bool tuned_config(const wl_ctx *c, int prec, int nrhs, wl_tune_result *r) {
    const TuneKey key = tune_key(c, prec, nrhs);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) return false;
    *r = it->second;
    return true;
}

void set_lds_reserve(wl_ctx *c, int value) {
    c->lds_auto = value == 0;
    c->g.lds = value > 0 ? value : 0;
}

int lds_reserve_of(const wl_ctx *c) { return c->lds_auto ? 0 : c->g.lds > 0 ? c->g.lds : -1; }

void apply_hop_config(wl_ctx *c, const wl_tune_result &r) {
    c->rhs_block = r.rhs_block;
    c->rhs_launch = r.rhs_launch;
    c->blk_auto = r.block == 0;
    c->g.blk = r.block == 0 ? dslash_block_auto(c->g.vh) : r.block;
    set_lds_reserve(c, r.lds_reserve);
}
}  // namespace wl

extern "C" int wl_autotune(wl_ctx *c, int prec, int nrhs, wl_tune_result *out) {
    if (!c || !out) return fail(WL_E_ARG, "null argument");
    if (prec != WL_F64 && prec != WL_F32 && prec != WL_F16) return fail(WL_E_NOTIMPL, "precision %d", prec);
    if (nrhs < 1 || nrhs > 64) return fail(WL_E_ARG, "nrhs %d out of range 1..64", nrhs);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc = ensure_gauge(c, prec);
    if (rc) return rc;
    const TuneKey key = tune_key(c, prec, nrhs);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            *out = it->second;
            out->candidates = 0;
            apply_hop_config(c, *out);
            return WL_OK;
        }
    }
    // scratch input (zeros: the kernel does the same work on any data) and output
    wl_field *fin = nullptr, *fout = nullptr;
    if ((rc = field_new(c, WL_EVEN, prec, nrhs, &fin))) return rc;
    if ((rc = field_new(c, WL_ODD, prec, nrhs, &fout))) {
        field_free(fin);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    wl_tune_result saved{};
    saved.rhs_block = c->rhs_block;
    saved.rhs_launch = c->rhs_launch;
    saved.block = c->g.blk;
    saved.lds_reserve = lds_reserve_of(c);
    const bool saved_blk_auto = c->blk_auto;
    auto restore = [&]() {
        apply_hop_config(c, saved);
        c->blk_auto = saved_blk_auto;
    };
    wl_tune_result best{};
    best.us_per_rhs = 1e300;
    int count = 0;
    const int iters = c->g.vh >= (1 << 18) ? 5 : 20;
    auto cleanup = [&]() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(c->stream);
        field_free(fin);
        field_free(fout);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        cleanup();
        return fail(WL_E_DEVICE, "event creation failed");
    }
    // warm-up: the default configuration (timed again against the choice at the end)
    c->rhs_block = 0;
    c->rhs_launch = 0;
    c->g.blk = dslash_block_auto(c->g.vh);
    c->blk_auto = true;
    set_lds_reserve(c, 0);
    double us_warm = 0;
    if ((rc = time_hop_mr(c, prec, fout->data, fin->data, nrhs, iters, e0, e1, &us_warm))) {
        restore();
        cleanup();
        return rc;
    }
    // LDS reserve: automatic (the occupancy cap where it applies), in FP64 and FP32 also none,
    // and a reserve the caller set
    std::vector<int> reserves{0};
    if (prec != WL_F16) reserves.push_back(-1);
    if (saved.lds_reserve > 0) reserves.push_back(saved.lds_reserve);
    const int nr_cap = prec == WL_F64 ? 2 : 4;
    static const int launches[] = {1, 2, 4, 6, 12, 16, 24, 32, 48, 64};
    for (int nr = 1; nr <= std::min(nr_cap, nrhs); nr *= 2) {
        std::vector<int> nls;  // effective rhs per launch (multiples of nr, or all)
        for (int want : launches) {
            if (want > nrhs) want = nrhs;
            const int nl = rhs_per_launch_of(want, nr, nrhs);
            if (std::find(nls.begin(), nls.end(), nl) == nls.end()) nls.push_back(nl);
            if (want == nrhs) break;
        }
        for (int nl : nls)
            for (int blk : {64, 128, 256})
                for (int lds : reserves) {
                    wl_tune_result cand{};
                    cand.rhs_block = nr;
                    cand.rhs_launch = nl;
                    cand.block = blk;
                    cand.lds_reserve = lds;
                    apply_hop_config(c, cand);
                    double us = 0;
                    if ((rc = time_hop_mr(c, prec, fout->data, fin->data, nrhs, iters, e0, e1, &us))) {
                        restore();
                        cleanup();
                        return rc;
                    }
                    ++count;
                    if (us < best.us_per_rhs * nrhs) {
                        cand.us_per_rhs = us / nrhs;
                        best = cand;
                    }
                }
    }
    // The search times each candidate once, briefly: time the choice and the default
    // configuration again, alternately, and keep the default unless the choice beats it.
    wl_tune_result def{};  // rhs per lane, per launch and workgroup automatic, LDS automatic
    std::vector<double> t_best, t_def;
    for (int rep = 0; rep < 3; ++rep)
        for (const wl_tune_result *cfg : {&best, &def}) {
            apply_hop_config(c, *cfg);
            double us = 0;
            if ((rc = time_hop_mr(c, prec, fout->data, fin->data, nrhs, 2 * iters, e0, e1, &us))) {
                restore();
                cleanup();
                return rc;
            }
            (cfg == &best ? t_best : t_def).push_back(us / nrhs);
        }
    cleanup();
    std::sort(t_best.begin(), t_best.end());
    std::sort(t_def.begin(), t_def.end());
    if (t_def[1] <= t_best[1]) {
        def.us_per_rhs = t_def[1];
        best = def;
    } else {
        best.us_per_rhs = t_best[1];
    }
    best.candidates = count;
    best.us_per_rhs_default = t_def[1];
    apply_hop_config(c, best);
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tuned.emplace(key, best);
    }
    *out = best;
    return WL_OK;
}
