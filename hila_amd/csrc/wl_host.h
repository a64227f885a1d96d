// Host-side helpers shared by the translation units of libhila_wilson.so that implement the C
// ABI: error reporting, device selection, device-memory accounting, work fields and the
// operator building blocks (one or many right-hand sides) that the solvers compose.
#pragma once

#include <cstdarg>

#include "wl_internal.h"

namespace wl {

// Store a formatted message as the calling thread's last error and return `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

#define HIPCHK(expr)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return ::wl::fail(WL_E_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                                              \
    } while (0)

// Make a device current for the calling thread; restore the previous one on exit.
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

// The operator shift (Hasenbusch_operator's h_parameter, Hasenbusch.h:40-48) that the
// launches of one entry-point call apply: every second half of M computes
// (1 + mh) in + kappa H_eo tmp, i.e. M + mh.  Reset to 0 when the call returns.
struct OpShift {
    wl_ctx *c;
    OpShift(wl_ctx *ctx, double mh) : c(ctx) { c->op_shift = mh; }
    ~OpShift() { c->op_shift = 0.0; }
    OpShift(const OpShift &) = delete;
    OpShift &operator=(const OpShift &) = delete;
};

// Device allocations of a context, counted in wl_ctx::dev_bytes.
int dev_alloc(wl_ctx *c, void **p, size_t bytes, const char *what);
void dev_free(wl_ctx *c, void *p, size_t bytes);

int ensure_staging(wl_ctx *c, size_t bytes);
int field_new(wl_ctx *c, int parity, int prec, int nrhs, wl_field **out);
void field_free(wl_field *f);
int ensure_tmp(wl_ctx *c, int prec, wl_field **slot, int parity);
RedArgs red_args(wl_ctx *c);
const void *gauge_for(wl_ctx *c, int prec);
int link_format(const wl_ctx *c, int prec);
// the links were overwritten: two-row copies and their check are stale
inline void links_changed(wl_ctx *c) {
    c->rows64_ok = c->rows32_ok = false;
    c->rows_check = 0;
}
int ensure_gauge(wl_ctx *c, int prec);
int check_sign(int sign);

// Device work memory of the single-rhs FP64 CG: the even fields r, p, Dp, DDp, the odd
// temporary t of M, the reduction partials, the CG state, the grid barrier of the one-launch
// solve and the target of stored reductions.  Either the context's own work fields (created
// on first use, cg_work_of) or a fixed workspace.
struct CgWork {
    void *r, *p, *Dp, *DDp, *t;
    double *partials;
    size_t partials_cap;  // doubles
    CgState *st;
    void *bar;            // GridBar of the one-launch solve
    double *red_out;
};
int cg_work_of(wl_ctx *c, CgWork *w);
// Partial sums one pass of the single-rhs CG writes at most: one per hopping workgroup or
// per 256-site BLAS workgroup.
size_t cg_partials_needed(const wl_ctx *c);
RedArgs red_args(const CgWork &w);
// The single-rhs FP64 CG (conjugate_gradient.h:46-104) on work memory w: in one cooperative
// launch when WL_OPT_CG_PERSISTENT asks for it and the launch fits, else kernel by kernel.
// x_bytes: bytes of the solution field (cleared for b = 0).  Makes no device allocation.
int cg_f64_run(wl_ctx *c, const CgWork &w, void *x, size_t x_bytes, const void *b, double kappa,
               double accuracy, int maxiters, wl_cg_stats *stats);
// The single-rhs FP64 CG in one cooperative launch (wl_cgp.hip), when WL_OPT_CG_PERSISTENT asks
// for it; *ran = false, with nothing done, when the launch does not fit the device.
bool cg_persistent_wanted(const wl_ctx *c);
int cg_f64_persistent(wl_ctx *c, const CgWork &w, void *x, size_t x_bytes, const void *b, double kappa,
                      double accuracy, int maxiters, wl_cg_stats *stats, bool *ran);
// Device address of the fixed workspace of wl_cg_solve_noalloc on `device` (wl_noalloc.hip),
// looked up once per device with that device current; wl_create calls it.
int noalloc_setup(int device);
// words 32-bit words at p = 0, by a kernel (no memset node if a caller captures the stream)
hipError_t launch_zero_words(void *p, int words, hipStream_t s);
int ensure_partials(wl_ctx *c, size_t n);
int ensure_st_mr(wl_ctx *c, int N);

// out_par = a * H_sign in (+ xin), one rhs; reductions through the context's partials, or
// through ra
int dslash_one(wl_ctx *c, int prec, void *out, int out_par, const void *in, const void *xin,
               double a, int sign, int epi, const int *done, int fin);
int dslash_one(wl_ctx *c, const RedArgs &ra, int prec, void *out, int out_par, const void *in, const void *xin,
               double a, int sign, int epi, const int *done, int fin);
// M_s in -> out (one rhs) through the odd temporary tmp
int meo_one(wl_ctx *c, int prec, void *out, const void *in, void *tmp, double kappa, int sign,
            int epi2, const int *done, int fin2);
int meo_one(wl_ctx *c, const RedArgs &ra, int prec, void *out, const void *in, void *tmp, double kappa,
            int sign, int epi2, const int *done, int fin2);

// Work fields of the solvers, one per (slot, precision, parity), created on first use with at
// least the requested number of right-hand sides and kept resident with the context.
enum WorkSlot { W_R = 0, W_P, W_DP, W_T, W_DP64, W_DDP64, W_T64, W_X, W_Y, W_SRC, W_BZ };
int work_field(wl_ctx *c, int slot, int prec, int parity, int nrhs, wl_field **out);

int rhs_per_lane(const wl_ctx *c, int prec, int N);
// rhs per launch for a requested count: a multiple of nr (at least nr), at most N
int rhs_per_launch_of(int want, int nr, int N);
int rhs_per_launch(const wl_ctx *c, int nr, int N);
// out = a H_sign in (+ epilogue) for N right-hand sides; per-rhs partial sums land in
// c->partials rows, *pstride_out apart.
int dslash_mr_all(wl_ctx *c, int prec, void *out, int out_par, const void *in, const void *xin,
                  double a, int sign, int epi, int N, const int *all_done, CgState *st, void *x,
                  const void *p, void *r, int *pstride_out);
// M_s for N right-hand sides (even fields) through the odd work field t
int meo_mr(wl_ctx *c, int prec, void *out, const void *in, void *t, double kappa, int sign,
           int epi2, int N, const int *all_done, CgState *st, int *pstride_out);

// The configuration wl_autotune chose in this process for c's lattice size, link format and
// link-load and pipelining modes, at (precision, nrhs), if any: rhs per lane, rhs per launch,
// workgroup size.
bool tuned_config(const wl_ctx *c, int prec, int nrhs, wl_tune_result *r);
void apply_hop_config(wl_ctx *c, const wl_tune_result &r);   // rhs block/launch, block, LDS
void set_lds_reserve(wl_ctx *c, int value);                  // WL_OPT_LDS_RESERVE semantics
int lds_reserve_of(const wl_ctx *c);                         // ... and its current value

// Upload / download of right-hand side j of an N-rhs field from / to host AoS spinors
// (V/2 x 24 doubles, EVEN_SITES_FIRST order within the parity), enqueued on the context's
// stream through its staging buffer.  The host buffer must stay valid until the stream has
// passed the copy.
int field_upload_async(wl_ctx *c, void *data, int prec, int N, int j, const double *host);
int field_download_async(wl_ctx *c, const void *data, int prec, int N, int j, double *host);
// Gauge upload (U[mu]: V x 18 doubles), enqueued on the context's stream.
int gauge_upload_async(wl_ctx *c, const double *const U[4]);
// Gauge field from a HILA configuration file (wl_config.hip); returns after the upload.
int config_upload(wl_ctx *c, const char *path);

}  // namespace wl
