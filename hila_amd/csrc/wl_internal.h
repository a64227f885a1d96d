// Host-side internals shared by the translation units of libhila_wilson.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <map>
#include <tuple>
#include <unordered_set>

#include "wl_device.h"
#include "wilson_mi355x.h"

namespace wl {

// Device layout of a Wilson-vector field of one parity: for every rhs, a block of 12 x stride
// complex numbers (spin-major, colour-minor components), site-tiled: component c of site k at
// comp_elem(c, stride, k) (wl_store.h).  Gauge field: for parity p and direction mu a block of
// 9 x stride complex numbers starting at ((p*4 + mu)*9)*stride, entry e = 3*row + col of the
// link U_mu(site k of parity p) at link_entry(e, stride, k) in it.
inline size_t field_elems(const Geom &g, int nrhs) { return (size_t)nrhs * 12 * g.stride; }
inline size_t gauge_elems(const Geom &g) { return (size_t)2 * 4 * 9 * g.stride; }

inline int prec_bytes(int prec) { return prec == WL_F64 ? 16 : (prec == WL_F32 ? 8 : 4); }

// Which epilogue a Dslash launch applies to a = coefficient times the hopping sum H in.
enum Epi : int {
    EPI_SET = 0,        // out = a H in                           (Dirac_Wilson_hop_set)
    EPI_AXPY = 1,       // out = xin + a H in                     (second half of M)
    EPI_AXPY_NORM = 2,  // out = xin + a H in, reduce ||out||^2   (M p with pDp fused)
    EPI_CGUPD = 3,      // DDp = xin + a H in; x += alpha p; r -= alpha DDp; reduce ||r||^2
};

struct DslashLaunch {
    void *out;
    const void *in;
    const void *xin;
    const void *gauge;
    double a;
    double b = 1.0;  // coefficient of xin in the epilogues that read it: out = b xin + a H in
    int lf = 9;      // stored entries per link of `gauge`: 9, or 6 (two-row links)
    int out_par;   // 0 even, 1 odd
    int sign;      // +1 apply, -1 dagger
    int epi;
    const int *done;  // skip the launch body when *done != 0 (CG loop); may be null
    // EPI_CGUPD operands (FP64 even fields)
    void *x;
    const void *p;
    void *r;
    RedArgs red;
    int fin;               // applied by a finish kernel after a norm epilogue; FIN_NONE: none
    // EPI_CGUPD of the launch-folded CG: alpha from fold_n partials (DslashArgs); null: none
    const double *fold_parts = nullptr;
    int fold_n = 0;
    int fold_it = 0;
};

// Kernel launchers (wl_dslash.hip / wl_blas.hip).  All enqueue on `s` and return hipError_t.
hipError_t launch_dslash(int prec, const Geom &g, const DslashLaunch &L, hipStream_t s);

// host AoS block <-> right-hand side j of an N-rhs field
hipError_t launch_aos_to_soa(int prec, const Geom &g, const double *aos, void *planes, int N,
                             int j, hipStream_t s);
hipError_t launch_soa_to_aos(int prec, const Geom &g, const void *planes, double *aos, int N,
                             int j, hipStream_t s);
hipError_t launch_gauge_to_soa(const Geom &g, int mu, const double *aos_full, void *gauge64,
                               hipStream_t s);
// two-row link copy (prec F64 or F32; rows may be null) and, with dev, the per-workgroup
// largest deviation of the stored third row from conj(row 0 x row 1) (two_rows_blocks values)
hipError_t launch_two_rows(const Geom &g, int prec, const void *gauge64, void *rows, double *dev, hipStream_t s);
int two_rows_blocks(const Geom &g);
inline size_t rows_elems(const Geom &g) { return (size_t)2 * 4 * 6 * g.stride; }
hipError_t launch_gauge_convert(const Geom &g, int prec, const void *gauge64, void *gauge_lo,
                                hipStream_t s);
// direction mu of link blocks (gauge layout, FP64) -> 18 doubles per site, sites in
// lexicographic order (lexicographic != 0) or EVEN_SITES_FIRST order (wl_config.hip)
hipError_t launch_links_to_host(const Geom &g, const void *planes, int mu, double *out, int lexicographic,
                                hipStream_t s);
// fermion force (wl_force.hip): F of Dirac_Wilson_evenodd::force from chi, psi (even) and
// tA = -kappa H_{-sign} chi, tB = -kappa H_sign psi (odd), into `out` planes (momentum == 0),
// or out += project_antihermitean(U coef F) (momentum != 0)
hipError_t launch_force(const Geom &g, int sign, int momentum, const void *chi, const void *psi, const void *tA,
                        const void *tB, const void *gauge, void *out, double kappa, double coef, hipStream_t s);
// y -= x for even FP64 fields of one rhs (wl_force.hip)
hipError_t launch_sub(const Geom &g, void *y, const void *x, hipStream_t s);
// sum over sites of ||f_j||^2 for rhs j of an N-rhs field (FP64 accumulate), then `fin`
hipError_t launch_norm2(int prec, const Geom &g, const void *f, int N, int j, const RedArgs &ra,
                        int fin, const int *done, hipStream_t s);
// r = b - DDp; p = r; reduce ||r||^2 -> FIN_RR   (conjugate_gradient.h:67-72)
hipError_t launch_cg_init(const Geom &g, const void *b, const void *DDp, void *r, void *p,
                          const RedArgs &ra, hipStream_t s);
// p = beta p + r, beta from the CG state (conjugate_gradient.h:93-94); skipped when done != 0
hipError_t launch_cg_xpay(int prec, const Geom &g, void *p, const void *r, const CgState *st,
                          hipStream_t s);
// the same in FP64 for iteration `it` of the launch-folded CG, with the finish of the n partials
// of ||r||^2 at `parts` folded in (convergence test, beta, state; k_cg_xpay_fold)
hipError_t launch_cg_xpay_fold(const Geom &g, void *p, const void *r, CgState *st, const double *parts, int n,
                               int it, hipStream_t s);
// reliable update: r_lo = b - DDp (FP64 in, inner precision out), ||.||^2 -> FIN_UPDATE
hipError_t launch_residual_update(int prec, const Geom &g, const void *b, const void *DDp,
                                  void *r_lo, const RedArgs &ra, hipStream_t s);
// reduce ||b - v||^2 with `fin`
hipError_t launch_diff_norm2(const Geom &g, const void *b, const void *v, const RedArgs &ra,
                             int fin, hipStream_t s);
// dst = src for whole fields of nrhs right-hand sides
hipError_t launch_copy(int prec, const Geom &g, const void *src, void *dst, int nrhs,
                       hipStream_t s);
// dst (storage pout) = src (storage pin != pout), site by site, for nrhs right-hand sides
hipError_t launch_convert(int pin, int pout, const Geom &g, const void *src, void *dst, int nrhs,
                          hipStream_t s);

// ---- multi-rhs (wl_multi.hip) ----
// One launch of the register-blocked multi-rhs hopping kernel: right-hand sides [j0, j0+NB),
// NR per lane.  Per-rhs partial sums (norm epilogues) go to partials[rhs * mr_segments + seg].
struct MrLaunch {
    void *out;
    const void *in;
    const void *xin;
    const void *gauge;
    double a;
    double b = 1.0;           // coefficient of xin (see DslashLaunch)
    int lf = 9;               // stored entries per link (see DslashLaunch)
    int out_par, sign, epi;
    int j0, NB, NR;
    bool link_nt;             // nontemporal link loads
    bool lds_links;           // links staged in LDS, one workgroup per segment (<= 8 rhs chunks)
    bool pipe;                // software-pipelined hops where the registers allow
    const int *all_done;      // may be null
    CgState *st;              // per-rhs states (skip finished rhs, alpha); may be null
    void *x;                  // EPI_CGUPD accumulator: FP64 x (FP64 arithmetic) or FP32 y
    const void *p;
    void *r;
    double *partials;
};
hipError_t launch_dslash_mr(int prec, const Geom &g, const MrLaunch &L, hipStream_t s);
int mr_segments(const Geom &g);
hipError_t launch_finish_mr(double *partials, int pstride, int n, int N, CgState *st,
                            double *out, int fin, const int *all_done, hipStream_t s);
// BLAS-1 kernels over N right-hand sides: per-rhs partials at partials[rhs * nblocks(g) + b]
hipError_t launch_norm2_mr(int prec, const Geom &g, const void *f, int N, double *partials,
                           hipStream_t s);
hipError_t launch_cg_init_mr(const Geom &g, const void *b, const void *ddp, void *r, void *p,
                             int N, double *partials, hipStream_t s);
hipError_t launch_residual_update_mr(int prec, const Geom &g, const void *b, const void *ddp,
                                     void *r, const CgState *st, int first, int N,
                                     double *partials, const int *gate, hipStream_t s);
hipError_t launch_xpay_mr(int prec, const Geom &g, void *p, const void *r, const CgState *st,
                          const int *all_done, int N, hipStream_t s);
// flags[0] all finished, flags[1] update pending, flags[2] no update pending; gate: skip if set
hipError_t launch_batch_sync(CgState *st, int N, int mixed, int *flags, const int *gate,
                             hipStream_t s);
hipError_t launch_zero_rhs(const Geom &g, void *x, int j, hipStream_t s);
// *flag = 1 if every one of the N rhs has converged (done == 1), else 0
hipError_t launch_all_converged(const CgState *st, int N, int *flag, hipStream_t s);
// x_j += y_j, y_j = 0 (x FP64, y FP32) for rhs taking a reliable update, or all with all != 0
hipError_t launch_fold_mr(const Geom &g, void *x, void *y, const CgState *st, int all, int N, const int *gate,
                          hipStream_t s);

// one-workgroup kernel: sum nparts partials in order, apply `fin` (skipped when *done != 0)
hipError_t launch_finish(const RedArgs &ra, int nparts, int fin, const int *done, hipStream_t s);

int nblocks(const Geom &g);
int dslash_block_auto(int vh);
// workgroups of a single-rhs hopping launch (one partial sum each in the norm epilogues)
int hop_blocks(const Geom &g);
// Geom::dir by default: the small-lattice hopping kernel up to kDirAutoVh sites per parity
constexpr int kDirAutoVh = 1 << 15;
int dir_auto(int vh);

}  // namespace wl

// ---------------------------------------------------------------- ABI objects ---------

struct wl_field {
    wl_ctx *ctx;     // null once the context has been destroyed (orphaned handle)
    int parity;  // WL_EVEN / WL_ODD
    int prec;
    int nrhs;
    void *data;
    size_t rhs_elems;  // stored elements per rhs: 12 * stride (13 * stride for FP16: scales)
    size_t bytes;
    // Layout: element (rhs j, component c, site k) at (j * P + c) * stride + k: every rhs is
    // a set of 12 component planes (P = 12), plus a plane of FP32 site scales for FP16
    // fields (P = 13).
};

struct wl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    wl::Geom g{};
    int64_t V = 0;
    int antiperiodic = 1;
    void *gauge64 = nullptr;
    void *gauge32 = nullptr;
    void *gauge16 = nullptr;
    void *force64 = nullptr;           // fermion force planes (gauge layout), on first use
    void *mom64 = nullptr;             // HMC momentum planes (gauge layout), zero until uploaded
    bool gauge_ok = false;
    bool gauge32_ok = false;
    bool gauge16_ok = false;
    // two-row link copies (rows 0 and 1 of every link) for the FP64 and FP32 hopping kernels,
    // used when link_rows == 2 and the links passed the special-unitarity check
    void *rows64 = nullptr;
    void *rows32 = nullptr;
    bool rows64_ok = false;
    bool rows32_ok = false;
    int rows_check = 0;                // 0 not checked since the upload, 1 passed, -1 failed
    double rows_dev = 0;               // largest |U[2][j] - conj(row 0 x row 1)_j| found
    int link_rows = 3;                 // WL_OPT_LINK_ROWS
    void *staging = nullptr;
    size_t staging_bytes = 0;
    double *partials = nullptr;
    wl::CgState *st = nullptr;
    wl::CgState *st_host = nullptr;  // pinned
    double *scratch_d = nullptr;     // FIN_STORE target
    double *scratch_h = nullptr;     // pinned
    // work fields (created on first use)
    wl_field *tmp_odd64 = nullptr;   // odd temporary of M
    wl_field *tmp_odd32 = nullptr;
    wl_field *cg_r = nullptr, *cg_p = nullptr, *cg_Dp = nullptr, *cg_DDp = nullptr;
    wl_field *mx_r = nullptr, *mx_p = nullptr, *mx_Dp = nullptr;  // inner-precision Krylov vectors
    // solver work fields, keyed by (slot, precision, parity); each holds at least the rhs
    // count of its largest use so far
    std::map<std::tuple<int, int, int>, wl_field *> work;
    size_t partials_cap = 0;           // doubles in `partials`
    wl::CgState *st_mr = nullptr;      // per-rhs CG states (device) and pinned host mirror
    wl::CgState *st_mr_host = nullptr;
    int st_mr_cap = 0;
    int *flags_d = nullptr;            // batched CG: [0] all finished, [1] update pending,
                                       // [2] no update pending, [3] all converged
    int *flags_h = nullptr;
    int rhs_block = 0;                 // rhs per lane of the multi-rhs kernel (0: automatic)
    int rhs_launch = 0;                // rhs per launch (pass over the lattice; 0: automatic)
    int link_load = 0;                 // multi-rhs links: 0 automatic, 1 cached, 2 nontemporal, 3 LDS
    bool use_graphs = false;           // batched CG: replay each chunk from a captured HIP graph
    int pipeline = 0;                  // multi-rhs hops: 0 pipelined where possible, 1 never
    int cg_persistent = 0;             // single-rhs FP64 CG in one launch: 1 yes (where the
                                       // launch fits), 0 (default: kernel by kernel, launch-
                                       // folded on small lattices) or -1 kernel by kernel
    void *grid_bar = nullptr;          // grid barrier of the persistent CG (on first use)
    bool block_krylov = false;         // WL_OPT_BLOCK_KRYLOV: several rhs share one Krylov space
    int small_hop = 0;                 // WL_OPT_SMALL_HOP (-1 never, 0 automatic, 1 always);
                                       // the setting in effect is g.dir
    bool blk_auto = true;              // g.blk is the volume's default (not chosen by a caller)
    bool lds_auto = true;              // g.lds unset: multi-rhs hops take mr_geom's occupancy cap
    double op_shift = 0.0;             // Hasenbusch shift mh of the operator being applied,
                                       // M + mh (0 outside a shifted call; see OpShift)
    // fields created through wl_field_create and not yet destroyed; wl_destroy orphans them
    std::unordered_set<wl_field *> user_fields;
    size_t dev_bytes = 0;              // device memory held by the context and its fields
    std::mutex mu;
};
