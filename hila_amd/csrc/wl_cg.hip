// Conjugate gradient on M^dag M (libraries/dirac/conjugate_gradient.h:19-105): the FP64
// single-rhs solver with HILA's semantics, the single-rhs mixed-precision solver with reliable
// updates, and the resumable batched solver (BatchCG) behind the multi-rhs entry points and
// the request server.
//
// Scalars (rr, alpha, beta, the convergence test) live on the device in CgState and are
// updated by one-workgroup finish kernels, so an iteration needs no host round trip; the host
// checks the `done` flag once per chunk of iterations.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "wl_solve.h"

namespace wl {

namespace {

// Launch folding of the FP64 CG on small lattices: a hopping pass of at most this many
// workgroups leaves its partial sums to the next kernel, which adds them itself in the finish
// kernel's order, instead of a one-workgroup finish kernel between them.
constexpr int kFoldParts = 512;

bool cg_fold(const wl_ctx *c, const CgWork &w) {
    const int nhop = hop_blocks(c->g);
    return c->cg_persistent == 0 && nhop <= kFoldParts && w.partials_cap >= (size_t)2 * nhop;
}

// The kernel-by-kernel FP64 CG on work memory w.  An iteration is M p with ||Dp||^2 partials,
// the finish of alpha, M^dag Dp fused with the x, r update and ||r||^2 partials, the finish of
// rrnew (convergence, beta) and p = beta p + r: seven launches.  Folded (cg_fold), the alpha
// finish moves into the update pass and the rrnew finish into p = beta p + r: five launches,
// the same sums in the same order, so the same bits.
int cg_f64(wl_ctx *c, const CgWork &w, void *x, size_t x_bytes, const void *b, double kappa, double accuracy,
           int maxiters, wl_cg_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    void *r = w.r, *p = w.p, *Dp = w.Dp, *DDp = w.DDp, *t = w.t;
    const RedArgs ra = red_args(w);
    const int *done = &w.st->done;
    const bool fold = cg_fold(c, w);
    const int nhop = hop_blocks(c->g);
    double *parts2 = w.partials + nhop;  // folded: partials of the update pass

    CgState init{};
    init.acc2 = accuracy * accuracy;
    init.maxiters = maxiters;
    *c->st_host = init;
    HIPCHK(hipMemcpyAsync(w.st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    // source_norm and target_rr (conjugate_gradient.h:61-63)
    HIPCHK(launch_norm2(WL_F64, c->g, b, 1, 0, ra, FIN_SOURCE, nullptr, c->stream));
    // r = b - M^dag M x ; p = r ; rr = |r|^2  (:65-73)
    if ((rc = meo_one(c, ra, WL_F64, Dp, x, t, kappa, +1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    if ((rc = meo_one(c, ra, WL_F64, DDp, Dp, t, kappa, -1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    HIPCHK(launch_cg_init(c->g, b, DDp, r, p, ra, c->stream));
    HIPCHK(hipMemcpyAsync(c->st_host, w.st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->st_host->source_norm == 0.0) {
        // b = 0: the solution is x = 0 (HILA would divide 0 by 0 here)
        HIPCHK(hipMemsetAsync(x, 0, x_bytes, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->converged = 1;
        }
        return WL_OK;
    }
    if (maxiters <= 0) {
        c->st_host->done = 2;
        HIPCHK(hipMemcpyAsync(&w.st->done, &c->st_host->done, sizeof(int), hipMemcpyHostToDevice, c->stream));
    }

    // main loop (:75-96), kChunk iterations per enqueue; the kernels of an iteration after
    // convergence return at once.
    int enqueued = 0;
    int done_h = c->st_host->done;
    while (!done_h && enqueued < maxiters) {
        const int n = std::min(kChunk, maxiters - enqueued);
        for (int it = 0; it < n; ++it) {
            // Dp = M p, pDp = |Dp|^2, alpha = rr / pDp (folded: alpha in the update pass)
            if ((rc = meo_one(c, ra, WL_F64, Dp, p, t, kappa, +1, EPI_AXPY_NORM, done, fold ? FIN_NONE : FIN_ALPHA)))
                return rc;
            // DDp = M^dag Dp fused with x += alpha p, r -= alpha DDp, rrnew = |r|^2, test, beta
            if ((rc = dslash_one(c, ra, WL_F64, t, 1, Dp, nullptr, -kappa, -1, EPI_SET, done, FIN_NONE)))
                return rc;
            DslashLaunch L{};
            L.out = nullptr;
            L.in = t;
            L.xin = Dp;
            L.gauge = gauge_for(c, WL_F64);
            L.lf = link_format(c, WL_F64);
            if (!L.gauge) return fail(WL_E_STATE, "FP64 links not prepared");
            L.a = kappa;
            L.b = 1.0 + c->op_shift;
            L.out_par = 0;
            L.sign = -1;
            L.epi = EPI_CGUPD;
            L.done = done;
            L.x = x;
            L.p = p;
            L.r = r;
            L.red = ra;
            L.fin = FIN_RRNEW;
            if (fold) {
                L.red.partials = parts2;
                L.fin = FIN_NONE;
                L.fold_parts = w.partials;
                L.fold_n = nhop;
                L.fold_it = enqueued + it;
            }
            HIPCHK(launch_dslash(WL_F64, c->g, L, c->stream));
            // p = beta p + r (folded: with the convergence test and beta)
            if (fold)
                HIPCHK(launch_cg_xpay_fold(c->g, p, r, w.st, parts2, nhop, enqueued + it, c->stream));
            else
                HIPCHK(launch_cg_xpay(WL_F64, c->g, p, r, w.st, c->stream));
        }
        enqueued += n;
        HIPCHK(hipMemcpyAsync(&c->st_host->done, &w.st->done, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        done_h = c->st_host->done;
    }

    // true residual |b - M^dag M x|^2 recomputed in FP64
    if ((rc = meo_one(c, ra, WL_F64, Dp, x, t, kappa, +1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    if ((rc = meo_one(c, ra, WL_F64, DDp, Dp, t, kappa, -1, EPI_AXPY, nullptr, FIN_NONE))) return rc;
    HIPCHK(launch_diff_norm2(c->g, b, DDp, ra, FIN_TRUE, c->stream));
    HIPCHK(hipMemcpyAsync(c->st_host, w.st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
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
            L.lf = link_format(c, prec);
            if (!L.gauge) return fail(WL_E_STATE, "links of precision %d not prepared", prec);
            L.a = kappa;
            L.b = 1.0 + c->op_shift;
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


}  // namespace

// ------------------------------------------------------------------ BatchCG ------------

BatchCG::~BatchCG() {
    drop_graphs();
    for (hipEvent_t e : all_ev_) (void)hipEventDestroy(e);
}

void BatchCG::drop_graphs() {
    for (hipGraphExec_t &g : exec_)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int BatchCG::prepare(wl_ctx *c, int N, int prec, bool mixed) {
    int rc;
    // the inner operator's links, and the FP64 ones of the reliable updates and true residuals
    // (a separate copy when two-row links are on)
    if ((rc = ensure_gauge(c, prec))) return rc;
    if (prec != WL_F64 && (rc = ensure_gauge(c, WL_F64))) return rc;
    if ((rc = ensure_st_mr(c, N))) return rc;
    wl_field *f;
    if ((rc = work_field(c, W_R, prec, WL_EVEN, N, &f)) || (rc = work_field(c, W_P, prec, WL_EVEN, N, &f)) ||
        (rc = work_field(c, W_DP, prec, WL_EVEN, N, &f)) || (rc = work_field(c, W_T, prec, WL_ODD, N, &f)) ||
        (rc = work_field(c, W_DP64, WL_F64, WL_EVEN, N, &f)) ||
        (rc = work_field(c, W_DDP64, WL_F64, WL_EVEN, N, &f)) ||
        (rc = work_field(c, W_T64, WL_F64, WL_ODD, N, &f)))
        return rc;
    if (mixed && prec != WL_F64)
        if ((rc = work_field(c, W_Y, WL_F32, WL_EVEN, N, &f))) return rc;
    const size_t rows = (size_t)std::max(nblocks(c->g), mr_segments(c->g));
    if ((rc = ensure_partials(c, (size_t)N * rows))) return rc;
    return ensure_staging(c, (size_t)c->g.vh * 24 * sizeof(double));
}

int BatchCG::reserve(wl_ctx *c, int N, int prec, bool mixed) {
    int rc;
    if (!c->gauge64) {  // as gauge_upload_async allocates it
        if ((rc = dev_alloc(c, &c->gauge64, gauge_elems(c->g) * 16, "gauge field"))) return rc;
        HIPCHK(hipMemsetAsync(c->gauge64, 0, gauge_elems(c->g) * 16, c->stream));
    }
    if (prec == WL_F32 && !c->gauge32)
        if ((rc = dev_alloc(c, &c->gauge32, gauge_elems(c->g) * 8, "FP32 gauge copy"))) return rc;
    if (prec == WL_F16 && !c->gauge16)
        if ((rc = dev_alloc(c, &c->gauge16, gauge_elems(c->g) * 4, "FP16 gauge copy"))) return rc;
    if ((rc = ensure_st_mr(c, N))) return rc;
    wl_field *f;
    if ((rc = work_field(c, W_R, prec, WL_EVEN, N, &f)) || (rc = work_field(c, W_P, prec, WL_EVEN, N, &f)) ||
        (rc = work_field(c, W_DP, prec, WL_EVEN, N, &f)) || (rc = work_field(c, W_T, prec, WL_ODD, N, &f)) ||
        (rc = work_field(c, W_DP64, WL_F64, WL_EVEN, N, &f)) ||
        (rc = work_field(c, W_DDP64, WL_F64, WL_EVEN, N, &f)) ||
        (rc = work_field(c, W_T64, WL_F64, WL_ODD, N, &f)))
        return rc;
    if (mixed && prec != WL_F64)
        if ((rc = work_field(c, W_Y, WL_F32, WL_EVEN, N, &f))) return rc;
    const size_t rows = (size_t)std::max(nblocks(c->g), mr_segments(c->g));
    if ((rc = ensure_partials(c, (size_t)N * rows))) return rc;
    // the gauge upload's staging (one direction of HILA's AoS links) is the larger one
    return ensure_staging(c, std::max((size_t)c->g.vh * 24, (size_t)c->V * 18) * sizeof(double));
}

int BatchCG::get_event(hipEvent_t *ev) {
    if (!free_ev_.empty()) {
        *ev = free_ev_.back();
        free_ev_.pop_back();
        return WL_OK;
    }
    HIPCHK(hipEventCreateWithFlags(ev, hipEventDisableTiming));
    all_ev_.push_back(*ev);
    return WL_OK;
}

int BatchCG::begin(wl_ctx *c, void *x, const void *b, int N, const CgSpec &spec) {
    if (N < 1 || N > 64) return fail(WL_E_ARG, "nrhs %d out of range 1..64", N);
    if (!spec.mixed && spec.prec != WL_F64) return fail(WL_E_ARG, "plain batched CG runs in FP64");
    c_ = c;
    x_ = x;
    b_ = b;
    N_ = N;
    s_ = spec;
    c->op_shift = s_.mh;  // every launch this solve enqueues applies M + mh
    chunks_ = 0;
    stop_ = final_enqueued_ = done_ = false;
    inflight_.clear();
    drop_graphs();  // operands or shapes may differ from the previous solve
    use_graphs_ = c->use_graphs;
    t0_ = std::chrono::steady_clock::now();
    int rc;
    if ((rc = prepare(c, N, s_.prec, s_.mixed))) return rc;
    const int prec = s_.prec;
    if ((rc = work_field(c, W_R, prec, WL_EVEN, N, &R_)) || (rc = work_field(c, W_P, prec, WL_EVEN, N, &P_)) ||
        (rc = work_field(c, W_DP, prec, WL_EVEN, N, &DP_)) || (rc = work_field(c, W_T, prec, WL_ODD, N, &T_)) ||
        (rc = work_field(c, W_DP64, WL_F64, WL_EVEN, N, &DP64_)) ||
        (rc = work_field(c, W_DDP64, WL_F64, WL_EVEN, N, &DDP64_)) ||
        (rc = work_field(c, W_T64, WL_F64, WL_ODD, N, &T64_)))
        return rc;
    Y_ = nullptr;
    acc_ = x;
    if (s_.mixed && prec != WL_F64) {
        if ((rc = work_field(c, W_Y, WL_F32, WL_EVEN, N, &Y_))) return rc;
        acc_ = Y_->data;
    }
    pplain_ = nblocks(c->g);
    cap_chunks_ = s_.maxiters > 0 ? (s_.maxiters + kChunk - 1) / kChunk + 4 : 0;

    CgState *st = c->st_mr;
    for (int j = 0; j < N; ++j) {
        CgState s0{};
        s0.acc2 = s_.accuracy * s_.accuracy;
        s0.maxiters = s_.maxiters;
        s0.delta2 = s_.delta * s_.delta;
        s0.update_every = s_.update_every;
        c->st_mr_host[j] = s0;
    }
    hipStream_t S = c->stream;
    HIPCHK(hipMemcpyAsync(st, c->st_mr_host, N * sizeof(CgState), hipMemcpyHostToDevice, S));
    HIPCHK(hipMemsetAsync(c->flags_d, 0, 64, S));
    if (Y_) HIPCHK(hipMemsetAsync(Y_->data, 0, (size_t)N * Y_->rhs_elems * 8, S));
    double *part = c->partials;
    HIPCHK(launch_norm2_mr(WL_F64, c->g, b, N, part, S));
    HIPCHK(launch_finish_mr(part, pplain_, pplain_, N, st, nullptr, FIN_SOURCE, nullptr, S));
    // r = b - M^dag M x, p = r  (conjugate_gradient.h:65-73)
    if ((rc = meo_mr(c, WL_F64, DP64_->data, x, T64_->data, s_.kappa, +1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, DDP64_->data, DP64_->data, T64_->data, s_.kappa, -1, EPI_AXPY, N, nullptr,
                     nullptr, nullptr)))
        return rc;
    if (!s_.mixed) {
        HIPCHK(launch_cg_init_mr(c->g, b, DDP64_->data, R_->data, P_->data, N, part, S));
        HIPCHK(launch_finish_mr(part, pplain_, pplain_, N, st, nullptr, FIN_RR, nullptr, S));
    } else {
        HIPCHK(launch_residual_update_mr(prec, c->g, b, DDP64_->data, R_->data, st, 1, N, part, nullptr, S));
        HIPCHK(launch_finish_mr(part, pplain_, pplain_, N, st, nullptr, FIN_UPDATE_FIRST, nullptr, S));
        HIPCHK(launch_copy(prec, c->g, R_->data, P_->data, N, S));
    }
    HIPCHK(launch_batch_sync(st, N, 0, c->flags_d, nullptr, S));
    if (cap_chunks_ == 0) stop_ = true;
    bool fin = false;
    return pump(&fin);
}

// kChunk iterations (conjugate_gradient.h:75-95 per rhs), the gated reliable update, and the
// copy of the all-finished flag to host slot `slot`.
int BatchCG::enqueue_chunk_body(int slot) {
    wl_ctx *c = c_;
    hipStream_t S = c->stream;
    CgState *st = c->st_mr;
    const int *all_done = c->flags_d;
    const int N = N_, prec = s_.prec;
    int rc;
    for (int it = 0; it < kChunk; ++it) {
        int ps = 0;
        // Dp = M p, pDp_j = |Dp_j|^2, alpha_j = rr_j / pDp_j
        if ((rc = meo_mr(c, prec, DP_->data, P_->data, T_->data, s_.kappa, +1, EPI_AXPY_NORM, N, all_done, st, &ps)))
            return rc;
        HIPCHK(launch_finish_mr(c->partials, ps, ps, N, st, nullptr, FIN_ALPHA, all_done, S));
        // DDp = M^dag Dp fused with x_j += alpha_j p_j, r_j -= alpha_j DDp_j, rrnew_j = |r_j|^2
        if ((rc = dslash_mr_all(c, prec, T_->data, 1, DP_->data, nullptr, -s_.kappa, -1, EPI_SET, N, all_done,
                                st, nullptr, nullptr, nullptr, nullptr)))
            return rc;
        if ((rc = dslash_mr_all(c, prec, nullptr, 0, T_->data, DP_->data, s_.kappa, -1, EPI_CGUPD, N, all_done,
                                st, acc_, P_->data, R_->data, &ps)))
            return rc;
        HIPCHK(launch_finish_mr(c->partials, ps, ps, N, st, nullptr, s_.mixed ? FIN_RRNEW_MIXED2 : FIN_RRNEW,
                                all_done, S));
        HIPCHK(launch_batch_sync(st, N, s_.mixed ? 1 : 0, c->flags_d, nullptr, S));
        // p_j = beta_j p_j + r_j
        HIPCHK(launch_xpay_mr(prec, c->g, P_->data, R_->data, st, all_done, N, S));
    }
    if (s_.mixed)
        if ((rc = enqueue_update())) return rc;
    HIPCHK(hipMemcpyAsync(c->flags_h + 4 * slot, c->flags_d, 2 * sizeof(int), hipMemcpyDeviceToHost, S));
    return WL_OK;
}

int BatchCG::enqueue_chunk() {
    wl_ctx *c = c_;
    hipStream_t S = c->stream;
    const int slot = chunks_ & 1;  // at most two chunks in flight
    int rc;
    if (use_graphs_) {
        if (!exec_[slot]) {
            HIPCHK(hipStreamBeginCapture(S, hipStreamCaptureModeThreadLocal));
            rc = enqueue_chunk_body(slot);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(S, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(WL_E_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiate(&exec_[slot], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                exec_[slot] = nullptr;
                return fail(WL_E_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ei));
            }
        }
        HIPCHK(hipGraphLaunch(exec_[slot], S));
    } else if ((rc = enqueue_chunk_body(slot))) {
        return rc;
    }
    ++chunks_;
    hipEvent_t ev;
    if ((rc = get_event(&ev))) return rc;
    HIPCHK(hipEventRecord(ev, S));
    inflight_.push_back({ev, slot, false});
    return WL_OK;
}

// Reliable update of every running rhs, run only when one asked for it (flags[2] == 0): the
// FP64 residual b - M^dag M x replaces r, the recurrences restart from it (or the rhs has
// converged on its true residual).
int BatchCG::enqueue_update() {
    wl_ctx *c = c_;
    hipStream_t S = c->stream;
    CgState *st = c->st_mr;
    const int *gate = c->flags_d + 2;
    const int N = N_, prec = s_.prec;
    int rc;
    if (Y_) HIPCHK(launch_fold_mr(c->g, x_, Y_->data, st, 0, N, gate, S));
    if ((rc = meo_mr(c, WL_F64, DP64_->data, x_, T64_->data, s_.kappa, +1, EPI_AXPY, N, gate, nullptr, nullptr)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, DDP64_->data, DP64_->data, T64_->data, s_.kappa, -1, EPI_AXPY, N, gate, nullptr,
                     nullptr)))
        return rc;
    HIPCHK(launch_residual_update_mr(prec, c->g, b_, DDP64_->data, R_->data, st, 0, N, c->partials, gate, S));
    HIPCHK(launch_finish_mr(c->partials, pplain_, pplain_, N, st, nullptr, FIN_UPDATE, gate, S));
    HIPCHK(launch_xpay_mr(prec, c->g, P_->data, R_->data, st, gate, N, S));
    HIPCHK(launch_batch_sync(st, N, 0, c->flags_d, gate, S));
    return WL_OK;
}

// FP64 true residual of every rhs, and the states to the host.
int BatchCG::enqueue_final() {
    wl_ctx *c = c_;
    hipStream_t S = c->stream;
    CgState *st = c->st_mr;
    const int N = N_;
    int rc;
    if (Y_) HIPCHK(launch_fold_mr(c->g, x_, Y_->data, st, 1, N, nullptr, S));
    // In mixed precision an rhs converges only at a reliable update, which has just computed
    // its FP64 true residual for the final x: the recomputation runs only if some rhs ended
    // otherwise (maxiters).  Plain CG always recomputes it.
    const int *skip = nullptr;
    if (s_.mixed) {
        HIPCHK(launch_all_converged(st, N, c->flags_d + 3, S));
        skip = c->flags_d + 3;
    }
    if ((rc = meo_mr(c, WL_F64, DP64_->data, x_, T64_->data, s_.kappa, +1, EPI_AXPY, N, skip, nullptr, nullptr)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, DDP64_->data, DP64_->data, T64_->data, s_.kappa, -1, EPI_AXPY, N, skip, nullptr,
                     nullptr)))
        return rc;
    HIPCHK(launch_residual_update_mr(WL_F64, c->g, b_, DDP64_->data, DP64_->data, st, 1, N, c->partials, skip, S));
    HIPCHK(launch_finish_mr(c->partials, pplain_, pplain_, N, st, nullptr, FIN_TRUE, skip, S));
    HIPCHK(hipMemcpyAsync(c->st_mr_host, st, N * sizeof(CgState), hipMemcpyDeviceToHost, S));
    hipEvent_t ev;
    if ((rc = get_event(&ev))) return rc;
    HIPCHK(hipEventRecord(ev, S));
    inflight_.push_back({ev, 0, true});
    final_enqueued_ = true;
    return WL_OK;
}

int BatchCG::pump(bool *finished) {
    *finished = done_;
    if (done_) return WL_OK;
    c_->op_shift = s_.mh;
    while (!inflight_.empty()) {
        const Check ck = inflight_.front();
        const hipError_t q = hipEventQuery(ck.ev);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) return fail(WL_E_DEVICE, "CG check failed: %s", hipGetErrorString(q));
        inflight_.pop_front();
        free_ev_.push_back(ck.ev);
        if (ck.final) {
            done_ = true;
            ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
            for (int j = 0; j < N_; ++j)
                if (c_->st_mr_host[j].source_norm == 0.0)  // b_j = 0: x_j = 0
                    HIPCHK(launch_zero_rhs(c_->g, x_, j, c_->stream));
            *finished = true;
            return WL_OK;
        }
        if (c_->flags_h[4 * ck.slot]) stop_ = true;  // every rhs has finished
    }
    if (final_enqueued_) return WL_OK;
    int rc;
    while (!stop_ && chunks_ < cap_chunks_ && (int)inflight_.size() < depth_)
        if ((rc = enqueue_chunk())) return rc;
    if (stop_ || chunks_ >= cap_chunks_) return enqueue_final();
    return WL_OK;
}

int BatchCG::results(wl_cg_stats *stats) const {
    int worst = WL_OK;
    for (int j = 0; j < N_; ++j) {
        const CgState &S = c_->st_mr_host[j];
        const bool conv = S.done == 1;
        if (!conv) worst = WL_E_NOCONV;
        if (stats) {
            wl_cg_stats &o = stats[j];
            std::memset(&o, 0, sizeof(o));
            o.iterations = S.iter;
            o.converged = conv ? 1 : 0;
            o.rr_rel = S.source_norm > 0 ? S.rrnew / S.source_norm : 0.0;
            o.true_resid = S.source_norm > 0 ? std::sqrt(S.true_rr / S.source_norm) : 0.0;
            o.reliable_updates = s_.mixed ? std::max(0, S.nreliable - 1) : 0;
            o.ms = ms_;
        }
    }
    return worst;
}

namespace {

// Blocking driver of BatchCG for the multi-rhs entry points.
int cg_batched(wl_ctx *c, void *x, const void *b, int N, const CgSpec &spec, wl_cg_stats *stats) {
    BatchCG cg;
    int rc = cg.begin(c, x, b, N, spec);
    if (rc) return rc;
    for (;;) {
        bool fin = false;
        if ((rc = cg.pump(&fin))) return rc;
        if (fin) break;
        hipEvent_t ev = cg.oldest();
        if (ev) HIPCHK(hipEventSynchronize(ev));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    rc = cg.results(stats);
    if (rc) return fail(WL_E_NOCONV, "batched CG: not every right-hand side converged");
    return WL_OK;
}

}  // namespace

size_t cg_partials_needed(const wl_ctx *c) {
    const int nhop = hop_blocks(c->g);
    return (size_t)std::max({(c->g.vh + c->g.blk - 1) / c->g.blk, nhop, nblocks(c->g),
                             nhop <= kFoldParts ? 2 * nhop : 0});  // two regions when folded
}

int cg_work_of(wl_ctx *c, CgWork *w) {
    int rc;
    for (wl_field **f : {&c->cg_r, &c->cg_p, &c->cg_Dp, &c->cg_DDp})
        if ((rc = ensure_tmp(c, WL_F64, f, WL_EVEN))) return rc;
    if ((rc = ensure_tmp(c, WL_F64, &c->tmp_odd64, WL_ODD))) return rc;
    if ((rc = ensure_partials(c, cg_partials_needed(c)))) return rc;
    if (!c->grid_bar)
        if ((rc = dev_alloc(c, &c->grid_bar, 256, "grid barrier"))) return rc;
    w->r = c->cg_r->data;
    w->p = c->cg_p->data;
    w->Dp = c->cg_Dp->data;
    w->DDp = c->cg_DDp->data;
    w->t = c->tmp_odd64->data;
    w->partials = c->partials;
    w->partials_cap = c->partials_cap;
    w->st = c->st;
    w->bar = c->grid_bar;
    w->red_out = c->scratch_d;
    return WL_OK;
}

int cg_f64_run(wl_ctx *c, const CgWork &w, void *x, size_t x_bytes, const void *b, double kappa,
               double accuracy, int maxiters, wl_cg_stats *stats) {
    if (w.partials_cap < cg_partials_needed(c)) return fail(WL_E_STATE, "CG work memory: too few partial sums");
    if (cg_persistent_wanted(c)) {  // the same solve in one launch (wl_cgp.hip)
        bool ran = false;
        const int rc = cg_f64_persistent(c, w, x, x_bytes, b, kappa, accuracy, maxiters, stats, &ran);
        if (ran || rc) return rc;
    }
    return cg_f64(c, w, x, x_bytes, b, kappa, accuracy, maxiters, stats);
}

}  // namespace wl

using namespace wl;

namespace {

// CG<Op>::apply (conjugate_gradient.h:46-104) for Op = Dirac_Wilson_evenodd (mh = 0) or
// Hasenbusch_operator<Dirac_Wilson_evenodd> (Hasenbusch.h:15-53): plain FP64, or with reliable
// updates (mixed) around an inner precision.
int cg_solve(wl_ctx *c, wl_field *x, const wl_field *b, const wl_cg_params &p, bool mixed, wl_cg_stats *stats) {
    if (!c || !x || !b) return fail(WL_E_ARG, "null argument");
    if (x->ctx != c || b->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    if (x->parity != WL_EVEN || b->parity != WL_EVEN) return fail(WL_E_ARG, "CG acts on even fields");
    if (x == b) return fail(WL_E_ARG, "x and b must be different fields");
    if (x->prec != WL_F64 || b->prec != WL_F64) return fail(WL_E_ARG, "x and b must be FP64 fields");
    if (x->nrhs != b->nrhs) return fail(WL_E_ARG, "x and b differ in nrhs");
    if (!(p.accuracy > 0)) return fail(WL_E_ARG, "accuracy must be > 0");
    if (!std::isfinite(p.mh) || !std::isfinite(p.kappa)) return fail(WL_E_ARG, "kappa and the shift must be finite");
    if (mixed) {
        if (p.inner_prec != WL_F32 && p.inner_prec != WL_F64 && p.inner_prec != WL_F16)
            return fail(WL_E_NOTIMPL, "inner precision %d", p.inner_prec);
        if (!(p.delta > 0 && p.delta < 1) || p.update_every < 1)
            return fail(WL_E_ARG, "need 0 < delta < 1, update_every >= 1");
    }
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    OpShift os(c, p.mh);
    int rc = ensure_gauge(c, WL_F64);
    if (rc) return rc;
    if (!mixed && b->nrhs == 1) {
        CgWork w;
        if ((rc = cg_work_of(c, &w))) return rc;
        return cg_f64_run(c, w, x->data, x->bytes, b->data, p.kappa, p.accuracy, p.maxiters, stats);
    }
    if (mixed && b->nrhs == 1 && p.inner_prec != WL_F16)  // FP16 inner: the batched solver, any count
        return cg_mixed(c, x->data, b->data, p.kappa, p.accuracy, p.maxiters, p.inner_prec, p.delta,
                        p.update_every, stats);
    CgSpec s;
    s.kappa = p.kappa;
    s.mh = p.mh;
    s.accuracy = p.accuracy;
    s.maxiters = p.maxiters;
    s.prec = mixed ? p.inner_prec : WL_F64;
    s.mixed = mixed;
    s.delta = p.delta;
    s.update_every = p.update_every;
    if (c->block_krylov && b->nrhs > 1) return cg_block(c, x->data, b->data, b->nrhs, s, stats);
    return cg_batched(c, x->data, b->data, b->nrhs, s, stats);
}

}  // namespace

extern "C" int wl_cg_solve(wl_ctx *c, wl_field *x, const wl_field *b, double kappa, double accuracy,
                           int maxiters, wl_cg_stats *stats) {
    wl_cg_params p{};
    p.kappa = kappa;
    p.accuracy = accuracy;
    p.maxiters = maxiters;
    p.inner_prec = WL_F64;
    return cg_solve(c, x, b, p, false, stats);
}

extern "C" int wl_cg_solve_mixed(wl_ctx *c, wl_field *x, const wl_field *b, double kappa,
                                 double accuracy, int maxiters, int inner_prec, double delta,
                                 int update_every, wl_cg_stats *stats) {
    wl_cg_params p{};
    p.kappa = kappa;
    p.accuracy = accuracy;
    p.maxiters = maxiters;
    p.inner_prec = inner_prec;
    p.delta = delta;
    p.update_every = update_every;
    return cg_solve(c, x, b, p, true, stats);
}

extern "C" int wl_cg_solve_ex(wl_ctx *c, wl_field *x, const wl_field *b, const wl_cg_params *p,
                              wl_cg_stats *stats) {
    if (!p) return fail(WL_E_ARG, "null argument");
    return cg_solve(c, x, b, *p, p->inner_prec != WL_F64, stats);
}
