// Resumable batched conjugate gradient on M^dag M for N right-hand sides on one context.
//
// The host never waits inside a solve: begin() enqueues the start, pump() retires completed
// host checks and enqueues more work, so that one thread can drive many solves on many
// streams (the request server) or one solve to completion (wl_cg_solve for nrhs > 1).
//
// Work is enqueued in chunks of kChunk iterations.  With WL_OPT_GRAPHS a chunk (its kernels
// and the flag copy; no memsets) is captured once per solve into a HIP graph and replayed,
// one launch per chunk instead of ~75.  Every kernel of an iteration reads a
// device flag and returns at once when all right-hand sides have finished, and rhs that have
// finished (or wait for a reliable update) are skipped inside the kernels, so chunks can be
// enqueued ahead of the host's knowledge of convergence: up to `depth` chunks are in flight.
// In mixed precision a chunk ends with the reliable-update sequence (fold the inner solution
// into x, FP64 M^dag M x, new residual, restart), gated on the device by the update-pending
// flag, so that the host takes no decision inside the iteration either.  With an FP32 or FP16
// inner precision the iterations accumulate alpha p into an FP32 inner solution y, which each
// reliable update adds to the FP64 x (and clears): the iterations then never touch x.
#pragma once

#include <chrono>
#include <deque>
#include <vector>

#include "wl_host.h"

namespace wl {

constexpr int kChunk = 8;  // iterations per chunk (one host check each)

struct CgSpec {
    double kappa = 0.125;
    double mh = 0.0;       // Hasenbusch shift: CG on (M + mh)^dag (M + mh)
    double accuracy = 1e-12;
    int maxiters = 10000;
    int prec = WL_F64;     // precision of the inner operator and Krylov vectors
    bool mixed = false;    // reliable updates against the FP64 operator
    double delta = 0.1;    // update when ||r|| drops by this factor since the last one
    int update_every = 100;
};

// Block CG (wl_bcg.hip): N <= 24 right-hand sides sharing one Krylov space, FP64 or FP32 inner
// precision (reliable updates when spec.mixed); stats: N entries.
int cg_block(wl_ctx *c, void *x, const void *b, int N, const CgSpec &spec, wl_cg_stats *stats);

class BatchCG {
  public:
    BatchCG() = default;
    BatchCG(const BatchCG &) = delete;
    BatchCG &operator=(const BatchCG &) = delete;
    ~BatchCG();

    // Allocate the resident work memory of an N-rhs solve at inner precision prec (the only
    // allocations of a solve; a server calls this once per slot).
    static int prepare(wl_ctx *c, int N, int prec, bool mixed);
    // The same allocations before any gauge field is uploaded: the FP64 links and the inner
    // precision's copy are allocated but not filled (a server provisioning idle slots).  Full
    // links only (a context with two-row links builds its copies at first use).
    static int reserve(wl_ctx *c, int N, int prec, bool mixed);
    // Enqueue the start of a solve of M^dag M x = b (x: initial guess) on c's stream.
    int begin(wl_ctx *c, void *x, const void *b, int N, const CgSpec &spec);
    // Retire completed host checks and enqueue more work; never waits.  Sets *finished once the
    // final states are on the host (then results() may be called).
    int pump(bool *finished);
    // An event of the oldest outstanding check (to block on), or null when none is outstanding.
    hipEvent_t oldest() const { return inflight_.empty() ? nullptr : inflight_.front().ev; }
    // Per-rhs statistics (N entries) of a finished solve; WL_E_NOCONV when some rhs did not
    // converge.
    int results(wl_cg_stats *stats) const;
    int N() const { return N_; }

  private:
    struct Check {
        hipEvent_t ev;
        int slot;    // flags slot in the pinned host mirror
        bool final;  // final states copy
    };
    int enqueue_chunk();
    int enqueue_chunk_body(int slot);
    int enqueue_update();
    int enqueue_final();
    int get_event(hipEvent_t *ev);
    void drop_graphs();

    wl_ctx *c_ = nullptr;
    void *x_ = nullptr;
    const void *b_ = nullptr;
    int N_ = 0;
    CgSpec s_{};
    wl_field *R_ = nullptr, *P_ = nullptr, *DP_ = nullptr, *T_ = nullptr;
    wl_field *DP64_ = nullptr, *DDP64_ = nullptr, *T64_ = nullptr;
    wl_field *Y_ = nullptr;  // FP32 inner solution (FP32 / FP16 inner precision), folded into x
    void *acc_ = nullptr;    // what x += alpha p accumulates into: Y_ or x itself (FP64)
    int pplain_ = 0;
    int chunks_ = 0;     // chunks enqueued
    int cap_chunks_ = 0;
    bool stop_ = false;  // no more chunks (all finished or the chunk cap)
    bool final_enqueued_ = false;
    bool done_ = false;
    int depth_ = 2;
    std::deque<Check> inflight_;
    std::vector<hipEvent_t> free_ev_;
    std::vector<hipEvent_t> all_ev_;
    std::chrono::steady_clock::time_point t0_;
    double ms_ = 0;
    // chunk graphs (WL_OPT_GRAPHS): one per flags slot parity, captured on first use in a solve
    bool use_graphs_ = false;
    hipGraphExec_t exec_[2] = {nullptr, nullptr};
};

}  // namespace wl
