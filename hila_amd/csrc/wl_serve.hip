// Request server: continuous batching of independent CG solves on one device.
//
// Three host threads drive the device:
//   loader     admits queued requests in submission order when a solve slot is free and the
//              slot's device memory fits in the HBM budget (a request that can start at once
//              may pass a head that waits for memory, a bounded number of times), and enqueues
//              the request's gauge field and sources on the slot's stream;
//   scheduler  starts each admitted solve (BatchCG, batched over the request's right-hand
//              sides) and pumps every running solve without waiting inside any of them, so
//              the solves of all admitted requests overlap on the device, one stream each;
//   unloader   writes finished solutions back to the request's host buffers, completes the
//              request and returns the slot to the idle pool.
// A slot is a wl_ctx (gauge field, rhs and solution fields, CG work fields, stream) of one
// lattice size; it stays resident and serves the next request of that size.  Idle slots of
// other sizes are released (least recently used first) when a new slot would not fit the
// budget.
//
// A request is solved exactly as the blocking wl_cg_solve / wl_cg_solve_mixed of the same
// right-hand sides on a fresh context would solve it: the kernels and their launch shapes
// depend only on the lattice size, the precision and the rhs count, so a result does not
// depend on what else is being served.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "wl_solve.h"
#include "wl_store.h"

namespace wl {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t ticket = 0;
    wl_solve_request r{};
    std::string config_path;  // owned copy of r.config_path
    Clock::time_point t_submit, t_admit, t_done;
    wl_solve_result res{};
    std::vector<wl_cg_stats> stats;
    bool done = false;
    int bypassed = 0;  // later requests admitted before this one (admission order)
};

struct Slot {
    wl_ctx *ctx = nullptr;
    int dims[4] = {0, 0, 0, 0};
    int antiperiodic = 1;
    int prec = WL_F64;
    int cap = 0;  // right-hand sides the slot's fields hold
    wl_field *x = nullptr, *b = nullptr;
    std::shared_ptr<Request> req;
    BatchCG cg;
    Clock::time_point last_used;
    bool busy = false;
};

// Slot capacities: the request's rhs count rounded up to one of these, so that slots are
// reused across nearby counts.
int slot_capacity(int n) {
    static const int caps[] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64};
    for (int c : caps)
        if (n <= c) return c;
    return n;
}

// Device bytes of a slot (fields at block stride S, gauge copies, staging).  link_rows: the
// slots' WL_OPT_LINK_ROWS.  With two-row links BatchCG::prepare makes the FP64 two-row copy
// for the reliable updates at every inner precision, and the FP32 inner operator reads an
// FP32 two-row copy instead of the full FP32 links; FP16 always reads full FP16 links.
size_t slot_estimate(const int dims[4], int prec, int cap, int link_rows) {
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
    const int64_t vh = V / 2;
    const int64_t S = (vh + kBlock - 1) / kBlock * kBlock;
    const size_t f64 = (size_t)12 * S * 16, flo = (size_t)planes_of(prec) * S * elem_bytes(prec);
    size_t bytes = (size_t)cap * (2 * f64          // x, b
                                  + 4 * flo        // r, p, Dp (even), t (odd)
                                  + 3 * f64);      // FP64 update work
    if (prec != WL_F64) bytes += (size_t)cap * 12 * S * 8;  // FP32 inner solution
    bytes += (size_t)72 * S * 16;                  // FP64 gauge (always held)
    if (link_rows == 2) {
        bytes += (size_t)48 * S * 16;              // FP64 two-row copy
        if (prec == WL_F32) bytes += (size_t)48 * S * 8;   // FP32 two-row copy
        if (prec == WL_F16) bytes += (size_t)72 * S * 4;   // FP16 links
    } else if (prec != WL_F64) {
        bytes += (size_t)72 * S * elem_bytes(prec);
    }
    bytes += (size_t)V * 18 * 8;                   // staging (gauge upload)
    bytes += (size_t)cap * ((vh + 63) / 64) * 8 + (1 << 20);
    return bytes;
}

// Default concurrency, from the 200-request Poisson run at 20/s over {16^4, 24^3x48, 32^4}
// (FP16 inner, two repeats, one box): at most 4 / 6 / 8 / 12 / 16 solves at once gave p50
// 254-272 / 219-247 / 246-337 / 361-434 / 351-487 ms, p99 662-709 / 721-725 / 828-891 /
// 1005-1083 / 1192-1499 ms, and the 24-request closed batch took 0.88-0.90 / 0.92-0.94 /
// 0.94-1.03 / 0.99-1.03 / 1.02 s.  A request's multi-rhs launches already fill the GPU; more
// solves at once only share its caches.
constexpr int kDefaultMaxActive = 6;

int slot_link_rows(const std::vector<std::pair<int, int>> &opts) {
    int rows = 3;
    for (const auto &o : opts)
        if (o.first == WL_OPT_LINK_ROWS) rows = o.second;
    return rows;
}

}  // namespace
}  // namespace wl

using namespace wl;

struct wl_server {
    int device = 0;
    size_t budget = 0;
    int max_active = kDefaultMaxActive;

    std::mutex mu;
    std::condition_variable cv_load;    // loader: new request / slot freed
    std::condition_variable cv_sched;   // scheduler: new admitted slot
    std::condition_variable cv_unload;  // unloader: finished solve
    std::condition_variable cv_done;    // waiters
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> requests;
    std::list<std::unique_ptr<Slot>> slots;  // all resident slots
    std::deque<Slot *> admitted;             // loaded, to be started by the scheduler
    std::deque<Slot *> finished;             // solved, to be written back
    size_t bytes_in_use = 0;
    int active = 0;  // busy slots
    uint64_t next_ticket = 1;
    int64_t completed = 0, slots_created = 0, slots_released = 0;
    uint64_t generation = 0;  // completed requests so far (wakes a blocked loader)
    bool stop = false;        // no more submissions; the loader exits once the queue is empty
    bool stop_sched = false, stop_unload = false;
    bool submitted = false;                     // options are frozen after the first submission
    std::vector<std::pair<int, int>> slot_opts; // wl_set_option calls applied to every new slot
    std::thread loader, scheduler, unloader;
};

namespace {

void complete(wl_server *S, Slot *slot, int status, const char *msg) {
    std::shared_ptr<Request> q = slot->req;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        q->t_done = Clock::now();
        q->res.status = status;
        q->res.nrhs = q->r.nrhs;
        q->res.queue_ms = ms_between(q->t_submit, q->t_admit);
        q->res.latency_ms = ms_between(q->t_submit, q->t_done);
        if (msg) snprintf(q->res.message, sizeof(q->res.message), "%s", msg);
        q->done = true;
        slot->req.reset();
        slot->busy = false;
        slot->last_used = Clock::now();
        S->active -= 1;
        S->completed += 1;
        S->generation += 1;
    }
    S->cv_done.notify_all();
    S->cv_load.notify_all();
}

// A request that never got a slot.
void reject(wl_server *S, const std::shared_ptr<Request> &q, int status, const char *msg) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        q->t_admit = q->t_done = Clock::now();
        q->res.status = status;
        q->res.nrhs = q->r.nrhs;
        q->res.queue_ms = ms_between(q->t_submit, q->t_admit);
        q->res.latency_ms = ms_between(q->t_submit, q->t_done);
        snprintf(q->res.message, sizeof(q->res.message), "%s", msg);
        q->done = true;
        S->completed += 1;
    }
    S->cv_done.notify_all();
}

// Free a slot's device memory (its stream is synchronised first).  Call without the server lock.
void destroy_slot(Slot *s) {
    if (!s->ctx) return;
    field_free(s->x);
    field_free(s->b);
    s->x = s->b = nullptr;
    wl_destroy(s->ctx);
    s->ctx = nullptr;
}

int create_slot(wl_server *S, const wl_solve_request &r, int cap,
                const std::vector<std::pair<int, int>> &opts, std::unique_ptr<Slot> *out) {
    auto s = std::make_unique<Slot>();
    std::memcpy(s->dims, r.dims, sizeof(s->dims));
    s->antiperiodic = r.antiperiodic_t ? 1 : 0;
    s->prec = r.inner_prec;
    s->cap = cap;
    int rc = wl_create(r.dims, s->antiperiodic, S->device, &s->ctx);
    if (rc) return rc;
    wl_ctx *c = s->ctx;
    for (const auto &o : opts)
        if ((rc = wl_set_option(c, o.first, o.second))) {
            destroy_slot(s.get());
            return rc;
        }
    if ((rc = field_new(c, WL_EVEN, WL_F64, cap, &s->x)) || (rc = field_new(c, WL_EVEN, WL_F64, cap, &s->b))) {
        destroy_slot(s.get());
        return rc;
    }
    *out = std::move(s);
    return WL_OK;
}

// Enqueue the request's inputs on the slot's stream and allocate the slot's work memory.
int load_slot(Slot *s, const wl_solve_request &r) {
    wl_ctx *c = s->ctx;
    int rc;
    if (r.config_path) {
        if ((rc = config_upload(c, r.config_path))) return rc;
    } else {
        const double *U[4] = {r.gauge[0], r.gauge[1], r.gauge[2], r.gauge[3]};
        if ((rc = gauge_upload_async(c, U))) return rc;
    }
    if ((rc = BatchCG::prepare(c, s->cap, r.inner_prec, r.inner_prec != WL_F64))) return rc;
    const size_t blk = (size_t)c->g.vh * 24;
    for (int j = 0; j < r.nrhs; ++j)
        if ((rc = field_upload_async(c, s->b->data, WL_F64, s->cap, j, r.source + j * blk))) return rc;
    HIPCHK(hipMemsetAsync(s->x->data, 0, s->x->bytes, c->stream));
    return WL_OK;
}

bool slot_matches(const Slot &s, const wl_solve_request &r) {
    return !s.busy && std::memcmp(s.dims, r.dims, sizeof(s.dims)) == 0 &&
           s.antiperiodic == (r.antiperiodic_t ? 1 : 0) && s.prec == r.inner_prec && s.cap >= r.nrhs;
}

// Admission order.  When a solve slot is free, the next request is the one with the least
// estimated work (request_work) among the first kOrderWindow waiting, earlier first on ties --
// unless one of them has been passed over kMaxPassed times (a later request admitted before
// it), which then goes next: the oldest such.  When the chosen request has to wait for device
// memory (its slot does not fit the budget even after idle slots are released), another among
// the next kBypassWindow may go first if it can start at once -- an idle slot of its lattice is
// free, or its new slot fits the budget without releasing any -- until the chosen one has been
// passed over kMaxPassed times; then the queue waits for it.  A request's solve does not depend
// on when it is admitted, only its queue time does.
constexpr bool kShortestFirst = true;
constexpr size_t kOrderWindow = 16;
constexpr int kBypassWindow = 8;
constexpr int kMaxPassed = 8;

// Estimated work of a request: sites x right-hand sides x log10(1 / tolerance) (CG iterations
// grow with the digits asked for).
double request_work(const wl_solve_request &r) {
    const double V = (double)r.dims[0] * r.dims[1] * r.dims[2] * r.dims[3];
    const double digits = std::log10(1.0 / std::max(r.tolerance, 1e-300));
    return V * std::max(r.nrhs, 1) * std::max(digits, 1.0);
}

// The request to admit next (the queue is not empty; called with the server lock held).
std::shared_ptr<Request> next_request(wl_server *S);

// A slot for r that needs no memory beyond the budget: an idle slot of its lattice, or room
// for a new one (*need > 0) without releasing anything.
bool admissible_now(wl_server *S, const wl_solve_request &r, Slot **slot, size_t *need) {
    *slot = nullptr;
    *need = 0;
    for (auto &sp : S->slots)
        if (slot_matches(*sp, r) && (!*slot || sp->cap < (*slot)->cap)) *slot = sp.get();
    if (*slot) return true;
    *need = slot_estimate(r.dims, r.inner_prec, slot_capacity(r.nrhs), slot_link_rows(S->slot_opts));
    return S->bytes_in_use + *need <= S->budget;
}

std::shared_ptr<Request> next_request(wl_server *S) {
    const size_t n = std::min(S->pending.size(), kOrderWindow);
    for (size_t i = 0; i < n; ++i)
        if (S->pending[i]->bypassed >= kMaxPassed) return S->pending[i];
    if (!kShortestFirst) return S->pending.front();
    size_t best = 0;
    double wbest = request_work(S->pending[0]->r);
    for (size_t i = 1; i < n; ++i) {
        const double w = request_work(S->pending[i]->r);
        if (w < wbest) {
            best = i;
            wbest = w;
        }
    }
    return S->pending[best];
}

// Removes q from the queue; every request ahead of it has been passed over once more.
void take_from_queue(wl_server *S, const std::shared_ptr<Request> &q) {
    auto it = std::find(S->pending.begin(), S->pending.end(), q);
    for (auto j = S->pending.begin(); j != it; ++j) (*j)->bypassed += 1;
    S->pending.erase(it);
}

void loader_main(wl_server *S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    // the queue waits for memory: for a completion after generation blocked_at, or for a new
    // request that might be admitted ahead of the head
    uint64_t blocked_at = ~(uint64_t)0;
    size_t blocked_qlen = 0;
    for (;;) {
        S->cv_load.wait(lk, [&] {
            return (S->stop && S->pending.empty()) ||
                   (!S->pending.empty() && S->active < S->max_active &&
                    (S->generation != blocked_at || S->pending.size() > blocked_qlen));
        });
        if (S->pending.empty()) return;  // stopping, queue drained
        std::shared_ptr<Request> q = next_request(S);
        // an idle slot of this lattice (smallest sufficient capacity)
        Slot *slot = nullptr;
        for (auto &sp : S->slots)
            if (slot_matches(*sp, q->r) && (!slot || sp->cap < slot->cap)) slot = sp.get();
        size_t need = 0;
        if (!slot) {
            const wl_solve_request &r = q->r;
            need = slot_estimate(r.dims, r.inner_prec, slot_capacity(r.nrhs), slot_link_rows(S->slot_opts));
            // release idle slots, least recently used first, until the new one fits
            std::vector<std::unique_ptr<Slot>> victims;
            while (S->bytes_in_use + need > S->budget) {
                auto victim = S->slots.end();
                for (auto it = S->slots.begin(); it != S->slots.end(); ++it)
                    if (!(*it)->busy && (victim == S->slots.end() || (*it)->last_used < (*victim)->last_used))
                        victim = it;
                if (victim == S->slots.end()) break;
                S->bytes_in_use -= std::min(S->bytes_in_use, (*victim)->ctx->dev_bytes);
                S->slots_released += 1;
                victims.push_back(std::move(*victim));
                S->slots.erase(victim);
            }
            if (!victims.empty()) {
                lk.unlock();
                for (auto &v : victims) destroy_slot(v.get());
                victims.clear();
                lk.lock();
            }
            if (S->bytes_in_use + need > S->budget) {
                if (S->active == 0) {  // can never fit
                    S->pending.erase(std::find(S->pending.begin(), S->pending.end(), q));
                    lk.unlock();
                    char msg[160];
                    snprintf(msg, sizeof(msg), "a slot of %zu bytes exceeds the HBM budget of %zu bytes", need,
                             S->budget);
                    reject(S, q, WL_E_OOM, msg);
                    lk.lock();
                    continue;
                }
                // q waits for memory: another request that can start now goes first
                std::shared_ptr<Request> pass;
                if (q->bypassed < kMaxPassed) {
                    const size_t n = std::min(S->pending.size(), (size_t)kBypassWindow + 1);
                    for (size_t i = 0; i < n && !pass; ++i)
                        if (S->pending[i] != q && admissible_now(S, S->pending[i]->r, &slot, &need))
                            pass = S->pending[i];
                }
                if (!pass) {
                    blocked_at = S->generation;
                    blocked_qlen = S->pending.size();
                    continue;
                }
                q = pass;  // take_from_queue counts the pass-over of every request ahead
            } else {
                slot = nullptr;  // a new slot for the head
            }
        }
        const wl_solve_request &r = q->r;
        if (!slot) {
            std::unique_ptr<Slot> sp;
            const std::vector<std::pair<int, int>> opts = S->slot_opts;
            lk.unlock();
            int rc = create_slot(S, r, slot_capacity(r.nrhs), opts, &sp);
            lk.lock();
            if (rc) {
                S->pending.erase(std::find(S->pending.begin(), S->pending.end(), q));
                const std::string msg = last_error();
                lk.unlock();
                reject(S, q, rc, msg.c_str());
                lk.lock();
                continue;
            }
            slot = sp.get();
            S->bytes_in_use += slot->ctx->dev_bytes;
            S->slots.push_back(std::move(sp));
            S->slots_created += 1;
        }
        blocked_at = ~(uint64_t)0;
        take_from_queue(S, q);
        slot->busy = true;
        slot->req = q;
        S->active += 1;
        const size_t before = slot->ctx->dev_bytes;
        lk.unlock();
        const int rc = load_slot(slot, q->r);
        const std::string msg = rc ? last_error() : "";
        lk.lock();
        S->bytes_in_use += slot->ctx->dev_bytes - before;
        q->t_admit = Clock::now();
        if (rc) {
            lk.unlock();
            complete(S, slot, rc, msg.c_str());
            lk.lock();
            continue;
        }
        S->admitted.push_back(slot);
        S->cv_sched.notify_one();
    }
}

void scheduler_main(wl_server *S) {
    (void)hipSetDevice(S->device);
    std::vector<Slot *> running;
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (running.empty())
            S->cv_sched.wait(lk, [&] { return S->stop_sched || !S->admitted.empty(); });
        if (running.empty() && S->admitted.empty() && S->stop_sched) return;
        std::deque<Slot *> fresh;
        fresh.swap(S->admitted);
        lk.unlock();
        for (Slot *s : fresh) {
            const wl_solve_request &r = s->req->r;
            // kernel configuration: the autotuned one for this shape if wl_autotune chose one
            // in this process, else the defaults (never timed here)
            wl_ctx *c = s->ctx;
            wl_tune_result tuned{};
            if (tuned_config(c, r.inner_prec, r.nrhs, &tuned)) {
                apply_hop_config(c, tuned);
            } else {
                c->rhs_block = 0;
                c->rhs_launch = 0;
                c->g.blk = dslash_block_auto(c->g.vh);
                c->blk_auto = true;
                set_lds_reserve(c, 0);
            }
            CgSpec spec;
            spec.kappa = r.kappa;
            spec.mh = r.mh;
            spec.accuracy = r.tolerance;
            spec.maxiters = r.maxiters;
            spec.prec = r.inner_prec;
            spec.mixed = r.inner_prec != WL_F64;
            spec.delta = r.delta;
            spec.update_every = r.update_every;
            const int rc = s->cg.begin(s->ctx, s->x->data, s->b->data, r.nrhs, spec);
            if (rc) {
                const std::string msg = last_error();
                complete(S, s, rc, msg.c_str());
            } else {
                running.push_back(s);
            }
        }
        bool any_done = false;
        for (size_t i = 0; i < running.size();) {
            Slot *s = running[i];
            bool fin = false;
            const int rc = s->cg.pump(&fin);
            if (rc || fin) {
                running[i] = running.back();
                running.pop_back();
                if (rc) {
                    const std::string msg = last_error();
                    complete(S, s, rc, msg.c_str());
                } else {
                    std::lock_guard<std::mutex> g(S->mu);
                    S->finished.push_back(s);
                    any_done = true;
                }
                continue;
            }
            ++i;
        }
        if (any_done) S->cv_unload.notify_one();
        if (!running.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
    }
}

void unloader_main(wl_server *S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_unload.wait(lk, [&] { return S->stop_unload || !S->finished.empty(); });
        if (S->finished.empty()) {
            if (S->stop_unload) return;
            continue;
        }
        Slot *s = S->finished.front();
        S->finished.pop_front();
        lk.unlock();
        const wl_solve_request &r = s->req->r;
        std::shared_ptr<Request> q = s->req;
        q->stats.assign(r.nrhs, wl_cg_stats{});
        int status = s->cg.results(q->stats.data());
        int rc = WL_OK;
        const size_t blk = (size_t)s->ctx->g.vh * 24;
        for (int j = 0; j < r.nrhs && rc == WL_OK; ++j)
            rc = field_download_async(s->ctx, s->x->data, WL_F64, s->cap, j, r.solution + j * blk);
        if (rc == WL_OK && hipStreamSynchronize(s->ctx->stream) != hipSuccess)
            rc = fail(WL_E_DEVICE, "write-back of the solutions failed");
        std::string msg;
        if (rc) {
            status = rc;
            msg = last_error();
        } else {
            int it = 0, ru = 0;
            double tr = 0;
            for (const wl_cg_stats &st : q->stats) {
                it = std::max(it, st.iterations);
                ru = std::max(ru, st.reliable_updates);
                tr = std::max(tr, st.true_resid);
            }
            q->res.iterations = it;
            q->res.reliable_updates = ru;
            q->res.true_resid = tr;
            if (status == WL_E_NOCONV) msg = "not every right-hand side converged";
        }
        complete(S, s, status, msg.empty() ? nullptr : msg.c_str());
        lk.lock();
    }
}

int check_request(const wl_solve_request *r) {
    if (!r) return fail(WL_E_ARG, "null request");
    for (int i = 0; i < 4; ++i) {
        if (r->dims[i] < 2 || (r->dims[i] & 1))
            return fail(WL_E_ARG, "lattice size %d in direction %d: must be even and >= 2", r->dims[i], i);
        if (!r->config_path && !r->gauge[i]) return fail(WL_E_ARG, "null gauge direction %d", i);
    }
    if (r->nrhs < 1 || r->nrhs > 64) return fail(WL_E_ARG, "nrhs %d out of range 1..64", r->nrhs);
    if (!r->source || !r->solution) return fail(WL_E_ARG, "null source or solution buffer");
    if (!(r->tolerance > 0)) return fail(WL_E_ARG, "tolerance must be > 0");
    if (!std::isfinite(r->mh) || !std::isfinite(r->kappa)) return fail(WL_E_ARG, "kappa and mh must be finite");
    if (r->inner_prec != WL_F64 && r->inner_prec != WL_F32 && r->inner_prec != WL_F16)
        return fail(WL_E_NOTIMPL, "inner precision %d", r->inner_prec);
    if (r->inner_prec != WL_F64 && (!(r->delta > 0 && r->delta < 1) || r->update_every < 1))
        return fail(WL_E_ARG, "mixed precision needs 0 < delta < 1 and update_every >= 1");
    return WL_OK;
}

}  // namespace

extern "C" int wl_server_create(int device, size_t hbm_budget_bytes, int max_active, wl_server **out) {
    if (!out) return fail(WL_E_ARG, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WL_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(WL_E_ARG, "device %d of %d", device, ndev);
    DeviceGuard dg(device);
    if (hbm_budget_bytes == 0) {
        size_t freeb = 0, total = 0;
        HIPCHK(hipMemGetInfo(&freeb, &total));
        hbm_budget_bytes = freeb / 2;
    }
    wl_server *S = new (std::nothrow) wl_server;
    if (!S) return fail(WL_E_OOM, "host allocation");
    S->device = device;
    S->budget = hbm_budget_bytes;
    S->max_active = max_active > 0 ? max_active : kDefaultMaxActive;
    try {
        S->loader = std::thread(loader_main, S);
        S->scheduler = std::thread(scheduler_main, S);
        S->unloader = std::thread(unloader_main, S);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = S->stop_sched = S->stop_unload = true;
        }
        S->cv_load.notify_all();
        S->cv_sched.notify_all();
        S->cv_unload.notify_all();
        for (std::thread *t : {&S->loader, &S->scheduler, &S->unloader})
            if (t->joinable()) t->join();
        delete S;
        return fail(WL_E_OOM, "server threads could not be started");
    }
    *out = S;
    return WL_OK;
}

extern "C" int wl_server_submit(wl_server *S, const wl_solve_request *req, uint64_t *ticket) {
    if (!S || !ticket) return fail(WL_E_ARG, "null argument");
    if (int rc = check_request(req)) return rc;
    auto q = std::make_shared<Request>();
    q->r = *req;
    if (req->config_path) {
        q->config_path = req->config_path;
        q->r.config_path = q->config_path.c_str();
    }
    q->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return fail(WL_E_STATE, "server is shutting down");
        q->ticket = S->next_ticket++;
        S->submitted = true;
        S->requests[q->ticket] = q;
        S->pending.push_back(q);
        *ticket = q->ticket;
    }
    S->cv_load.notify_all();
    return WL_OK;
}

extern "C" int wl_server_provision(wl_server *S, const int dims[4], int antiperiodic_t, int inner_prec, int nrhs,
                                   int count, int *created) {
    if (!S || !dims || !created) return fail(WL_E_ARG, "null argument");
    *created = 0;
    for (int i = 0; i < 4; ++i)
        if (dims[i] < 2 || (dims[i] & 1))
            return fail(WL_E_ARG, "lattice size %d in direction %d: must be even and >= 2", dims[i], i);
    if (nrhs < 1 || nrhs > 64) return fail(WL_E_ARG, "nrhs %d out of range 1..64", nrhs);
    if (inner_prec != WL_F64 && inner_prec != WL_F32 && inner_prec != WL_F16)
        return fail(WL_E_NOTIMPL, "inner precision %d", inner_prec);
    if (count < 0) return fail(WL_E_ARG, "count %d < 0", count);
    wl_solve_request r{};
    std::memcpy(r.dims, dims, sizeof(r.dims));
    r.antiperiodic_t = antiperiodic_t;
    r.inner_prec = inner_prec;
    r.nrhs = nrhs;
    const int cap = slot_capacity(nrhs);
    // the whole call under the server lock: admissions wait until the slots exist
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(WL_E_STATE, "server is shutting down");
    S->submitted = true;  // the slots are made with the options set so far
    const size_t need = slot_estimate(dims, inner_prec, cap, slot_link_rows(S->slot_opts));
    for (int k = 0; k < count; ++k) {
        if (S->bytes_in_use + need > S->budget) break;
        std::unique_ptr<Slot> sp;
        int rc = create_slot(S, r, cap, S->slot_opts, &sp);
        if (!rc) rc = BatchCG::reserve(sp->ctx, cap, inner_prec, inner_prec != WL_F64);
        if (!rc && hipStreamSynchronize(sp->ctx->stream) != hipSuccess)
            rc = fail(WL_E_DEVICE, "provisioning a slot failed");
        if (rc) {
            if (sp) destroy_slot(sp.get());
            return rc;
        }
        sp->last_used = Clock::now();
        S->bytes_in_use += sp->ctx->dev_bytes;
        S->slots.push_back(std::move(sp));
        S->slots_created += 1;
        *created += 1;
    }
    S->cv_load.notify_all();
    return WL_OK;
}

extern "C" int wl_server_wait(wl_server *S, uint64_t ticket, int timeout_ms, wl_solve_result *res,
                              wl_cg_stats *per_rhs) {
    if (!S || !res) return fail(WL_E_ARG, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->requests.find(ticket);
    if (it == S->requests.end()) return fail(WL_E_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    std::shared_ptr<Request> q = it->second;
    auto ready = [&] { return q->done; };
    if (timeout_ms < 0)
        S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return fail(WL_E_BUSY, "request %llu not finished", (unsigned long long)ticket);
    *res = q->res;
    if (per_rhs)
        for (int j = 0; j < q->r.nrhs; ++j) per_rhs[j] = j < (int)q->stats.size() ? q->stats[j] : wl_cg_stats{};
    S->requests.erase(ticket);
    return WL_OK;
}

extern "C" int wl_server_info_get(wl_server *S, wl_server_info *info) {
    if (!S || !info) return fail(WL_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    info->budget_bytes = S->budget;
    info->bytes_in_use = S->bytes_in_use;
    info->slots = (int)S->slots.size();
    info->active = S->active;
    info->queued = (int)S->pending.size();
    info->max_active = S->max_active;
    info->completed = S->completed;
    info->slots_created = S->slots_created;
    info->slots_released = S->slots_released;
    return WL_OK;
}

extern "C" int wl_server_set_option(wl_server *S, int option, int value) {
    if (!S) return fail(WL_E_ARG, "null server");
    if (option < WL_OPT_TRAVERSAL_CHUNK || option > WL_OPT_LINK_ROWS) return fail(WL_E_ARG, "unknown option %d", option);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->submitted) return fail(WL_E_STATE, "server options must be set before the first submission");
    S->slot_opts.emplace_back(option, value);
    return WL_OK;
}

extern "C" int wl_server_destroy(wl_server *S) {
    if (!S) return WL_OK;
    // finish everything submitted: the loader drains the queue, then stops
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_load.notify_all();
    if (S->loader.joinable()) S->loader.join();
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_done.wait(lk, [&] { return S->active == 0; });
        S->stop_sched = S->stop_unload = true;
    }
    S->cv_sched.notify_all();
    S->cv_unload.notify_all();
    if (S->scheduler.joinable()) S->scheduler.join();
    if (S->unloader.joinable()) S->unloader.join();
    {
        DeviceGuard dg(S->device);
        for (auto &sp : S->slots) destroy_slot(sp.get());
        S->slots.clear();
        S->bytes_in_use = 0;
    }
    delete S;
    return WL_OK;
}

extern "C" int wl_host_alloc(size_t bytes, void **out) {
    if (!out) return fail(WL_E_ARG, "null argument");
    *out = nullptr;
    if (bytes == 0) return WL_OK;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return fail(WL_E_OOM, "pinned host allocation of %zu bytes", bytes);
    }
    return WL_OK;
}

extern "C" int wl_host_free(void *p) {
    if (p) (void)hipHostFree(p);
    return WL_OK;
}
