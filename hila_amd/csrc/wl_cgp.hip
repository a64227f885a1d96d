// Persistent conjugate gradient on M^dag M (libraries/dirac/conjugate_gradient.h:46-104): the
// whole FP64 single-rhs solve -- source norm, initial residual, every iteration and the final
// true residual -- in one cooperative launch, with no host round trip and no kernel boundary
// inside the solve (WL_OPT_CG_PERSISTENT = 1; the kernel-by-kernel solve is wl_cg.hip, cg_f64).
//
// The workgroups meet at grid barriers between the passes that read other workgroups'
// output (a hopping pass reads neighbours; a reduction needs every partial sum).  Each pass
// runs the per-site code of the kernel-by-kernel solve (wl_dslash.h, wl_blas.h) over the same
// sites, grouped into the same "virtual" workgroups, and writes the same per-workgroup partial
// sums; every workgroup then reduces the partials itself, in the order of the one-workgroup
// finish kernel (k_finish), and applies the same scalar update (fin_apply) to its own copy of
// the CG state in LDS.  So every workgroup holds identical scalars, takes the same branches,
// and the solution equals the kernel-by-kernel one bit for bit.
//
// Per iteration: five barriers (after t = -kappa H_oe p; after Dp = p + kappa H_eo t with the
// partials of ||Dp||^2; after t = -kappa H_oe^dag Dp; after the x, r update with the partials
// of ||r||^2; after p = beta p + r).  A barrier spins a bounded number of times; a workgroup
// that gives up sets an error flag that the others see, so every workgroup leaves the kernel,
// and the host reports the failure.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "wl_blas.h"
#include "wl_dslash.h"
#include "wl_host.h"

namespace wl {

struct GridBar {
    unsigned count;  // workgroups arrived at the current barrier
    unsigned gen;    // barriers completed
    int err;         // a workgroup gave up waiting
    int pad;
};

namespace {

// about 2^22 polls of up to ~1 us: seconds, far beyond any barrier of a running solve
constexpr unsigned kSpinLimit = 1u << 22;

// Grid-wide barrier of a cooperative launch; false when it timed out (here or elsewhere).
__device__ bool grid_sync(GridBar *bar) {
    __shared__ int ok;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this wave's writes, before arriving
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        const unsigned g = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1) {
            __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            unsigned spins = 0;
            while (__hip_atomic_load(&bar->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (++spins > kSpinLimit || __hip_atomic_load(&bar->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(&bar->err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other workgroups' writes
    return ok != 0;
}

struct CgpArgs {
    Geom g;                    // the hopping passes' geometry; g.blk = blockDim.x
    const cx<double> *gauge;   // FP64 links (two-row copy when LF = 6)
    double kappa;
    double bshift;             // 1 + the Hasenbusch shift
    cx<double> *x;
    const cx<double> *b;
    cx<double> *r, *p, *Dp, *DDp, *t;
    double *partials;
    CgState *st;               // initial state in, final state out
    GridBar *bar;
    int nbh;                   // workgroups of a hopping pass: ceil(vh / g.blk)
    int nb256;                 // workgroups of a 256-lane BLAS pass: ceil(vh / 256)
};

// k_finish: sum the n partials in its order and apply `fin` to this workgroup's state.
__device__ void finish(const CgpArgs &P, int n, int fin, CgState *sst) {
    const double s = vthreads256_sum([&](int v) {
        double a = 0;
        for (int i = v; i < n; i += kBlock) a += P.partials[i];
        return a;
    });
    if (threadIdx.x == 0) {
        RedArgs ra;
        ra.partials = P.partials;
        ra.st = sst;
        ra.out = nullptr;
        fin_apply(fin, s, ra);
    }
    __syncthreads();
}

// One hopping pass (k_dslash with epilogue EPI) over virtual workgroups vb = blockIdx.x,
// blockIdx.x + gridDim.x, ... of nbh; partial sums at partials[vb] for the norm epilogues.
template <int SIGN, int EPI, int LF>
__device__ void hop_pass(const CgpArgs &P, cx<double> *out, int out_par, const cx<double> *in,
                         const cx<double> *xin, double a, CgState *sst) {
    DslashArgs<double> A;
    A.out = out;
    A.in = in;
    A.xin = xin;
    A.gauge = P.gauge;
    A.a = a;
    A.b = P.bshift;
    A.g = P.g;
    A.out_par = out_par;
    A.done = nullptr;
    A.x = P.x;
    A.p = P.p;
    A.r = P.r;
    A.red.partials = P.partials;
    A.red.st = sst;
    A.red.out = nullptr;
    A.fin = FIN_NONE;
    A.fold_parts = nullptr;
    A.fold_n = 0;
    A.fold_it = 0;
    for (int vb = blockIdx.x; vb < P.nbh; vb += gridDim.x) {
        const int k = lane_site_of(P.g, vb, P.nbh);
        double nrm = 0;
        if (k < P.g.vh) nrm = dslash_site<double, SIGN, EPI, LF>(A, k, sst->alpha);
        if constexpr (EPI == EPI_AXPY_NORM || EPI == EPI_CGUPD) {
            const double bs = block_sum(nrm);
            if (threadIdx.x == 0) P.partials[vb] = bs;
        }
    }
}

// One 256-lane BLAS pass (f: site -> its term of the reduction, with its writes) over virtual
// workgroups of 256 sites; partial sums at partials[vb] when `reduce`.
template <bool reduce, typename F>
__device__ void blas_pass(const CgpArgs &P, F f) {
    for (int vb = blockIdx.x; vb < P.nb256; vb += gridDim.x) {
        if constexpr (reduce) {
            const double s = vthreads256_sum([&](int v) {
                const int k = vb * kBlock + v;
                return k < P.g.vh ? f(k) : 0.0;
            });
            if (threadIdx.x == 0) P.partials[vb] = s;
        } else {
            for (int v = threadIdx.x; v < kBlock; v += blockDim.x) {
                const int k = vb * kBlock + v;
                if (k < P.g.vh) f(k);
            }
        }
    }
}

// out = M_s in (wilson.h:272-289), two hopping passes through t, each followed by a barrier
#define WL_SYNC()                          \
    do {                                   \
        if (!grid_sync(P.bar)) return false; \
    } while (0)

template <int SIGN, int LF>
__device__ bool meo_pass(const CgpArgs &P, cx<double> *out, const cx<double> *in, CgState *sst) {
    hop_pass<SIGN, EPI_SET, LF>(P, P.t, 1, in, nullptr, -P.kappa, sst);
    WL_SYNC();
    hop_pass<SIGN, EPI_AXPY, LF>(P, out, 0, P.t, in, P.kappa, sst);
    WL_SYNC();
    return true;
}

template <int LF>
__device__ bool cg_body(const CgpArgs &P, CgState *sst) {
    const int S = P.g.stride;
    // source norm and target (conjugate_gradient.h:61-63)
    blas_pass<true>(P, [&](int k) { return site_norm2<double>(P.b, S, k); });
    WL_SYNC();
    finish(P, P.nb256, FIN_SOURCE, sst);
    if (sst->done) return true;  // b = 0: the host sets x = 0
    // r = b - M^dag M x, p = r, rr = |r|^2 (:65-73)
    if (!meo_pass<+1, LF>(P, P.Dp, P.x, sst)) return false;
    if (!meo_pass<-1, LF>(P, P.DDp, P.Dp, sst)) return false;
    blas_pass<true>(P, [&](int k) { return site_cg_init(P.b, P.DDp, P.r, P.p, S, k); });
    WL_SYNC();
    finish(P, P.nb256, FIN_RR, sst);
    if (threadIdx.x == 0 && sst->maxiters <= 0) sst->done = 2;
    __syncthreads();
    // iterations (:75-95)
    while (sst->done == 0) {
        hop_pass<+1, EPI_SET, LF>(P, P.t, 1, P.p, nullptr, -P.kappa, sst);
        WL_SYNC();
        hop_pass<+1, EPI_AXPY_NORM, LF>(P, P.Dp, 0, P.t, P.p, P.kappa, sst);
        WL_SYNC();
        finish(P, P.nbh, FIN_ALPHA, sst);
        hop_pass<-1, EPI_SET, LF>(P, P.t, 1, P.Dp, nullptr, -P.kappa, sst);
        WL_SYNC();
        hop_pass<-1, EPI_CGUPD, LF>(P, nullptr, 0, P.t, P.Dp, P.kappa, sst);
        WL_SYNC();
        finish(P, P.nbh, FIN_RRNEW, sst);
        if (sst->done) break;
        const double beta = sst->beta;
        blas_pass<false>(P, [&](int k) {
            site_xpay<double>(P.p, P.r, beta, S, k);
            return 0.0;
        });
        WL_SYNC();
    }
    // true residual |b - M^dag M x|^2 in FP64
    if (!meo_pass<+1, LF>(P, P.Dp, P.x, sst)) return false;
    if (!meo_pass<-1, LF>(P, P.DDp, P.Dp, sst)) return false;
    blas_pass<true>(P, [&](int k) { return site_diff_norm2(P.b, P.DDp, S, k); });
    WL_SYNC();
    finish(P, P.nb256, FIN_TRUE, sst);
    return true;
}
#undef WL_SYNC

template <int LF>
__global__ __launch_bounds__(kBlock) void k_cg_persistent(CgpArgs P) {
    __shared__ CgState sst;
    if (threadIdx.x == 0) sst = *P.st;
    __syncthreads();
    const bool ok = cg_body<LF>(P, &sst);
    if (ok && blockIdx.x == 0 && threadIdx.x == 0) *P.st = sst;
}

// Workgroups of the cooperative launch: one per CU at most, and at most one per virtual
// workgroup of a hopping pass, within what the device holds at once.
template <int LF>
int persistent_grid(const wl_ctx *c) {
    int ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_cg_persistent<LF>, c->g.blk, 0) !=
        hipSuccess)
        return 0;
    const int nbh = (c->g.vh + c->g.blk - 1) / c->g.blk;
    return std::min(nbh, std::min(ncu, ncu * per_cu));
}

}  // namespace

// Only on request: measured on MI355X the one-launch solve is not faster at any size (8^4:
// 1.05 against 0.96 ms per solve; 16^4: 18.0 against 4.1 ms).  A grid barrier must write back
// and invalidate each XCD's L2 (agent-scope release and acquire), which costs more than the
// kernel boundary it replaces.
bool cg_persistent_wanted(const wl_ctx *c) { return c->cg_persistent > 0; }

__global__ void k_zero_words(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

hipError_t launch_zero_words(void *p, int words, hipStream_t s) {
    if (words <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, (unsigned *)p, words);
    return hipGetLastError();
}

// The persistent FP64 CG on work memory w; *ran = false (and nothing done) when the launch does
// not fit the device, so that the caller solves kernel by kernel.
int cg_f64_persistent(wl_ctx *c, const CgWork &w, void *x, size_t x_bytes, const void *b, double kappa,
                      double accuracy, int maxiters, wl_cg_stats *stats, bool *ran) {
    *ran = false;
    const auto t0 = std::chrono::steady_clock::now();
    const int lf = link_format(c, WL_F64);
    const int G = lf == 6 ? persistent_grid<6>(c) : persistent_grid<9>(c);
    if (G <= 0) return WL_OK;
    const int nbh = (c->g.vh + c->g.blk - 1) / c->g.blk;
    CgpArgs P;
    P.g = c->g;
    P.gauge = (const cx<double> *)gauge_for(c, WL_F64);
    if (!P.gauge) return fail(WL_E_STATE, "FP64 links not prepared");
    P.kappa = kappa;
    P.bshift = 1.0 + c->op_shift;
    P.x = (cx<double> *)x;
    P.b = (const cx<double> *)b;
    P.r = (cx<double> *)w.r;
    P.p = (cx<double> *)w.p;
    P.Dp = (cx<double> *)w.Dp;
    P.DDp = (cx<double> *)w.DDp;
    P.t = (cx<double> *)w.t;
    P.partials = w.partials;
    P.st = w.st;
    P.bar = (GridBar *)w.bar;
    P.nbh = nbh;
    P.nb256 = nblocks(c->g);

    CgState init{};
    init.acc2 = accuracy * accuracy;
    init.maxiters = maxiters;
    *c->st_host = init;
    HIPCHK(hipMemcpyAsync(w.st, c->st_host, sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    HIPCHK(launch_zero_words(w.bar, (int)(sizeof(GridBar) / 4), c->stream));
    void *args[] = {&P};
    const void *fn = lf == 6 ? (const void *)k_cg_persistent<6> : (const void *)k_cg_persistent<9>;
    const hipError_t le = hipLaunchCooperativeKernel(fn, dim3(G), dim3(c->g.blk), args, 0, c->stream);
    if (le == hipErrorCooperativeLaunchTooLarge) {
        (void)hipGetLastError();
        return WL_OK;  // not run: the caller solves kernel by kernel
    }
    HIPCHK(le);
    *ran = true;
    HIPCHK(hipMemcpyAsync(c->st_host, w.st, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(c->scratch_h, w.bar, sizeof(GridBar), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    GridBar bar;
    std::memcpy(&bar, c->scratch_h, sizeof(bar));
    if (bar.err) return fail(WL_E_DEVICE, "persistent CG: a grid barrier timed out");
    const CgState &S = *c->st_host;
    if (S.source_norm == 0.0) {
        // b = 0: the solution is x = 0 (HILA would divide 0 by 0 here)
        HIPCHK(hipMemsetAsync(x, 0, x_bytes, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->converged = 1;
        }
        return WL_OK;
    }
    const bool conv = S.done == 1;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->iterations = S.iter;
        stats->converged = conv ? 1 : 0;
        stats->rr_rel = S.rrnew / S.source_norm;
        stats->true_resid = std::sqrt(S.true_rr / S.source_norm);
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (!conv)
        return fail(WL_E_NOCONV, "CG did not converge in %d iterations (rr_rel %.3e)", S.iter,
                    S.rrnew / S.source_norm);
    return WL_OK;
}

}  // namespace wl
