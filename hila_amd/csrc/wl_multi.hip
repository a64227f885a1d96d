// Batched-CG support kernels for N right-hand sides (gfx950), and the dispatch of the
// multi-right-hand-side hopping kernel (wl_rb.h) by storage format.
//
// Batched CG keeps one CgState per right-hand side: each runs HILA's recurrence
// (conjugate_gradient.h:75-95) with its own alpha, beta and convergence flag, and all share
// every operator application.  Per-rhs partial sums are reduced in a fixed order.

#include "wl_hop.h"
#include "wl_internal.h"
#include "wl_store.h"

namespace wl {

// ----------------------------------------------------------- batched-CG helpers ---------
// Grid (site blocks of kBlock, N right-hand sides): blockIdx.y = rhs.  Per-rhs partial sums
// land in partials[rhs * gridDim.x + blockIdx.x].

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm2_mr(const typename St<T>::E *__restrict__ f, int vh,
                                                     int S, double *partials) {
    const int k = blockIdx.x * kBlock + threadIdx.x, j = blockIdx.y;
    double s = 0;
    if (k < vh) {
        cx<typename St<T>::R> w[4][3];
        ld_site<T>(f + (size_t)j * St<T>::planes * S, S, k, w);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) s += (double)norm2(w[a][c]);
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[(size_t)j * gridDim.x + blockIdx.x] = s;
}

// plain CG start: r = b - DDp, p = r, ||r||^2 per rhs (conjugate_gradient.h:67-72)
__global__ __launch_bounds__(kBlock) void k_cg_init_mr(const cx<double> *__restrict__ b,
                                                       const cx<double> *__restrict__ ddp,
                                                       cx<double> *__restrict__ r,
                                                       cx<double> *__restrict__ p, int vh, int S,
                                                       double *partials) {
    const int k = blockIdx.x * kBlock + threadIdx.x, j = blockIdx.y;
    double s = 0;
    if (k < vh) {
        const size_t off = (size_t)j * 12 * S;
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const size_t e = off + comp_elem<double>(c, S, k);
            const cx<double> v = b[e] - ddp[e];
            r[e] = v;
            p[e] = v;
            s += norm2(v);
        }
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[(size_t)j * gridDim.x + blockIdx.x] = s;
}

// reliable update of the rhs that asked for one (done 3 or 4): r_lo = b - DDp (FP64 in),
// ||r||^2 per rhs; with first != 0 every rhs is updated (start of the solve).  Skipped when
// *gate != 0 (no update pending).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual_update_mr(
    const cx<double> *__restrict__ b, const cx<double> *__restrict__ ddp, typename St<T>::E *__restrict__ r,
    const CgState *st, int first, int vh, int S, double *partials, const int *gate) {
    using R = typename St<T>::R;
    if (gate && *gate) return;
    const int k = blockIdx.x * kBlock + threadIdx.x, j = blockIdx.y;
    const bool on = first || st[j].done == 3 || st[j].done == 4;  // block-uniform
    double s = 0;
    if (on && k < vh) {
        const size_t off = (size_t)j * 12 * S;
        cx<R> v[4][3];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const size_t e = off + comp_elem<double>(a * 3 + c, S, k);
                const cx<double> d = b[e] - ddp[e];
                v[a][c] = {(R)d.x, (R)d.y};
                s += norm2(d);
            }
        st_site<T>(r + (size_t)j * St<T>::planes * S, S, k, v);
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[(size_t)j * gridDim.x + blockIdx.x] = s;
}

// p = beta_j p + r for the rhs still running (done == 0)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_xpay_mr(typename St<T>::E *__restrict__ p,
                                                    const typename St<T>::E *__restrict__ r,
                                                    const CgState *st, const int *all_done,
                                                    int vh, int S) {
    using R = typename St<T>::R;
    if (all_done && *all_done) return;
    const int k = blockIdx.x * kBlock + threadIdx.x, j = blockIdx.y;
    if (k >= vh || st[j].done) return;
    const R beta = (R)st[j].beta;
    const size_t off = (size_t)j * St<T>::planes * S;
    cx<R> pv[4][3], rv[4][3];
    ld_site<T>(p + off, S, k, pv);
    ld_site<T>(r + off, S, k, rv);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) pv[a][c] = beta * pv[a][c] + rv[a][c];
    st_site<T>(p + off, S, k, pv);
}

// Fold the inner solution into the FP64 one: x_j += y_j, y_j = 0, for the rhs that take a
// reliable update (done 3 or 4), or for every rhs with all != 0.  Skipped when *gate != 0.
__global__ __launch_bounds__(kBlock) void k_fold_mr(cx<double> *__restrict__ x, cx<float> *__restrict__ y,
                                                    const CgState *st, int all, int vh, int S, const int *gate) {
    if (gate && *gate) return;
    const int k = blockIdx.x * kBlock + threadIdx.x, j = blockIdx.y;
    if (k >= vh || !(all || st[j].done == 3 || st[j].done == 4)) return;
    const size_t off = (size_t)j * 12 * S;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const size_t e = off + comp_elem<double>(c, S, k);
        const cx<float> v = y[e];
        x[e] = x[e] + cx<double>{(double)v.x, (double)v.y};
        y[e] = {0.f, 0.f};
    }
}

// flag = 1 when every rhs has converged (done == 1), else 0 (one workgroup of 64).
__global__ __launch_bounds__(64) void k_all_converged(const CgState *st, int N, int *flag) {
    __shared__ int all;
    if (threadIdx.x == 0) all = 1;
    __syncthreads();
    if ((int)threadIdx.x < N && st[threadIdx.x].done != 1) atomicAnd(&all, 0);
    __syncthreads();
    if (threadIdx.x == 0) *flag = all;
}

// x_j = 0 for one rhs j
__global__ __launch_bounds__(kBlock) void k_zero_rhs(cx<double> *__restrict__ x, int j, int vh,
                                                     int S) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    const size_t off = (size_t)j * 12 * S;
#pragma unroll
    for (int c = 0; c < 12; ++c) x[off + comp_elem<double>(c, S, k)] = {0.0, 0.0};
}

// Finish of a per-rhs reduction: block j sums partials[j * pstride + 0 .. n) in order and
// applies `fin` to rhs j if its state allows it (gate: running rhs for the loop reductions,
// rhs that asked for an update for FIN_UPDATE, every rhs otherwise).
__global__ __launch_bounds__(kBlock) void k_finish_mr(double *partials, int pstride, int n,
                                                      CgState *st, double *out, int fin,
                                                      const int *all_done) {
    if (all_done && *all_done) return;
    const int j = blockIdx.x;
    const int d = st ? st[j].done : 0;
    bool apply = true;
    if (fin == FIN_ALPHA || fin == FIN_RRNEW || fin == FIN_RRNEW_MIXED2) apply = d == 0;
    if (fin == FIN_UPDATE) apply = d == 3 || d == 4;
    if (fin == FIN_UPDATE_FIRST) apply = d != 1;  // rhs with b = 0 are finished already
    if (!apply) return;
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[(size_t)j * pstride + i];
    s = block_sum(s);
    if (threadIdx.x == 0) {
        RedArgs ra;
        ra.partials = partials;
        ra.st = st ? st + j : nullptr;
        ra.out = out ? out + j : nullptr;
        fin_apply(fin == FIN_UPDATE_FIRST ? FIN_UPDATE : fin, s, ra);
    }
}

// After the per-rhs convergence tests of one iteration (one workgroup, thread j = rhs j):
//   plain:  all_done = every rhs finished.
//   mixed:  if any running rhs asked for a reliable update, every running rhs joins it (the
//           update is one batched FP64 M^dag M); otherwise beta_j = rrnew_j / rr_j for the
//           running ones.
// flags[0] = all finished, flags[1] = an update is pending, flags[2] = no update pending (the
// gate of the update kernels, which run only when it is 0).  Skipped when *gate != 0.
__global__ __launch_bounds__(64) void k_batch_sync(CgState *st, int N, int mixed, int *flags,
                                                   const int *gate) {
    if (gate && *gate) return;
    __shared__ int any_req, all_fin;
    const int j = threadIdx.x;
    if (j == 0) {
        any_req = 0;
        all_fin = 1;
    }
    __syncthreads();
    if (j < N && mixed && (st[j].done == 3 || st[j].done == 4)) atomicOr(&any_req, 1);
    __syncthreads();
    if (j < N && mixed && st[j].done == 0) {
        if (any_req) {
            st[j].done = 3;
        } else {
            st[j].beta = st[j].rrnew / st[j].rr;
            st[j].rr = st[j].rrnew;
            if (st[j].iter >= st[j].maxiters) st[j].done = 2;
        }
    }
    __syncthreads();
    if (j < N && !(st[j].done == 1 || st[j].done == 2)) atomicAnd(&all_fin, 0);
    __syncthreads();
    if (j == 0) {
        flags[0] = all_fin;
        flags[1] = any_req;
        flags[2] = !any_req;
    }
}

// ------------------------------------------------------------------- launchers -----------

// Segments (64 sites) of a parity, i.e. the per-rhs partial count of a hopping launch.
int mr_segments(const Geom &g) { return (g.vh + 63) / 64; }

hipError_t launch_dslash_mr_f64(const Geom &g, const MrLaunch &L, hipStream_t s);
hipError_t launch_dslash_mr_f32(const Geom &g, const MrLaunch &L, hipStream_t s);
hipError_t launch_dslash_mr_f16(const Geom &g, const MrLaunch &L, hipStream_t s);
hipError_t launch_dslash_mr_f64_rows(const Geom &g, const MrLaunch &L, hipStream_t s);
hipError_t launch_dslash_mr_f32_rows(const Geom &g, const MrLaunch &L, hipStream_t s);

hipError_t launch_dslash_mr(int prec, const Geom &g, const MrLaunch &L, hipStream_t s) {
    if (L.lf != 9 && (L.lf != 6 || prec == WL_F16)) return hipErrorInvalidValue;
    if (prec == WL_F64) return L.lf == 6 ? launch_dslash_mr_f64_rows(g, L, s) : launch_dslash_mr_f64(g, L, s);
    if (prec == WL_F32) return L.lf == 6 ? launch_dslash_mr_f32_rows(g, L, s) : launch_dslash_mr_f32(g, L, s);
    if (prec == WL_F16) return launch_dslash_mr_f16(g, L, s);
    return hipErrorInvalidValue;
}

hipError_t launch_finish_mr(double *partials, int pstride, int n, int N, CgState *st,
                            double *out, int fin, const int *all_done, hipStream_t s) {
    hipLaunchKernelGGL(k_finish_mr, dim3(N), dim3(kBlock), 0, s, partials, pstride, n, st, out,
                       fin, all_done);
    return hipGetLastError();
}

hipError_t launch_norm2_mr(int prec, const Geom &g, const void *f, int N, double *partials,
                           hipStream_t s) {
    const dim3 grid(nblocks(g), N);
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_norm2_mr<double>, grid, dim3(kBlock), 0, s, (const cx<double> *)f,
                           g.vh, g.stride, partials);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_norm2_mr<float>, grid, dim3(kBlock), 0, s, (const cx<float> *)f, g.vh,
                           g.stride, partials);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_norm2_mr<f16s>, grid, dim3(kBlock), 0, s, (const h2 *)f, g.vh, g.stride,
                           partials);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_cg_init_mr(const Geom &g, const void *b, const void *ddp, void *r, void *p,
                             int N, double *partials, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init_mr, dim3(nblocks(g), N), dim3(kBlock), 0, s,
                       (const cx<double> *)b, (const cx<double> *)ddp, (cx<double> *)r,
                       (cx<double> *)p, g.vh, g.stride, partials);
    return hipGetLastError();
}

hipError_t launch_residual_update_mr(int prec, const Geom &g, const void *b, const void *ddp,
                                     void *r, const CgState *st, int first, int N,
                                     double *partials, const int *gate, hipStream_t s) {
    const dim3 grid(nblocks(g), N);
    const cx<double> *B = (const cx<double> *)b, *DD = (const cx<double> *)ddp;
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_residual_update_mr<double>, grid, dim3(kBlock), 0, s, B, DD, (cx<double> *)r,
                           st, first, g.vh, g.stride, partials, gate);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_residual_update_mr<float>, grid, dim3(kBlock), 0, s, B, DD, (cx<float> *)r, st,
                           first, g.vh, g.stride, partials, gate);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_residual_update_mr<f16s>, grid, dim3(kBlock), 0, s, B, DD, (h2 *)r, st, first,
                           g.vh, g.stride, partials, gate);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_xpay_mr(int prec, const Geom &g, void *p, const void *r, const CgState *st,
                          const int *all_done, int N, hipStream_t s) {
    const dim3 grid(nblocks(g), N);
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_xpay_mr<double>, grid, dim3(kBlock), 0, s, (cx<double> *)p,
                           (const cx<double> *)r, st, all_done, g.vh, g.stride);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_xpay_mr<float>, grid, dim3(kBlock), 0, s, (cx<float> *)p,
                           (const cx<float> *)r, st, all_done, g.vh, g.stride);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_xpay_mr<f16s>, grid, dim3(kBlock), 0, s, (h2 *)p, (const h2 *)r, st, all_done,
                           g.vh, g.stride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_batch_sync(CgState *st, int N, int mixed, int *flags, const int *gate,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_batch_sync, dim3(1), dim3(64), 0, s, st, N, mixed, flags, gate);
    return hipGetLastError();
}

hipError_t launch_fold_mr(const Geom &g, void *x, void *y, const CgState *st, int all, int N, const int *gate,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_fold_mr, dim3(nblocks(g), N), dim3(kBlock), 0, s, (cx<double> *)x, (cx<float> *)y, st, all,
                       g.vh, g.stride, gate);
    return hipGetLastError();
}

hipError_t launch_all_converged(const CgState *st, int N, int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_all_converged, dim3(1), dim3(64), 0, s, st, N, flag);
    return hipGetLastError();
}

hipError_t launch_zero_rhs(const Geom &g, void *x, int j, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_rhs, dim3(nblocks(g)), dim3(kBlock), 0, s, (cx<double> *)x, j,
                       g.vh, g.stride);
    return hipGetLastError();
}

}  // namespace wl
