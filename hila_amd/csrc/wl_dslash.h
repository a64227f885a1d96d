// The per-site body of the fused Wilson hopping kernel and its epilogues (see Epi in
// wl_internal.h): k_dslash (wl_dslash.hip) runs it once per lane, and the persistent CG kernel
// (wl_cgp.hip) runs the same code over the same sites, so that both give identical results.
#pragma once

#include "wl_hop.h"
#include "wl_internal.h"
#include "wl_store.h"

namespace wl {

template <typename R>
struct DslashArgs {
    cx<R> *__restrict__ out;
    const cx<R> *__restrict__ in;
    const cx<R> *__restrict__ xin;
    const cx<R> *__restrict__ gauge;
    R a;
    R b;    // coefficient of xin (1 + the Hasenbusch shift in the second half of M)
    Geom g;
    int out_par;
    const int *done;
    cx<double> *__restrict__ x;   // the solution is accumulated in FP64 at every precision
    const cx<R> *__restrict__ p;
    cx<R> *__restrict__ r;
    RedArgs red;
    int fin;
    // EPI_CGUPD in the launch-folded CG: alpha = rr / (the sum of fold_n partials of the norm
    // pass before, in the finish kernel's order), rr from red.st->rrb[fold_it & 1]; null: alpha
    // from red.st (set by the finish kernel)
    const double *fold_parts;
    int fold_n;
    int fold_it;
};

// alpha of the EPI_CGUPD epilogue (conjugate_gradient.h:79-80).  Folded: every workgroup adds
// the previous pass's partials in the finish kernel's order, so all get FIN_ALPHA's bits, and
// workgroup 0 records pDp and alpha.  Contains barriers when folded: every thread calls it.
template <typename R>
__device__ __forceinline__ double cg_alpha(const DslashArgs<R> &A) {
    if (!A.fold_parts) return A.red.st->alpha;
    const double s = finish_order_sum(A.fold_parts, A.fold_n);
    const double alpha = A.red.st->rrb[A.fold_it & 1] / s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.red.st->pDp = s;
        A.red.st->alpha = alpha;
    }
    return alpha;
}

// Output site k (< vh) of H_SIGN with epilogue EPI; returns the site's contribution to the
// norm the epilogue reduces (0 for the others).  LF: stored entries per link, 9 or 6
// (two-row links, wl_store.h).  alpha: the CG step of EPI_CGUPD.
template <typename R, int SIGN, int EPI, int LF>
__device__ __forceinline__ double dslash_site(const DslashArgs<R> &A, int k, double alpha = 0) {
    const int S = A.g.stride;
    double nrm = 0;
    {
        cx<R> acc[4][3];
        hop_sum<R, SIGN, true, LF>(acc, A.g, k, A.out_par, A.in, A.gauge, 1);

        const R a = A.a;
        if constexpr (EPI == EPI_SET) {
            cx<R> v[4][3];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) v[s][c] = a * acc[s][c];
            st_site<R, true>(A.out, S, k, v);
        } else if constexpr (EPI == EPI_AXPY || EPI == EPI_AXPY_NORM) {
            cx<R> v[4][3];
            ld_site<R, true>(A.xin, S, k, v);
            scale_unless_one(v, A.b);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    v[s][c] = v[s][c] + a * acc[s][c];
                    if constexpr (EPI == EPI_AXPY_NORM) nrm += (double)norm2(v[s][c]);
                }
            st_site<R, true>(A.out, S, k, v);
        } else {  // EPI_CGUPD: x (FP64) += alpha p, r -= alpha DDp in R
            const R alpha_r = (R)alpha;
            cx<R> ddp[4][3], pv[4][3], rv[4][3];
            ld_site<R, true>(A.xin, S, k, ddp);
            scale_unless_one(ddp, A.b);
            ld_site<R, true>(A.p, S, k, pv);
            ld_site<R, true>(A.r, S, k, rv);
            cx<double> xv[4][3];
            ld_site<double, true>(A.x, S, k, xv);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const cx<double> pd = {(double)pv[s][c].x, (double)pv[s][c].y};
                    xv[s][c] = xv[s][c] + alpha * pd;
                    const cx<R> dd = ddp[s][c] + a * acc[s][c];
                    rv[s][c] = rv[s][c] - alpha_r * dd;
                    nrm += (double)norm2(rv[s][c]);
                }
            st_site<double, true>(A.x, S, k, xv);
            st_site<R, true>(A.r, S, k, rv);
        }
    }
    return nrm;
}

}  // namespace wl
