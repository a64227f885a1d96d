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
};

// Output site k (< vh) of H_SIGN with epilogue EPI; returns the site's contribution to the
// norm the epilogue reduces (0 for the others).  LF: stored entries per link, 9 or 6
// (two-row links, wl_store.h).
template <typename R, int SIGN, int EPI, int LF>
__device__ __forceinline__ double dslash_site(const DslashArgs<R> &A, int k) {
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
            const double alpha = A.red.st->alpha;
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
