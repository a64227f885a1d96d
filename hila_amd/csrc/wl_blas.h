// Per-site bodies of the single-rhs BLAS-1 and reduction kernels of the CG (wl_blas.hip),
// shared with the persistent CG kernel (wl_cgp.hip) so that both compute every site, and every
// partial sum, identically.
#pragma once

#include "wl_internal.h"
#include "wl_store.h"

namespace wl {

// ||f(k)||^2 of site k, the components summed in order
template <typename T>
__device__ __forceinline__ double site_norm2(const typename St<T>::E *f, int S, int k) {
    cx<typename St<T>::R> w[4][3];
    ld_site<T>(f, S, k, w);
    double s = 0;
#pragma unroll
    for (int c = 0; c < 12; ++c) s += (double)norm2(w[c / 3][c % 3]);
    return s;
}

// r = p = b - ddp at site k (FP64); returns ||r(k)||^2 (conjugate_gradient.h:65-73)
__device__ __forceinline__ double site_cg_init(const cx<double> *__restrict__ b, const cx<double> *__restrict__ ddp,
                                               cx<double> *__restrict__ r, cx<double> *__restrict__ p, int S, int k) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const size_t e = comp_elem<double>(c, S, k);
        const cx<double> v = b[e] - ddp[e];
        r[e] = v;
        p[e] = v;
        s += norm2(v);
    }
    return s;
}

// p = beta p + r at site k (conjugate_gradient.h:94)
template <typename R>
__device__ __forceinline__ void site_xpay(cx<R> *__restrict__ p, const cx<R> *__restrict__ r, R beta, int S, int k) {
    cx<R> pv[4][3], rv[4][3];
    ld_site<R>(p, S, k, pv);
    ld_site<R>(r, S, k, rv);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) pv[s][c] = beta * pv[s][c] + rv[s][c];
    st_site<R>(p, S, k, pv);
}

// ||b(k) - v(k)||^2 (FP64)
__device__ __forceinline__ double site_diff_norm2(const cx<double> *__restrict__ b, const cx<double> *__restrict__ v,
                                                  int S, int k) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const size_t e = comp_elem<double>(c, S, k);
        s += norm2(b[e] - v[e]);
    }
    return s;
}

}  // namespace wl
