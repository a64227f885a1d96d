// Fused Wilson hopping kernel for gfx950: one wave64 lane per output site of one parity.
//
// For output site x (parity p) it computes H_s in(x) = sum_mu [ U_mu(x)(1 + s g_mu) in(x+mu)
// + U_mu(x-mu)^dag (1 - s g_mu) in(x-mu) ] in one pass: each neighbour spinor is gathered from
// its component planes (coalesced 16-byte loads in FP64), spin-projected in registers to a
// half spinor, multiplied by the link (or its adjoint) and reconstructed into the accumulator.
// This replaces HILA's 8 projection kernels, 8 half-vector temporaries and 4 read-modify-write
// accumulation kernels per half operator (libraries/dirac/wilson.h:15-95) with one kernel that
// reads every link once and writes the output once.  The antiperiodic time boundary is a sign
// on the two t hops that cross the edge (HILA fills negated halo copies instead,
// backend_gpu/field_storage_backend.h:333-384).
//
// The 8 hops are software-pipelined: the link and neighbour spinor of hop h+1 are requested
// before hop h is computed, so every wave keeps 15-25 16-byte loads in flight (the kernel is
// bound by HBM bandwidth and latency, not arithmetic: 0.86 flop/byte in FP64).  Links and the
// output are streamed with nontemporal accesses; neighbour spinors go through the caches, where
// their 8-fold reuse is caught.
//
// Epilogues fuse the even-odd operator and CG vector updates (wilson.h:272-289,
// conjugate_gradient.h:77-94) into the same pass; see Epi in wl_internal.h.

#include "wl_dslash.h"

namespace wl {

// LF: stored entries per link, 9 or 6 (two-row links, wl_store.h).
template <typename R, int SIGN, int EPI, int LF>
__global__ __launch_bounds__(kBlock) void k_dslash(DslashArgs<R> A) {
    if (A.done && *A.done) return;
    const int k = lane_site(A.g);
    double nrm = 0;
    if (k < A.g.vh) nrm = dslash_site<R, SIGN, EPI, LF>(A, k);
    if constexpr (EPI == EPI_AXPY_NORM || EPI == EPI_CGUPD) reduce_partial(nrm, A.red);
}

// Default workgroup size, from measurements on MI355X (256 CUs): 256 lanes from 2^19 sites
// per parity up, 64 lanes for mid-size lattices, 128 below (where a launch is latency-bound).
int dslash_block_auto(int vh) {
    if (vh >= 512 * 1024) return 256;
    if (vh >= 128 * 1024) return 64;
    return 128;
}

int nblocks(const Geom &g) { return (g.vh + kBlock - 1) / kBlock; }

template <typename R, int SIGN, int EPI>
static hipError_t launch_t(const Geom &g, const DslashLaunch &L, hipStream_t s) {
    DslashArgs<R> A;
    A.out = (cx<R> *)L.out;
    A.in = (const cx<R> *)L.in;
    A.xin = (const cx<R> *)L.xin;
    A.gauge = (const cx<R> *)L.gauge;
    A.a = (R)L.a;
    A.b = (R)L.b;
    A.g = g;
    A.out_par = L.out_par;
    A.done = L.done;
    A.x = (cx<double> *)L.x;
    A.p = (const cx<R> *)L.p;
    A.r = (cx<R> *)L.r;
    A.red = L.red;
    A.fin = L.fin;
    const int blk = g.blk;
    const int nb = (g.vh + blk - 1) / blk;
    if (L.lf == 6)
        hipLaunchKernelGGL((k_dslash<R, SIGN, EPI, 6>), dim3(nb), dim3(blk), g.lds, s, A);
    else
        hipLaunchKernelGGL((k_dslash<R, SIGN, EPI, 9>), dim3(nb), dim3(blk), g.lds, s, A);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || (EPI != EPI_AXPY_NORM && EPI != EPI_CGUPD)) return e;
    return launch_finish(L.red, nb, L.fin, L.done, s);
}

template <typename R, int SIGN>
static hipError_t launch_epi(const Geom &g, const DslashLaunch &L, hipStream_t s) {
    switch (L.epi) {
    case EPI_SET: return launch_t<R, SIGN, EPI_SET>(g, L, s);
    case EPI_AXPY: return launch_t<R, SIGN, EPI_AXPY>(g, L, s);
    case EPI_AXPY_NORM: return launch_t<R, SIGN, EPI_AXPY_NORM>(g, L, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dslash(int prec, const Geom &g, const DslashLaunch &L, hipStream_t s) {
    if (prec == WL_F64) {
        if (L.epi == EPI_CGUPD)
            return L.sign > 0 ? launch_t<double, 1, EPI_CGUPD>(g, L, s)
                              : launch_t<double, -1, EPI_CGUPD>(g, L, s);
        return L.sign > 0 ? launch_epi<double, 1>(g, L, s) : launch_epi<double, -1>(g, L, s);
    }
    if (prec == WL_F32) {
        if (L.epi == EPI_CGUPD)
            return L.sign > 0 ? launch_t<float, 1, EPI_CGUPD>(g, L, s)
                              : launch_t<float, -1, EPI_CGUPD>(g, L, s);
        return L.sign > 0 ? launch_epi<float, 1>(g, L, s) : launch_epi<float, -1>(g, L, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wl
