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
// Epilogues fuse the even-odd operator and CG vector updates (wilson.h:272-289,
// conjugate_gradient.h:77-94) into the same pass; see Epi in wl_internal.h.

#include "wl_internal.h"

namespace wl {

template <typename R>
struct DslashArgs {
    cx<R> *__restrict__ out;
    const cx<R> *__restrict__ in;
    const cx<R> *__restrict__ xin;
    const cx<R> *__restrict__ gauge;
    R a;
    Geom g;
    int out_par;
    const int *done;
    cx<double> *__restrict__ x;
    const cx<double> *__restrict__ p;
    cx<double> *__restrict__ r;
    RedArgs red;
    int fin;
};

template <typename R>
__device__ __forceinline__ void load_spinor(const cx<R> *__restrict__ f, int S, int k,
                                            cx<R> (&w)[4][3]) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) w[s][c] = f[(s * 3 + c) * S + k];
}

template <typename R>
__device__ __forceinline__ void load_link(const cx<R> *__restrict__ u0, int S, int k,
                                          cx<R> (&u)[3][3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) u[i][j] = u0[(i * 3 + j) * S + k];
}

// Both hops along direction MU.  BC: multiply by sf / sb (only the t direction can cross an
// antiperiodic edge).
template <int MU, int SIGN, bool BC, typename R>
__device__ __forceinline__ void hop_dir(cx<R> (&acc)[4][3], const cx<R> *__restrict__ in,
                                        const cx<R> *__restrict__ gout,
                                        const cx<R> *__restrict__ gin, int S, int k, int kf,
                                        int kb, R sf, R sb) {
    {
        cx<R> w[4][3], h[2][3], u[3][3], v[2][3];
        load_spinor(in, S, kf, w);
        load_link(gout + MU * 9 * S, S, k, u);
        project<MU, SIGN>(w, h);
        su3_mul(u, h, v);
        if constexpr (BC) {
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int c = 0; c < 3; ++c) v[q][c] = sf * v[q][c];
        }
        expand_add<MU, SIGN>(acc, v);
    }
    {
        cx<R> w[4][3], h[2][3], u[3][3], v[2][3];
        load_spinor(in, S, kb, w);
        load_link(gin + MU * 9 * S, S, kb, u);
        project<MU, -SIGN>(w, h);
        su3_adj_mul(u, h, v);
        if constexpr (BC) {
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int c = 0; c < 3; ++c) v[q][c] = sb * v[q][c];
        }
        expand_add<MU, -SIGN>(acc, v);
    }
}

template <typename R, int SIGN, int EPI>
__global__ __launch_bounds__(kBlock) void k_dslash(DslashArgs<R> A) {
    if (A.done && *A.done) return;
    const int k = xcd_tile(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const int S = A.g.stride;
    double nrm = 0;
    if (k < A.g.vh) {
        const SiteNb n = neighbours(A.g, k, A.out_par);
        const cx<R> *gout = A.gauge + (size_t)A.out_par * 36 * S;
        const cx<R> *gin = A.gauge + (size_t)(1 - A.out_par) * 36 * S;
        const R sf = (A.g.antiperiodic && n.fwd_edge_t) ? R(-1) : R(1);
        const R sb = (A.g.antiperiodic && n.bwd_edge_t) ? R(-1) : R(1);
        cx<R> acc[4][3];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[s][c] = {R(0), R(0)};
        hop_dir<0, SIGN, false>(acc, A.in, gout, gin, S, k, n.fwd[0], n.bwd[0], R(1), R(1));
        hop_dir<1, SIGN, false>(acc, A.in, gout, gin, S, k, n.fwd[1], n.bwd[1], R(1), R(1));
        hop_dir<2, SIGN, false>(acc, A.in, gout, gin, S, k, n.fwd[2], n.bwd[2], R(1), R(1));
        hop_dir<3, SIGN, true>(acc, A.in, gout, gin, S, k, n.fwd[3], n.bwd[3], sf, sb);

        const R a = A.a;
        if constexpr (EPI == EPI_SET) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) A.out[(s * 3 + c) * S + k] = a * acc[s][c];
        } else if constexpr (EPI == EPI_AXPY || EPI == EPI_AXPY_NORM) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int e = (s * 3 + c) * S + k;
                    const cx<R> o = A.xin[e] + a * acc[s][c];
                    A.out[e] = o;
                    if constexpr (EPI == EPI_AXPY_NORM) nrm += (double)norm2(o);
                }
        } else {  // EPI_CGUPD, FP64
            const double alpha = A.red.st->alpha;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int e = (s * 3 + c) * S + k;
                    const cx<double> ddp = A.xin[e] + a * acc[s][c];
                    A.x[e] = A.x[e] + alpha * A.p[e];
                    const cx<double> rn = A.r[e] - alpha * ddp;
                    A.r[e] = rn;
                    nrm += norm2(rn);
                }
        }
    }
    if constexpr (EPI == EPI_AXPY_NORM || EPI == EPI_CGUPD) reduce_finish(nrm, A.fin, A.red);
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
    A.g = g;
    A.out_par = L.out_par;
    A.done = L.done;
    A.x = (cx<double> *)L.x;
    A.p = (const cx<double> *)L.p;
    A.r = (cx<double> *)L.r;
    A.red = L.red;
    A.fin = L.fin;
    hipLaunchKernelGGL((k_dslash<R, SIGN, EPI>), dim3(nblocks(g)), dim3(kBlock), 0, s, A);
    return hipGetLastError();
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
        if (L.epi == EPI_CGUPD) return hipErrorInvalidValue;
        return L.sign > 0 ? launch_epi<float, 1>(g, L, s) : launch_epi<float, -1>(g, L, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wl
