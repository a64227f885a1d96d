// Fused Wilson hopping kernel for gfx950: one wave64 lane per output site of one parity.
//
// For output site x (parity p) it computes H_s in(x) = sum_mu [ U_mu(x)(1 + s g_mu) in(x+mu)
// + U_mu(x-mu)^dag (1 - s g_mu) in(x-mu) ] in one pass: each neighbour spinor is gathered from
// its site tile (coalesced 16-byte loads in FP64), spin-projected in registers to a
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
    double alpha = 0;
    if constexpr (EPI == EPI_CGUPD) alpha = cg_alpha(A);
    const int k = lane_site(A.g);
    double nrm = 0;
    if (k < A.g.vh) nrm = dslash_site<R, SIGN, EPI, LF>(A, k, alpha);
    if constexpr (EPI == EPI_AXPY_NORM || EPI == EPI_CGUPD) reduce_partial(nrm, A.red);
}

// ------------------------------------------------------------ small lattices ------------
// A launch of the kernel above has one lane per site, so on a small lattice it has few waves
// (8^4: 32 waves for 1024 SIMDs), and each lane waits for about four dependent memory round
// trips of its pipelined 8-hop chain: the launch is bound by memory latency.  Here a workgroup
// of 4 waves covers 64 sites and wave mu computes both hops of direction mu, with the loads of
// both issued before either is used: one round trip.  The four partial sums meet in LDS and
// are added in a fixed order; then wave s applies the epilogue to spin component s of the 64
// sites, with its operands loaded at the start of the kernel.

template <typename R, int SIGN, int MU, int LF>
__device__ __forceinline__ void dir_pair(cx<R> (&acc)[4][3], const DslashArgs<R> &A, int k, const SiteNb &n) {
    const int S = A.g.stride;
    const cx<R> *gout = A.gauge + ((size_t)A.out_par * 4 + MU) * LF * S;
    const cx<R> *gin = A.gauge + ((size_t)(1 - A.out_par) * 4 + MU) * LF * S;
    HopBuf<R> f, b;
    hop_load<true, LF>(f, A.in, gout, S, n.fwd[MU], S, k);
    hop_load<true, LF>(b, A.in, gin, S, n.bwd[MU], S, n.bwd[MU]);
    const R sf = (MU == 3 && A.g.antiperiodic && n.fwd_edge_t) ? R(-1) : R(1);
    const R sb = (MU == 3 && A.g.antiperiodic && n.bwd_edge_t) ? R(-1) : R(1);
    hop_compute<MU, SIGN, true>(acc, f, sf);
    hop_compute<MU, SIGN, false>(acc, b, sb);
}

template <typename R, int SIGN, int EPI, int LF>
__global__ __launch_bounds__(256) void k_dslash_dir(DslashArgs<R> A) {
    if (A.done && *A.done) return;
    double alpha = 0;
    if constexpr (EPI == EPI_CGUPD) alpha = cg_alpha(A);
    __shared__ cx<R> part[4][12][64];  // [direction][component][site]
    const int w = (int)(threadIdx.x >> 6), l = (int)(threadIdx.x & 63);
    const int k = xcd_tile(blockIdx.x, gridDim.x) * 64 + l;
    const bool live = k < A.g.vh;
    const int S = A.g.stride;
    size_t e[3], ex[3];  // elements of spin component w of site k (storage, FP64 solution)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e[c] = comp_elem<R>(3 * w + c, S, k);
        ex[c] = comp_elem<double>(3 * w + c, S, k);
    }
    cx<R> xi[3], pv[3], rv[3];
    cx<double> xv[3];
    if (live) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if constexpr (EPI != EPI_SET) xi[c] = ldc<true>(A.xin + e[c]);
            if constexpr (EPI == EPI_CGUPD) {
                pv[c] = ldc<true>(A.p + e[c]);
                rv[c] = ldc<true>(A.r + e[c]);
                xv[c] = ldc<true>(A.x + ex[c]);
            }
        }
        const SiteNb n = neighbours(A.g, k, A.out_par);
        cx<R> acc[4][3];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[s][c] = {R(0), R(0)};
        switch (w) {  // wave-uniform
        case 0: dir_pair<R, SIGN, 0, LF>(acc, A, k, n); break;
        case 1: dir_pair<R, SIGN, 1, LF>(acc, A, k, n); break;
        case 2: dir_pair<R, SIGN, 2, LF>(acc, A, k, n); break;
        default: dir_pair<R, SIGN, 3, LF>(acc, A, k, n); break;
        }
#pragma unroll
        for (int c = 0; c < 12; ++c) part[w][c][l] = acc[c / 3][c % 3];
    }
    __syncthreads();
    double nrm = 0;
    if (live) {
        const R a = A.a;
        cx<R> h[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const int q = 3 * w + c;
            h[c] = (part[0][q][l] + part[1][q][l]) + (part[2][q][l] + part[3][q][l]);
        }
        if constexpr (EPI == EPI_SET) {
#pragma unroll
            for (int c = 0; c < 3; ++c) stc<true>(A.out + e[c], a * h[c]);
        } else if constexpr (EPI == EPI_AXPY || EPI == EPI_AXPY_NORM) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cx<R> v = sizeof(R) == 8 ? cx<R>{(R)__builtin_fma(A.b, xi[c].x, 0.0), (R)__builtin_fma(A.b, xi[c].y, 0.0)}
                                         : cx<R>{__builtin_fmaf(A.b, xi[c].x, 0.0f), __builtin_fmaf(A.b, xi[c].y, 0.0f)};
                v = v + a * h[c];
                if constexpr (EPI == EPI_AXPY_NORM) nrm += (double)norm2(v);
                stc<true>(A.out + e[c], v);
            }
        } else {  // EPI_CGUPD: x (FP64) += alpha p, r -= alpha DDp
            const R alpha_r = (R)alpha;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const cx<R> ddp = sizeof(R) == 8
                                      ? cx<R>{(R)__builtin_fma(A.b, xi[c].x, 0.0), (R)__builtin_fma(A.b, xi[c].y, 0.0)}
                                      : cx<R>{__builtin_fmaf(A.b, xi[c].x, 0.0f), __builtin_fmaf(A.b, xi[c].y, 0.0f)};
                const cx<double> pd = {(double)pv[c].x, (double)pv[c].y};
                stc<true>(A.x + ex[c], xv[c] + alpha * pd);
                const cx<R> dd = ddp + a * h[c];
                const cx<R> r = rv[c] - alpha_r * dd;
                nrm += (double)norm2(r);
                stc<true>(A.r + e[c], r);
            }
        }
    }
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

int hop_blocks(const Geom &g) { return g.dir ? (g.vh + 63) / 64 : (g.vh + g.blk - 1) / g.blk; }

int dir_auto(int vh) { return vh <= kDirAutoVh ? 1 : 0; }

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
    A.fold_parts = L.fold_parts;
    A.fold_n = L.fold_n;
    A.fold_it = L.fold_it;
    const int nb = hop_blocks(g);
    if (g.dir) {
        if (L.lf == 6)
            hipLaunchKernelGGL((k_dslash_dir<R, SIGN, EPI, 6>), dim3(nb), dim3(256), 0, s, A);
        else
            hipLaunchKernelGGL((k_dslash_dir<R, SIGN, EPI, 9>), dim3(nb), dim3(256), 0, s, A);
    } else if (L.lf == 6) {
        hipLaunchKernelGGL((k_dslash<R, SIGN, EPI, 6>), dim3(nb), dim3(g.blk), g.lds, s, A);
    } else {
        hipLaunchKernelGGL((k_dslash<R, SIGN, EPI, 9>), dim3(nb), dim3(g.blk), g.lds, s, A);
    }
    hipError_t e = hipGetLastError();
    // FIN_NONE: the partials are left for the next kernel (launch-folded CG)
    if (e != hipSuccess || (EPI != EPI_AXPY_NORM && EPI != EPI_CGUPD) || L.fin == FIN_NONE) return e;
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
