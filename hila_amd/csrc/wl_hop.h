// The pipelined Wilson hopping sum shared by the single- and multi-rhs kernels.
//
// acc = sum_mu [ U_mu(x)(1 + s g_mu) in(x+mu) + U_mu(x-mu)^dag (1 - s g_mu) in(x-mu) ] for one
// output site: each neighbour spinor is gathered from its site tile, spin-projected in
// registers to a half spinor, multiplied by the link (or its adjoint) and reconstructed into
// the accumulator (wilson.h:15-95 in one pass; HalfWilsonVector projection/expand of
// wilson_vector.h:439-567).  The 8 hops are software-pipelined: hop h+1's link and spinor are
// requested before hop h is computed, so a wave keeps 15-25 16-byte loads in flight.  Links
// are streamed with nontemporal loads; neighbour spinors go through the caches, where their
// 8-fold reuse is caught.
#pragma once

#include "wl_device.h"
#include "wl_store.h"

namespace wl {

// Registers holding one hop's operands: neighbour spinor and link.
template <typename R>
struct HopBuf {
    cx<R> w[4][3];
    cx<R> u[3][3];
};

// The neighbour spinor at site kn of the field `in` and the link entries at site kl of the link
// block `link`, in the packed layouts of wl_store.h (FP64 and FP32 storage).
template <bool NTL = true, int LF = 9, typename R>
__device__ __forceinline__ void hop_load(HopBuf<R> &b, const cx<R> *__restrict__ in,
                                         const cx<R> *__restrict__ link, int Ssp, int kn,
                                         int Slk, int kl) {
    cx<R> w[12], u[9];
    ld_spinor_raw<false, R>(w, in, Ssp, kn);
    ld_link_raw<NTL, R, LF>(u, link, Slk, kl);
#pragma unroll
    for (int c = 0; c < 12; ++c) b.w[c / 3][c % 3] = w[c];
#pragma unroll
    for (int e = 0; e < 9; ++e) b.u[e / 3][e % 3] = u[e];
}

// acc += expand(U (1 + s g_mu) w) for a forward hop, expand(U^dag (1 - s g_mu) w) backward;
// sg = -1 for a t hop across the antiperiodic edge.
template <int MU, int SIGN, bool FWD, typename R>
__device__ __forceinline__ void hop_compute(cx<R> (&acc)[4][3], const HopBuf<R> &b, R sg) {
    cx<R> h[2][3], v[2][3];
    if constexpr (FWD) {
        project<MU, SIGN>(b.w, h);
        su3_mul(b.u, h, v);
    } else {
        project<MU, -SIGN>(b.w, h);
        su3_adj_mul(b.u, h, v);
    }
    if constexpr (MU == 3) {
        if constexpr (FWD)
            expand_add_scaled<MU, SIGN>(acc, v, sg);
        else
            expand_add_scaled<MU, -SIGN>(acc, v, sg);
    } else {
        if constexpr (FWD)
            expand_add<MU, SIGN>(acc, v);
        else
            expand_add<MU, -SIGN>(acc, v);
    }
}

// Hop H = 2 mu + (0 forward, 1 backward): issue hop H+1's loads, then compute hop H.
// nsp: neighbour element offsets of the spinor field (site index times the rhs count);
// nl: neighbour site indices (backward links live at the neighbour); S: link block stride;
// Ssp: spinor block stride; LF: stored entries per link (9, or 6 for two-row links).
template <typename R, int SIGN, int H, bool NTL = true, int LF = 9>
__device__ __forceinline__ void hop_chain(cx<R> (&acc)[4][3], HopBuf<R> (&buf)[2],
                                          const cx<R> *__restrict__ in, const cx<R> *gout,
                                          const cx<R> *gin, int Ssp, int S, int k,
                                          const SiteNb &nsp, const SiteNb &nl, R sf, R sb) {
    if constexpr (H < 8) {
        constexpr int HN = H + 1;
        if constexpr (HN < 8) {
            constexpr int MN = HN / 2;
            if constexpr (HN % 2 == 0)
                hop_load<NTL, LF>(buf[HN & 1], in, gout + MN * LF * S, Ssp, nsp.fwd[MN], S, k);
            else
                hop_load<NTL, LF>(buf[HN & 1], in, gin + MN * LF * S, Ssp, nsp.bwd[MN], S, nl.bwd[MN]);
        }
        hop_compute<H / 2, SIGN, (H % 2 == 0)>(acc, buf[H & 1], (H % 2 == 0) ? sf : sb);
        hop_chain<R, SIGN, H + 1, NTL, LF>(acc, buf, in, gout, gin, Ssp, S, k, nsp, nl, sf, sb);
    }
}

// acc = H_SIGN in at output site k (parity p): the whole pipelined 8-hop sum.  NTL: stream the
// links past the caches (single rhs, where each link is read once per application).  LF:
// stored entries per link, as hop_chain.
template <typename R, int SIGN, bool NTL = true, int LF = 9>
__device__ __forceinline__ void hop_sum(cx<R> (&acc)[4][3], const Geom &g, int k, int p,
                                        const cx<R> *__restrict__ in, const cx<R> *gauge, int N) {
    const int S = g.stride;
    const SiteNb n = neighbours(g, k, p);
    SiteNb nsp = n;
    if (N != 1) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            nsp.fwd[m] = n.fwd[m] * N;
            nsp.bwd[m] = n.bwd[m] * N;
        }
    }
    const cx<R> *gout = gauge + (size_t)p * 4 * LF * S;
    const cx<R> *gin = gauge + (size_t)(1 - p) * 4 * LF * S;
    const R sf = (g.antiperiodic && n.fwd_edge_t) ? R(-1) : R(1);
    const R sb = (g.antiperiodic && n.bwd_edge_t) ? R(-1) : R(1);
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[s][c] = {R(0), R(0)};
    HopBuf<R> buf[2];
    hop_load<NTL, LF>(buf[0], in, gout, S * N, nsp.fwd[0], S, k);
    hop_chain<R, SIGN, 0, NTL, LF>(acc, buf, in, gout, gin, S * N, S, k, nsp, n, sf, sb);
}


}  // namespace wl
