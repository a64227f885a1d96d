// The register-blocked multi-right-hand-side Wilson hopping kernel for gfx950 (templates):
// included by one translation unit per storage format (wl_rb_f64.hip, wl_rb_f32.hip,
// wl_rb_f16.hip), which instantiate its launchers.
//
// A field of N right-hand sides stores rhs j as a block of planes (wl_store.h).  A lane owns
// one output site and a chunk of NR right-hand sides; per hop it loads the link once and the
// NR neighbour spinors, all independent loads, and accumulates NR spinors.  A wave covers one
// 64-site segment of the checkerboard (coalesced plane accesses, same traversal as the
// single-rhs kernel) and the waves of one segment -- its right-hand-side chunks -- are
// scheduled next to each other, so the link re-reads of the second and later chunks hit the
// caches.
#pragma once

#include "wl_hop.h"
#include "wl_internal.h"
#include "wl_store.h"

namespace wl {

int mr_segments(const Geom &g);

// First site of segment o of the hopping-kernel traversal (>= vh: idle).
__device__ __forceinline__ int segment_base(const Geom &g, int o) {
    if (g.qc <= 0) return o << 6;
    if (o >= g.ns) return g.vh;
    const int per = g.qc * g.L[3];
    const int c = o / per, rem = o - c * per;
    const int t = rem / g.qc, q = rem - t * g.qc;
    return (t * g.sps + c * g.qc + q) << 6;
}

template <typename T>
struct RbArgs {
    using E = typename St<T>::E;
    using R = typename St<T>::R;
    E *__restrict__ out;
    const E *__restrict__ in;
    const E *__restrict__ xin;
    const E *__restrict__ gauge;
    R a;
    R b;                       // coefficient of xin (1 + the Hasenbusch shift)
    Geom g;
    int out_par;
    int j0, jend, C;           // right-hand sides [j0, jend) in chunks of NR; C chunks
    const int *all_done;       // skip the launch when set (batched CG loop); may be null
    CgState *st;               // per-rhs CG states (skip finished rhs, alpha); may be null
    cx<R> *__restrict__ x;     // solution accumulator: x itself (FP64) or the inner y (FP32)
    const E *__restrict__ p;
    E *__restrict__ r;
    double *partials;
    int pstride;               // partial sums: partials[rhs * pstride + segment]
};

// The projected half spinor of a stored neighbour spinor in FP32 or FP64 storage, widened
// first (FP16 storage is projected packed, hop_pk).
template <int MU, int S, typename T>
__device__ __forceinline__ void rb_project(const typename St<T>::E (&wr)[12], cx<typename St<T>::R> (&h)[2][3]) {
    cx<typename St<T>::R> w[4][3];
#pragma unroll
    for (int c = 0; c < 12; ++c) w[c / 3][c % 3] = widen(wr[c]);
    project<MU, S>(w, h);
}

// acc += f expand((1 + S gamma_mu) w, U or U^dag) for one rhs in FP16 storage: packed
// projection and U h (wl_store.h), FP32 accumulation; f = site scale x antiperiodic sign.
template <int MU, int SIGN, bool FWD>
__device__ __forceinline__ void hop_pk(cx<float> (&acc)[4][3], const h2 (&w)[12], const h2 (&u)[9], float f) {
    h2 hr[2][3], hi[2][3];
    cx<float> v[2][3];
    project_pk<MU, FWD ? SIGN : -SIGN, FWD>(w, hr, hi);
    su3_mul_pk<FWD>(u, hr, hi, v);
    expand_add_scaled<MU, FWD ? SIGN : -SIGN>(acc, v, f);
}

// One hop for the NR right-hand sides of this lane: link once, NR spinors.  NTL: load the
// link with nontemporal hints (it is not read again by this launch), so that the link stream
// does not displace the neighbour spinors from the caches.  For the scaled format the site
// scale of each neighbour is applied to the product U h (6 complex numbers) instead of to the
// 12 components of the spinor: the hop is linear.
template <int MU, int SIGN, bool FWD, int NR, bool NTL, typename T, int LF>
__device__ __forceinline__ void rb_hop(cx<typename St<T>::R> (&acc)[NR][4][3], const bool (&act)[NR],
                                       const typename St<T>::E *__restrict__ in, size_t rs,
                                       const typename St<T>::E *__restrict__ link, int S, int kn,
                                       int kl, typename St<T>::R sg) {
    using R = typename St<T>::R;
    typename St<T>::E ur[9];
    ld_link_raw<NTL, T, LF>(ur, link, S, kl);
    cx<R> u[3][3];
    if constexpr (!St<T>::scaled) {
#pragma unroll
        for (int e = 0; e < 9; ++e) u[e / 3][e % 3] = widen(ur[e]);
    }
    typename St<T>::E wr[NR][12];
    R sc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (act[r]) {
            ld_spinor_raw<false, T>(wr[r], in + r * rs, S, kn);
            if constexpr (St<T>::scaled)
                sc[r] = sg * scale_plane(in + r * rs, S)[kn];
            else
                sc[r] = sg;
        }
    if constexpr (St<T>::scaled) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (act[r]) hop_pk<MU, SIGN, FWD>(acc[r], wr[r], ur, sc[r]);
    } else {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (act[r]) {
                cx<R> h[2][3], v[2][3];
                rb_project<MU, FWD ? SIGN : -SIGN, T>(wr[r], h);
                if constexpr (FWD)
                    su3_mul(u, h, v);
                else
                    su3_adj_mul(u, h, v);
                if constexpr (MU == 3)
                    expand_add_scaled<MU, FWD ? SIGN : -SIGN>(acc[r], v, sc[r]);
                else
                    expand_add<MU, FWD ? SIGN : -SIGN>(acc[r], v);
            }
    }
}

// Software-pipelined form of the 8 hops (PIPE): hop h+1's link and NR neighbour spinors are
// requested, and held packed in the stored format, before hop h is computed, so that a wave
// keeps about two hops of loads in flight.  This lets few waves per CU saturate the memory
// system, and few waves in flight keep the neighbour spinors they share in the L2.  Used
// where the two register buffers fit without spilling (FP32 and FP16 arithmetic, NR <= 2).
template <typename T, int NR>
struct RbBuf {
    typename St<T>::E w[NR][12];
    typename St<T>::E u[9];
    typename St<T>::R sc[NR];
};

template <typename T, int NR>
constexpr bool rb_pipe_ok() {
    return sizeof(typename St<T>::R) == 4 && (NR <= 2 || St<T>::scaled);
}

template <int NR, bool NTL, typename T, int LF>
__device__ __forceinline__ void rb_load(RbBuf<T, NR> &b, const bool (&act)[NR],
                                        const typename St<T>::E *__restrict__ in, size_t rs,
                                        const typename St<T>::E *__restrict__ link, int S, int kn,
                                        int kl) {
    ld_link_raw<NTL, T, LF>(b.u, link, S, kl);
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (act[r]) {
            ld_spinor_raw<false, T>(b.w[r], in + r * rs, S, kn);
            if constexpr (St<T>::scaled) b.sc[r] = scale_plane(in + r * rs, S)[kn];
        }
}

// acc += the hop from the buffered spinors and the link u (FP32 or FP64 storage); sg:
// antiperiodic sign (t hops only)
template <int MU, int SIGN, bool FWD, int NR, typename T>
__device__ __forceinline__ void rb_compute_u(cx<typename St<T>::R> (&acc)[NR][4][3], const RbBuf<T, NR> &b,
                                             const cx<typename St<T>::R> (&u)[3][3], const bool (&act)[NR],
                                             typename St<T>::R sg) {
    using R = typename St<T>::R;
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (act[r]) {
            cx<R> h[2][3], v[2][3];
            rb_project<MU, FWD ? SIGN : -SIGN, T>(b.w[r], h);
            if constexpr (FWD)
                su3_mul(u, h, v);
            else
                su3_adj_mul(u, h, v);
            if constexpr (MU == 3)
                expand_add_scaled<MU, FWD ? SIGN : -SIGN>(acc[r], v, sg);
            else
                expand_add<MU, FWD ? SIGN : -SIGN>(acc[r], v);
        }
}

// acc += the hop from buffered operands (spinors and link)
template <int MU, int SIGN, bool FWD, int NR, typename T>
__device__ __forceinline__ void rb_compute(cx<typename St<T>::R> (&acc)[NR][4][3], const RbBuf<T, NR> &b,
                                           const bool (&act)[NR], typename St<T>::R sg) {
    using R = typename St<T>::R;
    if constexpr (St<T>::scaled) {
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (act[r]) hop_pk<MU, SIGN, FWD>(acc[r], b.w[r], b.u, (MU == 3 ? sg : R(1)) * b.sc[r]);
    } else {
        cx<R> u[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) u[i][j] = widen(b.u[i * 3 + j]);
        rb_compute_u<MU, SIGN, FWD, NR, T>(acc, b, u, act, sg);
    }
}

// Hop H = 2 mu + (0 forward, 1 backward): issue hop H+1's loads, then compute hop H.
template <typename T, int SIGN, int H, int NR, bool NTL, int LF>
__device__ __forceinline__ void rb_chain(cx<typename St<T>::R> (&acc)[NR][4][3], RbBuf<T, NR> (&buf)[2],
                                         const bool (&act)[NR], const typename St<T>::E *__restrict__ in,
                                         size_t rs, const typename St<T>::E *gout,
                                         const typename St<T>::E *gin, int S, int k, const SiteNb &n,
                                         typename St<T>::R sf, typename St<T>::R sb) {
    if constexpr (H < 8) {
        constexpr int HN = H + 1;
        if constexpr (HN < 8) {
            constexpr int MN = HN / 2;
            if constexpr (HN % 2 == 0)
                rb_load<NR, NTL, T, LF>(buf[HN & 1], act, in, rs, gout + MN * LF * S, S, n.fwd[MN], k);
            else
                rb_load<NR, NTL, T, LF>(buf[HN & 1], act, in, rs, gin + MN * LF * S, S, n.bwd[MN], n.bwd[MN]);
        }
        rb_compute<H / 2, SIGN, (H % 2 == 0), NR, T>(acc, buf[H & 1], act, (H % 2 == 0) ? sf : sb);
        rb_chain<T, SIGN, H + 1, NR, NTL, LF>(acc, buf, act, in, rs, gout, gin, S, k, n, sf, sb);
    }
}

// Epilogue for the NR rhs of this lane at output site k: see Epi in wl_internal.h.  Norms of
// the stored values are added to nrm.
template <typename T, int EPI, int NR>
__device__ __forceinline__ void rb_epilogue(const RbArgs<T> &A, cx<typename St<T>::R> (&acc)[NR][4][3],
                                            const bool (&act)[NR], int jf, int k, double (&nrm)[NR]) {
    using R = typename St<T>::R;
    const int S = A.g.stride;
    const size_t rs = (size_t)St<T>::planes * S;
    const R a = A.a;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (!act[r]) continue;
        const size_t eb = (size_t)(jf + r) * rs;
        if constexpr (EPI == EPI_SET) {
            cx<R> v[4][3];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) v[s][c] = a * acc[r][s][c];
            st_site<T, true>(A.out + eb, S, k, v);
        } else if constexpr (EPI == EPI_AXPY || EPI == EPI_AXPY_NORM) {
            cx<R> v[4][3];
            ld_site<T, true>(A.xin + eb, S, k, v);
            scale_unless_one(v, A.b);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) v[s][c] = v[s][c] + a * acc[r][s][c];
            st_site<T, true>(A.out + eb, S, k, v);
            if constexpr (EPI == EPI_AXPY_NORM) {
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int c = 0; c < 3; ++c) nrm[r] += (double)norm2(v[s][c]);
            }
        } else {  // EPI_CGUPD: x += alpha p (x: FP64 solution or FP32 inner y), r -= alpha DDp
            const double alpha = A.st[jf + r].alpha;
            const R alpha_r = (R)alpha;
            cx<R> ddp[4][3], pv[4][3], rv[4][3];
            ld_site<T, true>(A.xin + eb, S, k, ddp);
            scale_unless_one(ddp, A.b);
            ld_site<T, true>(A.p + eb, S, k, pv);
            ld_site<T, true>(A.r + eb, S, k, rv);
            const size_t xb = (size_t)(jf + r) * 12 * S;
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const size_t e = xb + comp_elem<double>(s * 3 + c, S, k);
                    stc<true>(A.x + e, ldc<true>(A.x + e) + alpha_r * pv[s][c]);
                    const cx<R> dd = ddp[s][c] + a * acc[r][s][c];
                    rv[s][c] = rv[s][c] - alpha_r * dd;
                }
            st_site<T, true>(A.r + eb, S, k, rv);
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) nrm[r] += (double)norm2(rv[s][c]);
        }
    }
}

// Per-wave partial sums of the norm epilogues: partials[rhs * pstride + segment o].
template <typename T, int EPI, int NR>
__device__ __forceinline__ void rb_partials(const RbArgs<T> &A, const double (&nrm)[NR], int jf, int o) {
    if constexpr (EPI == EPI_AXPY_NORM || EPI == EPI_CGUPD) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double s = wave_sum(nrm[r]);
            if ((threadIdx.x & 63) == 0 && jf + r < A.jend && o < A.pstride)
                A.partials[(size_t)(jf + r) * A.pstride + o] = s;
        }
    }
}

template <typename T, int SIGN, int EPI, int NR, bool NTL, bool PIPE, int LF>
__global__ __launch_bounds__(kBlock, 2) void k_dslash_rb(RbArgs<T> A) {
    using R = typename St<T>::R;
    using E = typename St<T>::E;
    if (A.all_done && *A.all_done) return;
    // the wave index, read from lane 0 so that the compiler knows it (and the segment, the rhs
    // chunk and the rhs blocks' addresses) to be wave-uniform: scalar registers
    const int W = xcd_tile(blockIdx.x, gridDim.x) * (int)(blockDim.x >> 6) +
                  __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int o = W / A.C, chunk = W - o * A.C;
    const int k = segment_base(A.g, o) + (int)(threadIdx.x & 63);
    const int jf = A.j0 + chunk * NR;
    const int S = A.g.stride;
    const size_t rs = (size_t)St<T>::planes * S;  // elements per right-hand side
    bool act[NR];  // wave-uniform
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int j = jf + r;
        act[r] = j < A.jend && (A.st == nullptr || A.st[j].done == 0);
    }
    double nrm[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) nrm[r] = 0;
    if (k < A.g.vh) {
        const SiteNb n = neighbours<true>(A.g, k, A.out_par);
        const E *gout = A.gauge + (size_t)A.out_par * 4 * LF * S;
        const E *gin = A.gauge + (size_t)(1 - A.out_par) * 4 * LF * S;
        const R sf = (A.g.antiperiodic && n.fwd_edge_t) ? R(-1) : R(1);
        const R sb = (A.g.antiperiodic && n.bwd_edge_t) ? R(-1) : R(1);
        const E *in = A.in + jf * rs;
        cx<R> acc[NR][4][3];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[r][s][c] = {R(0), R(0)};
        if constexpr (PIPE) {
            RbBuf<T, NR> buf[2];
            rb_load<NR, NTL, T, LF>(buf[0], act, in, rs, gout, S, n.fwd[0], k);
            rb_chain<T, SIGN, 0, NR, NTL, LF>(acc, buf, act, in, rs, gout, gin, S, k, n, sf, sb);
        } else {
            rb_hop<0, SIGN, true, NR, NTL, T, LF>(acc, act, in, rs, gout, S, n.fwd[0], k, R(1));
            rb_hop<0, SIGN, false, NR, NTL, T, LF>(acc, act, in, rs, gin, S, n.bwd[0], n.bwd[0], R(1));
            rb_hop<1, SIGN, true, NR, NTL, T, LF>(acc, act, in, rs, gout + LF * S, S, n.fwd[1], k, R(1));
            rb_hop<1, SIGN, false, NR, NTL, T, LF>(acc, act, in, rs, gin + LF * S, S, n.bwd[1], n.bwd[1], R(1));
            rb_hop<2, SIGN, true, NR, NTL, T, LF>(acc, act, in, rs, gout + 2 * LF * S, S, n.fwd[2], k, R(1));
            rb_hop<2, SIGN, false, NR, NTL, T, LF>(acc, act, in, rs, gin + 2 * LF * S, S, n.bwd[2], n.bwd[2], R(1));
            rb_hop<3, SIGN, true, NR, NTL, T, LF>(acc, act, in, rs, gout + 3 * LF * S, S, n.fwd[3], k, sf);
            rb_hop<3, SIGN, false, NR, NTL, T, LF>(acc, act, in, rs, gin + 3 * LF * S, S, n.bwd[3], n.bwd[3], sb);
        }

        rb_epilogue<T, EPI, NR>(A, acc, act, jf, k, nrm);
    }
    rb_partials<T, EPI, NR>(A, nrm, jf, o);
}

// ---------------------------------------------------------- LDS-staged links ---------------
// One workgroup per 64-site segment; its waves are the segment's rhs chunks (C <= 8).  The
// segment's links (8 hops x 9 entries x 64 sites) are read from HBM once, nontemporally, into
// LDS, shared by the C waves, so that link traffic is once per segment whatever the caches do;
// each wave then streams only its NR rhs' neighbour spinors (pipelined by one hop when PIPE).

template <int NR, typename T>
__device__ __forceinline__ void sp_load(RbBuf<T, NR> &b, const bool (&act)[NR],
                                        const typename St<T>::E *__restrict__ in, size_t rs, int S, int kn) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (act[r]) {
            ld_spinor_raw<false, T>(b.w[r], in + r * rs, S, kn);
            if constexpr (St<T>::scaled) b.sc[r] = scale_plane(in + r * rs, S)[kn];
        }
}

template <typename T, int SIGN, int H, int NR, bool PIPE>
__device__ __forceinline__ void rb_chain_lds(cx<typename St<T>::R> (&acc)[NR][4][3], RbBuf<T, NR> (&buf)[2],
                                             const bool (&act)[NR], const typename St<T>::E *__restrict__ in,
                                             size_t rs, const typename St<T>::E *lu, int lane, int S,
                                             const SiteNb &n, typename St<T>::R sf, typename St<T>::R sb) {
    using R = typename St<T>::R;
    if constexpr (H < 8) {
        if constexpr (PIPE) {
            constexpr int HN = H + 1;
            if constexpr (HN < 8)
                sp_load<NR, T>(buf[HN & 1], act, in, rs, S, (HN % 2 == 0) ? n.fwd[HN / 2] : n.bwd[HN / 2]);
        } else {
            sp_load<NR, T>(buf[H & 1], act, in, rs, S, (H % 2 == 0) ? n.fwd[H / 2] : n.bwd[H / 2]);
        }
        if constexpr (St<T>::scaled) {
            h2 ur[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) ur[e] = lu[(H * 9 + e) * 64 + lane];
            const R sg = (H / 2 == 3) ? ((H % 2 == 0) ? sf : sb) : R(1);
#pragma unroll
            for (int r = 0; r < NR; ++r)
                if (act[r]) hop_pk<H / 2, SIGN, (H % 2 == 0)>(acc[r], buf[H & 1].w[r], ur, sg * buf[H & 1].sc[r]);
        } else {
            cx<R> u[3][3];
#pragma unroll
            for (int e = 0; e < 9; ++e) u[e / 3][e % 3] = widen(lu[(H * 9 + e) * 64 + lane]);
            rb_compute_u<H / 2, SIGN, (H % 2 == 0), NR, T>(acc, buf[H & 1], u, act, (H % 2 == 0) ? sf : sb);
        }
        rb_chain_lds<T, SIGN, H + 1, NR, PIPE>(acc, buf, act, in, rs, lu, lane, S, n, sf, sb);
    }
}

template <typename T, int SIGN, int EPI, int NR, bool PIPE, int LF>
__global__ __launch_bounds__(512, 1) void k_dslash_rb_lds(RbArgs<T> A) {
    using R = typename St<T>::R;
    using E = typename St<T>::E;
    extern __shared__ unsigned char rb_smem[];
    E *lu = reinterpret_cast<E *>(rb_smem);  // [hop * 9 + entry][lane]
    if (A.all_done && *A.all_done) return;   // grid-uniform
    const int o = xcd_tile(blockIdx.x, gridDim.x);
    const int chunk = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63), nw = (int)(blockDim.x >> 6);
    const int k = segment_base(A.g, o) + lane;
    const int jf = A.j0 + chunk * NR;
    const int S = A.g.stride;
    const size_t rs = (size_t)St<T>::planes * S;
    bool act[NR];  // wave-uniform
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int j = jf + r;
        act[r] = j < A.jend && (A.st == nullptr || A.st[j].done == 0);
    }
    double nrm[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) nrm[r] = 0;
    const bool live = k < A.g.vh;
    SiteNb n{};
    if (live) n = neighbours<true>(A.g, k, A.out_par);
    const E *gout = A.gauge + (size_t)A.out_par * 4 * LF * S;
    const E *gin = A.gauge + (size_t)(1 - A.out_par) * 4 * LF * S;
    if (live) {
        if constexpr (LF == 9) {
            for (int idx = chunk; idx < 72; idx += nw) {
                const int h = idx / 9, e = idx - 9 * h, mu = h >> 1;
                const E *src = (h & 1) ? gin + (size_t)mu * 9 * S + link_entry<T>(e, S, n.bwd[mu])
                                       : gout + (size_t)mu * 9 * S + link_entry<T>(e, S, k);
                lu[idx * 64 + lane] = ldraw<true>(src);
            }
        } else {  // two-row links: a wave loads whole links and stages the rebuilt 9 entries
            for (int h = chunk; h < 8; h += nw) {
                const int mu = h >> 1;
                E u[9];
                ld_link_raw<true, T, LF>(u, ((h & 1) ? gin : gout) + (size_t)mu * LF * S, S,
                                         (h & 1) ? n.bwd[mu] : k);
#pragma unroll
                for (int e = 0; e < 9; ++e) lu[(h * 9 + e) * 64 + lane] = u[e];
            }
        }
    }
    __syncthreads();
    if (live) {
        const R sf = (A.g.antiperiodic && n.fwd_edge_t) ? R(-1) : R(1);
        const R sb = (A.g.antiperiodic && n.bwd_edge_t) ? R(-1) : R(1);
        const E *in = A.in + jf * rs;
        cx<R> acc[NR][4][3];
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) acc[r][s][c] = {R(0), R(0)};
        RbBuf<T, NR> buf[2];
        if constexpr (PIPE) sp_load<NR, T>(buf[0], act, in, rs, S, n.fwd[0]);
        rb_chain_lds<T, SIGN, 0, NR, PIPE>(acc, buf, act, in, rs, lu, lane, S, n, sf, sb);
        rb_epilogue<T, EPI, NR>(A, acc, act, jf, k, nrm);
    }
    rb_partials<T, EPI, NR>(A, nrm, jf, o);
}

template <typename T, int SIGN, int EPI, int NR, bool NTL, bool PIPE, int LF>
inline hipError_t rb_launch(const Geom &g, const MrLaunch &L, hipStream_t s) {
    using E = typename St<T>::E;
    using R = typename St<T>::R;
    RbArgs<T> A;
    A.out = (E *)L.out;
    A.in = (const E *)L.in;
    A.xin = (const E *)L.xin;
    A.gauge = (const E *)L.gauge;
    A.a = (R)L.a;
    A.b = (R)L.b;
    A.g = g;
    A.out_par = L.out_par;
    A.j0 = L.j0;
    A.jend = L.j0 + L.NB;
    A.C = (L.NB + NR - 1) / NR;
    A.all_done = L.all_done;
    A.st = L.st;
    A.x = (cx<R> *)L.x;
    A.p = (const E *)L.p;
    A.r = (E *)L.r;
    A.partials = L.partials;
    A.pstride = mr_segments(g);
    const long waves = (long)A.pstride * A.C;
    const int wpb = g.blk / 64;
    hipLaunchKernelGGL((k_dslash_rb<T, SIGN, EPI, NR, NTL, PIPE, LF>), dim3((unsigned)((waves + wpb - 1) / wpb)),
                       dim3(g.blk), g.lds, s, A);
    return hipGetLastError();
}

template <typename T, int SIGN, int EPI, int NR, bool NTL, int LF>
inline hipError_t rb_pipe(const Geom &g, const MrLaunch &L, hipStream_t s) {
    if constexpr (rb_pipe_ok<T, NR>())
        if (L.pipe) return rb_launch<T, SIGN, EPI, NR, NTL, true, LF>(g, L, s);
    return rb_launch<T, SIGN, EPI, NR, NTL, false, LF>(g, L, s);
}

// Spinor pipelining in the LDS-staged kernel (no link registers): FP32 arithmetic up to 4 per
// lane, FP64 up to 1.
template <typename T, int NR>
constexpr bool rb_lds_pipe_ok() {
    return sizeof(typename St<T>::R) == 4 ? NR <= 4 : NR <= 1;
}

template <typename T, int SIGN, int EPI, int NR, bool PIPE, int LF>
inline hipError_t rb_launch_lds(const Geom &g, const MrLaunch &L, hipStream_t s) {
    using E = typename St<T>::E;
    using R = typename St<T>::R;
    RbArgs<T> A;
    A.out = (E *)L.out;
    A.in = (const E *)L.in;
    A.xin = (const E *)L.xin;
    A.gauge = (const E *)L.gauge;
    A.a = (R)L.a;
    A.b = (R)L.b;
    A.g = g;
    A.out_par = L.out_par;
    A.j0 = L.j0;
    A.jend = L.j0 + L.NB;
    A.C = (L.NB + NR - 1) / NR;
    A.all_done = L.all_done;
    A.st = L.st;
    A.x = (cx<R> *)L.x;
    A.p = (const E *)L.p;
    A.r = (E *)L.r;
    A.partials = L.partials;
    A.pstride = mr_segments(g);
    const size_t lds = (size_t)72 * 64 * sizeof(E) + (size_t)g.lds;
    hipLaunchKernelGGL((k_dslash_rb_lds<T, SIGN, EPI, NR, PIPE, LF>), dim3((unsigned)A.pstride), dim3(64 * A.C),
                       lds, s, A);
    return hipGetLastError();
}

template <typename T, int SIGN, int EPI, int NR, int LF>
inline hipError_t rb_ntl(const Geom &g, const MrLaunch &L, hipStream_t s) {
    if (L.lds_links && (L.NB + NR - 1) / NR <= 8) {
        if constexpr (rb_lds_pipe_ok<T, NR>())
            if (L.pipe) return rb_launch_lds<T, SIGN, EPI, NR, true, LF>(g, L, s);
        return rb_launch_lds<T, SIGN, EPI, NR, false, LF>(g, L, s);
    }
    return L.link_nt ? rb_pipe<T, SIGN, EPI, NR, true, LF>(g, L, s) : rb_pipe<T, SIGN, EPI, NR, false, LF>(g, L, s);
}

template <typename T, int SIGN, int NR, int LF>
inline hipError_t rb_epi(const Geom &g, const MrLaunch &L, hipStream_t s) {
    switch (L.epi) {
    case EPI_SET: return rb_ntl<T, SIGN, EPI_SET, NR, LF>(g, L, s);
    case EPI_AXPY: return rb_ntl<T, SIGN, EPI_AXPY, NR, LF>(g, L, s);
    case EPI_AXPY_NORM: return rb_ntl<T, SIGN, EPI_AXPY_NORM, NR, LF>(g, L, s);
    case EPI_CGUPD: return rb_ntl<T, SIGN, EPI_CGUPD, NR, LF>(g, L, s);
    default: return hipErrorInvalidValue;
    }
}

// FP32 arithmetic takes up to 4 right-hand sides per lane; FP64 up to 2 (3 or 4 spill
// registers).  LF: stored entries per link (9, or 6 for two-row links).
template <typename T, int SIGN, int LF = 9>
inline hipError_t rb_nr(const Geom &g, const MrLaunch &L, hipStream_t s) {
    switch (L.NR) {
    case 1: return rb_epi<T, SIGN, 1, LF>(g, L, s);
    case 2: return rb_epi<T, SIGN, 2, LF>(g, L, s);
    default:
        if constexpr (sizeof(typename St<T>::R) == 4) {
            if (L.NR == 3) return rb_epi<T, SIGN, 3, LF>(g, L, s);
            if (L.NR == 4) return rb_epi<T, SIGN, 4, LF>(g, L, s);
        }
        return hipErrorInvalidValue;
    }
}

}  // namespace wl
