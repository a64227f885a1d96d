// Multi-right-hand-side Wilson kernels for gfx950 and the batched-CG support kernels.
//
// A field of N right-hand sides stores element (rhs j, component c, site k) at
// (j * 12 + c) * stride + k: every right-hand side is a set of 12 component planes.  The
// hopping kernel is register-blocked over right-hand sides: a lane owns one output site and a
// chunk of NR right-hand sides; per hop it loads the link once and the NR neighbour spinors,
// all independent loads (NR * 12 + 9 in flight per lane), and accumulates NR spinors.  A wave
// covers one 64-site segment of the checkerboard (coalesced plane accesses, same traversal as
// the single-rhs kernel) and the waves of one segment -- its right-hand-side chunks -- are
// scheduled next to each other, so the link re-reads of the second and later chunks hit the
// caches.  Per site and right-hand side the link traffic falls by the number of right-hand
// sides sharing a link, and the kernel, latency-bound per wave like the single-rhs one, keeps
// NR times more bytes in flight per wave.
//
// Batched CG keeps one CgState per right-hand side: each runs HILA's recurrence
// (conjugate_gradient.h:75-95) with its own alpha, beta and convergence flag, and all share
// every operator application.  Per-rhs partial sums are reduced in a fixed order.

#include "wl_hop.h"
#include "wl_internal.h"
#include "wl_store.h"

namespace wl {

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
    Geom g;
    int out_par;
    int j0, jend, C;           // right-hand sides [j0, jend) in chunks of NR; C chunks
    const int *all_done;       // skip the launch when set (batched CG loop); may be null
    CgState *st;               // per-rhs CG states (skip finished rhs, alpha); may be null
    cx<double> *__restrict__ x;
    const E *__restrict__ p;
    E *__restrict__ r;
    double *partials;
    int pstride;               // partial sums: partials[rhs * pstride + segment]
};

// One hop for the NR right-hand sides of this lane: link once, NR spinors.  NTL: load the
// link with nontemporal hints (it is not read again by this launch), so that the link stream
// does not displace the neighbour spinors from the caches.  For the scaled format the site
// scale of each neighbour is applied to the product U h (6 complex numbers) instead of to the
// 12 components of the spinor: the hop is linear.
template <int MU, int SIGN, bool FWD, int NR, bool NTL, typename T>
__device__ __forceinline__ void rb_hop(cx<typename St<T>::R> (&acc)[NR][4][3], const bool (&act)[NR],
                                       const typename St<T>::E *__restrict__ in, size_t rs,
                                       const typename St<T>::E *__restrict__ link, int S, int kn,
                                       int kl, typename St<T>::R sg) {
    using R = typename St<T>::R;
    cx<R> u[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) u[i][j] = lde<NTL>(link + (i * 3 + j) * S + kl);
    cx<R> w[NR][4][3];
    R sc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (act[r]) {
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int c = 0; c < 3; ++c) w[r][s][c] = lde<false>(in + r * rs + (s * 3 + c) * S + kn);
            if constexpr (St<T>::scaled)
                sc[r] = sg * scale_plane(in + r * rs, S)[kn];
            else
                sc[r] = sg;
        }
#pragma unroll
    for (int r = 0; r < NR; ++r)
        if (act[r]) {
            cx<R> h[2][3], v[2][3];
            if constexpr (FWD) {
                project<MU, SIGN>(w[r], h);
                su3_mul(u, h, v);
            } else {
                project<MU, -SIGN>(w[r], h);
                su3_adj_mul(u, h, v);
            }
            if constexpr (MU == 3 || St<T>::scaled) {
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int c = 0; c < 3; ++c) v[q][c] = sc[r] * v[q][c];
            }
            if constexpr (FWD)
                expand_add<MU, SIGN>(acc[r], v);
            else
                expand_add<MU, -SIGN>(acc[r], v);
        }
}

template <typename T, int SIGN, int EPI, int NR, bool NTL>
__global__ __launch_bounds__(kBlock, 2) void k_dslash_rb(RbArgs<T> A) {
    using R = typename St<T>::R;
    using E = typename St<T>::E;
    if (A.all_done && *A.all_done) return;
    const int W = xcd_tile(blockIdx.x, gridDim.x) * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
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
        const SiteNb n = neighbours(A.g, k, A.out_par);
        const E *gout = A.gauge + (size_t)A.out_par * 36 * S;
        const E *gin = A.gauge + (size_t)(1 - A.out_par) * 36 * S;
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
        rb_hop<0, SIGN, true, NR, NTL, T>(acc, act, in, rs, gout, S, n.fwd[0], k, R(1));
        rb_hop<0, SIGN, false, NR, NTL, T>(acc, act, in, rs, gin, S, n.bwd[0], n.bwd[0], R(1));
        rb_hop<1, SIGN, true, NR, NTL, T>(acc, act, in, rs, gout + 9 * S, S, n.fwd[1], k, R(1));
        rb_hop<1, SIGN, false, NR, NTL, T>(acc, act, in, rs, gin + 9 * S, S, n.bwd[1], n.bwd[1], R(1));
        rb_hop<2, SIGN, true, NR, NTL, T>(acc, act, in, rs, gout + 18 * S, S, n.fwd[2], k, R(1));
        rb_hop<2, SIGN, false, NR, NTL, T>(acc, act, in, rs, gin + 18 * S, S, n.bwd[2], n.bwd[2], R(1));
        rb_hop<3, SIGN, true, NR, NTL, T>(acc, act, in, rs, gout + 27 * S, S, n.fwd[3], k, sf);
        rb_hop<3, SIGN, false, NR, NTL, T>(acc, act, in, rs, gin + 27 * S, S, n.bwd[3], n.bwd[3], sb);

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
            } else {  // EPI_CGUPD: x (FP64) += alpha p, r -= alpha DDp
                const double alpha = A.st[jf + r].alpha;
                const R alpha_r = (R)alpha;
                cx<R> ddp[4][3], pv[4][3], rv[4][3];
                ld_site<T, true>(A.xin + eb, S, k, ddp);
                ld_site<T, true>(A.p + eb, S, k, pv);
                ld_site<T, true>(A.r + eb, S, k, rv);
                const size_t xb = (size_t)(jf + r) * 12 * S + k;
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const size_t e = xb + (size_t)(s * 3 + c) * S;
                        const cx<double> pd = {(double)pv[s][c].x, (double)pv[s][c].y};
                        stc<true>(A.x + e, ldc<true>(A.x + e) + alpha * pd);
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
    if constexpr (EPI == EPI_AXPY_NORM || EPI == EPI_CGUPD) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double s = wave_sum(nrm[r]);
            if ((threadIdx.x & 63) == 0 && jf + r < A.jend && o < A.pstride)
                A.partials[(size_t)(jf + r) * A.pstride + o] = s;
        }
    }
}

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
        const size_t off = (size_t)j * 12 * S + k;
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const size_t e = off + c * S;
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
        const size_t off = (size_t)j * 12 * S + k;
        cx<R> v[4][3];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const size_t e = off + (size_t)(a * 3 + c) * S;
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

// x_j = 0 for one rhs j
__global__ __launch_bounds__(kBlock) void k_zero_rhs(cx<double> *__restrict__ x, int j, int vh,
                                                     int S) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    const size_t off = (size_t)j * 12 * S + k;
#pragma unroll
    for (int c = 0; c < 12; ++c) x[off + c * S] = {0.0, 0.0};
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

template <typename T, int SIGN, int EPI, int NR, bool NTL>
static hipError_t rb_launch(const Geom &g, const MrLaunch &L, hipStream_t s) {
    using E = typename St<T>::E;
    using R = typename St<T>::R;
    RbArgs<T> A;
    A.out = (E *)L.out;
    A.in = (const E *)L.in;
    A.xin = (const E *)L.xin;
    A.gauge = (const E *)L.gauge;
    A.a = (R)L.a;
    A.g = g;
    A.out_par = L.out_par;
    A.j0 = L.j0;
    A.jend = L.j0 + L.NB;
    A.C = (L.NB + NR - 1) / NR;
    A.all_done = L.all_done;
    A.st = L.st;
    A.x = (cx<double> *)L.x;
    A.p = (const E *)L.p;
    A.r = (E *)L.r;
    A.partials = L.partials;
    A.pstride = mr_segments(g);
    const long waves = (long)A.pstride * A.C;
    const int wpb = g.blk / 64;
    hipLaunchKernelGGL((k_dslash_rb<T, SIGN, EPI, NR, NTL>), dim3((unsigned)((waves + wpb - 1) / wpb)),
                       dim3(g.blk), g.lds, s, A);
    return hipGetLastError();
}

template <typename T, int SIGN, int EPI, int NR>
static hipError_t rb_ntl(const Geom &g, const MrLaunch &L, hipStream_t s) {
    return L.link_nt ? rb_launch<T, SIGN, EPI, NR, true>(g, L, s) : rb_launch<T, SIGN, EPI, NR, false>(g, L, s);
}

template <typename T, int SIGN, int NR>
static hipError_t rb_epi(const Geom &g, const MrLaunch &L, hipStream_t s) {
    switch (L.epi) {
    case EPI_SET: return rb_ntl<T, SIGN, EPI_SET, NR>(g, L, s);
    case EPI_AXPY: return rb_ntl<T, SIGN, EPI_AXPY, NR>(g, L, s);
    case EPI_AXPY_NORM: return rb_ntl<T, SIGN, EPI_AXPY_NORM, NR>(g, L, s);
    case EPI_CGUPD: return rb_ntl<T, SIGN, EPI_CGUPD, NR>(g, L, s);
    default: return hipErrorInvalidValue;
    }
}

// FP32 arithmetic takes up to 4 right-hand sides per lane; FP64 up to 2 (3 or 4 spill
// registers).
template <typename T, int SIGN>
static hipError_t rb_nr(const Geom &g, const MrLaunch &L, hipStream_t s) {
    switch (L.NR) {
    case 1: return rb_epi<T, SIGN, 1>(g, L, s);
    case 2: return rb_epi<T, SIGN, 2>(g, L, s);
    default:
        if constexpr (sizeof(typename St<T>::R) == 4) {
            if (L.NR == 3) return rb_epi<T, SIGN, 3>(g, L, s);
            if (L.NR == 4) return rb_epi<T, SIGN, 4>(g, L, s);
        }
        return hipErrorInvalidValue;
    }
}

hipError_t launch_dslash_mr(int prec, const Geom &g, const MrLaunch &L, hipStream_t s) {
    if (prec == WL_F64)
        return L.sign > 0 ? rb_nr<double, 1>(g, L, s) : rb_nr<double, -1>(g, L, s);
    if (prec == WL_F32)
        return L.sign > 0 ? rb_nr<float, 1>(g, L, s) : rb_nr<float, -1>(g, L, s);
    if (prec == WL_F16)
        return L.sign > 0 ? rb_nr<f16s, 1>(g, L, s) : rb_nr<f16s, -1>(g, L, s);
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

hipError_t launch_zero_rhs(const Geom &g, void *x, int j, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_rhs, dim3(nblocks(g)), dim3(kBlock), 0, s, (cx<double> *)x, j,
                       g.vh, g.stride);
    return hipGetLastError();
}

}  // namespace wl
