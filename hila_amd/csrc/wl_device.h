// Device-side building blocks of the gfx950 Wilson kernels: complex arithmetic, the spin
// projection / reconstruction of the Wilson hop, SU(3) x half-spinor products, checkerboard
// geometry and wave64 reductions.
//
// Spin conventions follow libraries/datatypes/wilson_vector.h:439-567 (HalfWilsonVector
// projection and expand, unnormalised, gamma map {g1,g2,g3,g0}); SU(3) algebra follows
// matrix.h:1118-1137 (adjoint) and :2599-2622 (matrix x vector).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wl {

constexpr int kBlock = 256;  // threads per workgroup = 4 wave64

// ------------------------------------------------------------------ complex numbers ----

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
    return {a.x + b.x, a.y + b.y};
}
template <typename R>
__device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
    return {a.x - b.x, a.y - b.y};
}
template <typename R>
__device__ __forceinline__ cx<R> operator*(R s, cx<R> a) {
    return {s * a.x, s * a.y};
}
// v *= b for a spinor.  Each product is rounded on its own (an fma with a zero addend, which
// the compiler does not fuse with the addition that follows), so with b == 1 the values, and
// every later rounding, are exactly those of the unscaled spinor.  (A branch on b == 1 would
// keep both versions live and spill the FP64 kernels.)
template <typename R>
__device__ __forceinline__ void scale_unless_one(cx<R> (&v)[4][3], R b) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if constexpr (sizeof(R) == 8)
                v[s][c] = {__builtin_fma(b, v[s][c].x, 0.0), __builtin_fma(b, v[s][c].y, 0.0)};
            else
                v[s][c] = {__builtin_fmaf(b, v[s][c].x, 0.0f), __builtin_fmaf(b, v[s][c].y, 0.0f)};
        }
}
// a * b
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <typename R>
__device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {
    return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
// a * b + c, fused
__device__ __forceinline__ float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }
// acc + a * b.  In FP32 as four fused multiply-adds: written as acc + (a.x b.x - a.y b.y) the
// compiler forms the bracket first and adds it, two more instructions per term, which matters
// in the FP32/FP16 kernels, where the VALU is busy.  In FP64 (memory-bound) the bracket form is
// kept: the chained form's longer dependences raise the FP64 kernels' register use.
template <typename R>
__device__ __forceinline__ cx<R> cfma(cx<R> acc, cx<R> a, cx<R> b) {
    if constexpr (sizeof(R) == 4) {
        acc.x = fmar(a.x, b.x, acc.x);
        acc.x = fmar(-a.y, b.y, acc.x);
        acc.y = fmar(a.x, b.y, acc.y);
        acc.y = fmar(a.y, b.x, acc.y);
    } else {
        acc.x += a.x * b.x - a.y * b.y;
        acc.y += a.x * b.y + a.y * b.x;
    }
    return acc;
}
// acc + conj(a) * b, likewise
template <typename R>
__device__ __forceinline__ cx<R> cfmac(cx<R> acc, cx<R> a, cx<R> b) {
    if constexpr (sizeof(R) == 4) {
        acc.x = fmar(a.x, b.x, acc.x);
        acc.x = fmar(a.y, b.y, acc.x);
        acc.y = fmar(a.x, b.y, acc.y);
        acc.y = fmar(-a.y, b.x, acc.y);
    } else {
        acc.x += a.x * b.x + a.y * b.y;
        acc.y += a.x * b.y - a.y * b.x;
    }
    return acc;
}
// acc + f z (FP32: fused; FP64: product then sum, as cfma)
template <typename R>
__device__ __forceinline__ cx<R> caxpy(cx<R> acc, R f, cx<R> z) {
    if constexpr (sizeof(R) == 4)
        return {fmar(f, z.x, acc.x), fmar(f, z.y, acc.y)};
    else
        return acc + f * z;
}
// s * i * a  with s = +-1 folded at compile time
template <int S, typename R>
__device__ __forceinline__ cx<R> si(cx<R> a) {
    return S > 0 ? cx<R>{-a.y, a.x} : cx<R>{a.y, -a.x};
}
template <int S, typename R>
__device__ __forceinline__ cx<R> sr(cx<R> a) {
    return S > 0 ? a : cx<R>{-a.x, -a.y};
}
template <typename R>
__device__ __forceinline__ R norm2(cx<R> a) {
    return a.x * a.x + a.y * a.y;
}

// Loads / stores of one complex number.  NT = nontemporal: streamed data that this kernel does
// not read again (links, the output), so that it does not displace the neighbour spinors that
// the caches are meant to keep.
template <bool NT, typename R>
__device__ __forceinline__ cx<R> ldc(const cx<R> *p) {
    if constexpr (NT) {
        typedef R v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_nontemporal_load((const v2 *)p);
        return {v.x, v.y};
    } else {
        return *p;
    }
}
template <bool NT, typename R>
__device__ __forceinline__ void stc(cx<R> *p, cx<R> a) {
    if constexpr (NT) {
        typedef R v2 __attribute__((ext_vector_type(2)));
        const v2 v = {a.x, a.y};
        __builtin_nontemporal_store(v, (v2 *)p);
    } else {
        *p = a;
    }
}

template <typename R>
using Spinor = cx<R>[4][3];
template <typename R>
using Half = cx<R>[2][3];
template <typename R>
using Link = cx<R>[3][3];

// ------------------------------------------------------- spin projection (1 + S g_mu) ----
// HalfWilsonVector(w, dir, S): the two upper spin rows of (1 + S g_dir) w.
template <int MU, int S, typename R>
__device__ __forceinline__ void project(const cx<R> (&w)[4][3], cx<R> (&h)[2][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if constexpr (MU == 0) {  // g1
            h[0][c] = w[0][c] + si<S>(w[3][c]);
            h[1][c] = w[1][c] + si<S>(w[2][c]);
        } else if constexpr (MU == 1) {  // g2
            h[0][c] = w[0][c] - sr<S>(w[3][c]);
            h[1][c] = w[1][c] + sr<S>(w[2][c]);
        } else if constexpr (MU == 2) {  // g3
            h[0][c] = w[0][c] + si<S>(w[2][c]);
            h[1][c] = w[1][c] - si<S>(w[3][c]);
        } else {  // g0
            h[0][c] = w[0][c] + sr<S>(w[2][c]);
            h[1][c] = w[1][c] + sr<S>(w[3][c]);
        }
    }
}

// acc += expand(h, MU, S): rebuild the lower spin rows of (1 + S g_mu) from the upper two.
template <int MU, int S, typename R>
__device__ __forceinline__ void expand_add(cx<R> (&acc)[4][3], const cx<R> (&h)[2][3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc[0][c] = acc[0][c] + h[0][c];
        acc[1][c] = acc[1][c] + h[1][c];
        if constexpr (MU == 0) {
            acc[2][c] = acc[2][c] - si<S>(h[1][c]);
            acc[3][c] = acc[3][c] - si<S>(h[0][c]);
        } else if constexpr (MU == 1) {
            acc[2][c] = acc[2][c] + sr<S>(h[1][c]);
            acc[3][c] = acc[3][c] - sr<S>(h[0][c]);
        } else if constexpr (MU == 2) {
            acc[2][c] = acc[2][c] - si<S>(h[0][c]);
            acc[3][c] = acc[3][c] + si<S>(h[1][c]);
        } else {
            acc[2][c] = acc[2][c] + sr<S>(h[0][c]);
            acc[3][c] = acc[3][c] + sr<S>(h[1][c]);
        }
    }
}

// acc += f expand(h, MU, S): expand_add of the half spinor scaled by f (a site scale or the
// antiperiodic sign), with the scaling folded into the accumulation as fused multiply-adds.
template <int MU, int S, typename R>
__device__ __forceinline__ void expand_add_scaled(cx<R> (&acc)[4][3], const cx<R> (&h)[2][3], R f) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const cx<R> h0 = h[0][c], h1 = h[1][c];
        acc[0][c] = caxpy(acc[0][c], f, h0);
        acc[1][c] = caxpy(acc[1][c], f, h1);
        if constexpr (MU == 0) {
            acc[2][c] = caxpy(acc[2][c], -f, si<S>(h1));
            acc[3][c] = caxpy(acc[3][c], -f, si<S>(h0));
        } else if constexpr (MU == 1) {
            acc[2][c] = caxpy(acc[2][c], f, sr<S>(h1));
            acc[3][c] = caxpy(acc[3][c], -f, sr<S>(h0));
        } else if constexpr (MU == 2) {
            acc[2][c] = caxpy(acc[2][c], -f, si<S>(h0));
            acc[3][c] = caxpy(acc[3][c], f, si<S>(h1));
        } else {
            acc[2][c] = caxpy(acc[2][c], f, sr<S>(h0));
            acc[3][c] = caxpy(acc[3][c], f, sr<S>(h1));
        }
    }
}

// v = U h   (each of the two colour vectors)
template <typename R>
__device__ __forceinline__ void su3_mul(const cx<R> (&u)[3][3], const cx<R> (&h)[2][3],
                                        cx<R> (&v)[2][3]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            cx<R> a = cmul(u[i][0], h[s][0]);
            a = cfma(a, u[i][1], h[s][1]);
            v[s][i] = cfma(a, u[i][2], h[s][2]);
        }
}

// v = U^dag h
template <typename R>
__device__ __forceinline__ void su3_adj_mul(const cx<R> (&u)[3][3], const cx<R> (&h)[2][3],
                                            cx<R> (&v)[2][3]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            cx<R> a = cmulc(u[0][i], h[s][0]);
            a = cfmac(a, u[1][i], h[s][1]);
            v[s][i] = cfmac(a, u[2][i], h[s][2]);
        }
}

// --------------------------------------------------------------------- geometry -------
// Checkerboard site k of parity p (0 even, 1 odd) holds the site with lexicographic index
// 2k or 2k+1, whichever has parity p (HILA's EVEN_SITES_FIRST index, lattice.cpp:171-186).
struct Geom {
    int L[4];        // Lx, Ly, Lz, Lt (all even)
    int vh;          // sites per parity
    int stride;      // block stride S in sites: an rhs block holds 12 x S elements (>= vh, padded)
    int antiperiodic;
    // Traversal of the hopping kernel: a wave owns one 64-site segment of the checkerboard
    // (64 consecutive k: one site tile, coalesced accesses).  qc = 0 walks segments in memory order;
    // qc > 0 walks them chunk by chunk, each chunk = qc consecutive segments of a time slice,
    // taken for every t before the next chunk, so that the t neighbours of a site are read
    // while their spinors are still in the XCD's L2.  sps = segments per time slice,
    // ns = segments per parity (both valid only when qc > 0).
    int qc, sps, ns;
    int blk;         // workgroup size of the hopping kernel (64, 128 or 256)
    int dir;         // 1: single-rhs hops use the small-lattice kernel (one wave per direction,
                     // 64 sites per workgroup of 256 lanes), 0: one lane per site
    int lds;         // LDS bytes reserved per hopping workgroup (0: none); caps the workgroups
                     // sharing a CU, i.e. the sites in flight and their cache footprint
    unsigned long long dm[3];  // division multipliers for Lx, Ly, Lz (div_multiplier)
};

// Multiplier m of d for udiv: m = floor(2^64 / d) + 1, or 2^64 / d when d is a power of two.
inline constexpr unsigned long long div_multiplier(unsigned d) { return ~0ULL / d + 1; }

// floor(n / d) for d >= 2 given by m = div_multiplier(d), n < 2^31: floor(n m / 2^64), exact
// because n m / 2^64 = n / d + n e / 2^64 with 0 <= e < 1 and n e / 2^64 < 2^-32 < 1 / d, so
// the fraction of n / d (at most 1 - 1 / d) never reaches 1.  The 96-bit product is formed as
// n mh + floor(n ml / 2^32) (one multiply-high and one 64-bit multiply-add, below 2^63 + 2^32),
// whose high word is the floor: it replaces the ~15-instruction runtime division.
__device__ __forceinline__ unsigned udiv(unsigned n, unsigned long long m) {
    const unsigned lo = __umulhi(n, (unsigned)m);
    return (unsigned)(((unsigned long long)n * (unsigned)(m >> 32) + lo) >> 32);
}


struct SiteNb {
    int fwd[4], bwd[4];   // checkerboard index of x+mu / x-mu (opposite parity)
    bool fwd_edge_t, bwd_edge_t;
};

// MUL: the coordinates by multiply-high division (udiv), else by the compiler's division.  The
// issue-bound multi-rhs kernels take the shorter form (3-4% faster for FP16 at 4 rhs); the
// single-rhs kernels, bound by HBM and by their loads' issue order, keep the division, which
// measured 0.2% faster for them.
template <bool MUL = false>
__device__ __forceinline__ SiteNb neighbours(const Geom &g, int k, int p) {
    const unsigned Lx = g.L[0], Ly = g.L[1], Lz = g.L[2], Lt = g.L[3];
    const unsigned q = 2u * (unsigned)k;
    unsigned x0, y, z, t;
    if constexpr (MUL) {
        const unsigned r = udiv(q, g.dm[0]);
        x0 = q - r * Lx;
        const unsigned r2 = udiv(r, g.dm[1]);
        y = r - r2 * Ly;
        t = udiv(r2, g.dm[2]);
        z = r2 - t * Lz;
    } else {
        x0 = q % Lx;
        unsigned r = q / Lx;
        y = r % Ly;
        r /= Ly;
        z = r % Lz;
        t = r / Lz;
    }
    const unsigned odd = ((unsigned)p + y + z + t) & 1u;
    const unsigned x = x0 + odd;
    const unsigned idx = q + odd;
    const unsigned sy = Lx, sz = Lx * Ly, st = Lx * Ly * Lz;
    SiteNb n;
    n.fwd[0] = (x == Lx - 1 ? idx - (Lx - 1) : idx + 1) >> 1;
    n.bwd[0] = (x == 0 ? idx + (Lx - 1) : idx - 1) >> 1;
    n.fwd[1] = (y == Ly - 1 ? idx - sy * (Ly - 1) : idx + sy) >> 1;
    n.bwd[1] = (y == 0 ? idx + sy * (Ly - 1) : idx - sy) >> 1;
    n.fwd[2] = (z == Lz - 1 ? idx - sz * (Lz - 1) : idx + sz) >> 1;
    n.bwd[2] = (z == 0 ? idx + sz * (Lz - 1) : idx - sz) >> 1;
    n.fwd[3] = (t == Lt - 1 ? idx - st * (Lt - 1) : idx + st) >> 1;
    n.bwd[3] = (t == 0 ? idx + st * (Lt - 1) : idx - st) >> 1;
    n.fwd_edge_t = (t == Lt - 1);
    n.bwd_edge_t = (t == 0);
    return n;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b -> XCD b % 8).  Map them so
// that each XCD walks one contiguous eighth of the checkerboard: neighbouring sites (and the
// spinors they share) then stay in one XCD's L2.  XCD x takes nb / 8 tiles, one more when
// x < nb % 8; a bijection of [0, nb) for every nb.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Checkerboard site index handled by this lane as lane of workgroup b of a hopping-kernel
// launch of nb workgroups (>= vh: idle lane).
__device__ __forceinline__ int lane_site_of(const Geom &g, int b, int nb) {
    const int tile = xcd_tile(b, nb);
    if (g.qc <= 0) return tile * blockDim.x + threadIdx.x;
    const int o = tile * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (o >= g.ns) return g.vh;
    const int per = g.qc * g.L[3];
    const int c = o / per, rem = o - c * per;
    const int t = rem / g.qc, q = rem - t * g.qc;
    return ((t * g.sps + c * g.qc + q) << 6) + (int)(threadIdx.x & 63);
}

// ... of this launch's own workgroup
__device__ __forceinline__ int lane_site(const Geom &g) { return lane_site_of(g, blockIdx.x, gridDim.x); }

// ------------------------------------------------------------------- reductions ---------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the workgroup in a fixed order; the result is valid in thread 0.  Contains
// barriers: every thread of the workgroup must call it.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[kBlock / 64];
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) part[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
        const int nw = (int)(blockDim.x >> 6);
        for (int i = 0; i < nw; ++i) s += part[i];
    }
    __syncthreads();
    return s;
}

// Device-resident scalars of one CG solve (conjugate_gradient.h:55-57), updated by the last
// workgroup of each reduction kernel so that no scalar leaves the device inside the loop.
// `done` stops every gated kernel of the loop while non-zero:
//   1 converged, 2 maxiters reached, 3 reliable update requested (mixed precision),
//   4 iterated residual below the target, to be confirmed by a reliable update.
struct CgState {
    double source_norm, target_rr, rr, pDp, rrnew, alpha, beta, acc2;
    double true_rr, scratch;
    double rrb[2];      // rr by iteration parity (launch-folded single-rhs CG): slot i & 1 holds
                        // the rr iteration i starts from, so that the pass that finishes
                        // iteration i can record rr_(i+1) while its other workgroups still read
                        // rr_i
    double rr_update;   // ||r||^2 at the last reliable update
    double delta2;      // update when ||r||^2 < delta2 * rr_update
    int iter, done, maxiters, nreliable;
    int iter_since;     // iterations since the last reliable update
    int update_every;   // ... or after this many iterations
};

// What the last workgroup of a reduction does with the reduced sum.
enum Fin : int {
    FIN_NONE = 0,
    FIN_STORE = 1,     // *out = s
    FIN_SOURCE = 2,    // source_norm = s; target_rr = acc2 * s
    FIN_RR = 3,        // rr = s
    FIN_ALPHA = 4,     // pDp = s; alpha = rr / pDp
    FIN_RRNEW = 5,     // rrnew = s; convergence test and beta (conjugate_gradient.h:87-95)
    FIN_TRUE = 6,      // true_rr = s
    FIN_RRNEW_MIXED = 7,  // rrnew = s; convergence / reliable-update request / beta
    FIN_UPDATE = 8,    // s = ||b - M^dag M x||^2 after a reliable update: converged or restart
    FIN_RRNEW_MIXED2 = 9,  // batched: rrnew = s; convergence / update request (beta: sync kernel)
    FIN_UPDATE_FIRST = 10, // batched start: FIN_UPDATE applied to every rhs
};

struct RedArgs {
    double *partials;   // one slot per workgroup of the reducing kernel
    CgState *st;
    double *out;        // FIN_STORE target
};

__device__ __forceinline__ void fin_apply(int fin, double s, const RedArgs &ra) {
    CgState *st = ra.st;
    switch (fin) {
    case FIN_STORE: *ra.out = s; break;
    case FIN_SOURCE:
        st->source_norm = s;
        st->target_rr = st->acc2 * s;
        if (s == 0.0) st->done = 1;  // b = 0: x = 0 is the solution
        break;
    case FIN_RR:
        st->rr = s;
        st->rrb[0] = s;
        break;
    case FIN_ALPHA:
        st->pDp = s;
        st->alpha = st->rr / s;
        break;
    case FIN_RRNEW:
        st->rrnew = s;
        st->iter += 1;
        if (s < st->target_rr) {
            st->done = 1;
        } else {
            st->beta = s / st->rr;
            st->rr = s;
            if (st->iter >= st->maxiters) st->done = 2;
        }
        break;
    case FIN_TRUE: st->true_rr = s; break;
    case FIN_RRNEW_MIXED:
        st->rrnew = s;
        st->iter += 1;
        st->iter_since += 1;
        if (s < st->target_rr) {
            st->done = 4;
        } else if (s < st->delta2 * st->rr_update || st->iter_since >= st->update_every) {
            st->done = 3;
        } else {
            st->beta = s / st->rr;
            st->rr = s;
            if (st->iter >= st->maxiters) st->done = 2;
        }
        break;
    case FIN_RRNEW_MIXED2:
        st->rrnew = s;
        st->iter += 1;
        st->iter_since += 1;
        if (s < st->target_rr)
            st->done = 4;
        else if (s < st->delta2 * st->rr_update || st->iter_since >= st->update_every)
            st->done = 3;
        break;
    case FIN_UPDATE:
        st->true_rr = s;
        st->nreliable += 1;
        if (s < st->target_rr) {
            st->rrnew = s;
            st->done = 1;
        } else {
            // restart the recurrences from the true residual: beta = r2_true / r2_previous
            st->beta = s / st->rr;
            st->rr = s;
            st->rrnew = s;
            st->rr_update = s;
            st->iter_since = 0;
            st->done = st->iter >= st->maxiters ? 2 : 0;
        }
        break;
    default: break;
    }
}

// The sum that block_sum gives over 256 lanes holding f(0..255), computed by a workgroup of
// any multiple of 64 lanes (valid in thread 0): each 64-lane group is reduced by the same
// wave_sum, and the 4 group sums are added in order.  Contains barriers.
template <typename F>
__device__ double vthreads256_sum(F f) {
    __shared__ double part[kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int vw = w; vw < kBlock / 64; vw += nw) {
        const double v = wave_sum(f(vw * 64 + lane));
        if (lane == 0) part[vw] = v;
    }
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / 64; ++i) s += part[i];
    __syncthreads();
    return s;
}

// The sum of n partials exactly as the one-workgroup finish kernel k_finish adds them (256
// lanes, lane v adding partials v, v + 256, ...; then block_sum), returned to every thread of
// a workgroup of any multiple of 64 lanes.  Contains barriers.
__device__ __forceinline__ double finish_order_sum(const double *partials, int n) {
    __shared__ double out;
    const double s = vthreads256_sum([&](int v) {
        double a = 0;
        for (int i = v; i < n; i += kBlock) a += partials[i];
        return a;
    });
    if (threadIdx.x == 0) out = s;
    __syncthreads();
    return out;
}

// Each workgroup of a reducing kernel stores its partial sum in partials[blockIdx.x]; a
// one-workgroup finish kernel then adds them in workgroup order and applies the CG scalar
// update (deterministic; no atomics).  A last-workgroup-finishes scheme would need a
// device-scope release fence in every workgroup, which on this GPU writes back the XCD's L2
// and stalls a memory-bound kernel.  Every thread must call it.
__device__ __forceinline__ void reduce_partial(double v, const RedArgs &ra) {
    const double bs = block_sum(v);
    if (threadIdx.x == 0) ra.partials[blockIdx.x] = bs;
}

}  // namespace wl
