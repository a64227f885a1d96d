// Layout conversion, BLAS-1 and reduction kernels of the Wilson CG (gfx950).
//
// The reductions replace HILA's per-workgroup shared-memory tree followed by host-side
// summation of the partials (hilapp codegen_gpu.cpp:1033-1088, backend_gpu/
// gpu_templated_ops.h:130-162): the partials are summed on the device by a one-workgroup
// finish kernel, in workgroup order, and the result goes straight into the device-resident CG
// state.

#include "wl_blas.h"

namespace wl {

// ------------------------------------------------------------------- layout ------------

// HILA AoS (V/2 sites x 24 doubles) -> right-hand side j of an N-rhs field in storage T:
// component c of site k of rhs j at element j * planes * S + comp_elem(c, S, k).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_aos_to_soa(const double *__restrict__ aos,
                                                       typename St<T>::E *__restrict__ out, int vh,
                                                       int S, int N, int j) {
    using R = typename St<T>::R;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    const double *src = aos + (size_t)24 * k;
    cx<R> v[4][3];
#pragma unroll
    for (int c = 0; c < 12; ++c) v[c / 3][c % 3] = {(R)src[2 * c], (R)src[2 * c + 1]};
    st_site<T>(out + (size_t)j * St<T>::planes * S, S, k, v);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_soa_to_aos(const typename St<T>::E *__restrict__ in,
                                                       double *__restrict__ aos, int vh, int S,
                                                       int N, int j) {
    using R = typename St<T>::R;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    double *dst = aos + (size_t)24 * k;
    cx<R> v[4][3];
    ld_site<T>(in + (size_t)j * St<T>::planes * S, S, k, v);
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        dst[2 * c] = (double)v[c / 3][c % 3].x;
        dst[2 * c + 1] = (double)v[c / 3][c % 3].y;
    }
}

// One direction of HILA's Field<SU<3,double>> (V sites x 18 doubles, even sites first) ->
// gauge blocks of both parities (site-tiled, wl_store.h).
__global__ __launch_bounds__(kBlock) void k_gauge_to_soa(const double *__restrict__ aos,
                                                         cx<double> *__restrict__ gauge, int mu,
                                                         int vh, int S) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= 2 * vh) return;
    const int p = s >= vh;
    const int k = s - p * vh;
    const double *src = aos + (size_t)18 * s;
    cx<double> *dst = gauge + (size_t)((p * 4 + mu) * 9) * S;
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[link_entry<double>(e, S, k)] = {src[2 * e], src[2 * e + 1]};
}

// Entry e and site k of element `rem` of an FP64 (unpacked) block of ne entries per site.
__device__ __forceinline__ void tile_entry_site(size_t rem, int ne, int &e, int &k) {
    const size_t tile = rem / ((size_t)ne * kTile);
    const int r = (int)(rem - tile * ne * kTile);
    e = r / kTile;
    k = (int)tile * kTile + (r - e * kTile);
}

// FP64 links -> FP32, in the packed FP32 link order of wl_store.h: element i of the FP64 links
// is entry e of site k of block i / 9S (parity, direction).
__global__ __launch_bounds__(kBlock) void k_convert_f32(const cx<double> *__restrict__ in,
                                                        cx<float> *__restrict__ out, size_t n, int S) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const size_t blk = i / ((size_t)9 * S);
    int e, k;
    tile_entry_site(i - blk * 9 * S, 9, e, k);
    const cx<double> v = in[i];
    out[blk * 9 * S + link_entry<float>(e, S, k)] = {(float)v.x, (float)v.y};
}

// FP64 links -> FP16 (SU(3) entries are bounded by 1: no scale), in the packed FP16 link
// order of wl_store.h: element i of the FP64 links is entry e of site k of block i / 9S
// (parity, direction).
__global__ __launch_bounds__(kBlock) void k_convert_h2(const cx<double> *__restrict__ in,
                                                       h2 *__restrict__ out, size_t n, int S) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const size_t blk = i / ((size_t)9 * S);
    int e, k;
    tile_entry_site(i - blk * 9 * S, 9, e, k);
    const cx<double> v = in[i];
    h2 q;
    q.x = (_Float16)(float)v.x;
    q.y = (_Float16)(float)v.y;
    out[blk * 9 * S + link_entry<f16s>(e, S, k)] = q;
}

// Two-row link copy for the hopping kernels (rows 0 and 1 of every link, in the 6S-element
// blocks of ld_link_raw's LF = 6 layout) from the FP64 links; out may be null (check only).
// With dev non-null each workgroup also writes the largest |re|, |im| of
// U[2][j] - conj(row 0 x row 1)_j over its links to dev[workgroup]: how far the stored third
// row is from the one the two-row kernels rebuild.  Thread i: link block i / vh (parity,
// direction), site i % vh.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_two_rows(const cx<double> *__restrict__ in, cx<R> *__restrict__ out,
                                                     int vh, int S, double *dev) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    double d = 0;
    if (i < (size_t)8 * vh) {
        const int blk = (int)(i / vh), k = (int)(i - (size_t)blk * vh);
        const cx<double> *src = in + (size_t)blk * 9 * S;
        cx<double> u[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) u[e] = src[link_entry<double>(e, S, k)];
        if (out) {
#pragma unroll
            for (int e = 0; e < 6; ++e)
                out[(size_t)blk * 6 * S + link_entry<R, 6>(e, S, k)] = {(R)u[e].x, (R)u[e].y};
        }
        if (dev) {
            cx<double> w[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) w[e] = u[e];
            third_row(w);
            // fmax would drop a NaN operand: every difference is tested on its own, and a NaN
            // (from a NaN or infinite entry in any row) makes the deviation infinite
#pragma unroll
            for (int e = 6; e < 9; ++e)
#pragma unroll
                for (int part = 0; part < 2; ++part) {
                    const double x = part ? fabs(w[e].y - u[e].y) : fabs(w[e].x - u[e].x);
                    if (!(x <= d)) d = (x != x) ? INFINITY : x;
                }
        }
    }
    if (dev) {  // grid-uniform: every thread reaches the barriers
        __shared__ double part[kBlock];
        part[threadIdx.x] = d;
        __syncthreads();
        for (int o = kBlock / 2; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + o]);
            __syncthreads();
        }
        if (threadIdx.x == 0) dev[blockIdx.x] = part[0];
    }
}

// ------------------------------------------------------------------ reductions ---------

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm2(const typename St<T>::E *__restrict__ f, int vh,
                                                  int S, int N, int j, RedArgs ra, const int *done) {
    if (done && *done) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    double s = 0;
    if (k < vh) s = site_norm2<T>(f + (size_t)j * St<T>::planes * S, S, k);
    reduce_partial(s, ra);
}

__global__ __launch_bounds__(kBlock) void k_cg_init(const cx<double> *__restrict__ b,
                                                    const cx<double> *__restrict__ ddp,
                                                    cx<double> *__restrict__ r,
                                                    cx<double> *__restrict__ p, int vh, int S,
                                                    RedArgs ra) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    double s = 0;
    if (k < vh) s = site_cg_init(b, ddp, r, p, S, k);
    reduce_partial(s, ra);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_cg_xpay(cx<R> *__restrict__ p,
                                                    const cx<R> *__restrict__ r,
                                                    const CgState *st, int vh, int S) {
    if (st->done) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    site_xpay<R>(p, r, (R)st->beta, S, k);
}

// p = beta p + r of the launch-folded CG iteration `it` (0-based), with the finish kernel's
// FIN_RRNEW folded in (conjugate_gradient.h:87-95): every workgroup adds the n partials of
// ||r||^2 in the finish kernel's order, takes the same convergence decision and beta =
// rrnew / rr_it, and updates p unless the solve has ended; workgroup 0 records the state
// (rrnew, iter, done; beta and rr, with rr_(it+1) in slot (it + 1) & 1, unless converged).
// The done flag is read once per workgroup: workgroup 0 may set it meanwhile, and p is not
// needed once it is set.
__global__ __launch_bounds__(kBlock) void k_cg_xpay_fold(cx<double> *__restrict__ p,
                                                         const cx<double> *__restrict__ r, CgState *st,
                                                         const double *parts, int n, int it, int vh, int S) {
    __shared__ int ended;
    if (threadIdx.x == 0) ended = st->done;
    __syncthreads();
    if (ended) return;  // workgroup-uniform
    const double s = finish_order_sum(parts, n);
    const double rr = st->rrb[it & 1];
    const int iter = it + 1;
    const bool conv = s < st->target_rr;
    const double beta = s / rr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->rrnew = s;
        st->iter = iter;
        if (conv) {
            st->done = 1;
        } else {
            st->beta = beta;
            st->rr = s;
            st->rrb[(it + 1) & 1] = s;
            if (iter >= st->maxiters) st->done = 2;
        }
    }
    if (conv || iter >= st->maxiters) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < vh) site_xpay<double>(p, r, beta, S, k);
}

// Reliable update: r = b - DDp in FP64 (DDp = M^dag M x in FP64), stored in the inner
// precision R, ||r||^2 reduced in FP64.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_residual_update(const cx<double> *__restrict__ b,
                                                            const cx<double> *__restrict__ ddp,
                                                            cx<R> *__restrict__ r, int vh, int S,
                                                            RedArgs ra) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    double s = 0;
    if (k < vh) {
        cx<R> v[4][3];
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const size_t e = comp_elem<double>(c, S, k);
            const cx<double> d = b[e] - ddp[e];
            v[c / 3][c % 3] = {(R)d.x, (R)d.y};
            s += norm2(d);
        }
        st_site<R>(r, S, k, v);
    }
    reduce_partial(s, ra);
}

__global__ __launch_bounds__(kBlock) void k_diff_norm2(const cx<double> *__restrict__ b,
                                                       const cx<double> *__restrict__ v, int vh,
                                                       int S, RedArgs ra, int fin) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    double s = 0;
    if (k < vh) s = site_diff_norm2(b, v, S, k);
    reduce_partial(s, ra);
}

// whole-field copy of n complex numbers
template <typename R>
__global__ __launch_bounds__(kBlock) void k_copy(const cx<R> *__restrict__ src,
                                                 cx<R> *__restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_copy_words(const unsigned *__restrict__ src,
                                                       unsigned *__restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Right-hand side blockIdx.y of dst = the same of src in another storage format: HILA's
// assignment between fields of different element precision (c[ALL] = chi[X] with c a float
// field, libraries/hmc/fermion_field.h:77-78).  Each component is rounded once to the
// destination's arithmetic type; FP16 storage then takes its site scale as st_site does.
template <typename Ti, typename To>
__global__ __launch_bounds__(kBlock) void k_convert(const typename St<Ti>::E *__restrict__ in,
                                                    typename St<To>::E *__restrict__ out, int vh, int S) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    const size_t j = blockIdx.y;
    cx<typename St<Ti>::R> v[4][3];
    ld_site<Ti>(in + j * St<Ti>::planes * S, S, k, v);
    using Ro = typename St<To>::R;
    cx<Ro> w[4][3];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int c = 0; c < 3; ++c) w[s][c] = {(Ro)v[s][c].x, (Ro)v[s][c].y};
    st_site<To>(out + j * St<To>::planes * S, S, k, w);
}

// Sum the n partials in workgroup order and apply `fin` (skipped once the CG has finished).
__global__ __launch_bounds__(kBlock) void k_finish(RedArgs ra, int n, int fin, const int *done) {
    if (done && *done) return;
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += ra.partials[i];
    s = block_sum(s);
    if (threadIdx.x == 0) fin_apply(fin, s, ra);
}

// ------------------------------------------------------------------- launchers ---------

hipError_t launch_cg_xpay_fold(const Geom &g, void *p, const void *r, CgState *st, const double *parts, int n,
                               int it, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_xpay_fold, dim3(nblocks(g)), dim3(kBlock), 0, s, (cx<double> *)p, (const cx<double> *)r,
                       st, parts, n, it, g.vh, g.stride);
    return hipGetLastError();
}

hipError_t launch_finish(const RedArgs &ra, int nparts, int fin, const int *done, hipStream_t s) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, s, ra, nparts, fin, done);
    return hipGetLastError();
}

static inline dim3 grid_sites(const Geom &g) { return dim3(nblocks(g)); }

hipError_t launch_aos_to_soa(int prec, const Geom &g, const double *aos, void *planes, int N,
                             int j, hipStream_t s) {
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_aos_to_soa<double>, grid_sites(g), dim3(kBlock), 0, s, aos,
                           (cx<double> *)planes, g.vh, g.stride, N, j);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_aos_to_soa<float>, grid_sites(g), dim3(kBlock), 0, s, aos,
                           (cx<float> *)planes, g.vh, g.stride, N, j);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_aos_to_soa<f16s>, grid_sites(g), dim3(kBlock), 0, s, aos, (h2 *)planes,
                           g.vh, g.stride, N, j);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(int prec, const Geom &g, const void *planes, double *aos, int N,
                             int j, hipStream_t s) {
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_soa_to_aos<double>, grid_sites(g), dim3(kBlock), 0, s,
                           (const cx<double> *)planes, aos, g.vh, g.stride, N, j);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_soa_to_aos<float>, grid_sites(g), dim3(kBlock), 0, s,
                           (const cx<float> *)planes, aos, g.vh, g.stride, N, j);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_soa_to_aos<f16s>, grid_sites(g), dim3(kBlock), 0, s, (const h2 *)planes,
                           aos, g.vh, g.stride, N, j);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_gauge_to_soa(const Geom &g, int mu, const double *aos_full, void *gauge64,
                               hipStream_t s) {
    const int n = 2 * g.vh;
    hipLaunchKernelGGL(k_gauge_to_soa, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       aos_full, (cx<double> *)gauge64, mu, g.vh, g.stride);
    return hipGetLastError();
}

hipError_t launch_gauge_convert(const Geom &g, int prec, const void *gauge64, void *gauge_lo,
                                hipStream_t s) {
    const size_t n = gauge_elems(g);
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    if (prec == WL_F32)
        hipLaunchKernelGGL(k_convert_f32, grid, dim3(kBlock), 0, s, (const cx<double> *)gauge64,
                           (cx<float> *)gauge_lo, n, g.stride);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_convert_h2, grid, dim3(kBlock), 0, s, (const cx<double> *)gauge64,
                           (h2 *)gauge_lo, n, g.stride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

int two_rows_blocks(const Geom &g) { return (int)(((size_t)8 * g.vh + kBlock - 1) / kBlock); }

hipError_t launch_two_rows(const Geom &g, int prec, const void *gauge64, void *rows, double *dev, hipStream_t s) {
    const dim3 grid((unsigned)two_rows_blocks(g));
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_two_rows<double>, grid, dim3(kBlock), 0, s, (const cx<double> *)gauge64,
                           (cx<double> *)rows, g.vh, g.stride, dev);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_two_rows<float>, grid, dim3(kBlock), 0, s, (const cx<double> *)gauge64,
                           (cx<float> *)rows, g.vh, g.stride, dev);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_norm2(int prec, const Geom &g, const void *f, int N, int j, const RedArgs &ra,
                        int fin, const int *done, hipStream_t s) {
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_norm2<double>, grid_sites(g), dim3(kBlock), 0, s,
                           (const cx<double> *)f, g.vh, g.stride, N, j, ra, done);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_norm2<float>, grid_sites(g), dim3(kBlock), 0, s,
                           (const cx<float> *)f, g.vh, g.stride, N, j, ra, done);
    else if (prec == WL_F16)
        hipLaunchKernelGGL(k_norm2<f16s>, grid_sites(g), dim3(kBlock), 0, s, (const h2 *)f, g.vh,
                           g.stride, N, j, ra, done);
    else
        return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finish(ra, nblocks(g), fin, done, s);
}

hipError_t launch_cg_init(const Geom &g, const void *b, const void *DDp, void *r, void *p,
                          const RedArgs &ra, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_init, grid_sites(g), dim3(kBlock), 0, s, (const cx<double> *)b,
                       (const cx<double> *)DDp, (cx<double> *)r, (cx<double> *)p, g.vh,
                       g.stride, ra);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finish(ra, nblocks(g), FIN_RR, nullptr, s);
}

hipError_t launch_cg_xpay(int prec, const Geom &g, void *p, const void *r, const CgState *st,
                          hipStream_t s) {
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_cg_xpay<double>, grid_sites(g), dim3(kBlock), 0, s, (cx<double> *)p,
                           (const cx<double> *)r, st, g.vh, g.stride);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_cg_xpay<float>, grid_sites(g), dim3(kBlock), 0, s, (cx<float> *)p,
                           (const cx<float> *)r, st, g.vh, g.stride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_residual_update(int prec, const Geom &g, const void *b, const void *DDp,
                                  void *r_lo, const RedArgs &ra, hipStream_t s) {
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_residual_update<double>, grid_sites(g), dim3(kBlock), 0, s,
                           (const cx<double> *)b, (const cx<double> *)DDp, (cx<double> *)r_lo,
                           g.vh, g.stride, ra);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_residual_update<float>, grid_sites(g), dim3(kBlock), 0, s,
                           (const cx<double> *)b, (const cx<double> *)DDp, (cx<float> *)r_lo,
                           g.vh, g.stride, ra);
    else
        return hipErrorInvalidValue;
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finish(ra, nblocks(g), FIN_UPDATE, nullptr, s);
}

hipError_t launch_diff_norm2(const Geom &g, const void *b, const void *v, const RedArgs &ra,
                             int fin, hipStream_t s) {
    hipLaunchKernelGGL(k_diff_norm2, grid_sites(g), dim3(kBlock), 0, s, (const cx<double> *)b,
                       (const cx<double> *)v, g.vh, g.stride, ra, fin);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_finish(ra, nblocks(g), fin, nullptr, s);
}

hipError_t launch_copy(int prec, const Geom &g, const void *src, void *dst, int nrhs,
                       hipStream_t s) {
    const size_t n = (size_t)nrhs * planes_of(prec) * g.stride;
    const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
    if (prec == WL_F64)
        hipLaunchKernelGGL(k_copy<double>, grid, dim3(kBlock), 0, s, (const cx<double> *)src,
                           (cx<double> *)dst, n);
    else if (prec == WL_F32)
        hipLaunchKernelGGL(k_copy<float>, grid, dim3(kBlock), 0, s, (const cx<float> *)src,
                           (cx<float> *)dst, n);
    else if (prec == WL_F16)  // 4-byte elements: the FP16 pairs and the FP32 scale plane
        hipLaunchKernelGGL(k_copy_words, grid, dim3(kBlock), 0, s, (const unsigned *)src, (unsigned *)dst, n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename Ti>
static hipError_t convert_from(int pout, const Geom &g, const void *src, void *dst, int nrhs, hipStream_t s) {
    const dim3 grid((unsigned)nblocks(g), (unsigned)nrhs);
    const auto *in = (const typename St<Ti>::E *)src;
    if (pout == WL_F64)
        hipLaunchKernelGGL((k_convert<Ti, double>), grid, dim3(kBlock), 0, s, in, (cx<double> *)dst, g.vh, g.stride);
    else if (pout == WL_F32)
        hipLaunchKernelGGL((k_convert<Ti, float>), grid, dim3(kBlock), 0, s, in, (cx<float> *)dst, g.vh, g.stride);
    else if (pout == WL_F16)
        hipLaunchKernelGGL((k_convert<Ti, f16s>), grid, dim3(kBlock), 0, s, in, (h2 *)dst, g.vh, g.stride);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_convert(int pin, int pout, const Geom &g, const void *src, void *dst, int nrhs, hipStream_t s) {
    if (nrhs < 1 || nrhs > 65535) return hipErrorInvalidValue;
    if (pin == WL_F64) return convert_from<double>(pout, g, src, dst, nrhs, s);
    if (pin == WL_F32) return convert_from<float>(pout, g, src, dst, nrhs, s);
    if (pin == WL_F16) return convert_from<f16s>(pout, g, src, dst, nrhs, s);
    return hipErrorInvalidValue;
}

}  // namespace wl
