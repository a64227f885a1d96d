// Minimal-residual extrapolation of an initial guess from previous solutions: HILA's
// MRE_guess (libraries/hmc/MRE_guess.h:8-87) for CG on A = (M + mh)^dag (M + mh).
//
// HILA orthogonalises the n previous solutions with Gram-Schmidt under Re <a, b>, applies A to
// each orthogonalised vector, takes the n x n projected matrix and the projected source with
// one reduction pass per entry, and solves the small system by Gauss-Jordan elimination on the
// host.  Here every O(V) step is a streaming pass over whole fields:
//   1. A on all n basis vectors at once: the multi-rhs hopping kernel (M, then M^dag);
//   2. one pass reading the n basis vectors, the n products and chi, which sums every inner
//      product the small problem needs -- the Gram matrix G_ij = Re <b_i, b_j>, the projected
//      operator A_ij = Re <b_i, A b_j> (upper triangle; A is Hermitian) and v_i = Re <b_i, chi>;
//   3. the Gram-Schmidt and the elimination on those n x n numbers on the host, as HILA;
//   4. one pass psi = sum_i c_i b_i.
// Gram-Schmidt in the metric G gives the same orthogonalised vectors as on the fields (each is
// a combination of the b_i); they are normalised, so that the tests of the elimination compare
// numbers of the scale of A's eigenvalues.

#include <cmath>
#include <vector>

#include "wl_host.h"
#include "wl_store.h"

namespace wl {

namespace {

constexpr int kMaxBasis = 8;
constexpr int kMaxBlocks = 1024;
// A basis vector whose Gram-Schmidt remainder has squared norm at most this fraction of its
// own is taken as dependent on the vectors before it (HILA's test, MRE_guess.h:26, is the
// absolute norm^2 > 1e-32).
constexpr double kDependent = 1e-12;
// An elimination pivot with square at most this is dropped (MRE_guess.h:56, on the normalised
// basis).
constexpr double kPivot2 = 1e-32;

template <int NB>
struct MreCoef {
    double c[NB];
};

constexpr int n_coef(int nb) { return nb * (nb + 1) + nb; }

// Per-workgroup partial sums of G (upper triangle, row-major), A (upper triangle) and v:
// partials[i * gridDim.x + blockIdx.x] for coefficient i.
template <int NB>
__global__ __launch_bounds__(kBlock) void k_mre_dots(const cx<double> *__restrict__ B, const cx<double> *__restrict__ AB,
                                                     const cx<double> *__restrict__ chi, int vh, int S,
                                                     double *__restrict__ partials) {
    constexpr int NT = NB * (NB + 1) / 2;
    constexpr int NC = 2 * NT + NB;
    double acc[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) acc[i] = 0.0;
    const size_t rs = (size_t)12 * S;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < vh; k += gridDim.x * kBlock) {
#pragma unroll 1
        for (int c = 0; c < 12; ++c) {
            const size_t e = comp_elem<double>(c, S, k);
            cx<double> b[NB], ab[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                b[i] = ldc<true>(B + i * rs + e);
                ab[i] = ldc<true>(AB + i * rs + e);
            }
            const cx<double> h = ldc<true>(chi + e);
            int t = 0;
#pragma unroll
            for (int i = 0; i < NB; ++i)
#pragma unroll
                for (int j = i; j < NB; ++j, ++t) {
                    acc[t] += b[i].x * b[j].x + b[i].y * b[j].y;
                    acc[NT + t] += b[i].x * ab[j].x + b[i].y * ab[j].y;
                }
#pragma unroll
            for (int i = 0; i < NB; ++i) acc[2 * NT + i] += b[i].x * h.x + b[i].y * h.y;
        }
    }
    __shared__ double part[kBlock / 64][NC];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const double s = wave_sum(acc[i]);
        if ((threadIdx.x & 63) == 0) part[w][i] = s;
    }
    __syncthreads();
    if (threadIdx.x < NC) {
        double s = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) s += part[q][threadIdx.x];
        partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

// out[i] = sum over the nblk partials of coefficient i, in workgroup order (one workgroup per
// coefficient).
__global__ __launch_bounds__(kBlock) void k_mre_finish(const double *__restrict__ partials, int nblk,
                                                       double *__restrict__ out) {
    const double *p = partials + (size_t)blockIdx.x * nblk;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) s += p[b];
    s = block_sum(s);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// psi = sum_i c_i B_i on every site, each component summed over i in order.  A basis vector's
// 12 components of the site are loaded together (one contiguous stretch per wave in the
// site-tiled layout) into per-component sums.  With the sum over i innermost the compiler kept
// fewer than NB loads in flight per wave: 8 basis vectors at 32^4 took 203 us with 22 VGPRs,
// against 167 us with 66 VGPRs in this form (MI355X).
template <int NB>
__global__ __launch_bounds__(kBlock) void k_mre_combine(cx<double> *__restrict__ psi, const cx<double> *__restrict__ B,
                                                        MreCoef<NB> cf, int vh, int S) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
    const size_t rs = (size_t)12 * S;
    cx<double> s[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) s[c] = {0.0, 0.0};
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        cx<double> b[12];
#pragma unroll
        for (int c = 0; c < 12; ++c) b[c] = ldc<true>(B + i * rs + comp_elem<double>(c, S, k));
#pragma unroll
        for (int c = 0; c < 12; ++c) s[c] = s[c] + cf.c[i] * b[c];
    }
#pragma unroll
    for (int c = 0; c < 12; ++c) stc<true>(psi + comp_elem<double>(c, S, k), s[c]);
}

template <int NB>
hipError_t launch_dots(const Geom &g, const void *B, const void *AB, const void *chi, double *partials, int nblk,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_mre_dots<NB>, dim3(nblk), dim3(kBlock), 0, s, (const cx<double> *)B, (const cx<double> *)AB,
                       (const cx<double> *)chi, g.vh, g.stride, partials);
    return hipGetLastError();
}

template <int NB>
hipError_t launch_combine(const Geom &g, void *psi, const void *B, const std::vector<double> &c, hipStream_t s) {
    MreCoef<NB> cf;
    for (int i = 0; i < NB; ++i) cf.c[i] = c[i];
    hipLaunchKernelGGL(k_mre_combine<NB>, dim3((g.vh + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (cx<double> *)psi,
                       (const cx<double> *)B, cf, g.vh, g.stride);
    return hipGetLastError();
}

template <template <int> class F, typename... Args>
hipError_t dispatch_nb(int nb, Args &&...args) {
    switch (nb) {
        case 1: return F<1>::run(args...);
        case 2: return F<2>::run(args...);
        case 3: return F<3>::run(args...);
        case 4: return F<4>::run(args...);
        case 5: return F<5>::run(args...);
        case 6: return F<6>::run(args...);
        case 7: return F<7>::run(args...);
        case 8: return F<8>::run(args...);
        default: return hipErrorInvalidValue;
    }
}
template <int NB>
struct Dots {
    static hipError_t run(const Geom &g, const void *B, const void *AB, const void *chi, double *partials, int nblk,
                          hipStream_t s) {
        return launch_dots<NB>(g, B, AB, chi, partials, nblk, s);
    }
};
template <int NB>
struct Combine {
    static hipError_t run(const Geom &g, void *psi, const void *B, const std::vector<double> &c, hipStream_t s) {
        return launch_combine<NB>(g, psi, B, c, s);
    }
};

// Coefficients c of psi = sum_i c_i b_i from G, A (both symmetric, n x n, row-major) and v:
// MRE_guess.h:15-79 on the numbers, with the intended Gram-Schmidt coefficient (projection
// onto the vector projected out) and orthonormalised vectors.
std::vector<double> mre_coefficients(int n, const std::vector<double> &G, const std::vector<double> &A,
                                     const std::vector<double> &v) {
    // q_i = sum_k T[i][k] b_k, orthonormal under Re <., .>; kept[i] false for dropped vectors
    std::vector<double> T((size_t)n * n, 0.0);
    std::vector<char> kept(n, 0);
    auto gdot = [&](int i, int j) {  // <q_i, q_j> = T_i G T_j^T
        double s = 0;
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) s += T[(size_t)i * n + a] * G[(size_t)a * n + b] * T[(size_t)j * n + b];
        return s;
    };
    for (int i = 0; i < n; ++i) {
        T[(size_t)i * n + i] = 1.0;
        const double own = G[(size_t)i * n + i];
        if (!(own > 0)) continue;  // a zero vector
        for (int j = 0; j < i; ++j) {
            if (!kept[j]) continue;
            const double coef = gdot(j, i);  // <q_j, q_i> with |q_j| = 1
            for (int a = 0; a < n; ++a) T[(size_t)i * n + a] -= coef * T[(size_t)j * n + a];
        }
        const double rem = gdot(i, i);
        if (!(rem > kDependent * own)) continue;
        const double inv = 1.0 / std::sqrt(rem);
        for (int a = 0; a < n; ++a) T[(size_t)i * n + a] *= inv;
        kept[i] = 1;
    }
    // projected system M[j][i] = <q_j, A q_i>, rhs w[i] = <q_i, chi> (MRE_guess.h:35-51)
    std::vector<double> M((size_t)n * n, 0.0), w(n, 0.0);
    for (int i = 0; i < n; ++i) {
        if (!kept[i]) continue;
        for (int j = 0; j < n; ++j) {
            if (!kept[j]) continue;
            double s = 0;
            for (int a = 0; a < n; ++a)
                for (int b = 0; b < n; ++b) s += T[(size_t)j * n + a] * A[(size_t)a * n + b] * T[(size_t)i * n + b];
            M[(size_t)j * n + i] = s;
        }
        double s = 0;
        for (int a = 0; a < n; ++a) s += T[(size_t)i * n + a] * v[a];
        w[i] = s;
    }
    // Gauss-Jordan elimination (MRE_guess.h:53-79)
    for (int i = 0; i < n; ++i) {
        const double d = M[(size_t)i * n + i];
        if (d * d > kPivot2) {
            const double dinv = 1.0 / d;
            for (int j = 0; j < n; ++j) M[(size_t)i * n + j] *= dinv;
            w[i] *= dinv;
            for (int k = 0; k < n; ++k) {
                if (k == i) continue;
                const double wt = M[(size_t)k * n + i];
                for (int j = 0; j < n; ++j) M[(size_t)k * n + j] -= wt * M[(size_t)i * n + j];
                w[k] -= wt * w[i];
            }
        } else {
            w[i] = 0;
            for (int j = 0; j < n; ++j) M[(size_t)j * n + i] = 0;
        }
    }
    // psi = sum_i w_i q_i = sum_k (sum_i w_i T[i][k]) b_k  (MRE_guess.h:81-85, accumulated)
    std::vector<double> c(n, 0.0);
    for (int i = 0; i < n; ++i) {
        if (!kept[i] || std::isnan(w[i])) continue;
        for (int k = 0; k < n; ++k) c[k] += w[i] * T[(size_t)i * n + k];
    }
    return c;
}

}  // namespace

}  // namespace wl

using namespace wl;

extern "C" int wl_mre_guess(wl_ctx *c, wl_field *psi, const wl_field *chi, const wl_field *basis, int nbasis,
                            double kappa, double mh) {
    if (!c || !psi || !chi || !basis) return fail(WL_E_ARG, "null argument");
    if (psi->ctx != c || chi->ctx != c || basis->ctx != c) return fail(WL_E_ARG, "field belongs to another context");
    for (const wl_field *f : {(const wl_field *)psi, chi, basis})
        if (f->parity != WL_EVEN || f->prec != WL_F64) return fail(WL_E_ARG, "MRE acts on even FP64 fields");
    if (psi->nrhs != 1 || chi->nrhs != 1) return fail(WL_E_ARG, "psi and chi hold one right-hand side");
    if (psi == chi || psi == basis) return fail(WL_E_ARG, "psi must differ from chi and the basis");
    if (nbasis < 0 || nbasis > kMaxBasis || nbasis > basis->nrhs)
        return fail(WL_E_ARG, "nbasis %d: must be 0..%d and at most the basis field's %d rhs", nbasis, kMaxBasis,
                    basis->nrhs);
    if (!std::isfinite(mh) || !std::isfinite(kappa)) return fail(WL_E_ARG, "kappa and the shift must be finite");
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    const Geom &g = c->g;
    if (nbasis == 0) {  // MRE_size 0: fermion_action::initial_guess leaves psi = 0
        HIPCHK(hipMemsetAsync(psi->data, 0, psi->bytes, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return WL_OK;
    }
    int rc = ensure_gauge(c, WL_F64);
    if (rc) return rc;
    OpShift os(c, mh);
    // A b_i for every basis vector: Dp = (M + mh) b, DDp = (M + mh)^dag Dp
    wl_field *dp, *ddp, *t;
    if ((rc = work_field(c, W_DP64, WL_F64, WL_EVEN, nbasis, &dp)) ||
        (rc = work_field(c, W_DDP64, WL_F64, WL_EVEN, nbasis, &ddp)) ||
        (rc = work_field(c, W_T64, WL_F64, WL_ODD, nbasis, &t)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, dp->data, basis->data, t->data, kappa, +1, EPI_AXPY, nbasis, nullptr, nullptr,
                     nullptr)))
        return rc;
    if ((rc = meo_mr(c, WL_F64, ddp->data, dp->data, t->data, kappa, -1, EPI_AXPY, nbasis, nullptr, nullptr,
                     nullptr)))
        return rc;
    // every inner product in one pass, then the per-coefficient sums
    const int nc = n_coef(nbasis);
    const int nblk = std::min(kMaxBlocks, (g.vh + kBlock - 1) / kBlock);
    if ((rc = ensure_partials(c, (size_t)nc * nblk + nc))) return rc;
    double *sums = c->partials + (size_t)nc * nblk;
    HIPCHK(dispatch_nb<Dots>(nbasis, g, basis->data, ddp->data, chi->data, c->partials, nblk, c->stream));
    hipLaunchKernelGGL(k_mre_finish, dim3(nc), dim3(kBlock), 0, c->stream, c->partials, nblk, sums);
    HIPCHK(hipGetLastError());
    std::vector<double> h(nc);
    HIPCHK(hipMemcpyAsync(h.data(), sums, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int n = nbasis, nt = n * (n + 1) / 2;
    std::vector<double> G((size_t)n * n), A((size_t)n * n), v(n);
    for (int i = 0, q = 0; i < n; ++i)
        for (int j = i; j < n; ++j, ++q) {
            G[(size_t)i * n + j] = G[(size_t)j * n + i] = h[q];
            A[(size_t)i * n + j] = A[(size_t)j * n + i] = h[nt + q];
        }
    for (int i = 0; i < n; ++i) v[i] = h[2 * nt + i];
    const std::vector<double> coef = mre_coefficients(n, G, A, v);
    HIPCHK(dispatch_nb<Combine>(nbasis, g, psi->data, basis->data, coef, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}
