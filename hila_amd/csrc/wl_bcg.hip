// Block conjugate gradient on M^dag M for N right-hand sides that share one Krylov space
// (WL_OPT_BLOCK_KRYLOV): the block generalisation of HILA's single-rhs recurrence
// (libraries/dirac/conjugate_gradient.h:46-104), in the form with an orthonormalised residual
// block (Dubrulle, "Retooling the method of block conjugate gradients", ETNA 12 (2001)):
//
//   R = B - A X = Q C,  D = Q                       (Q: N orthonormal columns, C: N x N)
//   repeat:  Z = A D
//            M = (D^H Z)^+,  X += D (M C),  Q <- Q - Z M
//            Q S = Q  (orthonormalised: Q <- Q S^-1),  D <- Q + D S^H,  C <- S C
//   ||R_j|| = ||C e_j||.
//
// Orthonormalisation is Cholesky QR on the Gram matrix when it is well conditioned, else SVQB
// (eigendecomposition, wl_small.h), which drops directions whose Gram eigenvalue falls below
// 1e-12 of the largest: dependent or converged directions (a repeated source, a zero source)
// leave the block instead of breaking the recurrence; M is the Cholesky inverse or the matching
// thresholded pseudo-inverse.  The N x N algebra runs on
// the host in FP64 after each of the two Gram reductions of an iteration (2 x N^2 complex
// numbers each way); everything on the lattice runs on the device.
//
// Precision: the inner operator and the block vectors Q, D, Z are in FP64 or FP32; the Gram
// matrices are accumulated in FP64; with an FP32 inner precision X += D (M C) goes to an FP32
// inner solution Y, and reliable updates (r = B - A X with the FP64 operator, then
// Q C = r) restart the residual block when the estimated residuals have dropped by `delta`
// since the last update, every `update_every` iterations, and when every estimate meets the
// tolerance; convergence is declared only on the FP64 true residual of every right-hand side.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "wl_small.h"
#include "wl_solve.h"
#include "wl_store.h"
#include <type_traits>

namespace wl {

namespace {

constexpr int kBcgMax = 24;    // right-hand sides of one block

// The FP32 inner solution of the mixed solvers (BatchCG's y, folded into x by k_fold_mr) is
// laid out as the FP64 solution is, one complex number per component: unpacked FP32.
struct f32u {};
constexpr int kF32Unpacked = 100;  // its precision code in launch_bcombo
constexpr int kGramWg = 1024;  // workgroups of a Gram reduction (partials)

// The 16-byte vector q (components q P .. q P + P - 1, P = St<T>::pack) of site k of rhs block
// `base`, widened to the arithmetic type; VP = 16-byte vectors per spinor (12 / P).
template <typename T>
struct Vec16 {
    static constexpr int P = St<T>::pack, VP = 12 / P;
    cx<typename St<T>::R> e[P];
};
template <typename T>
__device__ __forceinline__ Vec16<T> ld_vec16(const typename St<T>::E *base, int q, int S, int k) {
    const auto raw = ld_vec<false, T>(base + comp_elem<T>(St<T>::pack * q, S, k));
    Vec16<T> v;
#pragma unroll
    for (int h = 0; h < St<T>::pack; ++h) v.e[h] = widen(raw.e[h]);
    return v;
}

// G[i][j] = sum over sites and components of conj(U_i) V_j (FP64 accumulation), i, j < N, for
// N-rhs fields U, V in storage T (FP64 or FP32).  A workgroup takes tiles of 64 sites; per
// 16-byte vector q of the spinor it stages the tile of every rhs of U and V in LDS (the next
// vector's loads are issued before the current one is used), and thread (bi, bj, g)
// accumulates the 3 x 3 block (3bi.., 3bj..) of G over the 16 sites of group g.  One partial
// G per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_bgram(const typename St<T>::E *__restrict__ U,
                                               const typename St<T>::E *__restrict__ V, int N, int vh, int S,
                                               cx<double> *__restrict__ partials) {
    using R = typename St<T>::R;
    using Vt = Vec16<T>;
    constexpr int P = Vt::P, VP = Vt::VP;
    constexpr int PER = kBcgMax * 64 / 256;  // vectors per field a thread stages
    // the staged tiles, then (after the last tile) the four site groups' sums
    // (rows padded to 65 vectors: the 8 rows a wave reads at once then fall in different banks)
    constexpr int RW = 65;
    constexpr size_t kTiles = (size_t)2 * kBcgMax * RW * sizeof(Vt), kRed = (size_t)4 * 64 * 9 * sizeof(cx<double>);
    __shared__ __align__(16) unsigned char smem[kTiles > kRed ? kTiles : kRed];
    auto su = reinterpret_cast<Vt(*)[RW]>(smem);
    auto sv = reinterpret_cast<Vt(*)[RW]>(smem + (size_t)kBcgMax * RW * sizeof(Vt));
    auto red = reinterpret_cast<cx<double>(*)[64][9]>(smem);
    const int t = threadIdx.x, blk = t & 63, g = t >> 6;
    const int bi = blk >> 3, bj = blk & 7;
    const size_t rs = (size_t)St<T>::planes * S;
    cx<double> acc[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) acc[a][b] = {0.0, 0.0};
    const int ntiles = (vh + 63) / 64;
    const int nsteps = ((ntiles - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x) * VP;
    // step = (tile, vector q); the loads of step + 1 are in flight while step is computed
    Vt pu[PER], pv[PER];
    auto load = [&](int step) {
        const int tile = (int)blockIdx.x + (step / VP) * (int)gridDim.x, q = step % VP;
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int e = t + 256 * r, i = e >> 6, k = tile * 64 + (e & 63);
            if (i < N && k < vh) {
                pu[r] = ld_vec16<T>(U + (size_t)i * rs, q, S, k);
                pv[r] = ld_vec16<T>(V + (size_t)i * rs, q, S, k);
            } else {
#pragma unroll
                for (int h = 0; h < P; ++h) pu[r].e[h] = pv[r].e[h] = {R(0), R(0)};
            }
        }
    };
    if (nsteps > 0) load(0);
    for (int step = 0; step < nsteps; ++step) {
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const int e = t + 256 * r;
            su[e >> 6][e & 63] = pu[r];
            sv[e >> 6][e & 63] = pv[r];
        }
        __syncthreads();
        if (step + 1 < nsteps) load(step + 1);
#pragma unroll 2
        for (int s2 = 0; s2 < 16; ++s2) {
            const int l = g * 16 + s2;
#pragma unroll
            for (int h = 0; h < P; ++h) {
                cx<double> u[3], v[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const cx<R> uu = su[3 * bi + a][l].e[h], vv = sv[3 * bj + a][l].e[h];
                    u[a] = {(double)uu.x, (double)uu.y};
                    v[a] = {(double)vv.x, (double)vv.y};
                }
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        // conj(u) v
                        acc[a][b].x += u[a].x * v[b].x + u[a].y * v[b].y;
                        acc[a][b].y += u[a].x * v[b].y - u[a].y * v[b].x;
                    }
            }
        }
        __syncthreads();
    }
    // sum the four site groups (fixed order) through LDS (the loop ends with a barrier)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) red[g][blk][3 * a + b] = acc[a][b];
    __syncthreads();
    if (g == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int q = 3 * a + b;
                const cx<double> s = (red[0][blk][q] + red[1][blk][q]) + (red[2][blk][q] + red[3][blk][q]);
                const int i = 3 * bi + a, j = 3 * bj + b;
                if (i < N && j < N) partials[((size_t)blockIdx.x * N + i) * N + j] = s;
            }
    }
}

// out[e] = sum over the nwg partial G of entry e: one workgroup per entry, a fixed-order
// reduction over the workgroup's lanes
__global__ __launch_bounds__(kBlock) void k_bgram_finish(const cx<double> *__restrict__ partials, int nwg, int N,
                                                         cx<double> *out) {
    const int e = blockIdx.x, NN = N * N;
    cx<double> s = {0.0, 0.0};
    for (int w = threadIdx.x; w < nwg; w += blockDim.x) s = s + partials[(size_t)w * NN + e];
    const double re = block_sum(s.x), im = block_sum(s.y);
    if (threadIdx.x == 0) out[e] = {re, im};
}

// Out_j = sum_i In_i A[i][j] + gamma Add_j for every site and component (i, j < N), with A an
// N x N complex FP64 matrix (row-major).  A thread takes one site and the component pair
// (2v, 2v + 1): one 16-byte vector of an FP32 field, two of an FP64 one.  Out may be In (every
// In_i of the thread's components is loaded before any Out_j is stored) and may be Add.
// Arithmetic in the wider of the In and Out storage types.
// Components (c0, c0 + 1) of site k in rhs block `base`: one 16-byte load where the storage
// packs them together (FP32), else two.
template <typename T, typename R>
__device__ __forceinline__ void ld_pair(const typename St<T>::E *base, int S, int k, int c0, cx<R> (&v)[2]) {
    if constexpr (St<T>::pack == 2) {
        const auto raw = ld_vec<false, T>(base + comp_elem<T>(c0, S, k));
        v[0] = {(R)raw.e[0].x, (R)raw.e[0].y};
        v[1] = {(R)raw.e[1].x, (R)raw.e[1].y};
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const auto a = base[comp_elem<T>(c0 + h, S, k)];
            v[h] = {(R)a.x, (R)a.y};
        }
    }
}
template <typename T, typename R>
__device__ __forceinline__ void st_pair(typename St<T>::E *base, int S, int k, int c0, const cx<R> (&v)[2]) {
    using RO = typename St<T>::R;
    if constexpr (St<T>::pack == 2) {
        Packed<typename St<T>::E, 2> raw;
        raw.e[0] = {(RO)v[0].x, (RO)v[0].y};
        raw.e[1] = {(RO)v[1].x, (RO)v[1].y};
        st_vec<false, T>(base + comp_elem<T>(c0, S, k), raw);
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) base[comp_elem<T>(c0 + h, S, k)] = {(RO)v[h].x, (RO)v[h].y};
    }
}

// Out_j = sum_i In_i A[i][j] + gamma Add_j for every site and component (i, j < N), with A an
// N x N complex FP64 matrix (row-major), held in LDS in the arithmetic type.  A thread takes
// one site and the component pair (2v, 2v + 1): one 16-byte vector of an FP32 field, two
// loads of an FP64 one.  Out may be In (every In_i of the thread's components is loaded before
// any Out_j is stored) and may be Add.  Arithmetic in the wider of the In and Out types.
template <typename TI, typename TO, typename TA>
__global__ __launch_bounds__(256) void k_bcombo(const typename St<TI>::E *In, typename St<TO>::E *Out,
                                                const typename St<TA>::E *Add, double gamma,
                                                const cx<double> *__restrict__ Am, int N, int vh, int S) {
    using RI = typename St<TI>::R;
    using RO = typename St<TO>::R;
    using R = typename std::conditional<(sizeof(RI) > sizeof(RO)), RI, RO>::type;
    __shared__ cx<R> A[kBcgMax * kBcgMax];
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) A[e] = {(R)Am[e].x, (R)Am[e].y};
    __syncthreads();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = 2 * (int)blockIdx.y;  // components c0, c0 + 1
    if (k >= vh) return;
    const size_t rsi = (size_t)St<TI>::planes * S, rso = (size_t)St<TO>::planes * S,
                 rsa = (size_t)St<TA>::planes * S;
    cx<R> w[kBcgMax][2];
#pragma unroll
    for (int i = 0; i < kBcgMax; ++i)
        if (i < N) ld_pair<TI, R>(In + (size_t)i * rsi, S, k, c0, w[i]);
    const R gm = (R)gamma;
    for (int j = 0; j < N; ++j) {
        cx<R> s[2] = {{R(0), R(0)}, {R(0), R(0)}};
        if (Add) {
            ld_pair<TA, R>(Add + (size_t)j * rsa, S, k, c0, s);
#pragma unroll
            for (int h = 0; h < 2; ++h) s[h] = {gm * s[h].x, gm * s[h].y};
        }
#pragma unroll
        for (int i = 0; i < kBcgMax; ++i)
            if (i < N) {
                const cx<R> m = A[i * N + j];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    s[h].x += w[i][h].x * m.x - w[i][h].y * m.y;
                    s[h].y += w[i][h].x * m.y + w[i][h].y * m.x;
                }
            }
        st_pair<TO, R>(Out + (size_t)j * rso, S, k, c0, s);
    }
}

}  // namespace

template <>
struct St<f32u> {
    using R = float;
    using E = cx<float>;
    static constexpr int planes = 12;
    static constexpr int pack = 1;
    static constexpr bool scaled = false;
};

namespace {

template <typename TI, typename TO, typename TA>
hipError_t combo_t(const Geom &g, const void *in, void *out, const void *add, double gamma, const void *A, int N,
                   hipStream_t s) {
    const dim3 grid((unsigned)((g.vh + 255) / 256), 6);
    hipLaunchKernelGGL((k_bcombo<TI, TO, TA>), grid, dim3(256), 0, s, (const typename St<TI>::E *)in,
                       (typename St<TO>::E *)out, (const typename St<TA>::E *)add, gamma, (const cx<double> *)A, N,
                       g.vh, g.stride);
    return hipGetLastError();
}

// Out = In A + gamma Add with the storage precisions of the three fields (FP64 / FP32)
hipError_t launch_bcombo(const Geom &g, int pin, const void *in, int pout, void *out, int padd, const void *add,
                         double gamma, const void *A, int N, hipStream_t s) {
    if (!add) padd = pout;
    if (pin == WL_F64 && pout == WL_F64 && padd == WL_F64) return combo_t<double, double, double>(g, in, out, add, gamma, A, N, s);
    if (pin == WL_F32 && pout == WL_F32 && padd == WL_F32) return combo_t<float, float, float>(g, in, out, add, gamma, A, N, s);
    if (pin == WL_F64 && pout == WL_F32 && padd == WL_F32) return combo_t<double, float, float>(g, in, out, add, gamma, A, N, s);
    if (pin == WL_F32 && pout == WL_F64 && padd == WL_F64) return combo_t<float, double, double>(g, in, out, add, gamma, A, N, s);
    if (pin == WL_F32 && pout == kF32Unpacked && padd == kF32Unpacked)
        return combo_t<float, f32u, f32u>(g, in, out, add, gamma, A, N, s);
    return hipErrorInvalidValue;
}

using small::cplx;
using small::Mat;

// The block solver's state on one context for one solve.
struct BlockCG {
    wl_ctx *c;
    int N, prec;
    CgSpec s;
    void *x;
    const void *b;
    wl_field *Q, *D, *Z, *DP, *T, *Y, *R64, *DP64, *T64;
    cx<double> *dmat;   // device: 4 matrices of N x N (operands of the combination kernels)
    cx<double> *dgram;  // device: the reduced Gram matrix
    std::vector<cx<double>> hbuf;

    int gram(int p, const void *U, const void *V, Mat &G) {
        const int nwg = std::min(kGramWg, (c->g.vh + 63) / 64);
        hipStream_t S = c->stream;
        if (p == WL_F64)
            hipLaunchKernelGGL(k_bgram<double>, dim3(nwg), dim3(256), 0, S, (const cx<double> *)U,
                               (const cx<double> *)V, N, c->g.vh, c->g.stride, (cx<double> *)c->partials);
        else
            hipLaunchKernelGGL(k_bgram<float>, dim3(nwg), dim3(256), 0, S, (const cx<float> *)U, (const cx<float> *)V,
                               N, c->g.vh, c->g.stride, (cx<double> *)c->partials);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_bgram_finish, dim3(N * N), dim3(kBlock), 0, S, (const cx<double> *)c->partials, nwg, N,
                           dgram);
        HIPCHK(hipGetLastError());
        hbuf.resize((size_t)N * N);
        HIPCHK(hipMemcpyAsync(hbuf.data(), dgram, (size_t)N * N * sizeof(cx<double>), hipMemcpyDeviceToHost, S));
        HIPCHK(hipStreamSynchronize(S));
        G.assign((size_t)N * N, cplx(0));
        for (int e = 0; e < N * N; ++e) G[e] = cplx(hbuf[e].x, hbuf[e].y);
        return WL_OK;
    }

    // upload the host matrices Ms[0..n) into device slots 0..n
    int upload(std::initializer_list<const Mat *> Ms) {
        hbuf.resize((size_t)4 * N * N);
        int slot = 0;
        for (const Mat *m : Ms) {
            for (int e = 0; e < N * N; ++e) hbuf[(size_t)slot * N * N + e] = {(*m)[e].real(), (*m)[e].imag()};
            ++slot;
        }
        HIPCHK(hipMemcpyAsync(dmat, hbuf.data(), (size_t)slot * N * N * sizeof(cx<double>), hipMemcpyHostToDevice,
                              c->stream));
        // the host buffer is reused by the next transfer: wait for this one
        HIPCHK(hipStreamSynchronize(c->stream));
        return WL_OK;
    }
    const cx<double> *mat(int slot) const { return dmat + (size_t)slot * N * N; }

    // R64 = B - A X with the FP64 operator; Gram of R64 on the host
    int true_residual(Mat &G) {
        int rc;
        if ((rc = meo_mr(c, WL_F64, DP64->data, x, T64->data, s.kappa, +1, EPI_AXPY, N, nullptr, nullptr, nullptr)))
            return rc;
        if ((rc = meo_mr(c, WL_F64, R64->data, DP64->data, T64->data, s.kappa, -1, EPI_AXPY, N, nullptr, nullptr,
                         nullptr)))
            return rc;
        // R64 = B - R64 (A = -I)
        Mat mI = small::identity(N);
        for (auto &v : mI) v = -v;
        if ((rc = upload({&mI}))) return rc;
        HIPCHK(launch_bcombo(c->g, WL_F64, R64->data, WL_F64, R64->data, WL_F64, b, 1.0, mat(0), N, c->stream));
        return gram(WL_F64, R64->data, R64->data, G);
    }
};

}  // namespace

int cg_block(wl_ctx *c, void *x, const void *b, int N, const CgSpec &spec, wl_cg_stats *stats) {
    if (N < 1 || N > kBcgMax) return fail(WL_E_ARG, "block CG takes 1..%d right-hand sides, not %d", kBcgMax, N);
    if (spec.prec != WL_F64 && spec.prec != WL_F32)
        return fail(WL_E_NOTIMPL, "block CG runs with an FP64 or FP32 inner precision");
    const auto t0 = std::chrono::steady_clock::now();
    BlockCG B;
    B.c = c;
    B.N = N;
    B.prec = spec.prec;
    B.s = spec;
    B.x = x;
    B.b = b;
    c->op_shift = spec.mh;
    int rc;
    const int p = spec.prec;
    const bool mixed = p != WL_F64;
    if ((rc = ensure_gauge(c, p)) || (rc = ensure_gauge(c, WL_F64))) return rc;
    if ((rc = work_field(c, W_R, p, WL_EVEN, N, &B.Q)) || (rc = work_field(c, W_P, p, WL_EVEN, N, &B.D)) ||
        (rc = work_field(c, W_BZ, p, WL_EVEN, N, &B.Z)) || (rc = work_field(c, W_DP, p, WL_EVEN, N, &B.DP)) ||
        (rc = work_field(c, W_T, p, WL_ODD, N, &B.T)) || (rc = work_field(c, W_DDP64, WL_F64, WL_EVEN, N, &B.R64)) ||
        (rc = work_field(c, W_DP64, WL_F64, WL_EVEN, N, &B.DP64)) ||
        (rc = work_field(c, W_T64, WL_F64, WL_ODD, N, &B.T64)))
        return rc;
    B.Y = nullptr;
    if (mixed && (rc = work_field(c, W_Y, WL_F32, WL_EVEN, N, &B.Y))) return rc;
    const int nwg = std::min(kGramWg, (c->g.vh + 63) / 64);
    if ((rc = ensure_partials(c, (size_t)nwg * N * N * 2 + 64))) return rc;
    if ((rc = ensure_staging(c, (size_t)5 * N * N * sizeof(cx<double>)))) return rc;
    B.dmat = (cx<double> *)c->staging;
    B.dgram = B.dmat + (size_t)4 * N * N;
    hipStream_t S = c->stream;
    // tiny: eigenvalues below it (relative) are dependent directions, dropped from the block;
    // chol_tol: Cholesky pivots must clear it (relative) for the fast path, else the
    // eigendecomposition decides what to drop
    const double tol = spec.accuracy, tiny = 1e-12, chol_tol = 1e-10;

    // source norms; zero sources give x = 0 and count as converged (HILA would divide 0 by 0)
    Mat G;
    if ((rc = B.gram(WL_F64, b, b, G))) return rc;
    std::vector<double> bn(N);
    for (int j = 0; j < N; ++j) {
        bn[j] = std::sqrt(std::fabs(G[(size_t)j * N + j].real()));
        if (bn[j] == 0.0) HIPCHK(launch_zero_rhs(c->g, x, j, S));
    }
    if (mixed) HIPCHK(hipMemsetAsync(B.Y->data, 0, B.Y->bytes, S));

    std::vector<double> est(N, 0.0), truer(N, 0.0), at_update(N, 0.0);
    Mat C, Sm, Sinv;
    int nrel = 0, it = 0, it_update = 0;
    bool converged = false;
    // restart the residual block from the FP64 true residual: R = Q C; true norms in truer
    auto restart = [&](bool first) -> int {
        Mat G64;
        int r;
        if ((r = B.true_residual(G64))) return r;
        bool all = true;
        for (int j = 0; j < N; ++j) {
            truer[j] = bn[j] > 0 ? std::sqrt(std::fabs(G64[(size_t)j * N + j].real())) / bn[j] : 0.0;
            all = all && truer[j] <= tol;
        }
        if (all) {
            converged = true;
            return WL_OK;
        }
        small::svqb(N, G64, tiny, Sm, Sinv);
        C = Sm;
        if ((r = B.upload({&Sinv}))) return r;
        // the new orthonormal residual block, in Z's storage (Z is recomputed before it is read)
        HIPCHK(launch_bcombo(c->g, WL_F64, B.R64->data, p, B.Z->data, p, nullptr, 0.0, B.mat(0), N, S));
        if (first) {
            HIPCHK(launch_copy(p, c->g, B.Z->data, B.D->data, N, S));
        } else {
            // The recurrence needs D^H Q = I.  The new block spans (to rounding) the old block's
            // space in another orthonormal basis, Q_new = Q_old U: rotate D with it, D <- D U,
            // U = Q_old^H Q_new.
            Mat Urot;
            if ((r = B.gram(p, B.Q->data, B.Z->data, Urot))) return r;
            if ((r = B.upload({&Urot}))) return r;
            HIPCHK(launch_bcombo(c->g, p, B.D->data, p, B.D->data, p, nullptr, 0.0, B.mat(0), N, S));
        }
        std::swap(B.Q, B.Z);
        for (int j = 0; j < N; ++j) at_update[j] = est[j] = truer[j];
        it_update = it;
        if (!first) ++nrel;
        return WL_OK;
    };
    if ((rc = restart(true))) return rc;
    while (!converged && it < spec.maxiters) {
        ++it;
        // Z = A D
        if ((rc = meo_mr(c, p, B.DP->data, B.D->data, B.T->data, spec.kappa, +1, EPI_AXPY, N, nullptr, nullptr,
                         nullptr)) ||
            (rc = meo_mr(c, p, B.Z->data, B.DP->data, B.T->data, spec.kappa, -1, EPI_AXPY, N, nullptr, nullptr,
                         nullptr)))
            return rc;
        Mat H;
        if ((rc = B.gram(p, B.D->data, B.Z->data, H))) return rc;
        Mat M;
        small::inverse_herm(N, H, chol_tol, M);
        Mat Coef = small::mul(N, M, C), mM = M;
        for (auto &v : mM) v = -v;
        if ((rc = B.upload({&Coef, &mM}))) return rc;
        // X (or Y) += D Coef;  Q += Z (-M)
        if (mixed)
            HIPCHK(launch_bcombo(c->g, p, B.D->data, kF32Unpacked, B.Y->data, kF32Unpacked, B.Y->data, 1.0, B.mat(0),
                                 N, S));
        else
            HIPCHK(launch_bcombo(c->g, WL_F64, B.D->data, WL_F64, x, WL_F64, x, 1.0, B.mat(0), N, S));
        HIPCHK(launch_bcombo(c->g, p, B.Z->data, p, B.Q->data, p, B.Q->data, 1.0, B.mat(1), N, S));
        // Q S = Q; D = Q + D S^H; C = S C
        Mat Gq;
        if ((rc = B.gram(p, B.Q->data, B.Q->data, Gq))) return rc;
        small::orthonormaliser(N, Gq, chol_tol, Sm, Sinv);
        Mat SH = small::adj(N, Sm);
        C = small::mul(N, Sm, C);
        if ((rc = B.upload({&Sinv, &SH}))) return rc;
        HIPCHK(launch_bcombo(c->g, p, B.Q->data, p, B.Q->data, p, nullptr, 0.0, B.mat(0), N, S));
        HIPCHK(launch_bcombo(c->g, p, B.D->data, p, B.D->data, p, B.Q->data, 1.0, B.mat(1), N, S));
        bool est_done = true, dropped = true;
        for (int j = 0; j < N; ++j) {
            double s2 = 0;
            for (int i = 0; i < N; ++i) s2 += std::norm(C[(size_t)i * N + j]);
            est[j] = bn[j] > 0 ? std::sqrt(s2) / bn[j] : 0.0;
            est_done = est_done && est[j] <= tol;
            // rhs already at the tolerance at the last update do not hold the others back
            dropped = dropped && (at_update[j] <= tol || est[j] <= spec.delta * at_update[j]);
        }
        const bool update = est_done || (mixed && (dropped || it - it_update >= spec.update_every));
        if (update) {
            if (mixed) HIPCHK(launch_fold_mr(c->g, x, B.Y->data, c->st_mr, 1, N, nullptr, S));
            if ((rc = restart(false))) return rc;
        }
    }
    if (!converged) {  // maxiters: the final true residual
        if (mixed) HIPCHK(launch_fold_mr(c->g, x, B.Y->data, c->st_mr, 1, N, nullptr, S));
        Mat G64;
        if ((rc = B.true_residual(G64))) return rc;
        for (int j = 0; j < N; ++j)
            truer[j] = bn[j] > 0 ? std::sqrt(std::fabs(G64[(size_t)j * N + j].real())) / bn[j] : 0.0;
    }
    HIPCHK(hipStreamSynchronize(S));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    bool all = true;
    for (int j = 0; j < N; ++j) {
        const bool conv = truer[j] <= tol;
        all = all && conv;
        if (stats) {
            std::memset(&stats[j], 0, sizeof(stats[j]));
            stats[j].iterations = it;
            stats[j].converged = conv ? 1 : 0;
            stats[j].rr_rel = est[j] * est[j];
            stats[j].true_resid = truer[j];
            stats[j].ms = ms;
            stats[j].reliable_updates = nrel;
        }
    }
    if (!all) return fail(WL_E_NOCONV, "block CG: not every right-hand side converged in %d iterations", it);
    return WL_OK;
}

}  // namespace wl
