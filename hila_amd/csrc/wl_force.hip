// Wilson fermion force on the GPU: Dirac_Wilson_evenodd::force (libraries/dirac/wilson.h:294-311)
// with Dirac_Wilson_calc_force (:108-137), and the HMC step fermion_action::force_step
// (libraries/hmc/fermion_field.h:146-167) with gauge_field::add_momentum (hmc/gauge_field.h:249-257).
//
// D.force(chi, psi, sign) needs the odd fields tA = -kappa H_{-s} chi and tB = -kappa H_s psi
// (two hopping launches), then for every link (X, mu), X + mu its forward neighbour:
//   X even:  F = kappa [ (1 - s g)tA(X+mu) (x) psi(X)^dag + (1 + s g)tB(X+mu) (x) chi(X)^dag ]
//   X odd:   F = kappa [ (1 + s g)psi(X+mu) (x) tA(X)^dag + (1 - s g)chi(X+mu) (x) tB(X)^dag ]
// with a (x) b^dag = sum over spin of a_s b_s^dag (Wilson_vector::outer_product) and the
// antiperiodic sign on a hop across the t edge.  Since (1 + S g) = 2 Pi with Pi a Hermitian
// projector whose range is fixed by the two upper spin rows, ((1 + S g) w) (x) b^dag equals
// the sum over the two upper rows of h_w (x) h_b^dag, h = HalfWilsonVector(., mu, S): two
// 3x3 outer products per term instead of four.
//
// One launch writes F for every link (force mode), or adds project_antihermitean(U c F) to
// the momentum (momentum mode: the force step's F and add_momentum fused, F never stored).

#include <cmath>

#include "wl_host.h"
#include "wl_store.h"

namespace wl {

namespace {

enum { FORCE_OUT = 0, FORCE_MOMENTUM = 1 };

struct ForceArgs {
    const cx<double> *chi, *psi;  // even fields
    const cx<double> *tA, *tB;    // odd fields
    const cx<double> *gauge;      // gauge blocks (momentum mode)
    cx<double> *out;              // force blocks, or momentum blocks (read-modify-write), site-tiled as the links
    double kappa;                 // F carries kappa
    double coef;                  // momentum mode: mom += P(U coef F)
    Geom g;
};

// acc += ((1 + SIG g_MU) a) (x) b^dag = sum_t h_a[t] (x) h_b[t]^dag, h = the upper rows of
// (1 + SIG g_MU) ., with a the spinor at site ka of `na` (times bc) and b the one at kb of
// `nb`.  Each spinor is reduced to its half vector as soon as it is loaded.
template <int MU, int SIG>
__device__ __forceinline__ void force_term(cx<double> (&acc)[3][3], const cx<double> *na, int ka,
                                           const cx<double> *nb, int kb, int S_, double bc) {
    cx<double> ha[2][3], hb[2][3];
    {
        cx<double> w[4][3];
        ld_site<double>(na, S_, ka, w);
        project<MU, SIG>(w, ha);
    }
    {
        cx<double> w[4][3];
        ld_site<double>(nb, S_, kb, w);
        project<MU, SIG>(w, hb);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j)  // a_i conj(b_j)
                acc[i][j] = cfma(acc[i][j], bc * ha[t][i], cx<double>{hb[t][j].x, -hb[t][j].y});
}

// F for link (X, MU) of a site of parity p at checkerboard index k, forward neighbour kn:
//   even X:  (1 - s g) tA(X+mu) (x) psi(X)^dag + (1 + s g) tB(X+mu) (x) chi(X)^dag
//   odd X:   (1 + s g) psi(X+mu) (x) tA(X)^dag + (1 - s g) chi(X+mu) (x) tB(X)^dag
template <int MU, int S>
__device__ __forceinline__ void link_force(cx<double> (&F)[3][3], const ForceArgs &A, int p, int k, int kn,
                                           double bc) {
    const int S_ = A.g.stride;
    if (p == 0) {
        force_term<MU, -S>(F, A.tA, kn, A.psi, k, S_, bc);
        force_term<MU, S>(F, A.tB, kn, A.chi, k, S_, bc);
    } else {
        force_term<MU, S>(F, A.psi, kn, A.tA, k, S_, bc);
        force_term<MU, -S>(F, A.chi, kn, A.tB, k, S_, bc);
    }
}

// A workgroup takes one 64-site segment of the checkerboard (both parities in sequence: sites
// s < vh even, s >= vh odd), its four waves the four directions, so that the site's own spinors
// are fetched from memory once and served from the caches to the other three waves; segments
// are spread over the XCDs in contiguous runs (xcd_tile) so that neighbouring segments share an
// L2.  (Alternating the parities segment by segment measured no faster.)
template <int S, int MODE>
__global__ __launch_bounds__(kBlock) void k_force(ForceArgs A) {
    const int s = xcd_tile(blockIdx.x, gridDim.x) * 64 + (int)(threadIdx.x & 63);
    const Geom &g = A.g;
    if (s >= 2 * g.vh) return;
    const int mu = (int)(threadIdx.x >> 6);
    const int p = s >= g.vh;
    const int k = s - p * g.vh;
    const SiteNb nb = neighbours(g, k, p);
    const int kn = nb.fwd[mu];
    const double bc = (mu == 3 && g.antiperiodic && nb.fwd_edge_t) ? -1.0 : 1.0;
    const int S_ = g.stride;
    cx<double> F[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[i][j] = {0.0, 0.0};
    switch (mu) {
        case 0: link_force<0, S>(F, A, p, k, kn, bc); break;
        case 1: link_force<1, S>(F, A, p, k, kn, bc); break;
        case 2: link_force<2, S>(F, A, p, k, kn, bc); break;
        default: link_force<3, S>(F, A, p, k, kn, bc); break;
    }
    const size_t base = (size_t)((p * 4 + mu) * 9) * S_;
    if constexpr (MODE == FORCE_OUT) {
#pragma unroll
        for (int e = 0; e < 9; ++e) stc<true>(A.out + base + link_entry<double>(e, S_, k), A.kappa * F[e / 3][e % 3]);
    } else {
        // M = U (coef kappa F); mom += (M - M^dag)/2 - tr((M - M^dag)/2)/3
        const double c = A.coef * A.kappa;
        cx<double> U[3][3], M[3][3];
#pragma unroll
        for (int e = 0; e < 9; ++e) U[e / 3][e % 3] = ldc<true>(A.gauge + base + link_entry<double>(e, S_, k));
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                cx<double> a = cmul(U[i][0], F[0][j]);
                a = cfma(a, U[i][1], F[1][j]);
                a = cfma(a, U[i][2], F[2][j]);
                M[i][j] = c * a;
            }
        const double tr = (M[0][0].y + M[1][1].y + M[2][2].y) / 3.0;  // Im tr M / 3
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                // ((M - M^dag)/2)_ij = ((M_ij.x - M_ji.x)/2, (M_ij.y + M_ji.y)/2)
                cx<double> ah = {0.5 * (M[i][j].x - M[j][i].x), 0.5 * (M[i][j].y + M[j][i].y)};
                if (i == j) ah.y -= tr;
                cx<double> *q = A.out + base + link_entry<double>(3 * i + j, S_, k);
                stc<false>(q, ldc<false>(q) + ah);
            }
    }
}

// y -= x on every site of an even FP64 field of one rhs
__global__ __launch_bounds__(kBlock) void k_sub(cx<double> *__restrict__ y, const cx<double> *__restrict__ x, int vh,
                                                int S) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= vh) return;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        const size_t e = comp_elem<double>(c, S, k);
        y[e] = y[e] - x[e];
    }
}

}  // namespace

hipError_t launch_sub(const Geom &g, void *y, const void *x, hipStream_t s) {
    hipLaunchKernelGGL(k_sub, dim3((g.vh + kBlock - 1) / kBlock), dim3(kBlock), 0, s, (cx<double> *)y,
                       (const cx<double> *)x, g.vh, g.stride);
    return hipGetLastError();
}

hipError_t launch_force(const Geom &g, int sign, int momentum, const void *chi, const void *psi, const void *tA,
                        const void *tB, const void *gauge, void *out, double kappa, double coef, hipStream_t s) {
    ForceArgs A;
    A.chi = (const cx<double> *)chi;
    A.psi = (const cx<double> *)psi;
    A.tA = (const cx<double> *)tA;
    A.tB = (const cx<double> *)tB;
    A.gauge = (const cx<double> *)gauge;
    A.out = (cx<double> *)out;
    A.kappa = kappa;
    A.coef = coef;
    A.g = g;
    static_assert(kBlock == 4 * 64, "one wave per direction");
    const dim3 grid((unsigned)((2 * (size_t)g.vh + 63) / 64));
    if (sign > 0) {
        if (momentum)
            hipLaunchKernelGGL((k_force<1, FORCE_MOMENTUM>), grid, dim3(kBlock), 0, s, A);
        else
            hipLaunchKernelGGL((k_force<1, FORCE_OUT>), grid, dim3(kBlock), 0, s, A);
    } else {
        if (momentum)
            hipLaunchKernelGGL((k_force<-1, FORCE_MOMENTUM>), grid, dim3(kBlock), 0, s, A);
        else
            hipLaunchKernelGGL((k_force<-1, FORCE_OUT>), grid, dim3(kBlock), 0, s, A);
    }
    return hipGetLastError();
}

}  // namespace wl

namespace wl {

namespace {

// Link blocks (gauge layout, FP64) held by the context, zero when first created.
int ensure_links(wl_ctx *c, void **p, const char *what) {
    if (*p) return WL_OK;
    const size_t bytes = gauge_elems(c->g) * 16;
    if (int rc = dev_alloc(c, p, bytes, what)) return rc;
    HIPCHK(hipMemsetAsync(*p, 0, bytes, c->stream));
    return WL_OK;
}

int links_download(wl_ctx *c, const void *planes, double *const P[4]) {
    const size_t bytes = (size_t)c->V * 18 * sizeof(double);
    if (int rc = ensure_staging(c, bytes)) return rc;
    for (int mu = 0; mu < 4; ++mu) {
        HIPCHK(launch_links_to_host(c->g, planes, mu, (double *)c->staging, 0, c->stream));
        HIPCHK(hipMemcpyAsync(P[mu], c->staging, bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}

int check_vec(wl_ctx *c, const wl_field *f, const char *name) {
    if (!f) return fail(WL_E_ARG, "null %s", name);
    if (f->ctx != c) return fail(WL_E_ARG, "%s belongs to another context", name);
    if (f->parity != WL_EVEN || f->prec != WL_F64 || f->nrhs != 1)
        return fail(WL_E_ARG, "%s must be an even FP64 field of one right-hand side", name);
    return WL_OK;
}

// D.force(chi, psi, sign): tA = -kappa H_{-sign} chi and tB = -kappa H_sign psi (wilson.h:300,
// :305, as Dirac_Wilson_hop_set into the odd temporary), then one force launch.
int force_launch(wl_ctx *c, const void *chi, const void *psi, double kappa, int sign, int momentum, void *out,
                 double coef) {
    wl_field *t;
    if (int rc = work_field(c, W_T64, WL_F64, WL_ODD, 2, &t)) return rc;
    void *tA = t->data;
    void *tB = (char *)t->data + t->rhs_elems * 16;
    if (int rc = dslash_one(c, WL_F64, tA, 1, chi, nullptr, -kappa, -sign, EPI_SET, nullptr, FIN_NONE)) return rc;
    if (int rc = dslash_one(c, WL_F64, tB, 1, psi, nullptr, -kappa, sign, EPI_SET, nullptr, FIN_NONE)) return rc;
    HIPCHK(launch_force(c->g, sign, momentum, chi, psi, tA, tB, c->gauge64, out, kappa, coef, c->stream));
    return WL_OK;
}

}  // namespace

}  // namespace wl

using namespace wl;

extern "C" int wl_fermion_force(wl_ctx *c, const wl_field *chi, const wl_field *psi, double kappa, int sign,
                                double *const F[4]) {
    if (!c || !F) return fail(WL_E_ARG, "null argument");
    if (int rc = check_vec(c, chi, "chi")) return rc;
    if (int rc = check_vec(c, psi, "psi")) return rc;
    if (check_sign(sign)) return WL_E_ARG;
    for (int mu = 0; mu < 4; ++mu)
        if (!F[mu]) return fail(WL_E_ARG, "null force direction %d", mu);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc;
    if ((rc = ensure_gauge(c, WL_F64)) || (rc = ensure_links(c, &c->force64, "fermion force"))) return rc;
    if ((rc = force_launch(c, chi->data, psi->data, kappa, sign, 0, c->force64, 0.0))) return rc;
    return links_download(c, c->force64, F);
}

extern "C" int wl_momentum_upload(wl_ctx *c, const double *const P[4]) {
    if (!c || !P) return fail(WL_E_ARG, "null argument");
    for (int mu = 0; mu < 4; ++mu)
        if (!P[mu]) return fail(WL_E_ARG, "null momentum direction %d", mu);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    if (int rc = ensure_links(c, &c->mom64, "momentum")) return rc;
    const size_t bytes = (size_t)c->V * 18 * sizeof(double);
    if (int rc = ensure_staging(c, bytes)) return rc;
    for (int mu = 0; mu < 4; ++mu) {
        HIPCHK(hipMemcpyAsync(c->staging, P[mu], bytes, hipMemcpyHostToDevice, c->stream));
        HIPCHK(launch_gauge_to_soa(c->g, mu, (const double *)c->staging, c->mom64, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return WL_OK;
}

extern "C" int wl_momentum_download(wl_ctx *c, double *const P[4]) {
    if (!c || !P) return fail(WL_E_ARG, "null argument");
    for (int mu = 0; mu < 4; ++mu)
        if (!P[mu]) return fail(WL_E_ARG, "null momentum direction %d", mu);
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    if (int rc = ensure_links(c, &c->mom64, "momentum")) return rc;
    return links_download(c, c->mom64, P);
}

extern "C" int wl_hasenbusch2_force_step(wl_ctx *c, const wl_field *chi, wl_field *psi, double eps,
                                         const wl_cg_params *p, wl_cg_stats *stats) {
    if (!c || !p) return fail(WL_E_ARG, "null argument");
    if (int rc = check_vec(c, chi, "chi")) return rc;
    if (int rc = check_vec(c, psi, "psi")) return rc;
    if (chi == psi) return fail(WL_E_ARG, "chi and psi must be different fields");
    if (!std::isfinite(eps) || !std::isfinite(p->mh)) return fail(WL_E_ARG, "eps and mh must be finite");
    // Dhchi = D_h^dag chi (fermion_field.h:331), kept in a work field of the context, wrapped as
    // a one-rhs field for the solver
    wl_field *src;
    {
        std::lock_guard<std::mutex> lk(c->mu);
        DeviceGuard dg(c->device);
        int rc;
        if ((rc = ensure_gauge(c, WL_F64)) || (rc = work_field(c, W_SRC, WL_F64, WL_EVEN, 1, &src)) ||
            (rc = ensure_tmp(c, WL_F64, &c->tmp_odd64, WL_ODD)))
            return rc;
        OpShift os(c, p->mh);
        if ((rc = meo_one(c, WL_F64, src->data, chi->data, c->tmp_odd64->data, p->kappa, -1, EPI_AXPY, nullptr,
                          FIN_NONE)))
            return rc;
    }
    // psi = (D^dag D)^-1 Dhchi with the unshifted operator (CG<DIRAC_OP> inverse(D), :327, :333-334)
    wl_cg_params q = *p;
    q.mh = 0.0;
    wl_field view = *src;
    view.nrhs = 1;
    const int solved = wl_cg_solve_ex(c, psi, &view, &q, stats);
    if (solved != WL_OK && solved != WL_E_NOCONV) return solved;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc;
    wl_field *mpsi;
    if ((rc = work_field(c, W_DP64, WL_F64, WL_EVEN, 1, &mpsi)) || (rc = ensure_links(c, &c->mom64, "momentum")))
        return rc;
    // Mpsi = D psi - chi (:337-339)
    if ((rc = meo_one(c, WL_F64, mpsi->data, psi->data, c->tmp_odd64->data, p->kappa, +1, EPI_AXPY, nullptr,
                      FIN_NONE)))
        return rc;
    HIPCHK(launch_sub(c->g, mpsi->data, chi->data, c->stream));
    // force = -eps (D.force(Mpsi, psi, +1) + D.force(psi, Mpsi, -1)) = -2 eps D.force(Mpsi, psi, +1)
    // (:341-344), add_momentum fused
    if ((rc = force_launch(c, mpsi->data, psi->data, p->kappa, +1, 1, c->mom64, -2.0 * eps))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return solved;
}

extern "C" int wl_fermion_force_step(wl_ctx *c, const wl_field *chi, wl_field *psi, double eps,
                                     const wl_cg_params *p, wl_cg_stats *stats) {
    if (!c || !p) return fail(WL_E_ARG, "null argument");
    if (int rc = check_vec(c, chi, "chi")) return rc;
    if (int rc = check_vec(c, psi, "psi")) return rc;
    if (chi == psi) return fail(WL_E_ARG, "chi and psi must be different fields");
    if (!std::isfinite(eps)) return fail(WL_E_ARG, "eps must be finite");
    // psi = (D^dag D)^-1 chi from the initial guess in psi (fermion_field.h:156-158); like HILA's
    // CG, running out of iterations does not stop the step, and is reported afterwards
    const int solved = wl_cg_solve_ex(c, psi, chi, p, stats);
    if (solved != WL_OK && solved != WL_E_NOCONV) return solved;
    std::lock_guard<std::mutex> lk(c->mu);
    DeviceGuard dg(c->device);
    int rc;
    wl_field *mpsi;
    if ((rc = work_field(c, W_DP64, WL_F64, WL_EVEN, 1, &mpsi)) ||
        (rc = ensure_tmp(c, WL_F64, &c->tmp_odd64, WL_ODD)) || (rc = ensure_links(c, &c->mom64, "momentum")))
        return rc;
    {
        OpShift os(c, p->mh);  // D.apply(psi, Mpsi) of the (shifted) operator, :160
        if ((rc = meo_one(c, WL_F64, mpsi->data, psi->data, c->tmp_odd64->data, p->kappa, +1, EPI_AXPY, nullptr,
                          FIN_NONE)))
            return rc;
    }
    // force = -eps (D.force(Mpsi, psi, +1) + D.force(psi, Mpsi, -1)) (:162-165); the second
    // call forms the same products in the other order, so the sum is 2 D.force(Mpsi, psi, +1)
    // exactly; add_momentum (gauge_field.h:249-257) fused into the launch
    if ((rc = force_launch(c, mpsi->data, psi->data, p->kappa, +1, 1, c->mom64, -2.0 * eps))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return solved;
}
