/*
 * wilson_mi355x.h -- C ABI of the MI355X (gfx950) Wilson even-odd Dirac operator and
 * conjugate-gradient inverter.
 *
 * This library replaces one path of HILA (CFT-HY/HILA):
 *   - Dirac_Wilson_evenodd<SU<3,double>>::apply / ::dagger  (libraries/dirac/wilson.h:272-289)
 *     together with the hopping kernels it calls, Dirac_Wilson_hop / Dirac_Wilson_hop_set
 *     (libraries/dirac/wilson.h:15-95);
 *   - CG<Op>::apply  (libraries/dirac/conjugate_gradient.h:46-104).
 *
 * Conventions (identical to HILA's on this path):
 *   - 4 dimensions, SU(3) colour, directions e_x,e_y,e_z,e_t = 0..3
 *     (libraries/plumbing/coordinates.h:34-44).
 *   - Host field buffers are HILA's CPU ("vanilla") AoS layout: one element per site, sites in
 *     the EVEN_SITES_FIRST order (libraries/plumbing/lattice.cpp:171-186): the lexicographic
 *     index i = x + Lx*(y + Ly*(z + Lz*t)) (x fastest) maps to i/2 for even sites and
 *     i/2 + V/2 for odd sites.  A Wilson vector is 24 doubles [spin 0..3][colour 0..2]{re,im}
 *     (libraries/datatypes/wilson_vector.h:56-63); an SU(3) link is 18 doubles, row-major
 *     3x3 complex (libraries/datatypes/matrix.h:130).
 *   - gamma basis of libraries/datatypes/wilson_vector.h:311-347, direction->gamma map
 *     {g1,g2,g3,g0} (:23), unnormalised projectors (1 +- gamma) (:390-567).
 *   - D_s psi(x) = psi(x) - kappa sum_mu [ U_mu(x)(1 + s g_mu) psi(x+mu)
 *                                       + U_mu(x-mu)^dag (1 - s g_mu) psi(x-mu) ],
 *     s = +1 for apply, -1 for dagger.  The backward hop uses the link and spinor of x-mu;
 *     HILA's wilson.h:47 reads the forward neighbour's temporary there instead, which makes
 *     dagger differ from the adjoint of apply.  This library implements the adjoint-consistent
 *     operator that HILA's own adjointness test (applications/test_cases/test_CG.cpp:236-245)
 *     requires.
 *   - M_s = 1 - kappa^2 D_s,eo D_s,oe acting on even sites; the odd half of an
 *     apply/dagger output is zero (wilson.h:278).
 *   - Boundary: periodic in x,y,z; antiperiodic in t when requested (psi(x + Lt e_t) = -psi(x)),
 *     as HILA's SPECIAL_BOUNDARY_CONDITIONS fermion fields (test_CG.cpp:222-226).
 *
 * Errors: every function returns WL_OK (0) or a negative WL_E_* code; no exceptions cross
 * this boundary.  wl_last_error() returns a per-thread message for the last failure.
 * Threading: one wl_ctx per host thread, or serialise calls on a shared ctx.  All device work
 * of a ctx is ordered on that ctx's own HIP stream.
 */
#ifndef WILSON_MI355X_H
#define WILSON_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
enum {
    WL_OK = 0,
    WL_E_ARG = -1,     /* invalid argument (bad size, null pointer, mismatched field)        */
    WL_E_OOM = -2,     /* device allocation failed / admission refused by the HBM budget     */
    WL_E_DEVICE = -3,  /* HIP runtime error                                                  */
    WL_E_NOCONV = -4,  /* CG reached maxiters without meeting the tolerance                  */
    WL_E_STATE = -5,   /* call out of order (e.g. no gauge field uploaded)                   */
    WL_E_NOTIMPL = -6, /* option not provided by this build                                  */
    WL_E_BUSY = -7     /* request not finished within the wait's time limit                  */
};

/* Parity, numbered as HILA's enum class Parity {none, even, odd, all}
 * (libraries/plumbing/coordinates.h:182). */
enum { WL_EVEN = 1, WL_ODD = 2 };

/* Storage / arithmetic precision of a device field or operator. */
enum { WL_F64 = 0, WL_F32 = 1, WL_F16 = 2 };

typedef struct wl_ctx wl_ctx;
typedef struct wl_field wl_field;

/* Statistics of one CG solve (an addition: HILA's CG prints steps, time and the squared
 * iterated residual ratio, conjugate_gradient.h:98-103, and returns nothing). */
typedef struct wl_cg_stats {
    int iterations;          /* CG iterations executed (HILA's printed step count + 1)      */
    int converged;           /* 1 if ||r||^2 < accuracy^2 ||b||^2 was reached                */
    double rr_rel;           /* final iterated ||r||^2 / ||b||^2 (HILA's "relative residue") */
    double true_resid;       /* ||b - M^dag M x|| / ||b||, recomputed in FP64 after the solve */
    double ms;               /* wall time of the solve on the host clock, milliseconds       */
    int reliable_updates;    /* mixed-precision solves: number of FP64 reliable updates      */
    int pad_;
} wl_cg_stats;

/* ---- library / context ---------------------------------------------------------------- */

/* Version string of this build. */
const char *wl_version(void);
/* Message for the last failing call on this host thread ("" if none). */
const char *wl_last_error(void);

/* Create a context for one L = dims[0..3] = (Lx,Ly,Lz,Lt) lattice on HIP device `device`.
 * Every L must be even and >= 2 (even-odd decomposition).  Replaces lattice.setup()
 * (libraries/plumbing/lattice.h:477-480) for this path. */
int wl_create(const int dims[4], int antiperiodic_t, int device, wl_ctx **out);
int wl_destroy(wl_ctx *ctx);
/* Block the host until all work queued on the context's stream has finished. */
int wl_sync(wl_ctx *ctx);
/* Number of lattice sites V, and V/2 (sites per parity). */
int64_t wl_volume(const wl_ctx *ctx);

/* Upload the gauge field.  U[mu] points at HILA's Field<SU<3,double>> gauge[mu] buffer
 * (V sites x 18 doubles, EVEN_SITES_FIRST order).  Replaces the non-owning gauge reference
 * that Dirac_Wilson_evenodd holds (wilson.h:237,261). */
int wl_gauge_upload(wl_ctx *ctx, const double *const U[4]);
/* The gauge field back in HILA's layout (U[mu]: V x 18 doubles, EVEN_SITES_FIRST), as it is
 * held on the device: what a GaugeField read from a file holds (gaugefield.h:219-272). */
int wl_gauge_download(wl_ctx *ctx, double *const U[4]);

/* ---- gauge configuration files ------------------------------------------------------ */

/* HILA's configuration file of a GaugeField<SU<3,double>> (libraries/plumbing/gaugefield.h:
 * 194-272): 7 native-endian int64 {394824242, 4, 144, Lx, Ly, Lz, Lt}, then for mu = 0..3 the
 * link U_mu(x) of every site in lexicographic order with x fastest (lattice.h:389-396,
 * Field<T>::write field_io.h:125-160), each 18 doubles (3x3 complex, row-major).  Errors are
 * WL_E_ARG with HILA's "CONFIG ERROR in file ..." text where config_read would terminate. */

/* The lattice size in the header of a configuration file, after config_read's checks of the
 * id, the dimensionality and the element size (gaugefield.h:219-262), and of the file's
 * length. */
int wl_config_header(const char *path, int dims[4]);
/* GaugeField::config_read (gaugefield.h:219-272): load the file's links as the context's
 * gauge field.  Every file dimension must divide the lattice's; a smaller file lattice is
 * replicated periodically to fill it (Field<T>::read(stream, insize), field_io.h:245-312).
 * A failure after the header checks leaves the context without a gauge field. */
int wl_config_read(wl_ctx *ctx, const char *path);
/* GaugeField::config_write (gaugefield.h:194-217): write the context's gauge field. */
int wl_config_write(wl_ctx *ctx, const char *path);

/* ---- device fields (one parity, nrhs right-hand sides) -------------------------------- */

/* A device Wilson-vector field on the sites of one parity (WL_EVEN / WL_ODD), holding nrhs
 * independent vectors (right-hand sides), stored in precision prec.  The operator applies to
 * all right-hand sides of a field in one pass, each lane loading a link once for several of
 * them (WL_OPT_RHS_BLOCK). */
int wl_field_create(wl_ctx *ctx, int parity, int prec, int nrhs, wl_field **out);
/* Fields should be destroyed before their context.  A field still alive when wl_destroy runs
 * loses its device memory and becomes an orphan: every call on it returns WL_E_STATE except
 * wl_field_destroy, which releases the handle. */
int wl_field_destroy(wl_field *f);
/* Copy rhs vector `rhs` of f from / to host AoS buffer `host` of V/2 sites x 24 doubles, i.e.
 * the parity block of a HILA Field<WilsonVector<3,double>> buffer: its first V/2 sites for
 * WL_EVEN, the next V/2 for WL_ODD. */
int wl_field_upload(wl_field *f, int rhs, const double *host);
int wl_field_download(const wl_field *f, int rhs, double *host);
int wl_field_zero(wl_field *f);
/* dst = src (same parity and nrhs): HILA's Field assignment dst[par] = src[X]
 * (libraries/plumbing/field.h:898-909), on the device.  When the precisions differ each
 * component is converted (rounded once to the destination's precision), as HILA's assignment
 * of a double field to a float one, c[ALL] = chi[X] (libraries/hmc/fermion_field.h:77-78). */
int wl_field_copy(wl_field *dst, const wl_field *src);
/* Right-hand side jd of dst = right-hand side js of src (same parity and precision): e.g.
 * keeping a solution in a slot of a basis of previous solutions, as
 * fermion_action::save_new_solution does (libraries/hmc/fermion_field.h:134-142). */
int wl_field_copy_rhs(wl_field *dst, int jd, const wl_field *src, int js);
/* sum over sites of ||f||^2 for one rhs, computed on the device in FP64 (the reduction of
 * conjugate_gradient.h:61,72,79,87). */
int wl_field_norm2(const wl_field *f, int rhs, double *out);

/* ---- operator ------------------------------------------------------------------------- */

/* Hopping half-operator, HILA's Dirac_Wilson_hop_set(gauge, kappa, in, out, par, sign)
 * (wilson.h:53-95): out = -kappa * H_s in, with H_s the sum over the 8 hops of
 * D_s above; out has parity opposite to in.  All nrhs vectors are processed together. */
int wl_dslash_eo(wl_ctx *ctx, wl_field *out, const wl_field *in, double kappa, int sign);

/* out = M_s in on even sites: sign=+1 is Dirac_Wilson_evenodd::apply (wilson.h:272-279),
 * sign=-1 is ::dagger (wilson.h:282-289).  in and out are even fields. */
int wl_meo_apply(wl_ctx *ctx, wl_field *out, const wl_field *in, double kappa, int sign);

/* out = (M_s + mh) in: Hasenbusch_operator<Dirac_Wilson_evenodd>::apply (s = +1) / ::dagger
 * (s = -1), libraries/dirac/Hasenbusch.h:40-48, the operator D + mh of the Hasenbusch split of
 * the fermion determinant.  The shift is applied inside the second hopping pass of M, so this
 * costs the same as wl_meo_apply. */
int wl_hasenbusch_apply(wl_ctx *ctx, wl_field *out, const wl_field *in, double kappa, double mh,
                        int sign);

/* ---- inverter ------------------------------------------------------------------------- */

/* CG on M^dag M, HILA's CG<Dirac_Wilson_evenodd>(D, accuracy, maxiters).apply(b, x)
 * (conjugate_gradient.h:33-104): x is the initial guess on entry and the solution on exit;
 * stops when ||r||^2 < accuracy^2 * ||b||^2.  FP64 throughout.  With nrhs > 1 the right-hand
 * sides are solved together (batched CG): each keeps its own recurrence and stopping test and
 * all share every operator application.  stats (may be NULL) points to nrhs entries.  Returns
 * WL_E_NOCONV (with x and stats filled) if maxiters is reached first for any rhs. */
int wl_cg_solve(wl_ctx *ctx, wl_field *x, const wl_field *b, double kappa, double accuracy,
                int maxiters, wl_cg_stats *stats);

/* wl_cg_solve for one right-hand side that allocates nothing: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (tight solve loops,
 * harnesses that forbid allocation during a run).  The same FP64 CG and kernels as
 * wl_cg_solve, with the same result bit for bit, but its work memory (Krylov fields, the odd
 * temporary, partial sums, CG state) is one fixed device array that the library holds on
 * every device it runs on (30 MiB, until the process ends).  Limit: lattices whose parity
 * block has at most 2^15 sites after rounding up to 256 (e.g. 16^4); larger ones return
 * WL_E_ARG.  Needs the gauge field uploaded and, with WL_OPT_LINK_ROWS = 2, the FP64 two-row
 * copy already made by an earlier operator or solver call (WL_E_STATE otherwise).  Calls on
 * one device share the array: within a process they are serialised by a lock; calls from
 * several processes each use their own copy. */
int wl_cg_solve_noalloc(wl_ctx *ctx, wl_field *x, const wl_field *b, double kappa, double accuracy,
                        int maxiters, wl_cg_stats *stats);

/* Mixed-precision CG on M^dag M with reliable updates (an addition: HILA's only mixed
 * precision is a full FP32 CG used as the initial guess of an FP64 CG,
 * libraries/hmc/fermion_field.h:65-86).  The operator and the Krylov vectors r, p, Dp are in
 * inner_prec (WL_F32; WL_F64 runs the same algorithm in FP64); the solution x and every dot /
 * norm reduction are FP64.  A reliable update -- r = b - M^dag M x with the FP64 operator --
 * is made when ||r|| has dropped by the factor delta since the last one, every update_every
 * iterations, and whenever the iterated residual meets the target; convergence is declared
 * only on that FP64 true residual: ||b - M^dag M x|| < accuracy ||b||.  x is the initial guess
 * on entry.  Same return convention as wl_cg_solve. */
int wl_cg_solve_mixed(wl_ctx *ctx, wl_field *x, const wl_field *b, double kappa, double accuracy,
                      int maxiters, int inner_prec, double delta, int update_every,
                      wl_cg_stats *stats);

/* All solver settings at once. */
typedef struct wl_cg_params {
    double kappa;
    double mh;           /* Hasenbusch shift: CG<Hasenbusch_operator<Dirac_Wilson_evenodd>>
                            solves (M + mh)^dag (M + mh) x = b (Hasenbusch.h:15-53); 0: plain M */
    double accuracy;
    int maxiters;
    int inner_prec;      /* WL_F64: FP64 CG as wl_cg_solve; WL_F32 / WL_F16: reliable updates
                            as wl_cg_solve_mixed                                               */
    double delta;        /* mixed only */
    int update_every;    /* mixed only */
} wl_cg_params;

/* CG<Op>::apply(b, x) for Op = Dirac_Wilson_evenodd or Hasenbusch_operator of it, with the
 * settings of *params; otherwise as wl_cg_solve / wl_cg_solve_mixed. */
int wl_cg_solve_ex(wl_ctx *ctx, wl_field *x, const wl_field *b, const wl_cg_params *params,
                   wl_cg_stats *stats);

/* MRE_guess (libraries/hmc/MRE_guess.h:8-87): initial guess psi for CG on
 * A = (M + mh)^dag (M + mh), A psi = chi, from previous solutions.  basis is an even FP64
 * field whose right-hand sides 0..nbasis-1 hold them (HILA's old_chi_inv); chi and psi are
 * even FP64 fields of one right-hand side.  psi = sum_i c_i basis_i with real c_i chosen as
 * HILA's algorithm intends: Gram-Schmidt on the basis under Re <a, b> (Vector::rdot), then
 * the Galerkin condition Re <basis_i, chi - A psi> = 0 for every kept basis vector, solved by
 * Gauss-Jordan elimination.  Zero and linearly dependent basis vectors (and directions on
 * which the projected A vanishes) are dropped, so the all-zero basis HILA starts with
 * (fermion_field.h:38-44) gives psi = 0.  Three defects of the source text are not copied:
 * its projection coefficient divides by the norm of the vector being orthogonalised instead
 * of the one projected out (:24-28), its final loop overwrites psi with basis_i (1 + c_i)
 * instead of accumulating c_i basis_i (:83-85), and save_new_solution's forward shift
 * (fermion_field.h:135-141) leaves every older slot holding the same solution.  1 <= nbasis <= 8, or 0
 * (psi = 0).  Cost: A on the nbasis vectors (one multi-rhs pass) plus two streaming passes. */
int wl_mre_guess(wl_ctx *ctx, wl_field *psi, const wl_field *chi, const wl_field *basis, int nbasis,
                 double kappa, double mh);

/* ---- fermion force and HMC momentum ------------------------------------------------------ */

/* Dirac_Wilson_evenodd::force(chi, psi, force, sign) (libraries/dirac/wilson.h:294-311, with
 * Dirac_Wilson_calc_force :108-137): the derivative of Re <chi, M_sign psi> with respect to
 * the links, F[mu] in HILA's layout of a Field<SquareMatrix<3,Complex<double>>> (V x 18
 * doubles, EVEN_SITES_FIRST, row-major): for every link, d/de Re <chi, M_sign(U + e E) psi>
 * = Re tr(E F_mu(x)) when U_mu(x) alone moves to U_mu(x) + e E.  chi and psi are even FP64
 * fields of one right-hand side.  The same for Hasenbusch_operator (its force is D's). */
int wl_fermion_force(wl_ctx *ctx, const wl_field *chi, const wl_field *psi, double kappa, int sign,
                     double *const F[4]);

/* The HMC momentum the context holds (gauge_field_base::momentum, hmc/gauge_field.h:62; one
 * 3x3 complex matrix per link), in HILA's layout as the gauge field.  Zero until uploaded. */
int wl_momentum_upload(wl_ctx *ctx, const double *const P[4]);
int wl_momentum_download(wl_ctx *ctx, double *const P[4]);

/* fermion_action::force_step(eps) (libraries/hmc/fermion_field.h:146-167) for the operator of
 * *params (Dirac_Wilson_evenodd, or its Hasenbusch_operator when params->mh != 0):
 *   psi = CG solution of (D^dag D) psi = chi, from the initial guess psi holds on entry
 *         (e.g. a wl_mre_guess); psi holds it on exit;
 *   force = -eps (D.force(D psi, psi, +1) + D.force(psi, D psi, -1));
 *   momentum += project_antihermitean(U force)  (gauge_field::add_momentum, gauge_field.h:249-257),
 * where project_antihermitean(M) = (M - M^dag)/2 - tr((M - M^dag)/2)/3.  The force and the
 * momentum update are one pass over the links.  Returns WL_E_NOCONV (after completing the
 * step, as HILA's CG returns silently) when CG ran out of iterations; stats as wl_cg_solve. */
int wl_fermion_force_step(wl_ctx *ctx, const wl_field *chi, wl_field *psi, double eps,
                          const wl_cg_params *params, wl_cg_stats *stats);

/* Hasenbusch_action_2::force_step(eps) (libraries/hmc/fermion_field.h:319-346), the second
 * term of the Hasenbusch split, action <chi, D_h (D^dag D)^-1 D_h^dag chi> with
 * D_h = D + params->mh:
 *   psi = CG solution of (D^dag D) psi = D_h^dag chi (unshifted D; psi holds the initial
 *         guess on entry and the solution on exit);
 *   Mpsi = D psi - chi;
 *   momentum += project_antihermitean(U (-eps)(D.force(Mpsi, psi, +1) + D.force(psi, Mpsi, -1))).
 * Returns as wl_fermion_force_step. */
int wl_hasenbusch2_force_step(wl_ctx *ctx, const wl_field *chi, wl_field *psi, double eps,
                              const wl_cg_params *params, wl_cg_stats *stats);

/* ---- kernel options ------------------------------------------------------------------- */

/* Tuning knobs of the hopping kernel (no counterpart in HILA, whose GPU loops are generated
 * with fixed N_threads = 256, libraries/plumbing/params.h:248):
 *   WL_OPT_TRAVERSAL_CHUNK  order in which the 64-site segments of the checkerboard are
 *       visited: 0 = memory order; n > 0 = chunks of n segments of one time slice, visited for
 *       every t before the next chunk (n must divide the segments per slice); -1 = default.
 *   WL_OPT_BLOCK  workgroup size 64, 128 or 256; 0 = default for the volume (multi-rhs FP32
 *       and FP16 launches of two or more rhs: 256).
 *   WL_OPT_RHS_BLOCK  right-hand sides per lane of the multi-rhs kernel (register blocking,
 *       1..4, at most 2 in FP64; 0 = default: for FP16 storage 1, or 2 when rhs x sites per
 *       parity reaches 3 x 2^20; for FP32 2 from 4 rhs on lattices of 2^17 sites per parity or
 *       more (from 2 rhs at 2^19 or more), else 1; for FP64 2, except 1 for up to 3 rhs below
 *       2^17 sites per parity).  A lane loads each link once for them.
 *   WL_OPT_RHS_LAUNCH  right-hand sides per launch, i.e. per pass over the lattice (rounded
 *       to a multiple of the per-lane count; 0 = default).  Fewer per pass keep the neighbour
 *       spinors of a pass in cache; each pass reads the links again.
 *   WL_OPT_PLANE_PAD  sites of padding added to the block stride of the device layout (a
 *       multiple of 64, 0..65536; -1 = default: 128 when the unpadded stride is a multiple of
 *       2^17 sites, as at 32^4, else 0), so that the rhs blocks of a field do not sit at large
 *       power-of-two distances.  Moves addresses only.  Only before the gauge field or any
 *       field of the context exists (else WL_E_STATE).
 *   WL_OPT_LINK_LOAD  link loads of the multi-rhs kernel: 0 = automatic (nontemporal when a
 *       launch reads each link once), 1 = cached, 2 = nontemporal, 3 = staged in LDS: one
 *       workgroup per 64-site segment whose waves (at most 8) are the rhs chunks share one read
 *       of the segment's links.
 *   WL_OPT_LDS_RESERVE  bytes of LDS reserved by every hopping workgroup (1..65536; -1 =
 *       none; 0 = automatic, the default: none, except multi-rhs FP64 launches of four or
 *       more rhs, or of two or more on lattices of 2^17 sites per parity or more, which reserve
 *       32 KB per 64 lanes (at most 5 workgroups per CU), and FP32 launches of four or more
 *       rhs at 3-4 per lane on lattices of 2^17 sites per parity or more, 24 KB (6); both,
 *       unless WL_OPT_BLOCK is set, in 64-lane workgroups).  The
 *       kernels use no LDS; the reservation caps how many workgroups share a CU, i.e. the
 *       sites in flight, whose neighbour spinors must fit the caches to be reused.
 *   WL_OPT_GRAPHS  1: the batched CG (nrhs > 1, FP16 inner, the server) captures each chunk
 *       of iterations into a HIP graph once per solve and replays it; 0 (default): launches
 *       kernel by kernel.  Results are identical either way.
 *   WL_OPT_PIPELINE  multi-rhs kernel: 0 (default) = software-pipelined hops where the
 *       registers allow (FP32 and FP16 storage, up to 2 rhs per lane); 1 = never.
 * Results do not depend on these options; only the speed does.
 *   WL_OPT_LINK_ROWS  3 (default): the FP64 and FP32 hopping kernels read all 9 entries of
 *       every link.  2: they read rows 0 and 1 and rebuild row 2 as conj(row 0 x row 1), 96
 *       instead of 144 bytes per FP64 link (a third less link traffic).  Allowed only for
 *       special unitary links: the uploaded links are checked (when the option is set, and
 *       after every upload) and every stored third row must agree with the rebuilt one to
 *       1e-13 in every real part, else WL_E_ARG.  Results then agree with the full-link ones
 *       to rounding, not bit for bit.  Two-row copies take 2/3 of the gauge field's memory per
 *       precision used.  FP16 storage always reads full links.  The fermion force reads the
 *       full links.
 *   WL_OPT_CG_PERSISTENT  single-rhs FP64 CG (wl_cg_solve, wl_cg_solve_ex with FP64 inner):
 *       1 = the whole solve -- source norm, initial residual, every iteration and the final
 *       true residual -- runs as one cooperative launch whose workgroups meet at grid
 *       barriers, with no host round trip inside the solve; -1 = kernel by kernel, seven
 *       launches per iteration, with the host checking convergence every 8 iterations; 0
 *       (default) = kernel by kernel, and on lattices of up to 2^15 sites per parity with
 *       the two finish kernels of an iteration folded into the kernels after them (five
 *       launches).  On MI355X the grid barriers of the one-launch solve (an L2 write-back and
 *       invalidate per XCD each) cost more than the kernel boundaries they replace, at every
 *       lattice size.
 *       Both give the same solution, bit for bit: the persistent launch runs the same per-site
 *       code and sums every reduction in the same order.  Falls back to kernel by kernel when
 *       the launch does not fit on the device at once.  The one-launch solve always runs one
 *       lane per site: it equals the kernel-by-kernel solve bit for bit with
 *       WL_OPT_SMALL_HOP = -1.
 *   WL_OPT_SMALL_HOP  single-rhs hopping launches (wl_dslash_eo, wl_meo_apply and the
 *       single-rhs solvers): 1 = the small-lattice kernel, in which each of 4 waves computes
 *       the two hops of one direction for 64 sites (loads of both issued together: one memory
 *       round trip) and the partial sums are added in LDS; -1 = one lane per site, the 8 hops
 *       pipelined; 0 (default) = the small-lattice kernel up to 2^15 sites per parity, where a
 *       one-lane launch has too few waves to hide memory latency.  Results agree to rounding
 *       (the hops are summed in another order).
 *   WL_OPT_BLOCK_KRYLOV  CG on a field of 2..24 right-hand sides (wl_cg_solve, wl_cg_solve_mixed
 *       and wl_cg_solve_ex with an FP64 or FP32 inner precision): 1 = block CG, the rhs share
 *       one Krylov space (the residual block is kept orthonormal, R = Q C; dependent or zero
 *       directions are dropped from the block), every rhs ends at its FP64 true residual, and
 *       every stats entry reports the block's iteration count; 0 (default) = an independent
 *       CG recurrence per rhs sharing the operator applications (batched CG).  With 1, a field
 *       of more than 24 rhs is refused (WL_E_ARG) and an FP16 inner precision too
 *       (WL_E_NOTIMPL); one rhs is solved as without it.  The request server always solves
 *       batched. */
enum {
    WL_OPT_TRAVERSAL_CHUNK = 1,
    WL_OPT_BLOCK = 2,
    WL_OPT_RHS_BLOCK = 3,
    WL_OPT_RHS_LAUNCH = 4,
    WL_OPT_PLANE_PAD = 5,
    WL_OPT_LINK_LOAD = 6,
    WL_OPT_LDS_RESERVE = 7,
    WL_OPT_GRAPHS = 8,
    WL_OPT_PIPELINE = 9,
    WL_OPT_LINK_ROWS = 10,
    WL_OPT_CG_PERSISTENT = 11,
    WL_OPT_SMALL_HOP = 12,
    WL_OPT_BLOCK_KRYLOV = 13
};
int wl_set_option(wl_ctx *ctx, int option, int value);
int wl_get_option(const wl_ctx *ctx, int option, int *value);

/* Autotuner of the multi-rhs hopping kernel (no counterpart in HILA).  For nrhs right-hand
 * sides at precision prec it times the hopping term on scratch fields of this context for
 * every candidate of rhs per lane (1, 2, 4 in FP32; 1, 2 in FP64), rhs per launch (1, 2, 4, 6,
 * 12, ..., nrhs: the RHS block of one pass over the lattice) and workgroup size (64, 128, 256),
 * and sets WL_OPT_RHS_BLOCK, WL_OPT_RHS_LAUNCH and WL_OPT_BLOCK to the fastest (in FP64 and
 * FP32 also WL_OPT_LDS_RESERVE: automatic, i.e. the occupancy cap where it applies, or none).
 * The choice is timed again against the default configuration, which is kept unless the choice
 * is faster.  The choice is
 * remembered per (lattice size, precision, nrhs, link format WL_OPT_LINK_ROWS, WL_OPT_LINK_LOAD,
 * WL_OPT_PIPELINE) for the life of the process, so that later calls, and other contexts of the
 * same size and options, take the same configuration without timing again (results then do
 * not depend on timing noise).  The gauge field must be uploaded. */
typedef struct wl_tune_result {
    int rhs_block;            /* 0 in all three: the default configuration was fastest      */
    int rhs_launch;
    int block;
    int candidates;           /* configurations timed (0 when the choice was remembered)     */
    double us_per_rhs;        /* hopping term of the chosen configuration, per rhs           */
    double us_per_rhs_default;/* ... and of the default configuration                        */
    int lds_reserve;          /* WL_OPT_LDS_RESERVE of the choice (0 automatic, -1 none)      */
} wl_tune_result;
int wl_autotune(wl_ctx *ctx, int prec, int nrhs, wl_tune_result *out);

/* ---- diagnostics ---------------------------------------------------------------------- */

/* Average device time per call of `iters` back-to-back calls (after `warmup` untimed ones),
 * measured with HIP events recorded on the context's stream, around the hopping term
 * (what = 0, as wl_dslash_eo), M (what = 1, as wl_meo_apply) or a plain copy out = in of two
 * fields of the context (what = 2, as wl_field_copy; kappa and sign unused), the device-copy
 * bandwidth the hopping term is compared with.  No counterpart in HILA
 * (its benchmark times the host loop, applications/benchmarks/bench_fermion.cpp:100-113). */
int wl_time_op(wl_ctx *ctx, int what, wl_field *out, const wl_field *in, double kappa, int sign,
               int warmup, int iters, double *ms_per_launch);

/* The packed FP16 projection's 16 operand forms on device `device`, for testing them bit for
 * bit.  No counterpart in HILA.  a[i], b[i] are complex numbers as two IEEE halves (re in the
 * low 16 bits); out[f * n + i] (16 n entries) is the form f = 4 k + F of a + C b with C = +1,
 * -1, +i, -i for k = 0..3, returned as (re, im), (im, re), (re, -im) or (im, -re) for F = 0..3.
 * Each half of a result is one FP16 addition of the selected, possibly negated, input halves
 * (the negation is applied to the operands, so -(x + y) is computed as (-x) + (-y)). */
int wl_fp16_forms(int device, const uint32_t *a, const uint32_t *b, int n, uint32_t *out);

/* ---- request server: continuous batching of CG solves --------------------------------- */

/* A server on one device runs CG solves of independent requests (each with its own lattice
 * size, gauge field, kappa, tolerance and 1..64 right-hand sides) concurrently.  No
 * counterpart in HILA, whose CG<Op>::apply is a blocking call on one operator
 * (conjugate_gradient.h:46-104); each request is solved as CG<Dirac_Wilson_evenodd> would on
 * its own (same operator, same stopping rule per rhs), batched over its right-hand sides.
 *
 * Admission: a request is admitted when a solve slot is free (at most max_active run at
 * once) and the device memory of a slot for its lattice fits in the HBM budget.  Slots (a
 * context with the gauge field, the rhs/solution fields and the CG work fields of a lattice
 * size) stay resident after a request and are reused by the next request of that size; idle
 * slots of other sizes are released when the budget requires it.  Order: the next request
 * admitted is the one with the least estimated work (sites x rhs x log10(1 / tolerance))
 * among the first 16 waiting, the earlier one on ties; a request passed over 8 times (later
 * requests admitted before it) goes next instead.  When the chosen request needs memory that
 * running requests hold, another of the next 8 that can start at once (an idle slot of its
 * lattice, or room for a new slot without releasing any) goes first, until the chosen one has
 * been passed over 8 times; then the queue waits for it.  The order changes queue times only:
 * a request's solution does not depend on it.  The solves of admitted requests overlap on the
 * device: each slot has its own stream, and the host drives all of them without waiting
 * inside a solve. */
typedef struct wl_server wl_server;

typedef struct wl_solve_request {
    int dims[4];
    int antiperiodic_t;
    int nrhs;                    /* 1..64                                                    */
    const double *gauge[4];      /* U[mu]: V x 18 doubles, HILA layout (see top of file)      */
    const double *source;        /* nrhs consecutive blocks of V/2 x 24 doubles: the even-site
                                    half of HILA's AoS field (EVEN_SITES_FIRST)               */
    double *solution;            /* nrhs blocks as `source`: the solutions (initial guess 0)  */
    double kappa;
    double tolerance;            /* CG accuracy: ||r||^2 < tolerance^2 ||b||^2 per rhs       */
    int maxiters;
    int inner_prec;              /* WL_F64: FP64 CG; WL_F32: mixed precision, reliable updates */
    double delta;                /* mixed: update when ||r|| drops by delta (0 < delta < 1)   */
    int update_every;            /* mixed: ... or after this many iterations                  */
    const char *config_path;     /* if not null: the gauge field is read from this HILA
                                    configuration file (wl_config_read) and gauge[] is unused;
                                    the path is copied at submission                          */
    double mh;                   /* Hasenbusch shift: solve with M + mh (0: M)                */
} wl_solve_request;

typedef struct wl_solve_result {
    int status;                  /* WL_OK, WL_E_NOCONV, or the error that ended the request  */
    int nrhs;
    int iterations;              /* largest over the rhs                                     */
    int reliable_updates;        /* largest over the rhs                                     */
    double true_resid;           /* largest FP64 ||b - M^dag M x|| / ||b|| over the rhs       */
    double queue_ms;             /* submission -> admitted (slot assigned, inputs enqueued)  */
    double latency_ms;           /* submission -> solutions written to host memory           */
    char message[160];           /* error text when status is an error                        */
} wl_solve_result;

typedef struct wl_server_info {
    size_t budget_bytes;
    size_t bytes_in_use;         /* device memory of all slots                               */
    int slots;                   /* resident slots                                           */
    int active;                  /* requests being loaded, solved or written back            */
    int queued;                  /* submitted, not yet admitted                              */
    int max_active;
    int64_t completed;
    int64_t slots_created;
    int64_t slots_released;
} wl_server_info;

/* hbm_budget_bytes = 0: half of the device memory free at creation.  max_active <= 0: 6 (on
 * MI355X fewer solves at once gave both lower latency and more throughput than 8-16). */
int wl_server_create(int device, size_t hbm_budget_bytes, int max_active, wl_server **out);
/* Queue a request.  The descriptor is copied; the host buffers it points to (gauge, source,
 * solution) must stay valid and unmodified until the request completes.  Buffers from
 * wl_host_alloc are copied by DMA without staging. */
int wl_server_submit(wl_server *srv, const wl_solve_request *req, uint64_t *ticket);
/* Provision idle slots before traffic (no counterpart in HILA): up to `count` slots for
 * requests of lattice dims, the time boundary antiperiodic_t, inner precision inner_prec and
 * at most nrhs right-hand sides, each with all the device memory such a request uses (gauge
 * field and the inner precision's link copy, source and solution fields, CG work fields), so
 * that the requests that later take them allocate nothing.  Stops early, with WL_OK, when the
 * next slot would exceed the HBM budget; *created = the slots made.  Holds the server lock
 * while it runs (admissions wait), and freezes the slot options as a submission does.  Slots
 * with two-row links (WL_OPT_LINK_ROWS = 2) build their two-row copies at first use. */
int wl_server_provision(wl_server *srv, const int dims[4], int antiperiodic_t, int inner_prec, int nrhs,
                        int count, int *created);
/* Wait up to timeout_ms (< 0: no limit, 0: poll) for a request.  On completion fills *res and,
 * if per_rhs is not null, its nrhs entries, releases the ticket and returns WL_OK (the
 * request's own outcome is res->status); WL_E_BUSY if not finished in time. */
int wl_server_wait(wl_server *srv, uint64_t ticket, int timeout_ms, wl_solve_result *res,
                   wl_cg_stats *per_rhs);
int wl_server_info_get(wl_server *srv, wl_server_info *info);
/* A kernel option (WL_OPT_*, as wl_set_option) for every slot of the server; only before the
 * first submission (else WL_E_STATE).  The multi-rhs kernel shape (WL_OPT_RHS_BLOCK,
 * WL_OPT_RHS_LAUNCH, WL_OPT_BLOCK) is taken per request from wl_autotune's remembered choice
 * for its lattice size, precision and rhs count, or the defaults.  With WL_OPT_LINK_ROWS = 2 a
 * request whose links are not special unitary to 1e-13 completes with status WL_E_ARG. */
int wl_server_set_option(wl_server *srv, int option, int value);
/* Finish every submitted request, then release all device memory and threads. */
int wl_server_destroy(wl_server *srv);

/* Page-locked host memory for request buffers (hipHostMalloc). */
int wl_host_alloc(size_t bytes, void **out);
int wl_host_free(void *p);

/* ---- seeded synthetic inputs (host side, counter-based, independent of thread count) --- */

/* Gauge field written in HILA's layout into U[mu] (V x 18 doubles each).
 * kind 0: unit links; 1: Haar-random SU(3); 2: exp(i*eps*H), H traceless Hermitian with
 * unit-variance complex Gaussian entries, reunitarized as SU<N>::reunitarize
 * (libraries/datatypes/sun_matrix.h:137-197); 3: HILA's SU<3>::random() -- 16 hits of random
 * SU(2) matrices on the three row pairs of the identity (sun_matrix.h:211-240).  eps is read
 * by kind 2 only. */
int wl_gen_gauge(const int dims[4], int kind, double eps, uint64_t seed, double *const U[4]);
/* Gaussian Wilson-vector field: HILA's AoS layout, V x 24 doubles (full lattice, EVEN_SITES_
 * FIRST); sites of the parities in parity_mask (1 even, 2 odd, 3 all) get re, im ~ N(0,1)
 * (Complex::gaussian_random, libraries/datatypes/cmplx.h:344-349); the others get 0. */
int wl_gen_spinor(const int dims[4], uint64_t seed, int parity_mask, double *out);
/* Point source: 1 at component (spin, colour) of one site chosen from seed, 0 elsewhere
 * (V x 24 doubles, full lattice).  *site_out receives the EVEN_SITES_FIRST index. */
int wl_gen_point(const int dims[4], uint64_t seed, int spin, int colour, double *out,
                 int64_t *site_out);

#ifdef __cplusplus
}
#endif
#endif /* WILSON_MI355X_H */
