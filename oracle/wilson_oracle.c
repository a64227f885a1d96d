/*
 * wilson_oracle.c -- CPU restatement of HILA's Wilson even-odd operator and CG.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark use it; the product library never links,
 * loads or calls it.
 *
 * HILA itself cannot be built here (its loops need the hilapp source-to-source compiler, which
 * needs Clang libtooling, and the Wilson headers of this snapshot reference types that do not
 * exist), so this is a restatement of the reference algorithm, pass by pass:
 *
 *   geometry / layout   libraries/plumbing/lattice.cpp:171-186 (EVEN_SITES_FIRST site index),
 *                       :394-407 (x-fastest lexicographic order), coordinates.h:34-104
 *                       (Direction numbering, -d = 7-d), coordinates.h:423-431 (parity);
 *   field storage       backend_cpu/field_storage_backend.h:10-19 (AoS, one element per site);
 *   antiperiodic BC     field_comm.h:115-124, lattice.cpp:995-1023,
 *                       backend_cpu/field_storage_backend.h:86-112: a hop across the t edge in
 *                       either direction reads the negated element;
 *   gamma / projection  wilson_vector.h:439-491 (HalfWilsonVector(w, dir, sign)),
 *                       :493-567 (expand(dir, sign)), no 1/2 in the projectors;
 *   SU(3) algebra       matrix.h:1118-1137 (adjoint), :2599-2622 (matrix x vector),
 *                       cmplx.h:761-763 (complex multiply);
 *   hopping term        wilson.h:15-50 (Dirac_Wilson_hop) and :53-95 (Dirac_Wilson_hop_set):
 *                       a projection pass over the opposite parity filling the 8 half-vector
 *                       temporaries vtemp[+dir] = P^{sign}_dir in and
 *                       vtemp[-dir] = U_dir^dag P^{-sign}_dir in, then one accumulation pass
 *                       per direction;
 *   even-odd operator   wilson.h:272-289 (apply / dagger), :98-106 (diag = copy, no clover);
 *   CG                  conjugate_gradient.h:46-104, literally.
 *
 * One deliberate difference from the source text: the backward hop reads
 * vtemp[-dir][X - dir] (the temporary gathered from -dir, which holds U^dag P psi of the site
 * x - dir).  wilson.h:47 and :83,:92 read vtemp[dir][X + dir] there, which leaves vtemp[-dir]
 * unused and breaks the adjointness that HILA's own test_CG.cpp:236-245 asserts.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NDIM 4
#define NDIRS 8
#define NC 3

/* Mutation-test builds only: -DORACLE_PLANT=n compiles one deliberate mistake into the
 * oracle so that the tests can show their pins reject it (tests/test_oracle_mutations.py).
 * 1: the backward hop as written in wilson.h:47 (reads vtemp[dir][X+dir]);
 * 2: no antiperiodic sign;  3: sign error in the e_z projection;
 * 4: U instead of U^dag on the backward hop;  5: gamma_t rows swapped in expand. */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

/* Parity as HILA's enum class Parity {none, even, odd, all} (coordinates.h:182). */
enum { P_NONE = 0, P_EVEN = 1, P_ODD = 2, P_ALL = 3 };

typedef struct {
    double re, im;
} cplx;

typedef struct {
    int L[NDIM];
    int64_t V, Vh;
    int antiperiodic_t;
    int *coord;          /* [V][4] */
    int64_t *nb[NDIRS];  /* neighbour site index in Direction d */
    signed char *sg[NDIRS]; /* -1 where the hop crosses the antiperiodic t edge, else +1 */
} olat;

/* Wilson vector: c[spin][colour]; half vector: c[2][colour] (wilson_vector.h:56-63,425-431). */
typedef struct {
    cplx c[4][NC];
} wvec;
typedef struct {
    cplx c[2][NC];
} hvec;
typedef struct {
    cplx e[NC][NC];
} su3;

static inline cplx cadd(cplx a, cplx b) { return (cplx){a.re + b.re, a.im + b.im}; }
static inline cplx csub(cplx a, cplx b) { return (cplx){a.re - b.re, a.im - b.im}; }
/* cmplx.h:761-763 */
static inline cplx cmul(cplx a, cplx b) {
    return (cplx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
static inline cplx cconj(cplx a) { return (cplx){a.re, -a.im}; }
/* I * a and -I * a (cmplx.h:1154-1175) */
static inline cplx ci(cplx a) { return (cplx){-a.im, a.re}; }
static inline cplx cmi(cplx a) { return (cplx){a.im, -a.re}; }
static inline cplx cneg(cplx a) { return (cplx){-a.re, -a.im}; }
static inline cplx cscale(double s, cplx a) { return (cplx){s * a.re, s * a.im}; }

/* ---------------------------------------------------------------- geometry (lattice.cpp) */

static int parity_of(const int *x) { return ((x[0] + x[1] + x[2] + x[3]) & 1) ? P_ODD : P_EVEN; }

/* lattice.cpp:171-186 with get_logical_index :394-407 */
static int64_t site_index(const olat *l, const int *x) {
    int64_t i = x[0] + (int64_t)l->L[0] * (x[1] + (int64_t)l->L[1] * (x[2] + (int64_t)l->L[2] * x[3]));
    return parity_of(x) == P_EVEN ? i / 2 : i / 2 + l->Vh;
}

olat *olat_create(const int L[4], int antiperiodic_t) {
    olat *l = (olat *)calloc(1, sizeof(olat));
    l->V = 1;
    for (int d = 0; d < NDIM; ++d) {
        if (L[d] < 2 || (L[d] & 1)) {
            free(l);
            return NULL;
        }
        l->L[d] = L[d];
        l->V *= L[d];
    }
    l->Vh = l->V / 2;
    l->antiperiodic_t = antiperiodic_t;
    l->coord = (int *)malloc(sizeof(int) * 4 * l->V);
    for (int d = 0; d < NDIRS; ++d) {
        l->nb[d] = (int64_t *)malloc(sizeof(int64_t) * l->V);
        l->sg[d] = (signed char *)malloc(l->V);
    }
    int x[4];
    for (int64_t lex = 0; lex < l->V; ++lex) {
        int64_t r = lex;
        for (int d = 0; d < NDIM; ++d) {
            x[d] = (int)(r % L[d]);
            r /= L[d];
        }
        const int64_t s = site_index(l, x);
        memcpy(l->coord + 4 * s, x, sizeof(x));
    }
    /* neighbours: Direction d < 4 is +e_d, d >= 4 is -e_{7-d} (coordinates.h:34-44,82-104) */
    for (int64_t s = 0; s < l->V; ++s) {
        for (int d = 0; d < NDIRS; ++d) {
            const int up = d < NDIM;
            const int mu = up ? d : NDIRS - 1 - d;
            int y[4];
            memcpy(y, l->coord + 4 * s, sizeof(y));
            int cross = 0;
            if (up) {
                if (++y[mu] == L[mu]) {
                    y[mu] = 0;
                    cross = 1;
                }
            } else {
                if (--y[mu] < 0) {
                    y[mu] = L[mu] - 1;
                    cross = 1;
                }
            }
            l->nb[d][s] = site_index(l, y);
            l->sg[d][s] = (cross && mu == 3 && antiperiodic_t && ORACLE_PLANT != 2) ? -1 : 1;
        }
    }
    return l;
}

void olat_destroy(olat *l) {
    if (!l) return;
    free(l->coord);
    for (int d = 0; d < NDIRS; ++d) {
        free(l->nb[d]);
        free(l->sg[d]);
    }
    free(l);
}

int64_t olat_volume(const olat *l) { return l->V; }
int64_t olat_neighbour(const olat *l, int dir, int64_t s) { return l->nb[dir][s]; }
void olat_coords(const olat *l, int64_t s, int *x) { memcpy(x, l->coord + 4 * s, 4 * sizeof(int)); }

static inline int site_parity(const olat *l, int64_t s) { return s < l->Vh ? P_EVEN : P_ODD; }
static inline int in_parity(const olat *l, int64_t s, int par) {
    return par == P_ALL || site_parity(l, s) == par;
}
static inline int opp_parity(int par) {
    return par == P_EVEN ? P_ODD : (par == P_ODD ? P_EVEN : par);
}

/* ----------------------------------------------------- half Wilson vector (wilson_vector.h) */

/* HalfWilsonVector(w, dir, sign), wilson_vector.h:439-491 */
static hvec half_project(const wvec *w, int dir, int sign) {
    hvec h;
    for (int a = 0; a < NC; ++a) {
        const cplx w0 = w->c[0][a], w1 = w->c[1][a], w2 = w->c[2][a], w3 = w->c[3][a];
        switch (dir) {
        case 0: /* e_x */
            h.c[0][a] = sign == 1 ? cadd(w0, ci(w3)) : csub(w0, ci(w3));
            h.c[1][a] = sign == 1 ? cadd(w1, ci(w2)) : csub(w1, ci(w2));
            break;
        case 1: /* e_y */
            h.c[0][a] = sign == 1 ? csub(w0, w3) : cadd(w0, w3);
            h.c[1][a] = sign == 1 ? cadd(w1, w2) : csub(w1, w2);
            break;
        case 2: /* e_z */
            h.c[0][a] = (sign == 1) != (ORACLE_PLANT == 3) ? cadd(w0, ci(w2)) : csub(w0, ci(w2));
            h.c[1][a] = sign == 1 ? csub(w1, ci(w3)) : cadd(w1, ci(w3));
            break;
        default: /* e_t */
            h.c[0][a] = sign == 1 ? cadd(w0, w2) : csub(w0, w2);
            h.c[1][a] = sign == 1 ? cadd(w1, w3) : csub(w1, w3);
            break;
        }
    }
    return h;
}

/* HalfWilsonVector::expand(dir, sign), wilson_vector.h:493-567 */
static wvec half_expand(const hvec *h, int dir, int sign) {
    wvec r;
    for (int a = 0; a < NC; ++a) {
        const cplx c0 = h->c[0][a], c1 = h->c[1][a];
        r.c[0][a] = c0;
        r.c[1][a] = c1;
        switch (dir) {
        case 0:
            r.c[2][a] = sign == 1 ? cmi(c1) : ci(c1);
            r.c[3][a] = sign == 1 ? cmi(c0) : ci(c0);
            break;
        case 1:
            r.c[2][a] = sign == 1 ? c1 : cneg(c1);
            r.c[3][a] = sign == 1 ? cneg(c0) : c0;
            break;
        case 2:
            r.c[2][a] = sign == 1 ? cmi(c0) : ci(c0);
            r.c[3][a] = sign == 1 ? ci(c1) : cmi(c1);
            break;
        default:
            r.c[2][a] = sign == 1 ? c0 : cneg(c0);
            r.c[3][a] = sign == 1 ? c1 : cneg(c1);
            if (ORACLE_PLANT == 5) {
                const cplx tmp = r.c[2][a];
                r.c[2][a] = r.c[3][a];
                r.c[3][a] = tmp;
            }
            break;
        }
    }
    return r;
}

/* matrix * half vector: the matrix acts on each colour vector (matrix.h:2599-2622) */
static hvec su3_mul_h(const su3 *u, const hvec *h) {
    hvec r;
    for (int s = 0; s < 2; ++s)
        for (int i = 0; i < NC; ++i) {
            cplx acc = {0, 0};
            for (int j = 0; j < NC; ++j) acc = cadd(acc, cmul(u->e[i][j], h->c[s][j]));
            r.c[s][i] = acc;
        }
    return r;
}

/* adjoint (matrix.h:1118-1137) */
static su3 su3_adjoint(const su3 *u) {
    su3 r;
    for (int i = 0; i < NC; ++i)
        for (int j = 0; j < NC; ++j) r.e[i][j] = cconj(u->e[j][i]);
    return r;
}

static su3 su3_scale(double s, const su3 *u) {
    su3 r;
    for (int i = 0; i < NC; ++i)
        for (int j = 0; j < NC; ++j) r.e[i][j] = cscale(s, u->e[i][j]);
    return r;
}

static hvec h_scale(double s, const hvec *h) {
    hvec r;
    for (int q = 0; q < 2; ++q)
        for (int a = 0; a < NC; ++a) r.c[q][a] = cscale(s, h->c[q][a]);
    return r;
}

static inline const wvec *W(const double *f, int64_t s) { return (const wvec *)(f + 24 * s); }
static inline wvec *Wm(double *f, int64_t s) { return (wvec *)(f + 24 * s); }
static inline const su3 *Ulink(const double *U, int64_t V, int mu, int64_t s) {
    return (const su3 *)(U + 18 * ((int64_t)mu * V + s));
}

static void wsub(wvec *o, const wvec *a) {
    for (int q = 0; q < 4; ++q)
        for (int c = 0; c < NC; ++c) o->c[q][c] = csub(o->c[q][c], a->c[q][c]);
}

/* ----------------------------------------------------------- hopping term (wilson.h:15-95) */

/* The 8 temporaries wilson_dirac_temp_vector[2*NDIM] (wilson.h:11-12): vtemp[d] is a half
 * vector per site.  Filled on the opposite parity, read at the neighbour (with the BC sign
 * the gather of an antiperiodic field applies). */
typedef struct {
    hvec *v[NDIRS];
    int64_t n;
} vtemps;

static vtemps vt_alloc(int64_t V) {
    vtemps t;
    for (int d = 0; d < NDIRS; ++d) t.v[d] = (hvec *)malloc(sizeof(hvec) * V);
    t.n = V;
    return t;
}
static void vt_free(vtemps *t) {
    for (int d = 0; d < NDIRS; ++d) free(t->v[d]);
}

/* projection pass: wilson.h:27-40 / :65-78 */
static void hop_project(const olat *l, const double *U, const double *in, int par, int sign,
                        vtemps *vt) {
    const int opp = opp_parity(par);
    for (int dir = 0; dir < NDIM; ++dir) {
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < l->V; ++s) {
            if (!in_parity(l, s, opp)) continue;
            hvec h = half_project(W(in, s), dir, -sign);
            su3 ua = ORACLE_PLANT == 4 ? *Ulink(U, l->V, dir, s) : su3_adjoint(Ulink(U, l->V, dir, s));
            vt->v[NDIRS - 1 - dir][s] = su3_mul_h(&ua, &h); /* vtemp[-dir][X] */
            vt->v[dir][s] = half_project(W(in, s), dir, sign); /* vtemp[dir][X] */
        }
    }
}

/* one direction of the accumulation pass, wilson.h:43-49 (with the backward-hop fix):
 * out[X] = out[X] - (kappa*U_dir(X)*vtemp[dir][X+dir]).expand(dir, sign)
 *                 - (kappa*vtemp[-dir][X-dir]).expand(dir, -sign) */
static void hop_accumulate_dir(const olat *l, const double *U, double kappa, const vtemps *vt,
                               double *out, int par, int sign, int dir, int set) {
    const int bdir = NDIRS - 1 - dir;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < l->V; ++s) {
        if (!in_parity(l, s, par)) continue;
        const int64_t sf = l->nb[dir][s], sb = l->nb[bdir][s];
        const hvec hf = h_scale((double)l->sg[dir][s], &vt->v[dir][sf]);
        const hvec hb = ORACLE_PLANT == 1 ? h_scale((double)l->sg[dir][s], &vt->v[dir][sf])
                                          : h_scale((double)l->sg[bdir][s], &vt->v[bdir][sb]);
        const su3 ku = su3_scale(kappa, Ulink(U, l->V, dir, s));
        const hvec uf = su3_mul_h(&ku, &hf);
        const hvec kb = h_scale(kappa, &hb);
        const wvec ef = half_expand(&uf, dir, sign);
        const wvec eb = half_expand(&kb, dir, -sign);
        wvec *o = Wm(out, s);
        if (set) memset(o, 0, sizeof(wvec));
        wsub(o, &ef);
        wsub(o, &eb);
    }
}

/* Dirac_Wilson_hop (wilson.h:15-50): out_par -= kappa * H_sign in */
void oracle_hop(const olat *l, const double *U, double kappa, const double *in, double *out,
                int par, int sign) {
    vtemps vt = vt_alloc(l->V);
    hop_project(l, U, in, par, sign, &vt);
    for (int dir = 0; dir < NDIM; ++dir) hop_accumulate_dir(l, U, kappa, &vt, out, par, sign, dir, 0);
    vt_free(&vt);
}

/* Dirac_Wilson_hop_set (wilson.h:53-95): out_par = -kappa * H_sign in */
void oracle_hop_set(const olat *l, const double *U, double kappa, const double *in, double *out,
                    int par, int sign) {
    vtemps vt = vt_alloc(l->V);
    hop_project(l, U, in, par, sign, &vt);
    for (int dir = 0; dir < NDIM; ++dir)
        hop_accumulate_dir(l, U, kappa, &vt, out, par, sign, dir, dir == 0);
    vt_free(&vt);
}

static void copy_par(const olat *l, const double *in, double *out, int par) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < l->V; ++s)
        if (in_parity(l, s, par)) memcpy(out + 24 * s, in + 24 * s, 24 * sizeof(double));
}

static void zero_par(const olat *l, double *out, int par) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < l->V; ++s)
        if (in_parity(l, s, par)) memset(out + 24 * s, 0, 24 * sizeof(double));
}

/* Dirac_Wilson_evenodd::apply (sign=+1) / ::dagger (sign=-1), wilson.h:272-289.
 * in and out are full-lattice fields; in and out must be different buffers. */
void oracle_meo(const olat *l, const double *U, double kappa, const double *in, double *out,
                int sign) {
    copy_par(l, in, out, P_EVEN);                       /* Dirac_Wilson_diag(in,out,EVEN) */
    oracle_hop_set(l, U, kappa, in, out, P_ODD, sign);  /* hop_set(..., ODD, sign) */
    /* Dirac_Wilson_diag_inverse(out, ODD): identity */
    oracle_hop(l, U, -kappa, out, out, P_EVEN, sign);   /* hop(..., -kappa, out, out, EVEN) */
    zero_par(l, out, P_ODD);                            /* out[ODD] = 0 */
}

/* Hasenbusch_operator<Dirac_Wilson_evenodd>::apply / ::dagger (Hasenbusch.h:40-48):
 * D.apply(in, out) (or dagger), then out[D.par] = out[X] + h_parameter * in[X]. */
void oracle_hasenbusch(const olat *l, const double *U, double kappa, double mh, const double *in,
                       double *out, int sign) {
    oracle_meo(l, U, kappa, in, out, sign);
    if (mh == 0.0) return;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < l->Vh; ++s)
        for (int k = 0; k < 24; ++k) out[24 * s + k] = out[24 * s + k] + mh * in[24 * s + k];
}

/* squarenorm summed over the sites of par (conjugate_gradient.h:61,72,79,87) */
double oracle_norm2(const olat *l, const double *f, int par) {
    double sum = 0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
    for (int64_t s = 0; s < l->V; ++s) {
        if (!in_parity(l, s, par)) continue;
        const double *p = f + 24 * s;
        double q = 0;
        for (int i = 0; i < 24; ++i) q += p[i] * p[i];
        sum += q;
    }
    return sum;
}

/* CG<Dirac_Wilson_evenodd>::apply(in, out), conjugate_gradient.h:46-104.  out is the initial
 * guess on entry.  Returns the loop counter i at exit (HILA's printed step count); *rr_rel
 * receives rrnew / source_norm as printed by conjugate_gradient.h:103. */
int oracle_cg_h(const olat *l, const double *U, double kappa, double mh, const double *in,
                double *out, double accuracy, int maxiters, double *rr_rel);

int oracle_cg(const olat *l, const double *U, double kappa, const double *in, double *out,
              double accuracy, int maxiters, double *rr_rel) {
    return oracle_cg_h(l, U, kappa, 0.0, in, out, accuracy, maxiters, rr_rel);
}

/* CG<Hasenbusch_operator<Dirac_Wilson_evenodd>>::apply (Hasenbusch.h:15-53 with
 * conjugate_gradient.h:46-104): the same recurrence on (M + mh)^dag (M + mh); mh = 0 is
 * CG<Dirac_Wilson_evenodd>. */
int oracle_cg_h(const olat *l, const double *U, double kappa, double mh, const double *in,
                double *out, double accuracy, int maxiters, double *rr_rel) {
    const int64_t n = 24 * l->V;
    double *r = (double *)calloc(n, sizeof(double));
    double *p = (double *)calloc(n, sizeof(double));
    double *Dp = (double *)calloc(n, sizeof(double));
    double *DDp = (double *)calloc(n, sizeof(double));
    const int par = P_EVEN; /* M.par */
    double pDp = 0, rr = 0, rrnew = 0;
    double alpha, beta, target_rr, source_norm = 0;
    int i;

    source_norm = oracle_norm2(l, in, par);
    target_rr = accuracy * accuracy * source_norm;

    oracle_hasenbusch(l, U, kappa, mh, out, Dp, +1);
    oracle_hasenbusch(l, U, kappa, mh, Dp, DDp, -1);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < l->Vh; ++s)
        for (int k = 0; k < 24; ++k) {
            r[24 * s + k] = in[24 * s + k] - DDp[24 * s + k];
            p[24 * s + k] = r[24 * s + k];
        }
    rr = oracle_norm2(l, r, par);

    for (i = 0; i < maxiters; i++) {
        pDp = rrnew = 0;
        oracle_hasenbusch(l, U, kappa, mh, p, Dp, +1);
        oracle_hasenbusch(l, U, kappa, mh, Dp, DDp, -1);
        pDp = oracle_norm2(l, Dp, par);
        alpha = rr / pDp;
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < l->Vh; ++s)
            for (int k = 0; k < 24; ++k) {
                out[24 * s + k] = out[24 * s + k] + alpha * p[24 * s + k];
                r[24 * s + k] = r[24 * s + k] - alpha * DDp[24 * s + k];
            }
        rrnew = oracle_norm2(l, r, par);
        if (rrnew < target_rr) break;
        beta = rrnew / rr;
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < l->Vh; ++s)
            for (int k = 0; k < 24; ++k) p[24 * s + k] = beta * p[24 * s + k] + r[24 * s + k];
        rr = rrnew;
    }
    if (rr_rel) *rr_rel = rrnew / source_norm;
    free(r);
    free(p);
    free(Dp);
    free(DDp);
    return i;
}

/* ------------------------------------------------------------- fermion force (wilson.h) */

/* a.outer_product(b) summed over spin (wilson_vector.h:229-235, matrix.h:1437-1448):
 * M(i, j) = sum_s a[s][i] conj(b[s][j]), scaled by f */
static su3 wouter(const wvec *a, const wvec *b, double f) {
    su3 m;
    for (int i = 0; i < NC; ++i)
        for (int j = 0; j < NC; ++j) {
            cplx acc = {0, 0};
            for (int q = 0; q < 4; ++q) acc = cadd(acc, cmul(a->c[q][i], cconj(b->c[q][j])));
            m.e[i][j] = cscale(f, acc);
        }
    return m;
}

/* Dirac_Wilson_calc_force (wilson.h:108-137).  For every direction, with
 * vtemp[0] = HalfWilsonVector(chi, dir, -sign) on opp(par) and vtemp[1] =
 * HalfWilsonVector(psi, dir, sign) on par, both gathered with chi's boundary condition:
 *   out[dir][par] = -kappa * vtemp[0][X+dir].expand(dir, -sign).outer_product(psi[X])
 *   out[dir][opp] = -kappa * vtemp[1][X+dir].expand(dir,  sign).outer_product(chi[X])
 * out: 4 x V x 18 doubles (SquareMatrix<3, Complex<double>>, row-major), every site set. */
void oracle_calc_force(const olat *l, double kappa, const double *chi, const double *psi, double *out,
                       int par, int sign) {
    for (int dir = 0; dir < NDIM; ++dir) {
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < l->V; ++s) {
            const int64_t sn = l->nb[dir][s];
            const double g = (double)l->sg[dir][s];
            su3 m;
            if (in_parity(l, s, par)) {
                hvec h = half_project(W(chi, sn), dir, -sign);
                h = h_scale(g, &h);
                const wvec e = half_expand(&h, dir, -sign);
                m = wouter(&e, W(psi, s), -kappa);
            } else {
                hvec h = half_project(W(psi, sn), dir, sign);
                h = h_scale(g, &h);
                const wvec e = half_expand(&h, dir, sign);
                m = wouter(&e, W(chi, s), -kappa);
            }
            memcpy(out + 18 * ((int64_t)dir * l->V + s), &m, sizeof(su3));
        }
    }
}

/* Dirac_Wilson_evenodd::force(chi, psi, force, sign) (wilson.h:294-311): chi, psi full fields
 * (their even halves are used); out as oracle_calc_force. */
void oracle_force(const olat *l, const double *U, double kappa, const double *chi, const double *psi,
                  double *out, int sign) {
    const int64_t n = 24 * l->V;
    double *tmp = (double *)calloc(n, sizeof(double));
    double *f2 = (double *)malloc(sizeof(double) * 18 * 4 * l->V);
    oracle_hop_set(l, U, kappa, chi, tmp, P_ODD, -sign);
    oracle_calc_force(l, -kappa, tmp, psi, out, P_EVEN, sign);
    memset(tmp, 0, n * sizeof(double));
    oracle_hop_set(l, U, kappa, psi, tmp, P_ODD, sign);
    oracle_calc_force(l, -kappa, chi, tmp, f2, P_ODD, sign);
    const int64_t m = 18 * 4 * l->V;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) out[i] = out[i] + f2[i];
    free(tmp);
    free(f2);
}

/* ---------------------------------------------------------- fused CPU variant (baseline) */

/* The same D_oe / D_eo half operator written as one pass per output site (gather, project,
 * multiply, expand): out_par = -kappa * H_sign in.  Reported separately from the HILA-structured
 * oracle_hop_set as the "fused CPU" figure; it is checked against oracle_hop_set in the tests. */
void oracle_hop_set_fused(const olat *l, const double *U, double kappa, const double *in,
                          double *out, int par, int sign) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < l->V; ++s) {
        if (!in_parity(l, s, par)) continue;
        wvec acc;
        memset(&acc, 0, sizeof(acc));
        for (int dir = 0; dir < NDIM; ++dir) {
            const int bdir = NDIRS - 1 - dir;
            const int64_t sf = l->nb[dir][s], sb = l->nb[bdir][s];
            hvec hf = half_project(W(in, sf), dir, sign);
            hf = h_scale((double)l->sg[dir][s], &hf);
            hvec hb = half_project(W(in, sb), dir, -sign);
            hb = h_scale((double)l->sg[bdir][s], &hb);
            const hvec uf = su3_mul_h(Ulink(U, l->V, dir, s), &hf);
            const su3 ua = su3_adjoint(Ulink(U, l->V, dir, sb));
            const hvec ub = su3_mul_h(&ua, &hb);
            const wvec ef = half_expand(&uf, dir, sign);
            const wvec eb = half_expand(&ub, dir, -sign);
            for (int q = 0; q < 4; ++q)
                for (int c = 0; c < NC; ++c)
                    acc.c[q][c] = cadd(acc.c[q][c], cadd(ef.c[q][c], eb.c[q][c]));
        }
        wvec *o = Wm(out, s);
        for (int q = 0; q < 4; ++q)
            for (int c = 0; c < NC; ++c) o->c[q][c] = cscale(-kappa, acc.c[q][c]);
    }
}

/* Threads of the OpenMP loops from now on (n > 0), for single-core timings; returns the count
 * before the call. */
int oracle_set_threads(int n) {
#ifdef _OPENMP
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
#else
    (void)n;
    return 1;
#endif
}
