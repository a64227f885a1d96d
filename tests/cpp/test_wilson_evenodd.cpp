// The Wilson even-odd checks of HILA's applications/test_cases/test_CG.cpp:215-288, written
// against the C++ mirror (hila_wilson.hpp) of Dirac_Wilson_evenodd and CG: D / D^dag
// adjointness, then D (DdgD)^-1 Ddg = 1, DdgD (DdgD)^-1 = 1 and (DdgD)^-1 DdgD = 1, for unit
// links (as test_CG.cpp:62-65) and for Haar-random links.  Prints PASS and returns 0 when all
// hold.  The 1e-16 bounds of test_CG.cpp on squared differences are applied relative to the
// squared norm of the vector compared.
#include <cmath>
#include <complex>
#include <cstdio>
#include <type_traits>
#include <unistd.h>
#include <vector>

#include "hila_wilson.hpp"

using namespace hila_wilson;

static int failures = 0;
#define CHECK(cond, what)                                                            \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::printf("FAIL: %s (%s:%d)\n", what, __FILE__, __LINE__);             \
            ++failures;                                                              \
        } else {                                                                     \
            std::printf("ok:   %s\n", what);                                         \
        }                                                                            \
    } while (0)

// sum over even sites of a[X].dot(b[X]) (matrix.h:1395-1403: conj(a) . b)
static std::complex<double> dot_even(const WilsonField &a, const WilsonField &b) {
    const std::vector<double> ha = a.host(), hb = b.host();
    const size_t n = ha.size() / 2;  // even sites come first
    std::complex<double> s = 0;
    for (size_t i = 0; i < n; i += 2)
        s += std::conj(std::complex<double>(ha[i], ha[i + 1])) * std::complex<double>(hb[i], hb[i + 1]);
    return s;
}

static double diff2_even(const WilsonField &a, const WilsonField &b) {
    const std::vector<double> ha = a.host(), hb = b.host();
    double s = 0;
    for (size_t i = 0; i < ha.size() / 2; ++i) s += (ha[i] - hb[i]) * (ha[i] - hb[i]);
    return s;
}

static void run(Lattice &lattice, int link_kind, const char *name) {
    std::printf("Checking with Dirac_Wilson_evenodd, %s links\n", name);
    GaugeField U(lattice);
    U.random(link_kind, 0.0, 7);
    using dirac = Dirac_Wilson_evenodd<SU3d>;
    dirac D(0.12, U);
    WilsonField a(lattice), b(lattice), Db(lattice), Ddaggera(lattice), DdaggerDb(lattice);

    a.zero(ODD);
    b.zero(ODD);
    a.gaussian_random(11, EVEN);
    b.gaussian_random(12, EVEN);

    D.apply(b, Db);
    D.dagger(a, Ddaggera);
    const std::complex<double> d = dot_even(a, Db) - dot_even(Ddaggera, b);
    const double scale = std::abs(dot_even(a, Db));
    CHECK(std::norm(d) < 1e-26 * scale * scale, "test Dirac_Wilson_evenodd (adjoint)");

    // Now check that D (DdgD)^-1 Ddg = 1
    CG<dirac> inverse(D);
    a.zero(ALL);
    b.zero(ALL);
    a.gaussian_random(13, EVEN);
    D.dagger(a, Ddaggera);
    inverse.apply(Ddaggera, b);
    D.apply(b, Db);
    double diffre = diff2_even(a, Db);
    std::printf("D inv Dg %g (%d iterations)\n", diffre, inverse.stats.iterations);
    CHECK(diffre < 1e-16 * a.squarenorm(EVEN), "test D (DdgD)^-1 Ddg");

    // Run CG on a, multiply by DdaggerD and check the same
    a.gaussian_random(14, EVEN);
    b.zero(ALL);
    inverse.apply(a, b);
    D.apply(b, Db);
    D.dagger(Db, DdaggerDb);
    diffre = diff2_even(a, DdaggerDb);
    CHECK(diffre < 1e-16 * a.squarenorm(EVEN), "test DdgD (DdgD)^-1");

    // The other way around, multiply first
    b.gaussian_random(15, EVEN);
    D.apply(b, Db);
    D.dagger(Db, DdaggerDb);
    a.zero(ALL);
    inverse.apply(DdaggerDb, a);
    diffre = diff2_even(a, b);
    CHECK(diffre < 1e-16 * b.squarenorm(EVEN), "test (DdgD)^-1 DdgD");
    CHECK(inverse.stats.converged == 1 && inverse.stats.true_resid < 1e-11,
          "CG true residual below accuracy");

    // the same round trip with the mixed-precision inverter (FP32 and FP16 inner, FP64
    // reliable updates): the true residual meets the accuracy
    for (int prec : {WL_F32, WL_F16}) {
        CG<dirac> mixed(D, 1e-10, 10000);
        mixed.set_precision(prec, 0.1, 100);
        a.zero(ALL);
        mixed.apply(DdaggerDb, a);
        diffre = diff2_even(a, b);
        CHECK(mixed.stats.converged == 1 && mixed.stats.true_resid <= 1e-10,
              prec == WL_F32 ? "mixed CG (FP32 inner) true residual" : "mixed CG (FP16 inner) true residual");
        CHECK(diffre < 1e-14 * b.squarenorm(EVEN), "mixed (DdgD)^-1 DdgD");
    }

    // two-row links (special unitary links: unit and Haar): the same round trip, the operator
    // within 1e-13 of the full-link one; then back to full links
    {
        lattice.set_option(WL_OPT_LINK_ROWS, 2);
        CHECK(lattice.option(WL_OPT_LINK_ROWS) == 2, "two-row links on");
        WilsonField Db2(lattice);
        D.apply(b, Db2);
        CHECK(diff2_even(Db2, Db) < 1e-26 * Db.squarenorm(EVEN), "two-row operator = full-link operator");
        CG<dirac> rows(D, 1e-10, 10000);
        a.zero(ALL);
        rows.apply(DdaggerDb, a);
        diffre = diff2_even(a, b);
        CHECK(rows.stats.converged == 1 && rows.stats.true_resid <= 1e-10 && diffre < 1e-16 * b.squarenorm(EVEN),
              "CG with two-row links");
        lattice.set_option(WL_OPT_LINK_ROWS, 3);
    }

    // GaugeField::config_write then config_read (gaugefield.h:194-272, as checkpoint.h:30,89)
    // gives back the same links, and the operator on them the same output, to the last bit
    char path[] = "/tmp/wl_config_XXXXXX";
    const int fd = mkstemp(path);
    CHECK(fd >= 0, "temporary configuration file");
    if (fd >= 0) {
        close(fd);
        U.config_write(path);
        GaugeField U2(lattice);
        U2.config_read(path);
        std::remove(path);
        bool same = true;
        for (int mu = 0; mu < 4; ++mu) same = same && U2.direction(mu) == U.direction(mu);
        CHECK(same, "config_write / config_read round trip");
        U.upload();
        WilsonField Db1(lattice), Db2(lattice);
        D.apply(b, Db1);
        dirac D2(0.12, U2);
        U2.upload();
        D2.apply(b, Db2);
        CHECK(Db1.host() == Db2.host(), "operator on the configuration read back");
    }

    // apply leaves the odd half of the output at zero (wilson.h:278)
    Db.gaussian_random(16, ALL);
    D.apply(b, Db);
    CHECK(Db.squarenorm(ODD) == 0.0, "out[ODD] = 0");
}

// test_CG.cpp:290-345 (the Hasenbusch_operator block), then MRE_guess (MRE_guess.h) on a
// chronological sequence of solves: with the three previous solutions of right-hand sides
// quadratic in k, the guess for the next one leaves CG little to do.
static void run_hasenbusch(Lattice &lattice) {
    std::printf("Checking with Hasenbusch_operator\n");
    GaugeField U(lattice);
    U.random(1, 0.0, 9);
    using dirac_base = Dirac_Wilson_evenodd<SU3d>;
    dirac_base Dbase(0.1, U);
    using dirac = Hasenbusch_operator<dirac_base>;
    dirac D(Dbase, 0.1);
    WilsonField a(lattice), b(lattice), Db(lattice), Ddaggera(lattice), DdaggerDb(lattice);
    a.zero(ODD);
    b.zero(ODD);
    a.gaussian_random(21, EVEN);
    b.gaussian_random(22, EVEN);
    D.apply(b, Db);
    D.dagger(a, Ddaggera);
    const std::complex<double> d = dot_even(a, Db) - dot_even(Ddaggera, b);
    CHECK(std::norm(d) < 1e-26 * std::norm(dot_even(a, Db)), "test Hasenbusch_operator (adjoint)");
    // b = 1/D a -> Db = a
    CG<dirac> inverse(D);
    b.zero(ALL);
    inverse.apply(Ddaggera, b);
    D.apply(b, Db);
    CHECK(diff2_even(a, Db) < 1e-16 * a.squarenorm(EVEN), "test CG (Hasenbusch)");
    // DdgD (DdgD)^-1 a = a
    a.gaussian_random(23, EVEN);
    b.zero(ALL);
    inverse.apply(a, b);
    D.apply(b, Db);
    D.dagger(Db, DdaggerDb);
    CHECK(diff2_even(a, DdaggerDb) < 1e-16 * a.squarenorm(EVEN), "test DdgD (DdgD)^-1 (Hasenbusch)");

    // chronological MRE guesses: chi_k = c0 + k c1 + k^2 c2
    WilsonField c0(lattice), c1(lattice), c2(lattice);
    c0.gaussian_random(24, EVEN);
    c1.gaussian_random(25, EVEN);
    c2.gaussian_random(26, EVEN);
    const std::vector<double> h0 = c0.host(), h1 = c1.host(), h2 = c2.host();
    std::vector<WilsonField> old_chi_inv;
    int iters[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        std::vector<double> h(h0.size());
        for (size_t i = 0; i < h.size(); ++i) h[i] = h0[i] + k * h1[i] + k * k * h2[i];
        WilsonField chi(lattice), psi(lattice);
        chi.set_host(h);
        psi.zero(ALL);
        MRE_guess(psi, chi, D, old_chi_inv);
        CG<dirac> solve(D, 1e-10);
        solve.apply(chi, psi);
        iters[k] = solve.stats.iterations;
        old_chi_inv.insert(old_chi_inv.begin(), psi);  // save_new_solution, newest first
    }
    std::printf("      CG iterations from the MRE guess: %d %d %d %d\n", iters[0], iters[1], iters[2], iters[3]);
    CHECK(iters[3] * 4 <= iters[0], "MRE guess from three previous solutions");
}

// fermion_action::force_step (fermion_field.h:145-167) against the same step composed from
// its parts: psi from CG, Dpsi = D psi, F = D.force(Dpsi, psi, +1) and
// momentum += project_antihermitean(U (-2 eps) F) on the host; then a second step, whose MRE
// guess from the first solution (same chi) leaves CG almost nothing to do.
static void run_force(Lattice &lattice) {
    std::printf("Checking fermion_action::force_step\n");
    GaugeField U(lattice);
    U.random(2, 0.3, 31);
    using dirac = Dirac_Wilson_evenodd<SU3d>;
    dirac D(0.124, U);
    fermion_action<dirac> fa(D, U, 3);
    fa.draw_gaussian_fields(32);
    const int64_t V = lattice.volume();
    GaugeField::LinkField zero;
    for (auto &v : zero) v.assign(18 * V, 0.0);
    U.upload_momentum(zero);
    const double eps = 0.01;
    fa.force_step(eps);
    const int first_iters = fa.stats.iterations;
    const GaugeField::LinkField P1 = U.download_momentum();

    WilsonField psi(lattice), Dpsi(lattice);
    psi.zero(ALL);
    CG<dirac> inverse(D, 1e-12);
    inverse.apply(fa.chi, psi);
    D.apply(psi, Dpsi);
    GaugeField::LinkField F;
    D.force(Dpsi, psi, F, +1);
    double err = 0, ref = 0, ah = 0;
    for (int mu = 0; mu < 4; ++mu) {
        const std::vector<double> &u = U.direction(mu);
        for (int64_t x = 0; x < V; ++x) {
            std::complex<double> m[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    std::complex<double> a = 0;
                    for (int k = 0; k < 3; ++k)
                        a += std::complex<double>(u[18 * x + 2 * (3 * i + k)], u[18 * x + 2 * (3 * i + k) + 1]) *
                             std::complex<double>(F[mu][18 * x + 2 * (3 * k + j)], F[mu][18 * x + 2 * (3 * k + j) + 1]);
                    m[i][j] = -2.0 * eps * a;
                }
            std::complex<double> tr = 0;
            std::complex<double> p[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) p[i][j] = 0.5 * (m[i][j] - std::conj(m[j][i]));
            for (int i = 0; i < 3; ++i) tr += p[i][i];
            for (int i = 0; i < 3; ++i) p[i][i] -= tr / 3.0;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const std::complex<double> got(P1[mu][18 * x + 2 * (3 * i + j)], P1[mu][18 * x + 2 * (3 * i + j) + 1]);
                    err += std::norm(got - p[i][j]);
                    ref += std::norm(p[i][j]);
                    const std::complex<double> gt(P1[mu][18 * x + 2 * (3 * j + i)], P1[mu][18 * x + 2 * (3 * j + i) + 1]);
                    ah += std::norm(got + std::conj(gt));
                }
        }
    }
    CHECK(ref > 0 && err < 1e-24 * ref, "force_step = P(U (-eps)(F + F2)) composed from its parts");
    CHECK(ah < 1e-26 * ref, "momentum change anti-Hermitian");
    fa.force_step(eps);
    std::printf("      CG iterations: first step %d, from the MRE guess %d\n", first_iters, fa.stats.iterations);
    CHECK(fa.stats.iterations * 4 <= first_iters, "second step starts from the MRE guess");
    CHECK(fa.action() > 0, "action > 0");
}

// The two Hasenbusch terms (fermion_field.h:168-368): after draw_gaussian_fields, the second
// term's action is the squared norm of its Gaussian draw (chi = D_h (D_h^dag D_h)^-1 D^dag psi
// gives D (D^dag D)^-1 D_h^dag chi = psi); both force steps move the momentum within the algebra.
static void run_hasenbusch_actions(Lattice &lattice) {
    std::printf("Checking Hasenbusch_action_1 / Hasenbusch_action_2\n");
    GaugeField U(lattice);
    U.random(2, 0.4, 41);
    using dirac = Dirac_Wilson_evenodd<SU3d>;
    dirac D(0.124, U);
    Hasenbusch_action_1<dirac> fa1(D, U, 0.2, 2);
    Hasenbusch_action_2<dirac> fa2(D, U, 0.2, 2);
    fa1.draw_gaussian_fields(42);
    fa2.draw_gaussian_fields(43);
    WilsonField drawn(lattice);
    drawn.gaussian_random(43, EVEN);
    const double a2 = fa2.action(), n2 = drawn.squarenorm(EVEN);
    CHECK(std::abs(a2 - n2) < 1e-8 * n2, "Hasenbusch_action_2: action = |drawn psi|^2");
    const int64_t V = lattice.volume();
    GaugeField::LinkField zero;
    for (auto &v : zero) v.assign(18 * V, 0.0);
    U.upload_momentum(zero);
    fa1.force_step(0.01);
    fa2.force_step(0.01);
    const GaugeField::LinkField P = U.download_momentum();
    double norm = 0, ah = 0;
    for (int mu = 0; mu < 4; ++mu)
        for (int64_t x = 0; x < V; ++x)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const std::complex<double> a(P[mu][18 * x + 2 * (3 * i + j)], P[mu][18 * x + 2 * (3 * i + j) + 1]);
                    const std::complex<double> b(P[mu][18 * x + 2 * (3 * j + i)], P[mu][18 * x + 2 * (3 * j + i) + 1]);
                    norm += std::norm(a);
                    ah += std::norm(a + std::conj(b));
                }
    CHECK(norm > 0 && ah < 1e-26 * norm, "Hasenbusch force steps: momentum change anti-Hermitian");
}

// The operator's constructors and its single-precision type (wilson.h:250-269):
// D(kappa, Field<matrix> (&U)[NDIM]) gives the operator of D(kappa, GaugeField) bit for bit,
// and sees a change of U at its next call (it holds a reference, wilson.h:237); type_flt built
// by the converting constructor on gauge.get_single_precision() applies within FP32 rounding
// of D and, used as fermion_action::initial_guess uses it (fermion_field.h:73-84: CG of the
// FP32 operator on float copies, the result copied back to psi), leaves the FP64 CG fewer
// iterations; fermion_action with that guess gives the same action.
static void run_single_precision(Lattice &lattice) {
    std::printf("Checking the operator constructors and type_flt\n");
    GaugeField U(lattice);
    U.random(2, 0.3, 51);
    using dirac = Dirac_Wilson_evenodd<SU3d>;
    static_assert(std::is_same<dirac::type_flt, Dirac_Wilson_evenodd<SU3f>>::value, "type_flt");
    static_assert(std::is_same<dirac::type_flt::vector_type, WilsonFieldT<float>>::value, "float fields");
    static_assert(std::is_same<Hasenbusch_operator<dirac>::type_flt, Hasenbusch_operator<dirac::type_flt>>::value,
                  "Hasenbusch type_flt");
    const double kappa = 0.124;
    dirac D(kappa, U);
    dirac Darr(kappa, U.fdir);
    WilsonField a(lattice), Da(lattice), Da2(lattice);
    a.gaussian_random(52, EVEN);
    D.apply(a, Da);
    Darr.apply(a, Da2);
    CHECK(Da.host() == Da2.host(), "D(kappa, U.fdir) = D(kappa, U)");

    // the operator holds the gauge fields by reference: a change of U shows at the next call
    {
        GaugeField V(lattice);
        V.random(0, 0.0, 1);  // unit links
        dirac Dv(kappa, V.fdir), Du(kappa, V);
        WilsonField unit_out(lattice), changed_out(lattice);
        Dv.apply(a, unit_out);
        V.direction(2) = U.direction(2);  // replace the z links
        Du.apply(a, changed_out);
        CHECK(!(unit_out.host() == changed_out.host()), "operator sees its gauge field change");
        V.direction(2).assign(V.direction(2).size(), 0.0);
        for (int64_t x = 0; x < lattice.volume(); ++x)
            for (int i = 0; i < 3; ++i) V.direction(2)[18 * x + 2 * 4 * i] = 1.0;  // unit again
        Dv.apply(a, changed_out);
        CHECK(unit_out.host() == changed_out.host(), "... and its restoration");
    }

    dirac::type_flt D_flt(D, U.get_single_precision());
    CHECK(D_flt.kappa == D.kappa, "type_flt takes kappa from the double operator");
    WilsonFieldT<float> af(lattice), Daf(lattice);
    af = a;
    D_flt.apply(af, Daf);
    WilsonField back(lattice);
    back = Daf;
    CHECK(diff2_even(back, Da) < 1e-11 * Da.squarenorm(EVEN), "type_flt.apply within FP32 rounding of apply");
    D_flt.dagger(af, Daf);
    D.dagger(a, Da);
    back = Daf;
    CHECK(diff2_even(back, Da) < 1e-11 * Da.squarenorm(EVEN), "type_flt.dagger within FP32 rounding of dagger");

    // initial_guess as fermion_field.h:73-84
    WilsonField chi(lattice), psi(lattice), psi0(lattice);
    chi.gaussian_random(53, EVEN);
    chi.zero(ODD);
    psi.zero(ALL);
    WilsonFieldT<float> c(lattice), p(lattice);
    c = chi;
    p = psi;
    CG<dirac::type_flt> inverse_flt(D_flt);
    inverse_flt.apply(c, p);
    psi = p;
    CHECK(inverse_flt.stats.converged == 1, "CG of the single-precision operator");
    CG<dirac> inverse(D, 1e-10);
    inverse.apply(chi, psi);
    const int it_guess = inverse.stats.iterations;
    psi0.zero(ALL);
    inverse.apply(chi, psi0);
    const int it_zero = inverse.stats.iterations;
    std::printf("      FP64 CG iterations: from zero %d, from the FP32 solve %d\n", it_zero, it_guess);
    CHECK(it_guess * 2 <= it_zero && inverse.stats.true_resid <= 1e-10, "FP64 CG from the single-precision guess");
    CHECK(diff2_even(psi, psi0) < 1e-16 * psi0.squarenorm(EVEN), "same solution from either start");

    fermion_action<dirac> fa(D, U), fb(D, U);
    fa.draw_gaussian_fields(54);
    fb.draw_gaussian_fields(54);
    fb.set_single_precision_guess(true);
    const double sa = fa.action(), sb = fb.action();
    CHECK(std::abs(sa - sb) < 1e-9 * std::abs(sa), "fermion_action with the single-precision guess");
    Hasenbusch_operator<dirac> Dh(D, 0.1);
    Hasenbusch_operator<dirac>::type_flt Dh_flt(Dh, U.get_single_precision());
    CHECK(Dh_flt.h_parameter == 0.1 && Dh_flt.kappa == kappa, "Hasenbusch type_flt converting constructor");
}

int main() {
    try {
        Lattice lattice({16, 8, 8, 8});  // test_CG.cpp:57, NDIM == 4
        run(lattice, 0, "unit");
        run(lattice, 1, "Haar random");
        run(lattice, 3, "SU<3>::random (16 SU(2) hits, as test_CG.cpp draws them)");
        run_single_precision(lattice);
        run_hasenbusch(lattice);
        run_force(lattice);
        run_hasenbusch_actions(lattice);
    } catch (const error &e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 2;
    }
    if (failures) {
        std::printf("%d FAILED\n", failures);
        return 1;
    }
    std::printf("PASS\n");
    return 0;
}
