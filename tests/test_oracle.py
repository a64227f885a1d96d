"""Pins the CPU oracle (oracle/wilson_oracle.c, the C restatement of HILA's wilson.h and
conjugate_gradient.h) before it is trusted as the checker of the HIP path.

Each function of the oracle is checked against something that a plausible mistake (a wrong
sign, a swapped gamma row, a dropped hop, a transposed link, a misplaced boundary sign) would
fail: the independent dense construction from the gamma tables, closed forms for free fields,
algebraic invariants HILA's own tests assert, brute-force solves on small lattices and the
committed fixtures.  CPU only.
"""
import glob
import math
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from hila_amd import wilson as W
from oracle import dense
from oracle.oracle import ALL, EVEN, ODD, Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cplx(f):
    return dense.field_complex(f)


# ------------------------------------------------------------ gamma algebra (a2-a3) --------

def test_gamma_algebra():
    """wilson_vector.h:311-347 tables: Hermitian, g_mu^2 = 1, anticommuting, g5 anticommutes
    with all four and equals g1 g2 g3 g0 up to the basis sign fixed by the table."""
    G = dense.GAMMA_DIR
    for m in range(4):
        assert np.allclose(G[m], G[m].conj().T)
        assert np.allclose(G[m] @ G[m], np.eye(4))
        assert np.allclose(G[m] @ dense.GAMMA5 + dense.GAMMA5 @ G[m], 0)
        for n in range(m + 1, 4):
            assert np.allclose(G[m] @ G[n] + G[n] @ G[m], 0)
    prod = G[0] @ G[1] @ G[2] @ G[3]
    assert np.allclose(prod, dense.GAMMA5) or np.allclose(prod, -dense.GAMMA5)


def test_projector_rank_and_norm_identities():
    """test_CG.cpp:17-41: ||w + g_d w||^2 = 2 ||h+||^2 where h+ are the upper two spin rows of
    (1 + g_d) w (HalfWilsonVector keeps rows 0,1 and expand rebuilds rows 2,3 from them)."""
    rng = np.random.default_rng(3)
    for m in range(4):
        for s in (1, -1):
            P = np.eye(4) + s * dense.GAMMA_DIR[m]
            assert np.linalg.matrix_rank(P) == 2
            w = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
            full = P @ w
            assert abs(np.vdot(full, full).real - 2 * np.vdot(full[:2], full[:2]).real) < 1e-12


# ----------------------------------------------- C restatement vs dense construction -------

LATTICES = [
    ((2, 2, 2, 4), "haar", 0.0, 1, 0.12, True),
    ((4, 4, 4, 4), "smooth", 0.3, 2, 0.125, True),
    ((4, 2, 6, 4), "haar", 0.0, 3, 0.11, False),
    ((6, 4, 2, 8), "smooth", 0.5, 4, 0.13, True),
]


@pytest.mark.parametrize("case", LATTICES, ids=lambda c: "x".join(map(str, c[0])))
def test_oracle_matches_dense(case):
    dims, kind, eps, seed, kappa, ap = case
    lat = dense.Lattice(dims, antiperiodic_t=ap)
    U = W.gen_gauge(dims, kind, eps, seed)
    Wd = dense.WilsonDense(lat, U, kappa)
    orc = Oracle(dims, antiperiodic_t=ap)
    pe = W.gen_spinor(dims, 10 + seed, EVEN)
    po = W.gen_spinor(dims, 20 + seed, ODD)
    Vh = lat.Vh
    for s in (1, -1):
        got = orc.meo(U, kappa, pe, s)
        assert np.all(got[Vh:] == 0)  # out[ODD] = 0 (wilson.h:278)
        assert rel(got[:Vh], dense.field_real(Wd.apply_M(cplx(pe)[:Vh], s))) < 1e-14
        ho = orc.hop_set(U, kappa, pe, ODD, s)
        assert rel(ho[Vh:], dense.field_real(Wd.hop_set_oe(cplx(pe)[:Vh], s))) < 1e-14
        he = orc.hop_set(U, kappa, po, EVEN, s)
        assert rel(he[:Vh], dense.field_real(Wd.hop_set_eo(cplx(po)[Vh:], s))) < 1e-14
        # the fused one-pass CPU variant computes the same operator
        assert rel(orc.hop_set(U, kappa, pe, ODD, s, fused=True), ho) < 1e-14


def test_hop_accumulates_like_hila():
    """Dirac_Wilson_hop subtracts from out (wilson.h:45-47); hop_set overwrites (:82)."""
    dims = (4, 4, 4, 4)
    U = W.gen_gauge(dims, "haar", 0, 5)
    orc = Oracle(dims)
    pe = W.gen_spinor(dims, 6, EVEN)
    base = W.gen_spinor(dims, 7, ODD)
    out = base.copy()
    orc.hop(U, 0.12, pe, out, ODD, 1)
    hs = orc.hop_set(U, 0.12, pe, ODD, 1)
    assert rel(out - base, hs) < 1e-14


# ------------------------------------------------------------- closed forms --------------

def _free_check(orc, lat, kappa, p, u):
    U = W.gen_gauge(lat.L, "unit")
    psi = dense.plane_wave(lat, p, u)
    Vh = lat.Vh
    psi_e = np.zeros_like(psi)
    psi_e[:Vh] = psi[:Vh]
    h2 = dense.free_field_h2(lat, p)
    expect = psi_e - kappa ** 2 * np.einsum("ab,vbc->vac", h2, psi_e)
    out_re = orc.meo(U, kappa, dense.field_real(psi_e.real + 0j), 1)
    out_im = orc.meo(U, kappa, dense.field_real(psi_e.imag + 0j), 1)
    got = cplx(out_re) + 1j * cplx(out_im)  # operator is real-linear: M(a + ib) = Ma + iMb
    return rel(got[:Vh], expect[:Vh])


@pytest.mark.parametrize("n", [(0, 0, 0, 0), (1, 0, 2, 1), (1, 1, 1, 2), (3, 2, 1, 0)])
def test_free_field_plane_waves_antiperiodic(n):
    """Unit links: M e^{ip.x} u = (1 - kappa^2 h(p)^2) e^{ip.x} u on even sites for
    p_{x,y,z} = 2 pi n / L and p_t = (2 n_t + 1) pi / L_t (antiperiodic time)."""
    dims = (4, 6, 4, 8)
    lat = dense.Lattice(dims, antiperiodic_t=True)
    orc = Oracle(dims, antiperiodic_t=True)
    p = np.array([2 * np.pi * n[0] / dims[0], 2 * np.pi * n[1] / dims[1],
                  2 * np.pi * n[2] / dims[2], (2 * n[3] + 1) * np.pi / dims[3]])
    rng = np.random.default_rng(sum(n))
    u = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    assert _free_check(orc, lat, 0.12, p, u) < 1e-13


def test_free_field_periodic_momentum_fails_with_antiperiodic_bc():
    """The same plane wave with a periodic time momentum is not an eigenvector when the time
    boundary is antiperiodic: this catches a missing or misplaced boundary sign."""
    dims = (4, 4, 4, 4)
    lat = dense.Lattice(dims, antiperiodic_t=True)
    orc = Oracle(dims, antiperiodic_t=True)
    p = np.array([2 * np.pi / 4, 0, 0, 2 * np.pi / 4])
    u = np.ones((4, 3)) + 0.5j
    assert _free_check(orc, lat, 0.12, p, u) > 1e-2
    # and with periodic time it is
    lat_p = dense.Lattice(dims, antiperiodic_t=False)
    orc_p = Oracle(dims, antiperiodic_t=False)
    assert _free_check(orc_p, lat_p, 0.12, p, u) < 1e-13


# ------------------------------------------------------------- invariants ----------------

@pytest.fixture(scope="module")
def rand8():
    dims = (8, 8, 8, 8)
    return dims, W.gen_gauge(dims, "haar", 0, 1), Oracle(dims)


def test_adjointness(rand8):
    """<a, M b> = <M^dag a, b> (test_CG.cpp:236-245)."""
    dims, U, orc = rand8
    a = W.gen_spinor(dims, 31, EVEN)
    b = W.gen_spinor(dims, 32, EVEN)
    lhs = np.vdot(cplx(a), cplx(orc.meo(U, 0.12, b, 1)))
    rhs = np.vdot(cplx(orc.meo(U, 0.12, a, -1)), cplx(b))
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)


def test_gamma5_hermiticity(rand8):
    """M^dag = g5 M g5 (g5 acts site-locally on the spin index)."""
    dims, U, orc = rand8
    a = W.gen_spinor(dims, 33, EVEN)
    g5 = lambda f: dense.field_real(np.einsum("ab,vbc->vac", dense.GAMMA5, cplx(f)))
    assert rel(g5(orc.meo(U, 0.12, g5(a), 1)), orc.meo(U, 0.12, a, -1)) < 1e-14


def test_gauge_covariance():
    """D[U^g](g psi) = g D[U] psi with U^g_mu(x) = g(x) U_mu(x) g(x+mu)^dag."""
    dims = (4, 4, 4, 6)
    lat = dense.Lattice(dims)
    U = W.gen_gauge(dims, "smooth", 0.4, 8)
    g = dense.links_complex(W.gen_gauge(dims, "haar", 0, 9))[0]  # one random SU(3) per site
    Uc = dense.links_complex(U)
    Ug = np.empty_like(Uc)
    for mu in range(4):
        fwd, _ = lat.hop(mu, +1)
        Ug[mu] = g @ Uc[mu] @ np.conj(np.transpose(g[fwd], (0, 2, 1)))
    Ug_r = np.stack([Ug.real, Ug.imag], -1).reshape(4, lat.V, 18)
    orc = Oracle(dims)
    psi = W.gen_spinor(dims, 10, EVEN)
    gpsi = dense.field_real(np.einsum("vij,vaj->vai", g, cplx(psi)))
    lhs = orc.meo(Ug_r, 0.12, gpsi, 1)
    rhs = dense.field_real(np.einsum("vij,vaj->vai", g, cplx(orc.meo(U, 0.12, psi, 1))))
    assert rel(lhs, rhs) < 1e-13


def test_translation_covariance():
    """Shifting source and gauge field by two sites in x (parity preserving) shifts M psi."""
    dims = (6, 4, 4, 4)
    lat = dense.Lattice(dims)
    U = W.gen_gauge(dims, "smooth", 0.3, 12)
    psi = W.gen_spinor(dims, 13, EVEN)
    sh = lat.site_of(lat.coords + np.array([2, 0, 0, 0]))  # site x -> x + 2 e_x
    Us = np.empty_like(U)
    Us[:, sh] = U
    ps = np.empty_like(psi)
    ps[sh] = psi
    orc = Oracle(dims)
    out = orc.meo(U, 0.12, psi, 1)
    outs = np.empty_like(out)
    outs[sh] = out
    assert rel(orc.meo(Us, 0.12, ps, 1), outs) < 1e-14


def test_schur_complement_against_full_dirac_solve():
    """Even-odd solve x_e = M^{-1}(b_e + kappa H_eo b_o), x_o = b_o + kappa H_oe x_e equals the
    solution of the full D x = b (the decomposition Dirac_Wilson_evenodd stands for)."""
    dims = (4, 4, 2, 4)
    lat = dense.Lattice(dims)
    U = W.gen_gauge(dims, "haar", 0, 14)
    Wd = dense.WilsonDense(lat, U, 0.12)
    rng = np.random.default_rng(5)
    b = rng.normal(size=12 * lat.V) + 1j * rng.normal(size=12 * lat.V)
    x = spla.spsolve(Wd.D(1).tocsc(), b)
    ne = Wd.ne
    Heo, Hoe = Wd.blocks(1)
    be, bo = b[:ne], b[ne:]
    xe = spla.spsolve(Wd.M(1).tocsc(), be + 0.12 * (Heo @ bo))
    xo = bo + 0.12 * (Hoe @ xe)
    assert np.linalg.norm(np.concatenate([xe, xo]) - x) < 1e-12 * np.linalg.norm(x)


# ---------------------------------------------------------------- CG ---------------------

def test_cg_round_trips(rand8):
    """test_CG.cpp:247-287 (the 1e-16 thresholds on ||.||^2 read as relative)."""
    dims, U, orc = rand8
    k = 0.12
    a = W.gen_spinor(dims, 40, EVEN)
    x, steps, rr = orc.cg(U, k, orc.meo(U, k, a, -1), accuracy=1e-12)
    assert rel(orc.meo(U, k, x, 1), a) < 1e-10          # M (M^dag M)^-1 M^dag a = a
    x2, _, _ = orc.cg(U, k, a, accuracy=1e-12)
    assert rel(orc.meo(U, k, orc.meo(U, k, x2, 1), -1), a) < 1e-11   # (M^dag M)(M^dag M)^-1 a
    b = W.gen_spinor(dims, 41, EVEN)
    x3, _, _ = orc.cg(U, k, orc.meo(U, k, orc.meo(U, k, b, 1), -1), accuracy=1e-12)
    assert rel(x3, b) < 1e-9                               # (M^dag M)^-1 (M^dag M) b = b
    assert rr < 1e-24 and steps > 0


def test_cg_iterated_residual_criterion():
    """The loop stops at the first i with ||r||^2 < accuracy^2 ||b||^2 (conjugate_gradient.h:91)
    and returns HILA's printed relative residue rrnew / source_norm."""
    dims = (4, 4, 4, 4)
    U = W.gen_gauge(dims, "haar", 0, 15)
    orc = Oracle(dims)
    b = W.gen_spinor(dims, 16, EVEN)
    for acc in (1e-4, 1e-8, 1e-12):
        x, steps, rr = orc.cg(U, 0.12, b, accuracy=acc)
        assert rr < acc * acc
        _, steps_short, rr_short = orc.cg(U, 0.12, b, accuracy=acc, maxiters=steps)
        assert steps_short == steps and rr_short >= acc * acc  # one iteration short: not there


def test_cg_matches_sparse_direct_solve():
    dims = (4, 4, 4, 4)
    lat = dense.Lattice(dims)
    U = W.gen_gauge(dims, "smooth", 0.3, 17)
    Wd = dense.WilsonDense(lat, U, 0.125)
    b = W.gen_spinor(dims, 18, EVEN)
    x, _, _ = Oracle(dims).cg(U, 0.125, b, accuracy=1e-12)
    xd = spla.spsolve((Wd.M(-1) @ Wd.M(1)).tocsc(), cplx(b)[: lat.Vh].reshape(-1))
    assert rel(cplx(x)[: lat.Vh].reshape(-1), xd) < 1e-9


def test_norm2_matches_compensated_sum():
    dims = (8, 8, 8, 8)
    f = W.gen_spinor(dims, 19, 3)
    orc = Oracle(dims)
    Vh = orc.Vh
    for par, sl in ((EVEN, slice(0, Vh)), (ODD, slice(Vh, None)), (ALL, slice(None))):
        ref = math.fsum((f[sl] ** 2).ravel())
        assert abs(orc.norm2(f, par) - ref) <= 1e-13 * ref


# ---------------------------------------------------------- golden fixtures --------------

FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_fixture(path):
    z = np.load(path)  # allow_pickle is False by default
    dims = tuple(int(d) for d in z["dims"])
    kappa, ap = float(z["kappa"]), bool(int(z["antiperiodic"]))
    orc = Oracle(dims, antiperiodic_t=ap)
    U, psi, psi_odd = z["U"], z["psi"], z["psi_odd"]
    Vh = orc.Vh
    assert rel(orc.meo(U, kappa, psi, 1)[:Vh], z["M_plus"]) < 1e-14
    assert rel(orc.meo(U, kappa, psi, -1)[:Vh], z["M_minus"]) < 1e-14
    assert rel(orc.hop_set(U, kappa, psi, ODD, 1)[Vh:], z["H_oe_plus"]) < 1e-14
    assert rel(orc.hop_set(U, kappa, psi, ODD, -1)[Vh:], z["H_oe_minus"]) < 1e-14
    assert rel(orc.hop_set(U, kappa, psi_odd, EVEN, 1)[:Vh], z["H_eo_plus"]) < 1e-14
    x, _, _ = orc.cg(U, kappa, psi, accuracy=1e-12)
    assert rel(x[:Vh], z["x_cg"]) < 1e-9


def test_fixtures_present():
    assert len(FIXTURES) >= 3


def test_thread_count_does_not_change_the_oracle():
    """The CPU baseline's single-core figure runs the same loops on one thread: per-site work,
    identical bits."""
    from oracle import oracle as O
    dims = (4, 4, 4, 8)
    orc = Oracle(dims)
    U = W.gen_gauge(dims, "haar", 0, 7)
    psi = W.gen_spinor(dims, 8, EVEN)
    many = orc.hop_set(U, 0.12, psi, ODD, 1)
    prev = O.set_threads(1)
    try:
        assert O.set_threads(0) == 1
        one = orc.hop_set(U, 0.12, psi, ODD, 1)
    finally:
        O.set_threads(prev)
    assert O.set_threads(0) == prev
    assert one.tobytes() == many.tobytes()
