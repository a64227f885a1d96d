"""Hasenbusch-shifted operator and CG (wl_hasenbusch_apply, wl_cg_solve_ex; Hasenbusch.h:15-53)
and the MRE initial guess (wl_mre_guess; MRE_guess.h:8-87) on the GPU, through the C ABI,
against the oracle (oracle_hasenbusch / oracle_cg_h, oracle/mre.py).

Floating point: the shifted operator to 1e-13 relative in FP64 (as the unshifted one), 1e-6 in
FP32, 2e-3 in FP16 storage; CG solutions to the FP64 true residual they are asked for; the MRE
guess to 1e-10 relative of the oracle's (the library sums the Gram and projected matrices in
one pass and orthogonalises in that metric; the oracle orthogonalises the fields), and to the
Galerkin condition it is defined by."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle import mre
from oracle.oracle import EVEN, Oracle

pytestmark = pytest.mark.gpu

KAPPA, MH = 0.1, 0.1   # test_CG.cpp:294-296


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup8():
    dims = (8, 8, 8, 8)
    U = W.gen_gauge(dims, "smooth", 0.4, 201)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    yield dims, U, lat, Oracle(dims)
    lat.close()


def _even(dims, seed):
    return W.gen_spinor(dims, seed, EVEN)


@pytest.mark.parametrize("prec,nrhs,tol", [(W.F64, 1, 1e-13), (W.F64, 3, 1e-13), (W.F32, 1, 1e-6),
                                           (W.F32, 4, 1e-6), (W.F16, 2, 2e-3)])
def test_shifted_operator_matches_oracle(setup8, prec, nrhs, tol):
    dims, U, lat, orc = setup8
    _say("hasenbusch apply", prec, nrhs)
    H = W.Hasenbusch_operator(W.Dirac_Wilson_evenodd(KAPPA, lat), MH)
    fin, fout = W.Field(lat, EVEN, prec, nrhs), W.Field(lat, EVEN, prec, nrhs)
    srcs = [_even(dims, 210 + j) for j in range(nrhs)]
    for j, s in enumerate(srcs):
        fin.upload(s, rhs=j)
    for sign, op in ((1, H.apply), (-1, H.dagger)):
        op(fin, fout)
        for j, s in enumerate(srcs):
            want = orc.meo(U, KAPPA, s, sign, mh=MH)
            assert rel(fout.to_full(j), want) <= tol, (sign, j)


def test_zero_shift_is_the_plain_operator_bit_for_bit(setup8):
    dims, U, lat, orc = setup8
    D = W.Dirac_Wilson_evenodd(KAPPA, lat)
    H = W.Hasenbusch_operator(D, 0.0)
    fin = W.Field(lat).upload(_even(dims, 220))
    a, b = W.Field(lat), W.Field(lat)
    D.apply(fin, a)
    H.apply(fin, b)
    assert a.download().tobytes() == b.download().tobytes()


def test_checks_of_test_cg_hasenbusch(setup8):
    """test_CG.cpp:290-345 on the GPU: adjointness of D + mh, then the CG round trips."""
    dims, U, lat, orc = setup8
    _say("test_CG Hasenbusch block")
    H = W.Hasenbusch_operator(W.Dirac_Wilson_evenodd(KAPPA, lat), MH)
    a, b = W.Field(lat).upload(_even(dims, 230)), W.Field(lat).upload(_even(dims, 231))
    Db, Dda = W.Field(lat), W.Field(lat)
    H.apply(b, Db)
    H.dagger(a, Dda)
    ca = lambda f: f.download()[:, 0::2] + 1j * f.download()[:, 1::2]
    lhs, rhs = np.vdot(ca(a), ca(Db)), np.vdot(ca(Dda), ca(b))
    assert abs(lhs - rhs) < 1e-13 * abs(lhs)
    x = W.Field(lat)
    st = W.CG(H, 1e-12).apply(Dda, x)[0]       # b = 1/D a -> D b = a
    assert st["converged"] == 1
    Dx = W.Field(lat)
    H.apply(x, Dx)
    assert rel(Dx.download(), a.download()) < 1e-10


@pytest.mark.parametrize("precision,nrhs", [("f64", 1), ("f64", 3), ("f32", 1), ("f32", 2), ("f16", 2)])
def test_shifted_cg_solves_the_shifted_system(setup8, precision, nrhs):
    dims, U, lat, orc = setup8
    _say("shifted CG", precision, nrhs)
    H = W.Hasenbusch_operator(W.Dirac_Wilson_evenodd(KAPPA, lat), MH)
    b = W.Field(lat, EVEN, W.F64, nrhs)
    srcs = [_even(dims, 240 + j) for j in range(nrhs)]
    for j, s in enumerate(srcs):
        b.upload(s, rhs=j)
    x = W.Field(lat, EVEN, W.F64, nrhs)
    stats = W.CG(H, 1e-10, precision=precision, delta=0.1 if precision != "f64" else 1e-3).apply(b, x)
    for j, s in enumerate(srcs):
        assert stats[j]["converged"] == 1
        xs = x.to_full(j)
        r = s - orc.meo(U, KAPPA, orc.meo(U, KAPPA, xs, 1, mh=MH), -1, mh=MH)
        assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(s)
        xo, _, _ = orc.cg(U, KAPPA, s, accuracy=1e-12, mh=MH)
        assert rel(xs, xo) < 1e-8


def _basis_field(lat, vecs, cap=None):
    B = W.Field(lat, EVEN, W.F64, cap or max(1, len(vecs)))
    for j, v in enumerate(vecs):
        B.upload(v, rhs=j)
    return B


def _A(orc, U, mh):
    return lambda v: orc.meo(U, KAPPA, orc.meo(U, KAPPA, v, 1, mh=mh), -1, mh=mh)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("mh", [0.0, MH])
def test_mre_matches_oracle(setup8, n, mh):
    dims, U, lat, orc = setup8
    _say("MRE", n, mh)
    vecs = [_even(dims, 300 + 10 * n + j) for j in range(n)]
    chi = _even(dims, 399)
    B = _basis_field(lat, vecs, cap=8)
    psi = W.Field(lat)
    D = W.Hasenbusch_operator(W.Dirac_Wilson_evenodd(KAPPA, lat), mh)
    W.MRE_guess(psi, W.Field(lat).upload(chi), D, B, n)
    got = psi.to_full()
    A = _A(orc, U, mh)
    want = mre.mre_guess(chi, vecs, A)
    assert rel(got, want) < 1e-10
    r = chi - A(got)
    for v in vecs:
        assert abs(mre.rdot(v, r)) < 1e-10 * np.linalg.norm(v) * np.linalg.norm(chi)


def test_mre_zero_duplicate_and_empty_bases(setup8):
    dims, U, lat, orc = setup8
    D = W.Dirac_Wilson_evenodd(KAPPA, lat)
    chi = W.Field(lat).upload(_even(dims, 410))
    psi = W.Field(lat).upload(_even(dims, 411))   # overwritten
    zeros = _basis_field(lat, [], cap=4)
    zeros.zero()
    W.MRE_guess(psi, chi, D, zeros, 3)             # HILA's initial all-zero basis
    assert not np.any(psi.download())
    psi.upload(_even(dims, 412))
    W.MRE_guess(psi, chi, D, zeros, 0)
    assert not np.any(psi.download())
    b = [_even(dims, 413), _even(dims, 414)]
    ref = W.Field(lat)
    W.MRE_guess(ref, chi, D, _basis_field(lat, b), 2)
    dup = _basis_field(lat, [b[0], np.zeros_like(b[0]), b[1], b[0], 2.0 * b[1]])
    W.MRE_guess(psi, chi, D, dup, 5)
    assert rel(psi.download(), ref.download()) < 1e-12
    with pytest.raises(W.WilsonError):
        W.MRE_guess(psi, chi, D, dup, 6)           # more than the basis field holds


def test_mre_chronological_guess_cuts_iterations(setup8):
    """Solutions x_k of A x_k = c0 + k c1 + k^2 c2 are quadratic in k: from x_0..x_2 kept with
    wl_field_copy_rhs, the guess for k = 3 is x_3 up to the solves' own errors, and CG from it
    needs a fraction of the iterations."""
    dims, U, lat, orc = setup8
    D = W.Dirac_Wilson_evenodd(0.124, lat)
    c = [_even(dims, 420 + i) for i in range(3)]
    B = W.Field(lat, EVEN, W.F64, 4)
    x = W.Field(lat)
    iters = []
    for k in range(4):
        chi = W.Field(lat).upload(c[0] + k * c[1] + k * k * c[2])
        x.zero()
        if k > 0:
            W.MRE_guess(x, chi, D, B, k)
        iters.append(W.CG(D, 1e-10).apply(chi, x)[0]["iterations"])
        B.copy_rhs_from(k, x)
    assert iters[0] > 20 and iters[3] * 4 <= iters[0], iters


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_served_shifted_request_matches_blocking_solver(precision):
    """A request with mh is solved with M + mh: the same bits as the blocking batched solver on
    the Hasenbusch operator, and the FP64 true residual of the shifted system meets the
    tolerance."""
    dims, nrhs, kappa, tol = (8, 8, 8, 8), 3, 0.124, 1e-10
    U = W.gen_gauge(dims, "smooth", 0.3, 430)
    Vh = int(np.prod(dims)) // 2
    src = np.ascontiguousarray(np.stack([_even(dims, 431 + j)[:Vh] for j in range(nrhs)]))
    sol = np.zeros_like(src)
    _say("served shifted request", precision)
    with W.Server() as S:
        t = S.submit(dims, U, src, sol, kappa, tol, precision=precision, delta=0.1, mh=MH)
        res, st = S.wait(t)
    assert res["status"] == W.WL_OK
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    H = W.Hasenbusch_operator(W.Dirac_Wilson_evenodd(kappa, lat), MH)
    b = W.Field(lat, EVEN, W.F64, nrhs)
    x = W.Field(lat, EVEN, W.F64, nrhs)
    for j in range(nrhs):
        b.upload(src[j], rhs=j)
    bst = W.CG(H, tol, precision=precision, delta=0.1, update_every=100).apply(b, x)
    orc = Oracle(dims)
    for j in range(nrhs):
        assert np.array_equal(x.download(j), sol[j]), j
        assert bst[j]["iterations"] == st[j]["iterations"]
        full = np.zeros((2 * Vh, 24))
        full[:Vh] = sol[j]
        sf = np.zeros_like(full)
        sf[:Vh] = src[j]
        r = sf - orc.meo(U, kappa, orc.meo(U, kappa, full, 1, mh=MH), -1, mh=MH)
        assert np.linalg.norm(r) <= 1.5 * tol * np.linalg.norm(sf)
    lat.close()
