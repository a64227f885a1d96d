"""Two-row links (WL_OPT_LINK_ROWS = 2) on the GPU, through the C ABI.

The FP64 and FP32 hopping kernels read rows 0 and 1 of every link and rebuild row 2 as
conj(row 0 x row 1).  For special unitary links that is the stored row to rounding, so the
operator matches the FP64 oracle to the same 1e-13 as with full links (1e-6 in FP32), and the
full-link result to 1e-14; links that are not special unitary are refused (WL_E_ARG), never
used.  Tolerances: the FP64 ones of the plain operator (a relative L2 difference); CG to the
FP64 true residual it is asked for."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, ODD, Oracle

pytestmark = pytest.mark.gpu

KAPPA = 0.12


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _say(*a):
    print(*a, flush=True)


def _lattice(dims, kind, seed, rows):
    U = W.gen_gauge(dims, kind, 0.3, seed)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    lat.set_option(W.OPT_LINK_ROWS, rows)
    return lat, U


@pytest.mark.parametrize("dims,kind", [((8, 8, 8, 8), "haar"), ((4, 6, 8, 10), "smooth"), ((8, 8, 8, 8), "unit")],
                         ids=["8^4-haar", "4x6x8x10-smooth", "8^4-unit"])
def test_operator_matches_oracle_and_full_links(dims, kind):
    _say("two-row operator", dims, kind)
    lat, U = _lattice(dims, kind, 31, 2)
    assert lat.get_option(W.OPT_LINK_ROWS) == 2
    full = W.Lattice(dims)
    full.upload_gauge(U)
    orc = Oracle(dims)
    psi = W.gen_spinor(dims, 32, EVEN)
    for prec, tol in ((W.F64, 1e-13), (W.F32, 1e-6)):
        D, Df = W.Dirac_Wilson_evenodd(KAPPA, lat), W.Dirac_Wilson_evenodd(KAPPA, full)
        fin, ffin = W.Field(lat, EVEN, prec).upload(psi), W.Field(full, EVEN, prec).upload(psi)
        for sign in (1, -1):
            fo, ffo = W.Field(lat, ODD, prec), W.Field(full, ODD, prec)
            D.hop_set(fin, fo, sign)
            Df.hop_set(ffin, ffo, sign)
            assert rel(fo.to_full(), orc.hop_set(U, KAPPA, psi, ODD, sign)) <= tol, (prec, sign)
            if prec == W.F64:
                assert rel(fo.to_full(), ffo.to_full()) <= 1e-14
            fe = W.Field(lat, EVEN, prec)
            (D.apply if sign == 1 else D.dagger)(fin, fe)
            assert rel(fe.to_full(), orc.meo(U, KAPPA, psi, sign)) <= tol, (prec, sign)
    full.close()
    lat.close()


@pytest.mark.parametrize("link_load", [0, 3], ids=["cached", "lds"])
@pytest.mark.parametrize("prec,tol", [(W.F64, 1e-13), (W.F32, 1e-6)], ids=["f64", "f32"])
def test_multi_rhs_operator_matches_oracle(prec, tol, link_load):
    dims, N = (8, 8, 8, 8), 5
    _say("two-row multi-rhs", prec, link_load)
    lat, U = _lattice(dims, "smooth", 33, 2)
    lat.set_option(W.OPT_LINK_LOAD, link_load)
    orc = Oracle(dims)
    D = W.Dirac_Wilson_evenodd(KAPPA, lat)
    srcs = [W.gen_spinor(dims, 340 + j, EVEN) for j in range(N)]
    fin, fe = W.Field(lat, EVEN, prec, N), W.Field(lat, EVEN, prec, N)
    for j, s in enumerate(srcs):
        fin.upload(s, rhs=j)
    for sign in (1, -1):
        (D.apply if sign == 1 else D.dagger)(fin, fe)
        for j, s in enumerate(srcs):
            assert rel(fe.to_full(j), orc.meo(U, KAPPA, s, sign)) <= tol, (sign, j)
    lat.close()


@pytest.mark.parametrize("precision,nrhs", [("f64", 1), ("f64", 3), ("f32", 1), ("f32", 2), ("f16", 2)])
def test_cg_reaches_the_true_residual(precision, nrhs):
    dims = (8, 8, 8, 8)
    _say("two-row CG", precision, nrhs)
    lat, U = _lattice(dims, "smooth", 35, 2)
    D = W.Dirac_Wilson_evenodd(0.124, lat)
    srcs = [W.gen_spinor(dims, 350 + j, EVEN) for j in range(nrhs)]
    b, x = W.Field(lat, EVEN, W.F64, nrhs), W.Field(lat, EVEN, W.F64, nrhs)
    for j, s in enumerate(srcs):
        b.upload(s, rhs=j)
    st = W.CG(D, 1e-10, precision=precision, delta=0.1 if precision != "f64" else 1e-3).apply(b, x)
    orc = Oracle(dims)
    for j, s in enumerate(srcs):
        assert st[j]["converged"] == 1
        xs = x.to_full(j)
        r = s - orc.meo(U, 0.124, orc.meo(U, 0.124, xs, 1), -1)
        assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(s)
    lat.close()


def test_links_that_are_not_special_unitary_are_refused():
    dims = (4, 4, 4, 4)
    U = W.gen_gauge(dims, "haar", 0, 36)
    bad = U.copy()
    bad[2, 17, :] *= 1.001                      # one link no longer unitary
    lat = W.Lattice(dims)
    lat.upload_gauge(bad)
    with pytest.raises(W.WilsonError):
        lat.set_option(W.OPT_LINK_ROWS, 2)
    assert lat.get_option(W.OPT_LINK_ROWS) == 3
    D = W.Dirac_Wilson_evenodd(KAPPA, lat)
    psi = W.gen_spinor(dims, 37, EVEN)
    fin, fe = W.Field(lat).upload(psi), W.Field(lat)
    D.apply(fin, fe)                            # full links: fine
    orc = Oracle(dims)
    assert rel(fe.to_full(), orc.meo(bad, KAPPA, psi, 1)) <= 1e-13
    # set before an upload: checked at the next operator call
    lat.upload_gauge(U)
    lat.set_option(W.OPT_LINK_ROWS, 2)
    lat.upload_gauge(bad)
    with pytest.raises(W.WilsonError):
        D.apply(fin, fe)
    with pytest.raises(W.WilsonError):
        lat.set_option(W.OPT_LINK_ROWS, 4)
    lat.close()


@pytest.mark.parametrize("where,value", [((8, 0), np.nan), ((7, 1), np.inf), ((0, 0), np.nan)],
                         ids=["nan_row2", "inf_row2", "nan_row0"])
def test_links_with_nan_or_inf_entries_are_refused(where, value):
    """A NaN or infinite entry in any row makes the stored third row fail the check: the
    full-link operator would return NaN there, so two-row links must not replace it."""
    dims = (4, 4, 4, 4)
    U = W.gen_gauge(dims, "haar", 0, 36)
    entry, part = where
    U[1, 5, 2 * entry + part] = value
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    with pytest.raises(W.WilsonError) as e:
        lat.set_option(W.OPT_LINK_ROWS, 2)
    assert e.value.code == W.WL_E_ARG
    assert lat.get_option(W.OPT_LINK_ROWS) == 3
    lat.close()


def test_a_new_upload_replaces_the_two_row_copies():
    dims = (8, 8, 8, 8)
    lat, U1 = _lattice(dims, "haar", 38, 2)
    D = W.Dirac_Wilson_evenodd(KAPPA, lat)
    psi = W.gen_spinor(dims, 39, EVEN)
    orc = Oracle(dims)
    for prec, tol in ((W.F64, 1e-13), (W.F32, 1e-6)):
        fin, fe = W.Field(lat, EVEN, prec).upload(psi), W.Field(lat, EVEN, prec)
        D.apply(fin, fe)
        assert rel(fe.to_full(), orc.meo(U1, KAPPA, psi, 1)) <= tol
    U2 = W.gen_gauge(dims, "smooth", 0.4, 40)
    lat.upload_gauge(U2)
    for prec, tol in ((W.F64, 1e-13), (W.F32, 1e-6)):
        fin, fe = W.Field(lat, EVEN, prec).upload(psi), W.Field(lat, EVEN, prec)
        D.apply(fin, fe)
        assert rel(fe.to_full(), orc.meo(U2, KAPPA, psi, 1)) <= tol
    lat.close()
