"""Block CG (WL_OPT_BLOCK_KRYLOV = 1: the right-hand sides share one Krylov space) against the
batched independent CG and the oracle: every rhs reaches the FP64 true residual, the solutions
agree with the single-rhs solves, in FP64 and with an FP32 inner precision and reliable
updates, with the Hasenbusch shift, and with a repeated and a zero source in the block
(conjugate_gradient.h:46-104 per rhs; the block form is the Krylov-space generalisation)."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def sources(dims, n, seed):
    Vh = int(np.prod(dims)) // 2
    out = []
    for j in range(n):
        if j < n // 2:  # point sources: spin-colour pair j at one even site
            s = np.zeros((Vh, 24))
            s[5, 2 * (j % 12)] = 1.0
            out.append(s)
        else:
            out.append(W.gen_spinor(dims, seed + j, EVEN)[:Vh])
    return out


def solve(lat, op, src, block, precision="f64", tol=1e-10, cg_kw=None):
    lat.set_option(W.OPT_BLOCK_KRYLOV, 1 if block else 0)
    n = len(src)
    b = W.Field(lat, EVEN, W.F64, n)
    x = W.Field(lat, EVEN, W.F64, n)
    for j, s in enumerate(src):
        b.upload(s, rhs=j)
    st = W.CG(op, tol, 10000, precision=precision, delta=1e-3, update_every=100, **(cg_kw or {})).apply(b, x)
    lat.set_option(W.OPT_BLOCK_KRYLOV, 0)
    return [x.download(j) for j in range(n)], st


@pytest.mark.parametrize("dims,n,precision", [((8, 8, 8, 8), 6, "f64"), ((8, 8, 8, 8), 12, "f32"),
                                              ((16, 16, 16, 16), 12, "f64"), ((16, 16, 16, 16), 24, "f32")],
                         ids=["8^4-6-f64", "8^4-12-f32", "16^4-12-f64", "16^4-24-f32"])
def test_block_cg_matches_batched_cg(dims, n, precision):
    print("block CG", dims, n, precision, flush=True)
    U = W.gen_gauge(dims, "smooth", 0.3, 5)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    src = sources(dims, n, 40)
    xb, sb = solve(lat, D, src, True, precision)
    xr, sr = solve(lat, D, src, False, precision)
    its_b, its_r = sb[0]["iterations"], max(s["iterations"] for s in sr)
    print(f"  block {its_b} iterations ({sb[0]['reliable_updates']} updates), batched {its_r}", flush=True)
    for j in range(n):
        assert sb[j]["converged"] == 1 and sb[j]["true_resid"] <= 1e-10, (j, sb[j])
        assert sb[j]["iterations"] == its_b
        assert rel(xb[j], xr[j]) <= 1e-8, (j, rel(xb[j], xr[j]))
    assert its_b <= its_r
    # the oracle's true residual of two of them
    orc = Oracle(dims)
    for j in (0, n - 1):
        full = np.zeros((lat.V, 24))
        full[: lat.Vh] = xb[j]
        bf = np.zeros((lat.V, 24))
        bf[: lat.Vh] = src[j]
        r = bf - orc.meo(U, 0.125, orc.meo(U, 0.125, full, 1), -1)
        t = np.linalg.norm(r) / np.linalg.norm(bf)
        assert t <= 1e-10 and abs(t - sb[j]["true_resid"]) <= 1e-3 * t + 1e-14
    lat.close()


def test_block_cg_repeated_and_zero_sources_and_shift():
    dims = (8, 8, 8, 8)
    U = W.gen_gauge(dims, "haar", 0.0, 1)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    src = sources(dims, 6, 70)
    src[3] = src[1].copy()           # a repeated source: a dependent direction
    src[4] = np.zeros_like(src[0])   # a zero source: x = 0
    for precision in ("f64", "f32"):
        xb, sb = solve(lat, D, src, True, precision)
        for j in range(6):
            assert sb[j]["converged"] == 1 and sb[j]["true_resid"] <= 1e-10, (precision, j, sb[j])
        assert rel(xb[3], xb[1]) <= 1e-9
        assert not xb[4].any()
    # Hasenbusch shift (M + mh): block against batched
    H = W.Hasenbusch_operator(D, 0.1)
    xb, sb = solve(lat, H, src[:3], True)
    xr, sr = solve(lat, H, src[:3], False)
    for j in range(3):
        assert sb[j]["converged"] == 1 and rel(xb[j], xr[j]) <= 1e-8
    # more rhs than one block holds
    lat.set_option(W.OPT_BLOCK_KRYLOV, 1)
    b = W.Field(lat, EVEN, W.F64, 25)
    with pytest.raises(W.WilsonError) as e:
        W.CG(D, 1e-10).apply(b, W.Field(lat, EVEN, W.F64, 25), raise_on_noconv=True)
    assert e.value.code == W.WL_E_ARG
    # block CG has no FP16 inner precision
    b3 = W.Field(lat, EVEN, W.F64, 3)
    with pytest.raises(W.WilsonError) as e:
        W.CG(D, 1e-10, precision="f16").apply(b3, W.Field(lat, EVEN, W.F64, 3), raise_on_noconv=True)
    assert e.value.code == W.WL_E_NOTIMPL
    lat.set_option(W.OPT_BLOCK_KRYLOV, 0)
    lat.close()
