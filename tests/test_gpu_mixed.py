"""Mixed precision on the GPU: FP32 hopping term / M against the FP64 oracle (FP32 tolerance),
and the reliable-update CG (FP32 inner, FP64 solution and reductions) reaching the FP64 true
residual target, with its solution compared with the FP64 CG's."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, ODD, Oracle

pytestmark = pytest.mark.gpu

# FP32 storage and arithmetic: relative L2 error of one application against the FP64 oracle.
F32_TOL = 2e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(dims, kind, eps, seed):
    U = W.gen_gauge(dims, kind, eps, seed)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    return lat, U


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 10)],
                         ids=lambda d: "x".join(map(str, d)))
def test_f32_operator_against_f64_oracle(dims):
    lat, U = _setup(dims, "smooth", 0.3, 5)
    orc = Oracle(dims)
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    psi = W.gen_spinor(dims, 6, EVEN)
    fin = W.Field(lat, EVEN, W.F32).upload(psi)
    fo = W.Field(lat, ODD, W.F32)
    fe = W.Field(lat, EVEN, W.F32)
    for sign in (1, -1):
        D.hop_set(fin, fo, sign)
        assert rel(fo.to_full(), orc.hop_set(U, 0.125, psi, ODD, sign)) <= F32_TOL
    D.apply(fin, fe)
    assert rel(fe.to_full(), orc.meo(U, 0.125, psi, 1)) <= F32_TOL
    D.dagger(fin, fe)
    assert rel(fe.to_full(), orc.meo(U, 0.125, psi, -1)) <= F32_TOL


@pytest.mark.parametrize("case", [
    ((8, 8, 8, 8), "haar", 0.0, 1, 0.12, 2),       # plain-method CPU config
    ((16, 16, 16, 16), "smooth", 0.3, 3, 0.125, 4),
    ((32, 32, 32, 32), "smooth", 0.3, 3, 0.125, 4),  # plain-method GPU config
], ids=["8^4", "16^4", "32^4"])
def test_mixed_cg_reaches_fp64_true_residual(case):
    dims, kind, eps, seed, kappa, sseed = case
    print("mixed cg", dims, flush=True)
    lat, U = _setup(dims, kind, eps, seed)
    del U
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, sseed, EVEN))
    x64, x32 = W.Field(lat), W.Field(lat)
    s64 = W.CG(D, 1e-10).apply(b, x64)[0]
    s32 = W.CG(D, 1e-10, precision="f32").apply(b, x32)[0]
    print(s64, s32, flush=True)
    assert s32["converged"] == 1 and s32["true_resid"] <= 1e-10
    assert s32["reliable_updates"] >= 1
    # both solutions satisfy the same tolerance: they differ by at most ~cond * 1e-10
    assert rel(x32.download(), x64.download()) <= 1e-8
    # the reported true residual is the real one (recomputed here with the FP64 operator)
    Mx, MdMx = W.Field(lat), W.Field(lat)
    D.apply(x32, Mx)
    D.dagger(Mx, MdMx)
    r = b.download() - MdMx.download()
    assert np.linalg.norm(r) / np.linalg.norm(b.download()) <= 1e-10


def test_mixed_solver_with_fp64_inner_tracks_plain_cg():
    """The reliable-update algorithm run with an FP64 inner operator converges like HILA's CG."""
    dims = (8, 8, 8, 8)
    lat, U = _setup(dims, "haar", 0, 1)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, 2, EVEN))
    x1, x2 = W.Field(lat), W.Field(lat)
    s1 = W.CG(D, 1e-10).apply(b, x1)[0]
    s2 = W.CG(D, 1e-10, precision="f64-mixed").apply(b, x2)[0]
    assert s2["converged"] == 1 and s2["true_resid"] <= 1e-10
    assert abs(s2["iterations"] - s1["iterations"]) <= 3
    assert rel(x2.download(), x1.download()) <= 1e-8


def test_mixed_cg_update_every_and_maxiters():
    dims = (8, 8, 8, 8)
    lat, U = _setup(dims, "haar", 0, 1)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, 2, EVEN))
    x = W.Field(lat)
    s = W.CG(D, 1e-10, precision="f32", delta=0.5, update_every=5).apply(b, x)[0]
    assert s["converged"] == 1 and s["true_resid"] <= 1e-10
    assert s["reliable_updates"] >= s["iterations"] // 5
    x.zero()
    s = W.CG(D, 1e-12, maxiters=4, precision="f32").apply(b, x)[0]
    assert s["converged"] == 0 and s["iterations"] <= 4 + 8
