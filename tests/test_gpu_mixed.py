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


# FP16 storage with a per-site scale: each stored component is within 2^-11 of the site's
# largest one; one application (input, links and output each rounded) stays within this
# relative L2 error of the FP64 oracle.
F16_TOL = 2e-3


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (8, 8, 8, 8), (4, 6, 8, 10)],
                         ids=lambda d: "x".join(map(str, d)))
def test_f16_operator_against_f64_oracle(dims):
    lat, U = _setup(dims, "smooth", 0.3, 5)
    orc = Oracle(dims)
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    psi = W.gen_spinor(dims, 6, EVEN)
    fin = W.Field(lat, EVEN, W.F16).upload(psi)
    fo = W.Field(lat, ODD, W.F16)
    fe = W.Field(lat, EVEN, W.F16)
    for sign in (1, -1):
        D.hop_set(fin, fo, sign)
        err = rel(fo.to_full(), orc.hop_set(U, 0.125, psi, ODD, sign))
        assert 1e-6 < err <= F16_TOL, err  # rounded to FP16, and within its tolerance
        (D.apply if sign == 1 else D.dagger)(fin, fe)
        assert rel(fe.to_full(), orc.meo(U, 0.125, psi, sign)) <= F16_TOL


def test_f16_storage_round_trip_and_norm():
    dims = (8, 8, 8, 8)
    lat, _ = _setup(dims, "unit", 0.0, 1)
    psi = W.gen_spinor(dims, 9, EVEN)[: lat.Vh]
    psi[5] = 0.0                      # a zero site: scale 0, stored exactly
    psi[7] *= 1e-30                   # a tiny site: the scale keeps its relative precision
    f = W.Field(lat, EVEN, W.F16).upload(psi)
    back = f.download()
    site_max = np.abs(psi).max(axis=1, keepdims=True)
    assert np.all(np.abs(back - psi) <= 2.0 ** -11 * site_max * (1 + 1e-6))
    assert not back[5].any()
    assert np.abs(back[7] - psi[7]).max() <= 2.0 ** -11 * np.abs(psi[7]).max()
    assert abs(f.norm2() - float((back ** 2).sum())) <= 1e-6 * float((back ** 2).sum())
    g = W.Field(lat, EVEN, W.F16).copy_from(f)
    assert np.array_equal(g.download(), back)


@pytest.mark.parametrize("nrhs", [1, 3])
def test_f16_mixed_cg_reaches_true_residual(nrhs):
    dims = (8, 8, 8, 8)
    lat, U = _setup(dims, "smooth", 0.3, 5)
    D = W.Dirac_Wilson_evenodd(0.124, lat)
    b = W.Field(lat, EVEN, W.F64, nrhs=nrhs)
    srcs = [W.gen_spinor(dims, 40 + j, EVEN) for j in range(nrhs)]
    for j, sj in enumerate(srcs):
        b.upload(sj, rhs=j)
    x16, x64 = W.Field(lat, EVEN, W.F64, nrhs=nrhs), W.Field(lat, EVEN, W.F64, nrhs=nrhs)
    st16 = W.CG(D, 1e-10, 10000, precision="f16", delta=0.1).apply(b, x16)
    st64 = W.CG(D, 1e-10, 10000, precision="f64").apply(b, x64)
    orc = Oracle(dims)
    for j in range(nrhs):
        assert st16[j]["converged"] and st16[j]["true_resid"] <= 1e-10
        assert st16[j]["reliable_updates"] >= 1
        x = x16.to_full(j)
        r = orc.meo(U, 0.124, orc.meo(U, 0.124, x, 1), -1) - srcs[j]
        assert np.linalg.norm(r) / np.linalg.norm(srcs[j]) <= 1.05e-10
        assert rel(x16.download(j), x64.download(j)) <= 1e-8
