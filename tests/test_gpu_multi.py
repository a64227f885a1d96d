"""Multi-right-hand-side operator and batched CG on the GPU.

Each right-hand side of a multi-rhs application performs the same arithmetic as the single-rhs
kernel (the compiler may fuse multiply-adds differently, so agreement is to rounding: 1e-14 in
FP64, 1e-6 in FP32) and matches the FP64 oracle to 1e-13; batched CG runs one HILA CG
recurrence per right-hand side and is compared with single-rhs solves."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, ODD, Oracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(dims, kind="smooth", eps=0.3, seed=5):
    U = W.gen_gauge(dims, kind, eps, seed)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    return lat, U


@pytest.mark.parametrize("N", [2, 3, 5, 12])
@pytest.mark.parametrize("prec", [W.F64, W.F32], ids=["f64", "f32"])
def test_multi_rhs_operator_equals_single_rhs(N, prec):
    dims = (8, 8, 8, 8)
    lat, U = _setup(dims)
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    srcs = [W.gen_spinor(dims, 300 + j, EVEN) for j in range(N)]
    fin = W.Field(lat, EVEN, prec, nrhs=N)
    for j, s in enumerate(srcs):
        fin.upload(s, rhs=j)
    fo, fe = W.Field(lat, ODD, prec, nrhs=N), W.Field(lat, EVEN, prec, nrhs=N)
    one_in, one_o, one_e = W.Field(lat, EVEN, prec), W.Field(lat, ODD, prec), W.Field(lat, EVEN, prec)
    orc = Oracle(dims)
    for sign in (1, -1):
        D.hop_set(fin, fo, sign)
        (D.apply if sign == 1 else D.dagger)(fin, fe)
        for j, s in enumerate(srcs):
            one_in.upload(s)
            D.hop_set(one_in, one_o, sign)
            (D.apply if sign == 1 else D.dagger)(one_in, one_e)
            tol = 1e-14 if prec == W.F64 else 1e-6
            assert rel(fo.download(j), one_o.download()) <= tol, (j, sign)
            assert rel(fe.download(j), one_e.download()) <= tol, (j, sign)
        if prec == W.F64:
            for j in (0, N - 1):
                ref = orc.meo(U, 0.125, srcs[j], sign)
                assert rel(fe.to_full(j), ref) <= 1e-13
                ref = orc.hop_set(U, 0.125, srcs[j], ODD, sign)
                assert rel(fo.to_full(j), ref) <= 1e-13


@pytest.mark.parametrize("dims", [(6, 10, 14, 12), (24, 24, 24, 48)], ids=["6x10x14x12", "24x24x24x48"])
def test_multi_rhs_operator_on_lattices_not_powers_of_two(dims):
    """The multi-rhs kernel finds a site's coordinates by multiply-high division by the x, y
    and z sizes; sizes 6, 10, 14 and 24 (not powers of two) check that division against the
    oracle's neighbours, at every rhs-per-lane setting.  6x10x14x12 also ends in a partly
    filled wave (2520 sites per parity); 24^3x48 runs the (z,t)-chunked traversal."""
    lat, U = _setup(dims)
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    orc = Oracle(dims)
    N = 3
    srcs = [W.gen_spinor(dims, 600 + j, EVEN) for j in range(N)]
    refs = {sign: [orc.hop_set(U, 0.125, s, ODD, sign) for s in srcs] for sign in (1, -1)}
    for prec, per_lane, tol in ((W.F64, (1, 2), 1e-13), (W.F32, (1, 2, 4), 2e-6), (W.F16, (1, 2, 4), 2e-3)):
        fin = W.Field(lat, EVEN, prec, nrhs=N)
        for j, s in enumerate(srcs):
            fin.upload(s, rhs=j)
        fo = W.Field(lat, ODD, prec, nrhs=N)
        for nr in per_lane:
            lat.set_option(W.OPT_RHS_BLOCK, nr)
            for sign in (1, -1):
                print(f"multi-rhs {dims} prec {prec} per lane {nr} sign {sign}", flush=True)
                D.hop_set(fin, fo, sign)
                for j in range(N):
                    assert rel(fo.to_full(j), refs[sign][j]) <= tol, (prec, nr, sign, j)
        lat.set_option(W.OPT_RHS_BLOCK, 0)
        fin.close()
        fo.close()
    lat.close()


def test_rhs_block_and_workgroup_options_do_not_change_results():
    dims, N = (16, 16, 16, 16), 12
    lat, U = _setup(dims)
    del U
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    fin = W.Field(lat, EVEN, W.F32, nrhs=N)
    for j in range(N):
        fin.upload(W.gen_spinor(dims, 400 + j, EVEN), rhs=j)
    fe = W.Field(lat, EVEN, W.F32, nrhs=N)
    ref = None
    for nb in (1, 2, 3, 4):
        for blk in (64, 256):
            lat.set_option(W.OPT_RHS_BLOCK, nb)
            lat.set_option(W.OPT_BLOCK, blk)
            D.apply(fin, fe)
            out = [fe.download(j) for j in range(N)]
            if ref is None:
                ref = out
            assert all(rel(a, b) <= 1e-6 for a, b in zip(out, ref)), (nb, blk)
    with pytest.raises(W.WilsonError):
        lat.set_option(W.OPT_RHS_BLOCK, 5)


@pytest.mark.parametrize("prec,N", [(W.F64, 3), (W.F32, 5)], ids=["f64", "f32"])
def test_occupancy_cap_and_lds_reserve_do_not_change_results(prec, N):
    """2^17 sites per parity: FP64 and FP32 multi-rhs hops take the automatic occupancy cap
    (64-lane workgroups with 32 / 24 KB of LDS each); no reserve, an explicit one and a chosen
    workgroup size give the same bits."""
    dims = (16, 16, 32, 32)
    lat, U = _setup(dims)
    del U
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    fin = W.Field(lat, EVEN, prec, nrhs=N)
    for j in range(N):
        fin.upload(W.gen_spinor(dims, 450 + j, EVEN), rhs=j)
    fe = W.Field(lat, EVEN, prec, nrhs=N)
    assert lat.get_option(W.OPT_LDS_RESERVE) == 0
    ref = None
    for lds in (0, -1, 32768, 0):
        for blk in (0, 128):
            lat.set_option(W.OPT_LDS_RESERVE, lds)
            lat.set_option(W.OPT_BLOCK, blk)
            assert lat.get_option(W.OPT_LDS_RESERVE) == lds
            D.apply(fin, fe)
            out = fe.download().tobytes()
            if ref is None:
                ref = out
            assert out == ref, (lds, blk)
    for bad in (-2, 65537):
        with pytest.raises(W.WilsonError):
            lat.set_option(W.OPT_LDS_RESERVE, bad)
    lat.close()


def test_autotune_remembers_its_choice_including_the_lds_reserve():
    dims, N = (8, 8, 8, 8), 2
    lat, U = _setup(dims)
    del U
    first = lat.autotune(W.F64, N)
    assert first["candidates"] > 0 and first["lds_reserve"] in (0, -1)
    assert first["us_per_rhs"] <= first["us_per_rhs_default"]   # the default is kept unless beaten
    assert lat.get_option(W.OPT_LDS_RESERVE) == first["lds_reserve"]
    assert lat.get_option(W.OPT_BLOCK) == first["block"] or first["block"] == 0
    lat.set_option(W.OPT_LDS_RESERVE, 4096)
    again = lat.autotune(W.F64, N)
    assert again["candidates"] == 0
    assert {k: again[k] for k in ("rhs_block", "rhs_launch", "block", "lds_reserve")} == \
        {k: first[k] for k in ("rhs_block", "rhs_launch", "block", "lds_reserve")}
    assert lat.get_option(W.OPT_LDS_RESERVE) == first["lds_reserve"]
    lat.close()


def test_batched_f64_cg_matches_single_rhs_solves():
    dims, N = (8, 8, 8, 8), 4
    lat, U = _setup(dims, "haar", 0, 1)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    srcs = [W.gen_spinor(dims, 500 + j, EVEN) for j in range(N)]
    B, X = W.Field(lat, nrhs=N), W.Field(lat, nrhs=N)
    for j, s in enumerate(srcs):
        B.upload(s, rhs=j)
    stats = W.CG(D, 1e-10).apply(B, X)
    b1, x1 = W.Field(lat), W.Field(lat)
    for j, s in enumerate(srcs):
        b1.upload(s)
        x1.zero()
        st1 = W.CG(D, 1e-10).apply(b1, x1)[0]
        assert stats[j]["converged"] == 1 and stats[j]["true_resid"] <= 1e-10
        assert abs(stats[j]["iterations"] - st1["iterations"]) <= 1
        assert rel(X.download(j), x1.download()) <= 1e-9


def test_batched_mixed_cg_point_and_gaussian_sources():
    """Batched multi-rhs layout: point sources (one per spin-colour pair) and Gaussian sources
    on one gauge field, FP32 inner solver with FP64 reliable updates, to 1e-10."""
    dims = (16, 16, 16, 16)
    lat, U = _setup(dims, "smooth", 0.3, 5)
    del U
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    srcs = []
    for sc in range(12):
        f, site = W.gen_point(dims, 9, sc // 3, sc % 3)
        if site >= lat.Vh:  # keep the point on an even site: shift the seed until it is
            for seed in range(10, 200):
                f, site = W.gen_point(dims, seed, sc // 3, sc % 3)
                if site < lat.Vh:
                    break
        srcs.append(f)
    srcs += [W.gen_spinor(dims, 600 + j, EVEN) for j in range(4)]
    N = len(srcs)
    B, X = W.Field(lat, nrhs=N), W.Field(lat, nrhs=N)
    for j, s in enumerate(srcs):
        B.upload(s, rhs=j)
    stats = W.CG(D, 1e-10, precision="f32").apply(B, X)
    for j, st in enumerate(stats):
        assert st["converged"] == 1 and st["true_resid"] <= 1e-10, (j, st)
    # spot-check two of them against single-rhs FP64 solves
    b1, x1 = W.Field(lat), W.Field(lat)
    for j in (0, N - 1):
        b1.upload(srcs[j])
        x1.zero()
        W.CG(D, 1e-10).apply(b1, x1)
        assert rel(X.download(j), x1.download()) <= 1e-8


def test_batched_cg_zero_source_and_maxiters():
    dims, N = (8, 8, 8, 8), 3
    lat, U = _setup(dims, "haar", 0, 1)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    B, X = W.Field(lat, nrhs=N), W.Field(lat, nrhs=N)
    B.upload(W.gen_spinor(dims, 700, EVEN), rhs=0)
    B.upload(W.gen_spinor(dims, 701, EVEN), rhs=2)   # rhs 1 stays zero
    X.upload(W.gen_spinor(dims, 702, EVEN), rhs=1)   # nonzero initial guess for the zero rhs
    for prec in ("f64", "f32"):
        stats = W.CG(D, 1e-10, precision=prec).apply(B, X)
        assert stats[1]["converged"] == 1 and np.all(X.download(1) == 0)
        assert stats[0]["converged"] == 1 and stats[2]["converged"] == 1
    X.zero()
    stats = W.CG(D, 1e-12, maxiters=5).apply(B, X)
    assert stats[0]["converged"] == 0 and stats[0]["iterations"] == 5


@pytest.mark.parametrize("precision", ["f64", "f32", "f16"])
def test_graph_replayed_chunks_equal_eager_launches(precision):
    """Each chunk of the batched CG is captured once per solve into a HIP graph and replayed
    every chunk; solutions and statistics must equal the kernel-by-kernel run exactly, on
    repeated solves with eager work in between."""
    dims, N = (8, 8, 8, 8), 3
    lat, U = _setup(dims)
    D = W.Dirac_Wilson_evenodd(0.124, lat)
    b = W.Field(lat, EVEN, W.F64, nrhs=N)
    for j in range(N):
        b.upload(W.gen_spinor(dims, 500 + j, EVEN), rhs=j)
    cg = W.CG(D, 1e-10, 10000, precision=precision, delta=0.1)
    x = W.Field(lat, EVEN, W.F64, nrhs=N)
    st_eager = cg.apply(b, x)
    eager = [x.download(j) for j in range(N)]
    lat.set_option(W.OPT_GRAPHS, 1)
    tmp = W.Field(lat, EVEN, W.F64, nrhs=N)
    for rep in range(3):
        D.apply(b, tmp)          # eager work between the graph runs
        lat.sync()
        x.zero()
        st = cg.apply(b, x)
        for j in range(N):
            assert np.array_equal(x.download(j), eager[j]), (rep, j)
            assert st[j]["iterations"] == st_eager[j]["iterations"]
            assert st[j]["true_resid"] == st_eager[j]["true_resid"]
    lat.set_option(W.OPT_GRAPHS, 0)


@pytest.mark.parametrize("prec", [W.F64, W.F32, W.F16], ids=["f64", "f32", "f16"])
def test_link_load_modes_give_the_same_operator(prec):
    """Cached, nontemporal and LDS-staged link loads of the multi-rhs kernel, with and without
    software pipelining, compute the same hopping term (to rounding: FMA contraction may differ
    between instantiations); FP64 also against the oracle."""
    dims, N = (8, 8, 8, 16), 6
    lat, U = _setup(dims)
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    srcs = [W.gen_spinor(dims, 700 + j, EVEN) for j in range(N)]
    fin = W.Field(lat, EVEN, prec, nrhs=N)
    for j, sj in enumerate(srcs):
        fin.upload(sj, rhs=j)
    fo = W.Field(lat, ODD, prec, nrhs=N)
    tol = {W.F64: 1e-14, W.F32: 1e-6, W.F16: 1e-4}[prec]  # FP16: a rounding step may flip a half
    ref = None
    for nr in (1, 2):
        for mode in (1, 2, 3):
            for pipe in (0, 1):
                lat.set_option(W.OPT_RHS_BLOCK, nr)
                lat.set_option(W.OPT_LINK_LOAD, mode)
                lat.set_option(W.OPT_PIPELINE, pipe)
                D.hop_set(fin, fo, -1)
                out = [fo.download(j) for j in range(N)]
                if ref is None:
                    ref = out
                for j in range(N):
                    assert rel(out[j], ref[j]) <= tol, (nr, mode, pipe, j)
    lat.set_option(W.OPT_LINK_LOAD, 0)
    lat.set_option(W.OPT_PIPELINE, 0)
    if prec == W.F64:
        orc = Oracle(dims)
        for j in (0, N - 1):
            assert rel(fo.to_full(j), orc.hop_set(U, 0.125, srcs[j], ODD, -1)) <= 1e-13
