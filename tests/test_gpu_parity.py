"""GPU parity of the HIP Wilson path against the CPU oracle (C restatement of HILA's
wilson.h / conjugate_gradient.h) on identical seeded inputs, through the C ABI.

Bar: relative L2 error <= 1e-13 for the FP64 hopping term and M / M^dag; CG solutions reach a
recomputed FP64 true residual <= tolerance and agree with the oracle CG.
"""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, ODD, Oracle

pytestmark = pytest.mark.gpu

DSLASH_TOL = 1e-13

# (dims, link kind, eps, gauge seed, kappa, antiperiodic)
SMALL_CASES = [
    ((2, 2, 2, 2), "haar", 0.0, 21, 0.12, True),
    ((4, 4, 4, 4), "haar", 0.0, 11, 0.12, True),
    ((4, 6, 8, 10), "smooth", 0.3, 7, 0.125, True),
    ((8, 4, 2, 6), "smooth", 0.5, 9, 0.11, False),
    ((8, 8, 8, 8), "haar", 0.0, 1, 0.12, True),   # plain method, CPU-runnable config
]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _setup(dims, kind, eps, seed, antiperiodic):
    U = W.gen_gauge(dims, kind, eps, seed)
    lat = W.Lattice(dims, antiperiodic_t=antiperiodic)
    lat.upload_gauge(U)
    return lat, U, Oracle(dims, antiperiodic_t=antiperiodic)


@pytest.mark.parametrize("case", SMALL_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_hop_set_both_parities(case):
    dims, kind, eps, seed, kappa, ap = case
    print("hop_set", case, flush=True)
    lat, U, orc = _setup(dims, kind, eps, seed, ap)
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    for in_par, out_par in ((EVEN, ODD), (ODD, EVEN)):
        psi = W.gen_spinor(dims, 100 + seed, in_par)
        fin = W.Field(lat, in_par).upload(psi)
        fout = W.Field(lat, out_par)
        for sign in (1, -1):
            D.hop_set(fin, fout, sign)
            got = fout.to_full()
            ref = orc.hop_set(U, kappa, psi, out_par, sign)
            assert rel(got, ref) <= DSLASH_TOL, (in_par, sign, rel(got, ref))


@pytest.mark.parametrize("case", SMALL_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_meo_apply_dagger(case):
    dims, kind, eps, seed, kappa, ap = case
    print("meo", case, flush=True)
    lat, U, orc = _setup(dims, kind, eps, seed, ap)
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    psi = W.gen_spinor(dims, 200 + seed, EVEN)
    fin = W.Field(lat, EVEN).upload(psi)
    fout = W.Field(lat, EVEN)
    D.apply(fin, fout)
    assert rel(fout.to_full(), orc.meo(U, kappa, psi, 1)) <= DSLASH_TOL
    D.dagger(fin, fout)
    assert rel(fout.to_full(), orc.meo(U, kappa, psi, -1)) <= DSLASH_TOL


def test_adjointness_gpu():
    """<a, M b> = <M^dag a, b> (applications/test_cases/test_CG.cpp:236-245)."""
    dims = (8, 8, 8, 8)
    lat, U, _ = _setup(dims, "haar", 0, 1, True)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    a = W.gen_spinor(dims, 31, EVEN)
    b = W.gen_spinor(dims, 32, EVEN)
    fa, fb = W.Field(lat).upload(a), W.Field(lat).upload(b)
    Mb, Mda = W.Field(lat), W.Field(lat)
    D.apply(fb, Mb)
    D.dagger(fa, Mda)
    c = lambda f: f[..., 0::2] + 1j * f[..., 1::2]
    lhs = np.vdot(c(a[: lat.Vh]), c(Mb.download()))
    rhs = np.vdot(c(Mda.download()), c(b[: lat.Vh]))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_norm2_matches_compensated_sum():
    dims = (8, 8, 8, 8)
    lat = W.Lattice(dims)
    psi = W.gen_spinor(dims, 5, EVEN)
    f = W.Field(lat).upload(psi)
    import math
    ref = math.fsum((psi[: lat.Vh] ** 2).ravel())
    assert abs(f.norm2() - ref) <= 1e-13 * ref


def test_plain_gpu_config_32_dslash():
    """Plain method on GPU: 32^4, links exp(i 0.3 H) (seed 3), kappa 0.125, Gaussian source
    (seed 4): FP64 hopping term and M against the oracle."""
    dims = (32, 32, 32, 32)
    print("32^4 parity", flush=True)
    lat, U, orc = _setup(dims, "smooth", 0.3, 3, True)
    kappa = 0.125
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    psi = W.gen_spinor(dims, 4, EVEN)
    fin = W.Field(lat, EVEN).upload(psi)
    fodd = W.Field(lat, ODD)
    D.hop_set(fin, fodd, 1)
    assert rel(fodd.to_full(), orc.hop_set(U, kappa, psi, ODD, 1)) <= DSLASH_TOL
    fout = W.Field(lat, EVEN)
    D.apply(fin, fout)
    assert rel(fout.to_full(), orc.meo(U, kappa, psi, 1)) <= DSLASH_TOL


def test_cg_8_matches_oracle():
    """Plain method, CPU-runnable config: 8^4 Haar links (seed 1), kappa 0.12, source seed 2,
    CG on M^dag M to 1e-10; compared with the oracle CG run on the same input."""
    dims = (8, 8, 8, 8)
    lat, U, orc = _setup(dims, "haar", 0, 1, True)
    kappa, tol = 0.12, 1e-10
    b = W.gen_spinor(dims, 2, EVEN)
    fb = W.Field(lat).upload(b)
    fx = W.Field(lat)
    st = W.CG(W.Dirac_Wilson_evenodd(kappa, lat), tol, 10000).apply(fb, fx)[0]
    x_ref, steps, rr_ref = orc.cg(U, kappa, b, accuracy=tol)
    assert st["converged"] == 1
    assert st["true_resid"] <= tol
    assert abs(st["iterations"] - (steps + 1)) <= 1
    x = fx.to_full()
    assert rel(x, x_ref) <= 1e-8
    # independent true residual on the CPU oracle
    r = b - orc.meo(U, kappa, orc.meo(U, kappa, x, 1), -1)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= tol


def test_cg_round_trips():
    """test_CG.cpp:247-287: M (M^dag M)^-1 M^dag a = a and (M^dag M)(M^dag M)^-1 a = a."""
    dims = (4, 4, 4, 8)
    lat, U, orc = _setup(dims, "smooth", 0.4, 17, True)
    kappa = 0.12
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    cg = W.CG(D, 1e-12, 10000)
    a = W.gen_spinor(dims, 41, EVEN)
    fa = W.Field(lat).upload(a)
    Mda, x, Mx = W.Field(lat), W.Field(lat), W.Field(lat)
    D.dagger(fa, Mda)
    cg.apply(Mda, x)
    D.apply(x, Mx)
    assert rel(Mx.download(), a[: lat.Vh]) <= 1e-10
    x.zero()
    cg.apply(fa, x)
    MdMx = W.Field(lat)
    D.apply(x, Mx)
    D.dagger(Mx, MdMx)
    assert rel(MdMx.download(), a[: lat.Vh]) <= 1e-11


def test_cg_initial_guess_and_zero_source():
    dims = (4, 4, 4, 4)
    lat, U, orc = _setup(dims, "haar", 0, 3, True)
    kappa = 0.12
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    b = W.gen_spinor(dims, 8, EVEN)
    fb, fx = W.Field(lat).upload(b), W.Field(lat)
    st1 = W.CG(D, 1e-10).apply(fb, fx)[0]
    # restarting from the converged solution needs (almost) no iterations
    st2 = W.CG(D, 1e-10).apply(fb, fx)[0]
    assert st2["iterations"] <= 2 and st2["true_resid"] <= 1e-10
    assert st1["iterations"] > st2["iterations"]
    # b = 0 gives x = 0
    fz = W.Field(lat)
    st3 = W.CG(D, 1e-10).apply(fz, fx)[0]
    assert st3["converged"] == 1 and np.all(fx.download() == 0)


def test_cg_maxiters_reports_noconv():
    dims = (4, 4, 4, 4)
    lat, U, _ = _setup(dims, "haar", 0, 3, True)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    fb, fx = W.Field(lat).upload(W.gen_spinor(dims, 8, EVEN)), W.Field(lat)
    st = W.CG(D, 1e-14, 3).apply(fb, fx)[0]
    assert st["converged"] == 0 and st["iterations"] == 3
    with pytest.raises(W.WilsonError):
        W.CG(D, 1e-14, 3).apply(fb, fx, raise_on_noconv=True)


def test_argument_errors():
    dims = (4, 4, 4, 4)
    lat = W.Lattice(dims)
    e, o = W.Field(lat, EVEN), W.Field(lat, ODD)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    with pytest.raises(W.WilsonError):  # no gauge uploaded yet
        D.apply(e, W.Field(lat, EVEN))
    lat.upload_gauge(W.gen_gauge(dims, "unit"))
    with pytest.raises(W.WilsonError):  # same parity
        D.hop_set(e, W.Field(lat, EVEN))
    with pytest.raises(W.WilsonError):  # odd field into M
        D.apply(o, e)
    with pytest.raises(W.WilsonError):
        W.Lattice((4, 4, 4, 5))


def test_timed_copy_copies_and_checks_its_fields():
    """wl_time_op what = 2 (the device-copy bandwidth the hopping term is compared with)."""
    dims = (8, 8, 8, 8)
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "unit"))
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    a = W.Field(lat, EVEN, W.F64, 2)
    for j in range(2):
        a.upload(W.gen_spinor(dims, 40 + j, EVEN), rhs=j)
    b = W.Field(lat, EVEN, W.F64, 2)
    assert D.time(a, b, "copy", warmup=1, iters=3) > 0
    assert a.download().tobytes() == b.download().tobytes()
    with pytest.raises(W.WilsonError):   # rhs counts differ
        D.time(a, W.Field(lat, EVEN), "copy")
    other = W.Lattice(dims)
    with pytest.raises(W.WilsonError):   # a field of another context
        D.time(a, W.Field(other, EVEN, W.F64, 2), "copy")
    other.close()
    lat.close()


@pytest.mark.parametrize("dims", [(32, 32, 32, 32), (24, 24, 24, 48), (16, 16, 16, 16)],
                         ids=lambda d: "x".join(map(str, d)))
def test_traversal_and_workgroup_options_do_not_change_results(dims):
    """The hopping kernel's site traversal (memory order or (chunk, t)-tiled) and workgroup
    size change only the speed: the output is bitwise identical."""
    lat, U, _ = _setup(dims, "smooth", 0.3, 3, True)
    del U
    lat.set_option(W.OPT_SMALL_HOP, -1)  # the one-lane kernel, whose traversal these options set
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    psi = W.gen_spinor(dims, 4, EVEN)
    fin, fo, fe = W.Field(lat).upload(psi), W.Field(lat, ODD), W.Field(lat)
    sps = dims[0] * dims[1] * dims[2] // 2 // 64
    ref_o = ref_e = None
    for blk in (64, 128, 256):
        for qc in sorted({0, -1} | {q for q in (4, 27, 32, 128) if sps % q == 0}):
            lat.set_option(W.OPT_TRAVERSAL_CHUNK, qc)
            lat.set_option(W.OPT_BLOCK, blk)
            D.hop_set(fin, fo, 1)
            D.apply(fin, fe)
            o, e = fo.download(), fe.download()
            if ref_o is None:
                ref_o, ref_e = o, e
            assert np.array_equal(o, ref_o) and np.array_equal(e, ref_e), (blk, qc)
    with pytest.raises(W.WilsonError):
        lat.set_option(W.OPT_BLOCK, 96)


def test_plane_padding_changes_no_bits():
    """Block padding (WL_OPT_PLANE_PAD) moves addresses only.  M in every storage format, with
    1 and 3 rhs, and the solutions of FP64 CG, batched FP16-inner CG and block CG are
    bitwise identical with no padding and with 192 sites.  The default pads 128 sites exactly
    where the unpadded stride is a multiple of 2^17 sites."""
    for dims, want in (((32, 32, 32, 32), 128), ((24, 24, 24, 48), 0), ((16, 16, 16, 16), 0)):
        lat = W.Lattice(dims)
        assert lat.get_option(W.OPT_PLANE_PAD) == want, dims
        lat.close()
    dims = (8, 8, 8, 8)
    print("plane padding", dims, flush=True)
    U = W.gen_gauge(dims, "smooth", 0.3, 5)
    srcs = [W.gen_spinor(dims, 70 + j, EVEN) for j in range(3)]
    outs = {}
    for pad in (0, 192):
        lat = W.Lattice(dims)
        lat.set_option(W.OPT_PLANE_PAD, pad)
        lat.upload_gauge(U)
        assert lat.get_option(W.OPT_PLANE_PAD) == pad
        D = W.Dirac_Wilson_evenodd(0.125, lat)
        res = []
        for prec in (W.F64, W.F32, W.F16):
            for N in (1, 3):
                fin, fe = W.Field(lat, EVEN, prec, N), W.Field(lat, EVEN, prec, N)
                for j in range(N):
                    fin.upload(srcs[j], rhs=j)
                D.apply(fin, fe)
                res += [fe.download(j) for j in range(N)]
        for precision, block, N in (("f64", False, 1), ("f16", False, 3), ("f64", True, 3)):
            lat.set_option(W.OPT_BLOCK_KRYLOV, 1 if block else 0)
            b, x = W.Field(lat, EVEN, W.F64, N), W.Field(lat, EVEN, W.F64, N)
            for j in range(N):
                b.upload(srcs[j], rhs=j)
            st = W.CG(D, 1e-10, 1000, precision=precision, delta=1e-2).apply(b, x)
            assert all(s["converged"] == 1 for s in st), (precision, block, st)
            res += [x.download(j) for j in range(N)]
        lat.set_option(W.OPT_BLOCK_KRYLOV, 0)
        outs[pad] = res
        lat.close()
    assert len(outs[0]) == len(outs[192])
    for i, (a, b) in enumerate(zip(outs[0], outs[192])):
        assert np.array_equal(a, b), i


def test_field_outliving_its_context_is_orphaned():
    """wl_destroy releases the device memory of fields still alive and orphans the handles:
    calls on them return WL_E_STATE, and destroying them afterwards is safe."""
    lat = W.Lattice((4, 4, 4, 4))
    f = W.Field(lat)
    f.upload(W.gen_spinor((4, 4, 4, 4), 1, EVEN))
    lat.close()
    with pytest.raises(W.WilsonError) as e:
        f.download()
    assert e.value.code == W.WL_E_STATE
    f.close()


@pytest.mark.parametrize("dims", [(6, 6, 6, 10), (10, 10, 10, 12), (20, 20, 20, 20)],
                         ids=lambda d: "x".join(map(str, d)))
def test_workgroup_counts_that_are_not_multiples_of_eight(dims):
    """Lattices whose hopping launches have a workgroup count that is not a multiple of 8
    (the XCD-aware workgroup map then gives some XCDs one tile more): single- and multi-rhs
    operator against the oracle, for each storage format."""
    print("odd workgroup counts", dims, flush=True)
    lat, U, orc = _setup(dims, "smooth", 0.3, 61, True)
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    srcs = [W.gen_spinor(dims, 620 + j, EVEN) for j in range(3)]
    for prec, tol in ((W.F64, 1e-13), (W.F32, 2e-6), (W.F16, 2e-3)):
        for N in (1, 3):
            fin, fe = W.Field(lat, EVEN, prec, N), W.Field(lat, EVEN, prec, N)
            for j in range(N):
                fin.upload(srcs[j], rhs=j)
            D.apply(fin, fe)
            for j in range(N):
                assert rel(fe.to_full(j), orc.meo(U, 0.12, srcs[j], 1)) <= tol, (prec, N, j)
    lat.close()
