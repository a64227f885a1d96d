"""GPU tests at the benchmark workloads themselves, through the C ABI.

- Plain method on the GPU: 32^4, links exp(i 0.3 H) (seed 3), kappa 0.125, Gaussian source
  (seed 4).  Both hopping blocks D_oe and D_eo with both signs, M and M^dag against the oracle
  (relative L2 <= 1e-13), and the CG solution against the oracle's CG on the same input.
- The committed golden fixtures (independent dense construction, tests/golden/) read directly:
  M, M^dag, H_oe (both signs), H_eo and the direct-solve CG solution.
- Mixed precision: 48^3x96, links exp(i 0.3 H) (seed 6), kappa 0.1245, source seed 7.  The FP64
  operator against the oracle at this size; FP32-inner and FP16-inner CG with FP64 reliable
  updates to 1e-10, their FP64 true residual recomputed with the FP64 operator, and their
  solutions bounded against the FP64 CG's.
- Batched multi-rhs: 32^4, links seed 5, kappa 0.125, 12 point sources (one even site, every
  spin-colour pair) + 12 Gaussian, FP32 inner with FP64 reliable updates, 24 rhs in one solve.
- Serving: a fixed seeded draw of requests over {16^4, 24^3x48, 32^4}, kappa in
  [0.120, 0.1245], tolerance in {1e-8, 1e-10, 1e-12}, 1-12 Gaussian rhs, gauge eps in
  [0.2, 0.5]; each solution identical bit for bit alone and in the mix, and each rhs's true
  residual within its tolerance.

The invariants restate applications/test_cases/test_CG.cpp:215-288 (operator parity, CG
round trips) at these sizes.
"""
import glob
import os

import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, ODD, Oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DSLASH_TOL = 1e-13


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(*a):
    print(*a, flush=True)


# ----------------------------------------------------------------- 32^4 plain method ----

@pytest.fixture(scope="module")
def c2():
    dims = (32, 32, 32, 32)
    say("32^4 FP64 setup", dims)
    U = W.gen_gauge(dims, "smooth", 0.3, 3)
    lat = W.Lattice(dims, antiperiodic_t=True)
    lat.upload_gauge(U)
    yield dims, lat, U, Oracle(dims, antiperiodic_t=True)
    lat.close()


def test_c2_both_hopping_blocks_both_signs(c2):
    """Dirac_Wilson_hop_set (wilson.h:53-95) from even to odd and from odd to even sites,
    sign +1 and -1, at 32^4 against the oracle."""
    dims, lat, U, orc = c2
    kappa = 0.125
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    psi_e = W.gen_spinor(dims, 4, EVEN)
    psi_o = W.gen_spinor(dims, 5, ODD)
    for psi, pin, pout in ((psi_e, EVEN, ODD), (psi_o, ODD, EVEN)):
        fin, fout = W.Field(lat, pin).upload(psi), W.Field(lat, pout)
        for sign in (1, -1):
            say("32^4 FP64 hop_set", pin, "->", pout, "sign", sign)
            D.hop_set(fin, fout, sign)
            err = rel(fout.to_full(), orc.hop_set(U, kappa, psi, pout, sign))
            assert err <= DSLASH_TOL, (pin, sign, err)


def test_c2_apply_and_dagger(c2):
    """Dirac_Wilson_evenodd::apply / dagger (wilson.h:272-289) at 32^4 against the oracle,
    and <a, M b> = <M^dag a, b> (test_CG.cpp:236-245)."""
    dims, lat, U, orc = c2
    kappa = 0.125
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    psi = W.gen_spinor(dims, 4, EVEN)
    fin, fout = W.Field(lat).upload(psi), W.Field(lat)
    for sign in (1, -1):
        say("32^4 FP64 meo sign", sign)
        (D.apply if sign == 1 else D.dagger)(fin, fout)
        err = rel(fout.to_full(), orc.meo(U, kappa, psi, sign))
        assert err <= DSLASH_TOL, (sign, err)
    a = W.gen_spinor(dims, 44, EVEN)
    fa, Mda = W.Field(lat).upload(a), W.Field(lat)
    D.apply(fin, fout)
    D.dagger(fa, Mda)
    c = lambda f: f[..., 0::2] + 1j * f[..., 1::2]
    lhs = np.vdot(c(a[: lat.Vh]), c(fout.download()))
    rhs = np.vdot(c(Mda.download()), c(psi[: lat.Vh]))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_c2_cg_matches_oracle_cg(c2):
    """CG on M^dag M to 1e-10 at 32^4 (conjugate_gradient.h:46-104), against the oracle's
    literal CG on the same input: the same iteration count within one, the same solution
    within the solver tolerance, and the oracle's own true residual of the GPU solution."""
    dims, lat, U, orc = c2
    kappa, tol = 0.125, 1e-10
    b = W.gen_spinor(dims, 4, EVEN)
    fb, fx = W.Field(lat).upload(b), W.Field(lat)
    say("32^4 FP64 GPU CG")
    st = W.CG(W.Dirac_Wilson_evenodd(kappa, lat), tol, 10000).apply(fb, fx)[0]
    assert st["converged"] == 1 and st["true_resid"] <= tol, st
    say("32^4 FP64 oracle CG (about 10 s on the host)")
    x_ref, steps, _ = orc.cg(U, kappa, b, accuracy=tol)
    say("32^4 FP64 GPU", st["iterations"], "it; oracle", steps + 1, "it")
    assert abs(st["iterations"] - (steps + 1)) <= 1
    x = fx.to_full()
    assert rel(x, x_ref) <= 1e-8
    r = b - orc.meo(U, kappa, orc.meo(U, kappa, x, 1), -1)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= tol
    assert abs(np.linalg.norm(r) / np.linalg.norm(b) - st["true_resid"]) <= 1e-3 * tol


# ------------------------------------------------------------------- golden fixtures ----

FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures_on_the_gpu(path):
    """The HIP operator against the committed fixtures of the independent dense construction
    (explicit gamma tables, sparse assembly), with no oracle in between."""
    say("golden", os.path.basename(path))
    d = np.load(path)
    dims = tuple(int(v) for v in d["dims"])
    kappa, ap = float(d["kappa"]), bool(int(d["antiperiodic"]))
    lat = W.Lattice(dims, antiperiodic_t=ap)
    lat.upload_gauge(d["U"])
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    fe, fo = W.Field(lat, EVEN).upload(d["psi"]), W.Field(lat, ODD).upload(d["psi_odd"])
    out_e, out_o = W.Field(lat, EVEN), W.Field(lat, ODD)
    D.apply(fe, out_e)
    assert rel(out_e.download(), d["M_plus"]) <= DSLASH_TOL
    D.dagger(fe, out_e)
    assert rel(out_e.download(), d["M_minus"]) <= DSLASH_TOL
    for sign, key in ((1, "H_oe_plus"), (-1, "H_oe_minus")):
        D.hop_set(fe, out_o, sign)
        assert rel(out_o.download(), d[key]) <= DSLASH_TOL, key
    D.hop_set(fo, out_e, 1)
    assert rel(out_e.download(), d["H_eo_plus"]) <= DSLASH_TOL
    x = W.Field(lat, EVEN)
    st = W.CG(D, 1e-13, 10000).apply(fe, x)[0]
    assert st["converged"] == 1
    assert rel(x.download(), d["x_cg"]) <= 1e-10
    lat.close()


# ------------------------------------------------------------- 48^3x96 mixed precision ----

@pytest.fixture(scope="module")
def c3():
    dims = (48, 48, 48, 96)
    say("48^3x96 mixed setup", dims)
    U = W.gen_gauge(dims, "smooth", 0.3, 6)
    lat = W.Lattice(dims, antiperiodic_t=True)
    lat.upload_gauge(U)
    yield dims, lat, U
    lat.close()


def test_c3_fp64_operator_against_oracle(c3):
    """M at 48^3x96 against the oracle: the FP64 operator that the reliable updates and the
    true residuals below rely on."""
    dims, lat, U = c3
    kappa = 0.1245
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    psi = W.gen_spinor(dims, 7, EVEN)
    fin, fout = W.Field(lat).upload(psi), W.Field(lat)
    D.apply(fin, fout)
    say("48^3x96 mixed oracle M (a few seconds on the host)")
    err = rel(fout.to_full(), Oracle(dims).meo(U, kappa, psi, 1))
    say("48^3x96 mixed M rel err", err)
    assert err <= DSLASH_TOL


@pytest.mark.parametrize("precision,delta", [("f32", 1e-3), ("f16", 1e-2)])
def test_c3_mixed_cg_to_fp64_true_residual(c3, precision, delta):
    """FP32- and FP16-inner CG with FP64 reliable updates to 1e-10 at 48^3x96: converged on
    the FP64 true residual, which is recomputed here with the FP64 operator, and within 1e-8
    of the FP64 CG solution (both meet the same tolerance)."""
    dims, lat, U = c3
    kappa, tol = 0.1245, 1e-10
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, 7, EVEN))
    x64, xm = W.Field(lat), W.Field(lat)
    s64 = W.CG(D, tol, 10000).apply(b, x64)[0]
    sm = W.CG(D, tol, 10000, precision=precision, delta=delta).apply(b, xm)[0]
    say("48^3x96 mixed", precision, sm, "FP64:", s64)
    assert s64["converged"] == 1 and sm["converged"] == 1
    assert sm["true_resid"] <= tol and sm["reliable_updates"] >= 1
    Mx, MdMx = W.Field(lat), W.Field(lat)
    D.apply(xm, Mx)
    D.dagger(Mx, MdMx)
    bh = b.download()
    tr = float(np.linalg.norm(bh - MdMx.download()) / np.linalg.norm(bh))
    assert tr <= 1.05 * tol, tr
    d = rel(xm.download(), x64.download())
    say("48^3x96 mixed", precision, "solution vs FP64 CG", d)
    assert d <= 1e-8


# -------------------------------------------------------------- 32^4 batched 24 rhs ----

def even_point_seed(dims, seed):
    Vh = int(np.prod(dims)) // 2
    while True:
        _, site = W.gen_point(dims, seed, 0, 0)
        if site < Vh:
            return seed
        seed += 1


def test_c4_batched_24_rhs():
    """32^4, links exp(i 0.3 H) seed 5, kappa 0.125: 12 point sources at one seeded even site
    (all spin-colour pairs) and 12 Gaussian sources solved together, FP32 inner with FP64
    reliable updates, to 1e-10.  Every rhs meets the FP64 true residual; a point and a
    Gaussian rhs are checked with the oracle's M^dag M and against single-rhs FP64 CG."""
    dims, kappa, N, tol = (32, 32, 32, 32), 0.125, 24, 1e-10
    say("32^4 24-rhs setup")
    U = W.gen_gauge(dims, "smooth", 0.3, 5)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    b = W.Field(lat, EVEN, W.F64, nrhs=N)
    pseed = even_point_seed(dims, 8)
    srcs = []
    for sc in range(12):
        pt, _ = W.gen_point(dims, pseed, sc // 3, sc % 3)
        srcs.append(pt)
    for j in range(12):
        srcs.append(W.gen_spinor(dims, 100 + j, EVEN))
    for j, s in enumerate(srcs):
        b.upload(s, rhs=j)
    x = W.Field(lat, EVEN, W.F64, nrhs=N)
    say("32^4 24-rhs batched solve")
    st = W.CG(D, tol, 10000, precision="f32", delta=1e-3, update_every=100).apply(b, x)
    assert len(st) == N
    for j, s in enumerate(st):
        assert s["converged"] == 1 and s["true_resid"] <= tol, (j, s)
    orc = Oracle(dims)
    b1, x1 = W.Field(lat), W.Field(lat)
    for j in (0, 12):
        say("32^4 24-rhs check rhs", j)
        xj = x.to_full(j)
        r = srcs[j] - orc.meo(U, kappa, orc.meo(U, kappa, xj, 1), -1)
        tr = float(np.linalg.norm(r) / np.linalg.norm(srcs[j]))
        assert tr <= 1.05 * tol and abs(tr - st[j]["true_resid"]) <= 0.05 * tol, (j, tr)
        b1.upload(srcs[j])
        x1.zero()
        s64 = W.CG(D, tol, 10000).apply(b1, x1)[0]
        assert s64["converged"] == 1
        assert rel(x.download(j), x1.download()) <= 1e-8
    # the same 24 rhs by block CG (one shared Krylov space), FP32 inner, FP64 reliable updates
    say("32^4 24-rhs block solve")
    lat.set_option(W.OPT_BLOCK_KRYLOV, 1)
    xb = W.Field(lat, EVEN, W.F64, nrhs=N)
    sb = W.CG(D, tol, 10000, precision="f32", delta=1e-3, update_every=100).apply(b, xb)
    lat.set_option(W.OPT_BLOCK_KRYLOV, 0)
    say("32^4 24-rhs block", sb[0]["iterations"], "iterations, batched", max(q["iterations"] for q in st))
    for j, q in enumerate(sb):
        assert q["converged"] == 1 and q["true_resid"] <= tol, (j, q)
        assert rel(xb.download(j), x.download(j)) <= 1e-8, j
    assert sb[0]["iterations"] < max(q["iterations"] for q in st)
    xj = xb.to_full(12)
    r = srcs[12] - orc.meo(U, kappa, orc.meo(U, kappa, xj, 1), -1)
    assert float(np.linalg.norm(r) / np.linalg.norm(srcs[12])) <= 1.05 * tol
    lat.close()


# ----------------------------------------------------------------- served request mix ----

SERVE_VOLUMES = ((16, 16, 16, 16), (24, 24, 24, 48), (32, 32, 32, 32))


def draw_requests(n, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dims = SERVE_VOLUMES[i % 3] if i < 3 else SERVE_VOLUMES[int(rng.integers(0, 3))]
        out.append(dict(dims=dims,
                        kappa=float(rng.uniform(0.120, 0.1245)),
                        tol=float((1e-8, 1e-10, 1e-12)[int(rng.integers(0, 3))]),
                        nrhs=int(rng.integers(1, 13)),
                        eps=float(rng.uniform(0.2, 0.5)),
                        seed=int(rng.integers(1, 1 << 30)),
                        prec=("f16", "f32", "f64")[i % 3]))
    return out


class ServedRequest:
    def __init__(self, d):
        self.__dict__.update(d)
        self.Vh = int(np.prod(self.dims)) // 2
        self.U = W.gen_gauge(self.dims, "smooth", self.eps, self.seed)
        self.src = np.ascontiguousarray(np.stack(
            [W.gen_spinor(self.dims, self.seed + 7 * j, EVEN)[: self.Vh] for j in range(self.nrhs)]))

    def submit(self, S):
        sol = np.zeros_like(self.src)
        t = S.submit(self.dims, self.U, self.src, sol, self.kappa, self.tol, 10000,
                     precision=self.prec, delta=1e-2 if self.prec == "f16" else 1e-3)
        return t, sol


def test_c5_served_mix_alone_and_together():
    """Requests at the serving workload's volumes: each solution is identical to the last bit
    whether the request runs alone or among the others (twice over, interleaved), every rhs
    converges with its reported FP64 true residual within its tolerance, and the oracle's
    M^dag M confirms the first and last rhs of each request."""
    reqs = [ServedRequest(d) for d in draw_requests(6)]
    alone = []
    for i, r in enumerate(reqs):
        say("served mix alone", i, r.dims, r.prec, "nrhs", r.nrhs, "tol", r.tol)
        with W.Server(max_active=1) as S:
            t, sol = r.submit(S)
            res, st = S.wait(t)
        assert res["status"] == W.WL_OK, res
        assert len(st) == r.nrhs and all(s["converged"] for s in st)
        lim = 20 * r.tol if r.prec == "f64" else r.tol
        assert all(s["true_resid"] <= lim for s in st), (i, [s["true_resid"] for s in st])
        alone.append(sol)
    order = list(range(len(reqs))) * 2
    order = order[::2] + order[1::2][::-1]
    say("served mix mix", order)
    with W.Server(max_active=8) as S:
        sub = [(i,) + reqs[i].submit(S) for i in order]
        for i, t, sol in sub:
            res, _ = S.wait(t)
            assert res["status"] == W.WL_OK, res
            assert np.array_equal(sol, alone[i]), i
        info = S.info()
    assert info["completed"] == len(order) and info["active"] == 0
    for i, r in enumerate(reqs):
        orc = Oracle(r.dims)
        V = 2 * r.Vh
        lim = (20 if r.prec == "f64" else 1.05) * r.tol
        for j in sorted({0, r.nrhs - 1}):
            say("served mix oracle check", i, j)
            x, bb = np.zeros((V, 24)), np.zeros((V, 24))
            x[: r.Vh], bb[: r.Vh] = alone[i][j], r.src[j]
            y = orc.meo(r.U, r.kappa, orc.meo(r.U, r.kappa, x, 1), -1)
            assert np.linalg.norm(y - bb) / np.linalg.norm(bb) <= lim, (i, j)
