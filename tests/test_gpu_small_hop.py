"""The small-lattice hopping kernel (WL_OPT_SMALL_HOP: one wave per direction, partial sums
added in LDS) against the oracle and against the one-lane-per-site kernel, in FP64 and FP32,
for every epilogue it serves: D_oe / D_eo of both signs (hop_set), M and M^dag (the second
half adds the input), the Hasenbusch shift, two-row links, and the CG's fused passes
(wilson.h:15-95, :272-289; conjugate_gradient.h:46-104)."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, ODD, Oracle

pytestmark = pytest.mark.gpu

# (dims, link kind, eps, gauge seed, kappa, antiperiodic); 4x6x8x10 and 6^3x10 have parity
# blocks that are not multiples of 64 sites (a partly filled last workgroup)
CASES = [
    ((2, 2, 2, 2), "haar", 0.0, 21, 0.12, True),
    ((4, 6, 8, 10), "smooth", 0.3, 7, 0.125, True),
    ((6, 6, 6, 10), "smooth", 0.3, 61, 0.12, False),
    ((8, 8, 8, 8), "haar", 0.0, 1, 0.12, True),
    ((16, 16, 16, 16), "smooth", 0.3, 3, 0.125, True),
]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_small_hop_matches_oracle_and_one_lane_kernel(case):
    dims, kind, eps, seed, kappa, ap = case
    print("small hop", dims, flush=True)
    U = W.gen_gauge(dims, kind, eps, seed)
    lat = W.Lattice(dims, antiperiodic_t=ap)
    lat.upload_gauge(U)
    assert lat.get_option(W.OPT_SMALL_HOP) == 0
    orc = Oracle(dims, antiperiodic_t=ap)
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    H = W.Hasenbusch_operator(D, 0.07)
    for prec, tol in ((W.F64, 1e-13), (W.F32, 2e-6)):
        for in_par, out_par in ((EVEN, ODD), (ODD, EVEN)):
            psi = W.gen_spinor(dims, 100 + seed, in_par)
            fin = W.Field(lat, in_par, prec).upload(psi)
            fout = W.Field(lat, out_par, prec)
            for sign in (1, -1):
                got = {}
                for mode in (1, -1):
                    lat.set_option(W.OPT_SMALL_HOP, mode)
                    D.hop_set(fin, fout, sign)
                    got[mode] = fout.to_full()
                ref = orc.hop_set(U, kappa, psi, out_par, sign)
                assert rel(got[1], ref) <= tol, (prec, in_par, sign, rel(got[1], ref))
                assert rel(got[1], got[-1]) <= (1e-15 if prec == W.F64 else 1e-6)
        psi = W.gen_spinor(dims, 200 + seed, EVEN)
        fin, fout = W.Field(lat, EVEN, prec).upload(psi), W.Field(lat, EVEN, prec)
        for op, mh in ((D, 0.0), (H, 0.07)):
            for sign in (1, -1):
                lat.set_option(W.OPT_SMALL_HOP, 1)
                (op.apply if sign == 1 else op.dagger)(fin, fout)
                ref = orc.meo(U, kappa, psi, sign, mh)
                assert rel(fout.to_full(), ref) <= tol, (prec, mh, sign)
    lat.set_option(W.OPT_SMALL_HOP, 0)
    lat.close()


@pytest.mark.parametrize("case", CASES[1:], ids=lambda c: "x".join(map(str, c[0])))
def test_small_hop_cg(case):
    """FP64 CG and FP32-inner mixed CG with the small-lattice kernel: FP64 true residual at the
    tolerance, the same iteration count as with one lane per site, and the oracle CG's
    solution."""
    dims, kind, eps, seed, kappa, ap = case
    print("small hop CG", dims, flush=True)
    U = W.gen_gauge(dims, kind, eps, seed)
    lat = W.Lattice(dims, antiperiodic_t=ap)
    lat.upload_gauge(U)
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    src = W.gen_spinor(dims, 300 + seed, EVEN)
    b = W.Field(lat).upload(src)
    sols, its = {}, {}
    for mode in (1, -1):
        lat.set_option(W.OPT_SMALL_HOP, mode)
        x = W.Field(lat)
        st = W.CG(D, 1e-10, 10000).apply(b, x)[0]
        assert st["converged"] == 1 and st["true_resid"] <= 1e-10, (mode, st)
        sols[mode], its[mode] = x.download(), st["iterations"]
    assert abs(its[1] - its[-1]) <= 1
    assert rel(sols[1], sols[-1]) <= 1e-9
    orc = Oracle(dims, antiperiodic_t=ap)
    xr, _, _ = orc.cg(U, kappa, src, accuracy=1e-10)
    assert rel(sols[1], xr[: lat.Vh]) <= 1e-8
    lat.set_option(W.OPT_SMALL_HOP, 1)
    x = W.Field(lat)
    st = W.CG(D, 1e-10, 10000, precision="f32").apply(b, x)[0]
    assert st["converged"] == 1 and st["true_resid"] <= 1e-10
    # two-row links
    lat.set_option(W.OPT_LINK_ROWS, 2)
    x2 = W.Field(lat)
    st2 = W.CG(D, 1e-10, 10000).apply(b, x2)[0]
    assert st2["converged"] == 1 and st2["true_resid"] <= 1e-10
    assert rel(x2.download(), sols[1]) <= 1e-9
    lat.close()
