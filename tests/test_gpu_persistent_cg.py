"""The persistent FP64 CG (one cooperative launch per solve, WL_OPT_CG_PERSISTENT = 1) against
the kernel-by-kernel CG (WL_OPT_CG_PERSISTENT = -1), which the oracle tests pin: the same
solution and statistics, bit for bit, on several lattices (including workgroup counts that are
not multiples of 8), with two-row links, with the Hasenbusch shift, when maxiters ends the
solve, and for a zero source (conjugate_gradient.h:46-104 in both).  The solution's true
residual is confirmed with the oracle."""
import time

import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, Oracle

pytestmark = pytest.mark.gpu

# (dims, link kind, eps, gauge seed, kappa, antiperiodic, source seed)
CASES = [
    ((8, 8, 8, 8), "haar", 0.0, 1, 0.12, True, 2),        # the CPU-runnable configuration
    ((16, 16, 16, 16), "smooth", 0.3, 3, 0.125, True, 4),
    ((6, 6, 6, 10), "smooth", 0.3, 61, 0.12, True, 5),
    ((4, 6, 8, 10), "smooth", 0.5, 9, 0.11, False, 6),
    ((20, 20, 20, 20), "smooth", 0.3, 7, 0.124, True, 8),
]


def solve(lat, op, b, mode, tol=1e-10, maxiters=10000):
    # the one-launch solve runs one lane per site: compare it with the same kernel by kernel
    lat.set_option(W.OPT_SMALL_HOP, -1)
    lat.set_option(W.OPT_CG_PERSISTENT, mode)
    x = W.Field(lat)
    t = time.perf_counter()
    st = W.CG(op, tol, maxiters).apply(b, x)[0]
    ms = 1e3 * (time.perf_counter() - t)
    return x.download(), st, ms


def same(a, b):
    xa, sa, _ = a
    xb, sb, _ = b
    assert np.array_equal(xa, xb)
    for key in ("iterations", "converged", "rr_rel", "true_resid"):
        assert sa[key] == sb[key], (key, sa[key], sb[key])


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_persistent_cg_equals_kernel_by_kernel(case):
    dims, kind, eps, seed, kappa, ap, sseed = case
    print("persistent CG", dims, flush=True)
    U = W.gen_gauge(dims, kind, eps, seed)
    lat = W.Lattice(dims, antiperiodic_t=ap)
    lat.upload_gauge(U)
    assert lat.get_option(W.OPT_CG_PERSISTENT) == 0
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    src = W.gen_spinor(dims, sseed, EVEN)
    b = W.Field(lat).upload(src)
    ref = solve(lat, D, b, -1)
    got = solve(lat, D, b, 1)
    print(f"  {ref[1]['iterations']} it: kernel by kernel {ref[2]:.2f} ms, one launch {got[2]:.2f} ms", flush=True)
    same(ref, got)
    assert got[1]["converged"] == 1 and got[1]["true_resid"] <= 1e-10
    orc = Oracle(dims, antiperiodic_t=ap)
    x = np.zeros_like(src)
    x[: lat.Vh] = got[0]
    r = src - orc.meo(U, kappa, orc.meo(U, kappa, x, 1), -1)
    assert abs(np.linalg.norm(r) / np.linalg.norm(src) - got[1]["true_resid"]) <= 1e-3 * got[1]["true_resid"]
    lat.close()


def test_persistent_cg_two_rows_shift_maxiters_and_zero_source():
    dims = (8, 8, 8, 8)
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "haar", 0.0, 1))
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, 2, EVEN))
    # two-row links
    lat.set_option(W.OPT_LINK_ROWS, 2)
    same(solve(lat, D, b, -1), solve(lat, D, b, 1))
    lat.set_option(W.OPT_LINK_ROWS, 3)
    # the Hasenbusch operator M + mh
    H = W.Hasenbusch_operator(D, 0.1)
    ref, got = solve(lat, H, b, -1), solve(lat, H, b, 1)
    same(ref, got)
    assert got[1]["converged"] == 1
    # maxiters reached: not converged, same x and statistics
    ref, got = solve(lat, D, b, -1, tol=1e-14, maxiters=3), solve(lat, D, b, 1, tol=1e-14, maxiters=3)
    same(ref, got)
    assert got[1]["converged"] == 0 and got[1]["iterations"] == 3
    # b = 0: x = 0
    z = W.Field(lat)
    xz, sz, _ = solve(lat, D, z, 1)
    assert sz["converged"] == 1 and not xz.any()
    # an initial guess: x holds the solution on entry
    lat.set_option(W.OPT_CG_PERSISTENT, 1)
    x = W.Field(lat)
    W.CG(D, 1e-10).apply(b, x)
    st = W.CG(D, 1e-10).apply(b, x)[0]
    assert st["iterations"] <= 2 and st["true_resid"] <= 1e-10
    with pytest.raises(W.WilsonError):
        lat.set_option(W.OPT_CG_PERSISTENT, 2)
    lat.close()


# The launch-folded kernel-by-kernel CG (WL_OPT_CG_PERSISTENT = 0, lattices up to 2^15 sites
# per parity: the alpha and rrnew finish kernels folded into the update pass and into
# p = beta p + r) against the unfolded one (-1), with the small-lattice kernel and with the
# one-lane kernel (128-lane workgroups: the partials are added by 256 virtual lanes), where the
# fold applies and where it does not (20^4).
FOLD_CASES = [
    ((8, 8, 8, 8), "haar", 0.0, 1, 0.12, True, 2),
    ((16, 16, 16, 16), "smooth", 0.3, 3, 0.125, True, 4),
    ((6, 6, 6, 10), "smooth", 0.3, 61, 0.12, True, 5),
    ((16, 16, 16, 32), "smooth", 0.3, 5, 0.124, False, 7),
    ((20, 20, 20, 20), "smooth", 0.3, 7, 0.124, True, 8),
]


def solve_mode(lat, op, b, mode, small_hop, tol=1e-10, maxiters=10000):
    lat.set_option(W.OPT_SMALL_HOP, small_hop)
    lat.set_option(W.OPT_CG_PERSISTENT, mode)
    x = W.Field(lat)
    t = time.perf_counter()
    st = W.CG(op, tol, maxiters).apply(b, x)[0]
    ms = 1e3 * (time.perf_counter() - t)
    lat.set_option(W.OPT_CG_PERSISTENT, 0)
    return x.download(), st, ms


@pytest.mark.parametrize("case", FOLD_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_launch_folded_cg_equals_unfolded(case):
    dims, kind, eps, seed, kappa, ap, sseed = case
    print("launch-folded CG", dims, flush=True)
    lat = W.Lattice(dims, antiperiodic_t=ap)
    lat.upload_gauge(W.gen_gauge(dims, kind, eps, seed))
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, sseed, EVEN))
    for small_hop in (0, -1):
        ref = solve_mode(lat, D, b, -1, small_hop)
        got = solve_mode(lat, D, b, 0, small_hop)
        print(f"  small_hop {small_hop}: {ref[1]['iterations']} it, unfolded {ref[2]:.2f} ms, "
              f"folded {got[2]:.2f} ms", flush=True)
        same(ref, got)
        assert got[1]["converged"] == 1 and got[1]["true_resid"] <= 1e-10
    lat.close()


def test_launch_folded_cg_maxiters_shift_two_rows_and_zero_source():
    dims = (8, 8, 8, 8)
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "haar", 0.0, 1))
    D = W.Dirac_Wilson_evenodd(0.12, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, 2, EVEN))
    # maxiters ends the solve (done = 2 set by p = beta p + r), also inside a chunk of 8
    for maxiters in (5, 8, 13):
        ref = solve_mode(lat, D, b, -1, 0, maxiters=maxiters)
        got = solve_mode(lat, D, b, 0, 0, maxiters=maxiters)
        same(ref, got)
        assert got[1]["converged"] == 0 and got[1]["iterations"] == maxiters
    # two-row links
    lat.set_option(W.OPT_LINK_ROWS, 2)
    same(solve_mode(lat, D, b, -1, 0), solve_mode(lat, D, b, 0, 0))
    lat.set_option(W.OPT_LINK_ROWS, 3)
    # the Hasenbusch operator M + mh
    H = W.Hasenbusch_operator(D, 0.2)
    same(solve_mode(lat, H, b, -1, 0), solve_mode(lat, H, b, 0, 0))
    # b = 0: x = 0
    z = W.Field(lat).zero()
    x0, st0, _ = solve_mode(lat, D, z, 0, 0)
    assert not np.any(x0) and st0["converged"] == 1
    lat.close()
