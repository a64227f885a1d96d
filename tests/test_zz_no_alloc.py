"""wl_cg_solve_noalloc (the single-rhs FP64 CG with its work memory in a fixed device array)
against wl_cg_solve, called through the C ABI: the same solution and statistics, bit for bit,
for two solves in a row on one device, the second on a smaller lattice (so the fixed array
still holds the first solve's contents), with the one-launch and the kernel-by-kernel solve,
and the limits it refuses (conjugate_gradient.h:46-104 in both)."""
import ctypes as C

import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN

pytestmark = pytest.mark.gpu


def setup(dims, kind, eps, seed, sseed):
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, kind, eps, seed))
    b = W.Field(lat).upload(W.gen_spinor(dims, sseed, EVEN))
    return lat, b


def call(fn, lat, b, kappa, tol=1e-10, maxiters=10000):
    x = W.Field(lat)
    st = W.CGStats()
    rc = fn(lat.handle, x.handle, b.handle, kappa, tol, maxiters, C.byref(st))
    W.check(rc, allow=(W.WL_E_NOCONV,))
    return x.download(), st.as_dict(), rc


def same(a, b):
    assert np.array_equal(a[0], b[0])
    assert a[2] == b[2]
    for key in ("iterations", "converged", "rr_rel", "true_resid"):
        assert a[1][key] == b[1][key], (key, a[1][key], b[1][key])


def test_noalloc_cg_equals_cg_solve_twice_second_smaller():
    L = W.lib()
    big = setup((16, 16, 16, 16), "smooth", 0.3, 3, 4)
    small = setup((8, 8, 8, 8), "haar", 0.0, 1, 2)
    print("noalloc CG 16^4 then 8^4", flush=True)
    got_big = call(L.wl_cg_solve_noalloc, *big, 0.125)
    got_small = call(L.wl_cg_solve_noalloc, *small, 0.12)
    ref_big = call(L.wl_cg_solve, *big, 0.125)
    ref_small = call(L.wl_cg_solve, *small, 0.12)
    same(got_big, ref_big)
    same(got_small, ref_small)
    assert got_big[1]["converged"] == 1 and got_big[1]["true_resid"] <= 1e-10
    assert got_small[1]["converged"] == 1 and got_small[1]["true_resid"] <= 1e-10
    # the kernel-by-kernel solve, and maxiters reached
    print("noalloc CG kernel by kernel, 8^4", flush=True)
    lat, b = small
    lat.set_option(W.OPT_CG_PERSISTENT, -1)
    same(call(L.wl_cg_solve_noalloc, lat, b, 0.12), call(L.wl_cg_solve, lat, b, 0.12))
    lat.set_option(W.OPT_CG_PERSISTENT, 0)
    got = call(L.wl_cg_solve_noalloc, lat, b, 0.12, 1e-14, 3)
    same(got, call(L.wl_cg_solve, lat, b, 0.12, 1e-14, 3))
    assert got[2] == W.WL_E_NOCONV and got[1]["iterations"] == 3
    for lat, _ in (big, small):
        lat.close()


def test_noalloc_cg_refuses_what_it_cannot_hold():
    L = W.lib()
    st = W.CGStats()
    # 20^4: 80000 sites per parity, beyond the fixed array (no compute is launched)
    lat = W.Lattice((20, 20, 20, 20))
    x, b = W.Field(lat), W.Field(lat)
    assert L.wl_cg_solve_noalloc(lat.handle, x.handle, b.handle, 0.12, 1e-10, 100, C.byref(st)) == W.WL_E_ARG
    assert b"exceed the fixed workspace" in L.wl_last_error()
    lat.close()
    lat = W.Lattice((8, 8, 8, 8))
    x, b = W.Field(lat), W.Field(lat)
    # no gauge field yet
    assert L.wl_cg_solve_noalloc(lat.handle, x.handle, b.handle, 0.12, 1e-10, 100, C.byref(st)) == W.WL_E_STATE
    # two-row links asked for before the upload: their FP64 copy is made by the first allocating
    # operator or solver call, not by wl_cg_solve_noalloc
    lat.set_option(W.OPT_LINK_ROWS, 2)
    lat.upload_gauge(W.gen_gauge((8, 8, 8, 8), "haar", 0.0, 1))
    b.upload(W.gen_spinor((8, 8, 8, 8), 2, EVEN))
    assert L.wl_cg_solve_noalloc(lat.handle, x.handle, b.handle, 0.12, 1e-10, 100, C.byref(st)) == W.WL_E_STATE
    assert b"two-row" in L.wl_last_error()
    ref = call(L.wl_cg_solve, lat, b, 0.12)
    same(call(L.wl_cg_solve_noalloc, lat, b, 0.12), ref)
    assert ref[1]["converged"] == 1
    # two right-hand sides
    x2, b2 = W.Field(lat, nrhs=2), W.Field(lat, nrhs=2)
    assert L.wl_cg_solve_noalloc(lat.handle, x2.handle, b2.handle, 0.12, 1e-10, 100, C.byref(st)) == W.WL_E_ARG
    lat.close()
