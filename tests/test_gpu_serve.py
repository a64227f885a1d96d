"""Request server (continuous batching of CG solves, wl_server_*) on the GPU.

A request is solved by its own batched CG on its own stream, with kernels and launch shapes
that depend only on its lattice size, precision and rhs count, and with reductions summed in a
fixed order.  So its solution must be identical to the last bit whether it runs alone or among
other requests, and identical to the blocking multi-rhs solver on the same inputs.  Solutions
are checked against the oracle's M^dag M: the FP64 true residual meets the tolerance (mixed
precision converges on it; plain FP64 CG on its iterated residual, as HILA's CG)."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle.oracle import EVEN, Oracle

pytestmark = pytest.mark.gpu


class Req:
    def __init__(self, dims, seed, nrhs, kappa, tol, prec, eps, maxiters=10000):
        self.dims, self.nrhs, self.kappa, self.tol, self.prec = dims, nrhs, kappa, tol, prec
        self.maxiters = maxiters
        V = int(np.prod(dims))
        self.Vh = V // 2
        self.U = W.gen_gauge(dims, "smooth", eps, seed)
        self.src = np.ascontiguousarray(
            np.stack([W.gen_spinor(dims, 1000 * seed + j, EVEN)[: self.Vh] for j in range(nrhs)]))

    def submit(self, S):
        sol = np.zeros_like(self.src)
        t = S.submit(self.dims, self.U, self.src, sol, self.kappa, self.tol, self.maxiters,
                     precision=self.prec, delta=0.1, update_every=100)
        return t, sol


MIX = [
    Req((4, 4, 4, 4), 1, 3, 0.120, 1e-10, "f32", 0.3),
    Req((8, 8, 8, 8), 2, 2, 0.1245, 1e-8, "f32", 0.5),
    Req((4, 4, 4, 8), 3, 1, 0.121, 1e-12, "f64", 0.2),
    Req((8, 8, 8, 8), 4, 5, 0.123, 1e-10, "f64", 0.4),
    Req((8, 8, 8, 16), 5, 12, 0.1235, 1e-10, "f32", 0.35),
    Req((4, 4, 4, 4), 6, 1, 0.1245, 1e-12, "f32", 0.25),
    Req((8, 8, 8, 8), 14, 4, 0.123, 1e-10, "f16", 0.3),
]


def true_resid(req, sol, j):
    orc = Oracle(req.dims)
    V = int(np.prod(req.dims))
    x = np.zeros((V, 24))
    x[: req.Vh] = sol[j]
    b = np.zeros((V, 24))
    b[: req.Vh] = req.src[j]
    y = orc.meo(req.U, req.kappa, orc.meo(req.U, req.kappa, x, 1), -1)
    return float(np.linalg.norm(y - b) / np.linalg.norm(b))


def solve_alone(req):
    with W.Server(max_active=1) as S:
        t, sol = req.submit(S)
        res, st = S.wait(t)
    return res, st, sol


def test_each_result_is_the_same_alone_and_in_the_mix():
    alone = [solve_alone(r) for r in MIX]
    for r, (res, st, sol) in zip(MIX, alone):
        assert res["status"] == W.WL_OK, res
        assert len(st) == r.nrhs and all(s["converged"] for s in st)
        lim = 20 * r.tol if r.prec == "f64" else 2 * r.tol
        for j in range(r.nrhs):
            assert true_resid(r, sol, j) <= lim, (r.dims, j)
            assert abs(st[j]["true_resid"] - true_resid(r, sol, j)) <= 0.05 * lim
    # the same requests twice over, all in flight together, in an interleaved order
    order = list(range(len(MIX))) * 2
    order = order[::2] + order[1::2][::-1]
    with W.Server(max_active=8) as S:
        sub = [(i,) + MIX[i].submit(S) for i in order]
        out = [(i, S.wait(t), sol) for i, t, sol in sub]
        info = S.info()
    assert info["completed"] == len(order) and info["active"] == 0
    for i, (res, st), sol in out:
        ares, ast, asol = alone[i]
        assert res["status"] == W.WL_OK
        assert np.array_equal(sol, asol), f"request {i} differs in the mix"
        assert [s["iterations"] for s in st] == [s["iterations"] for s in ast]
        assert res["latency_ms"] >= res["queue_ms"] >= 0


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_server_matches_blocking_multi_rhs_solver(prec):
    r = Req((8, 8, 8, 8), 7, 4, 0.124, 1e-10, prec, 0.3)
    res, st, sol = solve_alone(r)
    lat = W.Lattice(r.dims)
    lat.upload_gauge(r.U)
    D = W.Dirac_Wilson_evenodd(r.kappa, lat)
    b = W.Field(lat, EVEN, W.F64, nrhs=r.nrhs)
    x = W.Field(lat, EVEN, W.F64, nrhs=r.nrhs)
    for j in range(r.nrhs):
        b.upload(r.src[j], rhs=j)
    cg = W.CG(D, accuracy=r.tol, maxiters=10000, precision="f64" if prec == "f64" else "f32",
              delta=0.1, update_every=100)
    bst = cg.apply(b, x)
    for j in range(r.nrhs):
        assert np.array_equal(x.download(j), sol[j]), j
        assert bst[j]["iterations"] == st[j]["iterations"]


def test_queueing_beyond_max_active_and_zero_source():
    r = Req((4, 4, 4, 4), 8, 3, 0.122, 1e-10, "f32", 0.3)
    r.src[1] = 0.0
    with W.Server(max_active=2) as S:
        subs = [r.submit(S) for _ in range(7)]
        outs = [(S.wait(t), sol) for t, sol in subs]
    first = outs[0][1]
    for (res, st), sol in outs:
        assert res["status"] == W.WL_OK
        assert np.array_equal(sol, first)
        assert not sol[1].any() and st[1]["converged"] == 1 and st[1]["iterations"] == 0


def test_admission_least_work_first_with_aging():
    """One solve at a time: while a long request runs, the waiting requests are admitted by
    least estimated work (sites x rhs x digits of the tolerance), the earlier one first on ties,
    and a request passed over 8 times goes next.  The order is read from the queue times: with
    one solve at a time every admission waits for the completion before it."""
    import time
    blocker = Req((16, 16, 16, 16), 31, 12, 0.1245, 1e-12, "f64", 0.3)
    mid = Req((8, 8, 8, 16), 32, 6, 0.123, 1e-10, "f32", 0.3)
    small = Req((4, 4, 4, 4), 33, 1, 0.122, 1e-8, "f32", 0.3)
    with W.Server(max_active=1) as S:
        tb, _ = blocker.submit(S)
        t0 = time.perf_counter()
        while S.info()["active"] == 0:  # the blocker holds the one solve slot
            assert time.perf_counter() - t0 < 30
            time.sleep(0.001)
        tm, _ = mid.submit(S)
        ts = [small.submit(S)[0] for _ in range(10)]
        rb = S.wait(tb)[0]
        rm = S.wait(tm)[0]
        rs = [S.wait(t)[0] for t in ts]
    assert rb["status"] == rm["status"] == W.WL_OK and all(r["status"] == W.WL_OK for r in rs)
    qm, qs = rm["queue_ms"], [r["queue_ms"] for r in rs]
    # the small requests first, in submission order, until the mid-size one has been passed
    # over 8 times; then it, then the last two small ones
    assert qs[:8] == sorted(qs[:8]), qs
    assert max(qs[:8]) < qm < min(qs[8:]), (qm, qs)


def test_provisioned_slots_serve_without_allocating():
    """wl_server_provision: idle slots made before traffic hold all the device memory a
    request of their shape uses.  Requests that take them create no slot and allocate nothing
    (bytes in use unchanged), and their solutions are those of a server without provisioning,
    bit for bit.  Provisioning stops at the HBM budget."""
    r = Req((8, 8, 8, 8), 41, 3, 0.122, 1e-10, "f32", 0.3)
    base = solve_alone(r)
    with W.Server(max_active=2) as S:
        assert S.provision(r.dims, 4, 2, precision="f32") == 2
        i0 = S.info()
        assert i0["slots"] == 2 and i0["slots_created"] == 2 and i0["bytes_in_use"] > 0
        subs = [r.submit(S) for _ in range(2)]
        outs = [(S.wait(t), sol) for t, sol in subs]
        i1 = S.info()
    assert i1["slots_created"] == 2 and i1["bytes_in_use"] == i0["bytes_in_use"]
    for (res, st), sol in outs:
        assert res["status"] == W.WL_OK
        assert np.array_equal(sol, base[2])
    one = i0["bytes_in_use"] // 2
    with W.Server(budget_bytes=int(1.5 * one)) as S:
        assert S.provision(r.dims, 4, 3, precision="f32") == 1
    with W.Server() as S:
        with pytest.raises(W.WilsonError) as e:
            S.provision((8, 8, 8, 7), 4, 1)
        assert e.value.code == W.WL_E_ARG


def test_maxiters_reports_noconv():
    r = Req((8, 8, 8, 8), 9, 2, 0.1245, 1e-12, "f32", 0.3, maxiters=5)
    with W.Server() as S:
        t, sol = r.submit(S)
        res, st = S.wait(t)
    assert res["status"] == W.WL_E_NOCONV
    assert all(not s["converged"] and s["iterations"] <= 8 for s in st)


def test_hbm_budget_admission_and_slot_reuse():
    r8 = Req((8, 8, 8, 8), 10, 3, 0.122, 1e-8, "f32", 0.3)
    r4 = Req((8, 8, 8, 10), 11, 3, 0.122, 1e-8, "f32", 0.3)
    with W.Server() as S:
        t, _ = r8.submit(S)
        S.wait(t)
        one = S.info()["bytes_in_use"]
    assert one > 0
    # room for one slot of either size at a time: requests run one after another, the idle
    # slot of the other size is released, and results do not change (the least-work order
    # runs the three 8^4 requests before the 8^3 x 10 one: one release at least)
    with W.Server(budget_bytes=int(1.5 * one)) as S:
        subs = [q.submit(S) for q in (r8, r8, r4, r8)]
        outs = [(S.wait(t), sol) for t, sol in subs]
        info = S.info()
    assert all(res["status"] == W.WL_OK for (res, _), _ in outs)
    assert info["slots"] == 1 and info["slots_released"] >= 1
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][1], outs[3][1])
    # a budget below one slot: the request is refused, not queued forever
    with W.Server(budget_bytes=1 << 20) as S:
        t, _ = r8.submit(S)
        res, _ = S.wait(t, timeout_ms=60000)
    assert res["status"] == W.WL_E_OOM and "budget" in res["message"]


def test_invalid_requests_are_rejected_at_submission():
    r = Req((4, 4, 4, 4), 12, 1, 0.12, 1e-10, "f32", 0.3)
    with W.Server() as S:
        with pytest.raises(W.WilsonError) as e:
            S.submit(r.dims, r.U, r.src, np.zeros_like(r.src), r.kappa, -1.0)
        assert e.value.code == W.WL_E_ARG
        with pytest.raises(W.WilsonError) as e:
            S.submit(r.dims, r.U, r.src, np.zeros_like(r.src), r.kappa, 1e-8, precision="f32", delta=2.0)
        assert e.value.code == W.WL_E_ARG
        with pytest.raises(W.WilsonError) as e:
            S.wait(12345, timeout_ms=0)
        assert e.value.code == W.WL_E_ARG


def test_pinned_request_buffers():
    r = Req((8, 8, 8, 8), 13, 2, 0.123, 1e-10, "f32", 0.3)
    res_a, _, sol_a = solve_alone(r)
    gb = W.HostBuffer(r.U.shape)
    sb = W.HostBuffer(r.src.shape)
    xb = W.HostBuffer(r.src.shape)
    gb.array[:] = r.U
    sb.array[:] = r.src
    xb.array[:] = 0
    with W.Server() as S:
        t = S.submit(r.dims, gb.array, sb.array, xb.array, r.kappa, r.tol, precision="f32", delta=0.1)
        res, _ = S.wait(t)
    assert res["status"] == W.WL_OK
    assert np.array_equal(xb.array, sol_a)
    for b in (gb, sb, xb):
        b.close()


def test_server_with_graph_replay_matches():
    reqs = [MIX[1], MIX[3], MIX[6]]
    base = [solve_alone(r) for r in reqs]
    with W.Server(max_active=4) as S:
        S.set_option(W.OPT_GRAPHS, 1)
        subs = [r.submit(S) for r in reqs]
        outs = [(S.wait(t), sol) for t, sol in subs]
        with pytest.raises(W.WilsonError) as e:
            S.set_option(W.OPT_GRAPHS, 0)  # frozen after the first submission
        assert e.value.code == W.WL_E_STATE
    for ((res, st), sol), (bres, bst, bsol) in zip(outs, base):
        assert res["status"] == W.WL_OK
        assert np.array_equal(sol, bsol)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_two_row_links_server_matches_blocking_solver(prec):
    """WL_OPT_LINK_ROWS = 2 on every slot: the same bits as the blocking solver with two-row
    links, within the tolerance on the FP64 true residual; links that are not special unitary
    are refused with WL_E_ARG and do not disturb the next request."""
    r = Req((8, 8, 8, 8), 21, 3, 0.124, 1e-10, prec, 0.3)
    bad = Req((8, 8, 8, 8), 22, 1, 0.124, 1e-10, prec, 0.3)
    bad.U[1, 5, :] *= 1.001
    with W.Server(max_active=2) as S:
        S.set_option(W.OPT_LINK_ROWS, 2)
        tb, _ = bad.submit(S)
        resb, _ = S.wait(tb)
        t, sol = r.submit(S)
        res, st = S.wait(t)
    assert resb["status"] == W.WL_E_ARG
    assert res["status"] == W.WL_OK
    lat = W.Lattice(r.dims)
    lat.upload_gauge(r.U)
    lat.set_option(W.OPT_LINK_ROWS, 2)
    D = W.Dirac_Wilson_evenodd(r.kappa, lat)
    b = W.Field(lat, EVEN, W.F64, nrhs=r.nrhs)
    x = W.Field(lat, EVEN, W.F64, nrhs=r.nrhs)
    for j in range(r.nrhs):
        b.upload(r.src[j], rhs=j)
    bst = W.CG(D, accuracy=r.tol, maxiters=10000, precision=prec, delta=0.1, update_every=100).apply(b, x)
    for j in range(r.nrhs):
        assert np.array_equal(x.download(j), sol[j]), j
        assert bst[j]["iterations"] == st[j]["iterations"]
        assert true_resid(r, sol, j) <= 1.5 * r.tol   # FP64 CG stops on its iterated residual
    lat.close()
