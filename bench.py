"""Benchmark of the MI355X Wilson even-odd Dslash + CG path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cg-solves N] [--no-cpu]

Workload: the plain-method GPU configuration -- 32^4 lattice, FP64, one right-hand side,
links exp(i 0.3 H) (seed 3), kappa 0.125, Gaussian source (seed 4), antiperiodic in t.
A step is one half-lattice hopping application D_oe (out on V/2 sites), the unit of work of
the Dslash metric.  After the Dslash, CG on M^dag M is run to 1e-10 (relative, with the FP64
true residual recomputed) for --cg-solves solves: solves/s and p50/p99 latency.

Prints ONE JSON line.  `value` is the Dslash throughput in algorithmic GB/s: 1536 bytes per
output site in FP64 (8 links x 144 B + one input and one output spinor x 192 B), i.e. each
link and spinor counted once per application.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E, 8 TB/s
BYTES_PER_SITE_F64 = 1536        # 8*144 + 192 + 192
FLOPS_PER_SITE = 1320
BYTES_NO_REUSE_F64 = 2880        # 8*(24+18) + 24 reals: every neighbour spinor read per hop
DIMS = (32, 32, 32, 32)
KAPPA = 0.125
METRIC = "Dslash GB/s & %HBM roofline; CG solves/s + p99 latency @1e-10 true residual"
PMC_FILE = os.path.join(ROOT, "profiles", "dslash_f64_32x32x32x32_pmc.json")


def rel_diff(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def solve_series(W, D, fin, fx, n, precision, accuracy=1e-10, delta=1e-3):
    """n solves of M^dag M x = b from x = 0; latency percentiles, iterations, true residuals."""
    cg = W.CG(D, accuracy, 10000, precision=precision, delta=delta)
    lat_ms, iters, tres, upd = [], [], [], []
    for i in range(n):
        fx.zero()
        t = time.perf_counter()
        st = cg.apply(fin, fx)[0]
        lat_ms.append(1e3 * (time.perf_counter() - t))
        iters.append(st["iterations"])
        tres.append(st["true_resid"])
        upd.append(st["reliable_updates"])
        print(f"# cg {precision} solve {i}: {lat_ms[-1]:.1f} ms, {st['iterations']} it, "
              f"{st['reliable_updates']} updates, true resid {st['true_resid']:.2e}",
              file=sys.stderr, flush=True)
        if not st["converged"]:
            raise RuntimeError(f"CG ({precision}) did not converge: {st}")
    if not lat_ms:
        return None
    arr = np.array(lat_ms)
    return {
        "solves_per_s": round(1e3 / arr.mean(), 4),
        "p50_ms": round(float(np.percentile(arr, 50)), 2),
        "p99_ms": round(float(np.percentile(arr, 99)), 2),
        "iterations": int(np.median(iters)),
        "reliable_updates": int(np.median(upd)),
        "max_true_resid": float(max(tres)),
        "ms_per_iteration": round(float(arr.mean()) / float(np.mean(iters)), 4),
        "precision": precision,
        "solves": len(lat_ms),
    }


BYTES_PER_SITE_F64_ROWS = 8 * 96 + 2 * 192   # two-row links: rows 0 and 1 stored (12 reals)


def two_row_config(W, lat, D, fin, fodd, x64_full, n_solves):
    """The same 32^4 field with two-row links (WL_OPT_LINK_ROWS = 2): the hopping term's time
    and its roofline fraction on the bytes it actually needs (1152 per site), FP64 and
    FP32-inner CG, and the FP64 multi-rhs hop at 12 rhs, against full links."""
    vh = lat.Vh
    out = {"bytes_per_site": BYTES_PER_SITE_F64_ROWS}
    N = 12
    f12 = W.Field(lat, W.EVEN, W.F64, N)
    for j in range(N):
        f12.copy_rhs_from(j, fin)
    o12 = W.Field(lat, W.ODD, W.F64, N)
    ms_full = D.time(fin, fodd, "hop", 1, warmup=5, iters=100)
    ms12_full = D.time(f12, o12, "hop", 1, warmup=3, iters=20)
    lat.set_option(W.OPT_LINK_ROWS, 2)
    try:
        ms = D.time(fin, fodd, "hop", 1, warmup=5, iters=100)
        ms12 = D.time(f12, o12, "hop", 1, warmup=3, iters=20)
        gbs = BYTES_PER_SITE_F64_ROWS * vh / (ms * 1e-3) / 1e9
        out["dslash_us"] = round(ms * 1e3, 2)
        out["dslash_us_full_links"] = round(ms_full * 1e3, 2)
        out["dslash_speedup"] = round(ms_full / ms, 3)
        out["dslash_gbs"] = round(gbs, 1)
        out["dslash_hbm_frac"] = round(gbs / HBM_PEAK_GBS, 4)
        out["multi_rhs_12_us_per_rhs"] = round(ms12 * 1e3 / N, 2)
        out["multi_rhs_12_us_per_rhs_full_links"] = round(ms12_full * 1e3 / N, 2)
        x = W.Field(lat, W.EVEN)
        out["cg"] = solve_series(W, D, fin, x, n_solves, "f64")
        out["cg"]["solution_rel_diff_vs_full_links"] = rel_diff(x.download(), x64_full.download())
        out["cg_mixed_f32"] = solve_series(W, D, fin, x, n_solves, "f32")
    finally:
        lat.set_option(W.OPT_LINK_ROWS, 3)
    return out


SMALL_DIMS, SMALL_KAPPA = (8, 8, 8, 8), 0.12


def small_cg_config(W, n, dims=SMALL_DIMS, kind="haar", eps=0.0, gseed=1, kappa=SMALL_KAPPA, sseed=2):
    """The plain method's CPU-runnable configuration on the GPU: 8^4, Haar links (seed 1),
    kappa 0.12, Gaussian source (seed 2), FP64 CG to 1e-10.  3 MB of links and spinors sit in
    the caches, so this measures launch and memory latency, not bandwidth.  Solved kernel by
    kernel with a host check every 8 iterations (the default) and, for comparison, in one
    cooperative launch per solve (WL_OPT_CG_PERSISTENT = 1); the two solutions are identical."""
    U = W.gen_gauge(dims, kind, eps, gseed)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    fin = W.Field(lat, W.EVEN).upload(W.gen_spinor(dims, sseed, W.EVEN))
    xp, xk = W.Field(lat, W.EVEN), W.Field(lat, W.EVEN)
    out = solve_series(W, D, fin, xk, max(n, 5), "f64")
    lat.set_option(W.OPT_CG_PERSISTENT, 1)
    one = solve_series(W, D, fin, xp, max(n, 5), "f64")
    lat.set_option(W.OPT_CG_PERSISTENT, 0)
    out["lattice"] = list(dims)
    out["mode"] = "kernel by kernel"
    out["one_launch"] = one
    out["one_launch_speedup"] = round(out["p50_ms"] / one["p50_ms"], 3)
    out["solutions_bitwise_equal"] = bool(np.array_equal(xp.download(), xk.download()))
    lat.close()
    return out


def mixed_precision_config(W, n):
    """Mixed-precision configuration: 48^3x96, links exp(i 0.3 H) (seed 6), kappa 0.1245,
    Gaussian source (seed 7), CG to 1e-10 true residual in FP64, with the FP32 inner solver and
    with the FP16 (site-scaled) inner solver; each mixed solution's error is stated against the
    FP64 CG solution."""
    dims, kappa = (48, 48, 48, 96), 0.1245
    t0 = time.perf_counter()
    U = W.gen_gauge(dims, "smooth", 0.3, 6)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    del U
    b = W.Field(lat).upload(W.gen_spinor(dims, 7, W.EVEN))
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    print(f"# 48^3x96 setup {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    fo64, fo32 = W.Field(lat, W.ODD), W.Field(lat, W.ODD, W.F32)
    fin32 = W.Field(lat, W.EVEN, W.F32).upload(b.download())
    us64 = 1e3 * D.time(b, fo64, "hop", 1, warmup=3, iters=20)
    us32 = 1e3 * D.time(fin32, fo32, "hop", 1, warmup=3, iters=20)
    del fo64, fo32, fin32
    # device copy of the FP64 hop's byte count (two 4-rhs FP64 fields of 4.1 GB), the rate a
    # stream of that size reaches with no cache help
    ca, cb = W.Field(lat, W.EVEN, W.F64, 4), W.Field(lat, W.EVEN, W.F64, 4)
    copy_gbs = 2 * 4 * 192 * lat.Vh / (1e-3 * D.time(ca, cb, "copy", warmup=3, iters=20)) / 1e9
    ca.close()
    cb.close()
    x64, x32, x16 = W.Field(lat), W.Field(lat), W.Field(lat)
    r64 = solve_series(W, D, b, x64, n, "f64")
    r32 = solve_series(W, D, b, x32, n, "f32")
    r16 = solve_series(W, D, b, x16, n, "f16", delta=1e-2)
    r16["speedup_vs_f64"] = round(r16["solves_per_s"] / r64["solves_per_s"], 3)
    r16["solution_rel_diff_vs_f64"] = rel_diff(x16.download(), x64.download())
    out = {"lattice": list(dims), "kappa": kappa, "f64": r64, "f32_mixed": r32,
           "f16_mixed": r16,
           "speedup_vs_f64": round(r32["solves_per_s"] / r64["solves_per_s"], 3),
           "solution_rel_diff_vs_f64": rel_diff(x32.download(), x64.download()),
           "dslash_f64_us": round(us64, 1),
           "dslash_f64_hbm_frac": round(1536 * lat.Vh / (us64 * 1e-6) / 8e12, 4),
           "copy_gbs": round(copy_gbs, 1),
           "dslash_f64_frac_of_copy": round(1536 * lat.Vh / (us64 * 1e-6) / 1e9 / copy_gbs, 4),
           "dslash_f32_us": round(us32, 1),
           "dslash_f32_hbm_frac": round(768 * lat.Vh / (us32 * 1e-6) / 8e12, 4)}
    lat.close()
    return out


def even_point_seed(W, dims, seed):
    """First seed >= `seed` whose point-source site is even (CG acts on even sites)."""
    Vh = int(np.prod(dims)) // 2
    while True:
        _, site = W.gen_point(dims, seed, 0, 0)
        if site < Vh:
            return seed
        seed += 1


def multi_rhs_config(W, n_plain=4):
    """Batched multi-rhs configuration: 32^4, links exp(i 0.3 H) (seed 5), kappa 0.125;
    12 point sources at one seeded even site (one per spin-colour pair) and 12 Gaussian
    sources, solved together by the batched CG (FP32 inner, FP64 reliable updates) to 1e-10
    true residual, with the multi-rhs kernel autotuned for 24 rhs.  Compared per rhs with
    plain FP64 single-rhs CG (HILA's algorithm) on n_plain of the same sources."""
    dims, kappa, N = (32, 32, 32, 32), 0.125, 24
    U = W.gen_gauge(dims, "smooth", 0.3, 5)
    lat = W.Lattice(dims)
    lat.upload_gauge(U)
    del U
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    b = W.Field(lat, W.EVEN, W.F64, nrhs=N)
    pseed = even_point_seed(W, dims, 8)
    for sc in range(12):
        pt, site = W.gen_point(dims, pseed, sc // 3, sc % 3)
        b.upload(pt, rhs=sc)
    for j in range(12):
        b.upload(W.gen_spinor(dims, 100 + j, W.EVEN), rhs=12 + j)
    t0 = time.perf_counter()
    tune = lat.autotune(W.F32, N)
    tune["seconds"] = round(time.perf_counter() - t0, 2)
    print(f"# multi-rhs autotune: {tune}", file=sys.stderr, flush=True)
    x = W.Field(lat, W.EVEN, W.F64, nrhs=N)
    cg = W.CG(D, 1e-10, 10000, precision="f32", delta=1e-3, update_every=100)
    W.CG(D, 1e-10, 8, precision="f32").apply(b, x)  # warm-up: work fields, code paths
    x.zero()
    lat.sync()
    t = time.perf_counter()
    st = cg.apply(b, x)
    t_batch = time.perf_counter() - t
    print(f"# multi-rhs batched CG: {N} rhs in {t_batch * 1e3:.0f} ms", file=sys.stderr, flush=True)
    b1, x1 = W.Field(lat), W.Field(lat)
    cg64 = W.CG(D, 1e-10, 10000, precision="f64")
    picks = [0, 7, 12, 19][:n_plain]
    t_plain, diffs, it64 = [], [], []
    for j in picks:
        b1.upload(b.download(j))
        x1.zero()
        lat.sync()
        t = time.perf_counter()
        s64 = cg64.apply(b1, x1)[0]
        t_plain.append(time.perf_counter() - t)
        it64.append(s64["iterations"])
        diffs.append(rel_diff(x.download(j), x1.download()))
    per_rhs_batch = t_batch / N
    per_rhs_plain = float(np.mean(t_plain))
    out = {
        "lattice": list(dims), "kappa": kappa, "nrhs": N,
        "sources": "12 point (one even site, all spin-colour pairs) + 12 Gaussian",
        "autotune": tune,
        "batched_ms": round(t_batch * 1e3, 1),
        "rhs_solves_per_s": round(N / t_batch, 3),
        "iterations_max": max(s["iterations"] for s in st),
        "reliable_updates_max": max(s["reliable_updates"] for s in st),
        "max_true_resid": max(s["true_resid"] for s in st),
        "all_converged": all(s["converged"] for s in st),
        "plain_f64_single_rhs_ms_per_solve": round(per_rhs_plain * 1e3, 2),
        "plain_f64_iterations": it64,
        "speedup_per_rhs_vs_plain_f64": round(per_rhs_plain / per_rhs_batch, 3),
        "solution_rel_diff_vs_f64_max": max(diffs),
    }
    # the same 24 rhs by block CG (one shared Krylov space, FP32 inner, FP64 reliable updates)
    xb = W.Field(lat, W.EVEN, W.F64, nrhs=N)
    lat.set_option(W.OPT_BLOCK_KRYLOV, 1)
    W.CG(D, 1e-10, 2, precision="f32").apply(b, xb)  # warm-up: work fields, code paths
    xb.zero()
    lat.sync()
    t = time.perf_counter()
    sb = cg.apply(b, xb)
    t_block = time.perf_counter() - t
    lat.set_option(W.OPT_BLOCK_KRYLOV, 0)
    print(f"# multi-rhs block CG: {N} rhs in {t_block * 1e3:.0f} ms, {sb[0]['iterations']} iterations",
          file=sys.stderr, flush=True)
    out["block_cg"] = {
        "ms": round(t_block * 1e3, 1),
        "rhs_solves_per_s": round(N / t_block, 3),
        "iterations": sb[0]["iterations"],
        "reliable_updates": sb[0]["reliable_updates"],
        "max_true_resid": max(s["true_resid"] for s in sb),
        "all_converged": all(s["converged"] for s in sb),
        "speedup_per_rhs_vs_plain_f64": round(per_rhs_plain / (t_block / N), 3),
        "speedup_vs_batched": round(t_batch / t_block, 3),
        "solution_rel_diff_vs_batched_max": max(rel_diff(xb.download(j), x.download(j)) for j in range(N)),
    }
    lat.close()
    return out


SERVE_VOLUMES = ((16, 16, 16, 16), (24, 24, 24, 48), (32, 32, 32, 32))


class _SolutionPool:
    """Page-locked solution buffers (12 rhs each) per lattice size, recycled between requests."""

    def __init__(self, W):
        self.W = W
        self.free = {d: [] for d in SERVE_VOLUMES}
        self.lock = threading.Lock()

    def get(self, dims):
        with self.lock:
            if self.free[dims]:
                return self.free[dims].pop()
        return self.W.HostBuffer((12, int(np.prod(dims)) // 2, 24))

    def put(self, dims, buf):
        with self.lock:
            self.free[dims].append(buf)


def serving_pool(W, rng, per_volume):
    """Per lattice size: `per_volume` gauge fields exp(i eps H) with eps ~ U[0.2, 0.5] and
    their own seeds, and 12 Gaussian sources, all in page-locked host memory."""
    gauges, sources = {}, {}
    for dims in SERVE_VOLUMES:
        V = int(np.prod(dims))
        lst = []
        for _ in range(per_volume):
            eps, seed = float(rng.uniform(0.2, 0.5)), int(rng.integers(1, 1 << 30))
            hb = W.HostBuffer((4, V, 18))
            hb.array[:] = W.gen_gauge(dims, "smooth", eps, seed)
            lst.append((eps, seed, hb))
        sb = W.HostBuffer((12, V // 2, 24))
        for j in range(12):
            sb.array[j] = W.gen_spinor(dims, int(rng.integers(1, 1 << 30)), W.EVEN)[: V // 2]
        gauges[dims], sources[dims] = lst, sb
    return gauges, sources


def draw_requests(rng, n, per_volume):
    """Request parameters: volume uniform over the three sizes, kappa ~ U[0.120, 0.1245],
    tolerance from {1e-8, 1e-10, 1e-12}, 1..12 right-hand sides, gauge field from the pool."""
    out = []
    for _ in range(n):
        dims = SERVE_VOLUMES[int(rng.integers(len(SERVE_VOLUMES)))]
        out.append({"dims": dims, "g": int(rng.integers(per_volume)),
                    "kappa": float(rng.uniform(0.120, 0.1245)),
                    "tol": [1e-8, 1e-10, 1e-12][int(rng.integers(3))],
                    "nrhs": int(rng.integers(1, 13))})
    return out


SERVE_PRECISIONS = {"f32": 1e-3, "f16": 1e-2}  # inner precision -> reliable-update delta


def run_served(W, S, reqs, gauges, sources, pool, arrivals=None, precision="f16"):
    """Submit reqs (at the given arrival offsets in seconds, else all at once) from a client
    thread, wait for each; returns (results, wall seconds from first submission to last
    completion)."""
    import queue
    q = queue.Queue()
    results = [None] * len(reqs)
    t_start = time.perf_counter()

    def submitter():
        for i, r in enumerate(reqs):
            if arrivals is not None:
                dt = arrivals[i] - (time.perf_counter() - t_start)
                if dt > 0:
                    time.sleep(dt)
            d = r["dims"]
            sol = pool.get(d)
            t = S.submit(d, gauges[d][r["g"]][2].array, sources[d].array[: r["nrhs"]],
                         sol.array[: r["nrhs"]], r["kappa"], r["tol"], 10000, precision=precision,
                         delta=SERVE_PRECISIONS[precision], update_every=100)
            q.put((i, t, sol))
        q.put(None)

    th = threading.Thread(target=submitter)
    th.start()
    while True:
        item = q.get()
        if item is None:
            break
        i, t, sol = item
        res, st = S.wait(t)
        if res["status"] != 0:
            raise RuntimeError(f"served request {i} failed: {res}")
        results[i] = res
        pool.put(reqs[i]["dims"], sol)
    wall = time.perf_counter() - t_start
    th.join()
    return results, wall


def plain_f64_baseline(W, reqs, gauges, sources):
    """The same requests solved one right-hand side at a time by plain FP64 CG (HILA's
    algorithm, wl_cg_solve with one rhs): gauge upload, source upload, solve, download per rhs,
    contexts resident per lattice size.  Returns wall seconds."""
    lats = {}
    for d in SERVE_VOLUMES:
        lat = W.Lattice(d)
        lats[d] = (lat, W.Field(lat), W.Field(lat))
    t0 = time.perf_counter()
    for r in reqs:
        d = r["dims"]
        lat, b1, x1 = lats[d]
        lat.upload_gauge(gauges[d][r["g"]][2].array)
        D = W.Dirac_Wilson_evenodd(r["kappa"], lat)
        cg = W.CG(D, r["tol"], 10000, precision="f64")
        for j in range(r["nrhs"]):
            b1.upload(sources[d].array[j])
            x1.zero()
            st = cg.apply(b1, x1)[0]
            if not st["converged"]:
                raise RuntimeError(f"plain FP64 CG did not converge: {st}")
            x1.download()
    wall = time.perf_counter() - t0
    for lat, b1, x1 in lats.values():
        b1.close()
        x1.close()
        lat.close()
    return wall


def serving_config(W, n_poisson=200, n_closed=24, rate=20.0, per_volume=3, seed=11):
    """Serving configuration: Poisson arrivals at `rate`/s of CG requests (volume from
    {16^4, 24^3x48, 32^4}, kappa ~ U[0.120, 0.1245], tolerance from {1e-8, 1e-10, 1e-12},
    1..12 Gaussian rhs, own gauge field exp(i eps H), eps ~ U[0.2, 0.5]) on the request server
    (mixed precision FP32 inner, FP64 reliable updates, batched over each request's rhs).
    Reports p50/p99 latency, solves/s and true residuals; and, on a closed batch of
    n_closed requests submitted at once, served rhs-solves/s against plain FP64 single-rhs CG
    on the same requests."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    gauges, sources = serving_pool(W, rng, per_volume)
    print(f"# serving pool {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    pool = _SolutionPool(W)
    closed = draw_requests(rng, n_closed, per_volume)
    poisson = draw_requests(rng, n_poisson, per_volume)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_poisson))
    # The client's page-locked solution buffers are allocated before any timing: a large
    # hipHostMalloc blocks the HIP runtime, kernel launches of the server's threads included.
    for d in SERVE_VOLUMES:
        bufs = [pool.get(d) for _ in range(max(16, sum(1 for r in closed if r["dims"] == d)))]
        for b in bufs:
            pool.put(d, b)
    S = W.Server()  # the library's default concurrency (6 solves at once)
    # warm-up: one request per lattice size and precision (slots created, code paths loaded)
    warm = [{"dims": d, "g": 0, "kappa": 0.12, "tol": 1e-8, "nrhs": 12} for d in SERVE_VOLUMES]
    for prec in SERVE_PRECISIONS:
        run_served(W, S, warm, gauges, sources, pool, precision=prec)
    rhs_c = sum(r["nrhs"] for r in closed)
    closed_wall = {}
    for prec in SERVE_PRECISIONS:
        res_c, closed_wall[prec] = run_served(W, S, closed, gauges, sources, pool, precision=prec)
        bad = [i for i, (r, q) in enumerate(zip(res_c, closed)) if r["true_resid"] > q["tol"]]
        if bad:
            raise RuntimeError(f"served {prec}: requests {bad} missed their tolerance")
        print(f"# served closed batch ({prec}): {n_closed} requests, {rhs_c} rhs in "
              f"{closed_wall[prec]:.2f} s", file=sys.stderr, flush=True)
    # the same closed batch on a server whose slots read two-row links (WL_OPT_LINK_ROWS = 2;
    # the request links are special unitary): the FP64 reliable updates and the FP32 inner
    # operator read a third less link data
    S2 = W.Server()
    S2.set_option(W.OPT_LINK_ROWS, 2)
    for prec in SERVE_PRECISIONS:
        run_served(W, S2, warm, gauges, sources, pool, precision=prec)
    closed_wall_rows = {}
    for prec in SERVE_PRECISIONS:
        res_r, closed_wall_rows[prec] = run_served(W, S2, closed, gauges, sources, pool, precision=prec)
        bad = [i for i, (r, q) in enumerate(zip(res_r, closed)) if r["true_resid"] > q["tol"]]
        if bad:
            raise RuntimeError(f"served {prec} with two-row links: requests {bad} missed their tolerance")
        print(f"# served closed batch ({prec}, two-row links): {closed_wall_rows[prec]:.2f} s", file=sys.stderr,
              flush=True)
    S2.close()
    S.close()
    best = min(closed_wall, key=closed_wall.get)
    wall_c = closed_wall[best]
    # The Poisson run: a server for the faster inner precision whose slots are provisioned
    # before timing (wl_server_provision), as a deployed server holds its slots: for each
    # lattice size as many 12-rhs slots as the server runs solves at once, so that any request
    # finds one.  Slots the timed run still creates are reported.
    S = W.Server()
    per_size = S.info()["max_active"]
    for d in SERVE_VOLUMES:
        S.provision(d, 12, per_size, precision=best)
    run_served(W, S, warm, gauges, sources, pool, precision=best)  # code paths loaded
    created0 = S.info()["slots_created"]
    res_p, wall_p = run_served(W, S, poisson, gauges, sources, pool, arrivals=arrivals,
                               precision=best)
    info = S.info()
    S.close()
    lat_ms = np.array([r["latency_ms"] for r in res_p])
    q_ms = np.array([r["queue_ms"] for r in res_p])
    rhs_p = sum(r["nrhs"] for r in poisson)
    print(f"# served poisson: {n_poisson} requests in {wall_p:.2f} s, p50 "
          f"{np.percentile(lat_ms, 50):.0f} ms p99 {np.percentile(lat_ms, 99):.0f} ms",
          file=sys.stderr, flush=True)
    wall_plain = plain_f64_baseline(W, closed, gauges, sources)
    print(f"# plain FP64 single-rhs CG on the closed batch: {wall_plain:.2f} s", file=sys.stderr,
          flush=True)
    tres = [r["true_resid"] for r in res_p]
    tol_ok = sum(1 for r, q in zip(res_p, poisson) if r["true_resid"] <= q["tol"])
    for d in gauges:
        for _, _, hb in gauges[d]:
            hb.close()
        sources[d].close()
    return {
        "arrival_rate_per_s": rate,
        "requests": n_poisson,
        "rhs": rhs_p,
        "requests_per_s": round(n_poisson / wall_p, 3),
        "rhs_solves_per_s": round(rhs_p / wall_p, 2),
        "p50_ms": round(float(np.percentile(lat_ms, 50)), 1),
        "p99_ms": round(float(np.percentile(lat_ms, 99)), 1),
        "p90_ms": round(float(np.percentile(lat_ms, 90)), 1),
        "mean_queue_ms": round(float(q_ms.mean()), 1),
        "p99_queue_ms": round(float(np.percentile(q_ms, 99)), 1),
        "arrival_span_s": round(float(arrivals[-1]), 2),
        "max_true_resid": float(max(tres)),
        "requests_meeting_tolerance": tol_ok,
        "precision": f"{best} inner (reliable-update delta {SERVE_PRECISIONS[best]}), FP64 "
                     f"reliable updates, reductions, solution and true residual",
        "max_active": info["max_active"],
        "slots_created_in_run": info["slots_created"] - created0,
        "slots": info["slots"],
        "hbm_in_use_gb": round(info["bytes_in_use"] / 1e9, 2),
        "closed_batch": {
            "requests": n_closed, "rhs": rhs_c,
            "served_s": {p: round(w, 3) for p, w in closed_wall.items()},
            "served_rhs_solves_per_s": {p: round(rhs_c / w, 2) for p, w in closed_wall.items()},
            "plain_f64_single_rhs_s": round(wall_plain, 3),
            "plain_f64_rhs_solves_per_s": round(rhs_c / wall_plain, 2),
            "speedup_vs_plain_f64": {p: round(wall_plain / w, 3) for p, w in closed_wall.items()},
            "two_row_links_served_s": {p: round(w, 3) for p, w in closed_wall_rows.items()},
            "two_row_links_speedup_vs_plain_f64": {p: round(wall_plain / w, 3)
                                                   for p, w in closed_wall_rows.items()},
        },
        "gauge_pool": f"{per_volume} gauge fields per lattice size (each with its own eps and "
                      f"seed), 12 Gaussian sources per size; a request draws one gauge field "
                      f"and its first nrhs sources",
    }


def profile_multi(W, steps):
    """Autotuned multi-rhs hopping launches (12 rhs at 32^4, links exp(i 0.3 H) seed 5,
    Gaussian sources), FP16, FP32 then FP64 storage, then copies of a 12-rhs FP16 field (the
    byte counters' calibration: 12 x 13 planes x stride x 4 B each way)."""
    dims, N = (32, 32, 32, 32), 12
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "smooth", 0.3, 5))
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    src = W.gen_spinor(dims, 10, W.EVEN)
    for prec in (W.F16, W.F32, W.F64):
        fin = W.Field(lat, W.EVEN, prec, nrhs=N)
        for j in range(N):
            fin.upload(src, rhs=j)
        fo = W.Field(lat, W.ODD, prec, nrhs=N)
        tune = lat.autotune(prec, N)
        print(f"# profile-multi: prec {prec} tuned {tune}", file=sys.stderr, flush=True)
        for _ in range(steps):
            D.hop_set(fin, fo, 1)
        lat.sync()
        if prec == W.F16:
            fc = W.Field(lat, W.EVEN, prec, nrhs=N)
            for _ in range(steps):
                fc.copy_from(fin)
            lat.sync()
    lat.close()


def profile_cg_small(W, steps):
    """`steps` FP64 CG solves to 1e-10 at 8^4 (Haar links seed 1, kappa 0.12, source seed 2),
    kernel by kernel with the default small-lattice hopping kernel."""
    dims = SMALL_DIMS
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "haar", 0, 1))
    D = W.Dirac_Wilson_evenodd(SMALL_KAPPA, lat)
    b = W.Field(lat).upload(W.gen_spinor(dims, 2, W.EVEN))
    x = W.Field(lat)
    for _ in range(steps):
        x.zero()
        W.CG(D, 1e-10, 10000).apply(b, x)
    print(f"# profile-cg-small: {steps} solves", file=sys.stderr, flush=True)
    lat.close()


def profile_block(W, maxiters):
    """The 24-rhs configuration (32^4, links exp(i 0.3 H) seed 5, kappa 0.125, 12 point + 12
    Gaussian sources) solved by block CG, FP32 inner with FP64 reliable updates, for at most
    `maxiters` iterations (0: to 1e-10)."""
    dims, N = (32, 32, 32, 32), 24
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "smooth", 0.3, 5))
    D = W.Dirac_Wilson_evenodd(0.125, lat)
    b = W.Field(lat, W.EVEN, W.F64, nrhs=N)
    pseed = even_point_seed(W, dims, 8)
    for sc in range(12):
        pt, _ = W.gen_point(dims, pseed, sc // 3, sc % 3)
        b.upload(pt, rhs=sc)
    for j in range(12):
        b.upload(W.gen_spinor(dims, 100 + j, W.EVEN), rhs=12 + j)
    lat.set_option(W.OPT_BLOCK_KRYLOV, 1)
    x = W.Field(lat, W.EVEN, W.F64, nrhs=N)
    t = time.perf_counter()
    st = W.CG(D, 1e-10, maxiters if maxiters > 0 else 10000, precision="f32", delta=1e-3,
              update_every=100).apply(b, x)
    print(f"# profile-block: {st[0]['iterations']} iterations, {1e3 * (time.perf_counter() - t):.1f} ms",
          file=sys.stderr, flush=True)
    lat.close()


def profile_mre(W, steps):
    """MRE guesses at 32^4 (links exp(i 0.3 H) seed 3, kappa 0.124) from 4 and from 8 Gaussian
    basis vectors, `steps` times each: the multi-rhs M^dag M on the basis, the inner-product
    pass (k_mre_dots) and the combination (k_mre_combine)."""
    dims = DIMS
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "smooth", 0.3, 3))
    D = W.Dirac_Wilson_evenodd(0.124, lat)
    chi = W.Field(lat).upload(W.gen_spinor(dims, 600, W.EVEN))
    psi = W.Field(lat)
    for n in (4, 8):
        B = W.Field(lat, W.EVEN, W.F64, n)
        for j in range(n):
            B.upload(W.gen_spinor(dims, 610 + j, W.EVEN), rhs=j)
        for _ in range(steps):
            W.MRE_guess(psi, chi, D, B, n)
        print(f"# profile-mre: {steps} guesses from {n} vectors", file=sys.stderr, flush=True)
    lat.close()


def profile_force(W, steps):
    """Force steps at 32^4 (links exp(i 0.3 H) seed 3, kappa 0.124) from a solved psi, `steps`
    times, then `steps` force evaluations with the force written out: the hopping launches
    for the odd temporaries and the force kernel in its two modes."""
    dims = DIMS
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "smooth", 0.3, 3))
    D = W.Dirac_Wilson_evenodd(0.124, lat)
    chi = W.Field(lat).upload(W.gen_spinor(dims, 620, W.EVEN))
    psi = W.Field(lat)
    W.CG(D, 1e-12).apply(chi, psi)
    for _ in range(steps):
        W.force_step(D, chi, psi, 0.01, accuracy=1e-10)
    for _ in range(steps):
        W.fermion_force(D, chi, psi, 1)
    print(f"# profile-force: {steps} force steps and {steps} forces", file=sys.stderr, flush=True)
    lat.close()


def profile_serve(W):
    """The closed batch of the serving configuration (24 requests), FP16 inner."""
    rng = np.random.default_rng(11)
    gauges, sources = serving_pool(W, rng, 3)
    pool = _SolutionPool(W)
    closed = draw_requests(rng, 24, 3)
    for d in SERVE_VOLUMES:  # the client's page-locked solution buffers, before the batch
        bufs = [pool.get(d) for _ in range(sum(1 for r in closed if r["dims"] == d))]
        for b in bufs:
            pool.put(d, b)
    with W.Server() as S:
        res, wall = run_served(W, S, closed, gauges, sources, pool, precision="f16")
    print(f"# profile-serve: {len(closed)} requests, {sum(r['nrhs'] for r in closed)} rhs in "
          f"{wall:.2f} s", file=sys.stderr, flush=True)


def config_io(W, lat, U, reps=3):
    """HILA configuration file of the 32^4 gauge field (GaugeField::config_write / config_read,
    604 MB) written and read back through the library: the file is streamed through pinned
    chunks and reordered on the GPU.  Best of `reps`; the read follows the write, so the file
    comes from the page cache.  Checks that the field read back is the uploaded one bit for bit."""
    import tempfile
    fd, path = tempfile.mkstemp(suffix=".cfg")
    os.close(fd)
    try:
        tw, tr = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            lat.config_write(path)
            tw.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            lat.config_read(path)
            tr.append(time.perf_counter() - t0)
        size = os.path.getsize(path)
        exact = lat.download_gauge().tobytes() == U.tobytes()
    finally:
        os.remove(path)
    return {"file_bytes": size, "write_ms": round(1e3 * min(tw), 1), "read_ms": round(1e3 * min(tr), 1),
            "write_GBps": round(size / min(tw) / 1e9, 2), "read_GBps": round(size / min(tr) / 1e9, 2),
            "read_back_bit_exact": exact, "file": "temporary file, read from the page cache"}


def hmc_solver_config(W, nbasis=4, nseq=7, mh=0.1, noise=1e-3, kappa=0.124, tol=1e-10):
    """The solver steps of HILA's HMC fermion actions at 32^4 (fermion_field.h): CG on the
    Hasenbusch operator M + mh, and a chronological sequence of solves (right-hand sides
    c0 + k c1 + k^2 c2 + noise) each started from the MRE guess of the last `nbasis`
    solutions (wl_mre_guess), against starting from zero.  Links exp(i 0.3 H) seed 3."""
    dims = DIMS
    lat = W.Lattice(dims)
    lat.upload_gauge(W.gen_gauge(dims, "smooth", 0.3, 3))
    D = W.Dirac_Wilson_evenodd(kappa, lat)
    H = W.Hasenbusch_operator(D, mh)
    src = [W.gen_spinor(dims, 500 + i, W.EVEN) for i in range(3)]
    rng = np.random.default_rng(5)
    b = W.Field(lat)
    x = W.Field(lat)
    out = {"lattice": list(dims), "kappa": kappa, "tolerance": tol}
    # Hasenbusch-shifted CG
    b.upload(src[0])
    t = []
    for _ in range(3):
        x.zero()
        t0 = time.perf_counter()
        st = W.CG(H, tol).apply(b, x)[0]
        t.append(time.perf_counter() - t0)
    D_st = W.CG(D, tol).apply(b, x.zero())[0]
    out["hasenbusch_cg"] = {"mh": mh, "ms": round(1e3 * min(t), 2), "iterations": st["iterations"],
                            "true_resid": st["true_resid"], "unshifted_iterations": D_st["iterations"]}
    # chronological solves with and without the MRE guess
    basis = W.Field(lat, W.EVEN, W.F64, nbasis)
    seq = []
    for k in range(nseq):
        h = src[0] + k * src[1] + k * k * src[2]
        h[: lat.Vh] += noise * np.linalg.norm(h) / np.sqrt(lat.Vh * 24) * rng.standard_normal((lat.Vh, 24))
        seq.append(h)
    it_zero, it_mre, ms_zero, ms_mre, ms_guess = [], [], [], [], []
    for k, h in enumerate(seq):
        b.upload(h)
        x.zero()
        t0 = time.perf_counter()
        st0 = W.CG(D, tol).apply(b, x)[0]
        ms_zero.append(1e3 * (time.perf_counter() - t0))
        it_zero.append(st0["iterations"])
        n = min(k, nbasis)
        t0 = time.perf_counter()
        W.MRE_guess(x, b, D, basis, n)
        g = 1e3 * (time.perf_counter() - t0)
        st1 = W.CG(D, tol).apply(b, x)[0]
        ms_mre.append(1e3 * (time.perf_counter() - t0))
        ms_guess.append(g)
        it_mre.append(st1["iterations"])
        assert st1["converged"] == 1
        # save_new_solution, newest first: shift the ring down, insert at 0
        for j in range(min(k, nbasis - 1), 0, -1):
            basis.copy_rhs_from(j, basis, j - 1)
        basis.copy_rhs_from(0, x)
    # fermion_action::force_step with psi already solved (CG stops at once): the force pass
    mom = W.Field(lat)
    b.upload(seq[-1])
    chi = W.Field(lat).upload(seq[-1])
    W.CG(D, 1e-12).apply(chi, x.zero())
    fs = []
    for _ in range(5):
        t0 = time.perf_counter()
        st = W.force_step(D, chi, x, 0.01, accuracy=1e-10)
        fs.append(1e3 * (time.perf_counter() - t0))
    out["force_step"] = {"ms_with_solved_psi": round(min(fs), 3), "cg_iterations": st["iterations"]}
    del mom
    tail = slice(nbasis, nseq)
    out["mre_chronological"] = {
        "nbasis": nbasis, "sequence": nseq, "noise": noise,
        "iterations_from_zero": it_zero, "iterations_from_guess": it_mre,
        "guess_ms": round(float(np.median(ms_guess[1:])), 2),
        "solve_ms_from_zero": round(float(np.median(ms_zero[tail])), 2),
        "solve_ms_from_guess_incl_guess": round(float(np.median(ms_mre[tail])), 2),
        "speedup": round(float(np.median(ms_zero[tail]) / np.median(ms_mre[tail])), 3)}
    lat.close()
    return out


def cpu_baseline(W, U, psi, seconds=8.0):
    """The oracle's HILA-structured hopping term (projection pass + per-direction
    accumulation, wilson.h:53-95) on the host cores, on a bounded number of applications;
    then the same on one core, and the oracle's CG (conjugate_gradient.h:46-104) to 1e-10 at
    8^4 (Haar seed 1, kappa 0.12, source seed 2), at 16^4 (links exp(i 0.3 H) seed 3, kappa
    0.125, source seed 4: the GPU's 16^4 CG configuration) and at 32^4 (the Dslash's field and
    source)."""
    from oracle import oracle as O
    orc = O.Oracle(DIMS)
    orc.hop_set(U, KAPPA, psi, O.ODD, 1)  # warm-up (page-in, thread pool)

    def sample(limit_s, limit_n):
        n, t0 = 0, time.perf_counter()
        while True:
            orc.hop_set(U, KAPPA, psi, O.ODD, 1)
            n += 1
            el = time.perf_counter() - t0
            if el >= limit_s or n >= limit_n:
                return n, el

    n, el = sample(seconds, 200)
    vh = orc.Vh
    cores = O.set_threads(0)
    O.set_threads(1)
    n1, el1 = sample(3.0, 20)
    O.set_threads(cores)

    def cg(dims, Ug, kappa, src):
        o = O.Oracle(dims)
        t0 = time.perf_counter()
        x, steps, _ = o.cg(Ug, kappa, src, accuracy=1e-10)
        ms = 1e3 * (time.perf_counter() - t0)
        b = src.copy()
        b[o.Vh:] = 0
        r = b - o.meo(Ug, kappa, o.meo(Ug, kappa, x, 1), -1)
        return {"ms_per_solve": round(ms, 1), "iterations": int(steps),
                "true_resid": float(np.linalg.norm(r) / np.linalg.norm(b))}

    small = cg(SMALL_DIMS, W.gen_gauge(SMALL_DIMS, "haar", 0, 1), SMALL_KAPPA,
               W.gen_spinor(SMALL_DIMS, 2, W.EVEN))
    d16 = (16, 16, 16, 16)
    small16 = cg(d16, W.gen_gauge(d16, "smooth", 0.3, 3), 0.125, W.gen_spinor(d16, 4, W.EVEN))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    big = cg(DIMS, U, KAPPA, psi)
    return {
        "value": round(BYTES_PER_SITE_F64 * vh * n / el / 1e9, 3),
        "unit": "GB/s",
        "cores": cores,
        "cpu_model": model,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": f"{n} half-lattice hopping applications (Dirac_Wilson_hop_set restated, "
                  f"HILA pass structure) at 32^4 FP64, {el:.1f} s, OpenMP {cores} threads",
        "ms_per_apply": round(1e3 * el / n, 3),
        "single_core": {"value": round(BYTES_PER_SITE_F64 * vh * n1 / el1 / 1e9, 3),
                        "unit": "GB/s", "ms_per_apply": round(1e3 * el1 / n1, 1),
                        "sample": f"{n1} applications, {el1:.1f} s, one thread"},
        "cg_8x8x8x8": small,
        "cg_16x16x16x16": small16,
        "cg_32x32x32x32": big,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cg-solves", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-big", action="store_true",
                    help="skip the 48^3x96 mixed-precision configuration")
    ap.add_argument("--no-multi", action="store_true",
                    help="skip the batched multi-rhs configuration")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving configuration")
    ap.add_argument("--no-config-io", action="store_true",
                    help="skip writing and reading the 32^4 configuration file")
    ap.add_argument("--no-hmc", action="store_true",
                    help="skip the Hasenbusch CG and MRE-guess configuration")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--profile-multi", action="store_true",
                    help="short run for rocprofv3: --steps autotuned multi-rhs hopping launches "
                         "(12 rhs, FP16 then FP32 storage) at 32^4, then --steps copies of one "
                         "12-rhs FP16 field; prints no JSON line")
    ap.add_argument("--profile-cg-small", action="store_true",
                    help="short run for rocprofv3: --steps FP64 CG solves at 8^4 (kernel by kernel)")
    ap.add_argument("--profile-block", action="store_true",
                    help="short run for rocprofv3: the 24-rhs 32^4 configuration by block CG, at most "
                         "--steps iterations (0: to 1e-10)")
    ap.add_argument("--profile-mre", action="store_true",
                    help="short run for rocprofv3: MRE guesses at 32^4 from 4 and 8 vectors")
    ap.add_argument("--profile-force", action="store_true",
                    help="short run for rocprofv3: force steps and forces at 32^4")
    ap.add_argument("--profile-serve", action="store_true",
                    help="short run for rocprofv3: the closed batch of 24 served requests, FP16 "
                         "inner; prints no JSON line")
    ap.add_argument("--two-rows", action="store_true",
                    help="with --profile: the hopping launches and the CG read two-row links "
                         "(WL_OPT_LINK_ROWS = 2)")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: --steps hopping launches, then --steps "
                         "plain field copies (12 planes of V/2 complex doubles each way) as "
                         "the byte-counter calibration; prints no JSON line")
    ap.add_argument("--lattice", type=int, nargs=4, default=None, metavar=("LX", "LY", "LZ", "LT"),
                    help="with --profile: the lattice of the profiled launches (default 32 32 32 32)")
    a = ap.parse_args(argv)
    if a.lattice is not None and not a.profile:
        ap.error("--lattice applies to --profile only; the benchmark runs its fixed workloads")
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    from hila_amd import wilson as W

    if a.profile_multi:
        profile_multi(W, a.steps)
        return
    if a.profile_serve:
        profile_serve(W)
        return
    if a.profile_mre:
        profile_mre(W, a.steps)
        return
    if a.profile_force:
        profile_force(W, a.steps)
        return
    if a.profile_cg_small:
        profile_cg_small(W, a.steps)
        return
    if a.profile_block:
        profile_block(W, a.steps)
        return

    t_setup = time.perf_counter()
    dims = tuple(a.lattice) if a.lattice is not None else DIMS
    U = W.gen_gauge(dims, "smooth", 0.3, 3)
    psi = W.gen_spinor(dims, 4, W.EVEN)
    lat = W.Lattice(dims, antiperiodic_t=True)
    lat.upload_gauge(U)
    D = W.Dirac_Wilson_evenodd(KAPPA, lat)
    fin = W.Field(lat, W.EVEN).upload(psi)
    fodd = W.Field(lat, W.ODD)
    vh = lat.Vh
    print(f"# setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    if a.profile:
        if a.two_rows:
            lat.set_option(W.OPT_LINK_ROWS, 2)
        fcopy = W.Field(lat, W.EVEN)
        for _ in range(a.steps):
            D.hop_set(fin, fodd, 1)
        lat.sync()
        for _ in range(a.steps):
            fcopy.copy_from(fin)
        lat.sync()
        if a.cg_solves > 0:
            fM = W.Field(lat, W.EVEN)
            for _ in range(5):
                D.apply(fin, fM)
            st = W.CG(D, 1e-10, 10000).apply(fin, fM.zero())[0]
            lat.sync()
            print(f"# profile: CG {st['iterations']} iterations", file=sys.stderr, flush=True)
        print(f"# profile: {a.steps} hopping launches, {a.steps} copies of "
              f"{12 * 16 * vh} bytes", file=sys.stderr, flush=True)
        return

    # --- Dslash: W warm-up, K timed steps, synchronised on both sides ---
    for _ in range(a.warmup):
        D.hop_set(fin, fodd, 1)
    lat.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        D.hop_set(fin, fodd, 1)
    lat.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / a.steps
    # the kernel's own average duration, HIP events on the library's stream
    ms_kernel = D.time(fin, fodd, "hop", 1, warmup=5, iters=max(a.steps, 100))
    ms_M = D.time(fin, W.Field(lat, W.EVEN), "M", 1, warmup=3, iters=50)
    # device-copy bandwidth on the same byte count: one 4-rhs FP64 even field (402 MB) copied
    # into another, 805 MB moved, as many bytes as one hopping application
    ca, cb = W.Field(lat, W.EVEN, W.F64, 4), W.Field(lat, W.EVEN, W.F64, 4)
    ms_copy = D.time(ca, cb, "copy", warmup=5, iters=50)
    copy_bytes = 2 * 4 * 192 * vh
    del ca, cb
    alg_bytes = BYTES_PER_SITE_F64 * vh
    gbs_wall = alg_bytes / (ms_per_step * 1e-3) / 1e9
    gbs_kernel = alg_bytes / (ms_kernel * 1e-3) / 1e9
    print(f"# dslash {ms_kernel * 1e3:.1f} us/launch {gbs_kernel:.0f} GB/s", file=sys.stderr,
          flush=True)

    cfg_io = None
    if not a.no_config_io:
        cfg_io = config_io(W, lat, U)
        print(f"# config file: {cfg_io}", file=sys.stderr, flush=True)

    # --- CG solves to 1e-10: FP64 (HILA's algorithm) and mixed FP32 with reliable updates ---
    x64 = W.Field(lat, W.EVEN)
    cg_out = solve_series(W, D, fin, x64, a.cg_solves, "f64")
    xmx = W.Field(lat, W.EVEN)
    cg_mixed = solve_series(W, D, fin, xmx, a.cg_solves, "f32")
    if cg_out and cg_mixed:
        cg_mixed["speedup_vs_f64"] = round(cg_mixed["solves_per_s"] / cg_out["solves_per_s"], 3)
        cg_mixed["solution_rel_diff_vs_f64"] = rel_diff(xmx.download(), x64.download())
    rows = None
    if a.cg_solves > 0:
        rows = two_row_config(W, lat, D, fin, fodd, x64, max(2, a.cg_solves // 2))
        print(f"# two-row links: {rows}", file=sys.stderr, flush=True)
    del x64, xmx
    small = small_cg_config(W, a.cg_solves) if a.cg_solves > 0 else None
    small16 = (small_cg_config(W, a.cg_solves, (16, 16, 16, 16), "smooth", 0.3, 3, 0.125, 4)
               if a.cg_solves > 0 else None)
    big = multi = serve = None
    D = fin = fodd = None
    lat.close()
    if a.cg_solves > 0 and not a.no_big:
        big = mixed_precision_config(W, max(2, a.cg_solves // 2))
    if a.cg_solves > 0 and not a.no_multi:
        multi = multi_rhs_config(W)
    if a.cg_solves > 0 and not a.no_serve:
        serve = serving_config(W, n_poisson=a.serve_requests)
    hmc = None
    if a.cg_solves > 0 and not a.no_hmc:
        hmc = hmc_solver_config(W)
        print(f"# hmc solvers: {hmc}", file=sys.stderr, flush=True)

    traffic = None
    if os.path.exists(PMC_FILE):
        with open(PMC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(gbs_wall, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded exp(i*0.3*H) links, Gaussian source)",
        "config": {"workload": "wilson_evenodd_dslash_32x32x32x32_f64_1rhs",
                   "lattice": list(DIMS), "kappa": KAPPA, "bc_t": "antiperiodic",
                   "step": "one half-lattice hopping application D_oe (V/2 output sites)"},
        "roofline": {"bound": "hbm", "achieved": round(gbs_kernel, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs_kernel / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": "k_dslash<double,+1,EPI_SET>",
                     "kernel_us": round(ms_kernel * 1e3, 2),
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "copy_gbs": round(copy_bytes / (ms_copy * 1e-3) / 1e9, 1),
                     "frac_of_copy": round(gbs_kernel / (copy_bytes / (ms_copy * 1e-3) / 1e9), 4)},
        "effective_gbs_no_reuse": round(BYTES_NO_REUSE_F64 * vh / (ms_kernel * 1e-3) / 1e9, 1),
        "gflops": round(FLOPS_PER_SITE * vh / (ms_kernel * 1e-3) / 1e9, 1),
        "meo_us": round(ms_M * 1e3, 2),
        "cg": cg_out,
        "cg_mixed_f32": cg_mixed,
        "cg_8x8x8x8": small,
        "cg_16x16x16x16": small16,
        "two_row_links_32x32x32x32": rows,
        "mixed_48x48x48x96": big,
        "multi_rhs_32x32x32x32": multi,
        "serving": serve,
        "config_file_32x32x32x32": cfg_io,
        "hmc_solvers_32x32x32x32": hmc,
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(W, U, np.ascontiguousarray(psi))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
