"""ctypes loader for the C restatement of HILA's Wilson operator and CG (wilson_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by bench.py's
CPU-baseline leg, as the checker.  The product (hila_amd/) never imports it.

Fields are full-lattice numpy arrays in HILA's AoS layout: Wilson vectors (V, 24) float64 and
links (4, V, 18) float64, sites in EVEN_SITES_FIRST order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
EVEN, ODD, ALL = 1, 2, 3

_libs = {}


def _lib_path(variant=0):
    name = "libwilson_oracle.so" if variant == 0 else f"libwilson_oracle_plant{variant}.so"
    return os.path.join(_HERE, "_build", name)


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP); force rebuilds everything."""
    subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)


def lib(variant=0):
    """The oracle library; variant n > 0 is the mutation-test build with planted mistake n."""
    if variant not in _libs:
        if not os.path.exists(_lib_path(variant)):
            build()
        L = C.CDLL(_lib_path(variant))
        vp, dp, i64 = C.c_void_p, C.c_void_p, C.c_int64
        L.olat_create.restype = vp
        L.olat_create.argtypes = [C.POINTER(C.c_int), C.c_int]
        L.olat_destroy.argtypes = [vp]
        L.olat_volume.restype = i64
        L.olat_volume.argtypes = [vp]
        L.olat_neighbour.restype = i64
        L.olat_neighbour.argtypes = [vp, C.c_int, i64]
        for fn in ("oracle_hop", "oracle_hop_set", "oracle_hop_set_fused"):
            getattr(L, fn).argtypes = [vp, dp, C.c_double, dp, dp, C.c_int, C.c_int]
        L.oracle_meo.argtypes = [vp, dp, C.c_double, dp, dp, C.c_int]
        L.oracle_hasenbusch.argtypes = [vp, dp, C.c_double, C.c_double, dp, dp, C.c_int]
        L.oracle_calc_force.argtypes = [vp, C.c_double, dp, dp, dp, C.c_int, C.c_int]
        L.oracle_force.argtypes = [vp, dp, C.c_double, dp, dp, dp, C.c_int]
        L.oracle_cg_h.restype = C.c_int
        L.oracle_cg_h.argtypes = [vp, dp, C.c_double, C.c_double, dp, dp, C.c_double, C.c_int,
                                  C.POINTER(C.c_double)]
        L.oracle_norm2.restype = C.c_double
        L.oracle_norm2.argtypes = [vp, dp, C.c_int]
        L.oracle_set_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_cg.restype = C.c_int
        L.oracle_cg.argtypes = [vp, dp, C.c_double, dp, dp, C.c_double, C.c_int,
                                C.POINTER(C.c_double)]
        _libs[variant] = L
    return _libs[variant]


def set_threads(n):
    """OpenMP threads of the oracle's loops from now on; returns the previous count."""
    return int(lib().oracle_set_threads(int(n)))


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """One lattice geometry of the CPU restatement."""

    def __init__(self, dims, antiperiodic_t=True, variant=0):
        self.dims = tuple(int(d) for d in dims)
        self._L = lib(variant)
        arr = (C.c_int * 4)(*self.dims)
        self._h = self._L.olat_create(arr, 1 if antiperiodic_t else 0)
        if not self._h:
            raise ValueError(f"invalid lattice {self.dims}")
        self.V = int(self._L.olat_volume(self._h))
        self.Vh = self.V // 2

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.olat_destroy(h)
            self._h = None

    def neighbour(self, d, s):
        return int(self._L.olat_neighbour(self._h, d, s))

    def _field(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (self.V, 24)
        return f

    def _links(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4, self.V, 18)
        return U

    def hop_set(self, U, kappa, fin, par, sign, fused=False):
        """Dirac_Wilson_hop_set: out_par = -kappa H_sign in; other parity left at 0."""
        U, fin = self._links(U), self._field(fin)
        out = np.zeros_like(fin)
        fn = self._L.oracle_hop_set_fused if fused else self._L.oracle_hop_set
        fn(self._h, _p(U), kappa, _p(fin), _p(out), par, sign)
        return out

    def hop(self, U, kappa, fin, fout, par, sign):
        """Dirac_Wilson_hop: fout_par -= kappa H_sign in (fout updated in place)."""
        U, fin = self._links(U), self._field(fin)
        assert fout.dtype == np.float64 and fout.flags.c_contiguous
        self._L.oracle_hop(self._h, _p(U), kappa, _p(fin), _p(fout), par, sign)
        return fout

    def meo(self, U, kappa, fin, sign, mh=0.0):
        """Dirac_Wilson_evenodd::apply (sign=+1) / dagger (sign=-1) on a full field; with
        mh != 0, Hasenbusch_operator::apply / dagger (M + mh)."""
        U, fin = self._links(U), self._field(fin)
        out = np.zeros_like(fin)
        if mh == 0.0:
            self._L.oracle_meo(self._h, _p(U), kappa, _p(fin), _p(out), sign)
        else:
            self._L.oracle_hasenbusch(self._h, _p(U), kappa, mh, _p(fin), _p(out), sign)
        return out

    def calc_force(self, kappa, chi, psi, par, sign):
        """Dirac_Wilson_calc_force: (4, V, 18) float64, every site set."""
        chi, psi = self._field(chi), self._field(psi)
        out = np.zeros((4, self.V, 18))
        self._L.oracle_calc_force(self._h, kappa, _p(chi), _p(psi), _p(out), par, sign)
        return out

    def force(self, U, kappa, chi, psi, sign):
        """Dirac_Wilson_evenodd::force(chi, psi, force, sign): (4, V, 18) float64."""
        U, chi, psi = self._links(U), self._field(chi), self._field(psi)
        out = np.zeros((4, self.V, 18))
        self._L.oracle_force(self._h, _p(U), kappa, _p(chi), _p(psi), _p(out), sign)
        return out

    def norm2(self, f, par=EVEN):
        return float(self._L.oracle_norm2(self._h, _p(self._field(f)), par))

    def cg(self, U, kappa, b, x0=None, accuracy=1e-12, maxiters=10000, mh=0.0):
        """CG<Dirac_Wilson_evenodd>(D, accuracy, maxiters).apply(b, x); with mh != 0,
        CG<Hasenbusch_operator<...>>.  Returns (x, steps, rr_rel) with HILA's printed step
        count and relative residue."""
        U, b = self._links(U), self._field(b)
        x = np.zeros_like(b) if x0 is None else self._field(x0).copy()
        rr = C.c_double(0)
        steps = self._L.oracle_cg_h(self._h, _p(U), kappa, mh, _p(b), _p(x), accuracy,
                                    maxiters, C.byref(rr))
        return x, int(steps), float(rr.value)
