"""Python mirror of HILA's Wilson solve interface over the C ABI of libhila_wilson.so.

The class and method names follow the reference call surface:

    Dirac_Wilson_evenodd(kappa, gauge)          libraries/dirac/wilson.h:234-269
        .apply(in, out) / .dagger(in, out)      wilson.h:272-289
        .par == EVEN, .kappa
    CG(op, accuracy=1e-12, maxiters=10000)      libraries/dirac/conjugate_gradient.h:19-40
        .apply(in, out)   out = initial guess & result                     :46-104

Fields here are device-resident (`Field`), created on a `Lattice` context; host data is
exchanged in HILA's AoS layout (sites x 24 doubles, EVEN_SITES_FIRST order).  Every call goes
to the HIP library; there is no CPU fallback: without the built library, or without a GPU,
the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libhila_wilson.so")

# status codes and enums of include/wilson_mi355x.h
WL_OK, WL_E_ARG, WL_E_OOM, WL_E_DEVICE, WL_E_NOCONV, WL_E_STATE, WL_E_NOTIMPL, WL_E_BUSY = (
    0, -1, -2, -3, -4, -5, -6, -7)
EVEN, ODD = 1, 2
F64, F32, F16 = 0, 1, 2
OPT_TRAVERSAL_CHUNK, OPT_BLOCK, OPT_RHS_BLOCK, OPT_RHS_LAUNCH, OPT_PLANE_PAD, OPT_LINK_LOAD = 1, 2, 3, 4, 5, 6
OPT_LDS_RESERVE, OPT_GRAPHS, OPT_PIPELINE, OPT_LINK_ROWS, OPT_CG_PERSISTENT, OPT_SMALL_HOP = 7, 8, 9, 10, 11, 12
OPT_BLOCK_KRYLOV = 13

EXPORTED_SYMBOLS = [
    "wl_version", "wl_last_error", "wl_create", "wl_destroy", "wl_sync", "wl_volume",
    "wl_gauge_upload", "wl_field_create", "wl_field_destroy", "wl_field_upload",
    "wl_field_download", "wl_field_zero", "wl_field_copy", "wl_field_norm2", "wl_dslash_eo", "wl_meo_apply",
    "wl_cg_solve", "wl_cg_solve_mixed", "wl_set_option", "wl_get_option", "wl_time_op", "wl_gen_gauge", "wl_gen_spinor", "wl_gen_point",
    "wl_server_create", "wl_server_submit", "wl_server_wait", "wl_server_info_get", "wl_server_destroy",
    "wl_host_alloc", "wl_host_free", "wl_autotune", "wl_server_set_option",
    "wl_gauge_download", "wl_config_header", "wl_config_read", "wl_config_write",
    "wl_hasenbusch_apply", "wl_cg_solve_ex", "wl_mre_guess", "wl_field_copy_rhs",
    "wl_fermion_force", "wl_momentum_upload", "wl_momentum_download", "wl_fermion_force_step",
    "wl_hasenbusch2_force_step", "wl_cg_solve_noalloc", "wl_fp16_forms", "wl_server_provision",
]


class WilsonError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class CGStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("rr_rel", C.c_double),
                ("true_resid", C.c_double), ("ms", C.c_double), ("reliable_updates", C.c_int),
                ("pad_", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad_"}


class SolveRequest(C.Structure):
    _fields_ = [("dims", C.c_int * 4), ("antiperiodic_t", C.c_int), ("nrhs", C.c_int),
                ("gauge", C.c_void_p * 4), ("source", C.c_void_p), ("solution", C.c_void_p),
                ("kappa", C.c_double), ("tolerance", C.c_double), ("maxiters", C.c_int),
                ("inner_prec", C.c_int), ("delta", C.c_double), ("update_every", C.c_int),
                ("config_path", C.c_char_p), ("mh", C.c_double)]


class CGParams(C.Structure):
    _fields_ = [("kappa", C.c_double), ("mh", C.c_double), ("accuracy", C.c_double),
                ("maxiters", C.c_int), ("inner_prec", C.c_int), ("delta", C.c_double),
                ("update_every", C.c_int)]


class SolveResult(C.Structure):
    _fields_ = [("status", C.c_int), ("nrhs", C.c_int), ("iterations", C.c_int),
                ("reliable_updates", C.c_int), ("true_resid", C.c_double), ("queue_ms", C.c_double),
                ("latency_ms", C.c_double), ("message", C.c_char * 160)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["message"] = self.message.decode(errors="replace")
        return d


class TuneResult(C.Structure):
    _fields_ = [("rhs_block", C.c_int), ("rhs_launch", C.c_int), ("block", C.c_int),
                ("candidates", C.c_int), ("us_per_rhs", C.c_double), ("us_per_rhs_default", C.c_double),
                ("lds_reserve", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerInfo(C.Structure):
    _fields_ = [("budget_bytes", C.c_size_t), ("bytes_in_use", C.c_size_t), ("slots", C.c_int),
                ("active", C.c_int), ("queued", C.c_int), ("max_active", C.c_int),
                ("completed", C.c_int64), ("slots_created", C.c_int64), ("slots_released", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libhila_wilson.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hila_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        ip4 = C.POINTER(C.c_int)
        L.wl_version.restype = C.c_char_p
        L.wl_last_error.restype = C.c_char_p
        L.wl_create.argtypes = [ip4, i, i, C.POINTER(vp)]
        L.wl_destroy.argtypes = [vp]
        L.wl_sync.argtypes = [vp]
        L.wl_volume.argtypes = [vp]
        L.wl_volume.restype = i64
        L.wl_gauge_upload.argtypes = [vp, C.POINTER(vp)]
        L.wl_field_create.argtypes = [vp, i, i, i, C.POINTER(vp)]
        L.wl_field_destroy.argtypes = [vp]
        L.wl_field_upload.argtypes = [vp, i, vp]
        L.wl_field_download.argtypes = [vp, i, vp]
        L.wl_field_zero.argtypes = [vp]
        L.wl_field_copy.argtypes = [vp, vp]
        L.wl_field_norm2.argtypes = [vp, i, C.POINTER(d)]
        L.wl_dslash_eo.argtypes = [vp, vp, vp, d, i]
        L.wl_meo_apply.argtypes = [vp, vp, vp, d, i]
        L.wl_cg_solve.argtypes = [vp, vp, vp, d, d, i, C.POINTER(CGStats)]
        L.wl_cg_solve_noalloc.argtypes = [vp, vp, vp, d, d, i, C.POINTER(CGStats)]
        L.wl_fp16_forms.argtypes = [i, vp, vp, i, vp]
        L.wl_cg_solve_mixed.argtypes = [vp, vp, vp, d, d, i, i, d, i, C.POINTER(CGStats)]
        L.wl_set_option.argtypes = [vp, i, i]
        L.wl_get_option.argtypes = [vp, i, C.POINTER(i)]
        L.wl_time_op.argtypes = [vp, i, vp, vp, d, i, i, i, C.POINTER(d)]
        L.wl_gen_gauge.argtypes = [ip4, i, d, C.c_uint64, C.POINTER(vp)]
        L.wl_gen_spinor.argtypes = [ip4, C.c_uint64, i, vp]
        L.wl_gen_point.argtypes = [ip4, C.c_uint64, i, i, vp, C.POINTER(i64)]
        L.wl_server_create.argtypes = [i, C.c_size_t, i, C.POINTER(vp)]
        L.wl_server_submit.argtypes = [vp, C.POINTER(SolveRequest), C.POINTER(C.c_uint64)]
        L.wl_server_provision.argtypes = [vp, ip4, i, i, i, i, C.POINTER(i)]
        L.wl_server_wait.argtypes = [vp, C.c_uint64, i, C.POINTER(SolveResult), C.POINTER(CGStats)]
        L.wl_server_info_get.argtypes = [vp, C.POINTER(ServerInfo)]
        L.wl_server_destroy.argtypes = [vp]
        L.wl_host_alloc.argtypes = [C.c_size_t, C.POINTER(vp)]
        L.wl_host_free.argtypes = [vp]
        L.wl_autotune.argtypes = [vp, i, i, C.POINTER(TuneResult)]
        L.wl_server_set_option.argtypes = [vp, i, i]
        L.wl_gauge_download.argtypes = [vp, C.POINTER(vp)]
        L.wl_config_header.argtypes = [C.c_char_p, ip4]
        L.wl_config_read.argtypes = [vp, C.c_char_p]
        L.wl_config_write.argtypes = [vp, C.c_char_p]
        L.wl_hasenbusch_apply.argtypes = [vp, vp, vp, d, d, i]
        L.wl_cg_solve_ex.argtypes = [vp, vp, vp, C.POINTER(CGParams), C.POINTER(CGStats)]
        L.wl_mre_guess.argtypes = [vp, vp, vp, vp, i, d, d]
        L.wl_field_copy_rhs.argtypes = [vp, i, vp, i]
        L.wl_fermion_force.argtypes = [vp, vp, vp, d, i, C.POINTER(vp)]
        L.wl_momentum_upload.argtypes = [vp, C.POINTER(vp)]
        L.wl_momentum_download.argtypes = [vp, C.POINTER(vp)]
        L.wl_fermion_force_step.argtypes = [vp, vp, vp, d, C.POINTER(CGParams), C.POINTER(CGStats)]
        L.wl_hasenbusch2_force_step.argtypes = [vp, vp, vp, d, C.POINTER(CGParams), C.POINTER(CGStats)]
        for name in EXPORTED_SYMBOLS:
            if name not in ("wl_version", "wl_last_error", "wl_volume"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def check(rc, allow=()):
    if rc != WL_OK and rc not in allow:
        raise WilsonError(rc, lib().wl_last_error().decode())
    return rc


def _dims(dims):
    d = tuple(int(x) for x in dims)
    if len(d) != 4:
        raise ValueError("dims must have 4 entries (Lx, Ly, Lz, Lt)")
    return (C.c_int * 4)(*d)


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------ seeded inputs ------

def gen_gauge(dims, kind="smooth", eps=0.3, seed=1):
    """Gauge field in HILA's layout, (4, V, 18) float64.  kind: 'unit', 'haar', 'smooth'
    (exp(i eps H), H traceless Hermitian with unit-variance Gaussian entries) or 'su2_hits'
    (HILA's SU<3>::random(): 16 hits of random SU(2) subgroup matrices)."""
    kinds = {"unit": 0, "haar": 1, "smooth": 2, "su2_hits": 3}
    V = int(np.prod(dims))
    U = np.empty((4, V, 18), dtype=np.float64)
    ptrs = (C.c_void_p * 4)(*[U[m].ctypes.data for m in range(4)])
    check(lib().wl_gen_gauge(_dims(dims), kinds[kind], float(eps), int(seed), ptrs))
    return U


def gen_spinor(dims, seed, parity_mask=EVEN):
    """Gaussian Wilson-vector field (V, 24), zero outside parity_mask (1 even, 2 odd, 3 all)."""
    V = int(np.prod(dims))
    f = np.empty((V, 24), dtype=np.float64)
    check(lib().wl_gen_spinor(_dims(dims), int(seed), int(parity_mask), _ptr(f)))
    return f


def gen_point(dims, seed, spin, colour):
    """Point source (V, 24) at a seeded site; returns (field, EVEN_SITES_FIRST site index)."""
    V = int(np.prod(dims))
    f = np.empty((V, 24), dtype=np.float64)
    site = C.c_int64(0)
    check(lib().wl_gen_point(_dims(dims), int(seed), int(spin), int(colour), _ptr(f),
                             C.byref(site)))
    return f, int(site.value)


# ------------------------------------------------------------------------ objects ------

class Lattice:
    """A lattice geometry on one GPU (wl_ctx): HILA's lattice.setup() for this path."""

    def __init__(self, dims, antiperiodic_t=True, device=0):
        self.dims = tuple(int(x) for x in dims)
        self.V = int(np.prod(self.dims))
        self.Vh = self.V // 2
        h = C.c_void_p()
        check(lib().wl_create(_dims(self.dims), 1 if antiperiodic_t else 0, int(device),
                              C.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise WilsonError(WL_E_STATE, "lattice context closed")
        return self._h

    def upload_gauge(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != (4, self.V, 18):
            raise ValueError(f"gauge must have shape (4, {self.V}, 18)")
        ptrs = (C.c_void_p * 4)(*[U[m].ctypes.data for m in range(4)])
        check(lib().wl_gauge_upload(self.handle, ptrs))

    def download_gauge(self):
        """The device gauge field as (4, V, 18) float64, HILA layout (wl_gauge_download)."""
        U = np.empty((4, self.V, 18), dtype=np.float64)
        ptrs = (C.c_void_p * 4)(*[U[m].ctypes.data for m in range(4)])
        check(lib().wl_gauge_download(self.handle, ptrs))
        return U

    def _links_in(self, P, what):
        P = np.ascontiguousarray(P, dtype=np.float64)
        if P.shape != (4, self.V, 18):
            raise ValueError(f"{what} must have shape (4, {self.V}, 18)")
        return P, (C.c_void_p * 4)(*[P[m].ctypes.data for m in range(4)])

    def upload_momentum(self, P):
        """The HMC momentum (4, V, 18), HILA layout (wl_momentum_upload)."""
        P, ptrs = self._links_in(P, "momentum")
        check(lib().wl_momentum_upload(self.handle, ptrs))

    def download_momentum(self):
        P = np.empty((4, self.V, 18), dtype=np.float64)
        ptrs = (C.c_void_p * 4)(*[P[m].ctypes.data for m in range(4)])
        check(lib().wl_momentum_download(self.handle, ptrs))
        return P

    def config_read(self, path):
        """GaugeField::config_read (gaugefield.h:219-272): the gauge field from a HILA
        configuration file, replicated when the file's lattice divides this one."""
        check(lib().wl_config_read(self.handle, os.fsencode(path)))

    def config_write(self, path):
        """GaugeField::config_write (gaugefield.h:194-217)."""
        check(lib().wl_config_write(self.handle, os.fsencode(path)))

    def sync(self):
        check(lib().wl_sync(self.handle))

    def set_option(self, option, value):
        check(lib().wl_set_option(self.handle, int(option), int(value)))

    def get_option(self, option):
        v = C.c_int(0)
        check(lib().wl_get_option(self.handle, int(option), C.byref(v)))
        return v.value

    def autotune(self, prec, nrhs):
        """Time the multi-rhs kernel's configurations for nrhs right-hand sides and keep the
        fastest (wl_autotune); returns the choice."""
        r = TuneResult()
        check(lib().wl_autotune(self.handle, int(prec), int(nrhs), C.byref(r)))
        return r.as_dict()

    def close(self):
        if self._h is not None:
            lib().wl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Field:
    """A device Wilson-vector field on one parity (nrhs vectors, precision prec)."""

    def __init__(self, lat: Lattice, parity=EVEN, prec=F64, nrhs=1):
        self.lat = lat
        self.parity = parity
        self.prec = prec
        self.nrhs = nrhs
        h = C.c_void_p()
        check(lib().wl_field_create(lat.handle, parity, prec, nrhs, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def _block(self, host):
        host = np.asarray(host, dtype=np.float64)
        Vh = self.lat.Vh
        if host.shape == (self.lat.V, 24):  # full HILA field: take this parity's block
            host = host[:Vh] if self.parity == EVEN else host[Vh:]
        if host.shape != (Vh, 24):
            raise ValueError(f"host field must be (V, 24) or (V/2, 24); got {host.shape}")
        return np.ascontiguousarray(host)

    def upload(self, host, rhs=0):
        b = self._block(host)
        check(lib().wl_field_upload(self._h, rhs, _ptr(b)))
        return self

    def download(self, rhs=0):
        out = np.empty((self.lat.Vh, 24), dtype=np.float64)
        check(lib().wl_field_download(self._h, rhs, _ptr(out)))
        return out

    def to_full(self, rhs=0):
        """HILA full-lattice AoS array (V, 24), other parity zero."""
        full = np.zeros((self.lat.V, 24))
        blk = self.download(rhs)
        if self.parity == EVEN:
            full[: self.lat.Vh] = blk
        else:
            full[self.lat.Vh:] = blk
        return full

    def zero(self):
        check(lib().wl_field_zero(self._h))
        return self

    def copy_from(self, src: "Field"):
        """self = src on the device (HILA's field assignment)."""
        check(lib().wl_field_copy(self._h, src.handle))
        return self

    def copy_rhs_from(self, rhs, src: "Field", src_rhs=0):
        """Right-hand side `rhs` of self = right-hand side `src_rhs` of src."""
        check(lib().wl_field_copy_rhs(self._h, int(rhs), src.handle, int(src_rhs)))
        return self

    def norm2(self, rhs=0):
        d = C.c_double(0)
        check(lib().wl_field_norm2(self._h, rhs, C.byref(d)))
        return d.value

    def close(self):
        if self._h is not None:
            lib().wl_field_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Dirac_Wilson_evenodd:
    """HILA's Dirac_Wilson_evenodd<SU<3,double>> (wilson.h:234-312) on the GPU.
    `gauge` is the Lattice context the gauge field was uploaded to (the operator holds a
    non-owning reference, as HILA's holds Field<SU> (&gauge)[NDIM])."""

    par = EVEN

    def __init__(self, kappa, gauge: Lattice):
        self.kappa = float(kappa)
        self.lat = gauge

    mh = 0.0

    def apply(self, fin: Field, fout: Field):
        check(lib().wl_meo_apply(self.lat.handle, fout.handle, fin.handle, self.kappa, 1))

    def dagger(self, fin: Field, fout: Field):
        check(lib().wl_meo_apply(self.lat.handle, fout.handle, fin.handle, self.kappa, -1))

    def hop_set(self, fin: Field, fout: Field, sign=1):
        """Dirac_Wilson_hop_set: fout = -kappa H_sign fin, fout of the other parity."""
        check(lib().wl_dslash_eo(self.lat.handle, fout.handle, fin.handle, self.kappa, sign))

    def time(self, fin: Field, fout: Field, what="hop", sign=1, warmup=3, iters=100):
        """Average device milliseconds per call (HIP events on the library's stream) of the
        hopping term (what='hop'), of M (what='M') or of a plain copy fout = fin
        (what='copy')."""
        ms = C.c_double(0)
        code = {"hop": 0, "M": 1, "copy": 2}[what]
        check(lib().wl_time_op(self.lat.handle, code, fout.handle,
                               fin.handle, self.kappa, sign, warmup, iters, C.byref(ms)))
        return ms.value


class Hasenbusch_operator:
    """HILA's Hasenbusch_operator<Dirac_Wilson_evenodd> (Hasenbusch.h:15-53): D + mh, with
    apply / dagger = (M + mh) / (M^dag + mh) on even sites."""

    par = EVEN

    def __init__(self, D: Dirac_Wilson_evenodd, hasenbusch_parameter):
        self.D = D
        self.kappa = D.kappa
        self.lat = D.lat
        self.mh = float(hasenbusch_parameter)
        self.h_parameter = self.mh

    def apply(self, fin: Field, fout: Field):
        check(lib().wl_hasenbusch_apply(self.lat.handle, fout.handle, fin.handle, self.kappa,
                                        self.mh, 1))

    def dagger(self, fin: Field, fout: Field):
        check(lib().wl_hasenbusch_apply(self.lat.handle, fout.handle, fin.handle, self.kappa,
                                        self.mh, -1))


def MRE_guess(psi: Field, chi: Field, D, basis: Field, nbasis=None):
    """HILA's MRE_guess(psi, chi, D, old_chi_inv) (libraries/hmc/MRE_guess.h:8-87): psi =
    the minimal-residual combination, for A = D^dag D, of the first nbasis right-hand sides of
    `basis` (previous solutions).  D is a Dirac_Wilson_evenodd or Hasenbusch_operator."""
    n = basis.nrhs if nbasis is None else int(nbasis)
    check(lib().wl_mre_guess(D.lat.handle, psi.handle, chi.handle, basis.handle, n, D.kappa,
                             getattr(D, "mh", 0.0)))
    return psi


def fermion_force(D, chi: Field, psi: Field, sign=1):
    """D.force(chi, psi, force, sign) (wilson.h:294-311): (4, V, 18) float64, HILA layout; the
    derivative of Re <chi, M_sign psi> with respect to the links."""
    F = np.empty((4, D.lat.V, 18), dtype=np.float64)
    ptrs = (C.c_void_p * 4)(*[F[m].ctypes.data for m in range(4)])
    check(lib().wl_fermion_force(D.lat.handle, chi.handle, psi.handle, D.kappa, int(sign), ptrs))
    return F


def force_step(D, chi: Field, psi: Field, eps, accuracy=1e-12, maxiters=10000, precision="f64",
               delta=0.1, update_every=100):
    """fermion_action::force_step(eps) (fermion_field.h:146-167) on the momentum the lattice
    holds: psi (initial guess in, CG solution out), then momentum += P(U force).  Returns the
    CG stats dict."""
    p = CGParams(kappa=D.kappa, mh=getattr(D, "mh", 0.0), accuracy=float(accuracy),
                 maxiters=int(maxiters), inner_prec={"f64": F64, "f32": F32, "f16": F16}[precision],
                 delta=float(delta), update_every=int(update_every))
    st = CGStats()
    check(lib().wl_fermion_force_step(D.lat.handle, chi.handle, psi.handle, float(eps), C.byref(p),
                                      C.byref(st)), allow=(WL_E_NOCONV,))
    return st.as_dict()


def hasenbusch2_force_step(H, chi: Field, psi: Field, eps, accuracy=1e-12, maxiters=10000,
                           precision="f64", delta=0.1, update_every=100):
    """Hasenbusch_action_2::force_step(eps) (fermion_field.h:319-346) for the Hasenbusch
    operator H = D + mh: psi = (D^dag D)^-1 H^dag chi, Mpsi = D psi - chi, momentum +=
    P(U force).  Returns the CG stats dict."""
    p = CGParams(kappa=H.kappa, mh=H.mh, accuracy=float(accuracy), maxiters=int(maxiters),
                 inner_prec={"f64": F64, "f32": F32, "f16": F16}[precision],
                 delta=float(delta), update_every=int(update_every))
    st = CGStats()
    check(lib().wl_hasenbusch2_force_step(H.lat.handle, chi.handle, psi.handle, float(eps),
                                          C.byref(p), C.byref(st)), allow=(WL_E_NOCONV,))
    return st.as_dict()


class CG:
    """HILA's CG<Op> (conjugate_gradient.h:19-105): solves (M^dag M) out = in, for Op a
    Dirac_Wilson_evenodd or a Hasenbusch_operator of one (then M + mh).

    precision="f64" runs HILA's algorithm in FP64.  precision="f32" runs the mixed-precision
    solver with reliable updates (inner operator and Krylov vectors in FP32, solution and
    reductions in FP64, FP64 true residual recomputed at every update and at the end);
    precision="f16" the same with the inner fields stored in FP16 with a per-site scale (FP32
    arithmetic); delta and update_every set when reliable updates happen."""

    def __init__(self, op, accuracy=1e-12, maxiters=10000,
                 precision="f64", delta=1e-3, update_every=100):
        self.M = op
        self.accuracy = float(accuracy)
        self.maxiters = int(maxiters)
        self.precision = precision
        self.delta = float(delta)
        self.update_every = int(update_every)

    def apply(self, fin: Field, fout: Field, raise_on_noconv=False):
        """Runs CG; fout holds the initial guess on entry.  Returns per-rhs stats dicts.
        Like HILA, reaching maxiters is not an error unless raise_on_noconv is set."""
        stats = (CGStats * fin.nrhs)()
        mh = getattr(self.M, "mh", 0.0)
        if mh != 0.0:
            if self.precision not in ("f64", "f32", "f16"):
                raise ValueError(f"precision {self.precision!r} has no shifted-operator solver")
            p = CGParams(kappa=self.M.kappa, mh=mh, accuracy=self.accuracy,
                         maxiters=self.maxiters,
                         inner_prec={"f64": F64, "f32": F32, "f16": F16}[self.precision],
                         delta=self.delta, update_every=self.update_every)
            rc = lib().wl_cg_solve_ex(self.M.lat.handle, fout.handle, fin.handle, C.byref(p),
                                      stats)
        elif self.precision == "f64":
            rc = lib().wl_cg_solve(self.M.lat.handle, fout.handle, fin.handle, self.M.kappa,
                                   self.accuracy, self.maxiters, stats)
        else:
            inner = {"f32": F32, "f16": F16, "f64-mixed": F64}[self.precision]
            rc = lib().wl_cg_solve_mixed(self.M.lat.handle, fout.handle, fin.handle,
                                         self.M.kappa, self.accuracy, self.maxiters, inner,
                                         self.delta, self.update_every, stats)
        check(rc, allow=() if raise_on_noconv else (WL_E_NOCONV,))
        return [s.as_dict() for s in stats]


# ------------------------------------------------------------------------ serving ------

class HostBuffer:
    """Page-locked host memory (wl_host_alloc) viewed as a float64 numpy array; request
    buffers in it are copied by DMA without staging."""

    def __init__(self, shape):
        self.shape = tuple(int(x) for x in shape)
        n = int(np.prod(self.shape))
        p = C.c_void_p()
        check(lib().wl_host_alloc(n * 8, C.byref(p)))
        self._p = p
        self.array = np.ctypeslib.as_array((C.c_double * n).from_address(p.value)).reshape(self.shape)

    def close(self):
        if self._p is not None and self._p.value:
            self.array = None
            lib().wl_host_free(self._p)
        self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server: CG solves of independent requests (own lattice size, gauge field,
    kappa, tolerance and 1..64 right-hand sides) admitted by HBM budget and run concurrently
    on one GPU (wl_server_*).  The arrays of a request must stay alive and unmodified until
    wait() returns for it; the server keeps references to them until then."""

    def __init__(self, device=0, budget_bytes=0, max_active=0):
        h = C.c_void_p()
        check(lib().wl_server_create(int(device), int(budget_bytes), int(max_active), C.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, dims, gauge, source, solution, kappa, tolerance, maxiters=10000,
               precision="f32", delta=0.1, update_every=100, antiperiodic_t=True, mh=0.0):
        """gauge: (4, V, 18) float64, or the path of a HILA configuration file; source,
        solution: (nrhs, V/2, 24) float64 (even sites).  precision 'f64' (FP64 CG), 'f32' or
        'f16' (mixed, reliable updates).  Returns a ticket."""
        dims = tuple(int(x) for x in dims)
        V = int(np.prod(dims))
        config = None
        if isinstance(gauge, (str, bytes, os.PathLike)):
            config = os.fsencode(gauge)
        elif gauge.shape != (4, V, 18) or gauge.dtype != np.float64 or not gauge.flags.c_contiguous:
            raise ValueError("gauge must be a C-contiguous float64 array (4, V, 18) or a path")
        for a in (source, solution):
            if a.ndim != 3 or a.shape[1:] != (V // 2, 24) or a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("source/solution must be C-contiguous float64 arrays (nrhs, V/2, 24)")
        if source.shape[0] != solution.shape[0]:
            raise ValueError("source and solution differ in nrhs")
        r = SolveRequest()
        r.dims = (C.c_int * 4)(*dims)
        r.antiperiodic_t = 1 if antiperiodic_t else 0
        r.nrhs = int(source.shape[0])
        if config is None:
            for m in range(4):
                r.gauge[m] = gauge[m].ctypes.data
        else:
            r.config_path = config
        r.source = source.ctypes.data
        r.solution = solution.ctypes.data
        r.kappa = float(kappa)
        r.tolerance = float(tolerance)
        r.maxiters = int(maxiters)
        r.inner_prec = {"f64": F64, "f32": F32, "f16": F16}[precision]
        r.delta = float(delta)
        r.update_every = int(update_every)
        r.mh = float(mh)
        t = C.c_uint64(0)
        check(lib().wl_server_submit(self._h, C.byref(r), C.byref(t)))
        self._keep[t.value] = (gauge, source, solution)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """(result dict, per-rhs stats dicts) of a finished request; raises WilsonError with
        WL_E_BUSY if it does not finish within timeout_ms."""
        res = SolveResult()
        arrays = self._keep.get(ticket)
        nrhs = arrays[1].shape[0] if arrays is not None else 64
        stats = (CGStats * nrhs)()
        check(lib().wl_server_wait(self._h, int(ticket), int(timeout_ms), C.byref(res), stats))
        self._keep.pop(ticket, None)
        return res.as_dict(), [s.as_dict() for s in stats[: res.nrhs]]

    def set_option(self, option, value):
        """A kernel option (OPT_*) for every slot; only before the first submission."""
        check(lib().wl_server_set_option(self._h, int(option), int(value)))

    def provision(self, dims, nrhs, count, precision="f32", antiperiodic_t=True):
        """Create up to `count` idle slots, with all their device memory, for requests of this
        lattice, inner precision and at most nrhs rhs, within the HBM budget (wl_server_provision).
        Returns the number created."""
        made = C.c_int(0)
        check(lib().wl_server_provision(self._h, _dims(dims), 1 if antiperiodic_t else 0,
                                        {"f64": F64, "f32": F32, "f16": F16}[precision], int(nrhs),
                                        int(count), C.byref(made)))
        return made.value

    def info(self):
        i = ServerInfo()
        check(lib().wl_server_info_get(self._h, C.byref(i)))
        return i.as_dict()

    def close(self):
        if self._h is not None:
            lib().wl_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
