"""HILA's gauge-configuration file, restated in numpy.

TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py's CPU leg as the checker of the
library's wl_config_read / wl_config_write / wl_gauge_download.  The product never imports it.

The format, from GaugeField<T>::config_write / config_read
(libraries/plumbing/gaugefield.h:194-272) and Field<T>::write / read
(libraries/plumbing/field_io.h:125-160, 221-312):

  header   7 native-endian int64: config_flag = 394824242 (gaugefield.h:28), NDIM = 4,
           sizeof(T) = 144 for SU<3,double>, then the lattice size L_x, L_y, L_z, L_t
           (gaugefield.h:200-212);
  payload  fdir[0], fdir[1], fdir[2], fdir[3] one after the other (GaugeField::write,
           gaugefield.h:168-172), each the link of every site in lexicographic order with x
           fastest (lattice.h:389-396, field_io.h:144-158), one SU(3) matrix = 9 row-major
           complex doubles = 18 doubles (matrix.h:130).

Reading checks the flag, NDIM, the element size and that every file dimension divides the
lattice's (gaugefield.h:219-262); a smaller file lattice is replicated periodically to fill
the lattice (field_io.h:245-312).  Arrays here are in the library's host layout: links
(4, V, 18) float64, sites in HILA's EVEN_SITES_FIRST order (lattice.cpp:171-186).
"""
from __future__ import annotations

import numpy as np

CONFIG_FLAG = 394824242
NDIM = 4
ELEM_BYTES = 144
HEADER_BYTES = 8 * (3 + NDIM)


class ConfigError(ValueError):
    """Raised where HILA's config_read prints "CONFIG ERROR" and terminates."""


def lex_coords(dims):
    """(V, 4) coordinates of lexicographic index i = x + Lx (y + Ly (z + Lz t))."""
    dims = tuple(int(d) for d in dims)
    i = np.arange(int(np.prod(dims)), dtype=np.int64)
    out = np.empty((i.size, 4), dtype=np.int64)
    for d in range(4):
        out[:, d] = i % dims[d]
        i = i // dims[d]
    return out


def site_of_lex(dims):
    """EVEN_SITES_FIRST index of every lexicographic index: i/2 for even sites, i/2 + V/2 for
    odd ones (parity = sum of coordinates mod 2)."""
    c = lex_coords(dims)
    lex = np.arange(c.shape[0], dtype=np.int64)
    odd = c.sum(axis=1) % 2
    return lex // 2 + odd * (c.shape[0] // 2)


def config_bytes(U, dims):
    """The bytes config_write produces for links U (4, V, 18), even-first order."""
    dims = tuple(int(d) for d in dims)
    U = np.asarray(U, dtype=np.float64)
    V = int(np.prod(dims))
    assert U.shape == (4, V, 18)
    head = np.array([CONFIG_FLAG, NDIM, ELEM_BYTES, *dims], dtype=np.int64).tobytes()
    order = site_of_lex(dims)
    return head + b"".join(np.ascontiguousarray(U[mu][order]).tobytes() for mu in range(4))


def config_write(path, U, dims):
    with open(path, "wb") as f:
        f.write(config_bytes(U, dims))


def config_header(path):
    """The lattice size a configuration file holds, after config_read's header checks."""
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    if len(raw) < HEADER_BYTES:
        raise ConfigError(f"CONFIG ERROR in file {path}: header too short")
    h = np.frombuffer(raw, dtype=np.int64)
    if h[0] != CONFIG_FLAG:
        raise ConfigError(f"CONFIG ERROR in file {path}: wrong id, should be {CONFIG_FLAG} is {h[0]}")
    if h[1] != NDIM:
        raise ConfigError(f"CONFIG ERROR in file {path}: wrong dimensionality, should be {NDIM} is {h[1]}")
    if h[2] != ELEM_BYTES:
        raise ConfigError(f"CONFIG ERROR in file {path}: wrong size of field element, "
                          f"should be {ELEM_BYTES} is {h[2]}")
    if any(d < 1 for d in h[3:]):
        raise ConfigError(f"CONFIG ERROR in file {path}: incorrect lattice dimension")
    return tuple(int(d) for d in h[3:])


def config_read(path, dims):
    """Links (4, V, 18) in even-first order of a lattice of size dims, from the file at path
    (replicated when the file's lattice divides dims)."""
    dims = tuple(int(d) for d in dims)
    insize = config_header(path)
    for d in range(4):
        if dims[d] % insize[d] != 0:
            raise ConfigError(f"CONFIG ERROR in file {path}: incorrect lattice dimension {d} "
                              f"is {insize[d]} should be (or divide){dims[d]}")
    vin = int(np.prod(insize))
    payload = np.fromfile(path, dtype=np.float64, offset=HEADER_BYTES)
    if payload.size < 4 * vin * 18:
        raise ConfigError(f"CONFIG ERROR in file {path}: file too short")
    links_in = payload[: 4 * vin * 18].reshape(4, vin, 18)   # lexicographic in insize
    c = lex_coords(dims) % np.array(insize)
    src = c[:, 0] + insize[0] * (c[:, 1] + insize[1] * (c[:, 2] + insize[2] * c[:, 3]))
    V = int(np.prod(dims))
    U = np.empty((4, V, 18), dtype=np.float64)
    U[:, site_of_lex(dims)] = links_in[:, src]
    return U


def plaquette(U, dims):
    """Average Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag / 3 per plane (mu < nu),
    as a dict {(mu, nu): value}; links (4, V, 18) even-first."""
    dims = tuple(int(d) for d in dims)
    c = lex_coords(dims)
    site = site_of_lex(dims)
    M = np.empty((4, c.shape[0], 3, 3), dtype=complex)
    Uc = np.asarray(U)[..., 0::2] + 1j * np.asarray(U)[..., 1::2]
    M[:] = Uc[:, site].reshape(4, -1, 3, 3)   # now indexed by lexicographic site

    def shift(mu):
        cc = c.copy()
        cc[:, mu] = (cc[:, mu] + 1) % dims[mu]
        return cc[:, 0] + dims[0] * (cc[:, 1] + dims[1] * (cc[:, 2] + dims[2] * cc[:, 3]))

    out = {}
    for mu in range(4):
        for nu in range(mu + 1, 4):
            a = M[mu] @ M[nu][shift(mu)]
            b = M[nu] @ M[mu][shift(nu)]
            out[(mu, nu)] = float(np.einsum("sij,sij->", a, b.conj()).real / (3 * c.shape[0]))
    return out
