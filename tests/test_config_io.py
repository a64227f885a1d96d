"""CPU tests of HILA's gauge-configuration file format: the numpy restatement in
oracle/config_io.py (GaugeField::config_write / config_read, libraries/plumbing/gaugefield.h:
194-272; Field<T>::write / read, field_io.h:125-160, 221-312) and the library's host-side
header checks (wl_config_header; no GPU call).

HILA ships no configuration file to compare against, so the restatement is pinned by what the
format fixes on its own:
  * every link lands at the byte offset 56 + 144 (mu V + x + Lx (y + Ly (z + Lz t))) given by
    the header layout (gaugefield.h:200-212), the direction-major payload (gaugefield.h:168-
    172) and the lexicographic site order with x fastest (lattice.h:389-396); the links here
    carry their own (direction, coordinates, entry) as values, with coordinates taken from the
    dense oracle's independent site map;
  * reading inverts writing bit for bit;
  * a smaller file lattice is replicated: U(x) = U_file(x mod L_file), so every plaquette of
    the replicated field equals the file lattice's (the plaquette is pinned by unit links and
    gauge invariance);
  * each header check of config_read fails the file it is meant to fail.
"""
import ctypes
import os

import numpy as np
import pytest

from hila_amd import wilson as W
from oracle import config_io as cio
from oracle import dense


def tagged_links(dims):
    lat = dense.Lattice(dims)
    Lx, Ly, Lz, _ = lat.L
    c = lat.coords
    lex = c[:, 0] + Lx * (c[:, 1] + Ly * (c[:, 2] + Lz * c[:, 3]))
    U = np.empty((4, lat.V, 18))
    for mu in range(4):
        U[mu] = mu * 1e7 + lex[:, None] * 100.0 + np.arange(18)[None, :]
    return U


def test_byte_offsets_of_every_link(tmp_path):
    dims = (4, 2, 6, 4)
    U = tagged_links(dims)
    path = tmp_path / "tagged.cfg"
    cio.config_write(path, U, dims)
    raw = path.read_bytes()
    V = int(np.prod(dims))
    assert len(raw) == 56 + 4 * V * 144
    head = np.frombuffer(raw[:56], dtype=np.int64)
    assert head.tolist() == [394824242, 4, 144, 4, 2, 6, 4]
    Lx, Ly, Lz, Lt = dims
    for mu in range(4):
        for t in range(Lt):
            for z in range(Lz):
                for y in range(Ly):
                    for x in range(Lx):
                        lex = x + Lx * (y + Ly * (z + Lz * t))
                        off = 56 + 144 * (mu * V + lex)
                        v = np.frombuffer(raw[off: off + 144], dtype=np.float64)
                        assert np.array_equal(v, mu * 1e7 + lex * 100.0 + np.arange(18))


def test_round_trip_is_bit_exact(tmp_path):
    dims = (4, 4, 2, 6)
    U = W.gen_gauge(dims, "haar", 0.0, 7)
    path = tmp_path / "haar.cfg"
    cio.config_write(path, U, dims)
    assert cio.config_header(path) == dims
    back = cio.config_read(path, dims)
    assert back.tobytes() == U.tobytes()


def test_plaquette_pins():
    dims = (4, 4, 4, 4)
    unit = W.gen_gauge(dims, "unit", 0.0, 1)
    for v in cio.plaquette(unit, dims).values():
        assert v == pytest.approx(1.0, abs=1e-15)
    U = W.gen_gauge(dims, "smooth", 0.4, 11)
    g = W.gen_gauge(dims, "haar", 0.0, 12)[0]   # one SU(3) matrix per site
    lat = dense.Lattice(dims)
    Uc = U[..., 0::2] + 1j * U[..., 1::2]
    gc = (g[..., 0::2] + 1j * g[..., 1::2]).reshape(-1, 3, 3)
    Ug = np.empty_like(Uc)
    for mu in range(4):
        fwd, _ = lat.hop(mu, +1)
        m = gc @ Uc[mu].reshape(-1, 3, 3) @ np.conj(np.transpose(gc[fwd], (0, 2, 1)))
        Ug[mu] = m.reshape(-1, 9)
    Ugr = np.empty_like(U)
    Ugr[..., 0::2], Ugr[..., 1::2] = Ug.real, Ug.imag
    p0, p1 = cio.plaquette(U, dims), cio.plaquette(Ugr, dims)
    for k in p0:
        assert p1[k] == pytest.approx(p0[k], abs=1e-13)
        assert abs(p0[k]) < 1.0 - 1e-3   # not trivially 1


def test_replication_of_a_smaller_file_lattice(tmp_path):
    small, big = (2, 2, 2, 4), (4, 2, 6, 8)
    Us = W.gen_gauge(small, "smooth", 0.5, 21)
    path = tmp_path / "tile.cfg"
    cio.config_write(path, Us, small)
    Ub = cio.config_read(path, big)
    lat_b, lat_s = dense.Lattice(big), dense.Lattice(small)
    src = lat_s.site_of(lat_b.coords % np.array(small))
    assert np.array_equal(Ub, Us[:, src])
    ps, pb = cio.plaquette(Us, small), cio.plaquette(Ub, big)
    for k in ps:
        assert pb[k] == pytest.approx(ps[k], abs=1e-14)


def _write_raw(path, header, payload_links):
    with open(path, "wb") as f:
        f.write(np.array(header, dtype=np.int64).tobytes())
        f.write(np.zeros(payload_links * 18).tobytes())


BAD_HEADERS = [
    ("wrong id", [394824243, 4, 144, 2, 2, 2, 2], 64),
    ("wrong dimensionality", [394824242, 3, 144, 2, 2, 2, 2], 64),
    ("wrong size of field element", [394824242, 4, 72, 2, 2, 2, 2], 64),
    ("incorrect lattice dimension", [394824242, 4, 144, 2, 0, 2, 2], 64),
    ("needs", [394824242, 4, 144, 2, 2, 2, 2], 63),    # one link short
]


@pytest.mark.parametrize("what,header,links", BAD_HEADERS)
def test_header_checks(tmp_path, what, header, links):
    path = tmp_path / "bad.cfg"
    _write_raw(path, header, links)
    dims = (ctypes.c_int * 4)()
    rc = W.lib().wl_config_header(os.fsencode(str(path)), dims)
    assert rc == W.WL_E_ARG
    msg = W.lib().wl_last_error().decode()
    assert "CONFIG ERROR in file" in msg and what in msg, msg
    if what != "needs":
        with pytest.raises(cio.ConfigError, match=what):
            cio.config_header(path)
    else:
        with pytest.raises(cio.ConfigError):
            cio.config_read(path, (2, 2, 2, 2))


def test_header_of_a_good_file_and_a_dimension_that_does_not_divide(tmp_path):
    dims = (2, 4, 2, 6)
    path = tmp_path / "good.cfg"
    cio.config_write(path, W.gen_gauge(dims, "haar", 0.0, 3), dims)
    out = (ctypes.c_int * 4)()
    assert W.lib().wl_config_header(os.fsencode(str(path)), out) == W.WL_OK
    assert tuple(out) == dims
    with pytest.raises(cio.ConfigError, match="should be \\(or divide\\)"):
        cio.config_read(path, (4, 4, 4, 8))   # 6 does not divide 8
    assert W.lib().wl_config_header(os.fsencode(str(tmp_path / "missing.cfg")), out) == W.WL_E_ARG
