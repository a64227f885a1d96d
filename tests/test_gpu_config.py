"""HILA configuration files and the gauge download through the C ABI on the GPU
(wl_config_read / wl_config_write / wl_gauge_download; GaugeField::config_read / config_write,
libraries/plumbing/gaugefield.h:194-272).

Byte work, so every check is bit-exact against the numpy restatement (oracle/config_io.py):
the file the library writes is the oracle's file byte for byte, the field it reads is the
oracle's, including the periodic replication of a smaller file lattice; sizes are chosen so
that a direction spans several 16 MiB transfer chunks.  A request served from a file gives the
same solution, to the last bit, as the same request given the links in memory."""

import numpy as np
import pytest

from hila_amd import wilson as W
from oracle import config_io as cio
from oracle.oracle import EVEN, Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_download_returns_the_uploaded_links():
    dims = (8, 4, 4, 6)
    U = W.gen_gauge(dims, "haar", 0.0, 31)
    _say("download", dims)
    with W.Lattice(dims) as lat:
        lat.upload_gauge(U)
        assert lat.download_gauge().tobytes() == U.tobytes()


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (16, 16, 16, 32)])
def test_write_is_the_oracle_file(tmp_path, dims):
    U = W.gen_gauge(dims, "smooth", 0.3, 32)
    path = str(tmp_path / "w.cfg")
    _say("config_write", dims)
    with W.Lattice(dims) as lat:
        lat.upload_gauge(U)
        lat.config_write(path)
    assert open(path, "rb").read() == cio.config_bytes(U, dims)


@pytest.mark.parametrize("dims", [(4, 2, 6, 4), (16, 16, 16, 32)])
def test_read_is_the_oracle_field(tmp_path, dims):
    U = W.gen_gauge(dims, "haar", 0.0, 33)
    path = str(tmp_path / "r.cfg")
    cio.config_write(path, U, dims)
    _say("config_read", dims)
    with W.Lattice(dims) as lat:
        lat.config_read(path)
        assert lat.download_gauge().tobytes() == U.tobytes()


def test_read_replicates_a_smaller_file_lattice(tmp_path):
    small, big = (4, 2, 2, 4), (8, 4, 6, 8)
    Us = W.gen_gauge(small, "smooth", 0.4, 34)
    path = str(tmp_path / "tile.cfg")
    cio.config_write(path, Us, small)
    _say("config_read replicate", small, "->", big)
    with W.Lattice(big) as lat:
        lat.config_read(path)
        assert lat.download_gauge().tobytes() == cio.config_read(path, big).tobytes()


def test_operator_on_a_read_configuration(tmp_path):
    dims = (4, 4, 4, 8)
    U = W.gen_gauge(dims, "smooth", 0.3, 35)
    path = str(tmp_path / "op.cfg")
    cio.config_write(path, U, dims)
    psi = W.gen_spinor(dims, 36, EVEN)
    _say("operator after config_read", dims)
    with W.Lattice(dims) as lat:
        lat.config_read(path)
        fin = W.Field(lat, EVEN).upload(psi)
        fout = W.Field(lat, EVEN)
        W.Dirac_Wilson_evenodd(0.12, lat).apply(fin, fout)
        got = fout.to_full()
    want = Oracle(dims).meo(U, 0.12, psi, 1)
    assert np.linalg.norm(got - want) <= 1e-13 * np.linalg.norm(want)


def test_rejected_file_keeps_the_gauge_field(tmp_path):
    dims = (4, 4, 4, 8)
    U = W.gen_gauge(dims, "haar", 0.0, 37)
    bad = str(tmp_path / "bad.cfg")
    cio.config_write(bad, W.gen_gauge((4, 4, 4, 6), "haar", 0.0, 38), (4, 4, 4, 6))   # 6 ∤ 8
    _say("config_read rejected", dims)
    with W.Lattice(dims) as lat:
        lat.upload_gauge(U)
        with pytest.raises(W.WilsonError) as e:
            lat.config_read(bad)
        assert e.value.code == W.WL_E_ARG and "should be (or divide)" in str(e.value)
        with pytest.raises(W.WilsonError) as e:
            lat.config_read(str(tmp_path / "missing.cfg"))
        assert e.value.code == W.WL_E_ARG
        assert lat.download_gauge().tobytes() == U.tobytes()


def test_served_request_from_a_file(tmp_path):
    dims = (8, 8, 8, 8)
    U = W.gen_gauge(dims, "smooth", 0.3, 39)
    path = str(tmp_path / "serve.cfg")
    cio.config_write(path, U, dims)
    Vh = int(np.prod(dims)) // 2
    src = np.ascontiguousarray(np.stack([W.gen_spinor(dims, 40 + j, EVEN)[:Vh] for j in range(3)]))
    _say("served request from a configuration file", dims)
    with W.Server() as S:
        sol_mem = np.zeros_like(src)
        sol_file = np.zeros_like(src)
        t1 = S.submit(dims, U, src, sol_mem, 0.12, 1e-10, precision="f32", delta=0.1)
        t2 = S.submit(dims, path, src, sol_file, 0.12, 1e-10, precision="f32", delta=0.1)
        r1, _ = S.wait(t1)
        r2, _ = S.wait(t2)
        missing = np.zeros_like(src)
        t3 = S.submit(dims, str(tmp_path / "missing.cfg"), src, missing, 0.12, 1e-10)
        r3, _ = S.wait(t3)
    assert r1["status"] == W.WL_OK and r2["status"] == W.WL_OK
    assert sol_file.tobytes() == sol_mem.tobytes()
    assert r3["status"] == W.WL_E_ARG and "cannot open" in r3["message"]
