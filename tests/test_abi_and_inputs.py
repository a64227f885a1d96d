"""CPU tests of the C ABI library (it loads and exports every declared symbol; argument errors
are reported as codes) and of the seeded input generator.  No GPU compute is called."""
import hashlib
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from hila_amd import wilson as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "wilson_mi355x.h")]


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(wl_[a-z_0-9]+)\s*\(", text,
                            flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 18
    lib = W.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(W.EXPORTED_SYMBOLS) | set(names)
    assert set(W.EXPORTED_SYMBOLS) <= set(names), "binding lists a symbol the header lacks"


def test_python_constants_match_the_header():
    """Every WL_* enumerator of include/wilson_mi355x.h has the same value in the Python
    mirror (status codes, options, precisions, parities)."""
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(W.__file__), "..", "include", "wilson_mi355x.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)            # comments mention values too
    hdr = re.sub(r"//[^\n]*", "", hdr)
    enums = {m.group(1): int(m.group(2)) for m in re.finditer(r"\b(WL_[A-Z0-9_]+)\s*=\s*(-?\d+)", hdr)}
    assert len(enums) >= 20
    for name, value in enums.items():
        py = name[3:] if name.startswith(("WL_OPT_", "WL_F", "WL_EVEN", "WL_ODD")) else name
        if hasattr(W, py):
            assert getattr(W, py) == value, name
        else:
            assert hasattr(W, name) and getattr(W, name) == value, name


def test_version_and_error_reporting():
    assert b"gfx950" in W.lib().wl_version()
    with pytest.raises(W.WilsonError) as e:
        W.Lattice((4, 4, 3, 4))  # odd extent: rejected before any device call
    assert e.value.code == W.WL_E_ARG
    assert b"even" in W.lib().wl_last_error()


def test_generator_argument_errors():
    with pytest.raises(W.WilsonError):
        W.gen_spinor((4, 4, 4, 5), 1)
    with pytest.raises(W.WilsonError):
        W.gen_point((4, 4, 4, 4), 1, 4, 0)


def _su3_checks(U):
    m = W_complex(U)
    eye = np.eye(3)
    unit = np.abs(m @ np.conj(np.transpose(m, (0, 1, 3, 2))) - eye).max()
    det = np.abs(np.linalg.det(m) - 1).max()
    return unit, det


def W_complex(U):
    u = U.reshape(4, -1, 9, 2)
    return (u[..., 0] + 1j * u[..., 1]).reshape(4, -1, 3, 3)


@pytest.mark.parametrize("kind,eps", [("haar", 0.0), ("smooth", 0.3), ("smooth", 0.5), ("su2_hits", 0.0)])
def test_links_are_su3(kind, eps):
    U = W.gen_gauge((8, 8, 8, 8), kind, eps, 1)
    unit, det = _su3_checks(U)
    assert unit < 1e-13 and det < 1e-13


def test_haar_moments():
    """Haar SU(3): E[tr U] = 0, E[|tr U|^2] = 1 (one irreducible representation)."""
    m = W_complex(W.gen_gauge((8, 8, 8, 16), "haar", 0, 7))
    tr = np.trace(m, axis1=-2, axis2=-1).ravel()
    n = tr.size
    assert abs(tr.mean()) < 6 / np.sqrt(n)
    assert abs((np.abs(tr) ** 2).mean() - 1) < 6 * np.sqrt(2.0 / n) * 1.5


def test_su2_hits_links_are_haar_like():
    """HILA's SU<3>::random() (16 hits of random SU(2) subgroup matrices on the three row
    pairs) converges to the Haar measure: its links must show Haar SU(3)'s moments,
    E[tr U] = 0, E[|tr U|^2] = 1, E[|tr U|^4] = 2 (k = 2 < N: the U(N) value k!) and
    E[|U_ij|^2] = 1/3 for every entry.  A generator that mixed one row pair only leaves
    U_22 = 1 and fails the first and last.  The moments pin the limit distribution, not the
    sequence of hits: any SU(2) factors whose subgroups generate SU(3) tend to the Haar measure,
    so the individual draws stay unpinned (HILA's random streams cannot be reproduced anyway)."""
    m = W_complex(W.gen_gauge((8, 8, 8, 16), "su2_hits", 0, 11))
    tr = np.trace(m, axis1=-2, axis2=-1).ravel()
    n = tr.size
    a2 = np.abs(tr) ** 2
    assert abs(tr.mean()) < 6 / np.sqrt(n)
    assert abs(a2.mean() - 1) < 6 * a2.std() / np.sqrt(n)
    assert abs((a2 ** 2).mean() - 2) < 6 * (a2 ** 2).std() / np.sqrt(n)
    e2 = (np.abs(m) ** 2).reshape(-1, 9)
    assert np.all(np.abs(e2.mean(axis=0) - 1 / 3) < 6 * e2.std(axis=0) / np.sqrt(e2.shape[0]))
    # not the same draws as the Ginibre-based Haar links of the same seed
    assert not np.allclose(m, W_complex(W.gen_gauge((8, 8, 8, 16), "haar", 0, 11)))


def test_smooth_links_distance_from_identity():
    """exp(i eps H): Re tr U / 3 = 1 - eps^2 tr(H^2)/6 + O(eps^4); tr H^2 has mean 8 for the
    traceless Hermitian H with unit-variance entries (8 real Gaussian directions)."""
    for eps in (0.05, 0.1):
        m = W_complex(W.gen_gauge((4, 4, 4, 8), "smooth", eps, 3))
        plaq = np.trace(m, axis1=-2, axis2=-1).real.mean() / 3
        assert abs((1 - plaq) / (eps * eps * 8 / 6) - 1) < 0.05


def test_gaussian_spinor_moments_and_parity():
    dims = (8, 8, 8, 16)
    f = W.gen_spinor(dims, 2, W.EVEN)
    Vh = f.shape[0] // 2
    x = f[:Vh].ravel()
    n = x.size
    assert abs(x.mean()) < 6 / np.sqrt(n)
    assert abs(x.var() - 1) < 6 * np.sqrt(2.0 / n)
    assert np.all(f[Vh:] == 0)
    g = W.gen_spinor(dims, 2, 3)
    assert np.array_equal(g[:Vh], f[:Vh])  # value depends on (seed, site, component) only
    assert np.any(g[Vh:] != 0)


def test_point_source():
    f, site = W.gen_point((4, 4, 4, 4), 9, 2, 1)
    assert f[site, 2 * (3 * 2 + 1)] == 1.0 and np.count_nonzero(f) == 1
    f2, site2 = W.gen_point((4, 4, 4, 4), 9, 0, 0)
    assert site2 == site


def test_generator_deterministic_and_thread_independent():
    """Same bytes for the same seed, whatever the OpenMP thread count."""
    code = ("import sys, hashlib; sys.path.insert(0, %r); from hila_amd import wilson as W; "
            "U = W.gen_gauge((8, 8, 4, 6), 'smooth', 0.3, 5); "
            "f = W.gen_spinor((8, 8, 4, 6), 6, 3); "
            "print(hashlib.sha256(U.tobytes() + f.tobytes()).hexdigest())" % ROOT)
    hashes = set()
    for threads in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=threads)
        out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                             text=True, check=True)
        hashes.add(out.stdout.strip())
    U = W.gen_gauge((8, 8, 4, 6), "smooth", 0.3, 5)
    f = W.gen_spinor((8, 8, 4, 6), 6, 3)
    hashes.add(hashlib.sha256(U.tobytes() + f.tobytes()).hexdigest())
    assert len(hashes) == 1
    assert not np.array_equal(W.gen_gauge((4, 4, 4, 4), "smooth", 0.3, 1),
                              W.gen_gauge((4, 4, 4, 4), "smooth", 0.3, 2))
