"""Writes the golden fixtures of the Wilson path under tests/golden/.

    python tests/golden/make_fixtures.py

The expected outputs come from the independent dense construction (oracle/dense.py: explicit
gamma tables of libraries/datatypes/wilson_vector.h:311-347 and sparse-matrix assembly), not
from the C restatement or the HIP path, which the tests then check against these files.
HILA itself ships no golden vectors for this path (its applications/test_cases/test_CG.cpp
holds invariants only, at unit links), and it cannot be built here; the inputs are therefore
the project's seeded fields, stored in the fixture next to the outputs.

Each .npz holds: dims, kappa, antiperiodic, U (4,V,18), psi (V,24) even source,
psi_odd (V,24) odd source, M_plus / M_minus (V/2,24) = M_s psi on even sites,
H_oe_plus / H_oe_minus (V/2,24) = -kappa H_oe psi (Dirac_Wilson_hop_set(..., ODD, s)),
H_eo_plus (V/2,24) = -kappa H_eo psi_odd, and x_cg (V/2,24) = (M^dag M)^{-1} psi_e from a
sparse direct solve.
"""
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hila_amd import wilson as W  # noqa: E402  (seeded input generator)
from oracle import dense  # noqa: E402

CASES = [
    # name, dims, kind, eps, gauge seed, source seed, kappa, antiperiodic
    ("wilson_2x2x2x4_haar", (2, 2, 2, 4), "haar", 0.0, 101, 102, 0.12, True),
    ("wilson_4x4x4x4_smooth", (4, 4, 4, 4), "smooth", 0.3, 103, 104, 0.125, True),
    ("wilson_4x2x6x4_periodic", (4, 2, 6, 4), "haar", 0.0, 105, 106, 0.11, False),
]


def real(fc):
    return dense.field_real(fc)


def make(name, dims, kind, eps, gseed, sseed, kappa, ap):
    lat = dense.Lattice(dims, antiperiodic_t=ap)
    U = W.gen_gauge(dims, kind, eps, gseed)
    psi = W.gen_spinor(dims, sseed, W.EVEN)
    psi_odd = W.gen_spinor(dims, sseed + 1000, W.ODD)
    Wd = dense.WilsonDense(lat, U, kappa)
    pe = dense.field_complex(psi)[: lat.Vh]
    po = dense.field_complex(psi_odd)[lat.Vh:]
    out = dict(dims=np.array(dims), kappa=np.array(kappa), antiperiodic=np.array(int(ap)),
               U=U, psi=psi, psi_odd=psi_odd)
    for s, tag in ((1, "plus"), (-1, "minus")):
        out[f"M_{tag}"] = real(Wd.apply_M(pe, s))
        out[f"H_oe_{tag}"] = real(Wd.hop_set_oe(pe, s))
    out["H_eo_plus"] = real(Wd.hop_set_eo(po, 1))
    MdM = (Wd.M(-1) @ Wd.M(1)).tocsc()
    x = spla.spsolve(MdM, pe.reshape(-1))
    out["x_cg"] = real(x.reshape(-1, 4, 3))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return path


def main():
    for c in CASES:
        print(make(*c))


if __name__ == "__main__":
    main()
