"""Mutation check of the oracle's pins: each planted mistake (compiled into a separate build of
oracle/wilson_oracle.c with -DORACLE_PLANT=n, CPU only) must be rejected by the pins that
tests/test_oracle.py applies to the real oracle.

  1  HILA's backward hop as written (wilson.h:47 reads vtemp[dir][X+dir])
  2  antiperiodic time sign dropped
  3  sign error in the e_z spin projection
  4  U instead of U^dag on the backward hop
  5  gamma_t lower spin rows swapped in expand
"""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle import dense
from oracle.oracle import EVEN, ODD, Oracle

DIMS = (4, 4, 4, 6)


def _pins(variant):
    """Returns the relative errors of the three pins on the variant's operator."""
    lat = dense.Lattice(DIMS, antiperiodic_t=True)
    U = W.gen_gauge(DIMS, "smooth", 0.4, 51)
    Wd = dense.WilsonDense(lat, U, 0.12)
    orc = Oracle(DIMS, antiperiodic_t=True, variant=variant)
    pe = W.gen_spinor(DIMS, 52, EVEN)
    Vh = lat.Vh
    c = dense.field_complex
    # pin 1: against the dense construction (M and the hopping term, both signs)
    e_dense = 0.0
    for s in (1, -1):
        got = orc.meo(U, 0.12, pe, s)[:Vh]
        ref = dense.field_real(Wd.apply_M(c(pe)[:Vh], s))
        e_dense = max(e_dense, np.linalg.norm(got - ref) / np.linalg.norm(ref))
    # pin 2: adjointness
    a = W.gen_spinor(DIMS, 53, EVEN)
    lhs = np.vdot(c(a), c(orc.meo(U, 0.12, pe, 1)))
    rhs = np.vdot(c(orc.meo(U, 0.12, a, -1)), c(pe))
    e_adj = abs(lhs - rhs) / abs(lhs)
    # pin 3: free field, antiperiodic plane wave
    U1 = W.gen_gauge(DIMS, "unit")
    p = np.array([2 * np.pi / 4, 0, 2 * np.pi / 4, 3 * np.pi / 6])
    u = np.arange(12).reshape(4, 3) * (1 + 0.3j) + 1
    psi = dense.plane_wave(lat, p, u)
    psi[Vh:] = 0
    out = c(orc.meo(U1, 0.12, dense.field_real(psi.real + 0j), 1)) + \
        1j * c(orc.meo(U1, 0.12, dense.field_real(psi.imag + 0j), 1))
    expect = psi - 0.12 ** 2 * np.einsum("ab,vbc->vac", dense.free_field_h2(lat, p), psi)
    e_free = np.linalg.norm(out[:Vh] - expect[:Vh]) / np.linalg.norm(expect[:Vh])
    return e_dense, e_adj, e_free


def test_unmutated_oracle_passes_all_pins():
    e = _pins(0)
    assert max(e) < 1e-13, e


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_planted_mistake_is_caught(variant):
    e_dense, e_adj, e_free = _pins(variant)
    assert e_dense > 1e-6, "dense construction pin missed the planted mistake"
    if variant in (1, 4):  # wrong backward hop: breaks D^dag = adjoint(D)
        assert e_adj > 1e-6
    if variant in (1, 2, 3, 5):  # visible with unit links
        assert e_free > 1e-6
