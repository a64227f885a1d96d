"""HILA's HMC fermion force step, restated in numpy on top of the C oracle.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of wl_fermion_force_step.  The product
never imports it.

  force_step (libraries/hmc/fermion_field.h:146-167):
      psi = (D^dag D)^-1 chi;  Mpsi = D psi;
      force = -eps (D.force(Mpsi, psi, +1) + D.force(psi, Mpsi, -1));
      gauge.add_momentum(force)
  add_momentum (libraries/hmc/gauge_field.h:249-257):
      force = U force;  project_antihermitean(force);  momentum += force

project_antihermitean is called there but not defined anywhere in the reference snapshot; it is
taken as the traceless anti-Hermitian part, (M - M^dag)/2 - tr((M - M^dag)/2)/N, the projection
onto the su(N) algebra in which HILA's momenta live (gaussian_algebra, gauge_field.h:223-230).
Its choice is pinned by the step being the gradient of the action (tests/test_oracle_force.py).
Links and momenta are (4, V, 18) float64 arrays in HILA's layout.
"""
from __future__ import annotations

import numpy as np


def as_matrices(A):
    A = np.asarray(A)
    return (A[..., 0::2] + 1j * A[..., 1::2]).reshape(A.shape[:-1] + (3, 3))


def as_real(M):
    M = np.asarray(M)
    out = np.empty(M.shape[:-2] + (18,))
    flat = M.reshape(M.shape[:-2] + (9,))
    out[..., 0::2], out[..., 1::2] = flat.real, flat.imag
    return out


def project_antihermitean(M):
    A = 0.5 * (M - np.conj(np.swapaxes(M, -1, -2)))
    tr = np.trace(A, axis1=-2, axis2=-1) / 3.0
    return A - tr[..., None, None] * np.eye(3)


def add_momentum(U, force, mom):
    """gauge_field::add_momentum: mom + P(U force), all (4, V, 18)."""
    return as_real(as_matrices(mom) + project_antihermitean(as_matrices(U) @ as_matrices(force)))


def force_step(orc, U, kappa, mh, chi, psi, mom, eps):
    """The momentum after fermion_action::force_step, for psi already solved."""
    Mpsi = orc.meo(U, kappa, psi, 1, mh=mh)
    F1 = orc.force(U, kappa, Mpsi, psi, 1)
    F2 = orc.force(U, kappa, psi, Mpsi, -1)
    return add_momentum(U, -eps * (F1 + F2), mom)


def hasenbusch2_force_step(orc, U, kappa, mh, chi, psi, mom, eps):
    """The momentum after Hasenbusch_action_2::force_step (fermion_field.h:319-346), for psi
    already solved from (D^dag D) psi = D_h^dag chi: Mpsi = D psi - chi on D.par, then the
    same force and add_momentum as force_step."""
    Mpsi = orc.meo(U, kappa, psi, 1)
    Mpsi[: orc.Vh] -= chi[: orc.Vh]
    F1 = orc.force(U, kappa, Mpsi, psi, 1)
    F2 = orc.force(U, kappa, psi, Mpsi, -1)
    return add_momentum(U, -eps * (F1 + F2), mom)
