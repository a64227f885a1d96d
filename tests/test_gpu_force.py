"""Wilson fermion force and the HMC force step on the GPU (wl_fermion_force,
wl_fermion_force_step, wl_momentum_upload/download), against the oracle (oracle_force in
wilson_oracle.c, oracle/hmc.py), which tests/test_oracle_force.py pins as the derivative of
the action.

Floating point: the force to 1e-13 relative (the kernel forms each term from the two upper
spin rows of the projections; the oracle expands all four), the momentum change of a force
step to 1e-12 relative, computed by the oracle from the GPU's own CG solution.  The momentum
round trip is bit-exact."""
import numpy as np
import pytest

from hila_amd import wilson as W
from oracle import hmc
from oracle.oracle import EVEN, Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dims,ap", [((8, 8, 8, 8), True), ((4, 6, 4, 8), False), ((16, 8, 8, 16), True)])
@pytest.mark.parametrize("sign", [1, -1])
def test_force_matches_oracle(dims, ap, sign):
    _say("force", dims, ap, sign)
    U = W.gen_gauge(dims, "smooth", 0.4, 801)
    chi = W.gen_spinor(dims, 802, EVEN)
    psi = W.gen_spinor(dims, 803, EVEN)
    with W.Lattice(dims, antiperiodic_t=ap) as lat:
        lat.upload_gauge(U)
        D = W.Dirac_Wilson_evenodd(0.12, lat)
        F = W.fermion_force(D, W.Field(lat).upload(chi), W.Field(lat).upload(psi), sign)
    want = Oracle(dims, antiperiodic_t=ap).force(U, 0.12, chi, psi, sign)
    assert rel(F, want) < 1e-13


def test_momentum_round_trip():
    dims = (8, 4, 4, 6)
    P = np.random.default_rng(1).standard_normal((4, int(np.prod(dims)), 18))
    with W.Lattice(dims) as lat:
        assert not np.any(lat.download_momentum())      # zero until uploaded
        lat.upload_momentum(P)
        assert lat.download_momentum().tobytes() == P.tobytes()


@pytest.mark.parametrize("mh,precision", [(0.0, "f64"), (0.1, "f64"), (0.0, "f32")])
def test_force_step_matches_oracle(mh, precision):
    dims, kappa, eps = (8, 8, 8, 8), 0.124, 0.05
    _say("force step", mh, precision)
    U = W.gen_gauge(dims, "smooth", 0.3, 810)
    chi = W.gen_spinor(dims, 811, EVEN)
    P0 = hmc.as_real(hmc.project_antihermitean(
        hmc.as_matrices(np.random.default_rng(2).standard_normal((4, int(np.prod(dims)), 18)))))
    with W.Lattice(dims) as lat:
        lat.upload_gauge(U)
        lat.upload_momentum(P0)
        D = W.Dirac_Wilson_evenodd(kappa, lat)
        op = W.Hasenbusch_operator(D, mh) if mh else D
        psi = W.Field(lat)
        st = W.force_step(op, W.Field(lat).upload(chi), psi, eps, accuracy=1e-10, precision=precision)
        P1 = lat.download_momentum()
        psi_h = psi.to_full()
    assert st["converged"] == 1
    orc = Oracle(dims)
    r = chi - orc.meo(U, kappa, orc.meo(U, kappa, psi_h, 1, mh=mh), -1, mh=mh)
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(chi)
    want = hmc.force_step(orc, U, kappa, mh, chi, psi_h, P0, eps)
    assert rel(P1 - P0, want - P0) < 1e-12
    # the update is in the algebra: anti-Hermitian and traceless
    dP = hmc.as_matrices(P1 - P0)
    assert np.abs(dP + np.conj(np.swapaxes(dP, -1, -2))).max() < 1e-13 * np.abs(dP).max()
    assert np.abs(np.trace(dP, axis1=-2, axis2=-1)).max() < 1e-13 * np.abs(dP).max()


def test_hasenbusch2_force_step_matches_oracle():
    dims, kappa, mh, eps = (8, 8, 8, 8), 0.124, 0.2, 0.05
    _say("Hasenbusch_action_2 force step")
    U = W.gen_gauge(dims, "smooth", 0.3, 820)
    chi = W.gen_spinor(dims, 821, EVEN)
    with W.Lattice(dims) as lat:
        lat.upload_gauge(U)
        H = W.Hasenbusch_operator(W.Dirac_Wilson_evenodd(kappa, lat), mh)
        psi = W.Field(lat)
        st = W.hasenbusch2_force_step(H, W.Field(lat).upload(chi), psi, eps, accuracy=1e-10)
        P1 = lat.download_momentum()
        psi_h = psi.to_full()
    assert st["converged"] == 1
    orc = Oracle(dims)
    b = orc.meo(U, kappa, chi, -1, mh=mh)              # D_h^dag chi
    r = b - orc.meo(U, kappa, orc.meo(U, kappa, psi_h, 1), -1)
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(b)
    want = hmc.hasenbusch2_force_step(orc, U, kappa, mh, chi, psi_h, np.zeros_like(P1), eps)
    assert rel(P1, want) < 1e-12
