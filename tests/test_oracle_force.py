"""Pins the oracle's Wilson fermion force (oracle_calc_force / oracle_force in wilson_oracle.c:
Dirac_Wilson_calc_force and Dirac_Wilson_evenodd::force, wilson.h:108-137, 294-311) and its
HMC force step (oracle/hmc.py: fermion_action::force_step, add_momentum) before they check the
HIP path.  CPU only.

HILA holds no fixture for the force; what fixes it is that it is a derivative:
  * d/de Re <chi, M_s(U + e E) psi> = Re tr(E F_mu(x)) for a perturbation E of one link, by
    central differences, for both signs, links of both parities and links across the
    antiperiodic time edge.  As E ranges over all complex 3x3 matrices this determines F
    completely, so a transposed, conjugated, mis-signed or mis-projected term fails it;
  * gauge covariance F_mu(x) -> g(x+mu) F_mu(x) g(x)^dag;
  * D.force(a, b, s) = D.force(b, a, -s) (the two calls of force_step form the same products);
  * the force step's momentum change is the gradient of the action S = <chi, (D^dag D)^-1 chi>:
    for U_mu(x) -> exp(e T) U_mu(x), T traceless anti-Hermitian, dS/de = Re tr(T dP) / eps,
    which also fixes project_antihermitean (undefined in the reference snapshot); the same for
    the second Hasenbusch term, S2 = <chi, D_h (D^dag D)^-1 D_h^dag chi>, D_h = D + mh.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from hila_amd import wilson as W
from oracle import hmc
from oracle.oracle import EVEN, Oracle

KAPPA = 0.12


@pytest.fixture(scope="module")
def small():
    dims = (4, 4, 4, 6)
    return dims, W.gen_gauge(dims, "haar", 0, 701), Oracle(dims)


def _perturb(U, mu, x, E, e):
    U2 = U.copy()
    m = hmc.as_matrices(U2[mu, x]) + e * E
    U2[mu, x] = hmc.as_real(m)
    return U2


def _links_to_check(orc):
    """A link of each parity in each direction, and time links across the t edge."""
    lat = orc
    out = []
    rng = np.random.default_rng(5)
    for mu in range(4):
        out.append((mu, int(rng.integers(lat.Vh))))            # even site
        out.append((mu, lat.Vh + int(rng.integers(lat.Vh))))   # odd site
    V = lat.V
    # sites with t = Lt - 1 (the last time slice) in lexicographic order are the last slice
    # of each parity block; take one of each parity
    out.append((3, lat.Vh - 1))
    out.append((3, V - 1))
    return out


@pytest.mark.parametrize("sign", [1, -1])
def test_force_is_the_derivative(small, sign):
    dims, U, orc = small
    chi = W.gen_spinor(dims, 702, EVEN)
    psi = W.gen_spinor(dims, 703, EVEN)
    F = orc.force(U, KAPPA, chi, psi, sign)
    cv = lambda f: f[:, 0::2] + 1j * f[:, 1::2]
    S = lambda UU: np.vdot(cv(chi), cv(orc.meo(UU, KAPPA, psi, sign))).real
    rng = np.random.default_rng(7)
    eps = 1e-5
    for mu, x in _links_to_check(orc):
        E = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        dS = (S(_perturb(U, mu, x, E, eps)) - S(_perturb(U, mu, x, E, -eps))) / (2 * eps)
        want = np.trace(E @ hmc.as_matrices(F[mu, x])).real
        assert abs(dS - want) < 1e-7 * max(1.0, abs(want)), (mu, x, dS, want)


def test_force_symmetry_and_gauge_covariance(small):
    dims, U, orc = small
    a = W.gen_spinor(dims, 710, EVEN)
    b = W.gen_spinor(dims, 711, EVEN)
    Fab = orc.force(U, KAPPA, a, b, 1)
    assert np.array_equal(Fab, orc.force(U, KAPPA, b, a, -1))
    # gauge transformation g(x): U_mu(x) -> g(x) U_mu(x) g(x+mu)^dag, spinors -> g spinors
    from oracle import dense
    lat = dense.Lattice(dims)
    g = hmc.as_matrices(W.gen_gauge(dims, "haar", 0, 712)[0])
    Um = hmc.as_matrices(U)
    Ug = np.empty_like(Um)
    fwd = [lat.hop(mu, 1)[0] for mu in range(4)]
    for mu in range(4):
        Ug[mu] = g @ Um[mu] @ np.conj(np.swapaxes(g[fwd[mu]], -1, -2))
    rot = lambda f: np.einsum("vij,vsj->vsi", g, (f[:, 0::2] + 1j * f[:, 1::2]).reshape(-1, 4, 3))
    def real(fc):
        fc = fc.reshape(-1, 12)
        out = np.empty((fc.shape[0], 24))
        out[:, 0::2], out[:, 1::2] = fc.real, fc.imag
        return out
    Fg = orc.force(hmc.as_real(Ug), KAPPA, real(rot(a)), real(rot(b)), 1)
    for mu in range(4):
        want = g[fwd[mu]] @ hmc.as_matrices(Fab[mu]) @ np.conj(np.swapaxes(g, -1, -2))
        assert np.abs(hmc.as_matrices(Fg[mu]) - want).max() < 1e-12 * np.abs(want).max()


def test_projection():
    rng = np.random.default_rng(3)
    M = rng.normal(size=(5, 3, 3)) + 1j * rng.normal(size=(5, 3, 3))
    P = hmc.project_antihermitean(M)
    assert np.allclose(P, -np.conj(np.swapaxes(P, -1, -2)), atol=1e-15)
    assert np.allclose(np.trace(P, axis1=1, axis2=2), 0, atol=1e-15)
    assert np.allclose(hmc.project_antihermitean(P), P, atol=1e-15)


@pytest.mark.parametrize("mh", [0.0, 0.1])
def test_force_step_is_the_gradient_of_the_action(mh):
    dims = (4, 4, 4, 4)
    U = W.gen_gauge(dims, "smooth", 0.5, 720)
    orc = Oracle(dims)
    chi = W.gen_spinor(dims, 721, EVEN)

    def action(UU):
        x, _, _ = orc.cg(UU, KAPPA, chi, accuracy=1e-14, mh=mh)
        return float(np.sum(chi[: orc.Vh] * x[: orc.Vh]))   # Re <chi, (D^dag D)^-1 chi>

    psi, _, _ = orc.cg(U, KAPPA, chi, accuracy=1e-14, mh=mh)
    eps = 0.25
    dP = hmc.as_matrices(hmc.force_step(orc, U, KAPPA, mh, chi, psi, np.zeros_like(U), eps))
    rng = np.random.default_rng(9)
    h = 1e-5
    for mu, x in [(0, 3), (1, orc.Vh + 5), (3, orc.Vh - 1), (2, 17)]:
        A = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        T = hmc.project_antihermitean(A)
        Um = hmc.as_matrices(U)

        def moved(e):
            U2 = U.copy()
            U2[mu, x] = hmc.as_real(sla.expm(e * T) @ Um[mu, x])
            return U2

        dS = (action(moved(h)) - action(moved(-h))) / (2 * h)
        want = np.trace(T @ dP[mu, x]).real / eps
        assert abs(dS - want) < 1e-6 * max(1.0, abs(want)), (mu, x, dS, want)


def test_hasenbusch2_force_step_is_the_gradient_of_its_action():
    dims, mh = (4, 4, 4, 4), 0.2
    U = W.gen_gauge(dims, "smooth", 0.5, 730)
    orc = Oracle(dims)
    chi = W.gen_spinor(dims, 731, EVEN)

    def solve(UU):
        Dhchi = orc.meo(UU, KAPPA, chi, -1, mh=mh)
        x, _, _ = orc.cg(UU, KAPPA, Dhchi, accuracy=1e-14)
        return x

    def action(UU):   # Hasenbusch_action_2::action (:255-272): |D (D^dag D)^-1 D_h^dag chi|^2
        v = orc.meo(UU, KAPPA, solve(UU), 1)
        return float(np.sum(v[: orc.Vh] ** 2))

    eps = 0.5
    dP = hmc.as_matrices(hmc.hasenbusch2_force_step(orc, U, KAPPA, mh, chi, solve(U), np.zeros_like(U), eps))
    rng = np.random.default_rng(11)
    h = 1e-5
    Um = hmc.as_matrices(U)
    for mu, x in [(0, 7), (2, orc.Vh + 9), (3, orc.V - 1)]:
        T = hmc.project_antihermitean(rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3)))

        def moved(e):
            U2 = U.copy()
            U2[mu, x] = hmc.as_real(sla.expm(e * T) @ Um[mu, x])
            return U2

        dS = (action(moved(h)) - action(moved(-h))) / (2 * h)
        want = np.trace(T @ dP[mu, x]).real / eps
        assert abs(dS - want) < 1e-6 * max(1.0, abs(want)), (mu, x, dS, want)
