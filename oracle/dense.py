"""Independent dense/sparse construction of HILA's Wilson operator (numpy/scipy).

TEST INFRASTRUCTURE ONLY: used by tests/ and by tests/golden/make_fixtures.py to pin the C
restatement (wilson_oracle.c) and, through committed fixtures, the HIP path.  Nothing in the
product imports it.

It shares no code with wilson_oracle.c: instead of HILA's half-vector projection/expansion it
builds the 4x4 gamma matrices from the tables of libraries/datatypes/wilson_vector.h:311-347
(direction -> gamma map {g1,g2,g3,g0}, :23) and assembles

    D_s = 1 - kappa * H_s,
    H_s psi(x) = sum_mu [ U_mu(x) (1 + s g_mu) psi(x+mu) + U_mu(x-mu)^dag (1 - s g_mu) psi(x-mu) ]

as an explicit sparse matrix over (site, spin, colour), with the site order of
libraries/plumbing/lattice.cpp:171-186 (EVEN_SITES_FIRST, x fastest) and the antiperiodic
time boundary of HILA's fermion fields (field_comm.h:115-124).  M_s = 1 - kappa^2 H_eo H_oe
(wilson.h:272-289).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

# gamma matrices, wilson_vector.h:311-347, acting as (g w).c[i] = sum_j G[i, j] w.c[j]
_I = 1j
GAMMA0 = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)
GAMMA1 = np.array([[0, 0, 0, _I], [0, 0, _I, 0], [0, -_I, 0, 0], [-_I, 0, 0, 0]], dtype=complex)
GAMMA2 = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex)
GAMMA3 = np.array([[0, 0, _I, 0], [0, 0, 0, -_I], [-_I, 0, 0, 0], [0, _I, 0, 0]], dtype=complex)
GAMMA5 = np.diag([1, 1, -1, -1]).astype(complex)
# direction e_x, e_y, e_z, e_t -> gamma1, gamma2, gamma3, gamma0 (wilson_vector.h:23)
GAMMA_DIR = [GAMMA1, GAMMA2, GAMMA3, GAMMA0]


class Lattice:
    """Geometry: coordinates and neighbours in HILA's EVEN_SITES_FIRST order."""

    def __init__(self, dims, antiperiodic_t=True):
        self.L = tuple(int(d) for d in dims)
        assert len(self.L) == 4 and all(d >= 2 and d % 2 == 0 for d in self.L)
        self.V = int(np.prod(self.L))
        self.Vh = self.V // 2
        self.antiperiodic_t = bool(antiperiodic_t)
        lex = np.arange(self.V)
        Lx, Ly, Lz, _ = self.L
        x = lex % Lx
        y = (lex // Lx) % Ly
        z = (lex // (Lx * Ly)) % Lz
        t = lex // (Lx * Ly * Lz)
        coords_lex = np.stack([x, y, z, t], axis=1)
        par = coords_lex.sum(1) % 2
        site = np.where(par == 0, lex // 2, lex // 2 + self.Vh)
        self.coords = np.empty_like(coords_lex)
        self.coords[site] = coords_lex
        self.lex_to_site = site

    def site_of(self, c):
        c = np.asarray(c) % np.array(self.L)
        Lx, Ly, Lz, _ = self.L
        lex = c[..., 0] + Lx * (c[..., 1] + Ly * (c[..., 2] + Lz * c[..., 3]))
        return self.lex_to_site[lex]

    def hop(self, mu, step):
        """(neighbour site, sign) for every site; sign -1 where the hop crosses the
        antiperiodic t edge."""
        c = self.coords.copy()
        c[:, mu] += step
        cross = (c[:, mu] < 0) | (c[:, mu] >= self.L[mu])
        sign = np.ones(self.V)
        if mu == 3 and self.antiperiodic_t:
            sign[cross] = -1.0
        return self.site_of(c), sign


def links_complex(U):
    """HILA AoS links (4, V, 18) float64 -> (4, V, 3, 3) complex."""
    U = np.asarray(U, dtype=np.float64).reshape(4, -1, 9, 2)
    return (U[..., 0] + 1j * U[..., 1]).reshape(4, -1, 3, 3)


def field_complex(f):
    """HILA AoS Wilson vectors (V, 24) -> (V, 4, 3) complex."""
    f = np.asarray(f, dtype=np.float64).reshape(-1, 12, 2)
    return (f[..., 0] + 1j * f[..., 1]).reshape(-1, 4, 3)


def field_real(fc):
    fc = np.asarray(fc).reshape(-1, 12)
    out = np.empty(fc.shape + (2,))
    out[..., 0] = fc.real
    out[..., 1] = fc.imag
    return out.reshape(-1, 24)


def hopping_matrix(lat: Lattice, U, sign):
    """Sparse H_s over (site, spin, colour), index = 12*site + 3*spin + colour."""
    Uc = links_complex(U)
    rows, cols, vals = [], [], []
    eye4 = np.eye(4, dtype=complex)
    base = np.arange(lat.V)
    for mu in range(4):
        fwd, sf = lat.hop(mu, +1)
        bwd, sb = lat.hop(mu, -1)
        Pf = eye4 + sign * GAMMA_DIR[mu]
        Pb = eye4 - sign * GAMMA_DIR[mu]
        # forward: U_mu(x) (1 + s g) psi(x + mu)
        blk_f = np.einsum("ab,vij->vaibj", Pf, Uc[mu]) * sf[:, None, None, None, None]
        # backward: U_mu(x - mu)^dag (1 - s g) psi(x - mu)
        Ud = np.conj(np.transpose(Uc[mu][bwd], (0, 2, 1)))
        blk_b = np.einsum("ab,vij->vaibj", Pb, Ud) * sb[:, None, None, None, None]
        for blk, nb in ((blk_f, fwd), (blk_b, bwd)):
            r = (12 * base[:, None, None, None, None] + 3 * np.arange(4)[None, :, None, None, None]
                 + np.arange(3)[None, None, :, None, None])
            c = (12 * nb[:, None, None, None, None] + 3 * np.arange(4)[None, None, None, :, None]
                 + np.arange(3)[None, None, None, None, :])
            r, c = np.broadcast_arrays(r, c)
            rows.append(r.ravel())
            cols.append(c.ravel())
            vals.append(blk.ravel())
    n = 12 * lat.V
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    H.sum_duplicates()
    return H


class WilsonDense:
    """D_s, its even/odd blocks and M_s for one lattice and gauge field."""

    def __init__(self, lat: Lattice, U, kappa):
        self.lat = lat
        self.kappa = float(kappa)
        self.H = {s: hopping_matrix(lat, U, s) for s in (+1, -1)}
        ne = 12 * lat.Vh
        self.ne = ne

    def D(self, sign):
        n = 12 * self.lat.V
        return sp.identity(n, dtype=complex, format="csr") - self.kappa * self.H[sign]

    def blocks(self, sign):
        H = self.H[sign]
        ne = self.ne
        return H[:ne, ne:], H[ne:, :ne]  # H_eo (even rows, odd cols), H_oe

    def M(self, sign):
        Heo, Hoe = self.blocks(sign)
        I = sp.identity(self.ne, dtype=complex, format="csr")
        return I - self.kappa ** 2 * (Heo @ Hoe)

    def apply_M(self, fe, sign):
        """fe: complex (Vh,4,3) even-site field -> M_s fe."""
        v = np.asarray(fe).reshape(-1)
        return (self.M(sign) @ v).reshape(-1, 4, 3)

    def hop_set_oe(self, fe, sign):
        """Dirac_Wilson_hop_set(..., ODD, sign): out_o = -kappa H_oe in_e."""
        _, Hoe = self.blocks(sign)
        return (-self.kappa * (Hoe @ np.asarray(fe).reshape(-1))).reshape(-1, 4, 3)

    def hop_set_eo(self, fo, sign):
        Heo, _ = self.blocks(sign)
        return (-self.kappa * (Heo @ np.asarray(fo).reshape(-1))).reshape(-1, 4, 3)


def free_field_h2(lat: Lattice, p):
    """Closed form for unit links: on a plane wave e^{i p.x} u, H_+ acts as
    h(p) = 2 sum cos p_mu + 2i sum g_mu sin p_mu, and H_+^2 = a^2 - |b|^2 + 2i a (b.g) with
    a = 2 sum cos p_mu, b_mu = 2 sin p_mu (the g_mu anticommute and square to 1)."""
    a = 2 * np.sum(np.cos(p))
    b = 2 * np.sin(p)
    bg = sum(b[m] * GAMMA_DIR[m] for m in range(4))
    return (a * a - b @ b) * np.eye(4) + 2j * a * bg


def plane_wave(lat: Lattice, p, u):
    """e^{i p.x} u on every site, (V,4,3) complex; u is a (4,3) spin-colour vector."""
    ph = np.exp(1j * (lat.coords @ np.asarray(p, dtype=float)))
    return ph[:, None, None] * np.asarray(u)[None, :, :]
