"""HILA's MRE_guess (libraries/hmc/MRE_guess.h:8-87), restated in numpy.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the library's wl_mre_guess.  The
product never imports it.

The initial guess psi for CG on A psi = chi, A = D^dag D, as the minimal-residual combination
of previous solutions, following the source loop by loop on even-site fields (real arrays,
Re <a, b> = a . b as Vector::rdot):
  1. Gram-Schmidt over the basis (:17-33), each vector then normalised;
  2. the projected matrix M[j][i] = Re <b_j, A b_i> and vector v[i] = Re <b_i, chi> (:35-51);
  3. Gauss-Jordan elimination with HILA's pivot test M[i][i]^2 > 1e-32 (:53-79);
  4. psi = sum_i v[i] b_i (:81-85).

Where the source text differs from what it evidently intends, this follows the intent:
  * the projection coefficient (:24-28) is Re <b_j, b_i> / Re <b_j, b_j>, the textbook
    Gram-Schmidt step; the source divides by Re <b_i, b_i>, which leaves the basis
    non-orthogonal but spans the same space, so the result differs only where a vector is
    dropped;
  * a vector counts as dependent (and is dropped) when its Gram-Schmidt remainder has squared
    norm at most DEPENDENT times its own; the source's test is absolute (norm^2 > 1e-32);
  * the final loop accumulates psi += v[i] b_i; the source assigns psi = b_i + v[i] b_i
    (:84), which keeps only the last vector and is not a solution of anything.
`plant` reproduces a source-text form for the mutation tests: 1 = the source's projection
coefficient, 2 = the source's final loop.
"""
from __future__ import annotations

import math

import numpy as np

DEPENDENT = 1e-12
PIVOT2 = 1e-32


def rdot(a, b):
    return float(np.dot(np.ravel(a), np.ravel(b)))


def mre_guess(chi, basis, apply_A, plant=0):
    """psi for A psi = chi from the previous solutions `basis` (list of arrays shaped as chi);
    apply_A(x) returns A x."""
    n = len(basis)
    B = [np.array(b, dtype=np.float64, copy=True) for b in basis]
    kept = [False] * n
    for i in range(n):
        own = rdot(B[i], B[i])
        if not own > 0:
            continue
        for j in range(i):
            if not kept[j]:
                continue
            vdot = rdot(B[j], B[i])
            norm = rdot(B[i], B[i]) if plant == 1 else rdot(B[j], B[j])
            B[i] = B[i] - vdot / norm * B[j]
        rem = rdot(B[i], B[i])
        if not rem > DEPENDENT * own:
            continue
        B[i] = B[i] / math.sqrt(rem)
        kept[i] = True
    M = np.zeros((n, n))
    v = np.zeros(n)
    for i in range(n):
        if not kept[i]:
            continue
        AB = apply_A(B[i])
        for j in range(n):
            if kept[j]:
                M[j, i] = rdot(B[j], AB)
        v[i] = rdot(B[i], chi)
    for i in range(n):
        if M[i, i] * M[i, i] > PIVOT2:
            diag_inv = 1.0 / M[i, i]
            M[i, :] *= diag_inv
            v[i] *= diag_inv
            for k in range(n):
                if k != i:
                    weight = M[k, i]
                    M[k, :] -= weight * M[i, :]
                    v[k] -= weight * v[i]
        else:
            v[i] = 0.0
            M[:, i] = 0.0
    psi = np.zeros_like(np.asarray(chi, dtype=np.float64))
    for i in range(n):
        if kept[i] and not math.isnan(v[i]):
            psi = B[i] + v[i] * B[i] if plant == 2 else psi + v[i] * B[i]
    return psi
