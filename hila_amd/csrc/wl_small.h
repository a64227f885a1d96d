// Dense algebra on small complex matrices (the N x N blocks of block CG, N <= 64), on the host
// in FP64: Hermitian eigendecomposition by cyclic Jacobi rotations, the SVQB orthonormaliser, a
// thresholded pseudo-inverse, and their Cholesky fast paths for well-conditioned matrices.
// Matrices are row-major n x n arrays of complex<double>.
// Header-only, host C++ only (no HIP), so that the CPU tests compile it on their own.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace wl {
namespace small {

using cplx = std::complex<double>;
using Mat = std::vector<cplx>;  // row-major n x n

// a b without the C99 special-value handling of std::complex multiplication
inline cplx cmul(cplx a, cplx b) {
    return {a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real()};
}

inline Mat identity(int n) {
    Mat I((size_t)n * n, cplx(0));
    for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
    return I;
}

// C = A B
inline Mat mul(int n, const Mat &A, const Mat &B) {
    Mat C((size_t)n * n, cplx(0));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const cplx a = A[(size_t)i * n + k];
            if (a == cplx(0)) continue;
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += cmul(a, B[(size_t)k * n + j]);
        }
    return C;
}

// A^H
inline Mat adj(int n, const Mat &A) {
    Mat B((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) B[(size_t)j * n + i] = std::conj(A[(size_t)i * n + j]);
    return B;
}

// Eigendecomposition of the Hermitian part of A: A = V diag(lam) V^H, V unitary, lam ascending.
// Cyclic Jacobi: each rotation zeroes one off-diagonal pair after a phase that makes it real.
inline void herm_eig(int n, const Mat &Ain, std::vector<double> &lam, Mat &V) {
    Mat A((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            A[(size_t)i * n + j] = 0.5 * (Ain[(size_t)i * n + j] + std::conj(Ain[(size_t)j * n + i]));
    V = identity(n);
    double diag = 0;
    for (int i = 0; i < n; ++i) diag += std::norm(A[(size_t)i * n + i]);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j) off += std::norm(A[(size_t)i * n + j]);
        if (off <= 1e-30 * (diag + off) || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const cplx apq = A[(size_t)p * n + q];
                const double r = std::abs(apq);
                if (r == 0.0) continue;
                const double alpha = A[(size_t)p * n + p].real(), beta = A[(size_t)q * n + q].real();
                const cplx ph = apq / r;  // e^{i phi}
                // real rotation of [[alpha, r], [r, beta]]
                const double tau = (beta - alpha) / (2 * r);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
                const double c = 1 / std::sqrt(1 + t * t), s = t * c;
                // U = diag(1, conj(ph)) [[c, s], [-s, c]]: columns p, q mix as
                // new_p = c col_p - s conj(ph) col_q, new_q = s col_p + c conj(ph) col_q
                const cplx upp = c, uqp = -s * std::conj(ph), upq = s, uqq = c * std::conj(ph);
                for (int k = 0; k < n; ++k) {  // A <- A U
                    const cplx akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = akp * upp + akq * uqp;
                    A[(size_t)k * n + q] = akp * upq + akq * uqq;
                }
                for (int k = 0; k < n; ++k) {  // A <- U^H A
                    const cplx apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = std::conj(upp) * apk + std::conj(uqp) * aqk;
                    A[(size_t)q * n + k] = std::conj(upq) * apk + std::conj(uqq) * aqk;
                }
                A[(size_t)p * n + q] = A[(size_t)q * n + p] = 0.0;
                A[(size_t)p * n + p] = A[(size_t)p * n + p].real();
                A[(size_t)q * n + q] = A[(size_t)q * n + q].real();
                for (int k = 0; k < n; ++k) {  // V <- V U
                    const cplx vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = vkp * upp + vkq * uqp;
                    V[(size_t)k * n + q] = vkp * upq + vkq * uqq;
                }
            }
    }
    // ascending order
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(),
              [&](int a, int b) { return A[(size_t)a * n + a].real() < A[(size_t)b * n + b].real(); });
    lam.assign(n, 0.0);
    Mat W((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        lam[j] = A[(size_t)idx[j] * n + idx[j]].real();
        for (int k = 0; k < n; ++k) W[(size_t)k * n + j] = V[(size_t)k * n + idx[j]];
    }
    V = W;
}

// SVQB orthonormalisation from the Gram matrix G = W^H W of n columns W: W = Q S with
// S = diag(sqrt(lam)) V^H and Q = W Sinv, Sinv = V diag(1 / sqrt(lam)), Q's columns orthonormal.
// Directions with lam <= tol * max(lam) (dependent or zero columns) are dropped: their rows of S
// and columns of Sinv are zero.  Returns the number kept.
inline int svqb(int n, const Mat &G, double tol, Mat &S, Mat &Sinv) {
    std::vector<double> lam;
    Mat V;
    herm_eig(n, G, lam, V);
    const double lmax = std::max(0.0, lam.back());
    S.assign((size_t)n * n, cplx(0));
    Sinv.assign((size_t)n * n, cplx(0));
    int kept = 0;
    for (int j = 0; j < n; ++j) {
        if (!(lam[j] > tol * lmax) || lmax <= 0) continue;
        ++kept;
        const double sq = std::sqrt(lam[j]);
        for (int k = 0; k < n; ++k) {
            S[(size_t)j * n + k] = sq * std::conj(V[(size_t)k * n + j]);
            Sinv[(size_t)k * n + j] = V[(size_t)k * n + j] / sq;
        }
    }
    return kept;
}

// Thresholded pseudo-inverse of a Hermitian positive semi-definite H: V diag(1/lam) V^H over the
// eigenvalues above tol * max(lam).  Returns the number kept.
inline int pinv_herm(int n, const Mat &H, double tol, Mat &M) {
    std::vector<double> lam;
    Mat V;
    herm_eig(n, H, lam, V);
    const double lmax = std::max(0.0, lam.back());
    M.assign((size_t)n * n, cplx(0));
    int kept = 0;
    for (int k = 0; k < n; ++k) {
        if (!(lam[k] > tol * lmax) || lmax <= 0) continue;
        ++kept;
        const double inv = 1 / lam[k];
        for (int i = 0; i < n; ++i) {
            const cplx vik = V[(size_t)i * n + k] * inv;
            for (int j = 0; j < n; ++j) M[(size_t)i * n + j] += vik * std::conj(V[(size_t)j * n + k]);
        }
    }
    return kept;
}

// Cholesky factor of a Hermitian positive definite G: G = R^H R, R upper triangular.  Fails
// (returns false) when a pivot is not above tol times the largest diagonal entry: G is then
// (numerically) semi-definite and the caller takes the eigendecomposition instead.
inline bool chol_upper(int n, const Mat &G, double tol, Mat &R) {
    double dmax = 0;
    for (int i = 0; i < n; ++i) dmax = std::max(dmax, G[(size_t)i * n + i].real());
    if (!(dmax > 0)) return false;
    R.assign((size_t)n * n, cplx(0));
    for (int i = 0; i < n; ++i) {
        double d = G[(size_t)i * n + i].real();
        for (int k = 0; k < i; ++k) d -= std::norm(R[(size_t)k * n + i]);
        if (!(d > tol * dmax)) return false;
        const double rii = std::sqrt(d);
        R[(size_t)i * n + i] = rii;
        for (int j = i + 1; j < n; ++j) {
            cplx v = G[(size_t)i * n + j];
            for (int k = 0; k < i; ++k) v -= cmul(std::conj(R[(size_t)k * n + i]), R[(size_t)k * n + j]);
            R[(size_t)i * n + j] = v / rii;
        }
    }
    return true;
}

// inverse of an upper triangular R with a nonzero diagonal
inline Mat inv_upper(int n, const Mat &R) {
    Mat X((size_t)n * n, cplx(0));
    for (int j = 0; j < n; ++j) {
        X[(size_t)j * n + j] = 1.0 / R[(size_t)j * n + j];
        for (int i = j - 1; i >= 0; --i) {
            cplx v = 0;
            for (int k = i + 1; k <= j; ++k) v += cmul(R[(size_t)i * n + k], X[(size_t)k * n + j]);
            X[(size_t)i * n + j] = -v / R[(size_t)i * n + i];
        }
    }
    return X;
}

// Orthonormaliser from the Gram matrix G = W^H W: W = Q S, Q = W Sinv.  Cholesky QR (S = R,
// upper triangular) when every pivot clears tol; otherwise SVQB, which drops dependent
// directions.  Returns the number of directions kept.
inline int orthonormaliser(int n, const Mat &G, double tol, Mat &S, Mat &Sinv) {
    if (chol_upper(n, G, tol, S)) {
        Sinv = inv_upper(n, S);
        return n;
    }
    return svqb(n, G, tol, S, Sinv);
}

// Inverse of a Hermitian positive definite H through its Cholesky factor when every pivot
// clears tol, else the thresholded pseudo-inverse.  Returns the rank used.
inline int inverse_herm(int n, const Mat &H, double tol, Mat &M) {
    Mat R;
    if (chol_upper(n, H, tol, R)) {
        const Mat X = inv_upper(n, R);  // H^-1 = X X^H
        M.assign((size_t)n * n, cplx(0));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                cplx v = 0;
                for (int k = std::max(i, j); k < n; ++k) v += cmul(X[(size_t)i * n + k], std::conj(X[(size_t)j * n + k]));
                M[(size_t)i * n + j] = v;
            }
        return n;
    }
    return pinv_herm(n, H, tol, M);
}

}  // namespace small
}  // namespace wl
