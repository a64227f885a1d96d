// CPU test of the block-CG small-matrix algebra (hila_amd/csrc/wl_small.h): Hermitian Jacobi
// eigendecomposition, SVQB orthonormalisation and the thresholded pseudo-inverse, on seeded
// random complex matrices, against the defining identities.  Prints one line per check and
// exits non-zero on the first failure.
#include <cstdio>
#include <random>

#include "wl_small.h"

using namespace wl::small;

static double maxabs(const Mat &A) {
    double m = 0;
    for (auto &v : A) m = std::max(m, std::abs(v));
    return m;
}
static Mat sub(const Mat &A, const Mat &B) {
    Mat C(A.size());
    for (size_t i = 0; i < A.size(); ++i) C[i] = A[i] - B[i];
    return C;
}
static int fails = 0;
static void check(const char *what, double err, double tol) {
    std::printf("%-58s %.3e (tol %.0e) %s\n", what, err, tol, err <= tol ? "ok" : "FAIL");
    if (!(err <= tol)) ++fails;
}

int main() {
    std::mt19937_64 rng(12345);
    std::normal_distribution<double> g;
    for (int n : {1, 2, 5, 12, 24, 32}) {
        // W: 3n x n complex, then G = W^H W (Hermitian positive definite)
        const int m = 3 * n;
        std::vector<cplx> W((size_t)m * n);
        for (auto &w : W) w = cplx(g(rng), g(rng));
        Mat G((size_t)n * n, cplx(0));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                for (int k = 0; k < m; ++k) G[(size_t)i * n + j] += std::conj(W[(size_t)k * n + i]) * W[(size_t)k * n + j];
        std::vector<double> lam;
        Mat V;
        herm_eig(n, G, lam, V);
        Mat L((size_t)n * n, cplx(0));
        for (int i = 0; i < n; ++i) L[(size_t)i * n + i] = lam[i];
        const double scale = maxabs(G);
        char buf[96];
        std::snprintf(buf, sizeof buf, "n=%2d  G V - V diag(lam)", n);
        check(buf, maxabs(sub(mul(n, G, V), mul(n, V, L))) / scale, 1e-12);
        std::snprintf(buf, sizeof buf, "n=%2d  V^H V - I", n);
        check(buf, maxabs(sub(mul(n, adj(n, V), V), identity(n))), 1e-12);
        bool asc = true;
        for (int i = 1; i < n; ++i) asc = asc && lam[i - 1] <= lam[i];
        std::snprintf(buf, sizeof buf, "n=%2d  eigenvalues ascending and positive", n);
        check(buf, (asc && lam[0] > 0) ? 0.0 : 1.0, 0.0);
        // SVQB: S^H S = G, and (W Sinv)^H (W Sinv) = I
        Mat S, Sinv;
        const int kept = svqb(n, G, 1e-14, S, Sinv);
        std::snprintf(buf, sizeof buf, "n=%2d  svqb keeps every direction", n);
        check(buf, kept == n ? 0.0 : 1.0, 0.0);
        std::snprintf(buf, sizeof buf, "n=%2d  S^H S - G", n);
        check(buf, maxabs(sub(mul(n, adj(n, S), S), G)) / scale, 1e-12);
        std::snprintf(buf, sizeof buf, "n=%2d  Sinv^H G Sinv - I", n);
        check(buf, maxabs(sub(mul(n, mul(n, adj(n, Sinv), G), Sinv), identity(n))), 1e-10);
        std::snprintf(buf, sizeof buf, "n=%2d  S Sinv - I", n);
        check(buf, maxabs(sub(mul(n, S, Sinv), identity(n))), 1e-10);
        // Cholesky fast paths on the full-rank matrix
        Mat Sc, Scinv, Mc;
        std::snprintf(buf, sizeof buf, "n=%2d  Cholesky QR keeps every direction", n);
        check(buf, orthonormaliser(n, G, 1e-10, Sc, Scinv) == n ? 0.0 : 1.0, 0.0);
        std::snprintf(buf, sizeof buf, "n=%2d  Cholesky S^H S - G", n);
        check(buf, maxabs(sub(mul(n, adj(n, Sc), Sc), G)) / scale, 1e-12);
        std::snprintf(buf, sizeof buf, "n=%2d  Cholesky S Sinv - I", n);
        check(buf, maxabs(sub(mul(n, Sc, Scinv), identity(n))), 1e-10);
        bool upper = true;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) upper = upper && Sc[(size_t)i * n + j] == cplx(0);
        std::snprintf(buf, sizeof buf, "n=%2d  Cholesky factor upper triangular", n);
        check(buf, upper ? 0.0 : 1.0, 0.0);
        inverse_herm(n, G, 1e-10, Mc);
        std::snprintf(buf, sizeof buf, "n=%2d  Cholesky inverse: inv(G) G - I", n);
        check(buf, maxabs(sub(mul(n, Mc, G), identity(n))), 1e-10);
        // pseudo-inverse of a full-rank matrix is its inverse
        Mat M;
        pinv_herm(n, G, 1e-14, M);
        std::snprintf(buf, sizeof buf, "n=%2d  pinv(G) G - I", n);
        check(buf, maxabs(sub(mul(n, M, G), identity(n))), 1e-10);
        if (n >= 2) {
            // rank deficiency: column n-1 = column 0 (a repeated right-hand side)
            for (int k = 0; k < m; ++k) W[(size_t)k * n + n - 1] = W[(size_t)k * n + 0];
            Mat Gd((size_t)n * n, cplx(0));
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < m; ++k)
                        Gd[(size_t)i * n + j] += std::conj(W[(size_t)k * n + i]) * W[(size_t)k * n + j];
            const double sd = maxabs(Gd);
            const int kd = svqb(n, Gd, 1e-12, S, Sinv);
            std::snprintf(buf, sizeof buf, "n=%2d  dependent column dropped (kept %d)", n, kd);
            check(buf, kd == n - 1 ? 0.0 : 1.0, 0.0);
            std::snprintf(buf, sizeof buf, "n=%2d  rank-deficient S^H S - G", n);
            check(buf, maxabs(sub(mul(n, adj(n, S), S), Gd)) / sd, 1e-10);
            // the fast paths fall back to the eigendecomposition and drop the same direction
            std::snprintf(buf, sizeof buf, "n=%2d  orthonormaliser falls back, drops one", n);
            check(buf, orthonormaliser(n, Gd, 1e-10, S, Sinv) == n - 1 ? 0.0 : 1.0, 0.0);
            Mat Md;
            std::snprintf(buf, sizeof buf, "n=%2d  inverse falls back to the pseudo-inverse", n);
            check(buf, inverse_herm(n, Gd, 1e-10, Md) == n - 1 ? 0.0 : 1.0, 0.0);
            pinv_herm(n, Gd, 1e-12, Md);
            std::snprintf(buf, sizeof buf, "n=%2d  rank-deficient G pinv(G) G - G", n);
            check(buf, maxabs(sub(mul(n, mul(n, Gd, Md), Gd), Gd)) / sd, 1e-9);
        }
    }
    // a zero matrix: nothing kept, no division by zero
    Mat Z((size_t)4 * 4, cplx(0)), S, Sinv, M;
    check("zero Gram: svqb keeps nothing", svqb(4, Z, 1e-12, S, Sinv) == 0 && maxabs(S) == 0 ? 0.0 : 1.0, 0.0);
    check("zero matrix: pinv is zero", pinv_herm(4, Z, 1e-12, M) == 0 && maxabs(M) == 0 ? 0.0 : 1.0, 0.0);
    std::printf(fails ? "FAILED %d\n" : "all ok\n", fails);
    return fails ? 1 : 0;
}
