// Seeded synthetic inputs for the Wilson solve path: gauge links and Wilson-vector sources.
//
// Every random number is a pure function of (seed, global lexicographic site, component), so
// the bytes written do not depend on the thread count, the order of evaluation or the machine
// (same binary, same libm).  HILA draws its inputs from a per-rank mt19937_64 stream
// (libraries/plumbing/random.cpp:46-69) on the host or per-thread XORWOW on GPUs
// (libraries/plumbing/backend_gpu/hila_gpu.cpp:287-292), so neither stream can be reproduced
// bit for bit; what is kept from HILA is the distributions:
//   - complex Gaussian: re, im ~ N(0,1) from one Box-Muller draw (random.cpp:187-201,
//     cmplx.h:344-349);
//   - reunitarization = row Gram-Schmidt (make_unitary) then determinant phase fix (fix_det)
//     (libraries/datatypes/sun_matrix.h:137-197).
// Host layout written: HILA's AoS, EVEN_SITES_FIRST order (lattice.cpp:171-186).

#include "wilson_mi355x.h"

#include <cmath>
#include <complex>
#include <cstring>

namespace {

using cd = std::complex<double>;

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// 64 random bits keyed by (seed, site, component).
inline uint64_t rbits(uint64_t seed, uint64_t site, uint64_t comp) {
    uint64_t k = splitmix64(seed ^ 0x5851F42D4C957F2DULL);
    uint64_t x = splitmix64(k ^ (site * 0xD1B54A32D192ED03ULL));
    return splitmix64(x + comp * 0xA24BAED4963EE407ULL);
}

// Uniform in (0, 1]: 53 random mantissa bits.
inline double runif(uint64_t seed, uint64_t site, uint64_t comp) {
    return (double)((rbits(seed, site, comp) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

// One complex Gaussian with re, im ~ N(0,1) from uniforms 2c, 2c+1 (Box-Muller; the radius
// uses 1-u as HILA's gaussrand2 does, random.cpp:187-201).
inline cd cgauss(uint64_t seed, uint64_t site, uint64_t c) {
    const double phi = 2.0 * M_PI * runif(seed, site, 2 * c);
    const double u = runif(seed, site, 2 * c + 1);
    const double r = std::sqrt(-2.0 * std::log(u));
    return cd(r * std::sin(phi), r * std::cos(phi));
}

struct Geo {
    int L[4];
    int64_t V;
    bool ok;
    explicit Geo(const int *d) {
        ok = d != nullptr;
        V = 1;
        for (int i = 0; i < 4 && ok; ++i) {
            L[i] = d[i];
            if (d[i] < 2 || (d[i] & 1)) ok = false;
            V *= d[i];
        }
    }
    // lexicographic index -> EVEN_SITES_FIRST index
    int64_t site_index(int64_t lex, int x, int y, int z, int t) const {
        return ((x + y + z + t) & 1) ? lex / 2 + V / 2 : lex / 2;
    }
};

// SU<N>::make_unitary + fix_det (sun_matrix.h:137-197) for N = 3, on a row-major matrix.
void reunitarize(cd m[9]) {
    for (int r = 0; r < 3; ++r) {
        double n2 = 0;
        for (int i = 0; i < 3; ++i) n2 += std::norm(m[3 * r + i]);
        n2 = 1.0 / std::sqrt(n2);
        for (int i = 0; i < 3; ++i) m[3 * r + i] *= n2;
        for (int j = r + 1; j < 3; ++j) {
            cd d = 0;
            for (int i = 0; i < 3; ++i) d += std::conj(m[3 * r + i]) * m[3 * j + i];
            for (int i = 0; i < 3; ++i) m[3 * j + i] -= d * m[3 * r + i];
        }
    }
    const cd det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                   m[2] * (m[3] * m[7] - m[4] * m[6]);
    const double t = std::arg(det) / 3.0;
    const cd ph(std::cos(-t), std::sin(-t));
    for (int i = 0; i < 9; ++i) m[i] *= ph;
}

void matmul3(const cd a[9], const cd b[9], cd c[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            cd s = 0;
            for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
            c[3 * i + j] = s;
        }
}

// exp(X) by scaling and squaring with a degree-16 Taylor polynomial (truncation error below
// 1e-20 for ||X/2^k||_F <= 1/4).
void expm3(const cd X[9], cd out[9]) {
    double nf = 0;
    for (int i = 0; i < 9; ++i) nf += std::norm(X[i]);
    nf = std::sqrt(nf);
    int k = 0;
    while (nf > 0.25) {
        nf *= 0.5;
        ++k;
    }
    const double s = std::ldexp(1.0, -k);
    cd A[9], T[9], P[9], tmp[9];
    for (int i = 0; i < 9; ++i) {
        A[i] = X[i] * s;
        T[i] = (i % 4 == 0) ? 1.0 : 0.0;
        P[i] = T[i];
    }
    for (int n = 1; n <= 16; ++n) {
        matmul3(T, A, tmp);
        for (int i = 0; i < 9; ++i) {
            T[i] = tmp[i] / (double)n;
            P[i] += T[i];
        }
    }
    for (int q = 0; q < k; ++q) {
        matmul3(P, P, tmp);
        std::memcpy(P, tmp, sizeof(P));
    }
    std::memcpy(out, P, sizeof(P));
}

// SU<3>::random(nhits) (sun_matrix.h:211-240): from the identity, each hit multiplies rows
// (r, q) = (0,1), (0,2), (1,2) in turn from the left by a random SU(2) matrix, and the result is
// reunitarized after every 16th hit and at the end.  The SU(2) matrix is the normalized
// Gaussian 4-vector v in the Pauli form [[v0 + i v3, v2 + i v1], [-v2 + i v1, v0 - i v3]]
// (its N == 2 branch); (v0, v1) and (v2, v3) are the two Box-Muller pairs of one complex draw
// each, as Matrix::gaussian_random fills a real 4-vector (matrix.h:1507-1514).
constexpr int kSu2Hits = 16;

void su3_by_su2_hits(uint64_t seed, int64_t lex, int mu, cd m[9]) {
    for (int i = 0; i < 9; ++i) m[i] = (i % 4 == 0) ? 1.0 : 0.0;
    static const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    const uint64_t cbase = 1024 + (uint64_t)mu * 2 * 3 * kSu2Hits;  // own component range
    for (int h = 1; h <= kSu2Hits; ++h) {
        for (int k = 0; k < 3; ++k) {
            const uint64_t c = cbase + 2 * (3 * (uint64_t)(h - 1) + k);
            const cd g0 = cgauss(seed, lex, c), g1 = cgauss(seed, lex, c + 1);
            double v[4] = {g0.real(), g0.imag(), g1.real(), g1.imag()};
            const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]);
            for (double &x : v) x /= n;
            const cd a(v[0], v[3]), b(v[2], v[1]), cc(-v[2], v[1]), d(v[0], -v[3]);
            const int r = pr[k][0], q = pr[k][1];
            for (int i = 0; i < 3; ++i) {  // rows r, q <- [[a, b], [cc, d]] (rows r, q)
                const cd x = m[3 * r + i], y = m[3 * q + i];
                m[3 * r + i] = a * x + b * y;
                m[3 * q + i] = cc * x + d * y;
            }
        }
        if (h % 16 == 0) reunitarize(m);
    }
    if (kSu2Hits % 16 != 0) reunitarize(m);
}

void gen_link(int kind, double eps, uint64_t seed, int64_t lex, int mu, double *dst) {
    cd m[9];
    const uint64_t cbase = (uint64_t)mu * 16;
    if (kind == 0) {
        for (int i = 0; i < 9; ++i) m[i] = (i % 4 == 0) ? 1.0 : 0.0;
    } else if (kind == 3) {
        su3_by_su2_hits(seed, lex, mu, m);
    } else if (kind == 1) {
        // Haar on U(3): Gram-Schmidt of a complex Ginibre matrix (positive diagonal of R);
        // the determinant phase fix then maps it to Haar on SU(3).
        for (int i = 0; i < 9; ++i) m[i] = cgauss(seed, lex, cbase + i);
        reunitarize(m);
    } else {
        cd a[9], H[9], X[9];
        for (int i = 0; i < 9; ++i) a[i] = cgauss(seed, lex, cbase + i);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) H[3 * i + j] = 0.5 * (a[3 * i + j] + std::conj(a[3 * j + i]));
        const cd tr = (H[0] + H[4] + H[8]) / 3.0;
        for (int i = 0; i < 3; ++i) H[4 * i] -= tr;
        for (int i = 0; i < 9; ++i) X[i] = cd(0.0, eps) * H[i];
        expm3(X, m);
        reunitarize(m);
    }
    for (int i = 0; i < 9; ++i) {
        dst[2 * i] = m[i].real();
        dst[2 * i + 1] = m[i].imag();
    }
}

}  // namespace

extern "C" int wl_gen_gauge(const int dims[4], int kind, double eps, uint64_t seed,
                            double *const U[4]) {
    Geo g(dims);
    if (!g.ok || !U || kind < 0 || kind > 3) return WL_E_ARG;
    for (int mu = 0; mu < 4; ++mu)
        if (!U[mu]) return WL_E_ARG;
    const int Lx = g.L[0], Ly = g.L[1], Lz = g.L[2];
#pragma omp parallel for schedule(static)
    for (int64_t lex = 0; lex < g.V; ++lex) {
        const int x = (int)(lex % Lx), y = (int)((lex / Lx) % Ly);
        const int z = (int)((lex / ((int64_t)Lx * Ly)) % Lz);
        const int t = (int)(lex / ((int64_t)Lx * Ly * Lz));
        const int64_t s = g.site_index(lex, x, y, z, t);
        for (int mu = 0; mu < 4; ++mu) gen_link(kind, eps, seed, lex, mu, U[mu] + 18 * s);
    }
    return WL_OK;
}

extern "C" int wl_gen_spinor(const int dims[4], uint64_t seed, int parity_mask, double *out) {
    Geo g(dims);
    if (!g.ok || !out || parity_mask < 0 || parity_mask > 3) return WL_E_ARG;
    const int Lx = g.L[0], Ly = g.L[1], Lz = g.L[2];
#pragma omp parallel for schedule(static)
    for (int64_t lex = 0; lex < g.V; ++lex) {
        const int x = (int)(lex % Lx), y = (int)((lex / Lx) % Ly);
        const int z = (int)((lex / ((int64_t)Lx * Ly)) % Lz);
        const int t = (int)(lex / ((int64_t)Lx * Ly * Lz));
        const int par = ((x + y + z + t) & 1) ? 2 : 1;
        double *dst = out + 24 * g.site_index(lex, x, y, z, t);
        if (parity_mask & par) {
            for (int c = 0; c < 12; ++c) {
                const cd v = cgauss(seed, lex, (uint64_t)c);
                dst[2 * c] = v.real();
                dst[2 * c + 1] = v.imag();
            }
        } else {
            std::memset(dst, 0, 24 * sizeof(double));
        }
    }
    return WL_OK;
}

extern "C" int wl_gen_point(const int dims[4], uint64_t seed, int spin, int colour, double *out,
                            int64_t *site_out) {
    Geo g(dims);
    if (!g.ok || !out || spin < 0 || spin > 3 || colour < 0 || colour > 2) return WL_E_ARG;
    std::memset(out, 0, sizeof(double) * 24 * (size_t)g.V);
    const int64_t lex = (int64_t)(rbits(seed, 0x7FFFFFFFULL, 0) % (uint64_t)g.V);
    const int Lx = g.L[0], Ly = g.L[1], Lz = g.L[2];
    const int x = (int)(lex % Lx), y = (int)((lex / Lx) % Ly);
    const int z = (int)((lex / ((int64_t)Lx * Ly)) % Lz);
    const int t = (int)(lex / ((int64_t)Lx * Ly * Lz));
    const int64_t s = g.site_index(lex, x, y, z, t);
    out[24 * s + 2 * (3 * spin + colour)] = 1.0;
    if (site_out) *site_out = s;
    return WL_OK;
}
