// Storage formats of Wilson-vector and gauge fields on the device, and the per-site load and
// store of a spinor in each.
//
//   double  FP64 complex planes, FP64 arithmetic
//   float   FP32 complex planes, FP32 arithmetic
//   f16s    FP16 complex numbers (re, im as two halves in 4 bytes) with one FP32 scale per site
//           and rhs, FP32 arithmetic.  A site's 24 reals are stored as v / m, m = the largest
//           |re|, |im| of the site, so each stored half lies in [-1, 1] and the relative error
//           of a component is at most 2^-11 of the site's largest one.  Links are stored
//           without scale (SU(3) entries are bounded by 1).
// A right-hand side of an N-rhs field is a block of `planes` x `stride` elements (stride S, a
// multiple of 64 sites).  Its 12 spin-colour components are packed P to a 16-byte vector
// (P = 1 FP64, 2 FP32, 4 FP16), so that a lane reads a site's spinor with 12 / P vector loads,
// and stored in site tiles: the sites are cut into tiles of 64, and tile t holds, one after
// the other, vector 0 of its 64 sites, then vector 1, and so on (tile_elem).  A wave's 64
// sites thus read and write one contiguous 12 x 64 x 16 = 12 KB stretch (FP64) instead of 12
// pieces a plane apart.  f16s adds the plane of FP32 site scales at element 12S (13S elements
// of 4 bytes in all).  Links likewise, in a (parity, direction) block of 9S: entries 0-7 of a
// site packed P to a vector, and (for P > 1) entry 8, not packed, after them in the tile.
#pragma once

#include "wl_device.h"

namespace wl {

struct f16s {};  // tag of the FP16-with-site-scale format

struct alignas(4) h2 {
    _Float16 x, y;
};

template <typename T>
struct St;
template <>
struct St<double> {
    using R = double;       // arithmetic
    using E = cx<double>;   // stored element
    static constexpr int planes = 12;
    static constexpr int pack = 1;  // complex numbers per 16-byte vector
    static constexpr bool scaled = false;
};
template <>
struct St<float> {
    using R = float;
    using E = cx<float>;
    static constexpr int planes = 12;
    static constexpr int pack = 2;
    static constexpr bool scaled = false;
};
template <>
struct St<f16s> {
    using R = float;
    using E = h2;
    static constexpr int planes = 13;
    static constexpr int pack = 4;
    static constexpr bool scaled = true;
};

// One stored complex element, widened to the arithmetic type.  NT: nontemporal load.
template <bool NT>
__device__ __forceinline__ cx<double> lde(const cx<double> *p) {
    return ldc<NT>(p);
}
template <bool NT>
__device__ __forceinline__ cx<float> lde(const cx<float> *p) {
    return ldc<NT>(p);
}
template <bool NT>
__device__ __forceinline__ cx<float> lde(const h2 *p) {
    unsigned bits;
    if constexpr (NT)
        bits = __builtin_nontemporal_load((const unsigned *)p);
    else
        bits = *(const unsigned *)p;
    const h2 v = __builtin_bit_cast(h2, bits);
    return {(float)v.x, (float)v.y};
}

// One stored element as it is in memory (kept packed in registers until used), and its
// widening to the arithmetic type.
template <bool NT>
__device__ __forceinline__ cx<double> ldraw(const cx<double> *p) {
    return ldc<NT>(p);
}
template <bool NT>
__device__ __forceinline__ cx<float> ldraw(const cx<float> *p) {
    return ldc<NT>(p);
}
template <bool NT>
__device__ __forceinline__ h2 ldraw(const h2 *p) {
    unsigned bits;
    if constexpr (NT)
        bits = __builtin_nontemporal_load((const unsigned *)p);
    else
        bits = *(const unsigned *)p;
    return __builtin_bit_cast(h2, bits);
}
__device__ __forceinline__ cx<double> widen(cx<double> v) { return v; }
__device__ __forceinline__ cx<float> widen(cx<float> v) { return v; }
__device__ __forceinline__ cx<float> widen(h2 v) { return {(float)v.x, (float)v.y}; }

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
template <typename E, int P>
struct Packed {
    E e[P];
};

// One 16-byte vector of P packed elements at v (16-byte aligned).
template <bool NT, typename T>
__device__ __forceinline__ Packed<typename St<T>::E, St<T>::pack> ld_vec(const typename St<T>::E *v) {
    const u32x4 *p = reinterpret_cast<const u32x4 *>(v);
    u32x4 bits;
    if constexpr (NT)
        bits = __builtin_nontemporal_load(p);
    else
        bits = *p;
    return __builtin_bit_cast(Packed<typename St<T>::E, St<T>::pack>, bits);
}
template <bool NT, typename T>
__device__ __forceinline__ void st_vec(typename St<T>::E *v, const Packed<typename St<T>::E, St<T>::pack> &x) {
    const u32x4 bits = __builtin_bit_cast(u32x4, x);
    u32x4 *p = reinterpret_cast<u32x4 *>(v);
    if constexpr (NT)
        __builtin_nontemporal_store(bits, p);
    else
        *p = bits;
}

// Sites per tile of the site-tiled layout: one wave's sites.
constexpr int kTile = 64;

// Element of entry e of site k in a block of ne elements per site, packed P to a vector
// (entries e < ne - ne % P), the rest one per site after the packed vectors of the tile.
template <int P>
__host__ __device__ __forceinline__ size_t tile_elem(int e, int ne, int k) {
    const size_t tile = (size_t)(k / kTile) * ne * kTile;
    const int lane = k % kTile;
    const int npk = ne - ne % P;
    if (e < npk) return tile + (size_t)(e / P) * P * kTile + (size_t)P * lane + (e % P);
    return tile + (size_t)e * kTile + lane;
}

// Element of spinor component c of site k in an rhs block; of link entry e in a link block of
// LF entries per site (9, or 6 for two-row links).  S, the block stride, only sizes the block.
template <typename T>
__device__ __forceinline__ size_t comp_elem(int c, int S, int k) {
    (void)S;
    return tile_elem<St<T>::pack>(c, 12, k);
}
template <typename T, int LF = 9>
__device__ __forceinline__ size_t link_entry(int e, int S, int k) {
    (void)S;
    return tile_elem<St<T>::pack>(e, LF, k);
}

// Raw (packed) loads of a site's 12 spinor components and 9 link entries.
template <bool NT, typename T>
__device__ __forceinline__ void ld_spinor_raw(typename St<T>::E (&w)[12], const typename St<T>::E *base, int S,
                                              int k) {
    constexpr int P = St<T>::pack;
#pragma unroll
    for (int q = 0; q < 12 / P; ++q) {
        const auto v = ld_vec<NT, T>(base + tile_elem<P>(P * q, 12, k));
#pragma unroll
        for (int i = 0; i < P; ++i) w[P * q + i] = v.e[i];
    }
}

// Row 2 of a special unitary 3x3 matrix from rows 0 and 1 (row-major entries u[0..5]):
// U[2][j] = conj(eps_jkl U[0][k] U[1][l]), i.e. conj(row 0 x row 1).
template <typename R>
__device__ __forceinline__ void third_row(cx<R> (&u)[9]) {
    const cx<R> a = cmul(u[1], u[5]) - cmul(u[2], u[4]);
    const cx<R> b = cmul(u[2], u[3]) - cmul(u[0], u[5]);
    const cx<R> c = cmul(u[0], u[4]) - cmul(u[1], u[3]);
    u[6] = {a.x, -a.y};
    u[7] = {b.x, -b.y};
    u[8] = {c.x, -c.y};
}

// The 9 entries of the link at site k of a link block.  LF = 9: all entries stored (block of
// 9S elements).  LF = 6: rows 0 and 1 stored (block of 6S elements, tiles of 6 entries per
// site), row 2 rebuilt by third_row: two-row storage, for links the host has checked to be
// special unitary to rounding (FP64 and FP32 storage).
template <bool NT, typename T, int LF = 9>
__device__ __forceinline__ void ld_link_raw(typename St<T>::E (&u)[9], const typename St<T>::E *base, int S,
                                            int k) {
    constexpr int P = St<T>::pack;
    static_assert(LF == 9 || (LF == 6 && !St<T>::scaled), "two-row links: FP64 and FP32 storage");
#pragma unroll
    for (int q = 0; q < (LF == 9 ? 8 : 6) / P; ++q) {
        const auto v = ld_vec<NT, T>(base + tile_elem<P>(P * q, LF, k));
#pragma unroll
        for (int i = 0; i < P; ++i) u[P * q + i] = v.e[i];
    }
    if constexpr (LF == 9)
        u[8] = ldraw<NT>(base + tile_elem<P>(8, 9, k));
    else
        third_row(u);
}

// FP32 scale plane of the rhs block starting at `base` (f16s only).
__device__ __forceinline__ const float *scale_plane(const h2 *base, int S) {
    return reinterpret_cast<const float *>(base + (size_t)12 * S);
}
__device__ __forceinline__ float *scale_plane(h2 *base, int S) {
    return reinterpret_cast<float *>(base + (size_t)12 * S);
}

// Spinor of site k of the rhs block `base` (component c at base[c * S + k]), in the
// arithmetic type, scale applied.
template <typename T, bool NT = false>
__device__ __forceinline__ void ld_site(const typename St<T>::E *base, int S, int k,
                                        cx<typename St<T>::R> (&w)[4][3]) {
    typename St<T>::E raw[12];
    ld_spinor_raw<NT, T>(raw, base, S, k);
#pragma unroll
    for (int c = 0; c < 12; ++c) w[c / 3][c % 3] = widen(raw[c]);
    if constexpr (St<T>::scaled) {
        const float m = scale_plane(base, S)[k];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) w[s][c] = m * w[s][c];
    }
}

// Store the spinor v at site k of the rhs block `base`.  For f16s, v is replaced by the values
// as stored (dequantised), so that norms taken from it are those of the stored field.
template <typename T, bool NT = false>
__device__ __forceinline__ void st_site(typename St<T>::E *base, int S, int k,
                                        cx<typename St<T>::R> (&v)[4][3]) {
    constexpr int P = St<T>::pack;
    using E = typename St<T>::E;
    if constexpr (!St<T>::scaled) {
#pragma unroll
        for (int q = 0; q < 12 / P; ++q) {
            Packed<E, P> x;
#pragma unroll
            for (int i = 0; i < P; ++i) x.e[i] = v[(P * q + i) / 3][(P * q + i) % 3];
            st_vec<NT, T>(base + tile_elem<P>(P * q, 12, k), x);
        }
    } else {
        float m = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int c = 0; c < 3; ++c) m = fmaxf(m, fmaxf(fabsf(v[s][c].x), fabsf(v[s][c].y)));
        const float inv = m > 0.f ? 1.0f / m : 0.f;
#pragma unroll
        for (int q = 0; q < 12 / P; ++q) {
            Packed<E, P> x;
#pragma unroll
            for (int i = 0; i < P; ++i) {
                const int c = P * q + i;
                h2 h;
                h.x = (_Float16)(v[c / 3][c % 3].x * inv);
                h.y = (_Float16)(v[c / 3][c % 3].y * inv);
                x.e[i] = h;
                v[c / 3][c % 3] = {m * (float)h.x, m * (float)h.y};
            }
            st_vec<NT, T>(base + tile_elem<P>(P * q, 12, k), x);
        }
        if constexpr (NT)
            __builtin_nontemporal_store(m, scale_plane(base, S) + k);
        else
            scale_plane(base, S)[k] = m;
    }
}

// ------------------------------------ packed FP16 hop arithmetic, FP32 accumulation ------
// The FP16 hop keeps links and spinors packed as stored: the spin projection is one
// v_pk_add_f16 per complex number and operand form (one FP16 rounding of each half, the same
// size as the storage rounding of the operands), and each complex multiply-add of U h is two
// v_dot2_f32_f16 (products of halves, summed into an FP32 accumulator) instead of four FP32
// FMAs after widening.  The projection is written as VALU assembly because the compiler does
// not fold the half selections and negations into v_pk_add_f16's modifiers (op_sel /
// op_sel_hi pick the half of each source used for the low / high result, neg_lo / neg_hi
// negate it); the dot products are compiler builtins, so that the compiler sees them and keeps
// the wait states a dot-product result needs before another VALU instruction reads it.

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// a + C b for stored complex numbers, C = +1 (1), -1 (-1), +i (2) or -i (-2), as the packed
// operand of the dot products, in form F of the result z = (re, im):
//   F = 0: (re, im)   1: (im, re)   2: (re, -im)   3: (im, -re)
template <int C, int F>
__device__ __forceinline__ h2 pk_cadd(h2 a, h2 b) {
    const unsigned ua = __builtin_bit_cast(unsigned, a), ub = __builtin_bit_cast(unsigned, b);
    unsigned d;
#define WL_PK(MODS) asm("v_pk_add_f16 %0, %1, %2 " MODS : "=v"(d) : "v"(ua), "v"(ub))
    // same halves (C = +-1) or crossed (C = +-i); swapped result (F odd); negated high half
    // (F >= 2): neg_hi flipped for both sources
    if constexpr (C == 1) {
        if constexpr (F == 0) WL_PK("");
        else if constexpr (F == 1) WL_PK("op_sel:[1,1] op_sel_hi:[0,0]");
        else if constexpr (F == 2) WL_PK("neg_hi:[1,1]");
        else WL_PK("op_sel:[1,1] op_sel_hi:[0,0] neg_hi:[1,1]");
    } else if constexpr (C == -1) {
        if constexpr (F == 0) WL_PK("neg_lo:[0,1] neg_hi:[0,1]");
        else if constexpr (F == 1) WL_PK("op_sel:[1,1] op_sel_hi:[0,0] neg_lo:[0,1] neg_hi:[0,1]");
        else if constexpr (F == 2) WL_PK("neg_lo:[0,1] neg_hi:[1,0]");
        else WL_PK("op_sel:[1,1] op_sel_hi:[0,0] neg_lo:[0,1] neg_hi:[1,0]");
    } else if constexpr (C == 2) {   // z = (a.re - b.im, a.im + b.re)
        if constexpr (F == 0) WL_PK("op_sel:[0,1] op_sel_hi:[1,0] neg_lo:[0,1]");
        else if constexpr (F == 1) WL_PK("op_sel:[1,0] op_sel_hi:[0,1] neg_hi:[0,1]");
        else if constexpr (F == 2) WL_PK("op_sel:[0,1] op_sel_hi:[1,0] neg_lo:[0,1] neg_hi:[1,1]");
        else WL_PK("op_sel:[1,0] op_sel_hi:[0,1] neg_hi:[1,0]");
    } else {                         // z = (a.re + b.im, a.im - b.re)
        if constexpr (F == 0) WL_PK("op_sel:[0,1] op_sel_hi:[1,0] neg_hi:[0,1]");
        else if constexpr (F == 1) WL_PK("op_sel:[1,0] op_sel_hi:[0,1] neg_lo:[0,1]");
        else if constexpr (F == 2) WL_PK("op_sel:[0,1] op_sel_hi:[1,0] neg_hi:[1,0]");
        else WL_PK("op_sel:[1,0] op_sel_hi:[0,1] neg_lo:[0,1] neg_hi:[1,1]");
    }
#undef WL_PK
    return __builtin_bit_cast(h2, d);
}

// u.x v.x + u.y v.y + c in FP32 (v_dot2_f32_f16)
__device__ __forceinline__ float dot2(h2 u, h2 v, float c) {
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, u), __builtin_bit_cast(half2_t, v), c, false);
}

// project<MU, S> (wl_device.h) of a spinor in FP16 storage, kept packed, as the two operands
// of U h (FWD: forms 2 and 1) or of U^dag h (forms 0 and 3):
//   U h:      re = sum u.re h.re - u.im h.im = dot(u, (re, -im)),  im = dot(u, (im, re))
//   U^dag h:  re = sum u.re h.re + u.im h.im = dot(u, (re, im)),   im = dot(u, (im, -re))
// with u = U[i][j] for U h and u = U[j][i] for U^dag h.
template <int MU, int S, bool FWD>
__device__ __forceinline__ void project_pk(const h2 (&w)[12], h2 (&hr)[2][3], h2 (&hi)[2][3]) {
    constexpr int b0 = (MU == 0 || MU == 1) ? 3 : 2;
    constexpr int b1 = (MU == 0 || MU == 1) ? 2 : 3;
    constexpr int c0 = MU == 0 ? 2 * S : MU == 1 ? -S : MU == 2 ? 2 * S : S;
    constexpr int c1 = MU == 0 ? 2 * S : MU == 1 ? S : MU == 2 ? -2 * S : S;
    constexpr int FR = FWD ? 2 : 0, FI = FWD ? 1 : 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        hr[0][c] = pk_cadd<c0, FR>(w[c], w[b0 * 3 + c]);
        hi[0][c] = pk_cadd<c0, FI>(w[c], w[b0 * 3 + c]);
        hr[1][c] = pk_cadd<c1, FR>(w[3 + c], w[b1 * 3 + c]);
        hi[1][c] = pk_cadd<c1, FI>(w[3 + c], w[b1 * 3 + c]);
    }
}

// v = U h (FWD) or U^dag h for a link of 9 stored entries (row-major), from project_pk's
// operands.
template <bool FWD>
__device__ __forceinline__ void su3_mul_pk(const h2 (&u)[9], const h2 (&hr)[2][3], const h2 (&hi)[2][3],
                                           cx<float> (&v)[2][3]) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const h2 u0 = FWD ? u[i * 3 + 0] : u[0 * 3 + i];
            const h2 u1 = FWD ? u[i * 3 + 1] : u[1 * 3 + i];
            const h2 u2 = FWD ? u[i * 3 + 2] : u[2 * 3 + i];
            const float re = dot2(u2, hr[s][2], dot2(u1, hr[s][1], dot2(u0, hr[s][0], 0.f)));
            const float im = dot2(u2, hi[s][2], dot2(u1, hi[s][1], dot2(u0, hi[s][0], 0.f)));
            v[s][i] = {re, im};
        }
}

// Host side: precision code -> planes per rhs and bytes per stored element.
inline constexpr int planes_of(int prec) { return prec == 2 ? 13 : 12; }
inline constexpr int elem_bytes(int prec) { return prec == 0 ? 16 : (prec == 1 ? 8 : 4); }

}  // namespace wl
