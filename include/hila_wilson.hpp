// hila_wilson.hpp -- C++ mirror of HILA's Wilson even-odd operator and CG call surface over the
// C ABI of libhila_wilson.so (wilson_mi355x.h).  Header-only.
//
//   Dirac_Wilson_evenodd<SU3d> D(kappa, U);     libraries/dirac/wilson.h:234-269
//     ... D(kappa, U.fdir), D(D2), and typename decltype(D)::type_flt D_flt(D, U32)
//     (the single-precision operator, wilson.h:250-253,267-269)
//   D.apply(in, out);  D.dagger(in, out);       wilson.h:272-289   (out[ODD] = 0)
//   D.kappa, D.par == EVEN, vector_type          wilson.h:241-256
//   CG<Dirac_Wilson_evenodd<SU3d>> inv(D, accuracy, maxiters);   conjugate_gradient.h:33-40
//   inv.apply(in, out);   out = initial guess & result           conjugate_gradient.h:46-104
//
// Fields follow HILA's Field<WilsonVector<3,double>> on one lattice: full-lattice storage with
// EVEN_SITES_FIRST order; here the data of each parity lives on the GPU and a host copy in
// HILA's AoS layout is exchanged on request.  As in HILA, errors are fatal by default: a failing
// call throws hila_wilson::error (HILA calls hila::terminate, backend_gpu/defs.h:288-297).
#pragma once

#include <array>
#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "wilson_mi355x.h"

namespace hila_wilson {

// HILA's enum class Parity {none, even, odd, all} (libraries/plumbing/coordinates.h:182)
enum class Parity : int { none = 0, even = 1, odd = 2, all = 3 };
constexpr Parity EVEN = Parity::even;
constexpr Parity ODD = Parity::odd;
constexpr Parity ALL = Parity::all;

struct error : std::runtime_error {
    int code;
    error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline int check(int rc, const char *what) {
    if (rc != WL_OK) throw error(rc, std::string(what) + ": " + wl_last_error());
    return rc;
}

class MatrixField;

// The lattice geometry on one GPU (HILA's lattice.setup(size), lattice.h:477-480).
class Lattice {
  public:
    explicit Lattice(std::array<int, 4> size, bool antiperiodic_t = true, int device = 0)
        : size_(size) {
        check(wl_create(size.data(), antiperiodic_t ? 1 : 0, device, &ctx_), "wl_create");
        volume_ = wl_volume(ctx_);
    }
    ~Lattice() { wl_destroy(ctx_); }
    Lattice(const Lattice &) = delete;
    Lattice &operator=(const Lattice &) = delete;
    wl_ctx *ctx() const { return ctx_; }
    int64_t volume() const { return volume_; }
    const std::array<int, 4> &size() const { return size_; }
    void synchronize() { check(wl_sync(ctx_), "wl_sync"); }
    // kernel options (WL_OPT_*, wl_set_option): launch shapes, or WL_OPT_LINK_ROWS = 2 for
    // two-row links on special unitary gauge fields
    void set_option(int option, int value) { check(wl_set_option(ctx_, option, value), "wl_set_option"); }
    int option(int option) const {
        int v = 0;
        check(wl_get_option(ctx_, option, &v), "wl_get_option");
        return v;
    }

    // The device holds one gauge field per lattice.  upload_gauge puts the links of U there;
    // bind_gauge does so only when the device does not already hold U as it is now (other
    // fields, or fields changed since their upload), so that an operator holding a reference
    // to U sees U's current links at every call, as HILA's operator reads its gauge fields.
    inline void upload_gauge(const MatrixField (&U)[4]);
    inline void bind_gauge(const MatrixField (&U)[4]);
    // record that the device holds U as it is now (after a device-side load of U's links)
    inline void note_bound(const MatrixField (&U)[4]);

  private:
    std::array<int, 4> size_;
    int64_t volume_ = 0;
    wl_ctx *ctx_ = nullptr;
    const MatrixField *bound_[4] = {nullptr, nullptr, nullptr, nullptr};
    uint64_t bound_gen_[4] = {0, 0, 0, 0};
};

// Field<SU<3,double>> (or Field<SquareMatrix<3,Complex<double>>>) of one direction on one
// lattice: V x 18 doubles on the host in HILA's AoS layout (EVEN_SITES_FIRST, row-major 3x3
// complex).  Every non-const access to the data counts as a change (a new generation), which
// Lattice::bind_gauge compares with the generation it uploaded.
class MatrixField {
  public:
    explicit MatrixField(Lattice &lat) : lat_(&lat), d_((size_t)18 * lat.volume()), gen_(next_generation()) {}
    MatrixField(const MatrixField &o) : lat_(o.lat_), d_(o.d_), gen_(next_generation()) {}
    MatrixField &operator=(const MatrixField &o) {
        lat_ = o.lat_;
        d_ = o.d_;
        gen_ = next_generation();
        return *this;
    }
    Lattice &lattice() const { return *lat_; }
    std::vector<double> &data() {
        gen_ = next_generation();
        return d_;
    }
    const std::vector<double> &data() const { return d_; }
    uint64_t generation() const { return gen_; }

  private:
    static uint64_t next_generation() {
        static std::atomic<uint64_t> n{1};
        return n++;
    }
    Lattice *lat_;
    std::vector<double> d_;
    uint64_t gen_;
};

inline void Lattice::note_bound(const MatrixField (&U)[4]) {
    for (int mu = 0; mu < 4; ++mu) {
        bound_[mu] = &U[mu];
        bound_gen_[mu] = U[mu].generation();
    }
}
inline void Lattice::upload_gauge(const MatrixField (&U)[4]) {
    for (int mu = 0; mu < 4; ++mu)
        if (&U[mu].lattice() != this) throw error(WL_E_ARG, "gauge field of another lattice");
    const double *p[4] = {U[0].data().data(), U[1].data().data(), U[2].data().data(), U[3].data().data()};
    check(wl_gauge_upload(ctx_, p), "wl_gauge_upload");
    note_bound(U);
}
inline void Lattice::bind_gauge(const MatrixField (&U)[4]) {
    for (int mu = 0; mu < 4; ++mu)
        if (bound_[mu] != &U[mu] || bound_gen_[mu] != U[mu].generation()) {
            upload_gauge(U);
            return;
        }
}

// SU(3) gauge field: HILA's GaugeField<SU<3,double>> with its direction fields fdir[4]
// (libraries/plumbing/gaugefield.h:23-71; hmc/gauge_field.h:21), uploaded to the GPU.
class GaugeField {
  public:
    MatrixField fdir[4];

    explicit GaugeField(Lattice &lat)
        : fdir{MatrixField(lat), MatrixField(lat), MatrixField(lat), MatrixField(lat)}, lat_(lat) {}
    std::vector<double> &direction(int mu) { return fdir[mu].data(); }
    const std::vector<double> &direction(int mu) const { return fdir[mu].data(); }
    // kind 0 unit, 1 Haar random, 2 exp(i eps H), 3 SU<3>::random() (16 SU(2) subgroup hits)
    void random(int kind, double eps, uint64_t seed) {
        double *p[4] = {direction(0).data(), direction(1).data(), direction(2).data(), direction(3).data()};
        check(wl_gen_gauge(lat_.size().data(), kind, eps, seed, p), "wl_gen_gauge");
        upload();
    }
    void upload() { lat_.upload_gauge(fdir); }
    // gauge_field_base::get_single_precision (hmc/gauge_field.h:83-90).  The library derives
    // its FP32 copy of the device links itself on first use by a single-precision operator, so
    // the single-precision gauge field is this one: Dirac_Wilson_evenodd<SU3f> built on it
    // applies in FP32 with the FP32 copy of these links.
    GaugeField &get_single_precision() { return *this; }
    // GaugeField::config_write / config_read (libraries/plumbing/gaugefield.h:194-272): HILA's
    // configuration file, read straight into the device gauge field (replicated when the file's
    // lattice divides this one); the host copy is refreshed from the device.
    void config_write(const std::string &filename) {
        lat_.bind_gauge(fdir);
        check(wl_config_write(lat_.ctx(), filename.c_str()), "wl_config_write");
    }
    void config_read(const std::string &filename) {
        check(wl_config_read(lat_.ctx(), filename.c_str()), "wl_config_read");
        double *p[4] = {direction(0).data(), direction(1).data(), direction(2).data(), direction(3).data()};
        check(wl_gauge_download(lat_.ctx(), p), "wl_gauge_download");
        lat_.note_bound(fdir);
    }
    Lattice &lattice() { return lat_; }

    // gauge_field_base::momentum (hmc/gauge_field.h:62), held on the GPU with the lattice:
    // one 3x3 complex matrix per link, V x 18 doubles per direction.
    using LinkField = std::array<std::vector<double>, 4>;
    void upload_momentum(const LinkField &P) {
        const double *p[4] = {P[0].data(), P[1].data(), P[2].data(), P[3].data()};
        check(wl_momentum_upload(lat_.ctx(), p), "wl_momentum_upload");
    }
    LinkField download_momentum() const {
        LinkField P;
        double *p[4];
        for (int mu = 0; mu < 4; ++mu) {
            P[mu].resize(18 * lat_.volume());
            p[mu] = P[mu].data();
        }
        check(wl_momentum_download(lat_.ctx(), p), "wl_momentum_download");
        return P;
    }

  private:
    Lattice &lat_;
};

template <typename T>
struct precision_of;
template <>
struct precision_of<double> {
    static constexpr int value = WL_F64;
};
template <>
struct precision_of<float> {
    static constexpr int value = WL_F32;
};

// Field<WilsonVector<3,T>> (T = double or float): both parities on the GPU, stored in T.
template <typename T = double>
class WilsonFieldT {
  public:
    using radix = T;
    static constexpr int precision = precision_of<T>::value;

    explicit WilsonFieldT(Lattice &lat) : lat_(&lat) {
        check(wl_field_create(lat.ctx(), WL_EVEN, precision, 1, &even_), "wl_field_create");
        check(wl_field_create(lat.ctx(), WL_ODD, precision, 1, &odd_), "wl_field_create");
    }
    ~WilsonFieldT() {
        wl_field_destroy(even_);
        wl_field_destroy(odd_);
    }
    WilsonFieldT(const WilsonFieldT &o) : WilsonFieldT(*o.lat_) { *this = o; }
    WilsonFieldT &operator=(const WilsonFieldT &o) {  // deep copy (field.h:898-909)
        if (this != &o) {
            check(wl_field_copy(even_, o.even_), "wl_field_copy");
            check(wl_field_copy(odd_, o.odd_), "wl_field_copy");
        }
        return *this;
    }
    // assignment from a field of the other precision, converted on the device (HILA's
    // c[ALL] = chi[X] between float and double fields, fermion_field.h:77-78,84)
    template <typename U>
    WilsonFieldT &operator=(const WilsonFieldT<U> &o) {
        check(wl_field_copy(even_, o.part(EVEN)), "wl_field_copy");
        check(wl_field_copy(odd_, o.part(ODD)), "wl_field_copy");
        return *this;
    }
    wl_field *part(Parity p) const { return p == ODD ? odd_ : even_; }
    Lattice &lattice() const { return *lat_; }

    // f[par] = 0
    void zero(Parity par = ALL) {
        if (par != ODD) check(wl_field_zero(even_), "wl_field_zero");
        if (par != EVEN) check(wl_field_zero(odd_), "wl_field_zero");
    }
    // onsites(par) f[X].gaussian_random(), from the seeded counter-based generator
    void gaussian_random(uint64_t seed, Parity par = ALL) {
        std::vector<double> h(24 * lat_->volume());
        check(wl_gen_spinor(lat_->size().data(), seed, (int)par, h.data()), "wl_gen_spinor");
        set_host(h);
    }
    // whole field in HILA's AoS layout (V x 24 doubles, EVEN_SITES_FIRST)
    void set_host(const std::vector<double> &h) {
        const int64_t half = 12 * lat_->volume();
        check(wl_field_upload(even_, 0, h.data()), "wl_field_upload");
        check(wl_field_upload(odd_, 0, h.data() + half), "wl_field_upload");
    }
    std::vector<double> host() const {
        std::vector<double> h(24 * lat_->volume());
        const int64_t half = 12 * lat_->volume();
        check(wl_field_download(even_, 0, h.data()), "wl_field_download");
        check(wl_field_download(odd_, 0, h.data() + half), "wl_field_download");
        return h;
    }
    // sum over par of squarenorm (conjugate_gradient.h:61)
    double squarenorm(Parity par = ALL) const {
        double s = 0, t = 0;
        if (par != ODD) check(wl_field_norm2(even_, 0, &s), "wl_field_norm2");
        if (par != EVEN) check(wl_field_norm2(odd_, 0, &t), "wl_field_norm2");
        return s + t;
    }

  private:
    Lattice *lat_;
    wl_field *even_ = nullptr, *odd_ = nullptr;
};
using WilsonField = WilsonFieldT<double>;

// SU<3,double> and SU<3,float> as the operator's matrix type
struct SU3d {
    static constexpr int size = 3;
    using radix = double;
};
struct SU3f {
    static constexpr int size = 3;
    using radix = float;
};

// Dirac_Wilson_evenodd<SU<3,T>> (wilson.h:234-312).  It holds a non-owning reference to its
// gauge fields, as HILA's does (wilson.h:237), and makes them the device's links at every
// call (Lattice::bind_gauge).  force(chi, psi, force, sign) is wilson.h:294-311.
template <typename matrix = SU3d>
class Dirac_Wilson_evenodd {
  public:
    double kappa;
    static constexpr int N = matrix::size;
    using radix = typename matrix::radix;
    using vector_type = WilsonFieldT<radix>;
    using matrix_type = matrix;
    // The single-precision operator that preconditions the inversion of this one
    // (wilson.h:250-253; fermion_action::initial_guess, fermion_field.h:73-84): it applies in
    // FP32, on FP32 fields, with the library's FP32 copy of the same device links.
    using type_flt = Dirac_Wilson_evenodd<SU3f>;
    Parity par = EVEN;
    MatrixField (&gauge)[4];

    Dirac_Wilson_evenodd(Dirac_Wilson_evenodd &d) : kappa(d.kappa), gauge(d.gauge) {}  // wilson.h:259
    Dirac_Wilson_evenodd(double k, MatrixField (&U)[4]) : kappa(k), gauge(U) {}         // :261
    Dirac_Wilson_evenodd(double k, GaugeField &g) : kappa(k), gauge(g.fdir) {}          // :263-264
    // from an operator of another matrix type, on the gauge field g (:267-269): e.g. the
    // single-precision operator of a double one, on g.get_single_precision()
    template <typename M>
    Dirac_Wilson_evenodd(Dirac_Wilson_evenodd<M> &d, GaugeField &g) : kappa(d.kappa), gauge(g.fdir) {}

    void apply(const vector_type &in, vector_type &out) { meo(in, out, +1); }
    void dagger(const vector_type &in, vector_type &out) { meo(in, out, -1); }
    // force(chi, psi, force, sign) (wilson.h:294-311): the derivative of Re <chi, M_sign psi>
    // with respect to the links, V x 18 doubles per direction
    void force(const vector_type &chi, const vector_type &psi, GaugeField::LinkField &force, int sign) {
        double *p[4];
        for (int mu = 0; mu < 4; ++mu) {
            force[mu].resize(18 * lattice().volume());
            p[mu] = force[mu].data();
        }
        check(wl_fermion_force(bind(), chi.part(EVEN), psi.part(EVEN), kappa, sign, p), "wl_fermion_force");
    }
    Lattice &lattice() const { return gauge[0].lattice(); }
    // the device context with this operator's gauge fields as its links
    wl_ctx *bind() {
        lattice().bind_gauge(gauge);
        return lattice().ctx();
    }
    double shift() const { return 0.0; }  // the operator is M itself

  private:
    void meo(const vector_type &in, vector_type &out, int sign) {
        check(wl_meo_apply(bind(), out.part(EVEN), in.part(EVEN), kappa, sign), "wl_meo_apply");
        out.zero(ODD);  // out[ODD] = 0 (wilson.h:278,288)
    }
};

// Hasenbusch_operator<Dirac_type> (libraries/dirac/Hasenbusch.h:15-53): D + mh, whose apply /
// dagger add h_parameter * in on D.par after D's own (here inside the same GPU pass).
template <typename Dirac_type>
class Hasenbusch_operator {
  public:
    using vector_type = typename Dirac_type::vector_type;
    using type_flt = Hasenbusch_operator<typename Dirac_type::type_flt>;  // Hasenbusch.h:19
    Dirac_type D;
    Parity par;
    double h_parameter;
    double kappa;

    Hasenbusch_operator(Dirac_type &d, double hasenbusch_parameter)
        : D(d), par(d.par), h_parameter(hasenbusch_parameter), kappa(d.kappa) {}
    Hasenbusch_operator(Hasenbusch_operator &h)
        : D(h.D), par(h.par), h_parameter(h.h_parameter), kappa(h.kappa) {}
    // from a Hasenbusch operator of another Dirac type, on the gauge field mat
    // (Hasenbusch.h:31-35): e.g. type_flt from the double operator
    template <typename H_type, typename matrix>
    Hasenbusch_operator(H_type &h, matrix &mat)
        : D(h.D, mat), par(h.par), h_parameter(h.h_parameter), kappa(h.kappa) {}

    void apply(const vector_type &in, vector_type &out) { op(in, out, +1); }
    void dagger(const vector_type &in, vector_type &out) { op(in, out, -1); }
    void force(const vector_type &chi, const vector_type &psi, GaugeField::LinkField &f, int sign) {
        D.force(chi, psi, f, sign);  // Hasenbusch.h:50-53
    }
    Lattice &lattice() const { return D.lattice(); }
    wl_ctx *bind() { return D.bind(); }
    double shift() const { return h_parameter; }

  private:
    void op(const vector_type &in, vector_type &out, int sign) {
        check(wl_hasenbusch_apply(bind(), out.part(EVEN), in.part(EVEN), kappa, h_parameter, sign),
              "wl_hasenbusch_apply");
        out.zero(ODD);
    }
};

// MRE_guess(psi, chi, D, old_chi_inv) (libraries/hmc/MRE_guess.h:8-87): psi = the
// minimal-residual combination of the previous solutions old_chi_inv (at most 8) for
// D^dag D psi = chi, computed on the GPU (wl_mre_guess); psi's odd half is set to 0.
template <typename DIRAC_OP>
void MRE_guess(WilsonField &psi, const WilsonField &chi, DIRAC_OP &D,
               const std::vector<WilsonField> &old_chi_inv) {
    wl_ctx *c = D.bind();
    const int n = (int)old_chi_inv.size();
    wl_field *basis = nullptr;
    check(wl_field_create(c, WL_EVEN, WL_F64, n > 0 ? n : 1, &basis), "wl_field_create");
    int rc = WL_OK;
    for (int i = 0; i < n && rc == WL_OK; ++i) rc = wl_field_copy_rhs(basis, i, old_chi_inv[i].part(EVEN), 0);
    if (rc == WL_OK) rc = wl_mre_guess(c, psi.part(EVEN), chi.part(EVEN), basis, n, D.kappa, D.shift());
    wl_field_destroy(basis);
    check(rc, "MRE_guess");
    psi.zero(ODD);
}

// CG<Op> (conjugate_gradient.h:19-105): solves (M^dag M) out = in on M.par.
template <typename Op>
class CG {
  public:
    using vector_type = typename Op::vector_type;
    explicit CG(Op &op) : M(op) {}
    CG(Op &op, double accuracy_) : M(op), accuracy(accuracy_) {}
    CG(Op &op, double accuracy_, int maxiters_) : M(op), accuracy(accuracy_), maxiters(maxiters_) {}

    // out: initial guess on entry, solution on exit.  Reaching maxiters is not an error (as in
    // HILA); the statistics of the last call are kept in `stats`.
    //
    // With a single-precision operator (Op::type_flt of a double one) the inner operator and
    // the Krylov vectors are FP32 and the scalars FP64, as HILA's CG on float fields
    // (conjugate_gradient.h:55-57); the solve runs with FP64 reliable updates on FP64 copies
    // of in and out, and out receives the solution rounded to FP32.
    void apply(vector_type &in, vector_type &out) {
        wl_ctx *c = M.bind();
        wl_cg_params p{};
        p.kappa = M.kappa;
        p.mh = M.shift();  // Hasenbusch_operator: (M + mh)^dag (M + mh)
        p.accuracy = accuracy;
        p.maxiters = (int)maxiters;
        p.inner_prec = inner_prec_;
        p.delta = delta_;
        p.update_every = update_every_;
        int rc;
        if constexpr (vector_type::precision == WL_F64) {
            rc = wl_cg_solve_ex(c, out.part(EVEN), in.part(EVEN), &p, &stats);
        } else {
            WilsonField in64(M.lattice()), out64(M.lattice());
            in64 = in;
            out64 = out;
            p.inner_prec = vector_type::precision;
            rc = wl_cg_solve_ex(c, out64.part(EVEN), in64.part(EVEN), &p, &stats);
            if (rc == WL_OK || rc == WL_E_NOCONV) out = out64;
        }
        if (rc != WL_E_NOCONV) check(rc, "wl_cg_solve_ex");
    }
    // Mixed precision (not in HILA's CG): inner operator and Krylov vectors in WL_F32 or WL_F16
    // (with per-site scale), FP64 reliable updates when ||r|| drops by delta or every
    // update_every iterations, convergence on the FP64 true residual.  WL_F64 restores HILA's CG.
    void set_precision(int inner_prec, double delta = 1e-3, int update_every = 100) {
        inner_prec_ = inner_prec;
        delta_ = delta;
        update_every_ = update_every;
    }
    wl_cg_stats stats{};

  private:
    Op &M;
    double accuracy = 1e-12;   // CG_DEFAULT_ACCURACY (conjugate_gradient.h:16)
    double maxiters = 10000;   // CG_DEFAULT_MAXITERS (conjugate_gradient.h:15)
    int inner_prec_ = WL_F64;
    double delta_ = 1e-3;
    int update_every_ = 100;
};

// fermion_action<gauge_field, DIRAC_OP> (libraries/hmc/fermion_field.h:20-168): the
// pseudofermion chi of one HMC action term with its MRE basis of previous solutions.
template <typename DIRAC_OP>
class fermion_action {
  public:
    GaugeField &gauge;
    DIRAC_OP &D;
    WilsonField chi;
    int MRE_size = 0;
    std::vector<WilsonField> old_chi_inv;  // newest first
    wl_cg_stats stats{};

    fermion_action(DIRAC_OP &d, GaugeField &g, int mre_guess_size = 0)
        : gauge(g), D(d), chi(g.lattice()), MRE_size(mre_guess_size) {
        chi.zero(ALL);
    }
    // CG settings of the solves (HILA: CG_DEFAULT_ACCURACY, FP64)
    void set_solver(double accuracy, int inner_prec = WL_F64, double delta = 0.1, int update_every = 100) {
        accuracy_ = accuracy;
        inner_prec_ = inner_prec;
        delta_ = delta;
        update_every_ = update_every;
    }

    // initial_guess (:62-86): the MRE guess from the saved solutions; then, when
    // set_single_precision_guess(true), HILA's double-precision path (:73-84): a CG solve of
    // the single-precision operator DIRAC_OP::type_flt, built by its converting constructor on
    // gauge.get_single_precision(), from the guess, whose result becomes psi.  (By default the
    // FP32 pre-solve is left to the reliable-update solver, set_solver.)
    void initial_guess(const WilsonField &c, WilsonField &psi) {
        psi.zero(ALL);
        if (MRE_size > 0 && !old_chi_inv.empty()) MRE_guess(psi, c, D, old_chi_inv);
        if (single_precision_guess_) {
            using D_flt_type = typename DIRAC_OP::type_flt;
            D_flt_type D_flt(D, gauge.get_single_precision());
            typename D_flt_type::vector_type cf(gauge.lattice()), pf(gauge.lattice());
            cf = c;
            pf = psi;
            CG<D_flt_type> inverse(D_flt);
            inverse.apply(cf, pf);
            psi = pf;
        }
    }
    void set_single_precision_guess(bool on) { single_precision_guess_ = on; }
    // save_new_solution (:134-142): newest first, the oldest dropped
    void save_new_solution(const WilsonField &psi) {
        if (MRE_size <= 0) return;
        old_chi_inv.insert(old_chi_inv.begin(), psi);
        if ((int)old_chi_inv.size() > MRE_size) old_chi_inv.pop_back();
    }
    // draw_gaussian_fields (:119-132): psi Gaussian on D.par, chi = D^dag psi
    void draw_gaussian_fields(uint64_t seed) {
        WilsonField psi(gauge.lattice());
        psi.gaussian_random(seed, EVEN);
        D.dagger(psi, chi);
    }
    // action (:88-103): Re <chi, (D^dag D)^-1 chi> on D.par
    double action() {
        WilsonField psi(gauge.lattice());
        initial_guess(chi, psi);
        solve(psi);
        const std::vector<double> a = chi.host(), b = psi.host();
        double s = 0;
        for (size_t i = 0; i < a.size() / 2; ++i) s += a[i] * b[i];  // even sites come first
        return s;
    }
    // force_step (:145-167): psi = (D^dag D)^-1 chi, momentum += P(U (-eps)(F + F2)), one
    // GPU call after the initial guess
    void force_step(double eps) {
        WilsonField psi(gauge.lattice());
        initial_guess(chi, psi);
        const wl_cg_params p = params();
        const int rc = wl_fermion_force_step(D.bind(), chi.part(EVEN), psi.part(EVEN), eps, &p, &stats);
        if (rc != WL_E_NOCONV) check(rc, "wl_fermion_force_step");
        save_new_solution(psi);
    }

  private:
    wl_cg_params params() const {
        wl_cg_params p{};
        p.kappa = D.kappa;
        p.mh = D.shift();
        p.accuracy = accuracy_;
        p.maxiters = 10000;
        p.inner_prec = inner_prec_;
        p.delta = delta_;
        p.update_every = update_every_;
        return p;
    }
    void solve(WilsonField &psi) {
        const wl_cg_params p = params();
        const int rc = wl_cg_solve_ex(D.bind(), psi.part(EVEN), chi.part(EVEN), &p, &stats);
        if (rc != WL_E_NOCONV) check(rc, "wl_cg_solve_ex");
    }
    double accuracy_ = 1e-12;
    int inner_prec_ = WL_F64;
    double delta_ = 0.1;
    int update_every_ = 100;
    bool single_precision_guess_ = false;
};

// Hasenbusch_action_1 (libraries/hmc/fermion_field.h:168-196): the fermion action of
// D_h = D + mh.
template <typename DIRAC_OP>
class Hasenbusch_action_1 {
  public:
    Hasenbusch_operator<DIRAC_OP> D_h;
    fermion_action<Hasenbusch_operator<DIRAC_OP>> base_action;

    Hasenbusch_action_1(DIRAC_OP &d, GaugeField &g, double mh, int mre_guess_size = 0)
        : D_h(d, mh), base_action(D_h, g, mre_guess_size) {}
    double action() { return base_action.action(); }
    void draw_gaussian_fields(uint64_t seed) { base_action.draw_gaussian_fields(seed); }
    void force_step(double eps) { base_action.force_step(eps); }
};

// Hasenbusch_action_2 (libraries/hmc/fermion_field.h:198-368): the second term of the split,
// action <chi, D_h (D^dag D)^-1 D_h^dag chi>.
template <typename DIRAC_OP>
class Hasenbusch_action_2 {
  public:
    GaugeField &gauge;
    DIRAC_OP D;
    Hasenbusch_operator<DIRAC_OP> D_h;
    double mh;
    WilsonField chi;
    int MRE_size = 0;
    std::vector<WilsonField> old_chi_inv;  // newest first
    wl_cg_stats stats{};

    Hasenbusch_action_2(DIRAC_OP &d, GaugeField &g, double _mh, int mre_guess_size = 0)
        : gauge(g), D(d), D_h(d, _mh), mh(_mh), chi(g.lattice()), MRE_size(mre_guess_size) {
        chi.zero(ALL);
    }

    // action (:255-272): |D (D^dag D)^-1 D_h^dag chi|^2 on D.par
    double action() {
        WilsonField psi(gauge.lattice()), v(gauge.lattice());
        D_h.dagger(chi, psi);
        v.zero(ALL);
        CG<DIRAC_OP> inverse(D, accuracy_);
        inverse.apply(psi, v);
        D.apply(v, psi);
        stats = inverse.stats;
        return psi.squarenorm(EVEN);
    }
    // draw_gaussian_fields (:274-291): chi = D_h (D_h^dag D_h)^-1 D^dag psi, psi Gaussian on D.par
    void draw_gaussian_fields(uint64_t seed) {
        WilsonField psi(gauge.lattice()), v(gauge.lattice());
        psi.gaussian_random(seed, EVEN);
        D.dagger(psi, v);
        psi.zero(ALL);
        CG<Hasenbusch_operator<DIRAC_OP>> inverse_h(D_h, accuracy_);
        inverse_h.apply(v, psi);
        D_h.apply(psi, chi);
    }
    // initial_guess (:293-317): the MRE guess for (D^dag D) psi = c
    void initial_guess(const WilsonField &c, WilsonField &psi) {
        psi.zero(ALL);
        if (MRE_size > 0 && !old_chi_inv.empty()) MRE_guess(psi, c, D, old_chi_inv);
    }
    void save_new_solution(const WilsonField &psi) {
        if (MRE_size <= 0) return;
        old_chi_inv.insert(old_chi_inv.begin(), psi);
        if ((int)old_chi_inv.size() > MRE_size) old_chi_inv.pop_back();
    }
    // force_step (:319-346), one GPU call after the initial guess
    void force_step(double eps) {
        WilsonField Dhchi(gauge.lattice()), psi(gauge.lattice());
        D_h.dagger(chi, Dhchi);
        initial_guess(Dhchi, psi);
        wl_cg_params p{};
        p.kappa = D.kappa;
        p.mh = mh;
        p.accuracy = accuracy_;
        p.maxiters = 10000;
        p.inner_prec = WL_F64;
        const int rc = wl_hasenbusch2_force_step(D.bind(), chi.part(EVEN), psi.part(EVEN), eps, &p, &stats);
        if (rc != WL_E_NOCONV) check(rc, "wl_hasenbusch2_force_step");
        save_new_solution(psi);
    }
    void set_accuracy(double a) { accuracy_ = a; }

  private:
    double accuracy_ = 1e-12;
};

// Request server (wl_server_*): independent solves admitted by HBM budget and run
// concurrently on one GPU.  Host buffers of a request must outlive its wait().
class Server {
  public:
    explicit Server(int device = 0, size_t hbm_budget_bytes = 0, int max_active = 0) {
        check(wl_server_create(device, hbm_budget_bytes, max_active, &srv_), "wl_server_create");
    }
    ~Server() { wl_server_destroy(srv_); }
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;
    uint64_t submit(const wl_solve_request &r) {
        uint64_t t = 0;
        check(wl_server_submit(srv_, &r, &t), "wl_server_submit");
        return t;
    }
    // Idle slots, with their device memory, for up to nrhs-rhs requests of one lattice and inner
    // precision, made before traffic; returns how many the HBM budget allowed.
    int provision(const std::array<int, 4> &dims, int inner_prec, int nrhs, int count, bool antiperiodic_t = true) {
        int made = 0;
        check(wl_server_provision(srv_, dims.data(), antiperiodic_t ? 1 : 0, inner_prec, nrhs, count, &made),
              "wl_server_provision");
        return made;
    }
    // Blocks until the request finishes; per_rhs (if given) receives r.nrhs entries.
    wl_solve_result wait(uint64_t ticket, wl_cg_stats *per_rhs = nullptr) {
        wl_solve_result res{};
        check(wl_server_wait(srv_, ticket, -1, &res, per_rhs), "wl_server_wait");
        return res;
    }

  private:
    wl_server *srv_ = nullptr;
};

}  // namespace hila_wilson
