// hila_wilson.hpp -- C++ mirror of HILA's Wilson even-odd operator and CG call surface over the
// C ABI of libhila_wilson.so (wilson_mi355x.h).  Header-only.
//
//   Dirac_Wilson_evenodd<SU3d> D(kappa, U);     libraries/dirac/wilson.h:234-269
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

  private:
    std::array<int, 4> size_;
    int64_t volume_ = 0;
    wl_ctx *ctx_ = nullptr;
};

// SU(3) gauge field: HILA's Field<SU<3,double>> gauge[4] (hmc/gauge_field.h:21), uploaded to
// the GPU once; host layout V x 18 doubles per direction.
class GaugeField {
  public:
    explicit GaugeField(Lattice &lat) : lat_(lat), U_(4, std::vector<double>(18 * lat.volume())) {}
    std::vector<double> &direction(int mu) { return U_.at(mu); }
    // kind 0 unit, 1 Haar random, 2 exp(i eps H)
    void random(int kind, double eps, uint64_t seed) {
        double *p[4] = {U_[0].data(), U_[1].data(), U_[2].data(), U_[3].data()};
        check(wl_gen_gauge(lat_.size().data(), kind, eps, seed, p), "wl_gen_gauge");
        upload();
    }
    void upload() {
        const double *p[4] = {U_[0].data(), U_[1].data(), U_[2].data(), U_[3].data()};
        check(wl_gauge_upload(lat_.ctx(), p), "wl_gauge_upload");
    }
    // GaugeField::config_write / config_read (libraries/plumbing/gaugefield.h:194-272): HILA's
    // configuration file, read straight into the device gauge field (replicated when the file's
    // lattice divides this one); the host copy is refreshed from the device.
    void config_write(const std::string &filename) const {
        check(wl_config_write(lat_.ctx(), filename.c_str()), "wl_config_write");
    }
    void config_read(const std::string &filename) {
        check(wl_config_read(lat_.ctx(), filename.c_str()), "wl_config_read");
        double *p[4] = {U_[0].data(), U_[1].data(), U_[2].data(), U_[3].data()};
        check(wl_gauge_download(lat_.ctx(), p), "wl_gauge_download");
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
    std::vector<std::vector<double>> U_;
};

// Field<WilsonVector<3,double>>: both parities on the GPU.
class WilsonField {
  public:
    explicit WilsonField(Lattice &lat) : lat_(&lat) {
        check(wl_field_create(lat.ctx(), WL_EVEN, WL_F64, 1, &even_), "wl_field_create");
        check(wl_field_create(lat.ctx(), WL_ODD, WL_F64, 1, &odd_), "wl_field_create");
    }
    ~WilsonField() {
        wl_field_destroy(even_);
        wl_field_destroy(odd_);
    }
    WilsonField(const WilsonField &o) : WilsonField(*o.lat_) { *this = o; }
    WilsonField &operator=(const WilsonField &o) {  // deep copy (field.h:898-909)
        if (this != &o) {
            check(wl_field_copy(even_, o.even_), "wl_field_copy");
            check(wl_field_copy(odd_, o.odd_), "wl_field_copy");
        }
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

struct SU3d {
    static constexpr int size = 3;
};

// Dirac_Wilson_evenodd<SU<3,double>> (wilson.h:234-312); force() is not provided.
template <typename matrix = SU3d>
class Dirac_Wilson_evenodd {
  public:
    double kappa;
    static constexpr int N = matrix::size;
    using vector_type = WilsonField;
    using matrix_type = matrix;
    Parity par = EVEN;

    Dirac_Wilson_evenodd(double k, GaugeField &U) : kappa(k), gauge_(U) {}
    Dirac_Wilson_evenodd(Dirac_Wilson_evenodd &d) : kappa(d.kappa), gauge_(d.gauge_) {}

    void apply(const WilsonField &in, WilsonField &out) { meo(in, out, +1); }
    void dagger(const WilsonField &in, WilsonField &out) { meo(in, out, -1); }
    // force(chi, psi, force, sign) (wilson.h:294-311): the derivative of Re <chi, M_sign psi>
    // with respect to the links, V x 18 doubles per direction
    void force(const WilsonField &chi, const WilsonField &psi, GaugeField::LinkField &force, int sign) {
        double *p[4];
        for (int mu = 0; mu < 4; ++mu) {
            force[mu].resize(18 * gauge_.lattice().volume());
            p[mu] = force[mu].data();
        }
        check(wl_fermion_force(gauge_.lattice().ctx(), chi.part(EVEN), psi.part(EVEN), kappa, sign, p),
              "wl_fermion_force");
    }
    GaugeField &gauge() { return gauge_; }
    double shift() const { return 0.0; }  // the operator is M itself

  private:
    void meo(const WilsonField &in, WilsonField &out, int sign) {
        check(wl_meo_apply(gauge_.lattice().ctx(), out.part(EVEN), in.part(EVEN), kappa, sign),
              "wl_meo_apply");
        out.zero(ODD);  // out[ODD] = 0 (wilson.h:278,288)
    }
    GaugeField &gauge_;
};

// Hasenbusch_operator<Dirac_type> (libraries/dirac/Hasenbusch.h:15-53): D + mh, whose apply /
// dagger add h_parameter * in on D.par after D's own (here inside the same GPU pass).
template <typename Dirac_type>
class Hasenbusch_operator {
  public:
    using vector_type = typename Dirac_type::vector_type;
    Dirac_type D;
    Parity par;
    double h_parameter;
    double kappa;

    Hasenbusch_operator(Dirac_type &d, double hasenbusch_parameter)
        : D(d), par(d.par), h_parameter(hasenbusch_parameter), kappa(d.kappa) {}
    Hasenbusch_operator(Hasenbusch_operator &h)
        : D(h.D), par(h.par), h_parameter(h.h_parameter), kappa(h.kappa) {}

    void apply(const vector_type &in, vector_type &out) { op(in, out, +1); }
    void dagger(const vector_type &in, vector_type &out) { op(in, out, -1); }
    void force(const vector_type &chi, const vector_type &psi, GaugeField::LinkField &f, int sign) {
        D.force(chi, psi, f, sign);  // Hasenbusch.h:50-53
    }
    GaugeField &gauge() { return D.gauge(); }
    double shift() const { return h_parameter; }

  private:
    void op(const vector_type &in, vector_type &out, int sign) {
        check(wl_hasenbusch_apply(gauge().lattice().ctx(), out.part(EVEN), in.part(EVEN), kappa,
                                  h_parameter, sign),
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
    wl_ctx *c = D.gauge().lattice().ctx();
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
    void apply(vector_type &in, vector_type &out) {
        wl_ctx *c = M.gauge().lattice().ctx();
        wl_cg_params p{};
        p.kappa = M.kappa;
        p.mh = M.shift();  // Hasenbusch_operator: (M + mh)^dag (M + mh)
        p.accuracy = accuracy;
        p.maxiters = (int)maxiters;
        p.inner_prec = inner_prec_;
        p.delta = delta_;
        p.update_every = update_every_;
        const int rc = wl_cg_solve_ex(c, out.part(EVEN), in.part(EVEN), &p, &stats);
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

    // initial_guess (:62-86): the MRE guess from the saved solutions (the FP32 pre-solve of
    // HILA's double-precision builds is replaced by the reliable-update solver, set_solver)
    void initial_guess(const WilsonField &c, WilsonField &psi) {
        psi.zero(ALL);
        if (MRE_size > 0 && !old_chi_inv.empty()) MRE_guess(psi, c, D, old_chi_inv);
    }
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
        const int rc = wl_fermion_force_step(gauge.lattice().ctx(), chi.part(EVEN), psi.part(EVEN), eps, &p, &stats);
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
        const int rc = wl_cg_solve_ex(gauge.lattice().ctx(), psi.part(EVEN), chi.part(EVEN), &p, &stats);
        if (rc != WL_E_NOCONV) check(rc, "wl_cg_solve_ex");
    }
    double accuracy_ = 1e-12;
    int inner_prec_ = WL_F64;
    double delta_ = 0.1;
    int update_every_ = 100;
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
        const int rc = wl_hasenbusch2_force_step(gauge.lattice().ctx(), chi.part(EVEN), psi.part(EVEN), eps, &p, &stats);
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
