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
