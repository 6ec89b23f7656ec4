// thcgw.hpp - C++ host interface over the C ABI of libthcgw.
//
// Mirrors the reference's GW/RPA solver objects for the THC path so that a driver written
// against them (scf_loop / qp_scf_loop / rpa_loop) can call the MI355X library instead:
//
//   scr_coulomb_t::update_w(MBState&, THC&, long)   src/methods/scr_coulomb/scr_coulomb_t.h:82
//   gw_t::evaluate(MBState&, THC const&)            src/methods/GW/gw_t.h:55
//   gw_t::rpa_energy(G_tskij, THC&) -> double       src/methods/GW/gw_t.h:75
//
// MBState and THCView carry the device arrays those routines read and write (the fields
// of mb_state.hpp:79-84 and the accessors of thc_reader_t.hpp:590-829 used on the path).
// Header-only; link with libthcgw.so.  Errors throw thcgw_cxx::error.
#ifndef THCGW_HPP
#define THCGW_HPP

#include <hip/hip_runtime.h>

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "thcgw.h"

namespace thcgw_cxx {

using cplx = std::complex<double>;

class error : public std::runtime_error {
 public:
  int code;
  error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(const thcgw_ctx* ctx, int rc) {
  if (rc != THCGW_OK) throw error(rc, thcgw_last_error(ctx));
}

inline double* D(cplx* p) { return reinterpret_cast<double*>(p); }
inline const double* D(const cplx* p) { return reinterpret_cast<const double*>(p); }

// The MBState fields of the path (mb_state.hpp:79-84), device arrays; eps_inv_head on the host.
struct MBState {
  cplx* G_tskij = nullptr;      // sG_tskij     (nt, 1, nk, nbnd, nbnd)
  cplx* Sigma_tskij = nullptr;  // sSigma_tskij (nt, 1, nk, nbnd, nbnd)
  cplx* W_qtPQ = nullptr;       // dW_qtPQ      (nk, nt_half, Np, Np), W - V
  std::vector<cplx> eps_inv_head;  // eps_inv_head (nt_half), set by update_w unless ignore_g0
};

// The thc_reader_t and MF accessors of the path (thc_reader_t.hpp:590-604, 714-751, 803;
// MF.hpp:152-259).  The last five are read only for div_treatment != ignore_g0.
struct THCView {
  const cplx* X_skPa = nullptr;  // X(is=0, ip=0, ik) stacked over k: (1, nk, Np, nbnd)
  const cplx* Z_qPQ = nullptr;   // Z(iq) stacked over q: (nk, Np, Np), N_k factor included
  int Np = 0;
  const cplx* chi_head_P = nullptr;  // basis_head()(0, :), device (Np)
  const cplx* chi_bar_qP = nullptr;  // basis_bar_head(), device (nk, Np)
  const double* qpts_cart = nullptr; // MF::Qpts(), host (nk, 3), Cartesian
  double volume = 0.0;               // MF::volume()
  double madelung = 0.0;             // MF::madelung()
};

// div_treatment_e of the path (src/methods/ERI/div_treatment_e.hpp:7-8), thcgw numbering.
enum div_treatment { ignore_g0 = 0, gygi = 1, gygi_extrplt = 2, gygi_extrplt_2d = 3 };

// One system shape: the IAFT tables and the dimensions (thcgw_create / thcgw_destroy).
class context {
 public:
  context(const thcgw_dims& d, const thcgw_ir_tables& t) : dims_(d) { check(nullptr, thcgw_create(&d, &t, &ctx_)); }
  ~context() { thcgw_destroy(ctx_); }
  context(const context&) = delete;
  context& operator=(const context&) = delete;
  thcgw_ctx* get() const { return ctx_; }
  // thcgw_set_flag (e.g. 6: Dyson form, 7: precision, 9: time-reversal input)
  void set_flag(int flag, long value) { check(ctx_, thcgw_set_flag(ctx_, flag, value)); }
  // flag 6: LU with partial pivoting (the reference's algorithm), the two-Cholesky
  // Hermitian form, or the resolvent form (V^-1 - Pi)^-1
  enum dyson_form { dyson_lu = 0, dyson_hermitian = 1, dyson_resolvent = 2 };
  void set_dyson_form(dyson_form f) { set_flag(6, f); }
  const thcgw_dims& dims() const { return dims_; }
  int nk() const { return dims_.kgrid[0] * dims_.kgrid[1] * dims_.kgrid[2]; }
  int nt_half() const { return (dims_.nt + 1) / 2; }
  int nw_half() const { return (dims_.nw_b + 1) / 2; }

 private:
  thcgw_dims dims_;
  thcgw_ctx* ctx_ = nullptr;
};

// Device scratch owned by the solver objects (the reference allocates per call).
class scratch {
 public:
  ~scratch() {
    if (p_) (void)hipFree(p_);
  }
  cplx* get(size_t count) {
    if (count > n_) {
      if (p_) (void)hipFree(p_);
      p_ = nullptr;
      n_ = 0;
      if (hipMalloc(&p_, count * sizeof(cplx)) != hipSuccess) throw error(THCGW_ERR_ALLOC, "scratch hipMalloc failed");
      n_ = count;
    }
    return static_cast<cplx*>(p_);
  }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// scr_coulomb_t with screen_type = "rpa".
class scr_coulomb_t {
 public:
  explicit scr_coulomb_t(context& c, hipStream_t s = nullptr, div_treatment div = ignore_g0)
      : c_(c), s_(s), div_(div) {}

  // update_w (scr_coulomb_t.cpp:37-95): chi0 from mb.G_tskij, tau -> iW, Dyson, iW -> tau,
  // result (W - V) in mb.W_qtPQ (also the scratch for Pi(tau)); then, unless ignore_g0, the
  // head of the inverse dielectric function into mb.eps_inv_head (scr_coulomb_t.cpp:63-66).
  void update_w(MBState& mb, const THCView& thc, long /*h5_iter*/ = -1) {
    require(thc);
    const size_t M = size_t(c_.dims().nmu) * c_.dims().nmu;
    cplx* Piw = piw_.get(size_t(c_.nw_half()) * c_.nk() * M);
    check(c_.get(), thcgw_pi(c_.get(), D(mb.G_tskij), D(thc.X_skPa), D(mb.W_qtPQ), s_));
    check(c_.get(), thcgw_tau_to_w(c_.get(), D(mb.W_qtPQ), D(Piw), s_));
    check(c_.get(), thcgw_dyson_w(c_.get(), D(Piw), D(thc.Z_qPQ), nullptr, nullptr, s_));
    check(c_.get(), thcgw_w_to_tau(c_.get(), D(Piw), D(mb.W_qtPQ), s_));
    if (div_ != ignore_g0) {
      mb.eps_inv_head.assign(size_t(c_.nt_half()), cplx(0.0));
      check(c_.get(), thcgw_eps_inv_head(c_.get(), D(mb.W_qtPQ), D(thc.chi_bar_qP), thc.qpts_cart, thc.volume, div_,
                                         nullptr, D(mb.eps_inv_head.data()), s_));
    }
  }

 private:
  void require(const THCView& thc) const {
    if (thc.Np != c_.dims().nmu) throw error(THCGW_ERR_ARG, "scr_coulomb_t: THC Np does not match the context");
  }
  context& c_;
  hipStream_t s_;
  div_treatment div_;
  scratch piw_;
};

class gw_t {
 public:
  explicit gw_t(context& c, hipStream_t s = nullptr, div_treatment div = ignore_g0) : c_(c), s_(s), div_(div) {}

  // evaluate (thc_gw.cpp:26-85): Sigma_tskij from G_tskij and W_qtPQ (R-space, both passes),
  // then Sigma_div_correction (thc_gw.icc:446-526) with mb.eps_inv_head unless ignore_g0.
  // Like the reference, it does not hermitize; call hermitize() for that.
  void evaluate(MBState& mb, const THCView& thc) {
    check(c_.get(), thcgw_sigma(c_.get(), D(mb.G_tskij), D(mb.W_qtPQ), D(thc.X_skPa), D(mb.Sigma_tskij), 0, s_));
    if (div_ != ignore_g0) {
      if (mb.eps_inv_head.size() != size_t(c_.nt_half()))
        throw error(THCGW_ERR_ARG, "gw_t::evaluate: eps_inv_head is not set (run update_w with the same div_treatment)");
      check(c_.get(), thcgw_sigma_div_correction(c_.get(), D(mb.Sigma_tskij), D(mb.G_tskij), D(thc.X_skPa),
                                                 D(thc.chi_head_P), thc.madelung, D(mb.eps_inv_head.data()), div_, s_));
    }
  }

  // hermitize(Sigma) as called by scf_loop (scf_common.hpp:70-86, scf_driver.cpp:128).
  void hermitize(MBState& mb) {
    const thcgw_dims& d = c_.dims();
    check(c_.get(), thcgw_hermitize(c_.get(), long(d.nt) * c_.nk(), d.no, D(mb.Sigma_tskij), s_));
  }

  // rpa_energy (thc_rpa.cpp:22-61): chi0 -> tau_to_w -> Tr(Pi V) + ln Re det(I - Pi V).
  double rpa_energy(const cplx* G_tskij, const THCView& thc) {
    const size_t M = size_t(c_.dims().nmu) * c_.dims().nmu;
    cplx* Pit = pit_.get(size_t(c_.nt_half()) * c_.nk() * M);
    cplx* Piw = piw_.get(size_t(c_.nw_half()) * c_.nk() * M);
    check(c_.get(), thcgw_pi(c_.get(), D(G_tskij), D(thc.X_skPa), D(Pit), s_));
    check(c_.get(), thcgw_tau_to_w(c_.get(), D(Pit), D(Piw), s_));
    double ec = 0.0;
    check(c_.get(), thcgw_rpa_ec(c_.get(), D(Piw), D(thc.Z_qPQ), &ec, s_));
    return ec;
  }

 private:
  context& c_;
  hipStream_t s_;
  div_treatment div_;
  scratch pit_, piw_;
};

}  // namespace thcgw_cxx

#endif  // THCGW_HPP
