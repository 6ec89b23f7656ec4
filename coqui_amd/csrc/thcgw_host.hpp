// Host-side parts of libthcgw: IR tables, Ec assembly, quasiparticle solver.
#pragma once

#include <complex>
#include <vector>

#include "../../include/thcgw.h"

namespace thcgw {
namespace host {

using cd = std::complex<double>;

// IR tables rescaled with beta as in src/numerics/imag_axes_ft/ir/ir_driver.hpp:68-73.
struct Tables {
  int nt = 0, nw_b = 0, nw_f = 0;
  double beta = 0.0;
  std::vector<double> x;    // tau mesh in [-1, 1]
  std::vector<double> tau;  // (x + 1) beta / 2   (qp_scf_common.cpp:96-98)
  std::vector<long> wn_f;   // fermionic Matsubara indices
  std::vector<cd> Ttw_ff, Twt_ff, Ttw_bb, Twt_bb, T_beta;
};

// Copies and rescales the tables; checks that the tau mesh is symmetric about beta/2
// (scr_coulomb_t.cpp:26-34, thc_gw.cpp:58-67).
void load_tables(const thcgw_dims& d, const thcgw_ir_tables& t, Tables& out);

// Particle-hole folded bosonic transforms: Tp (nw_half x nt_half), IAFT.icc:62-70, and
// Tm (nt_half x nw_half), IAFT.icc:117-125.
void fold_ph(const Tables& tab, const thcgw_dims& d, std::vector<cd>& Tp, std::vector<cd>& Tm);

// Ec = Re[T_beta_t_ff . (Ttw_bb e_w)] / (2 nq)   (thc_rpa.icc:171-173)
double ec_from_ew(const Tables& tab, const std::vector<cd>& ew, int nq);

// GM energy from SigmaG(iw_n) = sum_k w_k Tr[Sigma(iw_n) G(iw_n)] (nw_f values), scf_common.cpp:119-128.
double gm_from_sigmag(const Tables& tab, const std::vector<cd>& sg);

// Head of the inverse dielectric function at q -> 0 from eps (nt_half x nq, tau-major)
// (g0_div_utils.hpp:55-111 eps_inv_head_t, 264-322 find_smallest_qabs / extrapolate_to_q0):
// mode 0 (ignore_g0) and 1 (gygi): the column of the smallest nonzero |q|; 2 (gygi_extrplt)
// and 3 (gygi_extrplt_2d, q_z = 0 only): per bosonic frequency a least-squares polynomial of
// order fit_order in |q| through the q with 0 < |q| <= q_max (distinct |q|), evaluated at
// 0 and brought back to tau; with fewer than 3 such q, the first of them.
std::vector<cd> eps_inv_q0(const Tables& tab, const thcgw_dims& d, const std::vector<cd>& eps, const double* qcart,
                           int nq, int mode, double q_max, int fit_order);

// Pade frequency selection of pade_driver::init (pade_driver.hpp:38-60) for a symmetric mesh.
std::vector<int> pade_fit_indices(int niw, int nfit);

}  // namespace host
}  // namespace thcgw
