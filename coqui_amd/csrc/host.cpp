// Host-side parts of libthcgw: IR table handling, RPA energy assembly and the Pade
// fit-frequency selection of the quasiparticle equation (solved on the device, qp.hip).
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "thcgw_host.hpp"
#include "thcgw_internal.hpp"

namespace thcgw {
namespace host {

void load_tables(const thcgw_dims& d, const thcgw_ir_tables& t, Tables& o) {
  THCGW_REQUIRE(t.tau_mesh_f && t.wn_mesh_f && t.Ttw_ff && t.Twt_ff && t.Ttw_bb && t.Twt_bb && t.T_beta_t_ff,
                "IR tables: null pointer");
  o.nt = d.nt;
  o.nw_b = d.nw_b;
  o.nw_f = d.nw_f;
  o.beta = d.beta;
  o.x.assign(t.tau_mesh_f, t.tau_mesh_f + d.nt);
  o.tau.resize(d.nt);
  for (int i = 0; i < d.nt; ++i) o.tau[i] = (o.x[i] + 1.0) * d.beta / 2.0;
  for (int i = 0; i < d.nt; ++i) {
    const double diff = std::fabs(o.x[i]) - std::fabs(o.x[d.nt - i - 1]);
    THCGW_REQUIRE(diff <= 1e-6, "IAFT grid is not compatible with particle-hole symmetry at tau index " +
                                    std::to_string(i));
  }
  o.wn_f.assign(t.wn_mesh_f, t.wn_mesh_f + d.nw_f);
  auto cp = [](const double* p, size_t n, double s) {
    std::vector<cd> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = cd(p[2 * i], p[2 * i + 1]) * s;
    return v;
  };
  const double r2 = std::sqrt(2.0);
  o.Ttw_ff = cp(t.Ttw_ff, size_t(d.nt) * d.nw_f, r2 / d.beta);
  o.Twt_ff = cp(t.Twt_ff, size_t(d.nw_f) * d.nt, d.beta / r2);
  o.Ttw_bb = cp(t.Ttw_bb, size_t(d.nt) * d.nw_b, r2 / d.beta);
  o.Twt_bb = cp(t.Twt_bb, size_t(d.nw_b) * d.nt, d.beta / r2);
  o.T_beta = cp(t.T_beta_t_ff, size_t(d.nt), 1.0);
}

void fold_ph(const Tables& tab, const thcgw_dims& d, std::vector<cd>& Tp, std::vector<cd>& Tm) {
  const int nt = d.nt, nwb = d.nw_b;
  const int nth = (nt % 2 == 0) ? nt / 2 : nt / 2 + 1;
  const int nwh = (nwb % 2 == 0) ? nwb / 2 : nwb / 2 + 1;
  Tp.assign(size_t(nwh) * nth, cd(0.0));
  Tm.assign(size_t(nth) * nwh, cd(0.0));
  for (int n = 0; n < nwh; ++n) {
    const int iw = nwb / 2 + n;
    for (int it = 0; it < nth; ++it) {
      const int imt = nt - it - 1;
      const cd a = tab.Twt_bb[size_t(iw) * nt + it];
      Tp[size_t(n) * nth + it] = (it == imt) ? a : a + tab.Twt_bb[size_t(iw) * nt + imt];
    }
  }
  for (int it = 0; it < nth; ++it)
    for (int n = 0; n < nwh; ++n) {
      const int iw = nwb / 2 + n, imw = nwb / 2 - n;
      const cd a = tab.Ttw_bb[size_t(it) * nwb + iw];
      Tm[size_t(it) * nwh + n] = (iw == imw) ? a : a + tab.Ttw_bb[size_t(it) * nwb + imw];
    }
}

std::vector<cd> eps_inv_q0(const Tables& tab, const thcgw_dims& d, const std::vector<cd>& eps, const double* qcart,
                           int nq, int mode, double q_max, int fit_order) {
  const int nth = (d.nt % 2 == 0) ? d.nt / 2 : d.nt / 2 + 1;
  const int nwh = (d.nw_b % 2 == 0) ? d.nw_b / 2 : d.nw_b / 2 + 1;
  std::vector<double> qa(nq);
  for (int q = 0; q < nq; ++q)
    qa[q] = std::sqrt(qcart[3 * q] * qcart[3 * q] + qcart[3 * q + 1] * qcart[3 * q + 1] + qcart[3 * q + 2] * qcart[3 * q + 2]);
  std::vector<cd> out(nth);
  auto column = [&](int q) {
    for (int t = 0; t < nth; ++t) out[t] = eps[size_t(t) * nq + q];
  };
  if (mode == 2 || mode == 3) {
    // filter_qpts: 0 < |q| <= q_max, distinct |q|, (2D: q_z == 0), in q order.  |q| values
    // within 1e-10 (relative) count as equal: images of equal length computed along
    // different axes differ in the last bit, and keeping both would make the fit singular.
    std::vector<int> sel;
    std::vector<double> seen;
    for (int q = 0; q < nq; ++q) {
      const double r = qa[q];
      const bool dup = std::any_of(seen.begin(), seen.end(), [r](double v) { return std::fabs(v - r) <= 1e-10 * std::max(v, r); });
      if (r > 0.0 && r <= q_max && (mode != 3 || qcart[3 * q + 2] == 0.0) && !dup) {
        seen.push_back(r);
        sel.push_back(q);
      }
    }
    if (sel.empty()) throw std::runtime_error("Gygi extrapolation: no q-point with 0 < |q| <= q_max");
    if (sel.size() < 3) {
      column(sel[0]);
      return out;
    }
    // tau -> iW (PH folded), fit per frequency, iW -> tau
    std::vector<cd> Tp, Tm;
    fold_ph(tab, d, Tp, Tm);
    const int ns = static_cast<int>(sel.size()), nc = fit_order + 1;
    std::vector<double> A(size_t(ns) * nc);
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < nc; ++j) A[size_t(i) * nc + j] = std::pow(qa[sel[i]], j);
    // (A^T A)^-1 by Gauss-Jordan with partial pivoting (A is real)
    std::vector<double> N(size_t(nc) * nc, 0.0), Ni(size_t(nc) * nc, 0.0);
    for (int a = 0; a < nc; ++a)
      for (int b = 0; b < nc; ++b)
        for (int i = 0; i < ns; ++i) N[size_t(a) * nc + b] += A[size_t(i) * nc + a] * A[size_t(i) * nc + b];
    for (int a = 0; a < nc; ++a) Ni[size_t(a) * nc + a] = 1.0;
    for (int c = 0; c < nc; ++c) {
      int piv = c;
      for (int r = c + 1; r < nc; ++r)
        if (std::fabs(N[size_t(r) * nc + c]) > std::fabs(N[size_t(piv) * nc + c])) piv = r;
      if (N[size_t(piv) * nc + c] == 0.0) throw std::runtime_error("Gygi extrapolation: singular normal equations");
      for (int j = 0; j < nc; ++j) {
        std::swap(N[size_t(c) * nc + j], N[size_t(piv) * nc + j]);
        std::swap(Ni[size_t(c) * nc + j], Ni[size_t(piv) * nc + j]);
      }
      const double inv = 1.0 / N[size_t(c) * nc + c];
      for (int j = 0; j < nc; ++j) {
        N[size_t(c) * nc + j] *= inv;
        Ni[size_t(c) * nc + j] *= inv;
      }
      for (int r = 0; r < nc; ++r)
        if (r != c) {
          const double f = N[size_t(r) * nc + c];
          for (int j = 0; j < nc; ++j) {
            N[size_t(r) * nc + j] -= f * N[size_t(c) * nc + j];
            Ni[size_t(r) * nc + j] -= f * Ni[size_t(c) * nc + j];
          }
        }
    }
    std::vector<cd> ew0(nwh);
    for (int n = 0; n < nwh; ++n) {
      std::vector<cd> atb(nc, 0.0);
      for (int i = 0; i < ns; ++i) {
        cd e = 0.0;  // eps(iW_n, q_i)
        for (int t = 0; t < nth; ++t) e += Tp[size_t(n) * nth + t] * eps[size_t(t) * nq + sel[i]];
        for (int a = 0; a < nc; ++a) atb[a] += A[size_t(i) * nc + a] * e;
      }
      cd x0 = 0.0;
      for (int a = 0; a < nc; ++a) x0 += Ni[a] * atb[a];
      ew0[n] = x0;
    }
    for (int t = 0; t < nth; ++t) {
      cd v = 0.0;
      for (int n = 0; n < nwh; ++n) v += Tm[size_t(t) * nwh + n] * ew0[n];
      out[t] = v;
    }
    return out;
  }
  // find_smallest_qabs (a single q-point is its own answer)
  int idx = 0;
  if (nq > 1) {
    idx = -1;
    for (int q = 0; q < nq; ++q)
      if (qa[q] > 0.0 && (idx < 0 || qa[q] < qa[idx])) idx = q;
    if (idx < 0) throw std::runtime_error("Gygi correction: no nonzero q-point");
  }
  column(idx);
  return out;
}

double ec_from_ew(const Tables& tab, const std::vector<cd>& ew, int nq) {
  cd e = 0.0;
  for (int it = 0; it < tab.nt; ++it) {
    cd et = 0.0;
    for (int w = 0; w < tab.nw_b; ++w) et += tab.Ttw_bb[size_t(it) * tab.nw_b + w] * ew[w];
    e += tab.T_beta[it] * et;
  }
  return e.real() / (2.0 * double(nq));
}

double gm_from_sigmag(const Tables& tab, const std::vector<cd>& sg) {
  cd e = 0.0;
  for (int it = 0; it < tab.nt; ++it) {
    cd st = 0.0;
    for (int w = 0; w < tab.nw_f; ++w) st += tab.Ttw_ff[size_t(it) * tab.nw_f + w] * sg[w];
    e += tab.T_beta[it] * st;
  }
  return (-0.5 * 2.0 * e).real();  // spin factor 2 for ns = 1 (scf_common.cpp:130-131)
}

std::vector<int> pade_fit_indices(int niw, int nfit) {
  THCGW_REQUIRE(niw % 2 == 0, "Pade: the fermionic mesh size must be even");
  const int niw_pos = niw / 2, shift = niw / 2;
  THCGW_REQUIRE(nfit >= 2 && nfit <= niw_pos, "Pade: nfit must be in [2, nw_f/2]");
  std::vector<int> idx;
  const int num_lchunk = niw_pos % nfit;
  if (num_lchunk != 0) {
    const int step = niw_pos / nfit + 1;
    for (int i = 0; i < step * num_lchunk; i += step) idx.push_back(i + shift);
    for (int i = step * num_lchunk; i < niw_pos; i += step - 1) idx.push_back(i + shift);
  } else {
    const int step = niw_pos / nfit;
    for (int i = 0; i < niw_pos; i += step) idx.push_back(i + shift);
  }
  THCGW_REQUIRE(static_cast<int>(idx.size()) == nfit, "Pade: fit-frequency selection size mismatch");
  return idx;
}

}  // namespace host
}  // namespace thcgw
