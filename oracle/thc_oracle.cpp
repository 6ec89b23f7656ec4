// FP64 CPU oracle for the THC imaginary-time GW/RPA inner loop.
//
// TEST INFRASTRUCTURE ONLY.  This file restates CoQui's host algorithm in plain
// C++/OpenMP so that tests can check the HIP library against it and bench.py can
// time it as the CPU baseline.  Nothing in the product (coqui_amd/, include/)
// links, loads or calls it.
//
// Every function follows a reference file:line (paths relative to the CoQui
// repository root).  Conventions kept exactly:
//   * arrays are C-order (row-major), complex<double>;
//   * the k<->R transforms are dense N_k x N_k DFT GEMMs (rpa_pi.icc:115-124,165-172),
//     not FFTs, with f_Rk = e^{-ik.R}/N_k and f_kR = e^{+ik.R} (kpoint_utils.hpp:461-545,588-680);
//   * Dyson is A = V Pi; A = I - A; A = A^{-1} (getrf+getri); A -= I; W = A V
//     (scr_coulomb_t.cpp:171-199);
//   * the RPA energy uses Tr(Pi V) + log(Re det(I - Pi V)) per (q, n) (thc_rpa.icc:130-173).
//
// Parity status of each function is documented in the header of oracle/oracle.py.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>
#include <quadmath.h>

using cd = std::complex<double>;

namespace {

inline cd cmul(cd a, cd b) {
  return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
}

// op(X)[i][k] for a row-major matrix with leading dimension ld.
// t = 'N': X[i][k]; 'T': X[k][i]; 'C': conj(X[k][i]).
inline cd opget(char t, const cd* X, long ld, long i, long k) {
  if (t == 'N') return X[i * ld + k];
  if (t == 'T') return X[k * ld + i];
  return std::conj(X[k * ld + i]);
}

// C = alpha * op(A) * op(B) + beta * C  (row-major, M x N result, inner dimension K).
// op(B) is packed into split re/im rows, then 4 rows of C are accumulated at a time
// so that the inner loop over columns vectorises.
void zgemm(char ta, char tb, long M, long N, long K, cd alpha, const cd* A, long lda,
           const cd* B, long ldb, cd beta, cd* C, long ldc) {
  std::vector<double> br(static_cast<size_t>(K) * N), bi(static_cast<size_t>(K) * N);
#pragma omp parallel for schedule(static)
  for (long k = 0; k < K; ++k)
    for (long j = 0; j < N; ++j) {
      cd b = opget(tb, B, ldb, k, j);
      br[k * N + j] = b.real();
      bi[k * N + j] = b.imag();
    }
  const long RB = 4;
#pragma omp parallel
  {
    std::vector<double> accr(RB * N), acci(RB * N);
#pragma omp for schedule(dynamic, 1)
    for (long i0 = 0; i0 < M; i0 += RB) {
      long nr = std::min(RB, M - i0);
      std::fill(accr.begin(), accr.end(), 0.0);
      std::fill(acci.begin(), acci.end(), 0.0);
      for (long k = 0; k < K; ++k) {
        const double* __restrict brk = &br[k * N];
        const double* __restrict bik = &bi[k * N];
        for (long r = 0; r < nr; ++r) {
          cd a = opget(ta, A, lda, i0 + r, k);
          double ar = a.real(), ai = a.imag();
          double* __restrict cr = &accr[r * N];
          double* __restrict ci = &acci[r * N];
          for (long j = 0; j < N; ++j) {
            cr[j] += ar * brk[j] - ai * bik[j];
            ci[j] += ar * bik[j] + ai * brk[j];
          }
        }
      }
      for (long r = 0; r < nr; ++r)
        for (long j = 0; j < N; ++j) {
          cd acc(accr[r * N + j], acci[r * N + j]);
          cd old = (beta == cd(0.0)) ? cd(0.0) : cmul(beta, C[(i0 + r) * ldc + j]);
          C[(i0 + r) * ldc + j] = cmul(alpha, acc) + old;
        }
    }
  }
}

// LU with partial pivoting of a row-major n x n matrix, A = P L U.  The pivot is
// chosen by |Re|+|Im| as LAPACK's izamax does (used through nda::lapack::getrf,
// slate_ops.hpp:324-333).  Returns 0 on success, k+1 if U(k,k) == 0.
int zgetrf(long n, cd* A, std::vector<long>& piv) {
  piv.resize(n);
  int info = 0;
  for (long c = 0; c < n; ++c) {
    long p = c;
    double best = -1.0;
    for (long r = c; r < n; ++r) {
      double v = std::fabs(A[r * n + c].real()) + std::fabs(A[r * n + c].imag());
      if (v > best) { best = v; p = r; }
    }
    piv[c] = p;
    if (p != c)
      for (long j = 0; j < n; ++j) std::swap(A[c * n + j], A[p * n + j]);
    cd d = A[c * n + c];
    if (d == cd(0.0)) { if (!info) info = static_cast<int>(c + 1); continue; }
    cd inv = 1.0 / d;
#pragma omp parallel for schedule(static) if (n - c > 64)
    for (long r = c + 1; r < n; ++r) {
      cd l = cmul(A[r * n + c], inv);
      A[r * n + c] = l;
      const cd* __restrict urow = &A[c * n];
      cd* __restrict arow = &A[r * n];
      for (long j = c + 1; j < n; ++j) arow[j] -= cmul(l, urow[j]);
    }
  }
  return info;
}

// Inverse from the LU factors (getri): solve A X = I column block by column block.
void zgetri(long n, const cd* LU, const std::vector<long>& piv, cd* Ainv) {
  // X = U^{-1} L^{-1} P I ; build P I then forward/back substitute all columns.
  std::vector<cd> X(static_cast<size_t>(n) * n, cd(0.0));
  for (long i = 0; i < n; ++i) X[i * n + i] = 1.0;
  for (long c = 0; c < n; ++c)
    if (piv[c] != c)
      for (long j = 0; j < n; ++j) std::swap(X[c * n + j], X[piv[c] * n + j]);
  // Columns of X are independent: each thread substitutes one block of columns.
  const long JB = 32;
#pragma omp parallel for schedule(dynamic, 1)
  for (long j0 = 0; j0 < n; j0 += JB) {
    const long j1 = std::min(n, j0 + JB);
    // L has unit diagonal: row r of X -= sum_{c<r} L[r][c] X[c]
    for (long r = 1; r < n; ++r) {
      cd* __restrict xr = &X[r * n];
      for (long c = 0; c < r; ++c) {
        cd l = LU[r * n + c];
        if (l == cd(0.0)) continue;
        const cd* __restrict xc = &X[c * n];
        for (long j = j0; j < j1; ++j) xr[j] -= cmul(l, xc[j]);
      }
    }
    for (long r = n - 1; r >= 0; --r) {
      cd* __restrict xr = &X[r * n];
      for (long c = r + 1; c < n; ++c) {
        cd u = LU[r * n + c];
        const cd* __restrict xc = &X[c * n];
        for (long j = j0; j < j1; ++j) xr[j] -= cmul(u, xc[j]);
      }
      cd inv = 1.0 / LU[r * n + r];
      for (long j = j0; j < j1; ++j) xr[j] = cmul(xr[j], inv);
    }
  }
  std::memcpy(Ainv, X.data(), sizeof(cd) * n * n);
}

// k-point and lattice-vector bookkeeping of the synthetic Gamma-centred grid:
// k index C-order over (i,j,l), k_crys = (i/n1, j/n2, l/n3); R index p -> (a,b,c)
// with each component wrapped into (-n/2, n/2] (kpoint_utils.hpp:524-533).
void r_index(const int* kg, long p, long* abc) {
  long nx = kg[0], ny = kg[1], nz = kg[2];
  long a = p / (ny * nz), b = (p / nz) % ny, c = p % nz;
  if (a > nx / 2) a -= nx;
  if (b > ny / 2) b -= ny;
  if (c > nz / 2) c -= nz;
  abc[0] = a; abc[1] = b; abc[2] = c;
}

double k_dot_R(const int* kg, long k, const long* abc) {
  long ny = kg[1], nz = kg[2];
  long i = k / (ny * nz), j = (k / nz) % ny, l = k % nz;
  // lattv = I, reciprocal vectors 2*pi*I  =>  k.R = 2 pi k_crys . (a,b,c)
  return 2.0 * M_PI * (double(i) / kg[0] * abc[0] + double(j) / kg[1] * abc[1] +
                       double(l) / kg[2] * abc[2]);
}

long kminus(const int* kg, long k, long q) {  // index of k - q (bz_symmetry.hpp:517-524)
  long ny = kg[1], nz = kg[2];
  long ki = k / (ny * nz), kj = (k / nz) % ny, kl = k % nz;
  long qi = q / (ny * nz), qj = (q / nz) % ny, ql = q % nz;
  long a = (ki - qi + kg[0]) % kg[0], b = (kj - qj + kg[1]) % kg[1], c = (kl - ql + kg[2]) % kg[2];
  return (a * ny + b) * nz + c;
}

}  // namespace

extern "C" {

int orc_num_threads() { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

void orc_zgemm(char ta, char tb, long M, long N, long K, const cd* alpha, const cd* A, long lda,
               const cd* B, long ldb, const cd* beta, cd* C, long ldc) {
  zgemm(ta, tb, M, N, K, *alpha, A, lda, B, ldb, *beta, C, ldc);
}

int orc_zinverse(long n, cd* A) {
  std::vector<long> piv;
  int info = zgetrf(n, A, piv);
  if (info) return info;
  zgetri(n, A, piv, A);
  return 0;
}

// f_Rk(R,k) = exp(-i k.R)/N_k  (kpoint_utils.hpp:461-491 with the R list of :524-533)
// f_kR(k,R) = exp(+i k.R)      (kpoint_utils.hpp:588-620, weights 1, :655-658)
void orc_kR_coefficients(const int* kgrid, cd* f_Rk, cd* f_kR) {
  long nk = long(kgrid[0]) * kgrid[1] * kgrid[2];
  for (long R = 0; R < nk; ++R) {
    long abc[3];
    r_index(kgrid, R, abc);
    for (long k = 0; k < nk; ++k) {
      double kR = k_dot_R(kgrid, k, abc);
      f_Rk[R * nk + k] = std::exp(cd(0.0, -kR)) / double(nk);
      f_kR[k * nk + R] = std::exp(cd(0.0, kR));
    }
  }
}

// a1: G_ij(tau,k) = sum_a C_ia g(e_a - mu, tau) conj(C_ja)
// qp_scf_common.cpp:82-118; tau = (x+1) beta/2 from the fermionic IR mesh x (:96-98).
void orc_update_G(int nt, int nk, int no, const double* tau, double beta, double mu,
                  const cd* C, const double* E, cd* G) {
  auto g0 = [&](double e, double t) {
    double x = e - mu;
    if (x > 0) return -std::exp(-x * t) / (1 + std::exp(-x * beta));
    return -std::exp(x * (beta - t)) / (1 + std::exp(x * beta));
  };
  std::vector<cd> GCdag(static_cast<size_t>(no) * no);
  for (int it = 0; it < nt; ++it)
    for (int ik = 0; ik < nk; ++ik) {
      const cd* Ck = C + static_cast<size_t>(ik) * no * no;
      for (int a = 0; a < no; ++a) {
        double ga = g0(E[ik * no + a], tau[it]);
        for (int j = 0; j < no; ++j) GCdag[a * no + j] = std::conj(Ck[j * no + a]) * ga;
      }
      cd* Gt = G + (static_cast<size_t>(it) * nk + ik) * no * no;
      zgemm('N', 'N', no, no, no, 1.0, Ck, no, GCdag.data(), no, 0.0, Gt, no);
    }
}

// a2: O_PQ = X_Pa O_ab conj(X_Qb): A = X O (thc_solver_comm.hpp:422), O_PQ = A X^H (:427)
void orc_primary_to_aux(long nb, int nk, int nmu, int no, const cd* X, const cd* O, cd* out) {
  std::vector<cd> A(static_cast<size_t>(nmu) * no);
  for (long b = 0; b < nb; ++b)
    for (int k = 0; k < nk; ++k) {
      const cd* Xk = X + static_cast<size_t>(k) * nmu * no;
      const cd* Ok = O + (static_cast<size_t>(b) * nk + k) * no * no;
      cd* Pk = out + (static_cast<size_t>(b) * nk + k) * nmu * nmu;
      zgemm('N', 'N', nmu, no, no, 1.0, Xk, no, Ok, no, 0.0, A.data(), no);
      zgemm('N', 'C', nmu, nmu, no, 1.0, A.data(), no, Xk, no, 0.0, Pk, nmu);
    }
}

// a10: O_ab = conj(X_Pa) O_PQ X_Qb: B = X^H O (thc_solver_comm.hpp:560), O_ab = B X (:563)
void orc_aux_to_primary(long nb, int nk, int nmu, int no, const cd* X, const cd* O, cd* out) {
  std::vector<cd> B(static_cast<size_t>(no) * nmu);
  for (long b = 0; b < nb; ++b)
    for (int k = 0; k < nk; ++k) {
      const cd* Xk = X + static_cast<size_t>(k) * nmu * no;
      const cd* Ok = O + (static_cast<size_t>(b) * nk + k) * nmu * nmu;
      cd* Ak = out + (static_cast<size_t>(b) * nk + k) * no * no;
      zgemm('C', 'N', no, nmu, nmu, 1.0, Xk, no, Ok, nmu, 0.0, B.data(), nmu);
      zgemm('N', 'N', no, no, nmu, 1.0, B.data(), nmu, Xk, no, 0.0, Ak, no);
    }
}

// a3+a4: R-space polarizability, rpa_pi.icc:102-176 (ns = 1, no symmetry):
//   Gp = X G(tau_i) X^H ; Gp_R = f_Rk Gp ; Gn = X G(beta - tau_i) X^H ; Gn_R = f_Rk Gn
//   GG_R = sp * Gp_R o conj(Gn_R), sp = -2 ; Pi(tau_i, q) = f_qR GG_R
// Output Pi_tqPQ has shape (nt_half, nk, nmu, nmu).
void orc_pi_rspace(int nt, const int* kgrid, int nmu, int no, const cd* G, const cd* X, cd* Pi) {
  long nk = long(kgrid[0]) * kgrid[1] * kgrid[2];
  long M = long(nmu) * nmu;
  int nt_half = (nt % 2 == 0) ? nt / 2 : nt / 2 + 1;
  const double sp = -2.0;
  std::vector<cd> f_Rk(nk * nk), f_kR(nk * nk);
  orc_kR_coefficients(kgrid, f_Rk.data(), f_kR.data());
  std::vector<cd> Gp(nk * M), Gn(nk * M), buf(nk * M);
  for (int it = 0; it < nt_half; ++it) {
    orc_primary_to_aux(1, nk, nmu, no, X, G + static_cast<size_t>(it) * nk * no * no, Gp.data());
    if (nk != 1) {
      zgemm('N', 'N', nk, M, nk, 1.0, f_Rk.data(), nk, Gp.data(), M, 0.0, buf.data(), M);
      Gp.swap(buf);
    }
    orc_primary_to_aux(1, nk, nmu, no, X, G + static_cast<size_t>(nt - it - 1) * nk * no * no, Gn.data());
    if (nk != 1) {
      zgemm('N', 'N', nk, M, nk, 1.0, f_Rk.data(), nk, Gn.data(), M, 0.0, buf.data(), M);
      Gn.swap(buf);
    }
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nk * M; ++x) Gp[x] = cmul(Gp[x], std::conj(Gn[x])) * sp;
    cd* Pit = Pi + static_cast<size_t>(it) * nk * M;
    if (nk != 1)
      zgemm('N', 'N', nk, M, nk, 1.0, f_kR.data(), nk, Gp.data(), M, 0.0, Pit, M);
    else
      std::memcpy(Pit, Gp.data(), sizeof(cd) * M);
  }
}

// k-space polarizability, rpa_pi.icc:182-336 (second algorithm for the same object):
//   Pi^k(tau_i, q) = (sp/N_k) sum_k conj(G_PQ(tau_i, k - q)) o G_PQ(beta - tau_i, k)
// For Hermitian G_PQ the two are related by Pi^k_PQ(q) = Pi^R_QP(-q).
void orc_pi_kspace(int nt, const int* kgrid, int nmu, int no, const cd* G, const cd* X, cd* Pi) {
  long nk = long(kgrid[0]) * kgrid[1] * kgrid[2];
  long M = long(nmu) * nmu;
  int nt_half = (nt % 2 == 0) ? nt / 2 : nt / 2 + 1;
  const double factor = -2.0 / double(nk);
  std::vector<cd> Gp(nk * M), Gn(nk * M);
  for (int it = 0; it < nt_half; ++it) {
    orc_primary_to_aux(1, nk, nmu, no, X, G + static_cast<size_t>(it) * nk * no * no, Gp.data());
    orc_primary_to_aux(1, nk, nmu, no, X, G + static_cast<size_t>(nt - it - 1) * nk * no * no, Gn.data());
    cd* Pit = Pi + static_cast<size_t>(it) * nk * M;
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nk * M; ++x) {
      long q = x / M, pq = x % M;
      cd acc = 0.0;
      for (long k = 0; k < nk; ++k) {
        long kmq = kminus(kgrid, k, q);
        acc += cmul(std::conj(Gp[kmq * M + pq]), Gn[k * M + pq]) * factor;
      }
      Pit[x] = acc;
    }
  }
}

// a5: IAFT::tau_to_w_PHsym, IAFT.icc:49-73.  Twt_bb is (nw_b x nt_b), already rescaled
// by beta/sqrt(2) (ir_driver.hpp:71).  X_t is (nt_half, dim), X_w is (nw_half, dim).
void orc_tau_to_w_phsym(int nt_b, int nw_b, const cd* Twt_bb, long dim, const cd* Xt, cd* Xw) {
  int nt = (nt_b % 2 == 0) ? nt_b / 2 : nt_b / 2 + 1;
  int nw = (nw_b % 2 == 0) ? nw_b / 2 : nw_b / 2 + 1;
  std::vector<cd> T(static_cast<size_t>(nw) * nt);
  for (int n = 0; n < nw; ++n) {
    int iw = nw_b / 2 + n;
    for (int it = 0; it < nt; ++it) {
      int imt = nt_b - it - 1;
      T[n * nt + it] = (it == imt) ? Twt_bb[iw * nt_b + it] : Twt_bb[iw * nt_b + it] + Twt_bb[iw * nt_b + imt];
    }
  }
  zgemm('N', 'N', nw, dim, nt, 1.0, T.data(), nt, Xt, dim, 0.0, Xw, dim);
}

// a7: IAFT::w_to_tau_PHsym, IAFT.icc:104-128.  Ttw_bb is (nt_b x nw_b), rescaled by sqrt(2)/beta.
void orc_w_to_tau_phsym(int nt_b, int nw_b, const cd* Ttw_bb, long dim, const cd* Xw, cd* Xt) {
  int nt = (nt_b % 2 == 0) ? nt_b / 2 : nt_b / 2 + 1;
  int nw = (nw_b % 2 == 0) ? nw_b / 2 : nw_b / 2 + 1;
  std::vector<cd> T(static_cast<size_t>(nt) * nw);
  for (int it = 0; it < nt; ++it)
    for (int n = 0; n < nw; ++n) {
      int iw = nw_b / 2 + n, imw = nw_b / 2 - n;
      T[it * nw + n] = (iw == imw) ? Ttw_bb[it * nw_b + iw] : Ttw_bb[it * nw_b + iw] + Ttw_bb[it * nw_b + imw];
    }
  zgemm('N', 'N', nt, dim, nw, 1.0, T.data(), nw, Xw, dim, 0.0, Xt, dim);
}

// Generic full-grid transforms (IAFT.icc:9-26, 81-98): out(n_out, dim) = T(n_out, n_in) in(n_in, dim)
void orc_transform(int n_out, int n_in, const cd* T, long dim, const cd* in, cd* out) {
  zgemm('N', 'N', n_out, dim, n_in, 1.0, T, n_in, in, dim, 0.0, out, dim);
}

// a6: scr_coulomb_t::dyson_W_in_place, scr_coulomb_t.cpp:171-199 (one rank):
//   A = V Pi ; A = I - A ; A = A^{-1} ; A = A - I ; W = A V
// PiW is (nw_half, nq, nmu, nmu): in = Pi(iW), out = W(iW) - V.  Returns getrf info.
int orc_dyson(int nw_half, int nq, int nmu, cd* PiW, const cd* V) {
  long M = long(nmu) * nmu;
  std::vector<cd> A(M);
  std::vector<long> piv;
  for (int n = 0; n < nw_half; ++n)
    for (int q = 0; q < nq; ++q) {
      cd* P = PiW + (static_cast<size_t>(n) * nq + q) * M;
      const cd* Vq = V + static_cast<size_t>(q) * M;
      zgemm('N', 'N', nmu, nmu, nmu, 1.0, Vq, nmu, P, nmu, 0.0, A.data(), nmu);
      for (long x = 0; x < M; ++x) A[x] = -A[x];
      for (int i = 0; i < nmu; ++i) A[i * nmu + i] += 1.0;
      int info = zgetrf(nmu, A.data(), piv);
      if (info) return info;
      zgetri(nmu, A.data(), piv, A.data());
      for (int i = 0; i < nmu; ++i) A[i * nmu + i] -= 1.0;
      zgemm('N', 'N', nmu, nmu, nmu, 1.0, A.data(), nmu, Vq, nmu, 0.0, P, nmu);
    }
  return 0;
}

// a9 + a10: R-space self-energy, thc_gw.cpp:152-178 -> thc_gw.icc:70-213, both passes:
//   pass 1 (tau_i, i < nt_half): Sigma_PQ(k) = f_kR[-(f_Rk G_PQ(tau_i)) o (f_Rk W(tau_i))]
//   pass 2 (beta - tau_i, i < nt/2): same with G(beta - tau_i), stored at nt - 1 - i
// followed by Sigma_ab = X^H Sigma_PQ X (setup_Sigma_primary, thc_gw.icc:37-65).
// W_qt is (nk, nt_half, nmu, nmu); Sigma is (nt, nk, no, no).
void orc_sigma_rspace(int nt, const int* kgrid, int nmu, int no, const cd* G, const cd* X,
                      const cd* W_qt, cd* Sigma) {
  long nk = long(kgrid[0]) * kgrid[1] * kgrid[2];
  long M = long(nmu) * nmu;
  int nt_half = (nt % 2 == 0) ? nt / 2 : nt / 2 + 1;
  std::vector<cd> f_Rk(nk * nk), f_kR(nk * nk);
  orc_kR_coefficients(kgrid, f_Rk.data(), f_kR.data());
  std::vector<cd> Gb(nk * M), Wb(nk * M), buf(nk * M);
  std::memset(static_cast<void*>(Sigma), 0, sizeof(cd) * size_t(nt) * nk * no * no);
  for (int pass = 0; pass < 2; ++pass) {
    int ntp = (pass == 0) ? nt_half : nt / 2;
    for (int it = 0; it < ntp; ++it) {
      int itg = (pass == 0) ? it : nt - it - 1;
      orc_primary_to_aux(1, nk, nmu, no, X, G + static_cast<size_t>(itg) * nk * no * no, Gb.data());
      for (long q = 0; q < nk; ++q)
        std::memcpy(&Wb[q * M], W_qt + (static_cast<size_t>(q) * nt_half + it) * M, sizeof(cd) * M);
      if (nk != 1) {
        zgemm('N', 'N', nk, M, nk, 1.0, f_Rk.data(), nk, Gb.data(), M, 0.0, buf.data(), M);
        Gb.swap(buf);
        zgemm('N', 'N', nk, M, nk, 1.0, f_Rk.data(), nk, Wb.data(), M, 0.0, buf.data(), M);
        Wb.swap(buf);
      }
#pragma omp parallel for schedule(static)
      for (long x = 0; x < nk * M; ++x) Gb[x] = -1.0 * cmul(Gb[x], Wb[x]);
      if (nk != 1) {
        zgemm('N', 'N', nk, M, nk, 1.0, f_kR.data(), nk, Gb.data(), M, 0.0, buf.data(), M);
        Gb.swap(buf);
      }
      orc_aux_to_primary(1, nk, nmu, no, X, Gb.data(), Sigma + static_cast<size_t>(itg) * nk * no * no);
    }
  }
}

// THC Hartree-Fock potential, hf_t::thc_hf_Xqindep (thc_hf.icc:18-300) for ns = npol = 1,
// no symmetry and ignore_g0 (HF_K_correction returns at once, hf_t.cpp:49-52):
//   Dm_kPQ = X_k Dm_k X_k^H                                  (primary_to_aux, :162-163)
//   Dm_RPQ = f_Rk Dm_kPQ                                      (:170-179)
//   Hartree: J_P = sum_Q U(q=0)_PQ * 2 Dm_{R=0,QQ}            (:75, :182-205, factor 2 for ns=npol=1)
//            F_kPP += J_P for every k
//   exchange: U_RPQ = f_Rk U_qPQ                              (:143-151)
//             F_kPQ += f_kR [ -Dm_RPQ o U_RPQ ]               (:211-232)
//   F_kij = X_k^H F_kPQ X_k                                   (aux_to_primary, :236-237)
// Dm and F are (nk, no, no); U = V is (nk, nmu, nmu).
void orc_thc_hf(const int* kgrid, int nmu, int no, const cd* Dm, const cd* X, const cd* V, int hartree, int exchange,
                cd* F) {
  long nk = long(kgrid[0]) * kgrid[1] * kgrid[2];
  long M = long(nmu) * nmu;
  std::vector<cd> f_Rk(nk * nk), f_kR(nk * nk);
  orc_kR_coefficients(kgrid, f_Rk.data(), f_kR.data());
  std::vector<cd> D(nk * M), U(V, V + nk * M), buf(nk * M), FP(nk * M, cd(0.0));
  orc_primary_to_aux(1, int(nk), nmu, no, X, Dm, D.data());
  if (nk != 1) {
    zgemm('N', 'N', nk, M, nk, 1.0, f_Rk.data(), nk, D.data(), M, 0.0, buf.data(), M);
    D.swap(buf);
  }
  if (hartree) {
    std::vector<cd> dQ(nmu), J(nmu, cd(0.0));
    for (int Q = 0; Q < nmu; ++Q) dQ[Q] = 2.0 * D[long(Q) * nmu + Q];  // R = 0 is the first R index
    for (int P = 0; P < nmu; ++P)
      for (int Q = 0; Q < nmu; ++Q) J[P] += V[long(P) * nmu + Q] * dQ[Q];  // U(q = 0)
    for (long k = 0; k < nk; ++k)
      for (int P = 0; P < nmu; ++P) FP[k * M + long(P) * nmu + P] += J[P];
  }
  if (exchange) {
    if (nk != 1) {
      zgemm('N', 'N', nk, M, nk, 1.0, f_Rk.data(), nk, U.data(), M, 0.0, buf.data(), M);
      U.swap(buf);
    }
#pragma omp parallel for schedule(static)
    for (long x = 0; x < nk * M; ++x) D[x] = -1.0 * cmul(D[x], U[x]);
    if (nk != 1) {
      zgemm('N', 'N', nk, M, nk, 1.0, f_kR.data(), nk, D.data(), M, 1.0, FP.data(), M);
    } else {
      for (long x = 0; x < M; ++x) FP[x] += D[x];
    }
  }
  orc_aux_to_primary(1, int(nk), nmu, no, X, FP.data(), F);
}

// Dyson equation for G on the Matsubara axis, simple_dyson::solve_dyson
// (simple_dyson.cpp:11-127), ns = 1, one rank:
//   Sigma(iw_n) = tau_to_w(Sigma(tau), fermi)                   (:39-40; IAFT.icc:9-26)
//   G(iw_n, k) = [(iw_n + mu) S_k - H0_k - F_k - Sigma(iw_n, k)]^-1   (:70-77)
//                with iw_n = i wn pi / beta                          (IAFT.hpp:55-57)
//   G(tau) = w_to_tau(G(iw), fermi)                              (:97-98; IAFT.icc:82-98)
//   Dm = -tau_to_beta(G(tau))                                     (:116-127; IAFT.icc:226-238)
// H0, F, S are (nk, no, no) (S = NULL: identity, H0 = NULL: zero); Sigma, G are
// (nt, nk, no, no); Dm (nk, no, no) may be NULL.  Returns 0 or the first singular (n, k) + 1.
int orc_dyson_g(int nt, int nw, const cd* Twt_ff, const cd* Ttw_ff, const cd* T_beta, const long* wn, double beta,
                int nk, int no, const cd* H0, const cd* F, const cd* S, const cd* Sigma, double mu, cd* G, cd* Dm) {
  const long m = long(no) * no, dim = long(nk) * m;
  std::vector<cd> Sw(size_t(nw) * dim), Gw(size_t(nw) * dim);
  orc_transform(nw, nt, Twt_ff, dim, Sigma, Sw.data());
  int bad = 0;
#pragma omp parallel for schedule(dynamic)
  for (long nkk = 0; nkk < long(nw) * nk; ++nkk) {
    const long n = nkk / nk, k = nkk % nk;
    const cd z = cd(mu, double(wn[n]) * M_PI / beta);
    cd* X = &Gw[size_t(n) * dim + size_t(k) * m];
    for (long i = 0; i < no; ++i)
      for (long j = 0; j < no; ++j) {
        const long e = k * m + i * no + j;
        const cd sij = S ? S[e] : cd(i == j ? 1.0 : 0.0, 0.0);
        X[i * no + j] = z * sij - (H0 ? H0[e] : cd(0.0)) - F[e] - Sw[size_t(n) * dim + e];
      }
    if (orc_zinverse(no, X) != 0) {
#pragma omp critical
      if (bad == 0) bad = int(nkk) + 1;
    }
  }
  orc_transform(nt, nw, Ttw_ff, dim, Gw.data(), G);
  if (Dm) {
    for (long e = 0; e < dim; ++e) {
      cd acc = 0.0;
      for (int t = 0; t < nt; ++t) acc += T_beta[t] * G[size_t(t) * dim + e];
      Dm[e] = -acc;
    }
  }
  return bad;
}

// Electron number at chemical potential mu, compute_Nelec (scf_common.cpp:10-42), with the
// spectral sum of compute_eigenspectra (simple_dyson.cpp:130-172) taken as a resolvent trace:
//   sum_b 1/(iw_n + mu - lambda_b) = Tr[((iw_n + mu) S - H0 - F - Sigma(iw_n))^-1 S],
//   lambda_b the eigenvalues of S^-1 (H0 + F + Sigma(iw_n));
//   Xw(n) = sum_k w_k (the above), w_k = 1/nk;  nelec = -2 Re tau_to_beta(w_to_tau(Xw)).
// Sw is Sigma(iw_n) (nw, nk, no, no).
double orc_nelec(int nt, int nw, const cd* Ttw_ff, const cd* T_beta, const long* wn, double beta, int nk, int no,
                 const cd* H0, const cd* F, const cd* S, const cd* Sw, double mu) {
  const long m = long(no) * no, dim = long(nk) * m;
  std::vector<cd> Xw(nw, cd(0.0));
#pragma omp parallel for schedule(dynamic)
  for (int n = 0; n < nw; ++n) {
    std::vector<cd> X(m), R(m);
    const cd z = cd(mu, double(wn[n]) * M_PI / beta);
    cd acc = 0.0;
    for (long k = 0; k < nk; ++k) {
      for (long i = 0; i < no; ++i)
        for (long j = 0; j < no; ++j) {
          const long e = k * m + i * no + j;
          const cd sij = S ? S[e] : cd(i == j ? 1.0 : 0.0, 0.0);
          X[i * no + j] = z * sij - (H0 ? H0[e] : cd(0.0)) - F[e] - Sw[size_t(n) * dim + e];
        }
      orc_zinverse(no, X.data());
      cd tr = 0.0;
      for (long i = 0; i < no; ++i)
        for (long j = 0; j < no; ++j) tr += X[i * no + j] * (S ? S[k * m + j * no + i] : cd(i == j ? 1.0 : 0.0));
      acc += tr / double(nk);
    }
    Xw[n] = acc;
  }
  cd nel = 0.0;
  for (int t = 0; t < nt; ++t) {
    cd xt = 0.0;
    for (int n = 0; n < nw; ++n) xt += Ttw_ff[long(t) * nw + n] * Xw[n];
    nel += T_beta[t] * xt;
  }
  return -2.0 * nel.real();
}

// update_mu (scf_common.cpp:147-202): bracket in steps of 0.2 from old_mu, then bisect until
// |nelec - target| < mu_tol (at most 200 bisection steps).
double orc_update_mu(int nt, int nw, const cd* Twt_ff, const cd* Ttw_ff, const cd* T_beta, const long* wn,
                     double beta, int nk, int no, const cd* H0, const cd* F, const cd* S, const cd* Sigma,
                     double nel_target, double old_mu, double mu_tol, double* nel_out) {
  const long dim = long(nk) * no * no;
  std::vector<cd> Sw(size_t(nw) * dim);
  orc_transform(nw, nt, Twt_ff, dim, Sigma, Sw.data());
  auto N = [&](double mu) { return orc_nelec(nt, nw, Ttw_ff, T_beta, wn, beta, nk, no, H0, F, S, Sw.data(), mu); };
  const double delta = 0.2;
  double nel = N(old_mu);
  if (std::abs(nel - nel_target) < mu_tol) {
    if (nel_out) *nel_out = nel;
    return old_mu;
  }
  double mu1, mu2;
  if (nel >= nel_target) {
    mu2 = old_mu;
    mu1 = old_mu - delta;
    for (int it = 0; it < 1000 && N(mu1) > nel_target; ++it) mu1 -= delta;
  } else {
    mu1 = old_mu;
    mu2 = old_mu + delta;
    for (int it = 0; it < 1000 && N(mu2) < nel_target; ++it) mu2 += delta;
  }
  double mid = 0.5 * (mu1 + mu2);
  nel = N(mid);
  for (int it = 0; it < 200 && std::abs(nel - nel_target) >= mu_tol; ++it) {
    if (nel >= nel_target) mu2 = mid;
    else mu1 = mid;
    mid = 0.5 * (mu1 + mu2);
    nel = N(mid);
  }
  if (nel_out) *nel_out = nel;
  return mid;
}

// k-space self-energy, thc_gw.icc:342-443 with isym = 0 and no time reversal:
//   Sigma_PQ(tau, k) = -(1/N_k) sum_q G_PQ(tau, k - q) o W_PQ(tau, q)
void orc_sigma_kspace(int nt, const int* kgrid, int nmu, int no, const cd* G, const cd* X,
                      const cd* W_qt, cd* Sigma) {
  long nk = long(kgrid[0]) * kgrid[1] * kgrid[2];
  long M = long(nmu) * nmu;
  int nt_half = (nt % 2 == 0) ? nt / 2 : nt / 2 + 1;
  const double factor = -1.0 / double(nk);
  std::vector<cd> Gb(nk * M), Sb(nk * M);
  std::memset(static_cast<void*>(Sigma), 0, sizeof(cd) * size_t(nt) * nk * no * no);
  for (int pass = 0; pass < 2; ++pass) {
    int ntp = (pass == 0) ? nt_half : nt / 2;
    for (int it = 0; it < ntp; ++it) {
      int itg = (pass == 0) ? it : nt - it - 1;
      orc_primary_to_aux(1, nk, nmu, no, X, G + static_cast<size_t>(itg) * nk * no * no, Gb.data());
#pragma omp parallel for schedule(static)
      for (long x = 0; x < nk * M; ++x) {
        long k = x / M, pq = x % M;
        cd acc = 0.0;
        for (long q = 0; q < nk; ++q) {
          long kmq = kminus(kgrid, k, q);
          acc += cmul(Gb[kmq * M + pq], W_qt[(q * nt_half + it) * M + pq]) * factor;
        }
        Sb[x] = acc;
      }
      orc_aux_to_primary(1, nk, nmu, no, X, Sb.data(), Sigma + static_cast<size_t>(itg) * nk * no * no);
    }
  }
}

// a12: hermitize, scf_common.hpp:70-86
void orc_hermitize(long nb, int n, cd* A) {
  std::vector<cd> buf(static_cast<size_t>(n) * n);
  for (long b = 0; b < nb; ++b) {
    cd* Ab = A + static_cast<size_t>(b) * n * n;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) buf[i * n + j] = (Ab[i * n + j] + std::conj(Ab[j * n + i])) * 0.5;
    std::memcpy(Ab, buf.data(), sizeof(cd) * n * n);
  }
}

// log(Re det) of an LU-factored matrix, written through log|u_ii| and the phase so
// that it cannot overflow: ln|Re det| = sum ln|u_ii| + ln|cos(sum arg u_ii + pi*swaps)|.
// thc_rpa.icc:151-163 takes log(Re det) after flipping a negative real part.
static double log_re_det(long n, const cd* LU, const std::vector<long>& piv) {
  double la = 0.0, ph = 0.0;
  long swaps = 0;
  for (long i = 0; i < n; ++i) {
    la += std::log(std::abs(LU[i * n + i]));
    ph += std::arg(LU[i * n + i]);
    if (piv[i] != i) ++swaps;
  }
  ph += M_PI * double(swaps % 2);
  return la + std::log(std::fabs(std::cos(ph)));
}

// a13: RPA correlation energy, thc_rpa.icc:72-181 (one rank, q_weight = 1):
//   for (q, n): A = Pi(iW_n, q) V_q ; e_w(+-n) += Tr A ; e_w(+-n) += log(Re det(I - A))
//   e_t = Ttw_bb e_w ; Ec = T_beta_t_ff e_t / (2 N_k)  (fermionic T_beta on bosonic tau data)
// Pi_w is (nw_half, nq, nmu, nmu).  Returns Ec and writes e_w (nw_b, complex) if non-null.
double orc_rpa_ec(int nt_b, int nw_b, const cd* Ttw_bb, const cd* T_beta_t_ff, int nq, int nmu,
                  const cd* Pi_w, const cd* V, cd* e_w_out) {
  int nw_half = (nw_b % 2 == 0) ? nw_b / 2 : nw_b / 2 + 1;
  long M = long(nmu) * nmu;
  std::vector<cd> e_w(nw_b, cd(0.0)), A(M);
  std::vector<long> piv;
  for (int q = 0; q < nq; ++q)
    for (int n = 0; n < nw_half; ++n) {
      int wn = nw_b / 2 + n, mwn = nw_b / 2 - n;
      const cd* P = Pi_w + (static_cast<size_t>(n) * nq + q) * M;
      zgemm('N', 'N', nmu, nmu, nmu, 1.0, P, nmu, V + static_cast<size_t>(q) * M, nmu, 0.0, A.data(), nmu);
      cd tr = 0.0;
      for (int i = 0; i < nmu; ++i) tr += A[i * nmu + i];
      e_w[wn] += tr;
      if (mwn != wn) e_w[mwn] += tr;
      for (long x = 0; x < M; ++x) A[x] = -A[x];
      for (int i = 0; i < nmu; ++i) A[i * nmu + i] += 1.0;
      zgetrf(nmu, A.data(), piv);
      double l = log_re_det(nmu, A.data(), piv);
      e_w[wn] += l;
      if (mwn != wn) e_w[mwn] += l;
    }
  if (e_w_out) std::memcpy(e_w_out, e_w.data(), sizeof(cd) * nw_b);
  cd e = 0.0;
  for (int it = 0; it < nt_b; ++it) {
    cd et = 0.0;
    for (int w = 0; w < nw_b; ++w) et += Ttw_bb[it * nw_b + w] * e_w[w];
    e += T_beta_t_ff[it] * et;
  }
  return e.real() / (2.0 * double(nq));
}

// ------------------------------------------------------------------------------------
// a15: quasiparticle equation with Thiele Pade continuation.
// pade_driver.hpp:24-65 (choice of fit frequencies), pade_t.hpp:34-168 (continued
// fraction; the reference works in 100-digit cpp_dec_float, this restatement in
// binary128), qp_scf_common.cpp:150-311 (MO projection, tau -> iw_n), :679-728 (bisection).
// ------------------------------------------------------------------------------------
namespace {
struct qc {
  __float128 r, i;
};
inline qc qadd(qc a, qc b) { return {a.r + b.r, a.i + b.i}; }
inline qc qsub(qc a, qc b) { return {a.r - b.r, a.i - b.i}; }
inline qc qmul(qc a, qc b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
inline qc qdiv(qc a, qc b) {
  __float128 d = b.r * b.r + b.i * b.i;
  return {(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}

struct Pade {
  int nfit = 0;
  std::vector<qc> z, a;
  void fit(const cd* zf, const cd* Af, int n) {
    nfit = n;
    z.resize(n);
    a.resize(n);
    std::vector<qc> g(static_cast<size_t>(n) * n);
    for (int w = 0; w < n; ++w) {
      z[w] = {zf[w].real(), zf[w].imag()};
      g[w] = {Af[w].real(), Af[w].imag()};
    }
    // g(i,j) = (g(i-1,i-1) - g(i-1,j)) / ((z_j - z_{i-1}) g(i-1,j))   (pade_t.hpp:186-192)
    for (int i = 1; i < n; ++i)
      for (int j = i; j < n; ++j)
        g[i * n + j] = qdiv(qsub(g[(i - 1) * n + (i - 1)], g[(i - 1) * n + j]),
                            qmul(qsub(z[j], z[i - 1]), g[(i - 1) * n + j]));
    for (int i = 0; i < n; ++i) a[i] = g[i * n + i];
  }
  cd eval(cd w) const {  // pade_t.hpp:109-128
    qc W{w.real(), w.imag()};
    qc one{1, 0};
    qc X = qmul(a[nfit - 1], qsub(W, z[nfit - 2]));
    for (int f = 1; f < nfit - 1; ++f) {
      int idx = nfit - f - 1;
      X = qdiv(qmul(a[idx], qsub(W, z[idx - 1])), qadd(one, X));
    }
    X = qdiv(a[0], qadd(one, X));
    return cd(double(X.r), double(X.i));
  }
};

// pade_driver::init frequency selection for a full (+-) fermionic mesh of size niw.
void pade_fit_indices(int niw, int nfit, std::vector<int>& idx) {
  int niw_pos = niw / 2, shift = niw / 2;
  idx.clear();
  int num_lchunk = niw_pos % nfit;
  if (num_lchunk != 0) {
    int step = niw_pos / nfit + 1;
    for (int i = 0; i < step * num_lchunk; i += step) idx.push_back(i + shift);
    for (int i = step * num_lchunk; i < niw_pos; i += step - 1) idx.push_back(i + shift);
  } else {
    int step = niw_pos / nfit;
    for (int i = 0; i < niw_pos; i += step) idx.push_back(i + shift);
  }
}

double qp_bisection(double vhf, const Pade& S, double mu, double eps0, double tol, double eta,
                    double* res_out) {
  auto res_f = [&](double e) { return (cd(e, eta) - vhf - S.eval(cd(e - mu, eta))).real(); };
  double eps1, eps2, delta = 0.01;
  double res = res_f(eps0);
  if (std::fabs(res) < tol) { *res_out = res; return eps0; }
  if (res >= 0) {
    eps2 = eps0; eps1 = eps0 - delta;
    double r1 = res_f(eps1);
    while (r1 > 0) { eps1 -= delta; r1 = res_f(eps1); }
  } else {
    eps1 = eps0; eps2 = eps0 + delta;
    double r2 = res_f(eps2);
    while (r2 < 0) { eps2 += delta; r2 = res_f(eps2); }
  }
  double mid = 0.5 * (eps1 + eps2);
  res = res_f(mid);
  int guard = 0;
  while (std::fabs(res) > tol && guard++ < 400) {
    if (res >= 0) eps2 = mid; else eps1 = mid;
    mid = 0.5 * (eps1 + eps2);
    res = res_f(mid);
  }
  *res_out = res;
  return mid;
}
}  // namespace

// Pade on one function: fit on (zf, Af) of length n, evaluate at w (for pinning tests).
void orc_pade_eval(int n, const cd* zf, const cd* Af, int nw, const cd* w, cd* out) {
  Pade p;
  p.fit(zf, Af, n);
  for (int i = 0; i < nw; ++i) out[i] = p.eval(w[i]);
}

// Full QP solve (ns = 1): Sigma_tkij (nt_f, nk, no, no), C (nk, no, no) MO coefficients
// (columns = orbitals), Vhf (nk, no, no) in the primary basis, E0 (nk, no) initial energies.
// Twt_ff (nw_f x nt_f) rescaled; wn_mesh_f (nw_f) Matsubara indices; E_out (nk, no).
void orc_solve_qp(int nt_f, int nw_f, const cd* Twt_ff, const long* wn_mesh_f, double beta,
                  int nk, int no, const cd* Sigma, const cd* C, const cd* Vhf, const double* E0,
                  double mu, int nfit, double eta, double tol, double* E_out) {
  std::vector<cd> St(static_cast<size_t>(nt_f) * nk * no), Sw(static_cast<size_t>(nw_f) * nk * no);
  std::vector<cd> tmp(static_cast<size_t>(no) * no), Sab(static_cast<size_t>(no) * no);
  for (int it = 0; it < nt_f; ++it)
    for (int k = 0; k < nk; ++k) {
      const cd* Ck = C + static_cast<size_t>(k) * no * no;
      zgemm('N', 'N', no, no, no, 1.0, Sigma + (static_cast<size_t>(it) * nk + k) * no * no, no, Ck, no, 0.0,
            tmp.data(), no);
      zgemm('C', 'N', no, no, no, 1.0, Ck, no, tmp.data(), no, 0.0, Sab.data(), no);
      for (int a = 0; a < no; ++a) St[(static_cast<size_t>(it) * nk + k) * no + a] = Sab[a * no + a];
    }
  zgemm('N', 'N', nw_f, long(nk) * no, nt_f, 1.0, Twt_ff, nt_f, St.data(), long(nk) * no, 0.0, Sw.data(),
        long(nk) * no);
  std::vector<cd> vd(static_cast<size_t>(nk) * no);
  for (int k = 0; k < nk; ++k) {
    const cd* Ck = C + static_cast<size_t>(k) * no * no;
    zgemm('N', 'N', no, no, no, 1.0, Vhf + static_cast<size_t>(k) * no * no, no, Ck, no, 0.0, tmp.data(), no);
    zgemm('C', 'N', no, no, no, 1.0, Ck, no, tmp.data(), no, 0.0, Sab.data(), no);
    for (int a = 0; a < no; ++a) vd[k * no + a] = Sab[a * no + a];
  }
  std::vector<int> idx;
  pade_fit_indices(nw_f, nfit, idx);
  std::vector<cd> zf(nfit), Af(nfit);
  for (int i = 0; i < nfit; ++i) zf[i] = cd(0.0, double(wn_mesh_f[idx[i]]) * M_PI / beta);
#pragma omp parallel for schedule(dynamic) firstprivate(Af)
  for (long I = 0; I < long(nk) * no; ++I) {
    for (int i = 0; i < nfit; ++i) Af[i] = Sw[static_cast<size_t>(idx[i]) * nk * no + I];
    Pade p;
    p.fit(zf.data(), Af.data(), nfit);
    double res;
    E_out[I] = qp_bisection(vd[I].real(), p, mu, E0[I], tol, eta, &res);
  }
}

// a16: Galitskii-Migdal correlation energy, scf_common.cpp:87-136 (ns = 1, k_weight = 1/N_k):
//   SigmaG(w) = sum_k w_k sum_ij Sigma_ij(iw,k) G_ji(iw,k) ; e = -0.5*2 * tau_to_beta(w_to_tau(SigmaG))
double orc_gm_energy(int nt_f, int nw_f, const cd* Twt_ff, const cd* Ttw_ff, const cd* T_beta_t_ff, int nk,
                     int no, const cd* G, const cd* Sigma) {
  long dim = long(nk) * no * no;
  std::vector<cd> Gw(nw_f * dim), Sw(nw_f * dim);
  zgemm('N', 'N', nw_f, dim, nt_f, 1.0, Twt_ff, nt_f, G, dim, 0.0, Gw.data(), dim);
  zgemm('N', 'N', nw_f, dim, nt_f, 1.0, Twt_ff, nt_f, Sigma, dim, 0.0, Sw.data(), dim);
  std::vector<cd> SG(nw_f, cd(0.0));
  for (int w = 0; w < nw_f; ++w)
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < no; ++i)
        for (int j = 0; j < no; ++j)
          SG[w] += (1.0 / nk) * Sw[(static_cast<size_t>(w) * nk + k) * no * no + i * no + j] *
                   Gw[(static_cast<size_t>(w) * nk + k) * no * no + j * no + i];
  cd e = 0.0;
  for (int it = 0; it < nt_f; ++it) {
    cd st = 0.0;
    for (int w = 0; w < nw_f; ++w) st += Ttw_ff[it * nw_f + w] * SG[w];
    e += T_beta_t_ff[it] * st;
  }
  return (-0.5 * 2.0 * e).real();
}

}  // extern "C"
