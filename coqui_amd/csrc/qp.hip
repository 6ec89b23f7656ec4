// Quasiparticle equation on the device: one thread per (k, n) state.
//
// Per state (solve_qp_eqn, qp_scf_common.cpp:150-311): Sigma_nn at the Pade fit frequencies
// from the diagonal Sigma_nn(tau) through the rows of Twt_ff (IAFT.hpp:55-57), a Thiele
// continued-fraction Pade through those points (pade_t.hpp:34-168, coefficients by the
// recursion of pade_t.hpp:176-192), and the root of
//   res(e) = Re[e + i eta - Vhf_nn - Sigma_nn(e - mu + i eta)]
// by the bracketing and bisection of qp_eqn_bisection (qp_scf_common.cpp:679-728).
// The reference runs the Pade in 100-digit arithmetic; here the fit and its evaluation run
// in double-double (Hida, Li & Bailey, ARITH-15, 2001), which the tests check against a
// binary128 restatement.
#include "thcgw_internal.hpp"

namespace thcgw {
namespace {

struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b, bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ dd quick_two_sum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
__device__ __forceinline__ dd two_prod(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
__device__ __forceinline__ dd add(dd x, dd y) {
  dd s = two_sum(x.hi, y.hi), t = two_sum(x.lo, y.lo);
  s.lo += t.hi;
  s = quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return quick_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ dd neg(dd x) { return {-x.hi, -x.lo}; }
__device__ __forceinline__ dd sub(dd x, dd y) { return add(x, neg(y)); }
__device__ __forceinline__ dd mul(dd x, dd y) {
  dd p = two_prod(x.hi, y.hi);
  p.lo += x.hi * y.lo + x.lo * y.hi;
  return quick_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd mul_d(dd x, double y) {
  dd p = two_prod(x.hi, y);
  p.lo += x.lo * y;
  return quick_two_sum(p.hi, p.lo);
}
__device__ __forceinline__ dd div(dd x, dd y) {
  const double q1 = x.hi / y.hi;
  dd r = sub(x, mul_d(y, q1));
  const double q2 = r.hi / y.hi;
  r = sub(r, mul_d(y, q2));
  const double q3 = r.hi / y.hi;
  return add(quick_two_sum(q1, q2), dd{q3, 0.0});
}

struct zdd {
  dd re, im;
};
__device__ __forceinline__ zdd zadd(zdd a, zdd b) { return {add(a.re, b.re), add(a.im, b.im)}; }
__device__ __forceinline__ zdd zsub(zdd a, zdd b) { return {sub(a.re, b.re), sub(a.im, b.im)}; }
__device__ __forceinline__ zdd zmul(zdd a, zdd b) {
  return {sub(mul(a.re, b.re), mul(a.im, b.im)), add(mul(a.re, b.im), mul(a.im, b.re))};
}
__device__ __forceinline__ zdd zdiv(zdd a, zdd b) {
  const dd den = add(mul(b.re, b.re), mul(b.im, b.im));
  return {div(add(mul(a.re, b.re), mul(a.im, b.im)), den), div(sub(mul(a.im, b.re), mul(a.re, b.im)), den)};
}
__device__ __forceinline__ zdd zd(double re, double im) { return {dd{re, 0.0}, dd{im, 0.0}}; }

// A(w) = a_0 / (1 + a_1 (w - z_0) / (1 + a_2 (w - z_1) / (1 + ...)))   (pade_t.hpp:109-128)
__device__ double2 pade_eval(const zdd* z, const zdd* a, int n, double wr, double wi) {
  const zdd W = zd(wr, wi), one = zd(1.0, 0.0);
  zdd X = zmul(a[n - 1], zsub(W, z[n - 2]));
  for (int f = 1; f < n - 1; ++f) {
    const int idx = n - f - 1;
    X = zdiv(zmul(a[idx], zsub(W, z[idx - 1])), zadd(one, X));
  }
  X = zdiv(a[0], zadd(one, X));
  return double2{X.re.hi + X.re.lo, X.im.hi + X.im.lo};
}

__global__ __launch_bounds__(64) void qp_solve_kernel(const c128* __restrict__ sig_t, const c128* __restrict__ vhf,
                                                      const double* __restrict__ E0, const c128* __restrict__ Twt,
                                                      QpFit fit, int nt, long dim, double mu, double eta, double tol,
                                                      double* __restrict__ E_out) {
  const long I = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (I >= dim) return;
  const int n = fit.nfit;
  zdd z[QP_MAX_FIT], a[QP_MAX_FIT], g[QP_MAX_FIT];
  // Sigma(iw_n) = Twt_ff Sigma(tau) at the fit frequencies; z_i = i w_{n_i}
  for (int i = 0; i < n; ++i) {
    const c128* row = Twt + static_cast<long>(fit.w[i]) * nt;
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < nt; ++t) {
      const c128 T = row[t], S = sig_t[static_cast<long>(t) * dim + I];
      sr += T.x * S.x - T.y * S.y;
      si += T.x * S.y + T.y * S.x;
    }
    g[i] = zd(sr, si);
    z[i] = zd(0.0, fit.zim[i]);
  }
  // Thiele coefficients a_i = g_i(z_i): g_i(z_j) = (g_{i-1}(z_{i-1}) - g_{i-1}(z_j)) / ((z_j - z_{i-1}) g_{i-1}(z_j))
  a[0] = g[0];
  for (int i = 1; i < n; ++i) {
    const zdd gd = g[i - 1];
    for (int j = n - 1; j >= i; --j) g[j] = zdiv(zsub(gd, g[j]), zmul(zsub(z[j], z[i - 1]), g[j]));
    a[i] = g[i];
  }
  const double v = vhf[I].x;
  auto res_f = [&](double e) {
    const double2 s = pade_eval(z, a, n, e - mu, eta);
    return e - v - s.x;
  };
  // qp_eqn_bisection: bracket in steps of 0.01 Ha from the starting energy, then bisect
  const double eps0 = E0[I], delta = 0.01;
  double res = res_f(eps0);
  if (fabs(res) < tol) {
    E_out[I] = eps0;
    return;
  }
  double e1, e2;
  int guard = 0;
  if (res >= 0) {
    e2 = eps0;
    e1 = eps0 - delta;
    double r1 = res_f(e1);
    while (r1 > 0 && guard++ < 100000) {
      e1 -= delta;
      r1 = res_f(e1);
    }
  } else {
    e1 = eps0;
    e2 = eps0 + delta;
    double r2 = res_f(e2);
    while (r2 < 0 && guard++ < 100000) {
      e2 += delta;
      r2 = res_f(e2);
    }
  }
  double mid = 0.5 * (e1 + e2);
  res = res_f(mid);
  guard = 0;
  while (fabs(res) > tol && guard++ < 200) {
    if (res >= 0) e2 = mid;
    else e1 = mid;
    mid = 0.5 * (e1 + e2);
    res = res_f(mid);
  }
  E_out[I] = mid;
}

}  // namespace

void qp_solve(hipStream_t s, const c128* sig_t, const c128* vhf, const double* E0, const c128* Twt, const QpFit& fit,
              int nt, long dim, double mu, double eta, double tol, double* E_out) {
  THCGW_REQUIRE(fit.nfit >= 2 && fit.nfit <= QP_MAX_FIT, "qp_solve: nfit must be in [2, 64]");
  if (dim <= 0) return;
  hipLaunchKernelGGL(qp_solve_kernel, dim3(static_cast<unsigned>((dim + 63) / 64)), dim3(64), 0, s, sig_t, vhf, E0,
                     Twt, fit, nt, dim, mu, eta, tol, E_out);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
