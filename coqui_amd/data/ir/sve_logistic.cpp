// Singular value expansion of the logistic kernel in binary128 arithmetic.
//
//   K(x, y) = exp(-L y (x + 1) / 2) / (1 + exp(-L y)),   x, y in [-1, 1],  L = Lambda = beta wmax
//
// is the kernel of the intermediate representation (IR) of imaginary-time Green's functions
// (Shinaoka et al., Phys. Rev. B 96, 035147 (2017); Chikano et al., Comput. Phys. Commun.
// 240, 181 (2019)).  K(-x, -y) = K(x, y), so the singular functions are even or odd and the
// expansion splits into the two kernels on [0, 1]^2
//
//   K_even/odd(x, y) = K(x, y) +/- K(x, -y)
//                    = [exp(-L y (1 + x) / 2) +/- exp(-L y (1 - x) / 2)] / (1 + exp(-L y)).
//
// Each is discretised by composite Gauss-Legendre quadrature (A_ij = sqrt(w_i) K(x_i, y_j)
// sqrt(w_j)) on segments graded geometrically towards x = 1 and y = 0, where the functions
// vary on the scale 1/Lambda.  The SVD of A is a Householder QR with column pivoting,
// truncated where the remaining columns fall below 1e-33 of the largest, followed by a
// one-sided Jacobi SVD of the small triangular factor.  Everything runs in __float128, so
// singular values down to ~1e-30 of the first and their singular functions are resolved.
//
// Output (binary, little endian): int32 n_seg, n_gauss, n_func; float64 knots[n_seg + 1]
// (segment ends on [0, 1]); then per function, sorted by decreasing singular value:
// float64 parity (+1 even, -1 odd), float64 s, float64 c[n_seg][n_gauss], the Legendre
// coefficients of u(x) on each segment (local variable t in [-1, 1]) for the function
// normalised on [0, 1]; the function on [-1, 1] is u(|x|) (odd: sign(x) u(|x|)) / sqrt(2).
//
// Build and run: g++ -O2 -fopenmp sve_logistic.cpp -lquadmath -o sve_logistic
//                ./sve_logistic LAMBDA OUT.bin [n_gauss]
#include <quadmath.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <vector>

typedef __float128 q;

namespace {

q qabs(q a) { return a < 0 ? -a : a; }

// Gauss-Legendre nodes and weights on [-1, 1] by Newton iteration on P_n.
void gauss_legendre(int n, std::vector<q>& t, std::vector<q>& w) {
  t.assign(n, 0);
  w.assign(n, 0);
  const q pi = M_PIq;
  for (int i = 0; i < n; ++i) {
    q x = cosq(pi * (i + 0.75Q) / (n + 0.5Q));
    for (int it = 0; it < 100; ++it) {
      q p0 = 1, p1 = x;
      for (int k = 2; k <= n; ++k) {
        q p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const q dp = n * (x * p1 - p0) / (x * x - 1);
      const q dx = p1 / dp;
      x -= dx;
      if (qabs(dx) < 1e-34Q) break;
    }
    q p0 = 1, p1 = x;
    for (int k = 2; k <= n; ++k) {
      q p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    const q dp = n * (x * p1 - p0) / (x * x - 1);
    t[n - 1 - i] = x;  // ascending
    w[n - 1 - i] = 2 / ((1 - x * x) * dp * dp);
  }
}

// Segment ends on [0, 1].  x: uniform steps of 1/16 on [0, 1/2], then geometric towards 1
// (ends 1 - 2^-k) down to a width below 0.005 / Lambda.  y: geometric from 0 (ends 2^-k, from
// a width below 0.005 / Lambda) up to 1/2, then uniform steps of 1/16.
std::vector<q> segments(double lambda, bool towards_one) {
  std::vector<q> g;
  int K = 1;
  while (std::ldexp(1.0, -K) > 0.005 / lambda) ++K;
  if (towards_one) {
    for (int j = 0; j < 8; ++j) g.push_back(q(j) / 16);
    for (int k = 1; k <= K; ++k) g.push_back(1 - ldexpq(1, -k));
    g.push_back(1);
  } else {
    g.push_back(0);
    for (int k = K; k >= 2; --k) g.push_back(ldexpq(1, -k));
    for (int j = 8; j <= 16; ++j) g.push_back(q(j) / 16);
  }
  return g;
}

struct Grid {
  std::vector<q> x, w;  // nodes and weights on [0, 1]
};

Grid composite(const std::vector<q>& ends, const std::vector<q>& t, const std::vector<q>& wt) {
  Grid g;
  for (size_t s = 0; s + 1 < ends.size(); ++s) {
    const q a = ends[s], b = ends[s + 1], h = (b - a) / 2, c = (a + b) / 2;
    for (size_t i = 0; i < t.size(); ++i) {
      g.x.push_back(c + h * t[i]);
      g.w.push_back(h * wt[i]);
    }
  }
  return g;
}

q kernel_part(q lam, q x, q y, int parity) {
  // [exp(-L y (1 + x)/2) +/- exp(-L y (1 - x)/2)] / (1 + exp(-L y)), y >= 0
  const q den = 1 + expq(-lam * y);
  const q a = expq(-lam * y * (1 + x) / 2), b = expq(-lam * y * (1 - x) / 2);
  if (parity > 0) return (a + b) / den;
  // a - b = -2 exp(-L y / 2) sinh(L y x / 2), without cancellation for small L y x
  return -2 * expq(-lam * y / 2) * sinhq(lam * y * x / 2) / den;
}

struct Func {
  int parity;
  q s;
  std::vector<q> u;  // values at the x nodes, normalised on [0, 1]
};

// SVD of the column-major nx x ny matrix A (destroyed): appends the singular triplets' x
// sides to out.
void svd_qrcp_jacobi(std::vector<q>& A, int nx, int ny, int parity, std::vector<Func>& out) {
  // Householder QR with column pivoting, truncated at 1e-33 of the largest column norm
  std::vector<int> perm(ny);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<q> vs;          // reflectors (column k has length nx - k), packed
  std::vector<q> taus;
  std::vector<long> voff;
  q norm0 = 0;
  for (int j = 0; j < ny; ++j) {
    q s = 0;
    for (int i = 0; i < nx; ++i) s += A[(long)j * nx + i] * A[(long)j * nx + i];
    norm0 = std::max(norm0, sqrtq(s));
  }
  const int kmax = std::min(nx, ny);
  int r = 0;
  std::vector<q> cn(ny);
  for (int k = 0; k < kmax; ++k) {
#pragma omp parallel for
    for (int j = k; j < ny; ++j) {
      q s = 0;
      for (int i = k; i < nx; ++i) s += A[(long)j * nx + i] * A[(long)j * nx + i];
      cn[j] = s;
    }
    int jm = k;
    for (int j = k + 1; j < ny; ++j)
      if (cn[j] > cn[jm]) jm = j;
    if (sqrtq(cn[jm]) < 1e-33Q * norm0) break;
    if (jm != k) {
      for (int i = 0; i < nx; ++i) std::swap(A[(long)k * nx + i], A[(long)jm * nx + i]);
      std::swap(perm[k], perm[jm]);
    }
    // reflector for A[k:, k]
    q* col = &A[(long)k * nx];
    const q alpha = col[k];
    const q nrm = sqrtq(cn[jm]);
    const q beta = alpha > 0 ? -nrm : nrm;
    std::vector<q> v(nx - k);
    v[0] = alpha - beta;
    for (int i = k + 1; i < nx; ++i) v[i - k] = col[i];
    q vn = 0;
    for (q e : v) vn += e * e;
    const q tau = vn > 0 ? 2 / vn : 0;
    col[k] = beta;
    for (int i = k + 1; i < nx; ++i) col[i] = 0;
#pragma omp parallel for
    for (int j = k + 1; j < ny; ++j) {
      q* c = &A[(long)j * nx];
      q d = 0;
      for (int i = k; i < nx; ++i) d += v[i - k] * c[i];
      d *= tau;
      for (int i = k; i < nx; ++i) c[i] -= d * v[i - k];
    }
    voff.push_back(vs.size());
    vs.insert(vs.end(), v.begin(), v.end());
    taus.push_back(tau);
    r = k + 1;
  }
  // C = R (r x ny, row major in c[i * ny + j]); M = C^T is ny x r: column p of M is row p of C
  std::vector<q> Mt((long)r * ny);  // Mt[p * ny + j] = C[p][j]
  for (int p = 0; p < r; ++p)
    for (int j = 0; j < ny; ++j) Mt[(long)p * ny + j] = j >= p ? A[(long)j * nx + p] : 0;
  std::vector<q> W((long)r * r, 0);  // column p of W at W[p * r ...]
  for (int p = 0; p < r; ++p) W[(long)p * r + p] = 1;
  // one-sided Jacobi on the columns of M (rows of C), cyclic by rows
  for (int sweep = 0; sweep < 60; ++sweep) {
    long rotations = 0;
    for (int p = 0; p < r - 1; ++p) {
      for (int qq = p + 1; qq < r; ++qq) {
        q* mp = &Mt[(long)p * ny];
        q* mq = &Mt[(long)qq * ny];
        q a = 0, b = 0, g = 0;
        for (int j = 0; j < ny; ++j) {
          a += mp[j] * mp[j];
          b += mq[j] * mq[j];
          g += mp[j] * mq[j];
        }
        if (qabs(g) <= 1e-33Q * sqrtq(a * b) || g == 0) continue;
        ++rotations;
        const q zeta = (b - a) / (2 * g);
        const q t = (zeta >= 0 ? 1 : -1) / (qabs(zeta) + sqrtq(1 + zeta * zeta));
        const q c = 1 / sqrtq(1 + t * t), s = c * t;
        for (int j = 0; j < ny; ++j) {
          const q x = mp[j], y = mq[j];
          mp[j] = c * x - s * y;
          mq[j] = s * x + c * y;
        }
        q* wp = &W[(long)p * r];
        q* wq = &W[(long)qq * r];
        for (int i = 0; i < r; ++i) {
          const q x = wp[i], y = wq[i];
          wp[i] = c * x - s * y;
          wq[i] = s * x + c * y;
        }
      }
    }
    std::fprintf(stderr, "sve: parity %+d, rank %d, sweep %d: %ld rotations\n", parity, r, sweep, rotations);
    if (rotations == 0) break;
  }
  // singular values: column norms of M; left vectors: Q W
  for (int p = 0; p < r; ++p) {
    q s = 0;
    for (int j = 0; j < ny; ++j) s += Mt[(long)p * ny + j] * Mt[(long)p * ny + j];
    Func f;
    f.parity = parity;
    f.s = sqrtq(s);
    // u = Q w_p: apply the reflectors H_0 ... H_{r-1} to (w_p, 0)
    std::vector<q> u(nx, 0);
    for (int i = 0; i < r; ++i) u[i] = W[(long)p * r + i];
    for (int k = r - 1; k >= 0; --k) {
      const q* v = &vs[voff[k]];
      q d = 0;
      for (int i = k; i < nx; ++i) d += v[i - k] * u[i];
      d *= taus[k];
      for (int i = k; i < nx; ++i) u[i] -= d * v[i - k];
    }
    f.u = u;
    out.push_back(std::move(f));
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s LAMBDA OUT.bin [n_gauss]\n", argv[0]);
    return 2;
  }
  const double lambda = std::atof(argv[1]);
  const int ng = argc > 3 ? std::atoi(argv[3]) : 32;
  if (!(lambda > 0) || ng < 4 || ng > 128) {
    std::fprintf(stderr, "sve: bad arguments\n");
    return 2;
  }
  std::vector<q> t, wt;
  gauss_legendre(ng, t, wt);
  const std::vector<q> xe = segments(lambda, true), ye = segments(lambda, false);
  const Grid gx = composite(xe, t, wt), gy = composite(ye, t, wt);
  const int nx = gx.x.size(), ny = gy.x.size();
  std::fprintf(stderr, "sve: Lambda %g, %d x %d quadrature points (%d Gauss points per segment)\n", lambda, nx, ny, ng);
  const q lam = lambda;
  std::vector<Func> funcs;
  for (int parity : {+1, -1}) {
    std::vector<q> A((long)nx * ny);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        A[(long)j * nx + i] = sqrtq(gx.w[i]) * kernel_part(lam, gx.x[i], gy.x[j], parity) * sqrtq(gy.w[j]);
    svd_qrcp_jacobi(A, nx, ny, parity, funcs);
  }
  std::sort(funcs.begin(), funcs.end(), [](const Func& a, const Func& b) { return a.s > b.s; });
  // nodal values u(x_i) = U_i / sqrt(w_i) -> Legendre coefficients per segment:
  // c_k = (2k + 1)/2 sum_i wt_i P_k(t_i) u(x_i)
  const int nseg = xe.size() - 1;
  std::vector<q> P((long)ng * ng);  // P[k * ng + i] = P_k(t_i)
  for (int i = 0; i < ng; ++i) {
    q p0 = 1, p1 = t[i];
    P[i] = 1;
    if (ng > 1) P[ng + i] = t[i];
    for (int k = 2; k < ng; ++k) {
      const q p2 = ((2 * k - 1) * t[i] * p1 - (k - 1) * p0) / k;
      P[(long)k * ng + i] = p2;
      p0 = p1;
      p1 = p2;
    }
  }
  FILE* fo = std::fopen(argv[2], "wb");
  if (!fo) {
    std::fprintf(stderr, "sve: cannot open %s\n", argv[2]);
    return 1;
  }
  const int32_t hdr[3] = {nseg, ng, (int32_t)funcs.size()};
  std::fwrite(hdr, sizeof(int32_t), 3, fo);
  for (const q& e : xe) {
    const double d = (double)e;
    std::fwrite(&d, sizeof(double), 1, fo);
  }
  std::vector<double> coef((long)nseg * ng);
  for (const Func& f : funcs) {
    for (int sgm = 0; sgm < nseg; ++sgm)
      for (int k = 0; k < ng; ++k) {
        q c = 0;
        for (int i = 0; i < ng; ++i) {
          const long n = (long)sgm * ng + i;
          c += wt[i] * P[(long)k * ng + i] * f.u[n] / sqrtq(gx.w[n]);
        }
        coef[(long)sgm * ng + k] = (double)(c * (2 * k + 1) / 2);
      }
    const double par = f.parity, s = (double)f.s;
    std::fwrite(&par, sizeof(double), 1, fo);
    std::fwrite(&s, sizeof(double), 1, fo);
    std::fwrite(coef.data(), sizeof(double), coef.size(), fo);
  }
  std::fclose(fo);
  std::fprintf(stderr, "sve: %zu singular functions, s_max %.6e, s_min/s_max %.3e\n", funcs.size(),
               (double)funcs.front().s, (double)(funcs.back().s / funcs.front().s));
  return 0;
}
