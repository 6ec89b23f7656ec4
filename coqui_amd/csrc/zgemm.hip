// Batched complex<double> GEMM on the CDNA4 FP64 matrix cores.
//
//   C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b] + gamma * I,   op in {N, T, C}
//
// All matrices are row-major with explicit leading dimensions.  One 256-thread
// workgroup (four waves, 2 x 2) computes a BM x BN tile of C; each wave owns a
// (BM/2) x (BN/2) sub-tile made of v_mfma_f64_16x16x4_f64 blocks with separate real
// and imaginary accumulators (a complex multiply-add is four real MFMAs).  Two tile
// shapes are built: 128 x 128 (4 x 4 MFMA blocks per wave, accumulators in 256 AGPRs,
// one wave per SIMD) for large products and 64 x 64 (2 x 2 blocks) for small ones.
// K is staged through LDS in chunks of 8 complex values, double buffered, so the next
// chunk's global loads overlap the current chunk's MFMAs.
//
// Workgroups are dispatched round-robin over the 8 XCDs; the block index is remapped
// so that each XCD receives a contiguous range of (batch, tile) work.  The tiles of one
// matrix then share their A row panels and B column panels through that XCD's L2.
#include "thcgw_internal.hpp"

#include <atomic>
#include <mutex>
#include <vector>

namespace thcgw {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int BK = 8;
constexpr int NT = 256;
constexpr int NXCD = 8;

struct Args {
  int M, N, K, batch, tiles_m, tiles_n;
  double ar, ai, br, bi, gr, gi;
  BatchPtr A, B;
  long lda, ldb;
  c128* C;
  long strideC, ldc;
};

__device__ __forceinline__ const c128* batch_ptr(const BatchPtr& p, int b) {
  return p.p + static_cast<long>(((b + p.off) / p.div) % p.mod) * p.stride;
}

__device__ __forceinline__ c128 ld_elem(const c128* X, long ld, int r, int c, int R, int Cn, bool conj) {
  if (r < R && c < Cn) {
    c128 v = X[static_cast<long>(r) * ld + c];
    if (conj) v.y = -v.y;
    return v;
  }
  return c128{0.0, 0.0};
}

// Tile of an operand: BX elements along its free dimension x (m for op(A), n for op(B))
// times BK along k, staged in LDS as S[k][x].  ROWS: element (x, k) is stored at X[x][k]
// (op(A) = A, op(B) = B^T or B^H); otherwise at X[k][x] (op(A) = A^T or A^H, op(B) = B).
// CONJ conjugates on load.  Pairs of adjacent elements along the contiguous direction
// are loaded together; each thread loads P = BX*BK/(2*NT) pairs.
template <bool ROWS, bool CONJ, int BX>
struct TileIO {
  static constexpr int P = BX * BK / (2 * NT);
  __device__ static void coords(int t, int p, int& x, int& k) {
    const int q = t + p * NT;  // pair id
    if (ROWS) {                // contiguous along k: BK/2 pairs per x
      x = q / (BK / 2);
      k = (q % (BK / 2)) * 2;
    } else {  // contiguous along x: BX/2 pairs per k
      k = q / (BX / 2);
      x = (q % (BX / 2)) * 2;
    }
  }
  __device__ static void load(const c128* X, long ld, int nx, int nk, int x0, int k0, c128 (&r)[2 * P]) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      int x, k;
      coords(threadIdx.x, p, x, k);
      if (ROWS) {
        r[2 * p] = ld_elem(X, ld, x0 + x, k0 + k, nx, nk, CONJ);
        r[2 * p + 1] = ld_elem(X, ld, x0 + x, k0 + k + 1, nx, nk, CONJ);
      } else {
        r[2 * p] = ld_elem(X, ld, k0 + k, x0 + x, nk, nx, CONJ);
        r[2 * p + 1] = ld_elem(X, ld, k0 + k, x0 + x + 1, nk, nx, CONJ);
      }
    }
  }
  __device__ static void store(c128 (*S)[BX], const c128 (&r)[2 * P]) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      int x, k;
      coords(threadIdx.x, p, x, k);
      if (ROWS) {
        S[k][x] = r[2 * p];
        S[k + 1][x] = r[2 * p + 1];
      } else {
        S[k][x] = r[2 * p];
        S[k][x + 1] = r[2 * p + 1];
      }
    }
  }
};

template <int OPA, int OPB, int BM, int BN>
__global__ __launch_bounds__(NT) void zgemm_kernel(Args p) {
  constexpr int TM = BM / 32, TN = BN / 32;  // 16x16 MFMA blocks per wave (waves are 2 x 2)
  using IOA = TileIO<OPA == OP_N, OPA == OP_C, BM>;  // op(A)(m,k): A[m][k] or A[k][m]
  using IOB = TileIO<OPB != OP_N, OPB == OP_C, BN>;  // op(B)(k,n): B[k][n] or B[n][k]
  __shared__ c128 As[2][BK][BM];
  __shared__ c128 Bs[2][BK][BN];

  // XCD-aware remap of the hardware block id to a logical (batch, tile) index.
  const int nblk = gridDim.x;
  const int hw = blockIdx.x;
  const int xcd = hw % NXCD, loc = hw / NXCD;
  const int base = nblk / NXCD, rem = nblk % NXCD;
  const int logical = xcd * base + (xcd < rem ? xcd : rem) + loc;

  const int tiles = p.tiles_m * p.tiles_n;
  const int b = logical / tiles;
  const int tile = logical % tiles;
  const int m0 = (tile / p.tiles_n) * BM;
  const int n0 = (tile % p.tiles_n) * BN;

  const c128* A = batch_ptr(p.A, b);
  const c128* B = batch_ptr(p.B, b);
  c128* C = p.C + static_cast<long>(b) * p.strideC;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  f64x4 cr[TM][TN], ci[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      cr[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
      ci[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    }

  const int nk = (p.K + BK - 1) / BK;
  c128 ra[2 * IOA::P], rb[2 * IOB::P];
  if (nk > 0) {
    IOA::load(A, p.lda, p.M, p.K, m0, 0, ra);
    IOB::load(B, p.ldb, p.N, p.K, n0, 0, rb);
    IOA::store(As[0], ra);
    IOB::store(Bs[0], rb);
  }
  __syncthreads();

  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) {
      IOA::load(A, p.lda, p.M, p.K, m0, (kt + 1) * BK, ra);
      IOB::load(B, p.ldb, p.N, p.K, n0, (kt + 1) * BK, rb);
    }
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      double a_r[TM], a_i[TM], a_n[TM], b_r[TN], b_i[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const c128 v = As[cur][ks * 4 + lk][wm * (BM / 2) + i * 16 + lr];
        a_r[i] = v.x;
        a_i[i] = v.y;
        a_n[i] = -v.y;
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const c128 v = Bs[cur][ks * 4 + lk][wn * (BN / 2) + j * 16 + lr];
        b_r[j] = v.x;
        b_i[j] = v.y;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_r[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_i[j], ci[i][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_n[i], b_i[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[i], b_r[j], ci[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) {
      IOA::store(As[cur ^ 1], ra);
      IOB::store(Bs[cur ^ 1], rb);
    }
    __syncthreads();
  }

  // Epilogue: for v_mfma_f64_16x16x4_f64 accumulator register v of lane l holds row
  // l/16 + 4v and column l%16 of its 16 x 16 block (checked on gfx950).
  const bool beta0 = (p.br == 0.0 && p.bi == 0.0);
  const bool gamma0 = (p.gr == 0.0 && p.gi == 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = m0 + wm * (BM / 2) + i * 16 + lk + 4 * v;
        const int c = n0 + wn * (BN / 2) + j * 16 + lr;
        if (r < p.M && c < p.N) {
          const double xr = cr[i][j][v], xi = ci[i][j][v];
          double outr = p.ar * xr - p.ai * xi;
          double outi = p.ar * xi + p.ai * xr;
          c128* dst = C + static_cast<long>(r) * p.ldc + c;
          if (!beta0) {
            const c128 old = *dst;
            outr += p.br * old.x - p.bi * old.y;
            outi += p.br * old.y + p.bi * old.x;
          }
          if (!gamma0 && r == c) {
            outr += p.gr;
            outi += p.gi;
          }
          *dst = c128{outr, outi};
        }
      }
}

std::atomic<long> g_launches{0};
std::atomic<long long> g_flops_milli{0};

// Optional per-launch timing with HIP events recorded on the launch stream.
struct ProfRec {
  hipEvent_t a, b;
  double flops;
  int cls;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_pool;

hipEvent_t pool_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  THCGW_HIP(hipEventCreate(&e));
  return e;
}

// Tile shapes: 0 = 64 x 64, 1 = 128 x 64, 2 = 128 x 128.
template <int OPA, int OPB>
void launch(hipStream_t s, Args a, int shape) {
  const int bm = shape == 0 ? 64 : 128, bn = shape == 2 ? 128 : 64;
  a.tiles_m = (a.M + bm - 1) / bm;
  a.tiles_n = (a.N + bn - 1) / bn;
  const long nblocks = static_cast<long>(a.batch) * a.tiles_m * a.tiles_n;
  THCGW_REQUIRE(nblocks < (1L << 31), "zgemm: grid too large");
  const dim3 g(static_cast<unsigned>(nblocks)), t(NT);
  if (shape == 0)
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, 64, 64>), g, t, 0, s, a);
  else if (shape == 1)
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, 128, 64>), g, t, 0, s, a);
  else
    hipLaunchKernelGGL((zgemm_kernel<OPA, OPB, 128, 128>), g, t, 0, s, a);
}

std::atomic<int> g_tile_mode{0};  // 0 = automatic, else shape + 1

// 64 x 64 tiles by default: at three waves per SIMD they hide load latency better than
// the larger shapes at one wave per SIMD (measured on cfg2: 45.8 vs 26.4 TFLOP/s).
int tile_shape(int, int, int) {
  const int forced = g_tile_mode.load();
  return forced > 0 ? forced - 1 : 0;
}

}  // namespace

void zgemm(hipStream_t s, int opA, int opB, int M, int N, int K, Cplx alpha, BatchPtr A, long lda,
           BatchPtr B, long ldb, Cplx beta, c128* C, long strideC, long ldc, int batch, Cplx gamma) {
  THCGW_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "zgemm: negative size");
  THCGW_REQUIRE(opA >= 0 && opA <= 2 && opB >= 0 && opB <= 2, "zgemm: bad op");
  if (M == 0 || N == 0 || batch == 0) return;
  Args a;
  a.M = M;
  a.N = N;
  a.K = K;
  a.batch = batch;
  a.ar = alpha.re;
  a.ai = alpha.im;
  a.br = beta.re;
  a.bi = beta.im;
  a.gr = gamma.re;
  a.gi = gamma.im;
  a.A = A;
  a.B = B;
  a.lda = lda;
  a.ldb = ldb;
  a.C = C;
  a.strideC = strideC;
  a.ldc = ldc;
  const int nb = tile_shape(M, N, K);
  const double flops = 8.0 * double(M) * N * K * batch;
  std::unique_lock<std::mutex> lk(g_prof_mu, std::defer_lock);
  ProfRec rec{nullptr, nullptr, flops, opA * 3 + opB};
  if (g_prof_on) {
    lk.lock();
    rec.a = pool_event();
    rec.b = pool_event();
    THCGW_HIP(hipEventRecord(rec.a, s));
  }
  switch (opA * 3 + opB) {
    case 0: launch<OP_N, OP_N>(s, a, nb); break;
    case 1: launch<OP_N, OP_T>(s, a, nb); break;
    case 2: launch<OP_N, OP_C>(s, a, nb); break;
    case 3: launch<OP_T, OP_N>(s, a, nb); break;
    case 4: launch<OP_T, OP_T>(s, a, nb); break;
    case 5: launch<OP_T, OP_C>(s, a, nb); break;
    case 6: launch<OP_C, OP_N>(s, a, nb); break;
    case 7: launch<OP_C, OP_T>(s, a, nb); break;
    default: launch<OP_C, OP_C>(s, a, nb); break;
  }
  THCGW_HIP(hipGetLastError());
  if (g_prof_on) {
    THCGW_HIP(hipEventRecord(rec.b, s));
    g_prof.push_back(rec);
  }
  g_launches.fetch_add(1);
  g_flops_milli.fetch_add(static_cast<long long>(flops / 1000.0));
}

GemmStats gemm_stats() { return GemmStats{g_launches.load(), 1000.0 * double(g_flops_milli.load())}; }
void gemm_stats_reset() {
  g_launches = 0;
  g_flops_milli = 0;
}

void gemm_set_tile_mode(int mode) { g_tile_mode = mode; }

void gemm_profile_enable(bool on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on;
}

// Per class (opA*3 + opB, 9 classes): summed launch durations (ms), modelled flops and
// launch counts of the recorded launches; the records are then dropped.
void gemm_profile_read(double* ms, double* flops, long* launches) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int c = 0; c < 9; ++c) {
    ms[c] = 0.0;
    flops[c] = 0.0;
    launches[c] = 0;
  }
  for (const ProfRec& r : g_prof) {
    THCGW_HIP(hipEventSynchronize(r.b));
    float e = 0.0f;
    THCGW_HIP(hipEventElapsedTime(&e, r.a, r.b));
    ms[r.cls] += e;
    flops[r.cls] += r.flops;
    launches[r.cls] += 1;
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_prof.clear();
}

}  // namespace thcgw
