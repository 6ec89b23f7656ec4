// Batched complex GEMM on the CDNA4 matrix cores, FP64 (v_mfma_f64_16x16x4_f64) or FP32
// (v_mfma_f32_16x16x4_f32) arithmetic, with operands and result stored as complex<double>
// or complex<float> independently of the arithmetic (conversion happens on the way
// through LDS and in the epilogue).
//
//   C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b] + gamma * I,   op in {N, T, C}
//
// All matrices are row-major with explicit leading dimensions.  One 256-thread
// workgroup (four waves, 2 x 2) computes a BM x BN tile of C; each wave owns a
// (BM/2) x (BN/2) sub-tile made of v_mfma_f64_16x16x4_f64 blocks with separate real
// and imaginary accumulators (a complex multiply-add is four real MFMAs, or three in the
// G3 variant: Ar Br, Ai Bi and (Ar + Ai)(Br + Bi), combined after the K loop).  Two tile
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
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;
constexpr int NXCD = 8;

struct Args {
  int M, N, K, batch, tiles_m, tiles_n, tri;
  double ar, ai, br, bi, gr, gi;
  BatchPtr A, B;
  BatchPtr H;  // GEMM_HERM_SUB: complex<double> matrix subtracted from the Hermitian result
  long lda, ldb;
  void* C;
  long strideC, ldc;
};

// Arithmetic traits: complex register type, accumulator, MFMA and accumulator layout.
template <typename R>
struct Arith;
template <>
struct Arith<double> {
  using cplx = c128;
  using acc = f64x4;
  static constexpr int BK = 8;     // K chunk staged per LDS buffer
  static constexpr int WAVES = 3;  // waves per SIMD the register budget is held to
  __device__ static acc mfma(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  // accumulator register v of lane l holds row l/16 + 4v, column l%16 (checked on gfx950)
  __device__ static int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Arith<float> {
  using cplx = c64;
  using acc = f32x4;
  static constexpr int BK = 16;
  static constexpr int WAVES = 4;
  __device__ static acc mfma(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  // accumulator register v of lane l holds row 4(l/16) + v, column l%16 (checked on gfx950)
  __device__ static int row(int lk, int v) { return 4 * lk + v; }
};

template <typename T>
__device__ __forceinline__ const T* batch_ptr(const BatchPtr& p, int b) {
  return static_cast<const T*>(p.p) + static_cast<long>(((b + p.off) / p.div) % p.mod) * p.stride;
}

template <typename RC, typename T>
__device__ __forceinline__ RC ld_elem(const T* X, long ld, int r, int c, int R, int Cn, bool conj) {
  using S = decltype(RC{}.x);
  if (r < R && c < Cn) {
    const T v = X[static_cast<long>(r) * ld + c];
    return RC{static_cast<S>(v.x), static_cast<S>(conj ? -v.y : v.y)};
  }
  return RC{0, 0};
}

// Tile of an operand: BX elements along its free dimension x (m for op(A), n for op(B))
// times BK along k, staged in LDS as S[k][x ^ swz(k)].  ROWS: element (x, k) is stored at X[x][k]
// (op(A) = A, op(B) = B^T or B^H); otherwise at X[k][x] (op(A) = A^T or A^H, op(B) = B).
// CONJ conjugates on load.  Pairs of adjacent elements along the contiguous direction
// are loaded together; each thread loads P = BX*BK/(2*NT) pairs.
// Column swizzle of LDS row k (rows are 1-2 KB apart, so every row starts in bank 0).
//   complex<double> (16 B): a K-contiguous store group (16 lanes) writes columns 0..3 of
//     rows 0, 2, 4, 6; XOR-ing the column with 4 (k/2) gives the 16 lanes distinct bank
//     groups.  The fragment reads (16 lanes, 16 consecutive columns, one row) stay within
//     an aligned 16-column block, i.e. all 64 banks.
//   complex<float> (8 B): a store group (32 lanes) writes columns 0..3 of rows 0, 2, ..., 14
//     (XOR with 4 (k/2) separates them), and one fragment read covers rows k and k + 1 with
//     16 lanes each, so odd rows are moved to the other half of a 32-column block.
template <typename RC>
__device__ __forceinline__ int swz(int k) {
  if constexpr (sizeof(RC) == 16) return ((k >> 1) & 3) << 2;
  else return (((k >> 1) & 7) << 2) ^ ((k & 1) << 4);
}

template <typename RC, typename T, bool ROWS, bool CONJ, int BX, int BK>
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
  __device__ static void load(const T* X, long ld, int nx, int nk, int x0, int k0, RC (&r)[2 * P]) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      int x, k;
      coords(threadIdx.x, p, x, k);
      if (ROWS) {
        r[2 * p] = ld_elem<RC>(X, ld, x0 + x, k0 + k, nx, nk, CONJ);
        r[2 * p + 1] = ld_elem<RC>(X, ld, x0 + x, k0 + k + 1, nx, nk, CONJ);
      } else {
        r[2 * p] = ld_elem<RC>(X, ld, k0 + k, x0 + x, nk, nx, CONJ);
        r[2 * p + 1] = ld_elem<RC>(X, ld, k0 + k, x0 + x + 1, nk, nx, CONJ);
      }
    }
  }
  __device__ static void store(RC (*S)[BX], const RC (&r)[2 * P]) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      int x, k;
      coords(threadIdx.x, p, x, k);
      if (ROWS) {
        S[k][x ^ swz<RC>(k)] = r[2 * p];
        S[k + 1][x ^ swz<RC>(k + 1)] = r[2 * p + 1];
      } else {
        S[k][x ^ swz<RC>(k)] = r[2 * p];
        S[k][(x + 1) ^ swz<RC>(k)] = r[2 * p + 1];
      }
    }
  }
};

// FOLD (FP32 arithmetic only): after every K chunk the FP32 accumulators are added into
// FP64 accumulators and cleared, so rounding grows with the chunk length (16) instead of
// with K.
template <typename R, typename TA, typename TB, typename TC, int OPA, int OPB, int BM, int BN, bool FOLD = false,
          bool G3 = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(
    BM == 64 ? (FOLD ? (G3 ? 2 : 3) : Arith<R>::WAVES) : (BN == 64 || sizeof(R) == 4 ? 2 : 1))))
void gemm_kernel(Args p) {
  using AR = Arith<R>;
  using RC = typename AR::cplx;
  constexpr int BK = AR::BK;
  constexpr int TM = BM / 32, TN = BN / 32;  // 16x16 MFMA blocks per wave (waves are 2 x 2)
  using IOA = TileIO<RC, TA, OPA == OP_N, OPA == OP_C, BM, BK>;  // op(A)(m,k): A[m][k] or A[k][m]
  using IOB = TileIO<RC, TB, OPB != OP_N, OPB == OP_C, BN, BK>;  // op(B)(k,n): B[k][n] or B[n][k]
  __shared__ RC As[2][BK][BM];
  __shared__ RC Bs[2][BK][BN];

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

  // Triangular structure (see GEMM_TRI_*): skip tiles strictly above the diagonal and
  // start the K loop where a triangular operand's band begins.
  if ((p.tri & GEMM_TRI_LOWER) && m0 + BM - 1 < n0) return;
  int k_lo = 0;
  if (p.tri & GEMM_TRI_KFROM_ROW) k_lo = m0;
  if (p.tri & GEMM_TRI_KFROM_COL) k_lo = n0 > k_lo ? n0 : k_lo;
  k_lo = (k_lo / BK) * BK;
  int k_hi = p.K;
  if ((p.tri & GEMM_TRI_KTO_ROW) && m0 + BM < k_hi) k_hi = m0 + BM;
  if ((p.tri & GEMM_TRI_KTO_COL) && n0 + BN < k_hi) k_hi = n0 + BN;

  const TA* A = batch_ptr<TA>(p.A, b);
  const TB* B = batch_ptr<TB>(p.B, b);
  TC* C = static_cast<TC*>(p.C) + static_cast<long>(b) * p.strideC;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  // G3: three real products per complex product (Gauss): cr accumulates Re a Re b, ci
  // Im a Im b and cs (Re a + Im a)(Re b + Im b); after the K loop (with FOLD: after every K
  // chunk, before the chunk is added into the FP64 accumulators) Re c = cr - ci and
  // Im c = cs - cr - ci.
  constexpr int GT = G3 ? TM : 1, GN = G3 ? TN : 1;
  typename AR::acc cr[TM][TN], ci[TM][TN], cs[GT][GN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      cr[i][j] = typename AR::acc{0, 0, 0, 0};
      ci[i][j] = typename AR::acc{0, 0, 0, 0};
    }
#pragma unroll
  for (int i = 0; i < GT; ++i)
#pragma unroll
    for (int j = 0; j < GN; ++j) cs[i][j] = typename AR::acc{0, 0, 0, 0};
  constexpr int FT = FOLD ? TM : 1, FN = FOLD ? TN : 1;
  f64x4 fr[FT][FN], fi[FT][FN];
  if constexpr (FOLD) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        fr[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        fi[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
      }
  }

  const int nk = (k_hi + BK - 1) / BK;
  const int kt0 = k_lo / BK;
  RC ra[2 * IOA::P], rb[2 * IOB::P];
  if (kt0 < nk) {
    IOA::load(A, p.lda, p.M, p.K, m0, k_lo, ra);
    IOB::load(B, p.ldb, p.N, p.K, n0, k_lo, rb);
    IOA::store(As[0], ra);
    IOB::store(Bs[0], rb);
  }
  __syncthreads();

  for (int kt = kt0; kt < nk; ++kt) {
    const int cur = (kt - kt0) & 1;
    if (kt + 1 < nk) {
      IOA::load(A, p.lda, p.M, p.K, m0, (kt + 1) * BK, ra);
      IOB::load(B, p.ldb, p.N, p.K, n0, (kt + 1) * BK, rb);
    }
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      if constexpr (G3) {
        R a_r[TM], a_i[TM], a_s[TM], b_r[TN], b_i[TN], b_s[TN];
        const int kk = ks * 4 + lk;
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const RC v = As[cur][kk][(wm * (BM / 2) + i * 16 + lr) ^ swz<RC>(kk)];
          a_r[i] = v.x;
          a_i[i] = v.y;
          a_s[i] = v.x + v.y;
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const RC v = Bs[cur][kk][(wn * (BN / 2) + j * 16 + lr) ^ swz<RC>(kk)];
          b_r[j] = v.x;
          b_i[j] = v.y;
          b_s[j] = v.x + v.y;
        }
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            cr[i][j] = AR::mfma(a_r[i], b_r[j], cr[i][j]);
            ci[i][j] = AR::mfma(a_i[i], b_i[j], ci[i][j]);
            cs[i][j] = AR::mfma(a_s[i], b_s[j], cs[i][j]);
          }
        continue;
      }
      R a_r[TM], a_i[TM], a_n[TM], b_r[TN], b_i[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int kk = ks * 4 + lk;
        const RC v = As[cur][kk][(wm * (BM / 2) + i * 16 + lr) ^ swz<RC>(kk)];
        a_r[i] = v.x;
        a_i[i] = v.y;
        a_n[i] = -v.y;
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int kk = ks * 4 + lk;
        const RC v = Bs[cur][kk][(wn * (BN / 2) + j * 16 + lr) ^ swz<RC>(kk)];
        b_r[j] = v.x;
        b_i[j] = v.y;
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] = AR::mfma(a_r[i], b_r[j], cr[i][j]);
          ci[i][j] = AR::mfma(a_r[i], b_i[j], ci[i][j]);
        }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          cr[i][j] = AR::mfma(a_n[i], b_i[j], cr[i][j]);
          ci[i][j] = AR::mfma(a_i[i], b_r[j], ci[i][j]);
        }
    }
    if constexpr (FOLD) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            if constexpr (G3) {
              const R t1 = cr[i][j][v], t2 = ci[i][j][v];
              fr[i][j][v] += static_cast<double>(t1 - t2);
              fi[i][j][v] += static_cast<double>(cs[i][j][v] - t1 - t2);
              cs[i][j][v] = 0;
            } else {
              fr[i][j][v] += static_cast<double>(cr[i][j][v]);
              fi[i][j][v] += static_cast<double>(ci[i][j][v]);
            }
            cr[i][j][v] = 0;
            ci[i][j][v] = 0;
          }
    }
    if (kt + 1 < nk) {
      IOA::store(As[cur ^ 1], ra);
      IOB::store(Bs[cur ^ 1], rb);
    }
    __syncthreads();
  }

  if constexpr (G3 && !FOLD) {
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const R t1 = cr[i][j][v], t2 = ci[i][j][v];
          cr[i][j][v] = t1 - t2;
          ci[i][j][v] = cs[i][j][v] - t1 - t2;
        }
  }

  // Epilogue (in double whatever the arithmetic).
  if (p.tri & GEMM_HERM_SUB) {
    // Lower elements (r >= c) of a Hermitian product: C = X - H on both triangles (C = X when
    // no H is given).  The lower element goes straight out.  The mirrored one, conj(X[r][c]) -
    // H[c][r]: in FP64 arithmetic a lane's four rows are consecutive, so the mirrored stores
    // already form 64-byte runs and go straight out too (an LDS transpose measured 7% slower
    // on the order-1600 Z Z^H).  In FP32 arithmetic the rows are 4 apart and every mirrored
    // element would be its own access: it goes through a wave-private 16 x 17 LDS block per
    // 16 x 16 accumulator block, so that its store and the read of H run along rows (the
    // N_mu = 1600 projections: 154.8 -> 140.1 ms per iteration).  The K loop's staging buffers
    // are free here (its last barrier has passed): waves 0-1 use As, 2-3 Bs.
    using S = decltype(TC{}.x);
    const c128* Hb = p.H.p ? batch_ptr<c128>(p.H, b) : nullptr;
    if constexpr (sizeof(R) == 8) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = m0 + wm * (BM / 2) + i * 16 + AR::row(lk, v);
            const int c = n0 + wn * (BN / 2) + j * 16 + lr;
            if (r < p.M && c < p.N && r >= c) {
              const double xr = cr[i][j][v], xi = ci[i][j][v];
              const double outr = p.ar * xr - p.ai * xi, outi = p.ar * xi + p.ai * xr;
              const c128 h = Hb ? Hb[static_cast<long>(r) * p.ldc + c] : c128{0.0, 0.0};
              C[static_cast<long>(r) * p.ldc + c] = TC{static_cast<S>(outr - h.x), static_cast<S>(outi - h.y)};
              if (r > c) {
                const c128 ht = Hb ? Hb[static_cast<long>(c) * p.ldc + r] : c128{0.0, 0.0};
                C[static_cast<long>(c) * p.ldc + r] = TC{static_cast<S>(outr - ht.x), static_cast<S>(-outi - ht.y)};
              }
            }
          }
      return;
    }
    c128* T = reinterpret_cast<c128*>(wave < 2 ? static_cast<void*>(&As[0][0][0]) : static_cast<void*>(&Bs[0][0][0])) +
              (wave & 1) * (16 * 17);
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        const int R0 = m0 + wm * (BM / 2) + i * 16, C0 = n0 + wn * (BN / 2) + j * 16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rl = AR::row(lk, v), r = R0 + rl, c = C0 + lr;
          double xr = cr[i][j][v], xi = ci[i][j][v];
          if constexpr (FOLD) {
            xr += fr[i][j][v];
            xi += fi[i][j][v];
          }
          const double outr = p.ar * xr - p.ai * xi, outi = p.ar * xi + p.ai * xr;
          if (r < p.M && c < p.N && r >= c) {
            const c128 h = Hb ? Hb[static_cast<long>(r) * p.ldc + c] : c128{0.0, 0.0};
            C[static_cast<long>(r) * p.ldc + c] = TC{static_cast<S>(outr - h.x), static_cast<S>(outi - h.y)};
          }
          T[lr * 17 + rl] = c128{outr, -outi};  // T[col][row] = conj(X[row][col])
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int cl = lk + 4 * v;  // source column: destination row C0 + cl
          const int r = R0 + lr, c = C0 + cl;
          if (r < p.M && c < p.N && r > c) {
            const c128 x = T[cl * 17 + lr];
            const c128 ht = Hb ? Hb[static_cast<long>(c) * p.ldc + r] : c128{0.0, 0.0};
            C[static_cast<long>(c) * p.ldc + r] = TC{static_cast<S>(x.x - ht.x), static_cast<S>(x.y - ht.y)};
          }
        }
        __syncthreads();
      }
    return;
  }
  const bool beta0 = (p.br == 0.0 && p.bi == 0.0);
  const bool gamma0 = (p.gr == 0.0 && p.gi == 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = m0 + wm * (BM / 2) + i * 16 + AR::row(lk, v);
        const int c = n0 + wn * (BN / 2) + j * 16 + lr;
        if (r < p.M && c < p.N) {
          double xr = cr[i][j][v], xi = ci[i][j][v];
          if constexpr (FOLD) {
            xr += fr[i][j][v];
            xi += fi[i][j][v];
          }
          double outr = p.ar * xr - p.ai * xi;
          double outi = p.ar * xi + p.ai * xr;
          TC* dst = C + static_cast<long>(r) * p.ldc + c;
          using S = decltype(TC{}.x);
          if (!beta0) {
            const TC old = *dst;
            const double o_r = old.x, o_i = old.y;
            outr += p.br * o_r - p.bi * o_i;
            outi += p.br * o_i + p.bi * o_r;
          }
          if (!gamma0 && r == c) {
            outr += p.gr;
            outi += p.gi;
          }
          *dst = TC{static_cast<S>(outr), static_cast<S>(outi)};
        }
      }
}

// ---------------------------------------------------------------------------------
// Split-BF16 product with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Every real operand value x is split into three bf16 parts x = x0 + x1 + x2 (x0 = bf16(x),
// x1 = bf16(x - x0), x2 = bf16(x - x0 - x1); the differences are exact in FP32), which
// carry 24 significant bits between them.  A real product keeps the six part products
// with s + t <= 2, a_s b_t, so the dropped terms are O(2^-24) relative: the result has
// FP32-class accuracy while the MFMA work is six bf16 products per real product, each
// 8 times the K depth of the FP32 instruction.
//
// Tile 64 x 64 (complex), four waves as 2 x 2, K chunks of 32 complex values.  The LDS
// holds, per operand, part and real/imaginary component, a 64 x 32 bf16 panel with the K
// index contiguous (rows padded to 40 values so that the 16-byte fragment reads of 16
// lanes fall into distinct banks).  Lane l of an MFMA reads row l%16, k = 8(l/16) .. +7.
// The next chunk's global loads are issued before the current chunk's MFMAs.
// ---------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int XBK = 32;  // K chunk, complex elements
constexpr int XLD = 40;  // LDS row stride, bf16 values

template <typename T>
__device__ __forceinline__ float2 ld_f2(const T* X, long ld, int r, int c, int R, int Cn, bool conj) {
  if (r < R && c < Cn) {
    const T v = X[static_cast<long>(r) * ld + c];
    return float2{static_cast<float>(v.x), static_cast<float>(conj ? -v.y : v.y)};
  }
  return float2{0.0f, 0.0f};
}

__device__ __forceinline__ void split3(float x, __bf16& h0, __bf16& h1, __bf16& h2) {
  h0 = static_cast<__bf16>(x);
  const float r1 = x - static_cast<float>(h0);
  h1 = static_cast<__bf16>(r1);
  const float r2 = r1 - static_cast<float>(h1);
  h2 = static_cast<__bf16>(r2);
}

// One operand's 64 x 32 chunk: 8 complex values per thread.  ROWS: element (x, k) at
// X[x][k] (8 consecutive k per thread); otherwise at X[k][x] (4 consecutive x for each
// of two consecutive k per thread).
template <typename T, bool ROWS, bool CONJ>
struct SplitIO {
  __device__ static void load(const T* X, long ld, int nx, int nk, int x0, int k0, float2 (&v)[8]) {
    const int t = threadIdx.x;
    if (ROWS) {
      const int x = t >> 2, k = (t & 3) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = ld_f2(X, ld, x0 + x, k0 + k + e, nx, nk, CONJ);
    } else {
      const int k = (t >> 4) * 2, x = (t & 15) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        v[e] = ld_f2(X, ld, k0 + k, x0 + x + e, nk, nx, CONJ);
        v[4 + e] = ld_f2(X, ld, k0 + k + 1, x0 + x + e, nk, nx, CONJ);
      }
    }
  }
  // S[part][re/im][x][k]
  __device__ static void store(__bf16 (*S)[2][64][XLD], const float2 (&v)[8]) {
    const int t = threadIdx.x;
    if (ROWS) {
      const int x = t >> 2, k = (t & 3) * 8;
      bf16x8 p[3][2];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 a0, a1, a2, b0, b1, b2;
        split3(v[e].x, a0, a1, a2);
        split3(v[e].y, b0, b1, b2);
        p[0][0][e] = a0;
        p[1][0][e] = a1;
        p[2][0][e] = a2;
        p[0][1][e] = b0;
        p[1][1][e] = b1;
        p[2][1][e] = b2;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int c = 0; c < 2; ++c) *reinterpret_cast<bf16x8*>(&S[q][c][x][k]) = p[q][c];
    } else {
      const int k = (t >> 4) * 2, x = (t & 15) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        __bf16 a[2][3], b[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          split3(v[4 * h + e].x, a[h][0], a[h][1], a[h][2]);
          split3(v[4 * h + e].y, b[h][0], b[h][1], b[h][2]);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          S[q][0][x + e][k] = a[0][q];
          S[q][0][x + e][k + 1] = a[1][q];
          S[q][1][x + e][k] = b[0][q];
          S[q][1][x + e][k + 1] = b[1][q];
        }
      }
    }
  }
};

__device__ __forceinline__ bf16x8 bneg(bf16x8 v) {
  u32x4 u = __builtin_bit_cast(u32x4, v);
  u ^= u32x4{0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
  return __builtin_bit_cast(bf16x8, u);
}

template <typename TA, typename TB, typename TC, int OPA, int OPB, bool FOLD>
__global__ __launch_bounds__(NT) void gemm_bf16x6_kernel(Args p) {
  using IOA = SplitIO<TA, OPA == OP_N, OPA == OP_C>;  // op(A)(m,k)
  using IOB = SplitIO<TB, OPB != OP_N, OPB == OP_C>;  // op(B)(k,n) stored as (n,k)
  __shared__ __attribute__((aligned(16))) __bf16 As[3][2][64][XLD];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[3][2][64][XLD];

  const int nblk = gridDim.x;
  const int hw = blockIdx.x;
  const int xcd = hw % NXCD, loc = hw / NXCD;
  const int base = nblk / NXCD, rem = nblk % NXCD;
  const int logical = xcd * base + (xcd < rem ? xcd : rem) + loc;
  const int tiles = p.tiles_m * p.tiles_n;
  const int b = logical / tiles;
  const int tile = logical % tiles;
  const int m0 = (tile / p.tiles_n) * 64;
  const int n0 = (tile % p.tiles_n) * 64;
  if ((p.tri & GEMM_TRI_LOWER) && m0 + 63 < n0) return;
  int k_lo = 0;
  if (p.tri & GEMM_TRI_KFROM_ROW) k_lo = m0;
  if (p.tri & GEMM_TRI_KFROM_COL) k_lo = n0 > k_lo ? n0 : k_lo;
  k_lo = (k_lo / XBK) * XBK;
  int k_hi = p.K;
  if ((p.tri & GEMM_TRI_KTO_ROW) && m0 + 64 < k_hi) k_hi = m0 + 64;
  if ((p.tri & GEMM_TRI_KTO_COL) && n0 + 64 < k_hi) k_hi = n0 + 64;

  const TA* A = batch_ptr<TA>(p.A, b);
  const TB* B = batch_ptr<TB>(p.B, b);
  TC* C = static_cast<TC*>(p.C) + static_cast<long>(b) * p.strideC;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;

  f32x4 cr[2][2], ci[2][2];
  f64x4 fr[FOLD ? 2 : 1][FOLD ? 2 : 1], fi[FOLD ? 2 : 1][FOLD ? 2 : 1];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = f32x4{0, 0, 0, 0};
      ci[i][j] = f32x4{0, 0, 0, 0};
      if constexpr (FOLD) {
        fr[i][j] = f64x4{0, 0, 0, 0};
        fi[i][j] = f64x4{0, 0, 0, 0};
      }
    }

  const int nk = (k_hi + XBK - 1) / XBK;
  const int kt0 = k_lo / XBK;
  float2 va[8], vb[8];
  if (kt0 < nk) {
    IOA::load(A, p.lda, p.M, p.K, m0, k_lo, va);
    IOB::load(B, p.ldb, p.N, p.K, n0, k_lo, vb);
  }
  for (int kt = kt0; kt < nk; ++kt) {
    IOA::store(As, va);
    IOB::store(Bs, vb);
    __syncthreads();
    if (kt + 1 < nk) {
      IOA::load(A, p.lda, p.M, p.K, m0, (kt + 1) * XBK, va);
      IOB::load(B, p.ldb, p.N, p.K, n0, (kt + 1) * XBK, vb);
    }
    bf16x8 fa[2][3][3], fb[2][3][2];  // [block][part][re, im, -im] and [block][part][re, im]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int row = wm * 32 + i * 16 + lr;
        fa[i][q][0] = *reinterpret_cast<const bf16x8*>(&As[q][0][row][lk * 8]);
        fa[i][q][1] = *reinterpret_cast<const bf16x8*>(&As[q][1][row][lk * 8]);
        fa[i][q][2] = bneg(fa[i][q][1]);
      }
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int col = wn * 32 + j * 16 + lr;
        fb[j][q][0] = *reinterpret_cast<const bf16x8*>(&Bs[q][0][col][lk * 8]);
        fb[j][q][1] = *reinterpret_cast<const bf16x8*>(&Bs[q][1][col][lk * 8]);
      }
    // part pairs (s, t) with s + t <= 2, smallest first
#pragma unroll
    for (int pp = 5; pp >= 0; --pp) {
      const int s_ = pp == 0 ? 0 : pp == 1 ? 0 : pp == 2 ? 1 : pp == 3 ? 0 : pp == 4 ? 1 : 2;
      const int t_ = pp == 0 ? 0 : pp == 1 ? 1 : pp == 2 ? 0 : pp == 3 ? 2 : pp == 4 ? 1 : 0;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][s_][0], fb[j][t_][0], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][s_][0], fb[j][t_][1], ci[i][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][s_][2], fb[j][t_][1], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i][s_][1], fb[j][t_][0], ci[i][j], 0, 0, 0);
        }
    }
    if constexpr (FOLD) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            fr[i][j][v] += static_cast<double>(cr[i][j][v]);
            fi[i][j][v] += static_cast<double>(ci[i][j][v]);
            cr[i][j][v] = 0.0f;
            ci[i][j][v] = 0.0f;
          }
    }
    __syncthreads();
  }

  const bool beta0 = (p.br == 0.0 && p.bi == 0.0);
  const bool gamma0 = (p.gr == 0.0 && p.gi == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // accumulator register v of lane l holds row 4(l/16) + v, column l%16 (checked on gfx950)
        const int r = m0 + wm * 32 + i * 16 + 4 * lk + v;
        const int c = n0 + wn * 32 + j * 16 + lr;
        if (r < p.M && c < p.N) {
          double xr = cr[i][j][v], xi = ci[i][j][v];
          if constexpr (FOLD) {
            xr += fr[i][j][v];
            xi += fi[i][j][v];
          }
          double outr = p.ar * xr - p.ai * xi;
          double outi = p.ar * xi + p.ai * xr;
          TC* dst = C + static_cast<long>(r) * p.ldc + c;
          if (!beta0) {
            const TC old = *dst;
            const double o_r = old.x, o_i = old.y;
            outr += p.br * o_r - p.bi * o_i;
            outi += p.br * o_i + p.bi * o_r;
          }
          if (!gamma0 && r == c) {
            outr += p.gr;
            outi += p.gi;
          }
          using S = decltype(TC{}.x);
          *dst = TC{static_cast<S>(outr), static_cast<S>(outi)};
        }
      }
}

std::atomic<long> g_launches{0};
std::atomic<long long> g_flops_milli{0};

// Optional per-launch timing with HIP events recorded on the launch stream.
struct ProfRec {
  hipEvent_t a, b;
  double flops, bytes;
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

// Three real products per complex product on 64 x 64 tiles: bit 0 FP64 arithmetic, bit 1 FP32
// without FP64 folding, bit 2 FP32 with FP64 folding (not the split-BF16 products); bit 3 is
// read by the callers (the FP32 projections of G for Sigma, through GEMM_G3).
std::atomic<int> g_complex3{0};

// Tile shapes: 0 = 64 x 64, 1 = 128 x 64, 2 = 128 x 128 (complex<double> only).
template <int OPA, int OPB>
void launch(hipStream_t s, Args a, int shape) {
  const int bm = shape == 0 ? 64 : 128, bn = shape == 2 ? 128 : 64;
  a.tiles_m = (a.M + bm - 1) / bm;
  a.tiles_n = (a.N + bn - 1) / bn;
  const long nblocks = static_cast<long>(a.batch) * a.tiles_m * a.tiles_n;
  THCGW_REQUIRE(nblocks < (1L << 31), "zgemm: grid too large");
  const dim3 g(static_cast<unsigned>(nblocks)), t(NT);
  if (shape == 0 && (g_complex3.load() & 1))
    hipLaunchKernelGGL((gemm_kernel<double, c128, c128, c128, OPA, OPB, 64, 64, false, true>), g, t, 0, s, a);
  else if (shape == 0)
    hipLaunchKernelGGL((gemm_kernel<double, c128, c128, c128, OPA, OPB, 64, 64>), g, t, 0, s, a);
  else if (shape == 1)
    hipLaunchKernelGGL((gemm_kernel<double, c128, c128, c128, OPA, OPB, 128, 64>), g, t, 0, s, a);
  else
    hipLaunchKernelGGL((gemm_kernel<double, c128, c128, c128, OPA, OPB, 128, 128>), g, t, 0, s, a);
}

std::atomic<int> g_tile_mode32{0};  // FP32 tile: 0 = 64 x 64, 1 = 128 x 64, 2 = 128 x 128

template <typename TA, typename TB, typename TC, int OPA, int OPB>
void launch_f32(hipStream_t s, Args a) {
  const int tb = (a.tri >> GEMM_TILE32_SHIFT) & 3;
  const int mode = ((a.tri & (GEMM_SPLIT_BF16 | GEMM_FOLD64)) == 0) ? (tb ? tb - 1 : g_tile_mode32.load()) : 0;
  const int bm = mode == 0 ? 64 : 128, bn = mode == 2 ? 128 : 64;
  a.tiles_m = (a.M + bm - 1) / bm;
  a.tiles_n = (a.N + bn - 1) / bn;
  const long nblocks = static_cast<long>(a.batch) * a.tiles_m * a.tiles_n;
  THCGW_REQUIRE(nblocks < (1L << 31), "cgemm: grid too large");
  if (mode == 1) {
    hipLaunchKernelGGL((gemm_kernel<float, TA, TB, TC, OPA, OPB, 128, 64>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(NT), 0, s, a);
  } else if (mode == 2) {
    hipLaunchKernelGGL((gemm_kernel<float, TA, TB, TC, OPA, OPB, 128, 128>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(NT), 0, s, a);
  } else if (a.tri & GEMM_SPLIT_BF16) {
    if (a.tri & GEMM_FOLD64)
      hipLaunchKernelGGL((gemm_bf16x6_kernel<TA, TB, TC, OPA, OPB, true>), dim3(static_cast<unsigned>(nblocks)),
                         dim3(NT), 0, s, a);
    else
      hipLaunchKernelGGL((gemm_bf16x6_kernel<TA, TB, TC, OPA, OPB, false>), dim3(static_cast<unsigned>(nblocks)),
                         dim3(NT), 0, s, a);
  } else if ((a.tri & GEMM_FOLD64) && ((g_complex3.load() & 4) || (a.tri & GEMM_G3)))
    hipLaunchKernelGGL((gemm_kernel<float, TA, TB, TC, OPA, OPB, 64, 64, true, true>),
                       dim3(static_cast<unsigned>(nblocks)), dim3(NT), 0, s, a);
  else if (a.tri & GEMM_FOLD64)
    hipLaunchKernelGGL((gemm_kernel<float, TA, TB, TC, OPA, OPB, 64, 64, true>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(NT), 0, s, a);
  else if ((g_complex3.load() & 2) || (a.tri & GEMM_G3))
    hipLaunchKernelGGL((gemm_kernel<float, TA, TB, TC, OPA, OPB, 64, 64, false, true>),
                       dim3(static_cast<unsigned>(nblocks)), dim3(NT), 0, s, a);
  else
    hipLaunchKernelGGL((gemm_kernel<float, TA, TB, TC, OPA, OPB, 64, 64>), dim3(static_cast<unsigned>(nblocks)),
                       dim3(NT), 0, s, a);
}

std::atomic<int> g_tile_mode{0};  // 0 = automatic, else shape + 1

// 64 x 64 tiles by default: at three waves per SIMD they hide load latency better than
// the larger shapes at one wave per SIMD (measured on the N_k = 64, N_mu = 800 iteration: 45.8 vs 26.4 TFLOP/s).
int tile_shape(int, int, int) {
  const int forced = g_tile_mode.load();
  return forced > 0 ? forced - 1 : 0;
}

}  // namespace

namespace {

Args make_args(int M, int N, int K, Cplx alpha, BatchPtr A, long lda, BatchPtr B, long ldb, Cplx beta, void* C,
               long strideC, long ldc, int batch, Cplx gamma, int tri, BatchPtr hsub) {
  THCGW_REQUIRE(!(tri & GEMM_HERM_SUB) || (M == N && beta.re == 0.0 && beta.im == 0.0 &&
                                           gamma.re == 0.0 && gamma.im == 0.0 && (tri & GEMM_TRI_LOWER) &&
                                           !(tri & GEMM_SPLIT_BF16)),
                "gemm: GEMM_HERM_SUB needs a square lower product without beta, gamma or split BF16");
  Args a;
  a.H = hsub;
  a.M = M;
  a.N = N;
  a.K = K;
  a.batch = batch;
  a.tri = tri;
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
  return a;
}

// Modelled flops; triangular structure (square operands) halves the work per flag.
double model_flops(int M, int N, int K, int batch, int tri) {
  // the other flags do not change the work
  tri &= GEMM_TRI_LOWER | GEMM_TRI_KFROM_ROW | GEMM_TRI_KFROM_COL | GEMM_TRI_KTO_ROW | GEMM_TRI_KTO_COL;
  double flops = 8.0 * double(M) * N * K * batch;
  if ((tri & GEMM_TRI_LOWER) && (tri & (GEMM_TRI_KFROM_ROW | GEMM_TRI_KFROM_COL))) flops /= 6.0;
  else if (tri) flops /= 2.0;
  return flops;
}

// Modelled algorithmic HBM bytes: op(A), op(B) read once, C written once (and read once
// when beta != 0), with element sizes ea, eb, ec.  A triangular output (GEMM_TRI_LOWER)
// counts half of C; a triangular operand (the K range of a tile bounded by its row or
// column) counts half of that operand.
double model_bytes(int M, int N, int K, int batch, int tri, bool beta_nonzero, int ea, int eb, int ec) {
  double a = double(M) * K * ea, b = double(K) * N * eb, c = double(M) * N * ec * (beta_nonzero ? 2 : 1);
  if (tri & GEMM_TRI_LOWER) c *= 0.5;
  if (tri & (GEMM_TRI_KFROM_ROW | GEMM_TRI_KTO_ROW)) a *= 0.5;
  if (tri & (GEMM_TRI_KFROM_COL | GEMM_TRI_KTO_COL)) b *= 0.5;
  return (a + b + c) * batch;
}

// Launch with optional event timing; cls indexes the profile classes (opA*3 + opB for
// FP64, 9 + ... for FP32).
template <class F>
void timed(hipStream_t s, double flops, double bytes, int cls, F&& launch_fn) {
  std::unique_lock<std::mutex> lk(g_prof_mu, std::defer_lock);
  ProfRec rec{nullptr, nullptr, flops, bytes, cls};
  if (g_prof_on) {
    lk.lock();
    rec.a = pool_event();
    rec.b = pool_event();
    THCGW_HIP(hipEventRecord(rec.a, s));
  }
  launch_fn();
  THCGW_HIP(hipGetLastError());
  if (g_prof_on) {
    THCGW_HIP(hipEventRecord(rec.b, s));
    g_prof.push_back(rec);
  }
  g_launches.fetch_add(1);
  g_flops_milli.fetch_add(static_cast<long long>(flops / 1000.0));
}

template <typename TA, typename TB, typename TC>
void dispatch_f32_ops(hipStream_t s, int opA, int opB, const Args& a) {
  if (opA == OP_N && opB == OP_N) launch_f32<TA, TB, TC, OP_N, OP_N>(s, a);
  else if (opA == OP_N && opB == OP_C) launch_f32<TA, TB, TC, OP_N, OP_C>(s, a);
  else if (opA == OP_C && opB == OP_N) launch_f32<TA, TB, TC, OP_C, OP_N>(s, a);
  else if (opA == OP_C && opB == OP_C) launch_f32<TA, TB, TC, OP_C, OP_C>(s, a);
  else throw Error(THCGW_ERR_UNSUPPORTED, "cgemm: transpose without conjugation is not built");
}

template <typename TA, typename TB>
void dispatch_f32_c(hipStream_t s, int opA, int opB, Dt tc, const Args& a) {
  if (tc == DT_C128) dispatch_f32_ops<TA, TB, c128>(s, opA, opB, a);
  else dispatch_f32_ops<TA, TB, c64>(s, opA, opB, a);
}

}  // namespace

void zgemm(hipStream_t s, int opA, int opB, int M, int N, int K, Cplx alpha, BatchPtr A, long lda,
           BatchPtr B, long ldb, Cplx beta, c128* C, long strideC, long ldc, int batch, Cplx gamma, int tri,
           BatchPtr hsub) {
  THCGW_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "zgemm: negative size");
  THCGW_REQUIRE(opA >= 0 && opA <= 2 && opB >= 0 && opB <= 2, "zgemm: bad op");
  if (M == 0 || N == 0 || batch == 0) return;
  const Args a = make_args(M, N, K, alpha, A, lda, B, ldb, beta, C, strideC, ldc, batch, gamma, tri, hsub);
  const int nb = tile_shape(M, N, K);
  const bool bnz = beta.re != 0.0 || beta.im != 0.0;
  timed(s, model_flops(M, N, K, batch, tri), model_bytes(M, N, K, batch, tri, bnz, 16, 16, 16), opA * 3 + opB, [&] {
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
  });
}

void cgemm(hipStream_t s, int opA, int opB, int M, int N, int K, Cplx alpha, BatchPtr A, Dt ta, long lda,
           BatchPtr B, Dt tb, long ldb, Cplx beta, void* C, Dt tc, long strideC, long ldc, int batch, Cplx gamma,
           int tri, BatchPtr hsub) {
  THCGW_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "cgemm: negative size");
  if (M == 0 || N == 0 || batch == 0) return;
  THCGW_REQUIRE(!(tri & GEMM_HERM_SUB) || tc == DT_C128 || !hsub.p,
                "cgemm: GEMM_HERM_SUB with a subtracted matrix writes complex<double>");
  const Args a = make_args(M, N, K, alpha, A, lda, B, ldb, beta, C, strideC, ldc, batch, gamma, tri, hsub);
  const int cls = ((tri & GEMM_SPLIT_BF16) ? 13 : 9) + (opA == OP_C ? 2 : 0) + (opB == OP_C ? 1 : 0);
  const bool bnz = beta.re != 0.0 || beta.im != 0.0;
  const int ea = ta == DT_C128 ? 16 : 8, eb = tb == DT_C128 ? 16 : 8, ec = tc == DT_C128 ? 16 : 8;
  timed(s, model_flops(M, N, K, batch, tri), model_bytes(M, N, K, batch, tri, bnz, ea, eb, ec), cls, [&] {
    if (ta == DT_C128 && tb == DT_C128) dispatch_f32_c<c128, c128>(s, opA, opB, tc, a);
    else if (ta == DT_C128) dispatch_f32_c<c128, c64>(s, opA, opB, tc, a);
    else if (tb == DT_C128) dispatch_f32_c<c64, c128>(s, opA, opB, tc, a);
    else dispatch_f32_c<c64, c64>(s, opA, opB, tc, a);
  });
}

GemmStats gemm_stats() { return GemmStats{g_launches.load(), 1000.0 * double(g_flops_milli.load())}; }
void gemm_stats_reset() {
  g_launches = 0;
  g_flops_milli = 0;
}

void gemm_set_tile_mode(int mode) { g_tile_mode = mode; }
void gemm_set_tile_mode32(int mode) { g_tile_mode32 = mode; }
void gemm_set_complex3(int on) { g_complex3 = on; }
int gemm_complex3() { return g_complex3.load(); }
int gemm_mode_key() { return g_tile_mode.load() | (g_tile_mode32.load() << 4) | (g_complex3.load() << 8); }

void gemm_profile_enable(bool on) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_on = on;
}

// Per class (FP64: opA*3 + opB, 0..8; FP32: 9 + 2*(opA == C) + (opB == C), 9..12; split
// BF16: 13 + 2*(opA == C) + (opB == C), 13..16): summed
// launch durations (ms), modelled flops, modelled algorithmic bytes (bytes may be null)
// and launch counts of the recorded launches; the records are then dropped.
void gemm_profile_read(double* ms, double* flops, double* bytes, long* launches) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int c = 0; c < GEMM_PROFILE_CLASSES; ++c) {
    ms[c] = 0.0;
    flops[c] = 0.0;
    if (bytes) bytes[c] = 0.0;
    launches[c] = 0;
  }
  for (const ProfRec& r : g_prof) {
    THCGW_HIP(hipEventSynchronize(r.b));
    float e = 0.0f;
    THCGW_HIP(hipEventElapsedTime(&e, r.a, r.b));
    ms[r.cls] += e;
    flops[r.cls] += r.flops;
    if (bytes) bytes[r.cls] += r.bytes;
    launches[r.cls] += 1;
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_prof.clear();
}

}  // namespace thcgw
