// Imaginary-axis transforms between tau and iW_n (IAFT.icc:49-73, 104-128): a short,
// wide product out[r][c] = sum_i T[r][i] in[i][c] with R <= 48, I <= 44 and up to
// N_q N_mu^2 columns.  As a 64 x 64 GEMM tile the R = 36..37 rows of the 72-point table
// fill little more than half the tile and every tile restages T through LDS.
//
// Here T is staged once per workgroup in LDS in the order of the A fragments of the FP64
// 16x16x4 MFMA (lane i + 16k reads T[16 rb + i][4 kb + k], consecutive lanes consecutive
// 16-byte words), and the columns stream from HBM straight into B fragments (lane j + 16k
// holds in[4 kb + k][c0 + j]; 16 lanes read 256 contiguous bytes).  One wave forms 16 columns by RB x KB x 4 real MFMAs and walks
// the column blocks with a grid stride.  Per column the useful work is 8 R I flops against
// 16 (R + I) bytes of HBM traffic, about 9 flops per byte at R = I = 36: close to the FP64
// ridge of MI355X (78.6 TFLOP/s over 8 TB/s).  Padding R to 16 RB rows costs 30% of the
// MFMA work at R = 37 against 73% for 64-row tiles.  The particle-hole folded tables are real
// up to rounding (|Im T| ~ 1e-14 |Re T|): with real_t the imaginary part is dropped, which
// halves the MFMAs (2 instead of 4 per complex product) and leaves the kernel HBM-bound.
#include "thcgw_internal.hpp"

#include <algorithm>

namespace thcgw {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int RB, int KB, bool CT>
__global__ __launch_bounds__(256) void ir_transform_kernel(const c128* __restrict__ T, int R, int I,
                                                           const c128* __restrict__ in, long ld_in, long bs_in,
                                                           c128* __restrict__ out, long ld_out, long bs_out,
                                                           long cols) {
  // T in A-fragment order: Tf[rb][kb][lane] = T[16 rb + lane % 16][4 kb + lane / 16] (zero padded)
  __shared__ c128 Tf[RB * KB * 64];
  for (int e = threadIdx.x; e < RB * KB * 64; e += 256) {
    const int l = e & 63, f = e >> 6, kb = f % KB, rb = f / KB;
    const int r = rb * 16 + (l & 15), k = kb * 4 + (l >> 4);
    Tf[e] = (r < R && k < I) ? T[static_cast<long>(r) * I + k] : c128{0.0, 0.0};
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const long b = blockIdx.y;
  const c128* inb = in + b * bs_in;
  c128* outb = out + b * bs_out;
  const long nblk = (cols + 15) / 16;
  const long stride = static_cast<long>(gridDim.x) * 4;
  for (long cb = static_cast<long>(blockIdx.x) * 4 + (threadIdx.x >> 6); cb < nblk; cb += stride) {
    // Up to 27 fragments the compiler keeps T in registers for the whole loop (2-3 waves per
    // SIMD); beyond that the fragments are re-read from LDS per block, which keeps 3 waves
    // per SIMD instead of 1.
    if constexpr (RB * KB > 27) asm volatile("" ::: "memory");
    const long col = cb * 16 + li;
    double xr[KB], xi[KB], xn[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int k = kb * 4 + lk;
      const c128 v = (k < I && col < cols) ? inb[static_cast<long>(k) * ld_in + col] : c128{0.0, 0.0};
      xr[kb] = v.x;
      xi[kb] = v.y;
      xn[kb] = -v.y;
    }
    f64x4 cr[RB], ci[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      cr[rb] = f64x4{0.0, 0.0, 0.0, 0.0};
      ci[rb] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const c128 t = Tf[(rb * KB + kb) * 64 + lane];
        cr[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(t.x, xr[kb], cr[rb], 0, 0, 0);
        ci[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(t.x, xi[kb], ci[rb], 0, 0, 0);
        if constexpr (CT) {  // complex T: the two products of its imaginary part
          cr[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(t.y, xn[kb], cr[rb], 0, 0, 0);
          ci[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(t.y, xr[kb], ci[rb], 0, 0, 0);
        }
      }
    // accumulator register v of lane l holds row l/16 + 4v, column l%16
    if (col < cols) {
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = rb * 16 + lk + 4 * v;
          if (r < R) outb[static_cast<long>(r) * ld_out + col] = c128{cr[rb][v], ci[rb][v]};
        }
    }
  }
}

template <int RB, int KB>
void launch_irt(hipStream_t s, const c128* T, int R, int I, const c128* in, long ld_in, long bs_in, c128* out,
                long ld_out, long bs_out, long cols, int batch, bool real_t) {
  const long nblk = (cols + 15) / 16;
  const long want = (nblk + 3) / 4;                  // one 16-column block per wave
  const long cap = std::max(1L, 4096L / batch);      // about 16 workgroups per CU in all
  const dim3 g(static_cast<unsigned>(std::min(want, cap)), batch);
  if (real_t)
    hipLaunchKernelGGL((ir_transform_kernel<RB, KB, false>), g, dim3(256), 0, s, T, R, I, in, ld_in, bs_in, out,
                       ld_out, bs_out, cols);
  else
    hipLaunchKernelGGL((ir_transform_kernel<RB, KB, true>), g, dim3(256), 0, s, T, R, I, in, ld_in, bs_in, out,
                       ld_out, bs_out, cols);
  THCGW_HIP(hipGetLastError());
}

}  // namespace

// At most 3 row blocks (R <= 48) and 11 K blocks (I <= 44) of the MFMA; T takes up to 33 KB
// of LDS.  The 80-point tables (nt_half = 40, nw_half = 41) need the eleventh K block for
// the iW -> tau direction.
bool ir_transform_fits(int R, int I) { return R >= 1 && I >= 1 && R <= 48 && I <= 44; }

void ir_transform(hipStream_t s, const c128* T, int R, int I, const c128* in, long ld_in, long bs_in, c128* out,
                  long ld_out, long bs_out, long cols, int batch, bool real_t) {
  THCGW_REQUIRE(ir_transform_fits(R, I), "ir_transform: at most 48 rows and 44 columns of T");
  THCGW_REQUIRE(batch >= 1 && batch <= 65535, "ir_transform: batch out of range");
  if (cols <= 0) return;
  if (R <= 32 && I <= 24)
    launch_irt<2, 6>(s, T, R, I, in, ld_in, bs_in, out, ld_out, bs_out, cols, batch, real_t);
  else if (I <= 36)
    launch_irt<3, 9>(s, T, R, I, in, ld_in, bs_in, out, ld_out, bs_out, cols, batch, real_t);
  else if (I <= 40)
    launch_irt<3, 10>(s, T, R, I, in, ld_in, bs_in, out, ld_out, bs_out, cols, batch, real_t);
  else
    launch_irt<3, 11>(s, T, R, I, in, ld_in, bs_in, out, ld_out, bs_out, cols, batch, real_t);
}

}  // namespace thcgw
