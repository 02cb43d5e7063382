// Rotation into the eigenbasis on MFMA: out[i][j] = sum_n A[i][n] * B[j][n], with A = rows to rotate (scan
// columns of (Y - X X^+ Y), or the SNPs of one chromosome) and B = the basis vectors as rows (orthonormal).
// This is U_eff^T Y_r and U_eff^T S of the reference (core.py:517-518, 628-629), computed as a split (or FP32)
// GEMM with FP32 accumulation.
//
// Epilogues write the next stage's operands directly (rows i, K index j):
//   MODE_UY:  one fp32 plane of the rotated columns, read by the delta search and the grid GEMM
//   MODE_US:  the scan's SNP-side planes S1 = x and S2 = x^2 (split pairs, or fp32)
// Split-FP16 operands arrive with per-row power-of-two scales (a_scale, b_scale) that the epilogue divides out
// exactly.  Because B's rows are unit vectors, |x[i][j]| <= ||A_i|| < 2^14 / a_scale[i], so the S1 / S2 planes are
// written with scales a_scale[i] and a_scale[i]^2 / 2^14 (out_scale1 / out_scale2).
#include "kernels.h"

namespace flmm {

enum RotMode : int { MODE_UY = 0, MODE_US = 1 };

template <int PREC, int MODE>
__global__ __launch_bounds__(NTHREADS, 1) void rotate_kernel(TileArgs<PREC, 1> args, int64_t n_jtiles,
                                                             int64_t n_tiles, const float* __restrict__ a_scale,
                                                             const float* __restrict__ b_scale, void* o0, void* o1,
                                                             void* o2, void* o3, float* __restrict__ out_scale1,
                                                             float* __restrict__ out_scale2) {
  extern __shared__ char lds[];
  const int64_t tile = xcd_remap(blockIdx.x, n_tiles);
  const int64_t jt = tile % n_jtiles;  // j fastest: the tiles sharing one A block run together on one XCD
  const int64_t it = tile / n_jtiles;
  const int64_t i0 = it * TILE, j0 = jt * TILE;

  TileAcc<PREC, 1> acc;
  tile_mainloop<PREC, 1>(args, i0, j0, acc, lds);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t rows_pad = args.a_rows_pad;
  // split-FP16 scales are powers of two, so multiplying by their reciprocals divides them out exactly
  double inv_b[2] = {1.0, 1.0};
  if constexpr (PREC == PREC_SPLIT_FP16) {
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) inv_b[bj] = 1.0 / (double)b_scale[j0 + acc_col(wn, bj, lane)];
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int64_t i = i0 + acc_row(wm, bi, lane, v);
      double inv_a = 1.0, s1 = 1.0, s2 = 1.0;
      if constexpr (PREC == PREC_SPLIT_FP16) {
        const double sa = (double)a_scale[i];
        inv_a = 1.0 / sa;
        s1 = sa;
        s2 = sa * sa * (1.0 / 16384.0);
        if (MODE == MODE_US && jt == 0 && wn == 0 && (lane & 31) == 0) {
          out_scale1[i] = (float)s1;
          out_scale2[i] = (float)s2;
        }
      }
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int64_t j = j0 + acc_col(wn, bj, lane);
        const int64_t off = ((j >> 5) * rows_pad + i) * 32 + (j & 31);
        double x = (double)acc.c[0][bi][bj][v];
        if constexpr (PREC == PREC_SPLIT_FP16) x *= inv_a * inv_b[bj];
        if constexpr (MODE == MODE_UY) {
          ((float*)o0)[off] = (float)x;
        } else if constexpr (is_split(PREC)) {
          u16 h, l;
          split2<PREC>(x * s1, h, l);
          ((u16*)o0)[off] = h;
          ((u16*)o1)[off] = l;
          split2<PREC>(x * x * s2, h, l);
          ((u16*)o2)[off] = h;
          ((u16*)o3)[off] = l;
        } else {
          ((float*)o0)[off] = (float)x;
          ((float*)o1)[off] = (float)(x * x);
        }
      }
    }
}

template <int PREC>
hipError_t launch_rotate(int mode, const TileArgs<PREC, 1>& args, const float* a_scale, const float* b_scale,
                         void* o0, void* o1, void* o2, void* o3, float* out_scale1, float* out_scale2,
                         hipStream_t stream) {
  const int64_t n_itiles = args.a_rows_pad / TILE;
  const int64_t n_jtiles = args.b_rows_pad / TILE;
  const int64_t n_tiles = n_itiles * n_jtiles;
  const int lds = tile_lds_bytes<PREC, 1>();
  if (mode == MODE_UY)
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_UY>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, a_scale, b_scale, o0, o1, o2, o3, out_scale1, out_scale2);
  else
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_US>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, a_scale, b_scale, o0, o1, o2, o3, out_scale1, out_scale2);
  return hipGetLastError();
}

template hipError_t launch_rotate<PREC_SPLIT_BF16>(int, const TileArgs<PREC_SPLIT_BF16, 1>&, const float*,
                                                    const float*, void*, void*, void*, void*, float*, float*,
                                                    hipStream_t);
template hipError_t launch_rotate<PREC_SPLIT_FP16>(int, const TileArgs<PREC_SPLIT_FP16, 1>&, const float*,
                                                    const float*, void*, void*, void*, void*, float*, float*,
                                                    hipStream_t);
template hipError_t launch_rotate<PREC_FP32>(int, const TileArgs<PREC_FP32, 1>&, const float*, const float*, void*,
                                              void*, void*, void*, float*, float*, hipStream_t);

}  // namespace flmm
