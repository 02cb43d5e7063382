// Rotation into the eigenbasis on MFMA: out[i][j] = sum_n A[i][n] * B[j][n], with A = rows to rotate (scan
// columns of (Y - X X^+ Y), or the SNPs of one chromosome) and B = U_eff^T (the kept eigenvectors as rows).
// This is U_eff^T Y_r and U_eff^T S of the reference (core.py:517-518, 628-629), computed as a split-BF16 (or
// FP32) GEMM with FP32 accumulation.
//
// Epilogues write the next stage's operand planes directly (rows i, K index j):
//   MODE_UY:  one fp32 plane of the rotated columns, read by the delta search and the grid GEMM
//   MODE_US:  the scan's SNP-side planes S1 = x and S2 = x^2 (bf16 hi/lo pairs, or fp32)
#include "kernels.h"

namespace flmm {

enum RotMode : int { MODE_UY = 0, MODE_US = 1 };

template <int PREC, int MODE>
__global__ __launch_bounds__(NTHREADS, 1) void rotate_kernel(TileArgs<PREC, 1> args, int64_t n_jtiles,
                                                             int64_t n_tiles, void* o0, void* o1, void* o2,
                                                             void* o3) {
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
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t i = i0 + acc_row(wm, bi, lane, v);
        const int64_t j = j0 + acc_col(wn, bj, lane);
        const int64_t off = ((j >> 5) * rows_pad + i) * 32 + (j & 31);
        const float x = acc.c[0][bi][bj][v];
        if constexpr (MODE == MODE_UY) {
          ((float*)o0)[off] = x;
        } else if constexpr (PREC == PREC_SPLIT_BF16) {
          u16 h, l;
          split_bf16((double)x, h, l);
          ((u16*)o0)[off] = h;
          ((u16*)o1)[off] = l;
          split_bf16((double)x * (double)x, h, l);
          ((u16*)o2)[off] = h;
          ((u16*)o3)[off] = l;
        } else {
          ((float*)o0)[off] = x;
          ((float*)o1)[off] = x * x;
        }
      }
}

template <int PREC>
hipError_t launch_rotate(int mode, const TileArgs<PREC, 1>& args, void* o0, void* o1, void* o2, void* o3,
                         hipStream_t stream) {
  const int64_t n_itiles = args.a_rows_pad / TILE;
  const int64_t n_jtiles = args.b_rows_pad / TILE;
  const int64_t n_tiles = n_itiles * n_jtiles;
  const int lds = tile_lds_bytes<PREC, 1>();
  if (mode == MODE_UY)
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_UY>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, o0, o1, o2, o3);
  else
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_US>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, o0, o1, o2, o3);
  return hipGetLastError();
}

template hipError_t launch_rotate<PREC_SPLIT_BF16>(int, const TileArgs<PREC_SPLIT_BF16, 1>&, void*, void*, void*,
                                                    void*, hipStream_t);
template hipError_t launch_rotate<PREC_FP32>(int, const TileArgs<PREC_FP32, 1>&, void*, void*, void*, void*,
                                              hipStream_t);

}  // namespace flmm
