// Kinship Gram on MFMA: G = Z Z^T for Z (n x m) held as split operand planes (rows = strains, K = SNPs), the
// product behind io.grm's K = Z Z^T / M (reference fasterlmm/io.py:257-265) and the LOCO kernels of
// core.py:708-716.
//
// Work split: one 256-thread workgroup per (upper-triangular 128 x 128 tile pair (ti <= tj), K split s).  The
// workgroup walks its K range in blocks of GRAM_KBLOCK chunks (1,024 k) through the tile engine: each block's
// FP32 MFMA sums (three split products, as in every split product of this library) are folded into FP64
// registers, so no FP32 accumulation chain runs over more than 1,024 k whatever m is.  Split-FP16 rows carry
// power-of-two scales that are divided out exactly.  Each split writes its FP64 tile to partial[s]; gram_reduce
// sums the splits in a fixed order and mirrors the upper triangle, so results are deterministic and symmetric.
#include "kernels.h"

namespace flmm {

constexpr int GRAM_KBLOCK = 32;  // K chunks (of 32) per FP32 block

template <int PREC>
__global__ __launch_bounds__(NTHREADS) void gram_kernel(TileArgs<PREC, 1> args, int n_tiles, int n_split,
                                                        const float* __restrict__ row_scale, double* __restrict__ partial,
                                                        int64_t n_pad) {
  extern __shared__ char lds[];
  // upper-triangular pair index -> (ti, tj), ti <= tj, row-major over ti
  int64_t pair = blockIdx.x, ti = 0;
  while (pair >= n_tiles - ti) {
    pair -= n_tiles - ti;
    ++ti;
  }
  const int64_t tj = ti + pair;
  const int s = blockIdx.y;
  const int64_t per = (args.nchunks + n_split - 1) / n_split;
  const int64_t c0 = s * per, c1 = c0 + per < args.nchunks ? c0 + per : args.nchunks;
  using T = PlaneTraits<PREC>;
  const int64_t chunk_bytes = args.a_rows_pad * T::ROWB;  // one K chunk of one plane

  double acc64[2][2][16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc64[i][j][v] = 0.0;

  for (int64_t kb = c0; kb < c1; kb += GRAM_KBLOCK) {
    TileArgs<PREC, 1> t = args;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      if (args.a[0][u]) t.a[0][u] = (const char*)args.a[0][u] + kb * chunk_bytes;
      if (args.b[0][u]) t.b[0][u] = (const char*)args.b[0][u] + kb * chunk_bytes;
    }
    t.nchunks = (int)(c1 - kb < GRAM_KBLOCK ? c1 - kb : GRAM_KBLOCK);
    TileAcc<PREC, 1> acc;
    tile_mainloop<PREC, 1>(t, ti * TILE, tj * TILE, acc, lds);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc64[i][j][v] += (double)acc.c[0][i][j][v];
    __syncthreads();  // the next block's first loads reuse the LDS
  }

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  double* out = partial + (int64_t)s * n_pad * n_pad;
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int64_t j = tj * TILE + acc_col(wn, bj, lane);
    const double sj = row_scale ? 1.0 / (double)row_scale[j] : 1.0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t i = ti * TILE + acc_row(wm, bi, lane, v);
        const double si = row_scale ? 1.0 / (double)row_scale[i] : 1.0;
        out[i * n_pad + j] = acc64[bi][bj][v] * si * sj;
      }
  }
}

// out[i][j] (n x n, leading dimension ld) = sum over s of partial[s][min(i, j)][max(i, j)], in split order (the
// upper element, so the result is exactly symmetric)
__global__ __launch_bounds__(256) void gram_reduce_kernel(const double* __restrict__ partial, int n_split,
                                                          int64_t n_pad, int64_t n, double* __restrict__ out,
                                                          int64_t ld) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * n) return;
  int64_t i = e / n, j = e % n;
  if (i > j) {
    const int64_t x = i;
    i = j;
    j = x;
  }
  double acc = 0.0;
  for (int s = 0; s < n_split; ++s) acc += partial[(int64_t)s * n_pad * n_pad + i * n_pad + j];
  out[(e / n) * ld + e % n] = acc;
}

int gram_tile_pairs(int64_t rows_pad) {
  const int64_t t = rows_pad / TILE;
  return (int)(t * (t + 1) / 2);
}

template <int PREC>
hipError_t launch_gram(const TileArgs<PREC, 1>& args, const float* row_scale, int n_split, double* partial, int64_t n,
                       double* out, int64_t ld, hipStream_t stream) {
  const int n_tiles = (int)(args.a_rows_pad / TILE);
  const dim3 grid((unsigned)gram_tile_pairs(args.a_rows_pad), (unsigned)n_split);
  constexpr int lds = tile_lds_bytes<PREC, 1>();
  hipLaunchKernelGGL(gram_kernel<PREC>, grid, dim3(NTHREADS), lds, stream, args, n_tiles, n_split, row_scale, partial,
                     args.a_rows_pad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, stream, partial, n_split,
                     args.a_rows_pad, n, out, ld);
  return hipGetLastError();
}

template hipError_t launch_gram<PREC_SPLIT_FP16>(const TileArgs<PREC_SPLIT_FP16, 1>&, const float*, int, double*,
                                                  int64_t, double*, int64_t, hipStream_t);
template hipError_t launch_gram<PREC_SPLIT_BF16>(const TileArgs<PREC_SPLIT_BF16, 1>&, const float*, int, double*,
                                                  int64_t, double*, int64_t, hipStream_t);
template hipError_t launch_gram<PREC_FP32>(const TileArgs<PREC_FP32, 1>&, const float*, int, double*, int64_t, double*,
                                            int64_t, hipStream_t);

}  // namespace flmm
