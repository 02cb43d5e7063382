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

// FP64 Gram on the FP64 matrix cores (v_mfma_f64_16x16x4f64), reading Z (n x m, row stride ldz) directly.  One
// 256-thread workgroup per (upper tile pair, K range), as above; the four waves split the 128 x 128 tile 2 x 2 and
// each holds 4 x 4 blocks of 16 x 16 FP64 accumulators.  K is staged through LDS 16 k at a time (rows padded to 18
// doubles: a half-wave's 8-byte reads of one k over 16 rows and of the next k cover the 64 banks once); the next
// stage's values are fetched into registers while the current one is multiplied.  Every product and sum is FP64,
// in a fixed order, so the result is deterministic and agrees with any FP64 evaluation of Z Z^T to FP64 rounding.
constexpr int G64_KT = 16;              // k per LDS stage
constexpr int G64_LD = G64_KT + 2;      // LDS row stride (doubles)
constexpr int G64_ROWS_PER_PASS = NTHREADS / G64_KT;  // rows one load instruction covers (16)
constexpr int G64_LOADS = TILE / G64_ROWS_PER_PASS;    // loads per thread per operand and stage (8)

typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(NTHREADS) void gram_f64_kernel(const double* __restrict__ Z, int64_t ldz, int64_t n,
                                                            int64_t m, int n_tiles, int64_t per,
                                                            double* __restrict__ partial, int64_t n_pad) {
  __shared__ double sa[TILE * G64_LD], sb[TILE * G64_LD];
  int64_t pair = blockIdx.x, ti = 0;
  while (pair >= n_tiles - ti) {
    pair -= n_tiles - ti;
    ++ti;
  }
  const int64_t tj = ti + pair;
  const bool diag = ti == tj;  // one operand tile serves both sides
  const int s = blockIdx.y;
  const int64_t k0 = (int64_t)s * per, k1 = k0 + per < m ? k0 + per : m;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int kk = tid % G64_KT, r0 = tid / G64_KT;

  f64x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  double ra[G64_LOADS], rb[G64_LOADS];
  auto fetch = [&](int64_t kb) {
    const int64_t k = kb + kk;
#pragma unroll
    for (int j = 0; j < G64_LOADS; ++j) {
      const int64_t ia = ti * TILE + r0 + j * G64_ROWS_PER_PASS, ib = tj * TILE + r0 + j * G64_ROWS_PER_PASS;
      ra[j] = (k < k1 && ia < n) ? Z[ia * ldz + k] : 0.0;
      if (!diag) rb[j] = (k < k1 && ib < n) ? Z[ib * ldz + k] : 0.0;
    }
  };
  const double* sbr = diag ? sa : sb;
  if (k0 < k1) fetch(k0);
  for (int64_t kb = k0; kb < k1; kb += G64_KT) {
#pragma unroll
    for (int j = 0; j < G64_LOADS; ++j) {
      sa[(r0 + j * G64_ROWS_PER_PASS) * G64_LD + kk] = ra[j];
      if (!diag) sb[(r0 + j * G64_ROWS_PER_PASS) * G64_LD + kk] = rb[j];
    }
    __syncthreads();
    if (kb + G64_KT < k1) fetch(kb + G64_KT);  // in flight during this stage's products
#pragma unroll
    for (int ks = 0; ks < G64_KT / 4; ++ks) {
      // A[i][k] of a 16 x 4 block at lane i + 16 k; B[k][j] = Z[j][k] at lane j + 16 k
      const int col = ks * 4 + lane / 16;
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sa[(wm * 64 + i * 16 + lane % 16) * G64_LD + col];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sbr[(wn * 64 + j * 16 + lane % 16) * G64_LD + col];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();  // the next stage overwrites the LDS
  }

  // accumulator item v of lane l holds row l / 16 + 4 v, column l % 16 of its 16 x 16 block
  double* out = partial + (int64_t)s * n_pad * n_pad;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = ti * TILE + wm * 64 + i * 16 + lane / 16 + 4 * v;
        const int64_t c = tj * TILE + wn * 64 + j * 16 + lane % 16;
        out[r * n_pad + c] = acc[i][j][v];
      }
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

hipError_t launch_gram_f64(const double* Z, int64_t ldz, int64_t n, int64_t m, int n_split, double* partial,
                           double* out, int64_t ld, hipStream_t stream) {
  const int64_t n_pad = (n + TILE - 1) / TILE * TILE;
  const int64_t stages = (m + G64_KT - 1) / G64_KT;
  const int64_t per = (stages + n_split - 1) / n_split * G64_KT;  // k per split, whole stages
  const dim3 grid((unsigned)gram_tile_pairs(n_pad), (unsigned)n_split);
  hipLaunchKernelGGL(gram_f64_kernel, grid, dim3(NTHREADS), 0, stream, Z, ldz, n, m, (int)(n_pad / TILE), per,
                     partial, n_pad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((unsigned)((n * n + 255) / 256)), dim3(256), 0, stream, partial, n_split,
                     n_pad, n, out, ld);
  return hipGetLastError();
}

template hipError_t launch_gram<PREC_SPLIT_FP16>(const TileArgs<PREC_SPLIT_FP16, 1>&, const float*, int, double*,
                                                  int64_t, double*, int64_t, hipStream_t);
template hipError_t launch_gram<PREC_SPLIT_BF16>(const TileArgs<PREC_SPLIT_BF16, 1>&, const float*, int, double*,
                                                  int64_t, double*, int64_t, hipStream_t);
template hipError_t launch_gram<PREC_FP32>(const TileArgs<PREC_FP32, 1>&, const float*, int, double*, int64_t, double*,
                                            int64_t, hipStream_t);

}  // namespace flmm
