// Rotation into the eigenbasis on MFMA: out[i][j] = sum_n A[i][n] * B[j][n], with A = rows to rotate (scan
// columns of (Y - X X^+ Y), or the SNPs of one chromosome) and B = the basis vectors as rows (orthonormal).
// This is U_eff^T Y_r and U_eff^T S of the reference (core.py:517-518, 628-629), computed as a split (or FP32)
// GEMM with FP32 accumulation.
//
// Epilogues write the next stage's operands directly (rows i, K index j):
//   MODE_UY:  one fp32 plane of the rotated columns, read by the delta search and the grid GEMM
//   MODE_US:  the scan's SNP-side planes S1 = x and S2 = x^2 (split pairs, or fp32)
//   MODE_US_X2: the S1 planes as MODE_US, and x^2 in FP64, row-major (a_rows_pad x b_rows_pad), for the factored
//             weights' S2 = x^2 A (a small FP64 GEMM by the caller)
// Split-FP16 operands arrive with per-row power-of-two scales (a_scale, b_scale) that the epilogue divides out
// exactly.  Because B's rows are unit vectors, |x[i][j]| <= ||A_i|| < 2^14 / a_scale[i], so the S1 / S2 planes are
// written with scales a_scale[i] and a_scale[i]^2 / 2^14 (out_scale1 / out_scale2).
#include <stdlib.h>
#include <string.h>

#include "kernels.h"

namespace flmm {

enum RotMode : int { MODE_UY = 0, MODE_US = 1, MODE_US_X2 = 2 };

template <int PREC, int MODE>
__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(3))) void rotate_kernel(TileArgs<PREC, 1> args, int64_t n_jtiles,
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
  tile_mainloop<PREC, 1, false, false>(args, i0, j0, acc, lds);

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
        if (MODE != MODE_UY && jt == 0 && wn == 0 && (lane & 31) == 0) {
          out_scale1[i] = (float)s1;
          if (MODE == MODE_US) out_scale2[i] = (float)s2;
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
        } else if constexpr (MODE == MODE_US_X2) {
          // lanes 0..31 of one store write 32 consecutive j of row i (256 contiguous bytes)
          ((double*)o2)[i * (n_jtiles * TILE) + j] = x * x;
          if constexpr (is_split(PREC)) {
            u16 h, l;
            split2<PREC>(x * s1, h, l);
            ((u16*)o0)[off] = h;
            ((u16*)o1)[off] = l;
          } else {
            ((float*)o0)[off] = (float)x;
          }
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

// MODE_UY in the split precisions on a 256 x 256 tile: 512 threads = 8 waves, two per SIMD; wave (wi, wj) owns rows
// 128 wi .. + 128 and columns 64 wj .. + 64 (4 x 2 blocks of 32 x 32, 128 accumulators), so a second wave per SIMD
// fills the first one's waits.  Each 32-k chunk of both operands (4 planes x 256 rows x 64 B = 64 KiB) is staged in
// LDS, double-buffered, one barrier per chunk; per MFMA this moves half the bytes of the 128 x 128 tile.  The products
// of an output element are summed in the same order as in rotate_kernel (per chunk and step: hi hi, hi lo, lo hi),
// so the results are bit-identical.
constexpr int RW_T = 256;                          // tile rows and columns
constexpr int RW_PLANE = RW_T * 64;                // bytes of one plane's chunk (256 rows x 32 two-byte values)
constexpr int RW_STAGE = 4 * RW_PLANE;             // A hi, A lo, B hi, B lo: 64 KiB

template <int PREC>
__global__ __launch_bounds__(512, 2) void rotate_uy_wide_kernel(TileArgs<PREC, 1> args, int64_t n_jtiles,
                                                                int64_t n_tiles, const float* __restrict__ a_scale,
                                                                const float* __restrict__ b_scale,
                                                                float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) char lds[2 * RW_STAGE];
  const int64_t tile = xcd_remap(blockIdx.x, n_tiles);
  const int64_t jt = tile % n_jtiles;  // j fastest: the tiles sharing one A block run together on one XCD
  const int64_t it = tile / n_jtiles;
  const int64_t i0 = it * RW_T, j0 = jt * RW_T;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wi = wave >> 2, wj = wave & 3;
  const int h = lane >> 5;
  const int nch = args.nchunks;

  // chunk loads: piece q = tid + 512 ps (ps = 0, 1) of a plane's chunk is row q / 4, 16-byte segment q % 4
  const char* const pl_base[4] = {(const char*)args.a[0][0] + i0 * 64, (const char*)args.a[0][1] + i0 * 64,
                                  (const char*)args.b[0][0] + j0 * 64, (const char*)args.b[0][1] + j0 * 64};
  const int64_t kstride[4] = {args.a_rows_pad * 64, args.a_rows_pad * 64, args.b_rows_pad * 64,
                              args.b_rows_pad * 64};
  auto gload = [&](int kc, u32x4 (&f)[8]) {
#pragma unroll
    for (int pl = 0; pl < 4; ++pl)
#pragma unroll
      for (int ps = 0; ps < 2; ++ps)
        f[pl * 2 + ps] = *gmem((const u32x4*)(pl_base[pl] + kc * kstride[pl] + (uint32_t)(tid + 512 * ps) * 16));
  };
  auto lstore = [&](const u32x4 (&f)[8], char* st) {
#pragma unroll
    for (int pl = 0; pl < 4; ++pl)
#pragma unroll
      for (int ps = 0; ps < 2; ++ps) {
        const int q = tid + 512 * ps, r = q >> 2, sg = q & 3;
        *(u32x4*)(st + pl * RW_PLANE + r * 64 + ((sg ^ ((r >> 2) & 3)) * 16)) = f[pl * 2 + ps];
      }
  };

  f32x16 acc[4][2];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[bi][bj][v] = 0.f;

  auto mma = [&](const char* st) {
#pragma unroll
    for (int step = 0; step < 2; ++step) {
      const int seg = h * 2 + step;
      bf16x8 af[4][2], bf[2][2];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) {
        const int r = wi * 128 + bi * 32 + (lane & 31);
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
          af[bi][sub] = *(const bf16x8*)(st + sub * RW_PLANE + r * 64 + ((seg ^ ((r >> 2) & 3)) * 16));
      }
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int r = wj * 64 + bj * 32 + (lane & 31);
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
          bf[bj][sub] = *(const bf16x8*)(st + (2 + sub) * RW_PLANE + r * 64 + ((seg ^ ((r >> 2) & 3)) * 16));
      }
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          f32x16 c = acc[bi][bj];
          c = mfma_split<PREC>(af[bi][0], bf[bj][0], c);
          c = mfma_split<PREC>(af[bi][0], bf[bj][1], c);
          c = mfma_split<PREC>(af[bi][1], bf[bj][0], c);
          acc[bi][bj] = c;
        }
    }
  };

  u32x4 pre[8];
  gload(0, pre);
  lstore(pre, lds);
  __syncthreads();
  for (int kc = 0; kc < nch; ++kc) {
    char* cur = lds + (kc & 1) * RW_STAGE;
    char* nxt = lds + ((kc & 1) ^ 1) * RW_STAGE;
    const bool more = kc + 1 < nch;
    if (more) gload(kc + 1, pre);
    mma(cur);
    if (more) lstore(pre, nxt);  // nxt was last read in the previous chunk, which ended with a barrier
    __syncthreads();
  }

  // epilogue: divide out the power-of-two row scales (split-FP16) and write the fp32 plane; lanes 0..31 of one store
  // write 32 consecutive j of row i (128 contiguous bytes)
  const int64_t rows_pad = args.a_rows_pad;
  double inv_b[2] = {1.0, 1.0};
  if constexpr (PREC == PREC_SPLIT_FP16) {
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) inv_b[bj] = 1.0 / (double)b_scale[j0 + wj * 64 + bj * 32 + (lane & 31)];
  }
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int64_t i = i0 + wi * 128 + bi * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
      const double inv_a = PREC == PREC_SPLIT_FP16 ? 1.0 / (double)a_scale[i] : 1.0;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int64_t j = j0 + wj * 64 + bj * 32 + (lane & 31);
        double x = (double)acc[bi][bj][v];
        if constexpr (PREC == PREC_SPLIT_FP16) x *= inv_a * inv_b[bj];
        out[((j >> 5) * rows_pad + i) * 32 + (j & 31)] = (float)x;
      }
    }
}

// MODE_UY in the split precisions on a 128 x 256 tile per 256-thread workgroup, two workgroups per CU: wave w owns
// the tile's 128 A rows (columns p) and B rows 64 w .. + 64 (basis directions k), 4 x 2 blocks of 32 x 32, 128
// accumulators.  Each 32-k chunk of the A planes (2 x 128 rows x 64 B = 16 KiB) is staged in LDS, double-buffered
// (32 KiB per workgroup); the B planes are small (the basis, L2-resident) and each wave loads its own fragments
// straight into registers one chunk ahead (lane (c, h) of block bj: 16 bytes at row 32 bj + c, segment 2 h + step).
// The products of an output element are summed in the same order as in rotate_kernel (per chunk and step: hi hi,
// hi lo, lo hi), so the results are bit-identical.  Two independent workgroups per CU drift apart, so one's load
// waits and epilogue overlap the other's MFMAs.
constexpr int RP_TA = 128;  // A rows (columns p) per tile
constexpr int RP_TB = 256;  // B rows (basis directions k) per tile

template <int PREC>
__global__ __launch_bounds__(256, 2) void rotate_uy_pair_kernel(TileArgs<PREC, 1> args, int64_t n_jtiles,
                                                                int64_t n_tiles, const float* __restrict__ a_scale,
                                                                const float* __restrict__ b_scale,
                                                                float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) char lds[2][2 * RP_TA * 64];  // [buffer][plane][row][64 B]
  const int64_t tile = xcd_remap(blockIdx.x, n_tiles);
  const int64_t jt = tile % n_jtiles;  // j fastest: the tiles sharing one A block run together on one XCD
  const int64_t it = tile / n_jtiles;
  const int64_t i0 = it * RP_TA, j0 = jt * RP_TB;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int h = lane >> 5;
  const int nch = args.nchunks;

  // A side: thread t copies 16-byte segment t % 4 of rows t / 4 and 64 + t / 4 of both planes of every chunk
  const char* const a_base[2] = {(const char*)args.a[0][0] + i0 * 64, (const char*)args.a[0][1] + i0 * 64};
  const int64_t a_kstride = args.a_rows_pad * 64;
  const uint32_t a_lane = tid * 16;
  // B side: lane (c, h) of block bj reads row j0 + 64 wave + 32 bj + c, bytes [32 h, 32 h + 32) of each chunk
  const char* const b_base[2] = {(const char*)args.b[0][0], (const char*)args.b[0][1]};
  const int64_t b_kstride = args.b_rows_pad * 64;
  uint32_t b_off[2];
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) b_off[bj] = (uint32_t)((j0 + wave * 64 + bj * 32 + (lane & 31)) * 64 + h * 32);

  struct ASet { u32x4 v[2][2]; };     // [plane][row half]
  struct BSet { u32x4 v[2][2][2]; };  // [block][step][plane]
  auto load_a = [&](int kc) {
    ASet a;
#pragma unroll
    for (int pl = 0; pl < 2; ++pl) {
      const char* sa = a_base[pl] + (int64_t)kc * a_kstride;
      a.v[pl][0] = *gmem((const u32x4*)(sa + a_lane));
      a.v[pl][1] = *gmem((const u32x4*)(sa + (a_lane + 4096)));
    }
    return a;
  };
  auto load_b = [&](int kc, BSet& b) {
#pragma unroll
    for (int pl = 0; pl < 2; ++pl) {
      const char* sb = b_base[pl] + (int64_t)kc * b_kstride;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int step = 0; step < 2; ++step) b.v[bj][step][pl] = *gmem((const u32x4*)(sb + (b_off[bj] + step * 16)));
    }
  };
  auto store = [&](const ASet& a, char* st) {
#pragma unroll
    for (int pl = 0; pl < 2; ++pl)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int r = (tid >> 2) + hf * 64;
        *(u32x4*)(st + pl * (RP_TA * 64) + r * 64 + (((tid & 3) ^ ((r >> 2) & 3)) * 16)) = a.v[pl][hf];
      }
  };

  f32x16 acc[4][2];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[bi][bj][v] = 0.f;

  auto mma = [&](const char* st, const BSet& b) {
#pragma unroll
    for (int step = 0; step < 2; ++step) {
      const int seg = h * 2 + step;
      bf16x8 af[4][2];
#pragma unroll
      for (int bi = 0; bi < 4; ++bi) {
        const int r = bi * 32 + (lane & 31);
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
          af[bi][sub] = *(const bf16x8*)(st + sub * (RP_TA * 64) + r * 64 + ((seg ^ ((r >> 2) & 3)) * 16));
      }
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const bf16x8 bh = __builtin_bit_cast(bf16x8, b.v[bj][step][0]);
          const bf16x8 bl = __builtin_bit_cast(bf16x8, b.v[bj][step][1]);
          f32x16 c = acc[bi][bj];
          c = mfma_split<PREC>(af[bi][0], bh, c);
          c = mfma_split<PREC>(af[bi][0], bl, c);
          c = mfma_split<PREC>(af[bi][1], bh, c);
          acc[bi][bj] = c;
        }
    }
  };
  auto chunk = [&](int kc, const char* cur, char* nxt, const BSet& bc, BSet& bn) {
    const int kn = kc + 1 < nch ? kc + 1 : kc;  // past the last chunk the prefetch repeats it (valid addresses)
    const ASet an = load_a(kn);
    load_b(kn, bn);
    mma(cur, bc);
    store(an, nxt);  // nxt was last read in the previous chunk, which ended with a barrier
    __syncthreads();
  };

  BSet b0, b1;
  load_b(0, b0);
  store(load_a(0), lds[0]);
  __syncthreads();
  int kc = 0;
  for (; kc + 2 <= nch; kc += 2) {
    chunk(kc, lds[0], lds[1], b0, b1);
    chunk(kc + 1, lds[1], lds[0], b1, b0);
  }
  if (kc < nch) chunk(kc, lds[0], lds[1], b0, b1);

  // epilogue: divide out the power-of-two row scales (split-FP16) and write the fp32 plane; lanes 0..31 of one store
  // write 32 consecutive j of row i (128 contiguous bytes)
  const int64_t rows_pad = args.a_rows_pad;
  double inv_b[2] = {1.0, 1.0};
  if constexpr (PREC == PREC_SPLIT_FP16) {
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) inv_b[bj] = 1.0 / (double)b_scale[j0 + wave * 64 + bj * 32 + (lane & 31)];
  }
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int64_t i = i0 + bi * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
      const double inv_a = PREC == PREC_SPLIT_FP16 ? 1.0 / (double)a_scale[i] : 1.0;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int64_t j = j0 + wave * 64 + bj * 32 + (lane & 31);
        double x = (double)acc[bi][bj][v];
        if constexpr (PREC == PREC_SPLIT_FP16) x *= inv_a * inv_b[bj];
        out[((j >> 5) * rows_pad + i) * 32 + (j & 31)] = (float)x;
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
  const int lds = tile_lds_bytes<PREC, 1, false>();
  if constexpr (is_split(PREC)) {
    // UY mode: the 128 x 256 tile at two workgroups per CU where the row counts allow it (a measured ~2.5% faster
    // than the 256 x 256 one at N = 1,000 x 655k columns); FLMM_ROTATE_VARIANT=wide keeps the 256 x 256 tile,
    // =square the 128 x 128 one (all three give the same bits)
    const char* var = getenv("FLMM_ROTATE_VARIANT");
    const bool square = var && strcmp(var, "square") == 0;
    const bool wide = var && strcmp(var, "wide") == 0;
    if (mode == MODE_UY && !square && !wide && args.a_rows_pad % RP_TA == 0 && args.b_rows_pad % RP_TB == 0) {
      const int64_t nj = args.b_rows_pad / RP_TB;
      const int64_t nt = (args.a_rows_pad / RP_TA) * nj;
      hipLaunchKernelGGL((rotate_uy_pair_kernel<PREC>), dim3((unsigned)nt), dim3(256), 0, stream, args, nj, nt,
                         a_scale, b_scale, (float*)o0);
      return hipGetLastError();
    }
    if (mode == MODE_UY && !square && args.a_rows_pad % RW_T == 0 && args.b_rows_pad % RW_T == 0) {
      const int64_t nj = args.b_rows_pad / RW_T;
      const int64_t nt = (args.a_rows_pad / RW_T) * nj;
      hipLaunchKernelGGL((rotate_uy_wide_kernel<PREC>), dim3((unsigned)nt), dim3(512), 0, stream, args, nj, nt,
                         a_scale, b_scale, (float*)o0);
      return hipGetLastError();
    }
  }
  if (mode == MODE_UY)
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_UY>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, a_scale, b_scale, o0, o1, o2, o3, out_scale1, out_scale2);
  else if (mode == MODE_US_X2)
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_US_X2>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, a_scale, b_scale, o0, o1, o2, o3, out_scale1, out_scale2);
  else
    hipLaunchKernelGGL((rotate_kernel<PREC, MODE_US>), dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args,
                       n_jtiles, n_tiles, a_scale, b_scale, o0, o1, o2, o3, out_scale1, out_scale2);
  return hipGetLastError();
}

// Row GEMM in FP64: C[i][r] = sum_{k < kpad} X2[i][k] A[k][r] (X2: rows x kpad, A: kpad x rpad, C: rows x rpad,
// all row-major; rows % 32 == 0, kpad % 32 == 0, rpad % 64 == 0) -- the factored weights' SNP side (X2 = US^2)
// and the FP64 rotations of the re-scoring (X2 = S^T or Y^T, A = U_eff).  Each element sums k in order, so a row's
// result does not depend on how many rows a launch holds.  A 256-thread block owns 32 rows x 64 r (a 4,608-row
// SNP block against 128 terms is 288 blocks): thread (g = tid / 32, c = tid % 32) accumulates rows 4 g .. 4 g + 3
// at r = c and c + 32, 8 FP64 sums, from K chunks of 32 staged in LDS (A chunk 32 x 64, X2 chunk 32 x 32 padded
// against bank conflicts).
__global__ __launch_bounds__(256) void gemm_f64_kernel(const double* __restrict__ x2, int64_t kpad,
                                                       const double* __restrict__ A, int64_t rpad,
                                                       double* __restrict__ C) {
  __shared__ double as[32][64];
  __shared__ double xs[32][33];
  const int tid = threadIdx.x, g = tid >> 5, c = tid & 31;
  const int64_t i0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 64;
  double acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = 0.0;
  for (int64_t k0 = 0; k0 < kpad; k0 += 32) {
#pragma unroll
    for (int e = tid; e < 32 * 64; e += 256) as[e >> 6][e & 63] = A[(k0 + (e >> 6)) * rpad + r0 + (e & 63)];
#pragma unroll
    for (int e = tid; e < 32 * 32; e += 256) xs[e >> 5][e & 31] = x2[(i0 + (e >> 5)) * kpad + k0 + (e & 31)];
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 32; ++k) {
      const double a0 = as[k][c], a1 = as[k][c + 32];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double x = xs[4 * g + i][k];
        acc[i][0] = fma(x, a0, acc[i][0]);
        acc[i][1] = fma(x, a1, acc[i][1]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    C[(i0 + 4 * g + i) * rpad + r0 + c] = acc[i][0];
    C[(i0 + 4 * g + i) * rpad + r0 + c + 32] = acc[i][1];
  }
}

hipError_t launch_gemm_f64(const double* x2, int64_t rows, int64_t kpad, const double* A, int64_t rpad, double* C,
                           hipStream_t stream) {
  const dim3 grid((unsigned)(rows / 32), (unsigned)(rpad / 64));
  hipLaunchKernelGGL(gemm_f64_kernel, grid, dim3(256), 0, stream, x2, kpad, A, rpad, C);
  return hipGetLastError();
}

__device__ inline void rescore_one(const double* __restrict__ US, const double* __restrict__ UY, int64_t ld,
                                   const double* __restrict__ s, int64_t neff, const double* __restrict__ log_delta,
                                   int64_t m, int64_t p, int64_t orow, double nu, double* __restrict__ f,
                                   double* __restrict__ beta, double* __restrict__ se, double* __restrict__ sfve,
                                   int64_t out_ld, int lane) {
  const double delta = exp(log_delta[p]);
  const double* u = US + m * ld;
  const double* y = UY + p * ld;
  double a = 0.0, d = 0.0, r = 0.0;
  for (int64_t k = lane; k < neff; k += 64) {
    const double w = 1.0 / (s[k] + delta);
    const double uk = u[k], yk = y[k];
    a += uk * w * yk;
    d += uk * uk * w;
    r += w * yk * yk;
  }
  a = wave_sum(a);
  d = fmax(wave_sum(d), FLOOR64);
  r = fmax(wave_sum(r), FLOOR64);
  if (lane == 0) {
    const double b = a / d;
    const double rss = fmax(r - a * a / d, FLOOR64);
    const double var = rss / nu / d;
    const int64_t o = orow * out_ld + p;
    f[o] = b * b / var;
    beta[o] = b;
    se[o] = sqrt(var);
    sfve[o] = sqrt(fmax(a * a / (d * r), 0.0));
  }
}

// FP64 re-evaluation of selected tests i < n, (SNP row us_row[i] of US, column col[i]): with w_k = 1 / (s_k + delta),
// delta = exp(log_delta[col]), the exact sums a = sum_k US w UY, d = sum_k US^2 w and rWr = sum_k w UY^2 over the
// FP64 operands (US: SNP rows, UY: column rows, both with leading dimension ld), then the scan epilogue's Wald
// arithmetic (core.py:649-667) in FP64, written at [row * out_ld + col[i]], row = out_row[i] (or out_row_base +
// us_row[i] without out_row).  One wave per test, lanes striding k, fixed-order wave sums.  With n_dev the number of
// tests is min(*n_dev, n), read on the device, and the waves stride over them (a fixed grid).
__global__ __launch_bounds__(256) void rescore_kernel(const double* __restrict__ US, const double* __restrict__ UY,
                                                      int64_t ld, const double* __restrict__ s, int64_t neff,
                                                      const double* __restrict__ log_delta,
                                                      const int64_t* __restrict__ us_row,
                                                      const int64_t* __restrict__ col,
                                                      const int64_t* __restrict__ out_row, int64_t out_row_base,
                                                      int64_t n, const unsigned long long* __restrict__ n_dev,
                                                      double nu, double* __restrict__ f, double* __restrict__ beta,
                                                      double* __restrict__ se, double* __restrict__ sfve,
                                                      int64_t out_ld) {
  const int lane = threadIdx.x & 63;
  int64_t nt = n;
  if (n_dev) {
    const unsigned long long c = *n_dev;
    nt = c < (unsigned long long)n ? (int64_t)c : n;
  }
  for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < nt; i += (int64_t)gridDim.x * 4)  // wave-uniform
    rescore_one(US, UY, ld, s, neff, log_delta, us_row[i], col[i], out_row ? out_row[i] : out_row_base + us_row[i],
                nu, f, beta, se, sfve, out_ld, lane);
}

// The tests the re-scoring takes: every (row, column) of f (rows x cols, leading dimension ld) with f > thr, appended
// at an atomic counter (one add per wave: ballot, then each lane's rank below it); entries past cap are counted, not
// written.  The order of the list is arbitrary; the re-scoring of each test does not depend on it.
__global__ __launch_bounds__(256) void select_above_kernel(const double* __restrict__ f, int64_t rows, int64_t cols,
                                                           int64_t ld, double thr, int64_t* __restrict__ out_row,
                                                           int64_t* __restrict__ out_col, int64_t cap,
                                                           unsigned long long* __restrict__ count) {
  const int lane = threadIdx.x & 63;
  const int64_t n = rows * cols;
  for (int64_t base = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63); base < n; base += (int64_t)gridDim.x * 256) {
    const int64_t i = base + lane;
    const int64_t r = i / cols, c = i - r * cols;
    const bool hit = i < n && gmem(f)[r * ld + c] > thr;
    const uint64_t b = __ballot(hit);
    if (!b) continue;
    unsigned long long first = 0;
    if (lane == __builtin_ctzll(b)) first = atomicAdd(count, (unsigned long long)__popcll(b));
    first = __shfl(first, __builtin_ctzll(b), 64);
    if (hit) {
      const int64_t slot = (int64_t)first + __popcll(b & ((1ull << lane) - 1ull));
      if (slot < cap) {
        out_row[slot] = r;
        out_col[slot] = c;
      }
    }
  }
}

hipError_t launch_select_above(const double* f, int64_t rows, int64_t cols, int64_t ld, double thr, int64_t* out_row,
                               int64_t* out_col, int64_t cap, unsigned long long* count, hipStream_t stream) {
  const int64_t n = rows * cols;
  const int64_t blocks = n / 256 + 1 < 4096 ? n / 256 + 1 : 4096;
  hipLaunchKernelGGL(select_above_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, f, rows, cols, ld, thr,
                     out_row, out_col, cap, count);
  return hipGetLastError();
}

hipError_t launch_rescore(const double* US, const double* UY, int64_t ld, const double* s, int64_t neff,
                          const double* log_delta, const int64_t* us_row, const int64_t* col, const int64_t* out_row,
                          int64_t out_row_base, int64_t n, const unsigned long long* n_dev, double nu, double* f,
                          double* beta, double* se, double* sfve, int64_t out_ld, hipStream_t stream) {
  // one wave per test, or with a device count a fixed grid of 4,096 waves striding over the tests
  const int64_t blocks = n_dev ? (n + 3) / 4 < 1024 ? (n + 3) / 4 : 1024 : (n + 3) / 4;
  hipLaunchKernelGGL(rescore_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, US, UY, ld, s, neff, log_delta,
                     us_row, col, out_row, out_row_base, n, n_dev, nu, f, beta, se, sfve, out_ld);
  return hipGetLastError();
}

// FP64 rotation of SNP columns on the FP64 matrix cores (v_mfma_f64_16x16x4f64): out[i][j] = sum_n z[n][c(i)]
// u[n][j], c(i) = col_index[i] (or i), for i < mc, j < neff, zeros elsewhere in the rows_pad x kpad output -- U_eff^T
// S_r of core.py:628-629 evaluated in FP64, whose products enter the scan weighted by 1 / (s_k + delta), large where
// both are small.  One 256-thread workgroup per 128 x 128 output tile, the four waves 2 x 2, each with 4 x 4 blocks
// of 16 x 16 FP64 accumulators; both operands are strain-major (k = strain), staged 16 strains at a time in
// k-major LDS rows 144 doubles apart (a half-wave's 8-byte reads of 16 consecutive entries of two k rows cover the
// 64 banks once); the next stage is fetched into registers while the current one is multiplied.  The 64 FP64
// accumulators per lane live in VGPRs capped at two waves per SIMD, so a second block's products fill the first
// one's barriers (one wave per SIMD left the matrix pipe idle through every LDS stage).
constexpr int R64_KT = 16;
constexpr int R64_LD = TILE + 16;
constexpr int R64_LOADS = R64_KT / (NTHREADS / TILE);  // k rows per thread and stage (8)

typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(NTHREADS) __attribute__((amdgpu_waves_per_eu(2))) void rotate_f64_kernel(
    const double* __restrict__ z, int64_t ldz,
                                                              const int64_t* __restrict__ col_index, int64_t mc,
                                                              int64_t n, const double* __restrict__ u, int64_t ldu,
                                                              int64_t neff, double* __restrict__ out, int64_t ldo) {
  __shared__ double sa[R64_KT * R64_LD], sb[R64_KT * R64_LD];
  // one-dimensional grid over (row tile, column tile), column tiles fastest and XCD-contiguous, so the blocks
  // that share a row tile of z run together on one L2
  const int64_t n_jt = ldo / TILE;
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t i0 = (tile / n_jt) * TILE, j0 = (tile % n_jt) * TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = tid % TILE, lk = tid / TILE;  // this thread stages entry li of k rows lk, lk + 2, ..., lk + 14
  const bool has_a = i0 + li < mc, has_b = j0 + li < neff;
  const int64_t zc = has_a ? (col_index ? col_index[i0 + li] : i0 + li) : 0;
  const int64_t uc = j0 + li;

  f64x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  double ra[R64_LOADS], rb[R64_LOADS];
  auto fetch = [&](int64_t kb) {
#pragma unroll
    for (int q = 0; q < R64_LOADS; ++q) {
      const int64_t k = kb + lk + 2 * q;
      ra[q] = (has_a && k < n) ? z[k * ldz + zc] : 0.0;
      rb[q] = (has_b && k < n) ? u[k * ldu + uc] : 0.0;
    }
  };
  fetch(0);
  for (int64_t kb = 0; kb < n; kb += R64_KT) {
#pragma unroll
    for (int q = 0; q < R64_LOADS; ++q) {
      sa[(lk + 2 * q) * R64_LD + li] = ra[q];
      sb[(lk + 2 * q) * R64_LD + li] = rb[q];
    }
    __syncthreads();
    if (kb + R64_KT < n) fetch(kb + R64_KT);  // in flight during this stage's products
#pragma unroll
    for (int ks = 0; ks < R64_KT / 4; ++ks) {
      // A[i][k] of a 16 x 4 block at lane i + 16 k, B[k][j] at lane j + 16 k
      const int row = (ks * 4 + lane / 16) * R64_LD + lane % 16;
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sa[row + wm * 64 + i * 16];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sb[row + wn * 64 + j * 16];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();  // the next stage overwrites the LDS
  }

  // accumulator item v of lane l holds row l / 16 + 4 v, column l % 16 of its 16 x 16 block
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = i0 + wm * 64 + i * 16 + lane / 16 + 4 * v;
        const int64_t c = j0 + wn * 64 + j * 16 + lane % 16;
        out[r * ldo + c] = acc[i][j][v];
      }
}

hipError_t launch_rotate_f64(const double* z, int64_t ldz, const int64_t* col_index, int64_t mc, int64_t n,
                             const double* u, int64_t ldu, int64_t neff, double* out, int64_t rows_pad, int64_t kpad,
                             hipStream_t stream) {
  const dim3 grid((unsigned)((rows_pad / TILE) * (kpad / TILE)));
  hipLaunchKernelGGL(rotate_f64_kernel, grid, dim3(NTHREADS), 0, stream, z, ldz, col_index, mc, n, u, ldu, neff, out,
                     kpad);
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
