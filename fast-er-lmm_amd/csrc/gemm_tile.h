// MFMA tile engine for "NT" products out[i][j] = sum_k A[i][k] * B[j][k] on gfx950.
//
// One workgroup = 256 threads = 4 wave64s computes a 128 x 128 output tile; wave w owns the 64 x 64 quadrant
// (w / 2, w % 2), i.e. 2 x 2 blocks of the 32 x 32 MFMA shape.  NG independent products that share the tile
// (the Wald scan needs two: sum w*y*s and sum w*s^2) accumulate side by side.
//
// K is walked in chunks of 32.  Each chunk of every operand plane is one contiguous block in HBM (see
// flmm_common.h), copied by all 256 threads with 16-byte loads into a register prefetch buffer while the
// previous chunk is multiplied out of LDS.  LDS rows are XOR-swizzled in 16-byte units so that the 16 lanes of
// a ds_read_b128 group (16 consecutive rows, same logical segment) hit 16 distinct bank groups.
//
// Precision modes
//   PREC_SPLIT_BF16: operands are bf16 (hi, lo) plane pairs; acc += hi*hi + hi*lo + lo*hi through
//                    v_mfma_f32_32x32x16_bf16 (FP32 accumulation).  The lo*lo term (~2^-16 relative) is dropped.
//   PREC_SPLIT_FP16: the same three products on power-of-two-scaled fp16 (hi, lo) pairs through
//                    v_mfma_f32_32x32x16_f16; ~22 significant bits per operand at the same MFMA rate.
//   PREC_FP32:       operands are fp32 planes; acc += a*b through v_mfma_f32_32x32x2_f32.
//
// K-slot mapping.  MFMA sums over the k-slots of its A and B fragments pairwise (A lane l, element e with B
// lane l, element e), so any assignment of physical k to (lane half, element) that is the SAME for A and B
// computes the same sum.  Each lane therefore reads 16 consecutive physical k of its row: lane half h = l / 32
// covers k = 16h .. 16h+15 of the chunk, which makes every fragment read a contiguous ds_read_b128.
//
// Output layout of a 32 x 32 block (f32x16 acc): lane l holds column j = l % 32 and rows
// i = 8 * (v / 4) + 4 * (l / 32) + v % 4 for v = 0..15.
#pragma once
#include "flmm_common.h"

namespace flmm {

constexpr int TILE = 128;
constexpr int NTHREADS = 256;

template <int PREC>
struct PlaneTraits;
template <>
struct PlaneTraits<PREC_SPLIT_BF16> {
  static constexpr int NSUB = 2;     // planes per operand (hi, lo)
  static constexpr int ESIZE = 2;    // bytes per element
  static constexpr int ROWB = 64;    // bytes per row per chunk
  static constexpr int SPR = 4;      // 16-byte segments per row
  static __device__ inline int swz(int row) { return (row >> 2) & 3; }
};
template <>
struct PlaneTraits<PREC_SPLIT_FP16> : PlaneTraits<PREC_SPLIT_BF16> {};
template <>
struct PlaneTraits<PREC_FP32> {
  static constexpr int NSUB = 1;
  static constexpr int ESIZE = 4;
  static constexpr int ROWB = 128;
  static constexpr int SPR = 8;
  static __device__ inline int swz(int row) { return (row >> 1) & 7; }
};

// one split product term: v_mfma_f32_32x32x16_{bf16,f16} on the raw 16-byte fragments
template <int PREC>
__device__ inline f32x16 mfma_split(bf16x8 a, bf16x8 b, f32x16 c) {
  if constexpr (PREC == PREC_SPLIT_FP16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0,
                                                   0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

template <int PREC, int NG>
struct TileArgs {
  const void* a[NG][2];  // A-side planes, [g][sub]
  const void* b[NG][2];  // B-side planes
  int64_t a_rows_pad;    // rows_pad of every A plane
  int64_t b_rows_pad;
  int nchunks;           // K / 32
  int nchunks2;          // K / 32 of the second product where it is shorter (low-rank scan only)
};

// LDS bytes needed by one workgroup: both sides' planes of one K chunk, twice when double-buffered (DB)
template <int PREC, int NG, bool DB = (NG == 1)>
constexpr int tile_lds_bytes() {
  return (DB ? 2 : 1) * 2 * NG * PlaneTraits<PREC>::NSUB * TILE * PlaneTraits<PREC>::ROWB;
}

template <int PREC, int NG>
struct TileAcc {
  f32x16 c[NG][2][2];
};

// Accumulate the full K extent of the 128 x 128 tile whose first rows are a_row0 (A side) and b_row0 (B side).
template <int PREC, int NG, bool SQUARE_B = false, bool DB = (NG == 1)>
__device__ inline void tile_mainloop(const TileArgs<PREC, NG>& args, int64_t a_row0, int64_t b_row0,
                                     TileAcc<PREC, NG>& acc, char* lds) {
  using T = PlaneTraits<PREC>;
  constexpr int NSUB = T::NSUB;
  constexpr int NPL = NG * NSUB;                    // planes per side
  constexpr int PLANE_LDS = TILE * T::ROWB;         // bytes per plane in LDS
  constexpr int ROWS_PER_PASS = NTHREADS / T::SPR;  // rows covered by one 16-byte load per thread
  constexpr int PASSES = TILE / ROWS_PER_PASS;
  constexpr int NLOAD = 2 * NPL * PASSES;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc.c[g][i][j][v] = 0.f;

  // per-thread load slot inside a chunk: row r_in = tid / SPR + pass * ROWS_PER_PASS, segment s_in = tid % SPR
  const int s_in = tid % T::SPR;
  const int r_in0 = tid / T::SPR;

  uint4 pre[NLOAD];
  auto gload = [&](int kc) {
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl) {
        const int g = pl / NSUB, sub = pl % NSUB;
        const char* base = (const char*)(side == 0 ? args.a[g][sub] : args.b[g][sub]);
        const int64_t rows_pad = side == 0 ? args.a_rows_pad : args.b_rows_pad;
        const int64_t row0 = side == 0 ? a_row0 : b_row0;
        const char* chunk = base + ((int64_t)kc * rows_pad + row0) * T::ROWB;
#pragma unroll
        for (int ps = 0; ps < PASSES; ++ps) {
          const int r = r_in0 + ps * ROWS_PER_PASS;
          pre[(side * NPL + pl) * PASSES + ps] =
              __builtin_bit_cast(uint4, *gmem((const u32x4*)(chunk + (int64_t)r * T::ROWB + s_in * 16)));
        }
      }
  };
  // DB (the default for NG == 1) double-buffers the chunks in LDS: one barrier per chunk, chunk k + 1 stored into
  // the other buffer while chunk k is multiplied.  NG == 2 keeps one buffer (its planes fill twice the LDS), and so
  // does the delta grid, whose small FP32 tiles run faster at the higher occupancy one buffer leaves.
  constexpr int BUF = 2 * NPL * PLANE_LDS;
  auto lstore = [&](int buf) {
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
        for (int ps = 0; ps < PASSES; ++ps) {
          const int r = r_in0 + ps * ROWS_PER_PASS;
          char* dst = lds + buf * BUF + (side * NPL + pl) * PLANE_LDS + r * T::ROWB + ((s_in ^ T::swz(r)) * 16);
          uint4 v = pre[(side * NPL + pl) * PASSES + ps];
          if constexpr (SQUARE_B && PREC == PREC_FP32) {
            if (side == 1) {
              float4 f = *(float4*)&v;
              f.x *= f.x; f.y *= f.y; f.z *= f.z; f.w *= f.w;
              v = *(uint4*)&f;
            }
          }
          *(uint4*)dst = v;
        }
  };

  gload(0);
  if constexpr (DB) {
    if (args.nchunks > 0) {
      __syncthreads();  // the caller's previous use of the LDS is over
      lstore(0);
      __syncthreads();
    }
  }
  for (int kc = 0; kc < args.nchunks; ++kc) {
    const int cur = DB ? (kc & 1) : 0;
    if constexpr (!DB) {
      __syncthreads();  // previous chunk fully consumed
      lstore(0);
      __syncthreads();
    }
    if (kc + 1 < args.nchunks) gload(kc + 1);
    const char* cl = lds + cur * BUF;

    const int h = lane >> 5;
    if constexpr (is_split(PREC)) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const int seg = h * 2 + ks;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          bf16x8 af[2][2], bfr[2][2];  // [block][sub]
#pragma unroll
          for (int bi = 0; bi < 2; ++bi) {
            const int r = wm * 64 + bi * 32 + (lane & 31);
#pragma unroll
            for (int sub = 0; sub < 2; ++sub) {
              const char* src = cl + (0 * NPL + g * 2 + sub) * PLANE_LDS + r * 64 + ((seg ^ T::swz(r)) * 16);
              af[bi][sub] = *(const bf16x8*)src;
            }
          }
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            const int r = wn * 64 + bj * 32 + (lane & 31);
#pragma unroll
            for (int sub = 0; sub < 2; ++sub) {
              const char* src = cl + (1 * NPL + g * 2 + sub) * PLANE_LDS + r * 64 + ((seg ^ T::swz(r)) * 16);
              bfr[bj][sub] = *(const bf16x8*)src;
            }
          }
#pragma unroll
          for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) {
              f32x16 c = acc.c[g][bi][bj];
              c = mfma_split<PREC>(af[bi][0], bfr[bj][0], c);
              c = mfma_split<PREC>(af[bi][0], bfr[bj][1], c);
              c = mfma_split<PREC>(af[bi][1], bfr[bj][0], c);
              acc.c[g][bi][bj] = c;
            }
        }
      }
    } else {
      // 4 sub-steps of 4 physical k each; lane half h covers k = 16h .. 16h+15
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int seg = h * 4 + q;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          f32x4 af[2], bfr[2];
#pragma unroll
          for (int bi = 0; bi < 2; ++bi) {
            const int r = wm * 64 + bi * 32 + (lane & 31);
            af[bi] = *(const f32x4*)(cl + (0 * NPL + g) * PLANE_LDS + r * 128 + ((seg ^ T::swz(r)) * 16));
          }
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            const int r = wn * 64 + bj * 32 + (lane & 31);
            bfr[bj] = *(const f32x4*)(cl + (1 * NPL + g) * PLANE_LDS + r * 128 + ((seg ^ T::swz(r)) * 16));
          }
#pragma unroll
          for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
              for (int bj = 0; bj < 2; ++bj)
                acc.c[g][bi][bj] =
                    __builtin_amdgcn_mfma_f32_32x32x2f32(af[bi][e], bfr[bj][e], acc.c[g][bi][bj], 0, 0, 0);
        }
      }
    }
    if constexpr (DB) {
      if (kc + 1 < args.nchunks) lstore(cur ^ 1);  // that buffer was last read before the previous barrier
      __syncthreads();
    }
  }
}

// row / column of accumulator element v of block (bi, bj) for this lane, relative to the tile origin
__device__ inline int acc_row(int wm, int bi, int lane, int v) {
  return wm * 64 + bi * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
}
__device__ inline int acc_col(int wn, int bj, int lane) { return wn * 64 + bj * 32 + (lane & 31); }

}  // namespace flmm
