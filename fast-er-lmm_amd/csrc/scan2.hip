// Wald scan, variant "wide": 128 SNP rows x 256 columns per workgroup.  Same arithmetic and epilogue as
// scan.hip (reference fasterlmm/core.py:649-667); different data movement:
//   * the SNP-side planes (S1 = US, S2 = US^2) are shared by the four waves and staged through LDS, double
//     buffered (one barrier per K chunk);
//   * each wave owns 64 columns and streams its column-side fragments (Q1 = w o UY, Q2 = w) straight from global
//     memory into registers, one K chunk ahead -- a 32-row block of one chunk is 2 KiB contiguous in the plane
//     layout, so each fragment load is a fully coalesced 32-byte-per-lane read and needs no LDS at all;
//   * a wave covers all 128 rows of the tile, so the per-column maximum needs no cross-wave exchange.
// LDS traffic per MFMA is ~40% of the square-tile variant's, which is what bounds it.
#include "kernels.h"

namespace flmm {

constexpr int W_TM = 128;   // SNP rows per workgroup
constexpr int W_TP = 256;   // columns per workgroup (64 per wave)

template <int PREC>
struct WideTraits;
template <>
struct WideTraits<PREC_SPLIT_BF16> {
  static constexpr int NSUB = 2;
  static constexpr int ROWB = 64;      // bytes per row per K chunk
  static constexpr int BREGS = 2;      // uint4 per lane per (g, bj, sub) per chunk: 32 bytes = 16 bf16
};
template <>
struct WideTraits<PREC_SPLIT_FP16> : WideTraits<PREC_SPLIT_BF16> {};
template <>
struct WideTraits<PREC_FP32> {
  static constexpr int NSUB = 1;
  static constexpr int ROWB = 128;
  static constexpr int BREGS = 4;      // 64 bytes = 16 floats
};

template <int PREC>
__device__ inline int wide_swz(int row);
template <>
__device__ inline int wide_swz<PREC_SPLIT_BF16>(int row) { return (row >> 2) & 3; }
template <>
__device__ inline int wide_swz<PREC_SPLIT_FP16>(int row) { return (row >> 2) & 3; }
template <>
__device__ inline int wide_swz<PREC_FP32>(int row) { return (row >> 1) & 7; }

template <int PREC>
__global__ __launch_bounds__(256, 1) void wald_scan_wide_kernel(
    TileArgs<PREC, 2> args, ScanScales sc, int64_t n_mtiles, int64_t n_tiles, int64_t Mc, int64_t P, int64_t n_real,
    const double* __restrict__ rWr, const double* __restrict__ h2, double nu, const int64_t* __restrict__ row_index,
    ScanOut out, double* __restrict__ pmax, int64_t pmax_ld) {
  using T = WideTraits<PREC>;
  constexpr int NSUB = T::NSUB;
  constexpr int NPL = 2 * NSUB;                       // SNP-side planes
  constexpr int PLANE_LDS = W_TM * T::ROWB;           // bytes per plane per stage
  constexpr int STAGE = NPL * PLANE_LDS;              // bytes per stage (32 KiB)
  constexpr int SPR = T::ROWB / 16;
  constexpr int ROWS_PER_PASS = 256 / SPR;
  constexpr int PASSES = W_TM / ROWS_PER_PASS;
  constexpr int NA = NPL * PASSES;                    // uint4 per thread per chunk (A staging)
  constexpr int NB = 2 * 2 * NSUB * T::BREGS;         // uint4 per lane per chunk (B fragments)

  extern __shared__ char lds[];
  const int64_t tile = xcd_remap(blockIdx.x, n_tiles);
  const int64_t mt = tile % n_mtiles;
  const int64_t pt = tile / n_mtiles;
  const int64_t m0 = mt * W_TM, p0 = pt * W_TP;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = lane >> 5;
  const int nchunks = args.nchunks;

  f32x16 acc[2][4][2];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[g][i][j][v] = 0.f;

  const int s_in = tid % SPR;
  const int r_in0 = tid / SPR;
  uint4 astage[NA];
  uint4 bcur[NB], bnext[NB];

  auto aload = [&](int kc) {
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl) {
      const char* base = (const char*)args.a[pl / NSUB][pl % NSUB];
      const char* chunk = base + ((int64_t)kc * args.a_rows_pad + m0) * T::ROWB;
#pragma unroll
      for (int ps = 0; ps < PASSES; ++ps) {
        const int r = r_in0 + ps * ROWS_PER_PASS;
        astage[pl * PASSES + ps] = *(const uint4*)(chunk + (int64_t)r * T::ROWB + s_in * 16);
      }
    }
  };
  auto astore = [&](int buf) {
    char* st = lds + buf * STAGE;
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
      for (int ps = 0; ps < PASSES; ++ps) {
        const int r = r_in0 + ps * ROWS_PER_PASS;
        *(uint4*)(st + pl * PLANE_LDS + r * T::ROWB + ((s_in ^ wide_swz<PREC>(r)) * 16)) = astage[pl * PASSES + ps];
      }
  };
  // this lane's fragment of column block (g, bj, sub) for chunk kc: row p0 + 64 wave + 32 bj + lane % 32,
  // bytes [h * ROWB / 2, (h + 1) * ROWB / 2) of the row's chunk -- 16 consecutive physical k
  auto bload = [&](int kc, uint4* dst) {
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int sub = 0; sub < NSUB; ++sub) {
          const char* base = (const char*)args.b[g][sub];
          const int64_t row = p0 + wave * 64 + bj * 32 + (lane & 31);
          const char* src = base + ((int64_t)kc * args.b_rows_pad + row) * T::ROWB + h * (T::ROWB / 2);
#pragma unroll
          for (int q = 0; q < T::BREGS; ++q)
            dst[((g * 2 + bj) * NSUB + sub) * T::BREGS + q] = *(const uint4*)(src + q * 16);
        }
  };

  aload(0);
  bload(0, bcur);
  astore(0);
  __syncthreads();
  for (int kc = 0; kc < nchunks; ++kc) {
    const int buf = kc & 1;
    const bool more = kc + 1 < nchunks;
    if (more) {
      aload(kc + 1);
      bload(kc + 1, bnext);
    }
    const char* st = lds + buf * STAGE;
    if constexpr (is_split(PREC)) {
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        const int seg = h * 2 + ks;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          bf16x8 af[4][2];
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) {
            const int r = bi * 32 + (lane & 31);
#pragma unroll
            for (int sub = 0; sub < 2; ++sub)
              af[bi][sub] = *(const bf16x8*)(st + (g * 2 + sub) * PLANE_LDS + r * 64 + ((seg ^ wide_swz<PREC>(r)) * 16));
          }
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            const bf16x8 bh = *(const bf16x8*)&bcur[((g * 2 + bj) * 2 + 0) * 2 + ks];
            const bf16x8 bl = *(const bf16x8*)&bcur[((g * 2 + bj) * 2 + 1) * 2 + ks];
#pragma unroll
            for (int bi = 0; bi < 4; ++bi) {
              f32x16 c = acc[g][bi][bj];
              c = mfma_split<PREC>(af[bi][0], bh, c);
              c = mfma_split<PREC>(af[bi][0], bl, c);
              c = mfma_split<PREC>(af[bi][1], bh, c);
              acc[g][bi][bj] = c;
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int seg = h * 4 + q;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
          f32x4 af[4];
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) {
            const int r = bi * 32 + (lane & 31);
            af[bi] = *(const f32x4*)(st + g * PLANE_LDS + r * 128 + ((seg ^ wide_swz<PREC>(r)) * 16));
          }
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            const f32x4 bf = *(const f32x4*)&bcur[(g * 2 + bj) * 4 + q];
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
              for (int bi = 0; bi < 4; ++bi)
                acc[g][bi][bj] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[bi][e], bf[e], acc[g][bi][bj], 0, 0, 0);
          }
        }
      }
    }
    if (more) {
      astore(buf ^ 1);
#pragma unroll
      for (int q = 0; q < NB; ++q) bcur[q] = bnext[q];
    }
    __syncthreads();
  }

  // epilogue: this wave owns columns p0 + 64 wave + [0, 64) and all 128 rows
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int64_t p = p0 + wave * 64 + bj * 32 + (lane & 31);
    const bool pv = p < P;
    const double rwr = pv ? rWr[p] : 1.0;
    const bool real = p < n_real;
    const double hh = real ? h2[p] : 0.0;
    const double inv_q1 = (sc.q1 && pv) ? 1.0 / (double)sc.q1[p] : 1.0;
    const double inv_q2 = (sc.q2 && pv) ? 1.0 / (double)sc.q2[p] : 1.0;
    double ba = 0.0, bd = 1.0;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t m = m0 + bi * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
        if (m < Mc) {
          double a = (double)acc[0][bi][bj][v];
          double d = (double)acc[1][bi][bj][v];
          unscale_sums(sc, m, inv_q1, inv_q2, a, d);
          d = fmax(d, FLOOR64);
          if (a * a * bd > ba * ba * d) { ba = a; bd = d; }
          if (real) {
            const double beta = a / d;
            const double rss = fmax(rwr - a * a / d, FLOOR64);
            const double var = rss / nu / d;
            const int64_t o = row_index[m] * out.ld + p;
            out.f[o] = beta * beta / var;
            out.beta[o] = beta;
            out.se[o] = sqrt(var);
            out.sfve[o] = sqrt(fmax(a * a / (d * rwr), 0.0));
            out.nullh2[o] = hh;
          }
        }
      }
    }
    const double oa = __shfl_xor(ba, 32, 64), od = __shfl_xor(bd, 32, 64);
    if (oa * oa * bd > ba * ba * od) { ba = oa; bd = od; }
    if (h == 0 && pv) {
      const double denom = fmax(bd, FLOOR64);
      const double beta = ba / denom;
      const double rss = fmax(rwr - ba * ba / denom, FLOOR64);
      pmax[mt * pmax_ld + p] = beta * beta / (rss / nu / denom);
    }
  }
}

template <int PREC>
hipError_t launch_wald_scan_wide(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P, int64_t n_real,
                                 const double* rWr, const double* h2, double nu, const int64_t* row_index,
                                 const ScanOut& out, double* pmax, hipStream_t stream) {
  const int64_t n_mtiles = (Mc + W_TM - 1) / W_TM;
  const int64_t n_ptiles = (P + W_TP - 1) / W_TP;
  const int64_t n_tiles = n_mtiles * n_ptiles;
  constexpr int lds = 2 * 2 * WideTraits<PREC>::NSUB * W_TM * WideTraits<PREC>::ROWB;
  hipLaunchKernelGGL(wald_scan_wide_kernel<PREC>, dim3((unsigned)n_tiles), dim3(256), lds, stream, args, sc, n_mtiles,
                     n_tiles, Mc, P, n_real, rWr, h2, nu, row_index, out, pmax, args.b_rows_pad);
  return hipGetLastError();
}

template hipError_t launch_wald_scan_wide<PREC_SPLIT_FP16>(const TileArgs<PREC_SPLIT_FP16, 2>&, const ScanScales&,
                                                            int64_t, int64_t, int64_t, const double*, const double*,
                                                            double, const int64_t*, const ScanOut&, double*,
                                                            hipStream_t);
template hipError_t launch_wald_scan_wide<PREC_SPLIT_BF16>(const TileArgs<PREC_SPLIT_BF16, 2>&, const ScanScales&, int64_t, int64_t,
                                                            int64_t, const double*, const double*, double,
                                                            const int64_t*, const ScanOut&, double*, hipStream_t);
template hipError_t launch_wald_scan_wide<PREC_FP32>(const TileArgs<PREC_FP32, 2>&, const ScanScales&, int64_t, int64_t, int64_t,
                                                      const double*, const double*, double, const int64_t*,
                                                      const ScanOut&, double*, hipStream_t);

}  // namespace flmm
