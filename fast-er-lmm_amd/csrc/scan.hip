// Fused Wald scan for the projected (covariate-free) basis on gfx950.
//
// For one chromosome's SNP block (rows m of the A side) against every scan column (rows p of the B side):
//     sWy[m][p] = sum_k US[m][k] * (w_p o UY_p)[k]        (GEMM 1: planes S1 = US,    Q1 = w o UY)
//     sWs[m][p] = sum_k US[m][k]^2 * w_p[k]               (GEMM 2: planes S2 = US^2,  Q2 = w)
// with w_p[k] = 1 / (s_k + delta_p), followed in registers by the Wald epilogue in FP64
// (reference fasterlmm/core.py:649-667):
//     denom = max(sWs, floor); beta = sWy / denom; rss = max(rWr - sWy^2 / denom, floor)
//     var = rss / nu / denom; F = beta^2 / var; se = sqrt(var); sfve = sqrt(max(sWy^2 / (denom rWr), 0))
// Real columns (p < n_real) store F, beta, se, sfve and nullh2 at row row_index[m] of the (M, n_real)
// outputs; every column contributes its tile-wise maximum F to pmax[m_tile][p] (core.py:657), reduced over
// tiles and chromosomes by flmm_max_reduce.  F is a non-decreasing function of t = sWy^2 / denom, so the
// maximum is tracked on t by FP64 cross-multiplication (no division per test) and F is formed once per column.
#include "kernels.h"

namespace flmm {


__device__ inline double wald_f(double a, double d, double rwr, double nu) {
  const double denom = fmax(d, FLOOR64);
  const double beta = a / denom;
  const double rss = fmax(rwr - a * a / denom, FLOOR64);
  const double var = rss / nu / denom;
  return beta * beta / var;
}

template <int PREC>
__global__ __launch_bounds__(NTHREADS, 1) void wald_scan_kernel(
    TileArgs<PREC, 2> args, ScanScales sc, int64_t n_mtiles, int64_t n_tiles, int64_t Mc, int64_t P, int64_t n_real,
    const double* __restrict__ rWr, const double* __restrict__ h2, double nu, const int64_t* __restrict__ row_index,
    ScanOut out, double* __restrict__ pmax, int64_t pmax_ld) {
  extern __shared__ char lds[];
  const int64_t tile = xcd_remap(blockIdx.x, n_tiles);
  const int64_t mt = tile % n_mtiles;
  const int64_t pt = tile / n_mtiles;
  const int64_t m0 = mt * TILE, p0 = pt * TILE;

  TileAcc<PREC, 2> acc;
  tile_mainloop<PREC, 2>(args, m0, p0, acc, lds);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  __syncthreads();  // LDS is reused below for the cross-wave max exchange
  double* xa = (double*)lds;        // [2 (bj)][64 cols] best sWy of the wm = 1 waves
  double* xd = xa + 2 * 64 * 2;     // best sWs

#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int cl = acc_col(wn, bj, lane);
    const int64_t p = p0 + cl;
    const bool pv = p < P;
    const double rwr = pv ? rWr[p] : 1.0;
    const bool real = p < n_real;
    const double hh = real ? h2[p] : 0.0;
    const double inv_q1 = (sc.q1 && pv) ? 1.0 / (double)sc.q1[p] : 1.0;
    const double inv_q2 = (sc.q2 && pv) ? 1.0 / (double)sc.q2[p] : 1.0;
    double ba = 0.0, bd = 1.0;  // best (sWy, sWs) so far: t = ba^2 / bd
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t m = m0 + acc_row(wm, bi, lane, v);
        if (m < Mc) {
          double a = (double)acc.c[0][bi][bj][v];
          double d = (double)acc.c[1][bi][bj][v];
          unscale_sums(sc, m, inv_q1, inv_q2, a, d);
          d = fmax(d, FLOOR64);
          // t = a^2/d > ba^2/bd  <=>  a^2 bd > ba^2 d   (d, bd > 0)
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
    // fold the other lane half (same column, other rows)
    {
      const double oa = __shfl_xor(ba, 32, 64), od = __shfl_xor(bd, 32, 64);
      if (oa * oa * bd > ba * ba * od) { ba = oa; bd = od; }
    }
    if (wm == 1 && lane < 32) {
      xa[(wn * 2 + bj) * 32 + lane] = ba;
      xd[(wn * 2 + bj) * 32 + lane] = bd;
    }
    __syncthreads();
    if (wm == 0 && lane < 32) {
      const double oa = xa[(wn * 2 + bj) * 32 + lane], od = xd[(wn * 2 + bj) * 32 + lane];
      if (oa * oa * bd > ba * ba * od) { ba = oa; bd = od; }
      if (pv) pmax[mt * pmax_ld + p] = wald_f(ba, bd, rwr, nu);
    }
    __syncthreads();
  }
}

// max_F[p] = max(max_F[p], max over tiles of pmax[t][p])
__global__ void max_reduce_kernel(const double* __restrict__ pmax, int64_t n_mtiles, int64_t ld, int64_t P,
                                  double* __restrict__ max_F) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double m = max_F[p];
  for (int64_t t = 0; t < n_mtiles; ++t) m = fmax(m, pmax[t * ld + p]);
  max_F[p] = m;
}

template <int PREC>
hipError_t launch_wald_scan(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P, int64_t n_real,
                            const double* rWr, const double* h2, double nu, const int64_t* row_index,
                            const ScanOut& out, double* pmax, hipStream_t stream) {
  const int64_t n_mtiles = (Mc + TILE - 1) / TILE;
  const int64_t n_ptiles = (P + TILE - 1) / TILE;
  const int64_t n_tiles = n_mtiles * n_ptiles;
  const int lds = tile_lds_bytes<PREC, 2>();
  hipLaunchKernelGGL(wald_scan_kernel<PREC>, dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args, sc, n_mtiles,
                     n_tiles, Mc, P, n_real, rWr, h2, nu, row_index, out, pmax, args.b_rows_pad);
  return hipGetLastError();
}

hipError_t launch_max_reduce(const double* pmax, int64_t n_mtiles, int64_t ld, int64_t P, double* max_F,
                             hipStream_t stream) {
  hipLaunchKernelGGL(max_reduce_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream, pmax, n_mtiles,
                     ld, P, max_F);
  return hipGetLastError();
}

template hipError_t launch_wald_scan<PREC_SPLIT_FP16>(const TileArgs<PREC_SPLIT_FP16, 2>&, const ScanScales&, int64_t,
                                                       int64_t, int64_t, const double*, const double*, double,
                                                       const int64_t*, const ScanOut&, double*, hipStream_t);
template hipError_t launch_wald_scan<PREC_SPLIT_BF16>(const TileArgs<PREC_SPLIT_BF16, 2>&, const ScanScales&, int64_t, int64_t,
                                                       int64_t, const double*, const double*, double,
                                                       const int64_t*, const ScanOut&, double*, hipStream_t);
template hipError_t launch_wald_scan<PREC_FP32>(const TileArgs<PREC_FP32, 2>&, const ScanScales&, int64_t, int64_t, int64_t,
                                                 const double*, const double*, double, const int64_t*,
                                                 const ScanOut&, double*, hipStream_t);

}  // namespace flmm
