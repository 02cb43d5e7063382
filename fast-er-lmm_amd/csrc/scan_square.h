// Fused Wald scan tile for the projected (covariate-free) basis on gfx950.
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
#pragma once
#include "kernels.h"

namespace flmm {



__device__ inline double wald_f(double a, double d, double rwr, double nu) {
  const double denom = fmax(d, FLOOR64);
  const double beta = a / denom;
  const double rss = fmax(rwr - a * a / denom, FLOOR64);
  const double var = rss / nu / denom;
  return beta * beta / var;
}

// One 128 x 128 tile (logical index `tile`, SNP tile fastest) of the square Wald scan, run by a 256-thread
// workgroup with tile_lds_bytes<PREC, 2>() of dynamic LDS.
template <int PREC>
__device__ inline void square_tile(
    const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t n_mtiles, int64_t tile, int64_t Mc, int64_t P,
    int64_t n_real,
    const double* __restrict__ rWr, const double* __restrict__ h2, double nu, const int64_t* __restrict__ row_index,
    ScanOut out, double* __restrict__ pmax, int64_t pmax_ld) {
  extern __shared__ char lds[];
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
    const double rwr = pv ? gmem(rWr)[p] : 1.0;
    const bool real = p < n_real;
    const double hh = real ? gmem(h2)[p] : 0.0;
    const double inv_q1 = (sc.q1 && pv) ? 1.0 / (double)gmem(sc.q1)[p] : 1.0;
    const double inv_q2 = (sc.q2 && pv) ? 1.0 / (double)gmem(sc.q2)[p] : 1.0;
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
            const int64_t o = gmem(row_index)[m] * out.ld + p;
            gmem(out.f)[o] = beta * beta / var;
            gmem(out.beta)[o] = beta;
            gmem(out.se)[o] = sqrt(var);
            gmem(out.sfve)[o] = sqrt(fmax(a * a / (d * rwr), 0.0));
            gmem(out.nullh2)[o] = hh;
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
      if (pv) gmem(pmax)[mt * pmax_ld + p] = wald_f(ba, bd, rwr, nu);
    }
    __syncthreads();
  }
}

}  // namespace flmm
