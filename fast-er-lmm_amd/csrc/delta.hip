// Per-column log-delta search for the projected-basis LMM (reference fasterlmm/core.py:525-595).
//
// Loss (-2 log-likelihood on the covariate-free projected basis, core.py:525-542):
//     L(t) = n (log 2 pi + log(max(R(t), floor) / n) + 1) + Lambda(t),   delta = e^t
//     R(t) = sum_k y_k^2 / (s_k + delta),   Lambda(t) = sum_j log(s_lam_j + delta)
// The compat path has n = Neff and s_lam = s.  The GLS path's ML loss (core.py:148-187) is the same function with
// n = N and s_lam = the full spectrum of K; its REML loss is the compat form up to a constant.
//
// Stage 1, grid: R at every grid point for every column is one GEMM, W_grid (G x K) times (UY o UY)^T, run on
// FP32 MFMA (the squaring is fused into the LDS store of the tile engine).  Lambda at the grid points is
// column independent and computed once.  The epilogue takes the per-column argmin (first index on ties, as
// torch.argmin does, core.py:571).
//
// Stage 2, refinement, one wave per column, FP64.  Every search mode first walks from the FP32 grid argmin to the
// FP64 grid minimum (fp64_grid_argmin), so it brackets what the reference's FP64 argmin brackets.
//   MODE_GOLDEN: the reference algorithm step for step -- bracket [g[i-1], g[i+1]], 50 golden-section
//                iterations, (a+b)/2 clamped (core.py:577-595).
//   MODE_FIXED:  no search; the caller's log delta is used as given.
//   MODE_GRID:   no refinement: the first minimiser of the FP64 loss over the grid nodes (refine=False).
//   MODE_LEARNED: a proposal t0 per column (from the MLP) finished by Newton within t0 +- radius; accepted when it
//                converges there and its loss is within judge_tol of the grid optimum, otherwise (flags[p] = 1)
//                the column falls back to MODE_NEWTON's grid bracket.
//   MODE_NEWTON: the same local minimiser in the same bracket, found by safeguarded Newton on
//                L'(t) = delta (W1 - n R2 / R), W1 = sum 1/(s_lam+delta), R2 = sum y^2 w^2, w = 1/(s+delta):
//                no logarithms per iteration and quadratic convergence; bisection keeps the iterate inside
//                the bracket.  A bracket whose left slope is >= 0 (right slope <= 0) has its minimum at that
//                edge, which is where golden section converges too.  Where the FP64 grid minimum is interior
//                and its two neighbours are higher, FP64 Newton starts at the vertex of the parabola through the
//                three node losses (within half a step of the minimiser, so a few FP64 steps); otherwise an FP32
//                Newton phase from the node comes first.  A grid minimum at an end node with the loss still
//                falling towards that end (h2 ~ 0 puts most permutation columns at the upper end) is settled
//                by one sweep that also yields the column's outputs.
// The refinement then writes the scan's column-side operand planes Q1 = w o y and Q2 = w for the final delta
// (split-FP16: each column scaled by the power of two of its maximum, recorded in q1s / q2s; Q2 skipped for
// factored weights, whose planes factored_planes_kernel writes),
// the null residual rWr = max(sum w y^2, floor) (core.py:641-643) and nullh2 = 1 / (1 + delta).
#include "kernels.h"

namespace flmm {

enum DeltaMode : int { MODE_GOLDEN = 0, MODE_NEWTON = 1, MODE_FIXED = 2, MODE_LEARNED = 3, MODE_GRID = 4 };

__device__ inline double grid_value(int i, int G, double lo, double hi) {
  if (G <= 1) return lo;
  if (i == G - 1) return hi;
  return (double)i * ((hi - lo) / (double)(G - 1)) + lo;
}

// One block per grid row g (128 blocks): W plane row g (fp32, zero beyond Neff and for g >= G) and
// lam[g] = sum_k log(s_lam_k + delta_g) over the log-det spectrum.
__global__ void delta_weights_kernel(const double* __restrict__ s, int64_t neff, const double* __restrict__ s_lam,
                                     int64_t nlam, int64_t kpad, int G, double lo, double hi,
                                     float* __restrict__ wplane, float* __restrict__ lam,
                                     double* __restrict__ lam64) {
  __shared__ double red[256];
  const int g = blockIdx.x;
  const double delta = exp(grid_value(g, G, lo, hi));
  double acc = 0.0;
  for (int64_t k = threadIdx.x; k < kpad; k += blockDim.x) {
    float w = 0.f;
    if (g < G && k < neff) w = (float)(1.0 / (s[k] + delta));
    wplane[((k >> 5) * TILE + g) * 32 + (k & 31)] = w;
  }
  if (g < G)
    for (int64_t k = threadIdx.x; k < nlam; k += blockDim.x) acc += log(s_lam[k] + delta);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    lam[g] = (float)red[0];
    if (lam64) lam64[g] = red[0];
  }
}

// argmin over the grid rows of one 128-column tile
__global__ __launch_bounds__(NTHREADS, 1) void delta_grid_kernel(TileArgs<PREC_FP32, 1> args, int64_t n_tiles,
                                                                  const float* __restrict__ lam, int G, float nco,
                                                                  int64_t P, int* __restrict__ idx_out) {
  extern __shared__ char lds[];
  const int64_t pt = xcd_remap(blockIdx.x, n_tiles);
  const int64_t p0 = pt * TILE;
  TileAcc<PREC_FP32, 1> acc;
  tile_mainloop<PREC_FP32, 1, true, false>(args, 0, p0, acc, lds);

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  __syncthreads();
  float* xl = (float*)lds;
  int* xi = (int*)(xl + 128);
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    float best = INFINITY;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int g = acc_row(wm, bi, lane, v);
        if (g < G) {
          const float loss = nco * __logf(fmaxf(acc.c[0][bi][bj][v], 1e-30f)) + lam[g];
          if (loss < best || (loss == best && g < bidx)) { best = loss; bidx = g; }
        }
      }
    {
      const float ob = __shfl_xor(best, 32, 64);
      const int oi = __shfl_xor(bidx, 32, 64);
      if (ob < best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
    }
    if (wm == 1 && lane < 32) { xl[(wn * 2 + bj) * 32 + lane] = best; xi[(wn * 2 + bj) * 32 + lane] = bidx; }
    __syncthreads();
    if (wm == 0 && lane < 32) {
      const float ob = xl[(wn * 2 + bj) * 32 + lane];
      const int oi = xi[(wn * 2 + bj) * 32 + lane];
      if (ob < best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
      const int64_t p = p0 + acc_col(wn, bj, lane);
      if (p < P) idx_out[p] = (bidx == 0x7fffffff) ? 0 : bidx;
    }
    __syncthreads();
  }
}

// The same grid argmin on split-FP16 MFMA.  Both operands arrive as fp32 planes and are scaled and split into fp16
// (hi, lo) planes as each 32-k chunk is stored into LDS: grid row g (weights w_gk = 1 / (s_k + delta_g) <=
// 1 / (s_min + delta_g)) by the power of two sa_g of that bound, column p (y_k^2 <= |y_p|^2) by
// sb_p = col_scale[p]^2 / 2^14, col_scale[p] a power of two with col_scale[p] |y_p| <= 2^14 (the rotated columns'
// own plane scale: the basis is orthonormal).  Three fp16 products per step (hi hi, hi lo, lo hi) with FP32
// accumulation give the weighted sums to ~2^-21 relative, as FP32 products with FP32 accumulation do; sb_p adds
// the same constant to every grid loss of column p (no change to its argmin) and sa_g is taken out through lam.
// 128 grid rows x 128 columns per workgroup, 4 waves of 64 x 64, chunks double-buffered in LDS.
// The grid weights' scaled split (hi, lo) planes, once per call (PRE): element (kc, r, kk) of the FP32 W plane at
// ((kc 128 + r) 32 + kk) -> the same index of the hi plane and of the lo plane (128 kpad values after it).
__global__ __launch_bounds__(256) void delta_wsplit_kernel(const float* __restrict__ wplane, double s_min, int G,
                                                           double lo, double hi, int64_t kpad, u16* __restrict__ wsp) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= TILE * kpad) return;
  const int r = (int)((e >> 5) % TILE);
  const float sc = (float)pow2_scale(1.0 / (s_min + exp(grid_value(r, G, lo, hi))));
  u16 h, l;
  split2f<PREC_SPLIT_FP16>(wplane[e] * sc, h, l);
  wsp[e] = h;
  wsp[TILE * kpad + e] = l;
}

template <bool PRE>
__global__ __launch_bounds__(NTHREADS, 2) void delta_grid_split_kernel(
    const float* __restrict__ wplane, const u16* __restrict__ wsp, const float* __restrict__ uy,
    const float* __restrict__ col_scale, double s_min, int G, double lo, double hi, const float* __restrict__ lam,
    float nco, int64_t P, int64_t p_pad, int64_t kpad, int64_t n_tiles, int* __restrict__ idx_out) {
  constexpr int PLANE = TILE * 64;  // one fp16 plane of a chunk: 128 rows x 32 values
  constexpr int STAGE = 4 * PLANE;  // A hi, A lo, B hi, B lo
  __shared__ __attribute__((aligned(16))) char lds[2 * STAGE];
  __shared__ float lam_s[TILE];
  const int64_t pt = xcd_remap(blockIdx.x, n_tiles);
  const int64_t p0 = pt * TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1, h = lane >> 5;
  const int nch = (int)(kpad / KCHUNK);
  // staging: thread t covers (row r = t / 4 + 64 ps, 8-value segment sg = t % 4), ps = 0, 1, of each side
  const int sg = tid & 3;
  float sc[2][2];  // [side][ps]
#pragma unroll
  for (int ps = 0; ps < 2; ++ps) {
    const int r = (tid >> 2) + 64 * ps;
    sc[0][ps] = (float)pow2_scale(1.0 / (s_min + exp(grid_value(r, G, lo, hi))));
    const int64_t p = p0 + r;
    const float cs = p < P ? gmem(col_scale)[p] : 0.f;
    sc[1][ps] = cs * cs * (1.0f / 16384.0f);
  }
  if (tid < TILE) {
    const double sa = pow2_scale(1.0 / (s_min + exp(grid_value(tid, G, lo, hi))));
    lam_s[tid] = tid < G ? (float)((double)gmem(lam)[tid] - (double)nco * log(sa)) : 0.f;
  }
  auto gload = [&](int kc, f32x4 (&f)[2][2][2]) {
#pragma unroll
    for (int ps = 0; ps < 2; ++ps) {
      const int r = (tid >> 2) + 64 * ps;
      const float* b = uy + ((int64_t)kc * p_pad + p0 + r) * 32 + sg * 8;
      if constexpr (PRE) {  // the split planes: 8 hi values, then the same 8 lo values
        const u16* a = wsp + ((int64_t)kc * TILE + r) * 32 + sg * 8;
        f[0][ps][0] = __builtin_bit_cast(f32x4, *gmem((const u32x4*)a));
        f[0][ps][1] = __builtin_bit_cast(f32x4, *gmem((const u32x4*)(a + TILE * kpad)));
      } else {
        const float* a = wplane + ((int64_t)kc * TILE + r) * 32 + sg * 8;
        f[0][ps][0] = *gmem((const f32x4*)a);
        f[0][ps][1] = *gmem((const f32x4*)(a + 4));
      }
      f[1][ps][0] = *gmem((const f32x4*)b);
      f[1][ps][1] = *gmem((const f32x4*)(b + 4));
    }
  };
  auto lstore = [&](const f32x4 (&f)[2][2][2], char* st) {
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int ps = 0; ps < 2; ++ps) {
        const int r = (tid >> 2) + 64 * ps;
        if (PRE && side == 0) {  // already scaled and split
          const int off = r * 64 + ((sg ^ ((r >> 2) & 3)) * 16);
          *(u32x4*)(st + off) = __builtin_bit_cast(u32x4, f[0][ps][0]);
          *(u32x4*)(st + PLANE + off) = __builtin_bit_cast(u32x4, f[0][ps][1]);
          continue;
        }
        u16 hv[8], lv[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float v = f[side][ps][e >> 2][e & 3];
          split2f<PREC_SPLIT_FP16>(side ? v * v * sc[1][ps] : v * sc[0][ps], hv[e], lv[e]);
        }
        const int off = r * 64 + ((sg ^ ((r >> 2) & 3)) * 16);
        u32x4 ph, pl;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ph[q] = hv[2 * q] | ((uint32_t)hv[2 * q + 1] << 16);
          pl[q] = lv[2 * q] | ((uint32_t)lv[2 * q + 1] << 16);
        }
        *(u32x4*)(st + (side * 2) * PLANE + off) = ph;
        *(u32x4*)(st + (side * 2 + 1) * PLANE + off) = pl;
      }
  };

  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  f32x4 pre[2][2][2];
  gload(0, pre);
  lstore(pre, lds);
  __syncthreads();
  for (int kc = 0; kc < nch; ++kc) {
    const char* cur = lds + (kc & 1) * STAGE;
    char* nxt = lds + ((kc & 1) ^ 1) * STAGE;
    const bool more = kc + 1 < nch;
    if (more) gload(kc + 1, pre);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int seg = h * 2 + ks;
      bf16x8 af[2][2], bfr[2][2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) {
        const int r = wm * 64 + bi * 32 + (lane & 31);
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
          af[bi][sub] = *(const bf16x8*)(cur + sub * PLANE + r * 64 + ((seg ^ ((r >> 2) & 3)) * 16));
      }
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int r = wn * 64 + bj * 32 + (lane & 31);
#pragma unroll
        for (int sub = 0; sub < 2; ++sub)
          bfr[bj][sub] = *(const bf16x8*)(cur + (2 + sub) * PLANE + r * 64 + ((seg ^ ((r >> 2) & 3)) * 16));
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          f32x16 c = acc[bi][bj];
          c = mfma_split<PREC_SPLIT_FP16>(af[bi][0], bfr[bj][0], c);
          c = mfma_split<PREC_SPLIT_FP16>(af[bi][0], bfr[bj][1], c);
          c = mfma_split<PREC_SPLIT_FP16>(af[bi][1], bfr[bj][0], c);
          acc[bi][bj] = c;
        }
    }
    if (more) lstore(pre, nxt);  // nxt was last read in the previous chunk, which ended with a barrier
    __syncthreads();
  }

  // argmin over the grid rows, as delta_grid_kernel (first index of the minimum)
  float* xl = (float*)lds;
  int* xi = (int*)(xl + 128);
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    float best = INFINITY;
    int bidx = 0x7fffffff;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int g = acc_row(wm, bi, lane, v);
        if (g < G) {
          const float loss = nco * __logf(fmaxf(acc[bi][bj][v], 1e-30f)) + lam_s[g];
          if (loss < best || (loss == best && g < bidx)) { best = loss; bidx = g; }
        }
      }
    {
      const float ob = __shfl_xor(best, 32, 64);
      const int oi = __shfl_xor(bidx, 32, 64);
      if (ob < best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
    }
    if (wm == 1 && lane < 32) { xl[(wn * 2 + bj) * 32 + lane] = best; xi[(wn * 2 + bj) * 32 + lane] = bidx; }
    __syncthreads();
    if (wm == 0 && lane < 32) {
      const float ob = xl[(wn * 2 + bj) * 32 + lane];
      const int oi = xi[(wn * 2 + bj) * 32 + lane];
      if (ob < best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
      const int64_t p = p0 + acc_col(wn, bj, lane);
      if (p < P) idx_out[p] = (bidx == 0x7fffffff) ? 0 : bidx;
    }
    __syncthreads();
  }
}

struct ColSums {
  double R, R2, R3, W1, W2, LS;
};

// The likelihood of one column: y (neff values in the weight spectrum s) and the log-det spectrum s_lam (nlam
// values; the same as s on the projected basis, the full spectrum for the ML loss of the GLS path).  With
// KPL > 0 each lane keeps its KPL elements (k = lane + 64 i) of y in registers (as the FP32 values the rotation
// produced), and the block holds both spectra once in LDS, so the column is read from HBM once however many
// loss / slope evaluations follow; KPL = 0 re-reads everything from global memory every evaluation.
// When s_lam is s (SAME, the projected compat basis), the log-det weights are the data weights: each element
// costs one reciprocal.
template <int KPL, bool SAME>
struct ColLik {
  const float* uy;
  int64_t p_pad;
  const double* s;
  int64_t neff;
  const double* s_lam;
  int64_t nlam;
  double nco;          // coefficient of log R in the loss
  const double* ls;    // LDS copies (KPL > 0): s, s_lam (padded with 1), and their FP32 roundings
  const double* ll;
  const float* lsf;
  const float* llf;
  float yr[KPL > 0 ? KPL : 1];

  __device__ void load(int64_t p, int lane) {
    if constexpr (KPL > 0) {
#pragma unroll
      for (int i = 0; i < KPL; ++i) {
        const int64_t k = lane + 64 * i;
        yr[i] = k < neff ? uy[((k >> 5) * p_pad + p) * 32 + (k & 31)] : 0.f;
      }
    }
  }

  // sums over k at delta; with_log adds Lambda = sum log(s_lam + delta)
  __device__ ColSums sums(int64_t p, double delta, bool with_log, int lane) const {
    ColSums c{0, 0, 0, 0, 0, 0};
    if constexpr (KPL > 0) {
#pragma unroll
      for (int i = 0; i < KPL; ++i) {
        const int k = lane + 64 * i;
        const double y = (double)yr[i];
        const double sk = ls[k];
        const double w = rcp_pos(sk + delta);
        const double yw = y * y * w;
        c.R += yw;
        c.R2 += yw * w;
        c.R3 += yw * w * w;
        if (k < nlam) {
          const double sd = (SAME ? sk : ll[k]) + delta;
          const double wl = SAME ? w : rcp_pos(sd);
          c.W1 += wl;
          c.W2 += wl * wl;
          if (with_log) c.LS += log(sd);
        }
      }
    } else {
      for (int64_t k = lane; k < neff; k += 64) {
        const double y = (double)uy[((k >> 5) * p_pad + p) * 32 + (k & 31)];
        const double w = rcp_pos(s[k] + delta);
        const double yw = y * y * w;
        c.R += yw;
        c.R2 += yw * w;
        c.R3 += yw * w * w;
        if constexpr (SAME) {
          c.W1 += w;
          c.W2 += w * w;
          if (with_log) c.LS += log(s[k] + delta);
        }
      }
      if constexpr (!SAME) for (int64_t k = lane; k < nlam; k += 64) {
        const double sd = s_lam[k] + delta;
        const double w = rcp_pos(sd);
        c.W1 += w;
        c.W2 += w * w;
        if (with_log) c.LS += log(sd);
      }
    }
    c.R = wave_sum(c.R);
    c.R2 = wave_sum(c.R2);
    c.R3 = wave_sum(c.R3);
    c.W1 = wave_sum(c.W1);
    c.W2 = wave_sum(c.W2);
    if (with_log) c.LS = wave_sum(c.LS);
    return c;
  }

  // R = sum_k y_k^2 / (s_k + delta) alone
  __device__ double R_only(int64_t p, double delta, int lane) const {
    double r = 0.0;
    if constexpr (KPL > 0) {
#pragma unroll
      for (int i = 0; i < KPL; ++i) {
        const double y = (double)yr[i];
        r += y * y * rcp_pos(ls[lane + 64 * i] + delta);
      }
    } else {
      for (int64_t k = lane; k < neff; k += 64) {
        const double y = (double)uy[((k >> 5) * p_pad + p) * 32 + (k & 31)];
        r += y * y * rcp_pos(s[k] + delta);
      }
    }
    return wave_sum(r);
  }

  // R at three deltas in one sweep (register path): each sum in the same order as R_only, so the values are
  // bit-identical to three R_only calls
  __device__ void R_three(const double (&delta)[3], double (&r)[3], int lane) const {
    double a[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < (KPL > 0 ? KPL : 1); ++i) {
      const double y = (double)yr[i];
      const double sk = ls[lane + 64 * i];
#pragma unroll
      for (int j = 0; j < 3; ++j) a[j] += y * y * rcp_pos(sk + delta[j]);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) r[j] = wave_sum(a[j]);
  }

  // FP32 slope L'(t) and curvature L''(t) (register path only): a cheap first approximation of the root
  __device__ void slope32(float t, int lane, float& g, float& gp) const {
    const float delta = __expf(t);
    float R = 0.f, R2 = 0.f, R3 = 0.f, W1 = 0.f, W2 = 0.f;
#pragma unroll
    for (int i = 0; i < (KPL > 0 ? KPL : 1); ++i) {
      const int k = lane + 64 * i;
      const float w = __builtin_amdgcn_rcpf(lsf[k] + delta);
      const float yw = yr[i] * yr[i] * w;
      R += yw;
      R2 += yw * w;
      R3 += yw * w * w;
      if (k < nlam) {
        const float wl = SAME ? w : __builtin_amdgcn_rcpf(llf[k] + delta);
        W1 += wl;
        W2 += wl * wl;
      }
    }
    R = wave_sum(R);
    R2 = wave_sum(R2);
    R3 = wave_sum(R3);
    W1 = wave_sum(W1);
    W2 = wave_sum(W2);
    const float n = (float)nco;
    const float h = W1 - n * R2 / R;
    const float hp = -W2 + n * (2.f * R3 * R - R2 * R2) / (R * R);
    g = delta * h;
    gp = delta * h + delta * delta * hp;
  }
};

constexpr double LOG2PI = 1.8378770664093453;

template <class LK>
__device__ inline double col_loss(const LK& L, int64_t p, double t, int lane) {
  const ColSums c = L.sums(p, exp(t), true, lane);
  return L.nco * (LOG2PI + log(fmax(c.R, FLOOR64) / L.nco) + 1.0) + c.LS;
}

// slope L'(t) and curvature L''(t)
template <class LK>
__device__ inline void col_slope(const LK& L, int64_t p, double t, int lane, double& g, double& gp) {
  const double delta = exp(t);
  const ColSums c = L.sums(p, delta, false, lane);
  const double n = L.nco;
  if (c.R <= FLOOR64) {  // all-zero column: R is floored, only Lambda moves
    g = delta * c.W1;
    gp = g - delta * delta * c.W2;
    return;
  }
  const double h = c.W1 - n * c.R2 / c.R;
  const double hp = -c.W2 + n * (2.0 * c.R3 * c.R - c.R2 * c.R2) / (c.R * c.R);
  g = delta * h;
  gp = delta * h + delta * delta * hp;
}

// The grid argmin handed in comes from the FP32 grid kernel; where neighbouring grid losses differ by less than
// FP32 resolves (flat likelihoods, h2 near 0) it can sit one or many nodes away from the FP64 argmin the reference
// brackets (core.py:570-575) -- on a profile that falls monotonically to the grid's edge by less than FP32 resolves,
// the FP32 argmin is the first of a run of equal values.  Walk from it to the FP64 local minimum of the grid: left
// while the left neighbour is not higher (ties go to the first index, as argmin's do), right while the right
// neighbour is lower; one FP64 loss per step, and only flat columns take more than one or two.
// lam64: the column-independent log-det term at every grid node in FP64 (from the weights kernel), so a step costs
// one R sum and one log; NULL: each node's log-det is summed per column.
template <class LK>
__device__ inline double node_loss(const LK& L, int64_t p, int g, int G, double lo, double hi,
                                   const double* __restrict__ lam64, int lane) {
  const double t = grid_value(g, G, lo, hi);
  if (!lam64) return col_loss(L, p, t, lane);
  return L.nco * (LOG2PI + log(fmax(L.R_only(p, exp(t), lane), FLOOR64) / L.nco) + 1.0) + lam64[g];
}

template <class LK>
__device__ inline double node_loss_from_R(const LK& L, double R, int g, const double* __restrict__ lam64) {
  return L.nco * (LOG2PI + log(fmax(R, FLOOR64) / L.nco) + 1.0) + lam64[g];
}

template <class LK, bool FUSED>
__device__ inline int fp64_grid_argmin(const LK& L, int64_t p, int i, int G, double lo, double hi,
                                       const double* __restrict__ lam64, int lane, double& t_vertex) {
  t_vertex = NAN;
  if constexpr (FUSED) {
    // the usual outcome -- the node is the minimum of its two neighbours -- costs one sweep for all three losses;
    // a move continues node by node below, exactly as the unfused walk
    if (lam64 && i > 0 && i + 1 < G) {
      const double d[3] = {exp(grid_value(i - 1, G, lo, hi)), exp(grid_value(i, G, lo, hi)),
                           exp(grid_value(i + 1, G, lo, hi))};
      double r[3];
      L.R_three(d, r, lane);
      const double lm = node_loss_from_R(L, r[0], i - 1, lam64), l0 = node_loss_from_R(L, r[1], i, lam64),
                   lp = node_loss_from_R(L, r[2], i + 1, lam64);
      if (lm <= l0) {
        --i;
        double li = lm;
        for (int it = 0; it < G && i > 0; ++it) {
          const double l = node_loss(L, p, i - 1, G, lo, hi, lam64, lane);
          if (!(l <= li)) break;
          --i;
          li = l;
        }
        return i;
      }
      if (!(lp < l0)) {
        // the node is the minimum of its neighbours (lm > l0 <= lp): the vertex of the parabola through the three
        // FP64 node losses lies within half a grid step of it and starts the search there
        const double h = grid_value(i + 1, G, lo, hi) - grid_value(i, G, lo, hi);
        const double den = lm - 2.0 * l0 + lp;
        if (den > 0.0) t_vertex = grid_value(i, G, lo, hi) + 0.5 * h * (lm - lp) / den;
        return i;
      }
      ++i;
      double li = lp;
      for (int it = 0; it < G && i + 1 < G; ++it) {
        const double l = node_loss(L, p, i + 1, G, lo, hi, lam64, lane);
        if (!(l < li)) break;
        ++i;
        li = l;
      }
      return i;
    }
  }
  double li = node_loss(L, p, i, G, lo, hi, lam64, lane);
  bool moved_left = false;
  for (int it = 0; it < G && i > 0; ++it) {
    const double l = node_loss(L, p, i - 1, G, lo, hi, lam64, lane);
    if (!(l <= li)) break;
    --i;
    li = l;
    moved_left = true;
  }
  if (!moved_left) {
    for (int it = 0; it < G && i + 1 < G; ++it) {
      const double l = node_loss(L, p, i + 1, G, lo, hi, lam64, lane);
      if (!(l < li)) break;
      ++i;
      li = l;
    }
  }
  return i;
}

template <class LK>
__device__ inline double golden_bracket(const LK& L, int64_t p, int i, int G, double lo, double hi, int lane) {
  const double inv_phi = 2.0 / (1.0 + sqrt(5.0));
  const double inv_phi2 = inv_phi * inv_phi;
  double a = grid_value(i > 0 ? i - 1 : 0, G, lo, hi);
  double b = grid_value(i + 1 < G ? i + 1 : G - 1, G, lo, hi);
  double c = a + inv_phi2 * (b - a), d = a + inv_phi * (b - a);
  double fc = col_loss(L, p, c, lane), fd = col_loss(L, p, d, lane);
  for (int it = 0; it < 50; ++it) {
    if (fc < fd) b = d; else a = c;
    c = a + inv_phi2 * (b - a);
    d = a + inv_phi * (b - a);
    fc = col_loss(L, p, c, lane);
    fd = col_loss(L, p, d, lane);
  }
  return 0.5 * (a + b);
}

// safeguarded Newton on L'(t) in the grid bracket [g[i-1], g[i+1]], started at the grid optimum g[i].  The slopes
// at the bracket ends are only needed for the edge cases (loss rising at the left end -> the left end; falling at
// the right end -> the right end), so they are evaluated the first time an iterate tries to leave the bracket;
// an interior minimum that Newton reaches directly costs no end evaluations.
//
// On the register path an FP32 Newton phase runs first (cheap passes, root to ~1e-5); its result, when it stays
// inside the bracket and converges, is only the starting point of the FP64 iteration, which decides the answer.
// the grid bracket [g[i-1], g[i+1]] (clamped at the grid's ends) and the search's starting point g[i] (the
// bracket's midpoint where g[i] is one of its ends)
__device__ inline void grid_bracket(int i, int G, double lo, double hi, double& a0, double& b0, double& t) {
  a0 = grid_value(i > 0 ? i - 1 : 0, G, lo, hi);
  b0 = grid_value(i + 1 < G ? i + 1 : G - 1, G, lo, hi);
  t = grid_value(i, G, lo, hi);
  if (!(t > a0 && t < b0)) t = 0.5 * (a0 + b0);
}

// FP32 Newton phase (register path) from t inside (a0, b0): on convergence t becomes the FP32 root, else it stays
template <class LK>
__device__ inline void newton_fp32(const LK& L, double a0, double b0, double& t, int lane) {
  float tf = (float)t;
  for (int it = 0; it < 12; ++it) {
    float g, gp;
    L.slope32(tf, lane, g, gp);
    const float tn = tf - g / gp;
    if (!(gp > 0.f) || !((double)tn > a0 && (double)tn < b0)) return;
    const bool conv = fabsf(tn - tf) <= 1e-5f * fmaxf(1.f, fabsf(tf));
    tf = tn;
    if (conv) {
      t = (double)tf;
      return;
    }
  }
}

template <class LK>
__device__ inline double newton_fp64(const LK& L, int64_t p, double a0, double b0, double t, int lane);

template <class LK, bool FP32_START>
__device__ inline double newton_bracket(const LK& L, int64_t p, int i, int G, double lo, double hi, int lane,
                                        double t_vertex) {
  double a0, b0, t;
  grid_bracket(i, G, lo, hi, a0, b0, t);
  if (!(b0 > a0)) return a0;
  // grid minimum at an end node: where the loss still falls towards that end the search ends there (what the
  // end-slope check inside newton_fp64 returns, here without the interior iterations first)
  if (i == 0 || i == G - 1) {
    double g, gp;
    col_slope(L, p, i == 0 ? a0 : b0, lane, g, gp);
    if (i == 0 ? g >= 0.0 : g <= 0.0) return i == 0 ? a0 : b0;
  }
  if (t_vertex > a0 && t_vertex < b0) t = t_vertex;
  else if constexpr (FP32_START) newton_fp32(L, a0, b0, t, lane);
  return newton_fp64(L, p, a0, b0, t, lane);
}

// safeguarded FP64 Newton on L'(t) in (a0, b0) from t
template <class LK>
__device__ inline double newton_fp64(const LK& L, int64_t p, double a0, double b0, double t, int lane) {
  double a = a0, b = b0;
  bool ends_checked = false;
  for (int it = 0; it < 100; ++it) {
    double g, gp;
    col_slope(L, p, t, lane, g, gp);
    if (g == 0.0) break;
    if (g < 0.0) a = t; else b = t;
    double tn = t - g / gp;
    bool newton = true;
    if (!(gp > 0.0) || !(tn > a && tn < b)) {
      if (!ends_checked) {
        ends_checked = true;
        double ge, gpe;
        col_slope(L, p, a0, lane, ge, gpe);
        if (ge >= 0.0) return a0;
        col_slope(L, p, b0, lane, ge, gpe);
        if (ge <= 0.0) return b0;
      }
      tn = 0.5 * (a + b);
      newton = false;
    }
    // a Newton step h below 1e-5 leaves the new iterate within ~K h^2 of the root (quadratic convergence, K the
    // ratio of L's third to twice its second derivative, ~1 in log delta): far below the ~1e-7 to which the loss
    // itself fixes the minimiser, so it is accepted without another evaluation; bisection steps stop on 1e-13
    const double step = fabs(tn - t), scale = fmax(1.0, fabs(t));
    const bool done = step <= 1e-13 * scale || (b - a) <= 1e-13 || (newton && step <= 1e-5 * scale);
    t = tn;
    if (done) break;
  }
  return t;
}

// NEWTON_ONLY: the launch of the usual Newton search (register path), with the other modes compiled out so their
// registers do not count against its occupancy
template <int PREC, int KPL, bool SAME, bool NEWTON_ONLY>
__device__ __forceinline__ void delta_refine_body(
    int mode, const float* __restrict__ uy, int64_t P, int64_t p_pad, int64_t kpad, const double* __restrict__ s,
    int64_t neff, const double* __restrict__ s_lam, int64_t nlam, double nco, int G, double lo, double hi,
    const int* __restrict__ idx, const double* __restrict__ lam64, double* __restrict__ log_delta,
    double* __restrict__ rWr, double* __restrict__ h2, void* q1a, void* q1b, void* q2a, void* q2b,
    float* __restrict__ q1s, float* __restrict__ q2s, double radius, double judge_tol, int* __restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  // both spectra once per block in LDS (register path), padded with 1 beyond their lengths
  constexpr int NS = KPL > 0 ? 64 * KPL : 1;
  __shared__ double sh_s[NS], sh_l[SAME ? 1 : NS];
  __shared__ float sh_sf[NS], sh_lf[SAME ? 1 : NS];
  if constexpr (KPL > 0) {
    for (int k = threadIdx.x; k < NS; k += blockDim.x) {
      const double v = k < neff ? s[k] : 1.0;
      sh_s[k] = v;
      sh_sf[k] = (float)v;
      if constexpr (!SAME) {
        const double u = k < nlam ? s_lam[k] : 1.0;
        sh_l[k] = u;
        sh_lf[k] = (float)u;
      }
    }
    __syncthreads();
  }
  if (p >= p_pad) return;
  if (p >= P) {  // padding columns (their plane rows are zeroed by delta_planes_kernel)
    if (PREC == PREC_SPLIT_FP16 && lane == 0) {
      q1s[p] = 1.f;
      if (q2s) q2s[p] = 1.f;
    }
    return;
  }
  ColLik<KPL, SAME> L{uy, p_pad, s, neff, s_lam, nlam, nco, sh_s, SAME ? sh_s : sh_l, sh_sf, SAME ? sh_sf : sh_lf, {}};
  L.load(p, lane);
  if constexpr (KPL > 0 && SAME) {
    // Grid minimum at an end node (most permutation columns: h2 ~ 0 puts delta at the upper end).  One sweep gives
    // everything the search and the outputs need there: R at the end node and at its neighbour (the FP64 walk's
    // comparison), the slope sums at the end node (the end check of newton_bracket) and the end node's rWr and
    // plane maxima (pass 1 below), each summed exactly as those steps sum it, so the result is theirs bit for bit.
    // A walk that moves, or a slope that turns back into the bracket, takes the general path.  MODE_LEARNED takes
    // the same sweep (the grid settles these columns without the proposal; they count as grid columns).
    const int ie = ((mode == MODE_NEWTON || mode == MODE_LEARNED) && lam64 && G >= 2) ? idx[p] : -1;
    if (ie == 0 || ie == G - 1) {
      const bool up = ie == G - 1;
      const int in = up ? G - 2 : 1;
      const double de = exp(grid_value(ie, G, lo, hi)), dn = exp(grid_value(in, G, lo, hi));
      double Re = 0.0, Rn = 0.0, R2e = 0.0, W1e = 0.0, rw = 0.0, m1 = 0.0, m2 = 0.0;
#pragma unroll
      for (int j = 0; j < KPL; ++j) {
        const int k = lane + 64 * j;
        const double y = (double)L.yr[j];
        const double sk = sh_s[k];
        Rn += y * y * rcp_pos(sk + dn);
        const double we = rcp_pos(sk + de);
        const double yw = y * y * we;
        Re += yw;
        R2e += yw * we;
        if (k < nlam) W1e += we;
        if (k < neff) {
          const double q1 = we * y;
          rw += q1 * y;
          m1 = fmax(m1, fabs(q1));
          m2 = fmax(m2, we);
        }
      }
      Re = wave_sum(Re);
      Rn = wave_sum(Rn);
      const double le = node_loss_from_R(L, Re, ie, lam64), ln = node_loss_from_R(L, Rn, in, lam64);
      if (up ? !(ln <= le) : !(ln < le)) {
        R2e = wave_sum(R2e);
        W1e = wave_sum(W1e);
        const double g = Re <= FLOOR64 ? de * W1e : de * (W1e - nco * R2e / Re);
        if (up ? g <= 0.0 : g >= 0.0) {
          rw = wave_sum(rw);
          double sc1 = 1.0, sc2 = 1.0;
          if constexpr (PREC == PREC_SPLIT_FP16) {
            sc1 = pow2_scale(wave_max(m1));
            sc2 = pow2_scale(wave_max(m2));
          }
          if (lane == 0) {
            if (flags) flags[p] = 1;
            log_delta[p] = fmin(fmax(up ? hi : lo, lo), hi);
            rWr[p] = fmax(rw, FLOOR64);
            h2[p] = 1.0 / (1.0 + de);
            if constexpr (PREC == PREC_SPLIT_FP16) {
              q1s[p] = (float)sc1;
              if (q2s) q2s[p] = (float)sc2;
            }
          }
          return;
        }
      }
    }
  }
  double t;
  int i = (mode == MODE_FIXED || mode == MODE_GRID) ? 0 : idx[p];
  // every search brackets the FP64 grid argmin, as the reference does (the FP32 grid's pick differs only where
  // neighbouring grid losses tie to FP32 precision, i.e. on flat likelihoods)
  double t_vertex = NAN;
  if (mode != MODE_FIXED && mode != MODE_GRID)
    i = fp64_grid_argmin<ColLik<KPL, SAME>, (KPL > 0)>(L, p, i, G, lo, hi, lam64, lane, t_vertex);
  bool need_bracket = NEWTON_ONLY || mode == MODE_NEWTON;
  if constexpr (NEWTON_ONLY) {
    t = 0.0;
  } else {
    if (mode == MODE_FIXED) {
      t = log_delta[p];  // the caller's log delta, used as given (core.py:639)
    } else if (mode == MODE_GRID) {
      // refine=False: the first argmin of the FP64 loss over the whole grid (core.py:570-573)
      double best = INFINITY;
      int bi = 0;
      for (int g = 0; g < G; ++g) {
        const double l = col_loss(L, p, grid_value(g, G, lo, hi), lane);
        if (l < best) { best = l; bi = g; }
      }
      t = grid_value(bi, G, lo, hi);
    } else if (mode == MODE_GOLDEN) {
      t = golden_bracket(L, p, i, G, lo, hi, lane);
    } else if (mode == MODE_LEARNED) {
      // proposal t0 (in log_delta), undamped Newton inside [t0 - radius, t0 + radius].  Accepted only if it converges
      // there, to a point inside the grid optimum's bracket [g[i-1], g[i+1]] (the basin the grid search refines),
      // whose loss is not worse than the grid optimum's by more than judge_tol; otherwise the grid bracket is refined
      const double t0 = fmin(fmax(log_delta[p], lo), hi);
      const double a_br = grid_value(i > 0 ? i - 1 : 0, G, lo, hi);
      const double b_br = grid_value(i + 1 < G ? i + 1 : G - 1, G, lo, hi);
      double tt = t0;
      bool ok = false;
      for (int it = 0; it < 10; ++it) {
        double g, gp;
        col_slope(L, p, tt, lane, g, gp);
        if (!(gp > 0.0)) break;
        double step = -g / gp;
        step = fmin(fmax(step, -radius), radius);
        const double tn = fmin(fmax(tt + step, lo), hi);
        if (fabs(tn - t0) > radius) break;
        const bool conv = fabs(tn - tt) <= 1e-12 * fmax(1.0, fabs(tt));
        tt = tn;
        if (conv) { ok = true; break; }
      }
      ok = ok && tt >= a_br && tt <= b_br;
      if (ok) {
        const double lstar = col_loss(L, p, tt, lane);
        const double lgrid = col_loss(L, p, grid_value(i, G, lo, hi), lane);
        ok = lstar <= lgrid + judge_tol;
      }
      if (ok) {
        t = tt;
      } else {
        need_bracket = true;
        t = 0.0;
      }
      if (flags && lane == 0) flags[p] = ok ? 0 : 1;
    } else {
      t = 0.0;
    }
  }
  if (need_bracket) t = newton_bracket<ColLik<KPL, SAME>, (KPL > 0)>(L, p, i, G, lo, hi, lane, t_vertex);
  if (mode != MODE_FIXED) t = fmin(fmax(t, lo), hi);
  const double delta = exp(t);
  // pass 1: the null residual rWr and (split-FP16) the column maxima that set the planes' scales
  double rw = 0.0, m1 = 0.0, m2 = 0.0;
  auto elem = [&](int i, int64_t k, double& w, double& q1) {  // weight and w * y of element k (k < neff)
    double y, sk;
    if constexpr (KPL > 0) { y = (double)L.yr[i]; sk = sh_s[k]; }
    else { y = (double)uy[((k >> 5) * p_pad + p) * 32 + (k & 31)]; sk = s[k]; }
    w = rcp_pos(sk + delta);
    q1 = w * y;
  };
  auto pass1 = [&](int i, int64_t k) {
    if (k >= neff) return;
    double w, q1;
    elem(i, k, w, q1);
    const double y = KPL > 0 ? (double)L.yr[i] : (double)uy[((k >> 5) * p_pad + p) * 32 + (k & 31)];
    rw += q1 * y;
    m1 = fmax(m1, fabs(q1));
    m2 = fmax(m2, w);
  };
  if constexpr (KPL > 0) {
#pragma unroll
    for (int i = 0; i < KPL; ++i) pass1(i, lane + 64 * i);
  } else {
    for (int64_t k = lane; k < neff; k += 64) pass1(0, k);
  }
  rw = wave_sum(rw);
  double sc1 = 1.0, sc2 = 1.0;
  if constexpr (PREC == PREC_SPLIT_FP16) {
    sc1 = pow2_scale(wave_max(m1));
    sc2 = pow2_scale(wave_max(m2));
  }
  // the column-side scan planes Q1 = w o y and Q2 = w are written by delta_planes_kernel from log_delta and the
  // scales below (the same arithmetic, with every block's reads and writes contiguous)
  if (lane == 0) {
    log_delta[p] = t;
    rWr[p] = fmax(rw, FLOOR64);
    h2[p] = 1.0 / (1.0 + delta);
    if constexpr (PREC == PREC_SPLIT_FP16) {
      q1s[p] = (float)sc1;
      if (q2s) q2s[p] = (float)sc2;
    }
  }
}

template <int PREC, int KPL, bool SAME, bool NEWTON_ONLY = false>
__global__ __launch_bounds__(256) void delta_refine_kernel(
    int mode, const float* __restrict__ uy, int64_t P, int64_t p_pad, int64_t kpad, const double* __restrict__ s,
    int64_t neff, const double* __restrict__ s_lam, int64_t nlam, double nco, int G, double lo, double hi,
    const int* __restrict__ idx, const double* __restrict__ lam64, double* __restrict__ log_delta,
    double* __restrict__ rWr, double* __restrict__ h2, void* q1a, void* q1b, void* q2a, void* q2b,
    float* __restrict__ q1s, float* __restrict__ q2s, double radius, double judge_tol, int* __restrict__ flags) {
  delta_refine_body<PREC, KPL, SAME, NEWTON_ONLY>(mode, uy, P, p_pad, kpad, s, neff, s_lam, nlam, nco, G, lo, hi, idx, lam64, log_delta, rWr, h2, q1a, q1b, q2a, q2b, q1s, q2s, radius, judge_tol, flags);
}

// the Newton launch at three waves per SIMD (its register budget without the other modes)
template <int PREC, int KPL, bool SAME>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void delta_refine_newton_kernel(
    int mode, const float* __restrict__ uy, int64_t P, int64_t p_pad, int64_t kpad, const double* __restrict__ s,
    int64_t neff, const double* __restrict__ s_lam, int64_t nlam, double nco, int G, double lo, double hi,
    const int* __restrict__ idx, const double* __restrict__ lam64, double* __restrict__ log_delta,
    double* __restrict__ rWr, double* __restrict__ h2, void* q1a, void* q1b, void* q2a, void* q2b,
    float* __restrict__ q1s, float* __restrict__ q2s, double radius, double judge_tol, int* __restrict__ flags) {
  delta_refine_body<PREC, KPL, SAME, true>(mode, uy, P, p_pad, kpad, s, neff, s_lam, nlam, nco, G, lo, hi, idx, lam64, log_delta, rWr, h2, q1a, q1b, q2a, q2b, q1s, q2s, radius, judge_tol, flags);
}

// Q1 = w o y and Q2 = w, w = 1 / (s_k + exp(log_delta[p])), written as the scan's column-side planes (zero on the
// padding k in [neff, kpad) and on the padding columns p in [P, p_pad)); split-FP16 rows are multiplied by the
// scales the refinement chose.  q2a == NULL: Q2 is not written (factored weights, factored_planes_kernel).
// Columns hi_from <= p < P (the certified screen's, split precisions) get the hi planes only, from FP32 arithmetic:
// each value is then within ~5 FP32 rounding units of the FP64 one before its fp16 rounding, which the screen's
// error bound covers; with the bucketed layout (ows, a ScreenOrder workspace), their Q1 hi plane goes to their plane
// slots in the fragment-native layout (frag_off) and Q2 is not written.  Q planes have q_rows rows per K-chunk
// (p_pad, or more for the bucketed layout's padded slots).
// A block owns 32 adjacent columns and walks the K chunks; thread (column tid / 8, quad tid % 8) handles 4
// consecutive k of each chunk, so for one chunk the block's UY read is one contiguous 4 KiB span (a 16-byte load
// per lane) and its writes are contiguous 8-byte stores per plane (element (p, k) at
// ((k / 32) p_pad + p) 32 + k % 32).

constexpr int64_t PL_SF_MAX = 4096;  // spectra up to this kpad are staged in LDS by the bucketed-layout blocks

template <int PREC>
__global__ __launch_bounds__(256) void delta_planes_kernel(const float* __restrict__ uy, int64_t P, int64_t p_pad,
                                                           int64_t kpad, const double* __restrict__ s, int64_t neff,
                                                           const double* __restrict__ log_delta,
                                                           const float* __restrict__ q1s,
                                                           const float* __restrict__ q2s, void* q1a, void* q1b,
                                                           void* q2a, void* q2b, int64_t hi_from, int64_t q_rows,
                                                           const int* __restrict__ ows) {
  __shared__ __attribute__((aligned(16))) float sf_lds[PL_SF_MAX];
  const int j4 = (threadIdx.x & 7) * 4;
  const int64_t nch = kpad / 32;  // a multiple of 4 (kpad % 128 == 0)
  if constexpr (is_split(PREC)) {
    if (ows && (int64_t)blockIdx.x * 32 >= hi_from) {  // block-uniform
      // screened columns with the bucketed layout: this block writes plane slots b .. b + 31 (slot t: column
      // hi_from + order[t]; padding slots are skipped: only their own columns of the product read them) in the
      // fragment-native layout, the Q1 hi half only (the bucketed screen takes sWs from FP64 node sums; hi_from %
      // 256 == 0: blocks do not straddle it)
      // the spectrum in FP32 once per block (+inf past neff, where w = 1 / (inf + delta) = 0)
      const bool staged = kpad <= PL_SF_MAX;
      if (staged) {
        for (int64_t k = threadIdx.x; k < kpad; k += 256) sf_lds[k] = k < neff ? (float)s[k] : INFINITY;
        __syncthreads();
      }
      const ScreenOrder o = screen_order((void*)ows, P - hi_from);
      const int64_t slot = (int64_t)blockIdx.x * 32 - hi_from + (threadIdx.x >> 3);
      if (slot >= gmem(o.head)[0]) return;
      const int col = gmem(o.order)[slot];
      if (col < 0) return;
      const int64_t p = hi_from + col;
      const float d32 = (float)exp(log_delta[p]);
      const float s1f = PREC == PREC_SPLIT_FP16 ? q1s[p] : 1.f;
#pragma unroll 4
      for (int64_t c = 0; c < nch; ++c) {
        const int64_t k0 = c * 32 + j4;
        const f32x4 y4 = *(const f32x4*)(uy + (c * p_pad + p) * 32 + j4);
        f32x4 s4;
        if (staged) {
          s4 = *(const f32x4*)&sf_lds[k0];
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) s4[e] = k0 + e < neff ? (float)s[k0 + e] : INFINITY;
        }
        u16 h[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) h[e] = hi_of<PREC>(__builtin_amdgcn_rcpf(s4[e] + d32) * y4[e] * s1f);
        const int64_t woff = frag_off(c, q_rows, hi_from, slot, j4) / 2;
        *(uint2*)((u16*)q1a + woff) = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
      }
      return;
    }
  }
  const int64_t p = (int64_t)blockIdx.x * 32 + (threadIdx.x >> 3);
  if (p >= p_pad) return;  // rows past the columns' padding (taller Q planes) are only the layout's slots
  const bool live = p < P;
  const double delta = live ? exp(log_delta[p]) : 1.0;
  double sc1 = 1.0, sc2 = 1.0;
  if constexpr (PREC == PREC_SPLIT_FP16) {
    if (live) { sc1 = (double)q1s[p]; sc2 = q2s ? (double)q2s[p] : 1.0; }
  }
  const bool wq2 = q2a != nullptr;
  if constexpr (is_split(PREC)) {
    if (live && p >= hi_from) {
      const float d32 = (float)delta, s1f = (float)sc1, s2f = (float)sc2;
#pragma unroll 2
      for (int64_t c = 0; c < nch; ++c) {
        const int64_t k0 = c * 32 + j4;
        const int64_t off = (c * p_pad + p) * 32 + j4, qoff = (c * q_rows + p) * 32 + j4;
        const f32x4 y4 = *(const f32x4*)(uy + off);
        u16 h[4], g[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float w = 0.f;
          if (k0 + e < neff) w = __builtin_amdgcn_rcpf((float)s[k0 + e] + d32);
          h[e] = hi_of<PREC>(w * y4[e] * s1f);
          g[e] = hi_of<PREC>(w * s2f);
        }
        *(uint2*)((u16*)q1a + qoff) = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
        if (wq2)
          *(uint2*)((u16*)q2a + qoff) = make_uint2(g[0] | ((uint32_t)g[1] << 16), g[2] | ((uint32_t)g[3] << 16));
      }
      return;
    }
  }
#pragma unroll 2
  for (int64_t c = 0; c < nch; ++c) {
    const int64_t k0 = c * 32 + j4;
    const int64_t off = (c * p_pad + p) * 32 + j4, qoff = (c * q_rows + p) * 32 + j4;
    const f32x4 y4 = *(const f32x4*)(uy + off);
    double w[4], q1[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      w[e] = 0.0;
      q1[e] = 0.0;
      if (live && k0 + e < neff) {
        w[e] = rcp_pos(s[k0 + e] + delta);
        q1[e] = w[e] * (double)y4[e];
      }
    }
    if constexpr (is_split(PREC)) {
      u16 h[4], l[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) split2f<PREC>((float)(q1[e] * sc1), h[e], l[e]);
      *(uint2*)((u16*)q1a + qoff) = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
      *(uint2*)((u16*)q1b + qoff) = make_uint2(l[0] | ((uint32_t)l[1] << 16), l[2] | ((uint32_t)l[3] << 16));
      if (wq2) {
#pragma unroll
        for (int e = 0; e < 4; ++e) split2<PREC>(w[e] * sc2, h[e], l[e]);
        *(uint2*)((u16*)q2a + qoff) = make_uint2(h[0] | ((uint32_t)h[1] << 16), h[2] | ((uint32_t)h[3] << 16));
        *(uint2*)((u16*)q2b + qoff) = make_uint2(l[0] | ((uint32_t)l[1] << 16), l[2] | ((uint32_t)l[3] << 16));
      }
    } else {
      *(f32x4*)((float*)q1a + qoff) = f32x4{(float)q1[0], (float)q1[1], (float)q1[2], (float)q1[3]};
      if (wq2) *(f32x4*)((float*)q2a + qoff) = f32x4{(float)w[0], (float)w[1], (float)w[2], (float)w[3]};
    }
  }
}

// Factored weights.  1 / z = sum_r lam_r^2 exp(-t_r z) to a stated relative accuracy on the z range of a scan (a
// positive exponential sum, nodes t_r and lam_r from the caller), so with z = s_k + delta_p
//   sWs = sum_k US_k^2 / (s_k + delta_p) = sum_r C[m][r] B[p][r],  C = (US^2) A,  A[k][r] = lam_r exp(-t_r s_k),
//   B[p][r] = lam_r exp(-t_r delta_p),
// and the scan's second product runs over R ~ 100 terms instead of Neff.  This kernel writes B as the Q2 planes
// (rows p_pad, K = rpad; zero for r >= R and on the padding columns p >= P); split-FP16 rows are multiplied by the
// power of two of their L2 norm (q2_scale[p]), which bounds every entry by 2^14.  A block owns 8 columns; thread
// (column tid / 32, j = tid % 32) handles r = 32 c + j of every chunk c, so each chunk's 256 elements are
// contiguous in the planes.  The norm is a fixed-order sum over the column's 32 lanes.  Columns hi_from <= p < P
// (the certified screen's) get the hi plane only.
template <int PREC>
__global__ __launch_bounds__(256) void factored_planes_kernel(int64_t P, int64_t p_pad, int64_t rpad,
                                                              const double* __restrict__ t,
                                                              const double* __restrict__ lam, int64_t R,
                                                              const double* __restrict__ log_delta, void* qa,
                                                              void* qb, float* __restrict__ qs, int64_t hi_from,
                                                              bool layout) {
  const int64_t p = (int64_t)blockIdx.x * 8 + (threadIdx.x >> 5);  // p < p_pad (p_pad is a multiple of 8)
  const int j = threadIdx.x & 31;
  const bool live = p < P;
  const double delta = live ? exp(log_delta[p]) : 0.0;
  auto val = [&](int64_t r) { return (live && r < R) ? lam[r] * exp(-t[r] * delta) : 0.0; };
  // with the bucketed layout the screened columns' sWs comes from the FP64 node sums: no planes are written for them
  // and their scale is 1 (the screen's records and the resolution both read it, so any power of two agrees)
  if (is_split(PREC) && live && p >= hi_from && layout) {
    if (PREC == PREC_SPLIT_FP16 && j == 0) qs[p] = 1.f;
    return;
  }
  double sc = 1.0;
  if constexpr (PREC == PREC_SPLIT_FP16) {
    double ss = 0.0;
    for (int64_t r = j; r < R; r += 32) {
      const double v = val(r);
      ss += v * v;
    }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 32);
    sc = live ? pow2_scale(sqrt(ss)) : 1.0;
    if (j == 0) qs[p] = (float)sc;
  }
  const int64_t nch = rpad / 32;
  for (int64_t c = 0; c < nch; ++c) {
    const int64_t r = c * 32 + j;
    const int64_t off = (c * p_pad + p) * 32 + j;
    const double v = val(r);
    if constexpr (is_split(PREC)) {
      u16 hh, ll;
      split2<PREC>(v * sc, hh, ll);
      ((u16*)qa)[off] = hh;
      if (!live || p < hi_from) ((u16*)qb)[off] = ll;
    } else {
      ((float*)qa)[off] = (float)v;
    }
  }
}

hipError_t launch_factored_planes(int prec, int64_t P, int64_t p_pad, int64_t rpad, const double* t, const double* lam,
                                  int64_t R, const double* log_delta, void* qa, void* qb, float* qs, int64_t hi_from,
                                  bool layout, hipStream_t stream) {
  const dim3 grid((unsigned)(p_pad / 8));
#define FLMM_FPL(PR)                                                                                                 \
  hipLaunchKernelGGL((factored_planes_kernel<PR>), grid, dim3(256), 0, stream, P, p_pad, rpad, t, lam, R, log_delta, \
                     qa, qb, qs, hi_from, layout)
  if (prec == PREC_SPLIT_BF16) FLMM_FPL(PREC_SPLIT_BF16);
  else if (prec == PREC_SPLIT_FP16) FLMM_FPL(PREC_SPLIT_FP16);
  else FLMM_FPL(PREC_FP32);
#undef FLMM_FPL
  return hipGetLastError();
}

hipError_t launch_delta_weights(const double* s, int64_t neff, const double* s_lam, int64_t nlam, int64_t kpad, int G,
                                double lo, double hi, float* wplane, float* lam, double* lam64, hipStream_t stream) {
  hipLaunchKernelGGL(delta_weights_kernel, dim3(TILE), dim3(256), 0, stream, s, neff, s_lam, nlam, kpad, G, lo, hi,
                     wplane, lam, lam64);
  return hipGetLastError();
}

hipError_t launch_delta_grid(const float* wplane, const float* lam, int G, double nco, const float* uy, int64_t P,
                             int64_t p_pad, int64_t kpad, int* idx, hipStream_t stream) {
  TileArgs<PREC_FP32, 1> args{};
  args.a[0][0] = wplane;
  args.b[0][0] = uy;
  args.a_rows_pad = TILE;
  args.b_rows_pad = p_pad;
  args.nchunks = (int)(kpad / KCHUNK);
  const int64_t n_tiles = p_pad / TILE;
  constexpr int lds = tile_lds_bytes<PREC_FP32, 1, false>();
  hipLaunchKernelGGL(delta_grid_kernel, dim3((unsigned)n_tiles), dim3(NTHREADS), lds, stream, args, n_tiles, lam, G,
                     (float)nco, P, idx);
  return hipGetLastError();
}

hipError_t launch_delta_grid_split(const float* wplane, const float* lam, int G, double nco, const float* uy,
                                   const float* col_scale, double s_min, double lo, double hi, int64_t P, int64_t p_pad,
                                   int64_t kpad, int* idx, void* wsplit, hipStream_t stream) {
  const int64_t n_tiles = p_pad / TILE;
  if (wsplit) {
    hipLaunchKernelGGL(delta_wsplit_kernel, dim3((unsigned)((TILE * kpad + 255) / 256)), dim3(256), 0, stream, wplane,
                       s_min, G, lo, hi, kpad, (u16*)wsplit);
    hipLaunchKernelGGL(delta_grid_split_kernel<true>, dim3((unsigned)n_tiles), dim3(NTHREADS), 0, stream, wplane,
                       (const u16*)wsplit, uy, col_scale, s_min, G, lo, hi, lam, (float)nco, P, p_pad, kpad, n_tiles,
                       idx);
  } else {
    hipLaunchKernelGGL(delta_grid_split_kernel<false>, dim3((unsigned)n_tiles), dim3(NTHREADS), 0, stream, wplane,
                       (const u16*)nullptr, uy, col_scale, s_min, G, lo, hi, lam, (float)nco, P, p_pad, kpad, n_tiles,
                       idx);
  }
  return hipGetLastError();
}

hipError_t launch_delta_refine(int prec, int mode, const float* uy, int64_t P, int64_t p_pad, int64_t kpad,
                               const double* s, int64_t neff, const double* s_lam, int64_t nlam, double nco, int G,
                               double lo, double hi, const int* idx, const double* lam64,
                               double* log_delta, double* rWr, double* h2, void* q1a, void* q1b, void* q2a,
                               void* q2b, float* q1s, float* q2s, double radius, double judge_tol, int* flags,
                               int64_t hi_from, int64_t q_rows, double t_lo, double t_h, int nb, void* order_ws,
                               hipStream_t stream) {
  const dim3 grid((unsigned)((p_pad + 3) / 4));
  // columns that fit 16 elements per lane (neff, nlam, kpad <= 1024) keep them in registers
  const bool regs = neff <= 1024 && nlam <= 1024 && kpad <= 1024;
  const bool same = s_lam == s && nlam == neff;
  // Newton on the register path runs the Newton-only kernel (three waves per SIMD)
  const bool newton = mode == MODE_NEWTON && regs;
#define FLMM_REFINE1(PR, K, SM)                                                                                     \
  do {                                                                                                              \
    if (newton && K > 0)                                                                                            \
      hipLaunchKernelGGL((delta_refine_newton_kernel<PR, K, SM>), grid, dim3(256), 0, stream, mode, uy, P, p_pad,    \
                         kpad, s, neff, s_lam, nlam, nco, G, lo, hi, idx, lam64, log_delta, rWr, h2, q1a, q1b, q2a,  \
                         q2b, q1s, q2s, radius, judge_tol, flags);                                                \
    else                                                                                                            \
      hipLaunchKernelGGL((delta_refine_kernel<PR, K, SM, false>), grid, dim3(256), 0, stream, mode, uy, P, p_pad,   \
                         kpad, s, neff, s_lam, nlam, nco, G, lo, hi, idx, lam64, log_delta, rWr, h2, q1a, q1b, q2a,  \
                         q2b, q1s, q2s, radius, judge_tol, flags);                                                \
  } while (0)
#define FLMM_REFINE(PR, K)                                                                                          \
  do {                                                                                                              \
    if (same) FLMM_REFINE1(PR, K, true); else FLMM_REFINE1(PR, K, false);                                           \
  } while (0)
  if (prec == PREC_SPLIT_BF16) {
    if (regs) FLMM_REFINE(PREC_SPLIT_BF16, 16); else FLMM_REFINE(PREC_SPLIT_BF16, 0);
  } else if (prec == PREC_SPLIT_FP16) {
    if (regs) FLMM_REFINE(PREC_SPLIT_FP16, 16); else FLMM_REFINE(PREC_SPLIT_FP16, 0);
  } else {
    if (regs) FLMM_REFINE(PREC_FP32, 16); else FLMM_REFINE(PREC_FP32, 0);
  }
#undef FLMM_REFINE
#undef FLMM_REFINE1
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the screened columns' layout from the fitted log delta: bucketed by log delta
  const int* ows = nullptr;
  if (order_ws && hi_from < P) {
    e = launch_screen_classify(log_delta + hi_from, P - hi_from, t_lo, t_h, nb, order_ws, stream);
    if (e != hipSuccess) return e;
    ows = (const int*)order_ws;
  }
  const dim3 pgrid((unsigned)(q_rows / 32));
#define FLMM_PLANES(PR)                                                                                              \
  hipLaunchKernelGGL((delta_planes_kernel<PR>), pgrid, dim3(256), 0, stream, uy, P, p_pad, kpad, s, neff, log_delta, \
                     q1s, q2s, q1a, q1b, q2a, q2b, hi_from, q_rows, ows)
  if (prec == PREC_SPLIT_BF16) FLMM_PLANES(PREC_SPLIT_BF16);
  else if (prec == PREC_SPLIT_FP16) FLMM_PLANES(PREC_SPLIT_FP16);
  else FLMM_PLANES(PREC_FP32);
#undef FLMM_PLANES
  return hipGetLastError();
}

}  // namespace flmm
