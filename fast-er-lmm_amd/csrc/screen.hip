// Certified screen of the permutation columns' Wald scan, and the exact resolution of each column's maximum F.
//
// Permutation columns feed only max_F, each column's largest Wald F over the SNPs (perm_threshold, reference
// fasterlmm/perms.py:60-68 over the scan of core.py:640-667).  For one column, F of a test depends on the sums
// a = sWy and d = sWs only through the weighted cosine c = |a| / sqrt(d rWr):  F = nu c^2 / (1 - c^2), increasing in
// c.  The screen tile runs the scan's two products on the hi halves of the split-FP16 planes only (one MFMA where the
// split takes three) and keeps, per column and 128-row tile, the largest and second-largest x = a^2 / d (up to the
// column constant, c^2 = x q2 / (q1^2 rWr)) and the row of the largest.
//
// Error bound.  Each hi value is its operand rounded to fp16 from an FP32 value within 2^-21 (relative) of the FP64
// one (the SNP planes round the FP64 value once; the screened columns' hi planes come from FP32 arithmetic): relative
// error at most u' = 2^-11 + 2^-21 in the normal range, and at most eps absolute below it (eps = 2^-25: gfx950 keeps
// fp16 subnormals in the conversion and in v_mfma_f32_32x32x16_f16; 2^-14 would cover flushing).  With w_k = 1 / (s_k + delta), Cauchy-Schwarz in the weights gives
// sum_k |us_k| w_k |uy_k| <= sqrt(d rWr), so for the FP64 values a, d of the test
//   |a_hi - a| <= kA sqrt(d rWr),   kA = 2u' + u'^2 + gamma(2K) (1 + 2u') + kArel R   (R = sqrt(w_max / w_min))
//   |d_hi - d| <= kD d,             kD = 2u' + u'^2 + gamma(2K2) (1 + 2u') + e_fac + eta + slack
// where gamma(n) = n 2^-24 / (1 - n 2^-24) bounds the FP32 accumulation, e_fac the factored weights' accuracy, the
// kArel R term the subnormal part of the first product (its operands' rows are scaled to norms in (2^13, 2^14]), and
// eta the subnormal part of the second product, which holds for rows whose scaled d is at least dmin (rows below are
// forced candidates).  Hence every test's exact c lies in
//   [cL, cU] = [c sqrt(1 - kD) - kA,  c sqrt(1 + kD) + kA sqrt((1 + kD) / (1 - kD))],
// both increasing in the screened c.  The resolve kernel takes per column L = the largest cL over the tiles and the
// running maximum; a test whose cU is below L cannot hold the new maximum.  It re-evaluates in FP64 (FP64 rotated
// SNPs, the rotated column, exact weights) the best row of every tile whose best cU >= L and every row of a tile whose
// second-best cU >= L (or that holds a forced row), and max_F is the maximum of the running value and those
// evaluations: the maximum of FP64 evaluations of the only tests that can hold it, whatever the screen's rounding.
#include "kernels.h"
#include "scan_wide.h"

namespace flmm {

// ScreenArgs: kernels.h

constexpr int64_t SC_GM = 16;                      // SNP tiles per scheduling group of the screen tiles
constexpr int SC_SLOTS = 4;                        // K-chunk products per pipeline stage
constexpr int SC_SLOT_LDS = W_TM * 64;             // one slot: 128 rows x 32 fp16
constexpr int SC_STAGE = SC_SLOTS * SC_SLOT_LDS;   // 32 KiB
static_assert(2 * EP_WAVE <= SC_STAGE, "epilogue staging of two waves fits one stage buffer");

struct ScASet {
  u32x4 v[SC_SLOTS][2];      // SNP-side blocks in flight: [slot][1 KiB block]
};
struct ScBSet {
  u32x4 v[SC_SLOTS][2][2];   // column-side fragments: [slot][column block][MFMA step]
};

__device__ inline int sc_swz(int row) { return (row >> 2) & 3; }

// One 128 x 256 screen tile.  Stage j holds four K-chunk products: while j < nSA the first product's chunks
// n2 + 4 j + i (i = 0..3), all into accumulator set 0; then both products' chunks 2 jb and 2 jb + 1 (jb = j - nSA)
// as slots (S1 Q1, S2 Q2, S1 Q1, S2 Q2) into sets 0, 1, 0, 1.  A stage's operands are fetched into registers one
// stage ahead, the SNP side then stored into the other LDS buffer.
__device__ inline void screen_tile(const ScreenArgs& g, int64_t n_mtiles, int64_t n_ptiles, int64_t tile, int64_t Mc,
                                   int64_t P, float dmin, uint64_t* __restrict__ rec, int64_t rec_ld) {
  __shared__ __attribute__((aligned(16))) char lds0[SC_STAGE];
  __shared__ __attribute__((aligned(16))) char lds1[SC_STAGE];
  __shared__ __attribute__((aligned(16))) float row_g[W_TM];   // s2 / s1^2 of each tile row (0 past Mc)
  __shared__ __attribute__((aligned(16))) uint32_t row_dm[W_TM];  // 0, or +inf bits past Mc (padding)

  const int64_t group = tile / (SC_GM * n_ptiles);
  const int64_t first_m = group * SC_GM;
  const int64_t gm = n_mtiles - first_m < SC_GM ? n_mtiles - first_m : SC_GM;
  const int64_t r_in = tile - group * (SC_GM * n_ptiles);
  const int64_t mt = first_m + r_in % gm;
  const int64_t pt = r_in / gm;
  const int64_t m0 = mt * W_TM, p0 = pt * W_TP;
  const int64_t n_cols = P;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int h = lane >> 5;

  float e_g = 0.f;
  uint32_t e_dm = 0x7f800000u;
  if (threadIdx.x < W_TM) {
    const int64_t m = m0 + threadIdx.x;
    if (m < Mc) {
      const double sa = gmem(g.s1s)[m];
      e_g = (float)((double)gmem(g.s2s)[m] / (sa * sa));
      e_dm = 0u;
    }
  }

  const int n2 = g.nchunks2;
  const int nSA = (g.nchunks - n2) / 4;
  const int nst = nSA + n2 / 2;
  const int64_t a_kstride = g.a_rows_pad * 64;
  const int64_t b_kstride = g.b_rows_pad * 64;
  // SNP side: wave w copies 1 KiB blocks 2w and 2w + 1 (16 rows each) of every slot; lane L: row 16 b + L / 4,
  // 16-byte segment L % 4, stored at segment (L % 4) ^ swz(row)
  // (32-bit lane offsets from uniform chunk bases, so every load takes the scalar-base + 32-bit-offset form; the
  // host bounds rows_pad * 64 bytes below 2^31)
  const char* const a_base = (const char*)g.s1 + m0 * 64;
  const char* const a_base2 = (const char*)g.s2 + m0 * 64;
  const char* const b_base = (const char*)g.q1;
  const char* const b_base2 = (const char*)g.q2;
  const uint32_t a_lane = wave * 2048 + lane * 16;
  const int a_row0 = wave * 32 + lane / 4;
  // column side: lane (column p0 + 64 wave + 32 bj + lane % 32, half h) reads bytes [32 h, 32 h + 32) of its row
  // (columns past n_cols read the last one; their records are not written)
  uint32_t b_off[2];
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int64_t j = p0 + wave * 64 + bj * 32 + (lane & 31);
    b_off[bj] = (uint32_t)((j < n_cols ? j : n_cols - 1) * 64 + h * 32);
  }

  // uniform chunk base of slot i of stage j, SNP side (a) and column side (b)
  auto slot_base = [&](int j, int i, bool a_side) -> const char* {
    const bool A = j < nSA;
    const int ck = A ? n2 + 4 * j + i : 2 * (j - nSA) + (i >> 1);
    const bool second = !A && (i & 1);
    return a_side ? (second ? a_base2 : a_base) + (int64_t)ck * a_kstride
                  : (second ? b_base2 : b_base) + (int64_t)ck * b_kstride;
  };
  auto load_a = [&](int j) {
    ScASet a;
#pragma unroll
    for (int i = 0; i < SC_SLOTS; ++i) {
      const char* sa = slot_base(j, i, true);
      a.v[i][0] = *gmem((const u32x4*)(sa + a_lane));
      a.v[i][1] = *gmem((const u32x4*)(sa + (a_lane + 1024)));
    }
    return a;
  };
  auto load_b = [&](int j) {
    ScBSet b;
#pragma unroll
    for (int i = 0; i < SC_SLOTS; ++i) {
      const char* sb = slot_base(j, i, false);
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int q = 0; q < 2; ++q) b.v[i][bj][q] = *gmem((const u32x4*)(sb + (b_off[bj] + q * 16)));
    }
    return b;
  };
  auto store = [&](const ScASet& a, char* st) {
#pragma unroll
    for (int i = 0; i < SC_SLOTS; ++i)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int r = a_row0 + b * 16;
        *(u32x4*)(st + i * SC_SLOT_LDS + r * 64 + (((lane & 3) ^ sc_swz(r)) * 16)) = a.v[i][b];
      }
  };

  f32x16 acc[2][4][2];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[s][i][j][v] = 0.f;

  // `both` (std::true_type / std::false_type): a stage of both products (slot i into set i % 2) or of the first only
  auto mma = [&](auto both, const char* st, const ScBSet& b) {
    constexpr bool B2 = decltype(both)::value;
#pragma unroll
    for (int i = 0; i < SC_SLOTS; ++i)
#pragma unroll
      for (int step = 0; step < 2; ++step) {
        const int seg = h * 2 + step;
        bf16x8 af[4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) {
          const int r = bi * 32 + (lane & 31);
          af[bi] = *(const bf16x8*)(st + i * SC_SLOT_LDS + r * 64 + ((seg ^ sc_swz(r)) * 16));
        }
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const bf16x8 bh = __builtin_bit_cast(bf16x8, b.v[i][bj][step]);
#pragma unroll
          for (int bi = 0; bi < 4; ++bi)
            acc[B2 ? (i & 1) : 0][bi][bj] = mfma_split<PREC_SPLIT_FP16>(af[bi], bh, acc[B2 ? (i & 1) : 0][bi][bj]);
        }
      }
    // interleave (masks: 0x008 MFMA, 0x020 VMEM read, 0x100 DS read): the first slot-step's 4 fragment reads up
    // front, each later slot-step's 4 reads and 3 of the stage's 24 global loads over the previous one's 8 MFMAs
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
    for (int ss = 0; ss < 2 * SC_SLOTS - 1; ++ss) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
  };
  auto stage = [&](auto both, int j, const char* cur, char* nxt, const ScBSet& bc, ScBSet& bn) {
    const int jn = j + 1 < nst ? j + 1 : j;  // past the last stage the prefetch repeats it (valid addresses)
    const ScASet an = load_a(jn);
    bn = load_b(jn);
    mma(both, cur, bc);
    __builtin_amdgcn_sched_barrier(0);
    store(an, nxt);  // nxt was last read in the previous stage, which ended with a barrier
    __syncthreads();
  };
  constexpr std::true_type both{};
  constexpr std::false_type first{};

  ScBSet b0 = load_b(0), b1;
  store(load_a(0), lds0);
  if (threadIdx.x < W_TM) {
    row_g[threadIdx.x] = e_g;
    row_dm[threadIdx.x] = e_dm;
  }
  __syncthreads();
  // Stages in pairs so every access names a fixed buffer: phase A pairs, then (for an odd phase A) its last stage
  // paired with the first phase-B stage, then phase-B pairs and a last single stage.  Straight-line phases: a phase
  // branch inside one loop made the register allocator spill the accumulators.
  int j = 0;
  for (; j + 2 <= nSA; j += 2) {
    stage(first, j, lds0, lds1, b0, b1);
    stage(first, j + 1, lds1, lds0, b1, b0);
  }
  if (j < nSA) {  // nst > nSA: nchunks2 >= 4
    stage(first, j, lds0, lds1, b0, b1);
    stage(both, j + 1, lds1, lds0, b1, b0);
    j += 2;
  }
  for (; j + 2 <= nst; j += 2) {
    stage(both, j, lds0, lds1, b0, b1);
    stage(both, j + 1, lds1, lds0, b1, b0);
  }
  if (j < nst) stage(both, j, lds0, lds1, b0, b1);

  // ---- epilogue: each 32-row block of the wave's accumulators is staged through the free LDS column-major (the
  // wide tile's swizzled layout); lane c then owns column p0 + 64 wave + c and walks the rows:
  //   x = A^2 g_m / max(D, dmin)  (A, D the raw sums; g_m = s2_m / s1_m^2, so x = q1^2 a^2 / (q2 d));
  // x and D are non-negative, so their float bit patterns order as unsigned integers, and the running maxima are
  // integer maxima (no NaN canonicalisation on the LDS values): m1 = the largest of (bits(x) & ~127) | row (the value
  // rounded down to 128 units in the last place, the row in the freed bits), m2 = the second largest, and dlo =
  // the smallest D of the tile's rows (padding rows count as +inf); a row with D < dmin makes the tile a forced
  // one.  Padding rows have A = D = 0 and g_m = 0, so x = 0.
  float* ep = (float*)((wave < 2 ? lds0 : lds1) + (wave & 1) * EP_WAVE);
  uint32_t m1 = 0u, m2 = 0u, dlo = 0x7f800000u;
  const uint32_t dmin_b = __float_as_uint(dmin);
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const f32x16& c = acc[s][bi][bj];
          const f32x4 w4 = {c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3]};
          *(f32x4*)(ep + ep_off(s * 64 + bj * 32 + (lane & 31), 2 * q + h)) = w4;
        }
    __builtin_amdgcn_wave_barrier();
#pragma unroll 1
    for (int t4 = 0; t4 < 8; ++t4) {  // not unrolled: unrolled, its LDS reads crowd the main loop's registers
      const f32x4 a4 = *(const f32x4*)(ep + ep_off(lane, t4));
      const u32x4 d4 = *(const u32x4*)(ep + ep_off(64 + lane, t4));
      const f32x4 g4 = *(const f32x4*)&row_g[bi * 32 + 4 * t4];
      const u32x4 p4 = *(const u32x4*)&row_dm[bi * 32 + 4 * t4];
      const uint32_t row0 = bi * 32 + 4 * t4;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t db = d4[e] > dmin_b ? d4[e] : dmin_b;
        const float x = (a4[e] * a4[e]) * g4[e] * __builtin_amdgcn_rcpf(__uint_as_float(db));
        const uint32_t xp = (__float_as_uint(x) & ~127u) | (row0 + e);
        const uint32_t lo = xp < m1 ? xp : m1;
        m2 = lo > m2 ? lo : m2;
        m1 = xp > m1 ? xp : m1;
        const uint32_t dp = d4[e] | p4[e];
        dlo = dp < dlo ? dp : dlo;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  const int64_t pc = p0 + wave * 64 + lane;
  if (pc < n_cols) {
    // high word: m2 rounded down to 128 units in the last place, bit 0 = forced
    const uint32_t x2 = (m2 & ~127u) | (dlo < dmin_b ? 1u : 0u);
    gmem(rec)[mt * rec_ld + pc] = ((uint64_t)x2 << 32) | m1;
  }
}

__global__ __launch_bounds__(256, 1) void wald_screen_kernel(ScreenArgs g, int64_t n_mtiles, int64_t n_ptiles,
                                                             int64_t n_tiles, int64_t Mc, int64_t P, float dmin,
                                                             uint64_t* __restrict__ rec, int64_t rec_ld) {
  screen_tile(g, n_mtiles, n_ptiles, xcd_remap(blockIdx.x, n_tiles), Mc, P, dmin, rec, rec_ld);
}

// ---- bucketed screen ----
// sWs of SNP row m for a column at log delta t is d_m(t) = sum_k c_k f_k(t), c_k = US_mk^2, f_k = 1 / (s_k + e^t):
// positive, decreasing and smooth in t.  The screened columns are bucketed by t (ScreenOrder: buckets of width
// h <= 1/5 from t_lo, their plane slots adjacent); for bucket b the FP64 sums D_j at the three nodes t_b + j h / 2
// (dnode, one FP64 GEMM of US^2 with the node weights per SNP block) give g = 1 / d by quadratic interpolation,
// g~ = G0 + u (C1 + u C2) (G_j = 1 / D_j, C1 = -3 G0 + 4 G1 - G2, C2 = 2 (G0 - 2 G1 + G2), u = (t - t_b) / h), so
// the tile's epilogue takes two FMAs and two products per test.  Error: let q_k = e^t / (s_k + e^t) in [0, 1] and m_i = E[q^i] under the weights c_k f_k / d, and write
// L1, L2, L3 for the first three derivatives of log d.  Then L1 = -m1, L2 = -m1' with m1' = E[q (1 - q)] - Var q
// (modulus <= 1/4), L3 = -m1'' with m1'' = m1' (1 + 2 m1) - 2 (E[2 q^2 (1 - q)] - Cov(q, q^2)) (modulus at most
// 3/4 + 16/27), and the third derivative of g = exp(-log d) is (m1^3 + 3 m1 L2 - L3) g, of modulus at most 3.1 g.
// g grows by at most e^h over a bucket, so |g~ - g| <= 3.1 e^h 0.0481 h^3 g / 6 (0.0481 h^3 = the nodes' largest
// |(t - t0)(t - t1)(t - t2)|): 2.4e-4 g at h = 1/5, and d~ = 1 / g~ is within 2.5e-4 of d, inside kD together with
// the FP32 evaluation (~1e-6).  The tile runs the sWy product alone: 128 accumulators per lane, so two workgroups
// share each CU and one's epilogue and load waits overlap the other's MFMAs.  With A the raw hi-plane sum
// (s1_m q1_p a),
//   x = A^2 g~ / (s1_m^2 q2_p)     (the tile keeps G0, C1, C2 over s1_m^2 per row, from FP64)
// is the general tile's x = q1^2 a^2 / (q2 d) with d~ for the exact d, so the records mean the same and one
// resolution serves both; 1 / q2_p is a power of two applied to the packed maxima at the end (exact; a result outside
// the normal range marks the record forced).  Rows whose node sums are not positive, or whose scaled reciprocals are
// not finite FP32 numbers, get G = 0 (x = 0) and make their tile forced; padding rows get G = 0.
constexpr int SU_SLOTS = 2;                        // K-chunks per stage
// SNP rows per scheduling group of the bucketed tiles: 12 tiles of 128 rows (at N = 1,000 x 648k screened columns
// x 4,546 SNPs, 36 row tiles: 6.47 ms and 13.6 GB fetched per call, against 6.53 ms / 15.9 GB for 16 and
// 6.60 ms / 11.2 GB for 8; fewer rows keep the group's SNP planes in L2, more share each column chunk more ways)
constexpr int64_t SB_ROWS_PER_GROUP = 12 * W_TM;
// 128-row halves per bucketed workgroup: 1 (two workgroups per CU, whose stages drift apart so one's epilogue and
// load waits overlap the other's MFMAs) measured 6.6 ms per chromosome at N = 1,000 x 648k screened columns against
// 7.4 ms for 2 (one 512-thread workgroup per CU, half the SNP-side bytes per flop, waves in lockstep at each stage)
constexpr int SB_WM = 1;

// the median of three unsigned values (one VALU instruction)
__device__ inline uint32_t umed3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
constexpr int SU_STAGE = SU_SLOTS * SC_SLOT_LDS;   // 16 KiB

// WM row halves: a workgroup of 256 WM threads owns 128 WM SNP rows x 256 slots; wave w computes rows
// 128 (w / 4) .. + 127 of the columns 64 (w % 4) .. + 63, as one 128-row tile of its own (records per 128 rows)
template <int WM>
__global__ __launch_bounds__(256 * WM, 2 / WM) void screen_bucketed_kernel(BucketArgs g, int64_t n_mtiles,
                                                                           uint64_t* __restrict__ rec,
                                                                           int64_t rec_ld) {
  constexpr int NT = 256 * WM;
  constexpr int64_t GM = SB_ROWS_PER_GROUP / (W_TM * WM);  // workgroup row tiles per scheduling group
  __shared__ __attribute__((aligned(16))) char lds0[WM * SU_STAGE];
  __shared__ __attribute__((aligned(16))) char lds1[WM * SU_STAGE];
  __shared__ __attribute__((aligned(16))) float fac[4 * WM][2][3][W_TM];  // [wave][column block][coefficient][row]
  __shared__ int blk_forced[4 * WM][2];

  const int64_t n_cols = gmem(g.head)[0];  // plane slots, padding included
  const int64_t n_ptiles = (n_cols + W_TP - 1) / W_TP;
  const int64_t n_wtiles = (n_mtiles + WM - 1) / WM;
  const int64_t n_live = n_wtiles * n_ptiles;
  if ((int64_t)blockIdx.x >= n_live) return;
  const int64_t tile = xcd_remap(blockIdx.x, n_live);
  const int64_t group = tile / (GM * n_ptiles);
  const int64_t first_m = group * GM;
  const int64_t gm = n_wtiles - first_m < GM ? n_wtiles - first_m : GM;
  const int64_t r_in = tile - group * (GM * n_ptiles);
  const int64_t mtw = first_m + r_in % gm;
  const int64_t pt = r_in / gm;
  const int64_t p0 = pt * W_TP;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cwave = wave & 3;                // the wave's 64 columns
  const int64_t mt = mtw * WM + (wave >> 2);  // the wave's 128-row tile
  const int64_t m0 = mtw * WM * W_TM;         // the workgroup's first row
  const int h = lane >> 5;

  // the buckets of the tile's eight 32-slot column blocks (every block lies in one bucket; -1 past the slots)
  auto block_bucket = [&](int w, int bj) -> int {
    const int64_t s0 = p0 + (w & 3) * 64 + bj * 32;
    return s0 < n_cols ? gmem(g.gbucket)[s0 / SO_PAD] : -1;
  };
  if (threadIdx.x < 8 * WM) blk_forced[threadIdx.x >> 1][threadIdx.x & 1] = 0;
  __syncthreads();
  // the interpolant of 1 / (s1_m^2 sWs) of each wave's 128 rows for each of its column blocks' buckets (G0, C1, C2
  // from the reciprocal node sums in FP64, rounded to FP32; 0 on padding and bad rows)
  for (int q = threadIdx.x; q < 8 * WM * W_TM; q += NT) {
    const int wb = q / W_TM, r = q % W_TM, w = wb >> 1, bj = wb & 1;
    const int bk = block_bucket(w, bj);
    const int64_t m = m0 + (w >> 2) * W_TM + r;
    float f[3] = {0.f, 0.f, 0.f};
    if (bk >= 0 && m < g.mc) {
      const double sc = (double)gmem(g.s1s)[m];
      bool bad = false;
      double gj[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double d = gmem(g.dnode)[m * SO_NODES + 2 * bk + j];
        gj[j] = 1.0 / (d * sc * sc);
        bad |= !(d > 0.0);
      }
      f[0] = (float)gj[0];
      f[1] = (float)(-3.0 * gj[0] + 4.0 * gj[1] - gj[2]);
      f[2] = (float)(2.0 * (gj[0] - 2.0 * gj[1] + gj[2]));
#pragma unroll
      for (int j = 0; j < 3; ++j) bad |= !(fabsf(f[j]) < INFINITY);
      if (bad) {
        blk_forced[w][bj] = 1;
        f[0] = f[1] = f[2] = 0.f;
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) fac[w][bj][j][r] = f[j];
  }

  const int64_t a_kstride = g.a_rows_pad * 64;
  const int64_t b_kstride = g.b_rows_pad * 64;
  const char* const a_base = (const char*)g.s1 + m0 * 64;
  const char* const b_base = (const char*)g.q1;
  // SNP side: thread t copies 16-byte segment t % 4 of rows t / 4 and 64 WM + t / 4 of every chunk; rows past the
  // planes (the second half of a workgroup past the last 128-row tile) read the first half's (their records are not
  // written)
  const uint32_t a_lane = threadIdx.x * 16;
  const uint32_t a_lane2 = m0 + 64 * WM + (threadIdx.x >> 2) < g.a_rows_pad ? a_lane + 4096 * WM : a_lane;
  const int a_row0 = threadIdx.x >> 2;
  // column side (fragment-native layout): lane (c, h) of column block bj reads its 16 bytes of the step's 1 KiB
  // fragment; slots past n_cols read the last one (their records are not written)
  uint32_t b_off[2];
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int64_t j = p0 + cwave * 64 + bj * 32 + (lane & 31);
    const int64_t sl = j < n_cols ? j : n_cols - 1;
    b_off[bj] = (uint32_t)((sl >> 5) * 2048 + (h * 32 + (sl & 31)) * 16);
  }

  struct ASet { u32x4 v[SU_SLOTS][2]; };
  struct BSet { u32x4 v[SU_SLOTS][2][2]; };
  auto load_a = [&](int j) {
    ASet a;
#pragma unroll
    for (int i = 0; i < SU_SLOTS; ++i) {
      const char* sa = a_base + (int64_t)(j * SU_SLOTS + i) * a_kstride;
      a.v[i][0] = *gmem((const u32x4*)(sa + a_lane));
      a.v[i][1] = *gmem((const u32x4*)(sa + a_lane2));
    }
    return a;
  };
  auto load_b = [&](int j, BSet& b) {
#pragma unroll
    for (int i = 0; i < SU_SLOTS; ++i) {
      const char* sb = b_base + (int64_t)(j * SU_SLOTS + i) * b_kstride;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int q = 0; q < 2; ++q) b.v[i][bj][q] = *gmem((const u32x4*)(sb + (b_off[bj] + q * 1024)));
    }
  };
  auto store = [&](const ASet& a, char* st) {
#pragma unroll
    for (int i = 0; i < SU_SLOTS; ++i)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int r = a_row0 + b * 64 * WM;
        *(u32x4*)(st + i * WM * SC_SLOT_LDS + r * 64 + (((threadIdx.x & 3) ^ sc_swz(r)) * 16)) = a.v[i][b];
      }
  };

  f32x16 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  auto mma = [&](const char* st, const BSet& b) {
#pragma unroll
    for (int i = 0; i < SU_SLOTS; ++i)
#pragma unroll
      for (int step = 0; step < 2; ++step) {
        const int seg = h * 2 + step;
        bf16x8 af[4];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) {
          const int r = (wave >> 2) * W_TM + bi * 32 + (lane & 31);
          af[bi] = *(const bf16x8*)(st + i * WM * SC_SLOT_LDS + r * 64 + ((seg ^ sc_swz(r)) * 16));
        }
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          const bf16x8 bh = __builtin_bit_cast(bf16x8, b.v[i][bj][step]);
#pragma unroll
          for (int bi = 0; bi < 4; ++bi) acc[bi][bj] = mfma_split<PREC_SPLIT_FP16>(af[bi], bh, acc[bi][bj]);
        }
      }
    // interleave as the general tile: the first slot-step's 4 fragment reads up front, then each later slot-step's
    // 4 reads and 4 of the stage's global loads over the previous one's 8 MFMAs
    constexpr int NLOAD = SU_SLOTS * 6;  // A: 2 per slot, B: 4 per slot
    constexpr int PER = NLOAD / (2 * SU_SLOTS - 1) > 0 ? NLOAD / (2 * SU_SLOTS - 1) : 1;
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
    for (int ss = 0; ss < 2 * SU_SLOTS - 1; ++ss) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        if (r < PER) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
    }
    __builtin_amdgcn_sched_group_barrier(0x020, NLOAD - PER * (2 * SU_SLOTS - 1), 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 8, 0);
  };
  const int nst = g.nchunks / SU_SLOTS;
  auto stage = [&](int j, const char* cur, char* nxt, const BSet& bc, BSet& bn) {
    const int jn = j + 1 < nst ? j + 1 : j;  // past the last stage the prefetch repeats it (valid addresses)
    const ASet an = load_a(jn);
    load_b(jn, bn);
    mma(cur, bc);
    __builtin_amdgcn_sched_barrier(0);
    store(an, nxt);  // nxt was last read in the previous stage, which ended with a barrier
    __syncthreads();
  };

  BSet b0, b1;
  load_b(0, b0);
  store(load_a(0), lds0);
  __syncthreads();
  for (int j = 0; j < nst; j += 2) {  // nchunks is a multiple of 4: stages pair up
    stage(j, lds0, lds1, b0, b1);
    stage(j + 1, lds1, lds0, b1, b0);
  }

  // each lane's column: u = (t - t_b) / h in its bucket, t its log delta; a column outside its bucket (log delta past
  // the buckets' range, clamped into the first or last) is forced
  const double t_lo = gmem(g.tb)[0], t_h = gmem(g.tb)[1];
  int col[2];
  bool out_b[2];
  float uu[2];
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int64_t j = p0 + cwave * 64 + bj * 32 + (lane & 31);
    const int bk = block_bucket(wave, bj);
    col[bj] = (j < n_cols && bk >= 0 && mt < n_mtiles) ? gmem(g.order)[j] : -1;
    double u = 0.0;
    if (col[bj] >= 0) u = (gmem(g.log_delta)[col[bj]] - (t_lo + bk * t_h)) / t_h;
    out_b[bj] = !(u >= -1e-9 && u <= 1.0 + 1e-9);
    uu[bj] = (float)fmin(fmax(u, 0.0), 1.0);
  }

  // epilogue from the accumulators: lane (h, c) holds column 32 bj + c of its wave and rows 32 bi + 8 q + 4 h + i;
  // x = A^2 g~ / s1^2, packed as in the general tile ((bits(x) & ~127) | row), top two per column over the lane's 64
  // rows, then merged with lane c + 32 (the column's other 64 rows)
  uint32_t m1[2] = {0u, 0u}, m2[2] = {0u, 0u};
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r0 = bi * 32 + 8 * q + 4 * h;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const f32x4 d0 = *(const f32x4*)&fac[wave][bj][0][r0];
        const f32x4 d1 = *(const f32x4*)&fac[wave][bj][1][r0];
        const f32x4 d2 = *(const f32x4*)&fac[wave][bj][2][r0];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float a = acc[bi][bj][4 * q + e];
          const float gg = fmaf(uu[bj], fmaf(uu[bj], d2[e], d1[e]), d0[e]);
          // gg >= 0 up to rounding (g~ is within 2.5e-4 of g > 0, or the row's coefficients are 0); the sign bit is
          // dropped with the low bits, so -0 packs as 0
          const float x = (a * a) * gg;
          const uint32_t xp = (__float_as_uint(x) & 0x7fffff80u) | (uint32_t)(r0 + e);
          m2[bj] = umed3(xp, m1[bj], m2[bj]);  // = max(m2, min(xp, m1)) as m2 <= m1
          m1[bj] = xp > m1[bj] ? xp : m1[bj];
        }
      }
    }
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const uint32_t o1 = (uint32_t)__shfl_xor((int)m1[bj], 32, 64), o2 = (uint32_t)__shfl_xor((int)m2[bj], 32, 64);
    const uint32_t lo = m1[bj] < o1 ? m1[bj] : o1;
    const uint32_t hi2 = m2[bj] > o2 ? m2[bj] : o2;
    m2[bj] = lo > hi2 ? lo : hi2;
    m1[bj] = m1[bj] > o1 ? m1[bj] : o1;
  }
  // lane (h, c) writes column block h's column c
  const uint32_t w1 = h ? m1[1] : m1[0], w2 = h ? m2[1] : m2[0];
  const int cw = h ? col[1] : col[0];
  if (cw >= 0) {
    // 1 / q2 = 2^e: add e to the exponent of both packed values (the low 7 bits carry the row / flag)
    const int e = 127 - (int)((__float_as_uint(gmem(g.q2s)[cw]) >> 23) & 255u);
    bool bad = blk_forced[wave][h] != 0 || (h ? out_b[1] : out_b[0]);
    auto scale = [&](uint32_t v) -> uint32_t {
      if ((v & ~127u) == 0u) return v & 127u;
      const int ex = (int)((v >> 23) & 255u);
      if (ex == 0 || ex == 255 || ex + e <= 0 || ex + e >= 255) {
        bad = true;
        return v;
      }
      return v + ((uint32_t)e << 23);
    };
    const uint32_t s1v = scale(w1);
    const uint32_t s2v = scale(w2 & ~127u);
    const uint32_t x2 = (s2v & ~127u) | (bad ? 1u : 0u);
    gmem(rec)[mt * rec_ld + cw] = ((uint64_t)x2 << 32) | s1v;
  }
}

// The screened columns' layout (ScreenOrder): order[] cleared to -1, per-1,024-column block counts of each bucket,
// one block's scan of them (bucket starts padded to SO_PAD, the slot groups' buckets), then each column's slot from
// its block's offset and a rank by LDS atomics.
constexpr int CL_BLOCK = 1024;

__device__ inline int bucket_of(double t, double t_lo, double t_h, int nb) {
  const double v = (t - t_lo) / t_h;
  const int b = v >= 0.0 ? (v < (double)nb ? (int)v : nb - 1) : 0;  // floor, clamped to [0, nb); NaN to 0
  return b;
}

__global__ __launch_bounds__(256) void screen_fill_kernel(int* __restrict__ order, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) order[i] = -1;
}

__global__ __launch_bounds__(CL_BLOCK) void screen_hist_kernel(const double* __restrict__ ld, int64_t n, double t_lo,
                                                              double t_h, int nb, int* __restrict__ blk) {
  __shared__ int hist[SO_NB_MAX];
  for (int b = threadIdx.x; b < SO_NB_MAX; b += CL_BLOCK) hist[b] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * CL_BLOCK + threadIdx.x;
  if (i < n) atomicAdd(&hist[bucket_of(gmem(ld)[i], t_lo, t_h, nb)], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < SO_NB_MAX; b += CL_BLOCK) blk[(int64_t)blockIdx.x * SO_NB_MAX + b] = hist[b];
}

// bucket b's blocks: exclusive prefix of their counts over the blocks (in place), and the bucket's count; one
// 256-thread workgroup per bucket, 256 blocks per pass (wave scans by lane shuffles, then the four wave totals)
__global__ __launch_bounds__(256) void screen_scan_kernel(int* __restrict__ blk, int64_t nblk,
                                                          int* __restrict__ bcount) {
  __shared__ int wsum[4];
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int carry = 0;
  for (int64_t k0 = 0; k0 < nblk; k0 += 256) {
    const int64_t k = k0 + threadIdx.x;
    const int c = k < nblk ? blk[k * SO_NB_MAX + b] : 0;
    int x = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int y = __shfl_up(x, d, 64);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = carry;
    for (int v = 0; v < w; ++v) before += wsum[v];
    if (k < nblk) blk[k * SO_NB_MAX + b] = before + x - c;
    carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();  // wsum is rewritten by the next pass
  }
  if (threadIdx.x == 0) bcount[b] = carry;
}

// the buckets' starts (each padded to SO_PAD slots), the slot count and the bucket parameters
__global__ __launch_bounds__(128) void screen_starts_kernel(int nb, double t_lo, double t_h,
                                                            const int* __restrict__ bcount, int* __restrict__ head,
                                                            double* __restrict__ tb, int* __restrict__ bstart) {
  __shared__ int padded[SO_NB_MAX];
  const int b = threadIdx.x;
  if (b < SO_NB_MAX) padded[b] = b < nb ? (bcount[b] + SO_PAD - 1) / SO_PAD * SO_PAD : 0;
  __syncthreads();
  if (b == 0) {
    int st = 0;
    for (int k = 0; k < SO_NB_MAX; ++k) {
      bstart[k] = st;
      st += padded[k];
    }
    bstart[SO_NB_MAX] = st;
    head[0] = st;
    head[1] = nb;
    tb[0] = t_lo;
    tb[1] = t_h;
  }
}

// each column's slot: its bucket's start + its block's offset in the bucket + its rank in the block (LDS atomics);
// and the bucket of every SO_PAD-slot group (-1 past the slots)
__global__ __launch_bounds__(CL_BLOCK) void screen_scatter_kernel(const double* __restrict__ ld, int64_t n,
                                                                 double t_lo, double t_h, int nb,
                                                                 const int* __restrict__ blk,
                                                                 const int* __restrict__ bstart,
                                                                 int* __restrict__ order, int* __restrict__ pos,
                                                                 int* __restrict__ gbucket, int64_t n_groups) {
  __shared__ int rank[SO_NB_MAX];
  __shared__ int st[SO_NB_MAX + 1];
  for (int b = threadIdx.x; b <= SO_NB_MAX; b += CL_BLOCK) {
    if (b < SO_NB_MAX) rank[b] = 0;
    st[b] = gmem(bstart)[b];
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * CL_BLOCK + threadIdx.x;
  if (i < n) {
    const int b = bucket_of(gmem(ld)[i], t_lo, t_h, nb);
    const int slot = st[b] + gmem(blk)[(int64_t)blockIdx.x * SO_NB_MAX + b] + atomicAdd(&rank[b], 1);
    order[slot] = (int)i;
    pos[i] = slot;
  }
  if (i < n_groups) {  // the launch covers the n columns and the n_groups groups
    const int s0 = (int)(i * SO_PAD);
    int g = -1;
    if (s0 < st[nb]) {
      int lo = 0, hi = nb - 1;  // the last bucket b with st[b] <= s0 (buckets may be empty: st repeats)
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (st[mid] <= s0) lo = mid; else hi = mid - 1;
      }
      g = lo;
    }
    gbucket[i] = g;
  }
}

hipError_t launch_screen_classify(const double* log_delta, int64_t n, double t_lo, double t_h, int nb,
                                  void* order_ws, hipStream_t stream) {
  const ScreenOrder o = screen_order(order_ws, n);
  const int64_t nblk = screen_order_nblk(n);
  if (nblk == 0) return hipSuccess;
  const int64_t ns = screen_order_slots(n);
  const int64_t fb = (ns + 255) / 256 < 4096 ? (ns + 255) / 256 : 4096;
  hipLaunchKernelGGL(screen_fill_kernel, dim3((unsigned)fb), dim3(256), 0, stream, o.order, ns);
  hipLaunchKernelGGL(screen_hist_kernel, dim3((unsigned)nblk), dim3(CL_BLOCK), 0, stream, log_delta, n, t_lo, t_h, nb,
                     o.blk);
  hipLaunchKernelGGL(screen_scan_kernel, dim3((unsigned)nb), dim3(256), 0, stream, o.blk, nblk, o.bcount);
  hipLaunchKernelGGL(screen_starts_kernel, dim3(1), dim3(128), 0, stream, nb, t_lo, t_h, o.bcount, o.head, o.tb,
                     o.bstart);
  const int64_t n_groups = ns / SO_PAD;
  const int64_t sblk = (n_groups + CL_BLOCK - 1) / CL_BLOCK > nblk ? (n_groups + CL_BLOCK - 1) / CL_BLOCK : nblk;
  hipLaunchKernelGGL(screen_scatter_kernel, dim3((unsigned)sblk), dim3(CL_BLOCK), 0, stream, log_delta, n, t_lo, t_h,
                     nb, o.blk, o.bstart, o.order, o.pos, o.gbucket, n_groups);
  return hipGetLastError();
}

// The node sums: wnode[k][j] = 1 / (s_k + exp(t_lo + j t_h / 2)) for k < neff and j <= 2 nb (0 elsewhere), then
// dnode[i][j] = sum_k us[i][k]^2 wnode[k][j] in FP64 (rows past mc give 0): a 256-thread block owns 32 rows x 64
// nodes, thread (g = tid / 32, c = tid % 32) accumulates rows 4 g .. 4 g + 3 at nodes c and c + 32, from K chunks of
// 32 staged in LDS (the squares of us in a padded 32 x 33 tile).  Node blocks past 2 nb exit.
__global__ __launch_bounds__(256) void screen_node_weights_kernel(const double* __restrict__ s, int64_t neff,
                                                                  const int* __restrict__ head,
                                                                  const double* __restrict__ tb,
                                                                  double* __restrict__ wnode) {
  const int64_t k = blockIdx.x;
  const int j = threadIdx.x;
  const int nb = gmem(head)[1];
  double w = 0.0;
  if (k < neff && j <= 2 * nb) w = 1.0 / (gmem(s)[k] + exp(gmem(tb)[0] + j * 0.5 * gmem(tb)[1]));
  wnode[k * SO_NODES + j] = w;
}

__global__ __launch_bounds__(256) void screen_nodes_kernel(const double* __restrict__ us, int64_t ld_us, int64_t mc,
                                                           int64_t kpad, const int* __restrict__ head,
                                                           const double* __restrict__ wnode,
                                                           double* __restrict__ dnode) {
  __shared__ double ws[32][64];
  __shared__ double xs[32][33];
  const int tid = threadIdx.x, g = tid >> 5, c = tid & 31;
  const int64_t i0 = (int64_t)blockIdx.x * 32;
  const int r0 = blockIdx.y * 64;
  if (r0 > 2 * gmem(head)[1]) return;  // block-uniform
  double acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = 0.0;
  for (int64_t k0 = 0; k0 < kpad; k0 += 32) {
#pragma unroll
    for (int e = tid; e < 32 * 64; e += 256) ws[e >> 6][e & 63] = wnode[(k0 + (e >> 6)) * SO_NODES + r0 + (e & 63)];
#pragma unroll
    for (int e = tid; e < 32 * 32; e += 256) {
      const int64_t i = i0 + (e >> 5);
      const double x = i < mc ? gmem(us)[i * ld_us + k0 + (e & 31)] : 0.0;
      xs[e >> 5][e & 31] = x * x;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 32; ++k) {
      const double a0 = ws[k][c], a1 = ws[k][c + 32];
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
    dnode[(i0 + 4 * g + i) * SO_NODES + r0 + c] = acc[i][0];
    dnode[(i0 + 4 * g + i) * SO_NODES + r0 + c + 32] = acc[i][1];
  }
}

hipError_t launch_screen_nodes(const double* us, int64_t ld_us, int64_t mc, const double* s, int64_t neff,
                               int64_t kpad, const void* order_ws, double* wnode, double* dnode, hipStream_t stream) {
  if (mc == 0) return hipSuccess;
  const ScreenOrder o = screen_order((void*)order_ws, 0);  // head and tb do not depend on the column count
  hipLaunchKernelGGL(screen_node_weights_kernel, dim3((unsigned)kpad), dim3(SO_NODES), 0, stream, s, neff, o.head,
                     o.tb, wnode);
  const dim3 grid((unsigned)((mc + 31) / 32), SO_NODES / 64);
  hipLaunchKernelGGL(screen_nodes_kernel, grid, dim3(256), 0, stream, us, ld_us, mc, kpad, o.head, wnode, dnode);
  return hipGetLastError();
}

hipError_t launch_screen_bucketed(const BucketArgs& g, int64_t Mc, int64_t P, uint64_t* rec, int64_t rec_ld,
                                  hipStream_t stream) {
  const int64_t n_mtiles = (Mc + W_TM - 1) / W_TM;
  // every possible slot tile
  const int64_t n_tiles = (n_mtiles + SB_WM - 1) / SB_WM * ((screen_order_slots(P) + W_TP - 1) / W_TP);
  hipLaunchKernelGGL(screen_bucketed_kernel<SB_WM>, dim3((unsigned)n_tiles), dim3(256 * SB_WM), 0, stream, g,
                     n_mtiles, rec, rec_ld);
  return hipGetLastError();
}

__device__ inline uint64_t readlane64(uint64_t v, int j) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), j);
  return ((uint64_t)hi << 32) | lo;
}

// Resolution.  Thread i (screened column i, global column p_first + i) reads the column's n_mt tile records and
// marks the tiles to re-evaluate (bit t of mask1: the tile's best row; of mask2: all its rows).  cL and cU are
// increasing in x, so L comes from the largest best x and the marks from one threshold X on x (cU(x) >= L iff
// x >= X), with no square root per tile.  Then each wave walks its lanes that marked any: the column's w_k y_k and
// w_k are formed once (held in registers when neff <= 1,024), and each marked test is evaluated with all 64 lanes
// (k strided) in FP64 as the re-scoring kernel does; the column's max_F is the maximum of the running value and
// those evaluations.
constexpr int RS_KREG = 16;  // k per lane held in registers: neff <= 64 * RS_KREG

__global__ __launch_bounds__(256) void screen_resolve_kernel(
    const uint64_t* __restrict__ rec, int64_t rec_ld, int64_t n_mt, int64_t Mc, int64_t P, int64_t p_first,
    const float* __restrict__ q1s, const float* __restrict__ q2s, const double* __restrict__ rWr,
    const double* __restrict__ log_delta, const double* __restrict__ US, int64_t ld_us, const float* __restrict__ uy,
    int64_t p_pad, const double* __restrict__ s, int64_t neff, double s_min, double s_max, double nu, double kA,
    double kArel, double kD, double* __restrict__ max_F, unsigned long long* __restrict__ n_eval) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63;
  uint64_t mask1 = 0, mask2 = 0;
  if (i < P) {
    const int64_t p = p_first + i;
    const double q1 = gmem(q1s)[p], q2 = gmem(q2s)[p], r = gmem(rWr)[p];
    const double cc = q2 / (q1 * q1 * r);  // c^2 = x cc
    const double delta = exp(gmem(log_delta)[p]);
    const double ka = kA + kArel * sqrt((s_max + delta) / (fmax(s_min, 0.0) + delta));
    const double up = sqrt(1.0 + kD), dn = sqrt(1.0 - kD), ur = ka * sqrt((1.0 + kD) / (1.0 - kD));
    const double F0 = gmem(max_F)[p];
    // record: low word m1 = (bits(x) & ~127) | row of the tile's best row; high word (m2 & ~127) | forced
    // (an x that overflowed FP32 counts as forced)
    auto forced = [](uint64_t v) { return ((v >> 32) & 1u) || ((uint32_t)v & ~127u) >= 0x7f800000u; };
    float xm = 0.f;
#pragma unroll 8
    for (int64_t t = 0; t < n_mt; ++t) {  // (unrolled: eight record loads in flight per thread)
      const uint64_t v = gmem(rec)[t * rec_ld + i];
      if (!forced(v)) xm = fmaxf(xm, __uint_as_float((uint32_t)v & ~127u));  // a lower bound of its x
    }
    // L: the largest lower bound, c of the running maximum (F = nu c^2 / (1 - c^2)) or the best tile's cL
    const double L = fmax(F0 > 0.0 ? sqrt(F0 / (nu + F0)) : 0.0, sqrt((double)xm * cc) * dn - ka);
    const double th = (L - ur) / up;  // cU(x) >= L  <=>  sqrt(x cc) >= th
    const double X = th > 0.0 ? th * th / cc : -1.0;
#pragma unroll 8
    for (int64_t t = 0; t < n_mt; ++t) {
      const uint64_t v = gmem(rec)[t * rec_ld + i];
      const uint32_t hw = (uint32_t)(v >> 32);
      // upper bounds: the best row's x and every other row's x (values rounded down by < 128 units in the last
      // place before packing)
      const float x1 = __uint_as_float((uint32_t)v | 127u), x2 = __uint_as_float(hw | 127u);
      if (forced(v) || (double)x2 >= X)
        mask2 |= 1ull << t;
      else if ((double)x1 >= X)
        mask1 |= 1ull << t;
    }
  }
  uint64_t pend = __ballot((mask1 | mask2) != 0);
  const int64_t wave_col0 = i - lane;
  const bool regs = neff <= 64 * RS_KREG;
  unsigned long long cnt = 0;  // this wave's evaluations (one counter update per wave)
  while (pend) {
    const int j = __builtin_ctzll(pend);
    pend &= pend - 1;
    const int64_t ij = wave_col0 + j;
    const int64_t p = p_first + ij;
    const uint64_t k1 = readlane64(mask1, j), k2 = readlane64(mask2, j);
    const double delta = exp(gmem(log_delta)[p]);
    auto y_at = [&](int64_t k) { return (double)gmem(uy)[((k >> 5) * p_pad + p) * 32 + (k & 31)]; };
    // the column's factors: w y and w per k (registers), and rWr = sum w y^2
    double wy[RS_KREG], ww[RS_KREG], rr = 0.0;
#pragma unroll
    for (int e = 0; e < RS_KREG; ++e) {
      const int64_t k = lane + 64 * e;
      wy[e] = 0.0;
      ww[e] = 0.0;
      if (regs && k < neff) {
        const double w = rcp_pos(gmem(s)[k] + delta), y = y_at(k);
        wy[e] = w * y;
        ww[e] = w;
        rr += wy[e] * y;
      }
    }
    if (!regs) {
      for (int64_t k = lane; k < neff; k += 64) {
        const double w = rcp_pos(gmem(s)[k] + delta), y = y_at(k);
        rr += w * y * y;
      }
    }
    rr = fmax(wave_sum(rr), FLOOR64);
    double fmx = -__builtin_inf();
    auto eval = [&](int64_t m) {
      const double* u = US + m * ld_us;
      double a = 0.0, d = 0.0;
      if (regs) {
#pragma unroll
        for (int e = 0; e < RS_KREG; ++e) {
          const int64_t k = lane + 64 * e;
          const double uk = k < neff ? gmem(u)[k] : 0.0;
          a += uk * wy[e];
          d += uk * uk * ww[e];
        }
      } else {
        for (int64_t k = lane; k < neff; k += 64) {
          const double w = rcp_pos(gmem(s)[k] + delta), uk = gmem(u)[k];
          a += uk * w * y_at(k);
          d += uk * uk * w;
        }
      }
      a = wave_sum(a);
      d = fmax(wave_sum(d), FLOOR64);
      const double b = a / d;
      const double rss = fmax(rr - a * a / d, FLOOR64);
      fmx = fmax(fmx, b * b / (rss / nu / d));
      ++cnt;
    };
    for (uint64_t b = k1; b; b &= b - 1) {
      const int64_t t = __builtin_ctzll(b);
      // a padding row can hold the packed maximum only where every row of the tile has x = 0; it is no test
      const int64_t m = t * W_TM + (int64_t)(gmem(rec)[t * rec_ld + ij] & 127u);
      if (m < Mc) eval(m);
    }
    for (uint64_t b = k2; b; b &= b - 1) {
      const int64_t t = __builtin_ctzll(b);
      const int64_t hi = (t + 1) * W_TM < Mc ? (t + 1) * W_TM : Mc;
      for (int64_t m = t * W_TM; m < hi; ++m) eval(m);
    }
    if (lane == 0) gmem(max_F)[p] = fmax(gmem(max_F)[p], fmx);
  }
  if (n_eval && lane == 0 && cnt) atomicAdd(n_eval, cnt);
}

hipError_t launch_wald_screen(const ScreenArgs& g, int64_t Mc, int64_t P, float dmin, uint64_t* rec, int64_t rec_ld,
                              hipStream_t stream) {
  const int64_t n_mtiles = (Mc + W_TM - 1) / W_TM;
  const int64_t n_ptiles = (P + W_TP - 1) / W_TP;
  const int64_t n_tiles = n_mtiles * n_ptiles;
  hipLaunchKernelGGL(wald_screen_kernel, dim3((unsigned)n_tiles), dim3(256), 0, stream, g, n_mtiles, n_ptiles,
                     n_tiles, Mc, P, dmin, rec, rec_ld);
  return hipGetLastError();
}

hipError_t launch_screen_resolve(const uint64_t* rec, int64_t rec_ld, int64_t Mc, int64_t P, int64_t p_first,
                                 const float* q1s, const float* q2s, const double* rWr, const double* log_delta,
                                 const double* US, int64_t ld_us, const float* uy, int64_t p_pad, const double* s,
                                 int64_t neff, double s_min, double s_max, double nu, double kA, double kArel,
                                 double kD, double* max_F, unsigned long long* n_eval, hipStream_t stream) {
  const int64_t n_mt = (Mc + W_TM - 1) / W_TM;
  hipLaunchKernelGGL(screen_resolve_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, stream, rec, rec_ld,
                     n_mt, Mc, P, p_first, q1s, q2s, rWr, log_delta, US, ld_us, uy, p_pad, s, neff, s_min, s_max, nu,
                     kA, kArel, kD, max_F, n_eval);
  return hipGetLastError();
}

}  // namespace flmm
