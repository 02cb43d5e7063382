// Grouped Wald scan: many independent (SNP block x column set) scans -- e.g. the phenotype groups of uneven
// retained N that a missingness-grouped run scans per chromosome (the reference loops over them:
// fasterlmm/cli.py:337-375 with one perm_threshold / loco_scan_compat each) -- in ONE persistent launch.
//
// The launch reads everything from a device table written by flmm_scan_table_fill: a header {n_groups,
// total_tiles, variant, 0}, the exclusive prefix sums of the groups' tile counts, then one flmm_scan_group per
// group.  Kernel arguments are only the table pointer, so a captured graph of the launch replays for a new set of
// groups (or a new chromosome) after the table is rewritten in place.  The grid is persistent (a multiple of 8
// workgroups); round r hands XCD x (blocks b with b % 8 == x) the contiguous tiles
// r G + x G / 8 + [0, G / 8), so the workgroups an XCD runs together share L2, and each tile is found by binary
// search over the prefix sums and run by the square or wide tile body, or (FLMM_SCAN_WIDE_FACTORED: factored
// weights, every group with kpad2 > 0) the wide two-phase body with its second product over kpad2.
#include "../../include/fasterlmm_amd.h"
#include "scan_square.h"
#include "scan_wide.h"

namespace flmm {

constexpr int TABLE_HDR = 4;  // int64 words before the prefix sums

__device__ inline const flmm_scan_group* table_groups(const int64_t* hdr) {
  return (const flmm_scan_group*)(hdr + TABLE_HDR + hdr[0] + 1);
}

template <int PREC>
__device__ inline void job_args(const flmm_scan_group& j, TileArgs<PREC, 2>& t, ScanScales& sc, ScanOut& out) {
  t.a[0][0] = j.s1_hi; t.a[0][1] = j.s1_lo; t.a[1][0] = j.s2_hi; t.a[1][1] = j.s2_lo;
  t.b[0][0] = j.q1_hi; t.b[0][1] = j.q1_lo; t.b[1][0] = j.q2_hi; t.b[1][1] = j.q2_lo;
  t.a_rows_pad = j.mc_pad;
  t.b_rows_pad = j.p_pad;
  t.nchunks = (int)(j.kpad / KCHUNK);
  t.nchunks2 = (int)(j.kpad2 / KCHUNK);
  sc = PREC == PREC_SPLIT_FP16 ? ScanScales{j.s1_scale, j.s2_scale, j.q1_scale, j.q2_scale}
                               : ScanScales{nullptr, nullptr, nullptr, nullptr};
  out = ScanOut{j.f, j.beta, j.se, j.sfve, j.nullh2, j.n_real};
}

template <int PREC, int VARIANT>
__global__ __launch_bounds__(256, 1) void wald_scan_grouped_kernel(const int64_t* __restrict__ hdr) {
  const int64_t n_groups = hdr[0], total = hdr[1];
  const int64_t* start = hdr + TABLE_HDR;
  const flmm_scan_group* groups = table_groups(hdr);
  const int64_t G = gridDim.x;
  const int64_t slot = (int64_t)(blockIdx.x % 8) * (G / 8) + blockIdx.x / 8;
  for (int64_t base = 0; base < total; base += G) {
    const int64_t t = base + slot;
    if (t < total) {
      int64_t lo = 0, hi = n_groups;  // start[lo] <= t < start[hi]
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (start[mid] <= t) lo = mid; else hi = mid;
      }
      const flmm_scan_group& j = groups[lo];
      TileArgs<PREC, 2> args;
      ScanScales sc;
      ScanOut out;
      job_args<PREC>(j, args, sc, out);
      const int64_t n_mtiles = (j.mc + TILE - 1) / TILE;
      if constexpr (VARIANT == FLMM_SCAN_SQUARE)
        square_tile<PREC>(args, sc, n_mtiles, t - start[lo], j.mc, j.P, j.n_real, j.rWr, j.h2, j.nu, j.row_index,
                          out, j.pmax, j.p_pad);
      else if constexpr (VARIANT == FLMM_SCAN_WIDE_FACTORED)  // second product over kpad2 (the low-rank tile)
        wide_tile<PREC, K_PAIRED>(args, sc, n_mtiles, (j.P + W_TP - 1) / W_TP, t - start[lo], j.mc, j.P, j.n_real,
                                  j.rWr, j.h2, j.nu, j.row_index, out, j.pmax, j.p_pad);
      else
        wide_tile<PREC>(args, sc, n_mtiles, (j.P + W_TP - 1) / W_TP, t - start[lo], j.mc, j.P, j.n_real, j.rWr, j.h2, j.nu,
                        j.row_index, out, j.pmax, j.p_pad);
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

// max_F[p] = max(max_F[p], max over the group's SNP tiles of pmax[t][p]), every group of the table
__global__ void max_reduce_grouped_kernel(const int64_t* __restrict__ hdr) {
  const int64_t g = blockIdx.y;
  if (g >= hdr[0]) return;
  const flmm_scan_group& j = table_groups(hdr)[g];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= j.P) return;
  const int64_t n_mt = (j.mc + TILE - 1) / TILE;
  double m = j.max_F[p];
  for (int64_t t = 0; t < n_mt; ++t) m = fmax(m, j.pmax[t * j.p_pad + p]);
  j.max_F[p] = m;
}

hipError_t launch_wald_scan_grouped(int prec, int variant, const int64_t* table, int blocks, hipStream_t stream) {
#define FLMM_GRP(PR)                                                                                                 \
  do {                                                                                                               \
    if (variant == FLMM_SCAN_SQUARE) {                                                                               \
      constexpr int lds = tile_lds_bytes<PR, 2>();                                                                   \
      hipLaunchKernelGGL((wald_scan_grouped_kernel<PR, FLMM_SCAN_SQUARE>), dim3(blocks), dim3(256), lds, stream,     \
                         table);                                                                                     \
    } else if (variant == FLMM_SCAN_WIDE) {                                                                          \
      hipLaunchKernelGGL((wald_scan_grouped_kernel<PR, FLMM_SCAN_WIDE>), dim3(blocks), dim3(256), 0, stream, table); \
    } else {                                                                                                         \
      hipLaunchKernelGGL((wald_scan_grouped_kernel<PR, FLMM_SCAN_WIDE_FACTORED>), dim3(blocks), dim3(256), 0, stream, \
                         table);                                                                                     \
    }                                                                                                                \
  } while (0)
  if (prec == PREC_SPLIT_BF16) FLMM_GRP(PREC_SPLIT_BF16);
  else if (prec == PREC_SPLIT_FP16) FLMM_GRP(PREC_SPLIT_FP16);
  else FLMM_GRP(PREC_FP32);
#undef FLMM_GRP
  return hipGetLastError();
}

hipError_t launch_max_reduce_grouped(const int64_t* table, int n_groups, int64_t max_P, hipStream_t stream) {
  const dim3 grid((unsigned)((max_P + 255) / 256), (unsigned)n_groups);
  hipLaunchKernelGGL(max_reduce_grouped_kernel, grid, dim3(256), 0, stream, table);
  return hipGetLastError();
}

}  // namespace flmm
