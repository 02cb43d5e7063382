// Wald scan, variant "square" (128 x 128 tiles, both operands staged through LDS): one launch over one SNP block,
// the tile body in scan_square.h; and the reduction of the per-tile maxima.
#include "scan_square.h"

namespace flmm {

template <int PREC>
__global__ __launch_bounds__(NTHREADS, 1) void wald_scan_kernel(
    TileArgs<PREC, 2> args, ScanScales sc, int64_t n_mtiles, int64_t n_tiles, int64_t Mc, int64_t P, int64_t n_real,
    const double* __restrict__ rWr, const double* __restrict__ h2, double nu, const int64_t* __restrict__ row_index,
    ScanOut out, double* __restrict__ pmax, int64_t pmax_ld) {
  square_tile<PREC>(args, sc, n_mtiles, xcd_remap(blockIdx.x, n_tiles), Mc, P, n_real, rWr, h2, nu, row_index, out,
                    pmax, pmax_ld);
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
