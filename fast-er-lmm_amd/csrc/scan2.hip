// Wald scan, variant "wide" (128 SNP rows x 256 columns per workgroup): one launch over one SNP block, the
// tile body in scan_wide.h.
#include "scan_wide.h"

namespace flmm {

template <int PREC>
__global__ __launch_bounds__(256, 1) void wald_scan_wide_kernel(
    TileArgs<PREC, 2> args, ScanScales sc, int64_t n_mtiles, int64_t n_ptiles, int64_t n_tiles, int64_t Mc, int64_t P,
    int64_t n_real, const double* __restrict__ rWr, const double* __restrict__ h2, double nu,
    const int64_t* __restrict__ row_index, ScanOut out, double* __restrict__ pmax, int64_t pmax_ld) {
  wide_tile<PREC>(args, sc, n_mtiles, n_ptiles, xcd_remap(blockIdx.x, n_tiles), Mc, P, n_real, rWr, h2, nu, row_index,
                  out, pmax, pmax_ld);
}

template <int PREC>
hipError_t launch_wald_scan_wide(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P, int64_t n_real,
                                 const double* rWr, const double* h2, double nu, const int64_t* row_index,
                                 const ScanOut& out, double* pmax, hipStream_t stream) {
  const int64_t n_mtiles = (Mc + W_TM - 1) / W_TM;
  const int64_t n_ptiles = (P + W_TP - 1) / W_TP;
  const int64_t n_tiles = n_mtiles * n_ptiles;
  hipLaunchKernelGGL(wald_scan_wide_kernel<PREC>, dim3((unsigned)n_tiles), dim3(256), 0, stream, args, sc, n_mtiles,
                     n_ptiles, n_tiles, Mc, P, n_real, rWr, h2, nu, row_index, out, pmax, args.b_rows_pad);
  return hipGetLastError();
}

// The wide tile with the second product limited to the first args.nchunks2 K-chunks (factored-weight compat scan,
// low-rank scan); KMODE: K_PAIRED for first-product K up to 8,192 past nchunks2, K_FLUSHED beyond.
template <int PREC, int KMODE>
__global__ __launch_bounds__(256, 1) void wald_scan_lowrank_kernel(
    TileArgs<PREC, 2> args, ScanScales sc, int64_t n_mtiles, int64_t n_ptiles, int64_t n_tiles, int64_t Mc, int64_t P,
    int64_t n_real, const double* __restrict__ rWr, const double* __restrict__ h2, double nu,
    const int64_t* __restrict__ row_index, ScanOut out, double* __restrict__ pmax, int64_t pmax_ld) {
  wide_tile<PREC, KMODE>(args, sc, n_mtiles, n_ptiles, xcd_remap(blockIdx.x, n_tiles), Mc, P, n_real, rWr, h2, nu,
                         row_index, out, pmax, pmax_ld);
}

template <int PREC>
hipError_t launch_wald_scan_lowrank(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P,
                                    int64_t n_real, const double* rWr, const double* h2, double nu,
                                    const int64_t* row_index, const ScanOut& out, double* pmax, hipStream_t stream) {
  const int64_t n_mtiles = (Mc + W_TM - 1) / W_TM;
  const int64_t n_ptiles = (P + W_TP - 1) / W_TP;
  const int64_t n_tiles = n_mtiles * n_ptiles;
  if (args.nchunks - args.nchunks2 <= 256)
    hipLaunchKernelGGL((wald_scan_lowrank_kernel<PREC, K_PAIRED>), dim3((unsigned)n_tiles), dim3(256), 0, stream, args,
                       sc, n_mtiles, n_ptiles, n_tiles, Mc, P, n_real, rWr, h2, nu, row_index, out, pmax,
                       args.b_rows_pad);
  else
    hipLaunchKernelGGL((wald_scan_lowrank_kernel<PREC, K_FLUSHED>), dim3((unsigned)n_tiles), dim3(256), 0, stream,
                       args, sc, n_mtiles, n_ptiles, n_tiles, Mc, P, n_real, rWr, h2, nu, row_index, out, pmax,
                       args.b_rows_pad);
  return hipGetLastError();
}

#define FLMM_LR_INST(PR)                                                                                            \
  template hipError_t launch_wald_scan_lowrank<PR>(const TileArgs<PR, 2>&, const ScanScales&, int64_t, int64_t,     \
                                                   int64_t, const double*, const double*, double, const int64_t*,   \
                                                   const ScanOut&, double*, hipStream_t);
FLMM_LR_INST(PREC_SPLIT_FP16)
FLMM_LR_INST(PREC_SPLIT_BF16)
FLMM_LR_INST(PREC_FP32)
#undef FLMM_LR_INST

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
