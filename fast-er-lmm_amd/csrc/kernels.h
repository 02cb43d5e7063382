// Host-side launch wrappers of the gfx950 kernels (defined in pack.hip, rotate.hip, delta.hip, scan.hip).
#pragma once
#include "gemm_tile.h"

namespace flmm {

struct ScanOut {
  double* f;
  double* beta;
  double* se;
  double* sfve;
  double* nullh2;
  int64_t ld;  // = n_real
};

hipError_t launch_pack_planes(int prec, const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride,
                              int64_t k_stride, const int64_t* row_index, void* out_a, void* out_b, int64_t rows_pad,
                              int64_t kpad, hipStream_t stream);
template <int PREC>
hipError_t launch_rotate(int mode, const TileArgs<PREC, 1>& args, void* o0, void* o1, void* o2, void* o3,
                         hipStream_t stream);
hipError_t launch_delta_weights(const double* s, int64_t neff, const double* s_lam, int64_t nlam, int64_t kpad, int G,
                                double lo, double hi, float* wplane, float* lam, hipStream_t stream);
hipError_t launch_delta_grid(const float* wplane, const float* lam, int G, double nco, const float* uy, int64_t P,
                             int64_t p_pad, int64_t kpad, int* idx, hipStream_t stream);
hipError_t launch_delta_refine(int prec, int mode, const float* uy, int64_t P, int64_t p_pad, int64_t kpad,
                               const double* s, int64_t neff, const double* s_lam, int64_t nlam, double nco, int G,
                               double lo, double hi, const int* idx,
                               double* log_delta, double* rWr, double* h2, void* q1a, void* q1b, void* q2a,
                               void* q2b, hipStream_t stream);
template <int PREC>
hipError_t launch_wald_scan(const TileArgs<PREC, 2>& args, int64_t Mc, int64_t P, int64_t n_real,
                            const double* rWr, const double* h2, double nu, const int64_t* row_index,
                            const ScanOut& out, double* pmax, hipStream_t stream);
hipError_t launch_standardise(int dtype, const void* src, int64_t n, int64_t m, int64_t lds, double* dst, int64_t ldd,
                              double* mean_out, double* std_out, hipStream_t stream);
hipError_t launch_max_reduce(const double* pmax, int64_t n_mtiles, int64_t ld, int64_t P, double* max_F,
                             hipStream_t stream);

}  // namespace flmm

namespace flmm {
hipError_t launch_neglog10p(const double* F, int64_t n, double df2, double* out, hipStream_t stream);
hipError_t launch_effect_size(const double* beta, int64_t M, int64_t n_real, const double* g_var,
                              const double* p_var, double* out, hipStream_t stream);
}  // namespace flmm
