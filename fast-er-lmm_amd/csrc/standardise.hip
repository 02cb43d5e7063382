// Column standardisation of a genotype matrix (reference fasterlmm/io.py:223-254, the pysnptools Unit
// standardiser): per column, NaN-aware mean and population standard deviation (ddof 0), (x - mean) / std,
// std == 0 -> the column becomes 0, and NaN -> 0 after standardising.
//
// Layout: src and dst are (n, m) row-major FP64 with leading dimension ld (the reference's Z (N, M)); one thread
// owns one column and walks the rows, so each row step is a coalesced read of consecutive columns.  Three
// passes (mean, centred sum of squares, write) keep the reference's two-pass variance arithmetic.  Int8 input
// (0/1/2 dosages, negative = missing) is standardised the same way.
#include "kernels.h"

namespace flmm {

template <typename T>
__device__ inline bool is_missing(T x);
template <>
__device__ inline bool is_missing<double>(double x) { return isnan(x); }
template <>
__device__ inline bool is_missing<int8_t>(int8_t x) { return x < 0; }

template <typename T>
__global__ __launch_bounds__(256) void standardise_kernel(const T* __restrict__ src, int64_t n, int64_t m,
                                                          int64_t lds, double* __restrict__ dst, int64_t ldd,
                                                          double* __restrict__ mean_out, double* __restrict__ std_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double sum = 0.0;
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    const T x = src[i * lds + j];
    if (!is_missing<T>(x)) { sum += (double)x; ++cnt; }
  }
  const double mean = cnt > 0 ? sum / (double)cnt : NAN;
  double ss = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const T x = src[i * lds + j];
    if (!is_missing<T>(x)) { const double d = (double)x - mean; ss += d * d; }
  }
  double sd = cnt > 0 ? sqrt(ss / (double)cnt) : NAN;
  if (sd == 0.0) sd = INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    const T x = src[i * lds + j];
    dst[i * ldd + j] = is_missing<T>(x) ? 0.0 : ((double)x - mean) / sd;
  }
  if (mean_out) mean_out[j] = mean;
  if (std_out) std_out[j] = sd;
}

hipError_t launch_standardise(int dtype, const void* src, int64_t n, int64_t m, int64_t lds, double* dst, int64_t ldd,
                              double* mean_out, double* std_out, hipStream_t stream) {
  const dim3 grid((unsigned)((m + 255) / 256));
  if (dtype == 0)
    hipLaunchKernelGGL(standardise_kernel<double>, grid, dim3(256), 0, stream, (const double*)src, n, m, lds, dst, ldd,
                       mean_out, std_out);
  else
    hipLaunchKernelGGL(standardise_kernel<int8_t>, grid, dim3(256), 0, stream, (const int8_t*)src, n, m, lds, dst, ldd,
                       mean_out, std_out);
  return hipGetLastError();
}

}  // namespace flmm
