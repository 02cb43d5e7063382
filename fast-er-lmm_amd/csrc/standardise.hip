// Column standardisation of a genotype matrix (reference fasterlmm/io.py:223-254, the pysnptools Unit
// standardiser): per column, NaN-aware mean and population standard deviation (ddof 0), (x - mean) / std,
// std == 0 -> the column becomes 0, and NaN -> 0 after standardising.
//
// Layout: src and dst are (n, m) row-major with leading dimensions lds / ldd (the reference's Z (N, M)).  A
// 1,024-thread block owns 32 adjacent columns; thread (r, c) = (threadIdx.x / 32, threadIdx.x % 32) walks the rows
// i = r, r + 32, ... of column c, so every row step of a wave is a coalesced read of 32 consecutive columns and a
// block keeps 32 rows per column in flight (a tall block of 16,384 columns x 50,000 rows runs 524k threads instead
// of 16k).  Three passes (mean, centred sum of squares, write) keep the reference's two-pass variance arithmetic;
// the 32 row-slice partials of each pass are summed in a fixed order (deterministic).  Int8 input (0/1/2 dosages,
// negative = missing) is standardised the same way.
#include "kernels.h"

namespace flmm {

template <typename T>
__device__ inline bool is_missing(T x);
template <>
__device__ inline bool is_missing<double>(double x) { return isnan(x); }
template <>
__device__ inline bool is_missing<int8_t>(int8_t x) { return x < 0; }

template <typename T>
__global__ __launch_bounds__(1024) void standardise_kernel(const T* __restrict__ src, int64_t n, int64_t m,
                                                           int64_t lds, double* __restrict__ dst, int64_t ldd,
                                                           double* __restrict__ mean_out,
                                                           double* __restrict__ std_out) {
  __shared__ double part[32][33];  // [row slice][column], padded against bank conflicts
  __shared__ double cntp[32][33];
  __shared__ double col_mean[32], col_sd[32];
  const int c = threadIdx.x & 31, r = threadIdx.x >> 5;
  const int64_t j = (int64_t)blockIdx.x * 32 + c;
  const bool live = j < m;
  double sum = 0.0, cnt = 0.0;
  if (live) {
#pragma unroll 4
    for (int64_t i = r; i < n; i += 32) {
      const T x = src[i * lds + j];
      if (!is_missing<T>(x)) { sum += (double)x; cnt += 1.0; }
    }
  }
  part[r][c] = sum;
  cntp[r][c] = cnt;
  __syncthreads();
  if (r == 0) {
    double s = 0.0, k = 0.0;
    for (int q = 0; q < 32; ++q) { s += part[q][c]; k += cntp[q][c]; }
    col_mean[c] = k > 0.0 ? s / k : NAN;
    cntp[0][c] = k;
  }
  __syncthreads();
  const double mean = col_mean[c];
  double ss = 0.0;
  if (live) {
#pragma unroll 4
    for (int64_t i = r; i < n; i += 32) {
      const T x = src[i * lds + j];
      if (!is_missing<T>(x)) { const double d = (double)x - mean; ss += d * d; }
    }
  }
  __syncthreads();  // every thread has read cntp[0][c] / col_mean before part is reused
  part[r][c] = ss;
  __syncthreads();
  if (r == 0) {
    double t = 0.0;
    for (int q = 0; q < 32; ++q) t += part[q][c];
    const double k = cntp[0][c];
    double sd = k > 0.0 ? sqrt(t / k) : NAN;
    if (sd == 0.0) sd = INFINITY;
    col_sd[c] = sd;
    if (live && mean_out) mean_out[j] = mean;
    if (live && std_out) std_out[j] = sd;
  }
  __syncthreads();
  const double sd = col_sd[c];
  if (live) {
#pragma unroll 4
    for (int64_t i = r; i < n; i += 32) {
      const T x = src[i * lds + j];
      dst[i * ldd + j] = is_missing<T>(x) ? 0.0 : ((double)x - mean) / sd;
    }
  }
}

hipError_t launch_standardise(int dtype, const void* src, int64_t n, int64_t m, int64_t lds, double* dst, int64_t ldd,
                              double* mean_out, double* std_out, hipStream_t stream) {
  const dim3 grid((unsigned)((m + 31) / 32));
  if (dtype == 0)
    hipLaunchKernelGGL(standardise_kernel<double>, grid, dim3(1024), 0, stream, (const double*)src, n, m, lds, dst,
                       ldd, mean_out, std_out);
  else
    hipLaunchKernelGGL(standardise_kernel<int8_t>, grid, dim3(1024), 0, stream, (const int8_t*)src, n, m, lds, dst,
                       ldd, mean_out, std_out);
  return hipGetLastError();
}

}  // namespace flmm
