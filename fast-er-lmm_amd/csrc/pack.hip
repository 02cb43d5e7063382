// FP64 matrix -> MFMA operand planes (bf16 / fp16 hi-lo pair, or fp32), with optional row gather; row norms.
//
// Source element (row r, k) is src[src_row(r) * row_stride + k * k_stride], src_row(r) = row_index ? row_index[r]
// : r, so the same kernel packs row-major SNP blocks, the transposed view of an (N, P) phenotype matrix and the
// kept eigenvector columns of U.  Output rows [n_rows, rows_pad) and k in [n_k, kpad) are written as zeros.
// A 64 x 64 tile goes through LDS so that both the read (along whichever source stride is 1) and the plane
// write (32 consecutive k per row) are coalesced.  Split-FP16 rows are multiplied by the power-of-two scale of
// their L2 norm (row_norm, from flmm_row_norms), which bounds every entry, and the scale is recorded.
#include "kernels.h"

namespace flmm {

template <int PREC>
__global__ __launch_bounds__(256) void pack_planes_kernel(const double* __restrict__ src, int64_t n_rows, int64_t n_k,
                                                          int64_t row_stride, int64_t k_stride,
                                                          const int64_t* __restrict__ row_index,
                                                          const double* __restrict__ row_norm, void* out_a,
                                                          void* out_b, float* __restrict__ row_scale,
                                                          int64_t rows_pad, int64_t kpad) {
  __shared__ double tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const int tid = threadIdx.x;
  const bool kfast = (k_stride == 1);
#pragma unroll 4
  for (int e = tid; e < 64 * 64; e += 256) {
    const int rr = kfast ? (e >> 6) : (e & 63);
    const int kk = kfast ? (e & 63) : (e >> 6);
    const int64_t r = r0 + rr, k = k0 + kk;
    double v = 0.0;
    if (r < n_rows && k < n_k) {
      const int64_t sr = row_index ? row_index[r] : r;
      v = src[sr * row_stride + k * k_stride];
    }
    tile[rr][kk] = v;
  }
  __syncthreads();
  // thread -> (row, 16 consecutive k)
  const int rr = tid >> 2;
  const int kq = (tid & 3) * 16;
  const int64_t r = r0 + rr;
  if (r >= rows_pad) return;
  double sc = 1.0;
  if constexpr (PREC == PREC_SPLIT_FP16) {
    sc = (r < n_rows) ? pow2_scale(row_norm[r]) : 1.0;
    if (blockIdx.y == 0 && (tid & 3) == 0) row_scale[r] = (float)sc;
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int64_t k = k0 + kq + e;
    if (k >= kpad) break;
    const int64_t off = ((k >> 5) * rows_pad + r) * 32 + (k & 31);
    const double x = tile[rr][kq + e];
    if constexpr (is_split(PREC)) {
      u16 h, l;
      split2<PREC>(x * sc, h, l);
      ((u16*)out_a)[off] = h;
      ((u16*)out_b)[off] = l;
    } else {
      ((float*)out_a)[off] = (float)x;
    }
  }
}

// norm[r] = sqrt(sum_k src(r, k)^2).  Rows along the unit stride (the transposed view of an (N, P) matrix):
// a 1,024-thread block owns 64 consecutive rows, lane = row, and its 16 waves split k (wave w sums k = w, w + 16,
// ...), so every load is 64 consecutive doubles and a few thousand rows against N ~ 10^4..10^5 strains still fill
// the chip; the 16 partial sums are added in wave order.  Otherwise a wave per row, lanes striding k.
__global__ __launch_bounds__(1024) void row_norms_strided_kernel(const double* __restrict__ src, int64_t n_rows,
                                                                 int64_t n_k, int64_t k_stride,
                                                                 const int64_t* __restrict__ row_index,
                                                                 double* __restrict__ norm) {
  __shared__ double part[16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * 64 + lane;
  double s = 0.0;
  if (r < n_rows) {
    const int64_t sr = row_index ? row_index[r] : r;
    for (int64_t k = w; k < n_k; k += 16) {
      const double x = src[sr + k * k_stride];
      s += x * x;
    }
  }
  part[w][lane] = s;
  __syncthreads();
  if (w == 0 && r < n_rows) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) t += part[q][lane];
    norm[r] = sqrt(t);
  }
}

__global__ __launch_bounds__(256) void row_norms_kernel(const double* __restrict__ src, int64_t n_rows, int64_t n_k,
                                                        int64_t row_stride, int64_t k_stride,
                                                        const int64_t* __restrict__ row_index,
                                                        double* __restrict__ norm) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n_rows) return;
  const int64_t sr = row_index ? row_index[r] : r;
  double s = 0.0;
  for (int64_t k = lane; k < n_k; k += 64) {
    const double x = src[sr * row_stride + k * k_stride];
    s += x * x;
  }
  s = wave_sum(s);
  if (lane == 0) norm[r] = sqrt(s);
}

hipError_t launch_pack_planes(int prec, const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride,
                              int64_t k_stride, const int64_t* row_index, const double* row_norm, void* out_a,
                              void* out_b, float* row_scale, int64_t rows_pad, int64_t kpad, hipStream_t stream) {
  const dim3 grid((unsigned)((rows_pad + 63) / 64), (unsigned)((kpad + 63) / 64));
#define FLMM_PACK(PR)                                                                                              \
  hipLaunchKernelGGL(pack_planes_kernel<PR>, grid, dim3(256), 0, stream, src, n_rows, n_k, row_stride, k_stride,    \
                     row_index, row_norm, out_a, out_b, row_scale, rows_pad, kpad)
  if (prec == PREC_SPLIT_BF16) FLMM_PACK(PREC_SPLIT_BF16);
  else if (prec == PREC_SPLIT_FP16) FLMM_PACK(PREC_SPLIT_FP16);
  else FLMM_PACK(PREC_FP32);
#undef FLMM_PACK
  return hipGetLastError();
}

hipError_t launch_row_norms(const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                            const int64_t* row_index, double* norm, hipStream_t stream) {
  if (row_stride == 1 && k_stride != 1)
    hipLaunchKernelGGL(row_norms_strided_kernel, dim3((unsigned)((n_rows + 63) / 64)), dim3(1024), 0, stream, src,
                       n_rows, n_k, k_stride, row_index, norm);
  else
    hipLaunchKernelGGL(row_norms_kernel, dim3((unsigned)((n_rows + 3) / 4)), dim3(256), 0, stream, src, n_rows, n_k,
                       row_stride, k_stride, row_index, norm);
  return hipGetLastError();
}

}  // namespace flmm
