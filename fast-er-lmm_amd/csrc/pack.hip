// FP64 matrix -> MFMA operand planes (bf16 hi/lo pair or fp32), with optional row gather.
//
// Source element (row r, k) is src[src_row(r) * row_stride + k * k_stride], src_row(r) = row_index ? row_index[r]
// : r, so the same kernel packs row-major SNP blocks, the transposed view of an (N, P) phenotype matrix and the
// kept eigenvector columns of U.  Output rows [n_rows, rows_pad) and k in [n_k, kpad) are written as zeros.
// A 64 x 64 tile goes through LDS so that both the read (along whichever source stride is 1) and the plane
// write (32 consecutive k per row) are coalesced.
#include "kernels.h"

namespace flmm {

template <int PREC>
__global__ __launch_bounds__(256) void pack_planes_kernel(const double* __restrict__ src, int64_t n_rows, int64_t n_k,
                                                          int64_t row_stride, int64_t k_stride,
                                                          const int64_t* __restrict__ row_index, void* out_a,
                                                          void* out_b, int64_t rows_pad, int64_t kpad) {
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
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    const int64_t k = k0 + kq + e;
    if (k >= kpad) break;
    const int64_t off = ((k >> 5) * rows_pad + r) * 32 + (k & 31);
    const double x = tile[rr][kq + e];
    if constexpr (PREC == PREC_SPLIT_BF16) {
      u16 h, l;
      split_bf16(x, h, l);
      ((u16*)out_a)[off] = h;
      ((u16*)out_b)[off] = l;
    } else {
      ((float*)out_a)[off] = (float)x;
    }
  }
}

hipError_t launch_pack_planes(int prec, const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride,
                              int64_t k_stride, const int64_t* row_index, void* out_a, void* out_b, int64_t rows_pad,
                              int64_t kpad, hipStream_t stream) {
  const dim3 grid((unsigned)((rows_pad + 63) / 64), (unsigned)((kpad + 63) / 64));
  if (prec == PREC_SPLIT_BF16)
    hipLaunchKernelGGL(pack_planes_kernel<PREC_SPLIT_BF16>, grid, dim3(256), 0, stream, src, n_rows, n_k, row_stride,
                       k_stride, row_index, out_a, out_b, rows_pad, kpad);
  else
    hipLaunchKernelGGL(pack_planes_kernel<PREC_FP32>, grid, dim3(256), 0, stream, src, n_rows, n_k, row_stride,
                       k_stride, row_index, out_a, out_b, rows_pad, kpad);
  return hipGetLastError();
}

}  // namespace flmm
