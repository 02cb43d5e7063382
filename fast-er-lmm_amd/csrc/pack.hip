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

// Projected source (PROJ): element (r, k) is src(r, k) - sum_c px[k * pc + c] pb[c * n_rows + r], the residual of an
// (N, P) matrix's columns r on an (N, C <= PROJ_CMAX) matrix X (px row-major) with coefficients B = X^+ Y (pb: (C, P)
// row-major); each thread keeps its row's coefficients in registers (bb, 0 past pc).
constexpr int PROJ_CMAX = 8;
template <bool PROJ>
__device__ inline void proj_coef(const double* __restrict__ pb, int pc, int64_t r, int64_t n_rows,
                                 double (&bb)[PROJ_CMAX]) {
#pragma unroll
  for (int c = 0; c < PROJ_CMAX; ++c) bb[c] = (PROJ && c < pc && r < n_rows) ? pb[(int64_t)c * n_rows + r] : 0.0;
}
// k is the same on every lane of a wave at both call sites (the rows, not the strains, run along the lanes), so the
// X row is read through the scalar cache
template <bool PROJ>
__device__ inline double proj_value(const double* __restrict__ src, int64_t idx, int64_t k,
                                    const double* __restrict__ px, int pc, const double (&bb)[PROJ_CMAX]) {
  double v = src[idx];
  if constexpr (PROJ) {
    const int64_t ku = (int64_t)__builtin_amdgcn_readfirstlane((int)k);
#pragma unroll
    for (int c = 0; c < PROJ_CMAX; ++c)
      if (c < pc) v -= gmem(px)[ku * pc + c] * bb[c];
  }
  return v;
}

template <int PREC, bool PROJ = false>
__global__ __launch_bounds__(256) void pack_planes_kernel(const double* __restrict__ src, int64_t n_rows, int64_t n_k,
                                                          int64_t row_stride, int64_t k_stride,
                                                          const int64_t* __restrict__ row_index,
                                                          const double* __restrict__ row_norm, void* out_a,
                                                          void* out_b, float* __restrict__ row_scale,
                                                          int64_t rows_pad, int64_t kpad,
                                                          const double* __restrict__ px = nullptr,
                                                          const double* __restrict__ pb = nullptr, int pc = 0) {
  __shared__ double tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const int tid = threadIdx.x;
  const bool kfast = !PROJ && k_stride == 1;  // a projected source has its rows along the unit stride (also P = 1)
  double bb[PROJ_CMAX];  // projected source: r = r0 + tid % 64 in every step below
  proj_coef<PROJ>(pb, pc, r0 + (tid & 63), n_rows, bb);
#pragma unroll 4
  for (int e = tid; e < 64 * 64; e += 256) {
    const int rr = kfast ? (e >> 6) : (e & 63);
    const int kk = kfast ? (e & 63) : (e >> 6);
    const int64_t r = r0 + rr, k = k0 + kk;
    double v = 0.0;
    if (r < n_rows && k < n_k) {
      const int64_t sr = row_index ? row_index[r] : r;
      v = proj_value<PROJ>(src, sr * row_stride + k * k_stride, k, px, pc, bb);
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
  const int64_t kb = k0 + kq;
  if (kb >= kpad) return;
  // the 16 k lie in one 32-k chunk (kpad is a multiple of 32): 32-byte-aligned runs of the planes, written as
  // 16-byte stores
  const int64_t off = ((kb >> 5) * rows_pad + r) * 32 + (kb & 31);
  if constexpr (is_split(PREC)) {
    uint32_t hw[8], lw[8];
#pragma unroll
    for (int e = 0; e < 16; e += 2) {
      u16 h0, l0, h1, l1;
      split2<PREC>(tile[rr][kq + e] * sc, h0, l0);
      split2<PREC>(tile[rr][kq + e + 1] * sc, h1, l1);
      hw[e / 2] = h0 | ((uint32_t)h1 << 16);
      lw[e / 2] = l0 | ((uint32_t)l1 << 16);
    }
    *(uint4*)((u16*)out_a + off) = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    *(uint4*)((u16*)out_a + off + 8) = make_uint4(hw[4], hw[5], hw[6], hw[7]);
    *(uint4*)((u16*)out_b + off) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    *(uint4*)((u16*)out_b + off + 8) = make_uint4(lw[4], lw[5], lw[6], lw[7]);
  } else {
#pragma unroll
    for (int e = 0; e < 16; e += 4)
      *(f32x4*)((float*)out_a + off + e) = f32x4{(float)tile[rr][kq + e], (float)tile[rr][kq + e + 1],
                                                 (float)tile[rr][kq + e + 2], (float)tile[rr][kq + e + 3]};
  }
}

// norm[r] = sqrt(sum_k src(r, k)^2).  Rows along the unit stride (the transposed view of an (N, P) matrix):
// a 1,024-thread block owns 64 consecutive rows, lane = row, and its 16 waves split k (wave w sums k = w, w + 16,
// ...), so every load is 64 consecutive doubles and a few thousand rows against N ~ 10^4..10^5 strains still fill
// the chip; the 16 partial sums are added in wave order.  Otherwise a wave per row, lanes striding k.
template <bool PROJ = false>
__global__ __launch_bounds__(1024) void row_norms_strided_kernel(const double* __restrict__ src, int64_t n_rows,
                                                                 int64_t n_k, int64_t k_stride,
                                                                 const int64_t* __restrict__ row_index,
                                                                 double* __restrict__ norm,
                                                                 const double* __restrict__ px = nullptr,
                                                                 const double* __restrict__ pb = nullptr,
                                                                 int pc = 0) {
  __shared__ double part[16][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * 64 + lane;
  double s = 0.0;
  if (r < n_rows) {
    const int64_t sr = row_index ? row_index[r] : r;
    double bb[PROJ_CMAX];
    proj_coef<PROJ>(pb, pc, r, n_rows, bb);
    for (int64_t k = w; k < n_k; k += 16) {
      const double x = proj_value<PROJ>(src, sr + k * k_stride, k, px, pc, bb);
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
  hipLaunchKernelGGL((pack_planes_kernel<PR, false>), grid, dim3(256), 0, stream, src, n_rows, n_k, row_stride,     \
                     k_stride, row_index, row_norm, out_a, out_b, row_scale, rows_pad, kpad, nullptr, nullptr, 0)
  if (prec == PREC_SPLIT_BF16) FLMM_PACK(PREC_SPLIT_BF16);
  else if (prec == PREC_SPLIT_FP16) FLMM_PACK(PREC_SPLIT_FP16);
  else FLMM_PACK(PREC_FP32);
#undef FLMM_PACK
  return hipGetLastError();
}

hipError_t launch_row_norms(const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                            const int64_t* row_index, double* norm, hipStream_t stream) {
  if (row_stride == 1 && k_stride != 1)
    hipLaunchKernelGGL(row_norms_strided_kernel<false>, dim3((unsigned)((n_rows + 63) / 64)), dim3(1024), 0, stream,
                       src, n_rows, n_k, k_stride, row_index, norm, nullptr, nullptr, 0);
  else
    hipLaunchKernelGGL(row_norms_kernel, dim3((unsigned)((n_rows + 3) / 4)), dim3(256), 0, stream, src, n_rows, n_k,
                       row_stride, k_stride, row_index, norm);
  return hipGetLastError();
}

// The columns of Y (N x P row-major, FP64) less their projection on X (N x C): Y_r = Y - X B, B = X^+ Y (C x P), as
// planes of the transposed view (row p, k = n), without forming Y_r: split-FP16 row norms of Y_r into norm (P
// doubles), then the planes.  The residual is formed in the same order in both passes (y - x_0 b_0 - x_1 b_1 - ...).
hipError_t launch_pack_columns_projected(int prec, const double* y, int64_t N, int64_t P, const double* x, int C,
                                         const double* b, double* norm, void* out_a, void* out_b, float* row_scale,
                                         int64_t rows_pad, int64_t kpad, hipStream_t stream) {
  if (prec == PREC_SPLIT_FP16 && P > 0)
    hipLaunchKernelGGL(row_norms_strided_kernel<true>, dim3((unsigned)((P + 63) / 64)), dim3(1024), 0, stream, y, P,
                       N, P, (const int64_t*)nullptr, norm, x, b, C);
  const dim3 grid((unsigned)((rows_pad + 63) / 64), (unsigned)((kpad + 63) / 64));
#define FLMM_PACKP(PR)                                                                                             \
  hipLaunchKernelGGL((pack_planes_kernel<PR, true>), grid, dim3(256), 0, stream, y, P, N, (int64_t)1, P,            \
                     (const int64_t*)nullptr, (const double*)norm, out_a, out_b, row_scale, rows_pad, kpad, x, b, C)
  if (prec == PREC_SPLIT_BF16) FLMM_PACKP(PREC_SPLIT_BF16);
  else if (prec == PREC_SPLIT_FP16) FLMM_PACKP(PREC_SPLIT_FP16);
  else FLMM_PACKP(PREC_FP32);
#undef FLMM_PACKP
  return hipGetLastError();
}

}  // namespace flmm
