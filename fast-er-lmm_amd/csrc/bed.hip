// PLINK 1 .bed -> genotype matrix on the device, with the unit standardisation fused in.
//
// A SNP-major .bed holds, after its 3 magic bytes, one record of ceil(n_file / 4) bytes per variant; individual i
// of a record sits in byte i / 4 at bit 2 (i % 4).  The 2-bit codes are 00 = homozygous first allele (A1), 01 =
// missing, 10 = heterozygous, 11 = homozygous second allele, so with count_A1 the dosage is 00 -> 2, 10 -> 1,
// 11 -> 0, 01 -> missing (pysnptools' Bed(count_A1=True), which the reference reads through:
// fasterlmm/io.py:54-67, io_stream.py:60-71), and 2 minus that without count_A1.
//
// One thread owns one selected variant and walks the selected individuals twice: the first pass takes exact
// integer sums (count, sum, sum of squares) of the dosages, the second writes the output row by row, so for a
// fixed individual the threads of a wave write consecutive variants (coalesced).  Outputs:
//   BED_I8:  the raw dosage as int8, -1 for missing;
//   BED_F64 / BED_F32: (x - mean) / std with the NaN-aware mean and population std of the selected individuals,
//            std == 0 -> the column becomes 0, missing -> 0 (the reference's unit standardiser, io.py:247-254).
// With integer dosages the centred sum of squares is the exact integer (cnt * S2 - S1^2) / cnt, so the
// statistics carry one rounding each (the reference's floating two-pass sum agrees to an ulp).
#include "kernels.h"

namespace flmm {

__device__ inline int bed_dosage(const uint8_t* __restrict__ rec, int64_t i, bool count_a1) {
  const int code = (rec[i >> 2] >> (2 * (i & 3))) & 3;
  if (code == 1) return -1;
  const int a1 = code == 0 ? 2 : (code == 2 ? 1 : 0);
  return count_a1 ? a1 : 2 - a1;
}

template <int OUT>
__global__ __launch_bounds__(256) void bed_decode_kernel(const uint8_t* __restrict__ bed, int64_t rec_bytes,
                                                         const int64_t* __restrict__ row_idx, int64_t n,
                                                         const int64_t* __restrict__ col_idx, int64_t m,
                                                         bool count_a1, void* __restrict__ out, int64_t ld,
                                                         double* __restrict__ mean_out, double* __restrict__ std_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint8_t* rec = bed + (col_idx ? col_idx[j] : j) * rec_bytes;
  int64_t cnt = 0, s1 = 0, s2 = 0;
  if constexpr (OUT != BED_I8) {
    for (int64_t r = 0; r < n; ++r) {
      const int g = bed_dosage(rec, row_idx ? row_idx[r] : r, count_a1);
      if (g >= 0) { ++cnt; s1 += g; s2 += g * g; }
    }
  }
  double mean = NAN, sd = NAN;
  if (cnt > 0) {
    mean = (double)s1 / (double)cnt;
    sd = sqrt((double)(cnt * s2 - s1 * s1)) / (double)cnt;
    if (sd == 0.0) sd = INFINITY;
  }
  for (int64_t r = 0; r < n; ++r) {
    const int g = bed_dosage(rec, row_idx ? row_idx[r] : r, count_a1);
    if constexpr (OUT == BED_I8) {
      ((int8_t*)out)[r * ld + j] = (int8_t)g;
    } else {
      const double z = g < 0 ? 0.0 : ((double)g - mean) / sd;
      if constexpr (OUT == BED_F64) ((double*)out)[r * ld + j] = z;
      else ((float*)out)[r * ld + j] = (float)z;
    }
  }
  if constexpr (OUT != BED_I8) {
    if (mean_out) mean_out[j] = mean;
    if (std_out) std_out[j] = sd;
  }
}

hipError_t launch_bed_decode(int out_kind, const uint8_t* bed, int64_t rec_bytes, const int64_t* row_idx, int64_t n,
                             const int64_t* col_idx, int64_t m, bool count_a1, void* out, int64_t ld,
                             double* mean_out, double* std_out, hipStream_t stream) {
  const dim3 grid((unsigned)((m + 255) / 256));
  if (out_kind == BED_I8)
    hipLaunchKernelGGL(bed_decode_kernel<BED_I8>, grid, dim3(256), 0, stream, bed, rec_bytes, row_idx, n, col_idx, m,
                       count_a1, out, ld, mean_out, std_out);
  else if (out_kind == BED_F64)
    hipLaunchKernelGGL(bed_decode_kernel<BED_F64>, grid, dim3(256), 0, stream, bed, rec_bytes, row_idx, n, col_idx,
                       m, count_a1, out, ld, mean_out, std_out);
  else
    hipLaunchKernelGGL(bed_decode_kernel<BED_F32>, grid, dim3(256), 0, stream, bed, rec_bytes, row_idx, n, col_idx,
                       m, count_a1, out, ld, mean_out, std_out);
  return hipGetLastError();
}

}  // namespace flmm
