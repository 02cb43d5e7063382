// Shared device helpers for the gfx950 LMM-scan kernels.
//
// Operand planes.  Every GEMM operand that feeds an MFMA tile is stored as a "plane": a rows x K matrix laid
// out in K-chunks of 32, chunk-major: element (row, k) lives at plane[((k / 32) * rows_pad + row) * 32 + k % 32].
// One chunk of 128 rows is then a single contiguous 8 KiB (bf16) or 16 KiB (fp32) block, so a workgroup's tile
// load is a straight coalesced copy, and rows_pad / K are padded with zeros (a zero contributes nothing to a
// sum over K).  Split-BF16 operands carry two planes, hi = bf16(x) and lo = bf16(x - hi), so hi + lo holds ~16
// significant bits of x; a product is recovered as hi*hi + hi*lo + lo*hi with FP32 accumulation.
// Split-FP16 operands are the same construction in IEEE half precision (hi + lo ~ 22 significant bits) on a row
// scaled by a power of two, row_scale = 2^(14 - ceil(log2(bound))) for a bound on the row's magnitudes, so every
// entry stays inside FP16's exponent range; consumers divide the accumulated sums by the (exact) scales.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flmm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // raw 16-byte fragment (a native vector, unlike HIP's
                                                             // uint4 wrapper class, so register arrays of it stay in
                                                             // registers)
typedef unsigned short u16;

// Global-memory view of a pointer.  Accesses through a generic pointer that the compiler cannot trace to a kernel
// argument (e.g. one read from a device table) compile to flat_* instructions, which count against the LDS wait
// counter as well and so stall every LDS wait behind in-flight HBM loads; through this view they are global_*.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

constexpr int KCHUNK = 32;  // K extent of one plane chunk
constexpr double FLOOR64 = 1e-300;  // positive clamp floor of the FP64 reference arithmetic (core.py:80-86)

enum Prec : int { PREC_SPLIT_BF16 = 0, PREC_FP32 = 1, PREC_SPLIT_FP16 = 2 };

// two-byte split precisions share one plane format and tile code path
__host__ __device__ constexpr bool is_split(int prec) { return prec == PREC_SPLIT_BF16 || prec == PREC_SPLIT_FP16; }

__host__ __device__ inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// plane element offset (in elements) of (row, k)
__device__ inline int64_t plane_off(int64_t row, int64_t k, int64_t rows_pad) {
  return ((k >> 5) * rows_pad + row) * 32 + (k & 31);
}

// round-to-nearest-even float -> bf16 bits (inputs here are finite)
__device__ inline u16 f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ inline float bf16_to_f32(u16 h) { return __uint_as_float(((uint32_t)h) << 16); }

// FP64 value -> (hi, lo) bf16 pair with x ~= hi + lo
__device__ inline void split_bf16(double x, u16& hi, u16& lo) {
  hi = f32_to_bf16_rne((float)x);
  double r = x - (double)bf16_to_f32(hi);
  lo = f32_to_bf16_rne((float)r);
}

// FP64 value (already scaled) -> (hi, lo) fp16 bit pair with x ~= hi + lo
__device__ inline void split_f16(double x, u16& hi, u16& lo) {
  const _Float16 h = (_Float16)(float)x;
  const _Float16 l = (_Float16)(float)(x - (double)(float)h);
  hi = __builtin_bit_cast(u16, h);
  lo = __builtin_bit_cast(u16, l);
}

// Fragment-native layout of the screened columns' hi planes (with the uniform-delta layout): within each 32-k chunk
// kc, the region of plane rows hi_from .. p_pad - 1 (p_pad - hi_from rows of 64 bytes in the row layout) holds, for
// each block of 32 plane slots and each 16-k MFMA step s, the 1 KiB B operand of v_mfma_f32_32x32x16_f16 in lane
// order: lane 32 h + c = slot 32 b + c, k segment 2 h + s (8 fp16 each).  One 16-byte load per lane then reads one
// contiguous KiB per wave.  Byte offset of element (slot, k = 32 kc + kk) from the plane's start:
__device__ inline int64_t frag_off(int64_t kc, int64_t p_pad, int64_t hi_from, int64_t slot, int kk) {
  const int seg = kk >> 3;
  return (kc * p_pad + hi_from) * 64 + ((slot >> 5) * 2 + (seg & 1)) * 1024 + ((seg >> 1) * 32 + (slot & 31)) * 16 +
         (kk & 7) * 2;
}

// power-of-two scale that maps magnitudes up to `bound` to at most 2^14 (1 for a zero bound)
__device__ inline double pow2_scale(double bound) {
  if (!(bound > 0.0) || isinf(bound)) return 1.0;
  int e;
  frexp(bound, &e);  // bound = m 2^e, m in [0.5, 1)  ->  bound <= 2^e
  return ldexp(1.0, 14 - e);
}

// split x into the plane pair of precision PREC (two-byte formats); x already scaled for FP16
template <int PREC>
__device__ inline void split2(double x, u16& hi, u16& lo) {
  if constexpr (PREC == PREC_SPLIT_FP16) split_f16(x, hi, lo);
  else split_bf16(x, hi, lo);
}

// the hi half alone: an FP32 value rounded to the two-byte format
template <int PREC>
__device__ inline u16 hi_of(float x) {
  if constexpr (PREC == PREC_SPLIT_FP16) return __builtin_bit_cast(u16, (_Float16)x);
  else return f32_to_bf16_rne(x);
}

// split of an FP32 value into the plane pair: hi = the value rounded to the two-byte format, lo = the exactly
// representable FP32 residual rounded the same way (hi + lo keeps ~22 bits of the value, as split2 does)
template <int PREC>
__device__ inline void split2f(float x, u16& hi, u16& lo) {
  if constexpr (PREC == PREC_SPLIT_FP16) {
    const _Float16 h = (_Float16)x;
    hi = __builtin_bit_cast(u16, h);
    lo = __builtin_bit_cast(u16, (_Float16)(x - (float)h));
  } else {
    hi = f32_to_bf16_rne(x);
    lo = f32_to_bf16_rne(x - bf16_to_f32(hi));
  }
}

// 1 / x for a positive normal x: the hardware reciprocal estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r), each of which squares the relative error (far below FP64 rounding after two); a few
// times cheaper than the IEEE division sequence, which also handles zeros, denormals and infinities
__device__ inline double rcp_pos(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// Wave-wide reductions in DPP steps, register to register (no LDS round trips, which a shuffle takes): within each
// quad (quad_perm [1,0,3,2] and [2,3,0,1]), half-row and row of 16 (row_half_mirror, row_mirror); then rows 1 and 3
// add lane 15 of rows 0 and 2 (row_bcast:15) and rows 2 and 3 add lane 31 (row_bcast:31), so lane 63 holds the
// total, which readlane makes wave-uniform.  Rows a step does not write keep `old` (0 for sums, the value itself
// for maxima).  64-bit values move as two 32-bit halves.
template <int CTRL, int ROWS>
__device__ inline float dpp_f32(float old, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v),
                                                               CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS>
__device__ inline double dpp_f64(double old, double v) {
  const unsigned long long o = __builtin_bit_cast(unsigned long long, old), x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)x, CTRL, ROWS, 0xF, false);
  const unsigned hi =
      (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(x >> 32), CTRL, ROWS, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline double lane63_f64(double v) {
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), 63);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ inline double wave_sum(double v) {
  v += dpp_f64<0xB1, 0xF>(0.0, v);
  v += dpp_f64<0x4E, 0xF>(0.0, v);
  v += dpp_f64<0x141, 0xF>(0.0, v);
  v += dpp_f64<0x140, 0xF>(0.0, v);
  v += dpp_f64<0x142, 0xA>(0.0, v);
  v += dpp_f64<0x143, 0xC>(0.0, v);
  return lane63_f64(v);
}
__device__ inline double wave_max(double v) {
  v = fmax(v, dpp_f64<0xB1, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x4E, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x141, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x140, 0xF>(v, v));
  v = fmax(v, dpp_f64<0x142, 0xA>(v, v));
  v = fmax(v, dpp_f64<0x143, 0xC>(v, v));
  return lane63_f64(v);
}
__device__ inline float wave_sum(float v) {
  v += dpp_f32<0xB1, 0xF>(0.f, v);
  v += dpp_f32<0x4E, 0xF>(0.f, v);
  v += dpp_f32<0x141, 0xF>(0.f, v);
  v += dpp_f32<0x140, 0xF>(0.f, v);
  v += dpp_f32<0x142, 0xA>(0.f, v);
  v += dpp_f32<0x143, 0xC>(0.f, v);
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// XCD-aware block remap: the hardware hands consecutive workgroup ids to the 8 XCDs round-robin.  Renumber so
// that XCD x runs a contiguous range of logical tiles, which keeps the tiles that share operands on one L2.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nblocks) {
  constexpr int NXCD = 8;
  int64_t per = nblocks / NXCD;
  int64_t rem = nblocks % NXCD;
  int64_t xcd = bid % NXCD;
  int64_t slot = bid / NXCD;
  // the first `rem` XCDs own per+1 logical tiles, the others own per
  int64_t start = xcd * per + (xcd < rem ? xcd : rem);
  return start + slot;
}

}  // namespace flmm
