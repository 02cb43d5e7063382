// C ABI entry points: argument validation, error reporting and launch.  See include/fasterlmm_amd.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>

#include "../../include/fasterlmm_amd.h"
#include "kernels.h"

using namespace flmm;

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
static int check_launch(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(FLMM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  g_err.clear();
  return FLMM_OK;
}

#define REQUIRE(cond, ...) \
  do {                     \
    if (!(cond)) return fail(FLMM_ERR_ARG, __VA_ARGS__); \
  } while (0)

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }
static bool valid_prec(int p) { return p == FLMM_PREC_SPLIT_BF16 || p == FLMM_PREC_FP32 || p == FLMM_PREC_SPLIT_FP16; }
static bool two_planes(int p) { return p != FLMM_PREC_FP32; }

// Argument checks and launch of the two-phase wide scan; `who` names the entry point in error messages.
static int scan_lowrank(const char* who, int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi,
                        const void* s2_lo, const float* s1_scale, const float* s2_scale, int64_t mc_pad,
                        const void* q1_hi, const void* q1_lo, const void* q2_hi, const void* q2_lo,
                        const float* q1_scale, const float* q2_scale, int64_t p_pad, int64_t kpad, int64_t kpad2,
                        int64_t mc, int64_t P, int64_t n_real, const double* rWr, const double* h2, double nu,
                        const int64_t* row_index, double* f, double* beta, double* se, double* sfve, double* nullh2,
                        double* pmax, void* stream) {
  REQUIRE(valid_prec(prec), "%s: bad precision %d", who, prec);
  REQUIRE(mc >= 0 && mc <= mc_pad && mc_pad % TILE == 0, "%s: bad SNP extent", who);
  REQUIRE(P >= 0 && P <= p_pad && p_pad % 256 == 0, "%s: bad column extent (p_pad %% 256)", who);
  REQUIRE(kpad2 > 0 && kpad2 % TILE == 0 && kpad % TILE == 0 && kpad2 < kpad,
          "%s: need 0 < kpad2 < kpad, both multiples of 128", who);
  REQUIRE(n_real >= 0 && n_real <= P, "%s: n_real must be in [0, P]", who);
  REQUIRE(nu > 0.0, "%s: dof must be positive", who);
  const bool two = two_planes(prec);
  REQUIRE(s1_hi && s2_hi && q1_hi && q2_hi && (!two || (s1_lo && s2_lo && q1_lo && q2_lo)),
          "%s: null operand plane", who);
  REQUIRE(aligned16(s1_hi) && aligned16(s2_hi) && aligned16(q1_hi) && aligned16(q2_hi) &&
              (!two || (aligned16(s1_lo) && aligned16(s2_lo) && aligned16(q1_lo) && aligned16(q2_lo))),
          "%s: operand planes must be 16-byte aligned", who);
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || (s1_scale && s2_scale && q1_scale && q2_scale),
          "%s: split-FP16 needs the four row-scale arrays", who);
  REQUIRE(rWr && pmax, "%s: null rWr / pmax", who);
  REQUIRE(n_real == 0 || (f && beta && se && sfve && nullh2 && h2 && row_index),
          "%s: null output", who);
  if (mc == 0 || P == 0) return FLMM_OK;
  ScanOut out{f, beta, se, sfve, nullh2, n_real};
  ScanScales sc{nullptr, nullptr, nullptr, nullptr};
  if (prec == FLMM_PREC_SPLIT_FP16) sc = ScanScales{s1_scale, s2_scale, q1_scale, q2_scale};
  const hipStream_t st = (hipStream_t)stream;
  hipError_t e;
#define FLMM_SCAN(PR)                                                                                                \
  do {                                                                                                               \
    TileArgs<PR, 2> t{};                                                                                             \
    t.a[0][0] = s1_hi; t.a[0][1] = s1_lo; t.a[1][0] = s2_hi; t.a[1][1] = s2_lo;                                      \
    t.b[0][0] = q1_hi; t.b[0][1] = q1_lo; t.b[1][0] = q2_hi; t.b[1][1] = q2_lo;                                      \
    t.a_rows_pad = mc_pad; t.b_rows_pad = p_pad; t.nchunks = (int)(kpad / KCHUNK);                                  \
    t.nchunks2 = (int)(kpad2 / KCHUNK);                                                                              \
    e = launch_wald_scan_lowrank<PR>(t, sc, mc, P, n_real, rWr, h2, nu, row_index, out, pmax, st);                  \
  } while (0)
  if (prec == FLMM_PREC_SPLIT_BF16) FLMM_SCAN(PREC_SPLIT_BF16);
  else if (prec == FLMM_PREC_SPLIT_FP16) FLMM_SCAN(PREC_SPLIT_FP16);
  else FLMM_SCAN(PREC_FP32);
#undef FLMM_SCAN
  return check_launch(e, who);
}

// ---- scratch of the allocation-free scan (flmm_wald_scan_lowrank_noalloc) ----
// The per-tile column maxima of one scan live in this file-scope device array instead of a caller buffer.  Host
// code must not pass `scratch` itself to a kernel (in host code the name is a host-side stand-in, not a device
// address): scratch_ptr() looks the array's device address up with hipGetSymbolAddress once per device, with that
// device current, and every launch uses that address.  Each GPU holds its own copy until the process ends.
static __device__ __align__(256) double scratch[FLMM_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];  // per device ordinal

// Device address of this device's scratch array (the device that `stream` belongs to; the current device for the
// null stream), or NULL with the error set.
static double* scratch_ptr(hipStream_t stream, const char* who) {
  int dev = 0;
  hipError_t e = stream ? hipStreamGetDevice(stream, &dev) : hipGetDevice(&dev);
  if (e != hipSuccess) {
    fail(FLMM_ERR_HIP, "%s: device of the stream: %s", who, hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= 64) {
    fail(FLMM_ERR_ARG, "%s: device ordinal %d out of range", who, dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    if ((e = hipGetDevice(&cur)) != hipSuccess) {
      fail(FLMM_ERR_HIP, "%s: hipGetDevice: %s", who, hipGetErrorString(e));
      return nullptr;
    }
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) {
      fail(FLMM_ERR_HIP, "%s: hipSetDevice(%d): %s", who, dev, hipGetErrorString(e));
      return nullptr;
    }
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      fail(FLMM_ERR_HIP, "%s: hipGetSymbolAddress(scratch): %s", who, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return (double*)g_scratch_ptr[dev];
}

extern "C" {

int flmm_abi_version(void) { return 18; }

const char* flmm_last_error(void) { return g_err.c_str(); }

int flmm_bed_decode(const uint8_t* bed, int64_t rec_bytes, const int64_t* row_idx, int64_t n,
                    const int64_t* col_idx, int64_t m, int count_a1, int out_kind, void* out, int64_t ld,
                    double* mean, double* std, void* stream) {
  REQUIRE(out_kind == FLMM_BED_I8 || out_kind == FLMM_BED_F64 || out_kind == FLMM_BED_F32,
          "flmm_bed_decode: bad output kind %d", out_kind);
  REQUIRE(rec_bytes > 0 && n >= 0 && m >= 0 && ld >= m, "flmm_bed_decode: bad sizes");
  REQUIRE(row_idx || n <= 4 * rec_bytes, "flmm_bed_decode: n exceeds the individuals of a record");
  REQUIRE((bed && out) || n == 0 || m == 0, "flmm_bed_decode: null pointer");
  if (n == 0 || m == 0) return FLMM_OK;
  return check_launch(launch_bed_decode(out_kind, bed, rec_bytes, row_idx, n, col_idx, m, count_a1 != 0, out, ld,
                                        mean, std, (hipStream_t)stream),
                      "flmm_bed_decode");
}

int flmm_row_norms(const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                   const int64_t* row_index, double* norm, void* stream) {
  REQUIRE(n_rows >= 0 && n_k >= 0, "flmm_row_norms: bad sizes");
  REQUIRE((src && norm) || n_rows == 0, "flmm_row_norms: null pointer");
  if (n_rows == 0) return FLMM_OK;
  return check_launch(launch_row_norms(src, n_rows, n_k, row_stride, k_stride, row_index, norm, (hipStream_t)stream),
                      "flmm_row_norms");
}

int flmm_pack_planes(int prec, const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                     const int64_t* row_index, const double* row_norm, void* plane_hi, void* plane_lo,
                     float* row_scale, int64_t rows_pad, int64_t kpad, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_pack_planes: bad precision %d", prec);
  REQUIRE(n_rows >= 0 && n_k >= 0 && n_rows <= rows_pad && n_k <= kpad, "flmm_pack_planes: bad sizes");
  REQUIRE(rows_pad % TILE == 0 && kpad % KCHUNK == 0, "flmm_pack_planes: rows_pad %% 128 and kpad %% 32 required");
  REQUIRE(plane_hi && (!two_planes(prec) || plane_lo), "flmm_pack_planes: null plane");
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || (row_scale && (row_norm || n_rows == 0)),
          "flmm_pack_planes: split-FP16 needs row_norm and row_scale");
  REQUIRE(src || n_rows == 0 || n_k == 0, "flmm_pack_planes: null src");
  if (rows_pad == 0 || kpad == 0) return FLMM_OK;
  return check_launch(launch_pack_planes(prec, src, n_rows, n_k, row_stride, k_stride, row_index, row_norm, plane_hi,
                                         plane_lo, row_scale, rows_pad, kpad, (hipStream_t)stream),
                      "flmm_pack_planes");
}

int flmm_pack_columns_projected(int prec, const double* y, int64_t N, int64_t P, const double* x, int C,
                                const double* b, double* norm, void* plane_hi, void* plane_lo, float* row_scale,
                                int64_t rows_pad, int64_t kpad, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_pack_columns_projected: bad precision %d", prec);
  REQUIRE(N >= 0 && P >= 0 && C >= 0 && C <= 8 && P <= rows_pad && N <= kpad,
          "flmm_pack_columns_projected: bad sizes (C <= 8)");
  REQUIRE(rows_pad % TILE == 0 && kpad % KCHUNK == 0,
          "flmm_pack_columns_projected: rows_pad %% 128 and kpad %% 32 required");
  REQUIRE(plane_hi && (!two_planes(prec) || plane_lo), "flmm_pack_columns_projected: null plane");
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || (row_scale && (norm || P == 0)),
          "flmm_pack_columns_projected: split-FP16 needs norm and row_scale");
  REQUIRE((y && (C == 0 || (x && b))) || P == 0 || N == 0, "flmm_pack_columns_projected: null pointer");
  if (rows_pad == 0 || kpad == 0) return FLMM_OK;
  return check_launch(launch_pack_columns_projected(prec, y, N, P, x, C, b, norm, plane_hi, plane_lo, row_scale,
                                                    rows_pad, kpad, (hipStream_t)stream),
                      "flmm_pack_columns_projected");
}

int flmm_rotate(int prec, int mode, const void* a_hi, const void* a_lo, const float* a_scale, int64_t a_rows_pad,
                const void* b_hi, const void* b_lo, const float* b_scale, int64_t b_rows_pad, int64_t n_pad,
                void* out0, void* out1, void* out2, void* out3, float* out_scale1, float* out_scale2, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_rotate: bad precision %d", prec);
  REQUIRE(mode == FLMM_ROTATE_UY || mode == FLMM_ROTATE_US || mode == FLMM_ROTATE_US_X2, "flmm_rotate: bad mode %d",
          mode);
  REQUIRE(a_rows_pad > 0 && b_rows_pad > 0 && n_pad > 0, "flmm_rotate: empty operand");
  REQUIRE(a_rows_pad % TILE == 0 && b_rows_pad % TILE == 0 && n_pad % KCHUNK == 0,
          "flmm_rotate: rows_pad %% 128 and n_pad %% 32 required");
  const bool two = two_planes(prec);
  REQUIRE(a_hi && b_hi && (!two || (a_lo && b_lo)), "flmm_rotate: null operand plane");
  REQUIRE(aligned16(a_hi) && aligned16(b_hi) && (!two || (aligned16(a_lo) && aligned16(b_lo))),
          "flmm_rotate: operand planes must be 16-byte aligned");
  if (mode == FLMM_ROTATE_US_X2)
    REQUIRE(out0 && out2 && (!two || out1) && aligned16(out2), "flmm_rotate: null or misaligned output");
  else
    REQUIRE(out0 && (mode == FLMM_ROTATE_UY || (out1 && (!two || (out2 && out3)))), "flmm_rotate: null output");
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 ||
              (a_scale && b_scale &&
               (mode == FLMM_ROTATE_UY || (out_scale1 && (mode == FLMM_ROTATE_US_X2 || out_scale2)))),
          "flmm_rotate: split-FP16 needs the operand and output row scales");
  const int nch = (int)(n_pad / KCHUNK);
  hipError_t e;
  if (prec == FLMM_PREC_SPLIT_BF16) {
    TileArgs<PREC_SPLIT_BF16, 1> t{};
    t.a[0][0] = a_hi; t.a[0][1] = a_lo; t.b[0][0] = b_hi; t.b[0][1] = b_lo;
    t.a_rows_pad = a_rows_pad; t.b_rows_pad = b_rows_pad; t.nchunks = nch;
    e = launch_rotate<PREC_SPLIT_BF16>(mode, t, nullptr, nullptr, out0, out1, out2, out3, nullptr, nullptr,
                                       (hipStream_t)stream);
  } else if (prec == FLMM_PREC_SPLIT_FP16) {
    TileArgs<PREC_SPLIT_FP16, 1> t{};
    t.a[0][0] = a_hi; t.a[0][1] = a_lo; t.b[0][0] = b_hi; t.b[0][1] = b_lo;
    t.a_rows_pad = a_rows_pad; t.b_rows_pad = b_rows_pad; t.nchunks = nch;
    e = launch_rotate<PREC_SPLIT_FP16>(mode, t, a_scale, b_scale, out0, out1, out2, out3, out_scale1, out_scale2,
                                       (hipStream_t)stream);
  } else {
    TileArgs<PREC_FP32, 1> t{};
    t.a[0][0] = a_hi; t.b[0][0] = b_hi;
    t.a_rows_pad = a_rows_pad; t.b_rows_pad = b_rows_pad; t.nchunks = nch;
    e = launch_rotate<PREC_FP32>(mode, t, nullptr, nullptr, out0, out1, out2, out3, nullptr, nullptr,
                                 (hipStream_t)stream);
  }
  return check_launch(e, "flmm_rotate");
}

int flmm_delta_weights(const double* s, int64_t neff, const double* s_lam, int64_t nlam, int64_t kpad, int n_grid,
                       double lo, double hi, float* wplane, float* lam, double* lam64, void* stream) {
  REQUIRE(n_grid >= 1 && n_grid <= TILE, "flmm_delta_weights: n_grid must be in [1, 128], got %d", n_grid);
  REQUIRE(neff >= 1 && neff <= kpad && kpad % TILE == 0 && nlam >= 1, "flmm_delta_weights: bad neff / nlam / kpad");
  REQUIRE(s && s_lam && wplane && lam, "flmm_delta_weights: null pointer");
  return check_launch(
      launch_delta_weights(s, neff, s_lam, nlam, kpad, n_grid, lo, hi, wplane, lam, lam64, (hipStream_t)stream),
      "flmm_delta_weights");
}

int flmm_delta_grid(const float* wplane, const float* lam, int n_grid, double nco, const float* uy, int64_t P,
                    int64_t p_pad, int64_t kpad, int32_t* idx, void* stream) {
  REQUIRE(n_grid >= 1 && n_grid <= TILE, "flmm_delta_grid: n_grid must be in [1, 128]");
  REQUIRE(P >= 0 && P <= p_pad && p_pad % TILE == 0 && kpad % TILE == 0 && kpad > 0, "flmm_delta_grid: bad sizes");
  REQUIRE(nco > 0.0, "flmm_delta_grid: loss coefficient must be positive");
  REQUIRE(wplane && lam && uy && idx, "flmm_delta_grid: null pointer");
  REQUIRE(aligned16(wplane) && aligned16(uy), "flmm_delta_grid: planes must be 16-byte aligned");
  if (p_pad == 0) return FLMM_OK;
  return check_launch(launch_delta_grid(wplane, lam, n_grid, nco, uy, P, p_pad, kpad, idx, (hipStream_t)stream),
                      "flmm_delta_grid");
}

int flmm_delta_grid_split(const float* wplane, const float* lam, int n_grid, double nco, const float* uy,
                          const float* col_scale, double s_min, double lo, double hi, int64_t P, int64_t p_pad,
                          int64_t kpad, int32_t* idx, void* wsplit, void* stream) {
  REQUIRE(n_grid >= 1 && n_grid <= TILE, "flmm_delta_grid_split: n_grid must be in [1, 128]");
  REQUIRE(P >= 0 && P <= p_pad && p_pad % TILE == 0 && kpad % TILE == 0 && kpad > 0,
          "flmm_delta_grid_split: bad sizes");
  REQUIRE(nco > 0.0, "flmm_delta_grid_split: loss coefficient must be positive");
  REQUIRE(s_min >= 0.0, "flmm_delta_grid_split: s_min must be >= 0 (the FP32 grid takes other spectra)");
  REQUIRE(wplane && lam && uy && col_scale && idx, "flmm_delta_grid_split: null pointer");
  REQUIRE(aligned16(wplane) && aligned16(uy) && aligned16(wsplit),
          "flmm_delta_grid_split: planes and workspace must be 16-byte aligned");
  if (p_pad == 0) return FLMM_OK;
  return check_launch(launch_delta_grid_split(wplane, lam, n_grid, nco, uy, col_scale, s_min, lo, hi, P, p_pad, kpad,
                                              idx, wsplit, (hipStream_t)stream),
                      "flmm_delta_grid_split");
}

static int check_refine(const char* who, int prec, const float* uy, int64_t P, int64_t p_pad, int64_t kpad,
                        const double* s, int64_t neff, const double* s_lam, int64_t nlam, double nco, int n_grid,
                        double* log_delta, double* rWr, double* h2, void* q1_hi, void* q1_lo, void* q2_hi,
                        void* q2_lo, float* q1_scale, float* q2_scale) {
  REQUIRE(valid_prec(prec), "%s: bad precision %d", who, prec);
  REQUIRE(n_grid >= 1, "%s: n_grid >= 1 required", who);
  REQUIRE(P >= 0 && P <= p_pad && p_pad % TILE == 0 && kpad % TILE == 0 && neff >= 1 && neff <= kpad && nlam >= 1,
          "%s: bad sizes", who);
  REQUIRE(nco > 0.0, "%s: loss coefficient must be positive", who);
  const bool two = two_planes(prec);
  // q2_hi == NULL: the Q2 planes are not written (factored weights: flmm_factored_weight_planes writes them)
  REQUIRE(uy && s && s_lam && log_delta && rWr && h2 && q1_hi && (!two || q1_lo) && (!q2_hi || !two || q2_lo),
          "%s: null pointer", who);
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || (q1_scale && (!q2_hi || q2_scale)),
          "%s: split-FP16 needs q1_scale / q2_scale", who);
  return FLMM_OK;
}

// the bucketed layout's arguments: q_rows Q-plane rows (p_pad without a layout, at least
// flmm_screen_q_rows(p_pad, hi_from, P) with one), buckets of width t_h <= 1/5 (the interpolation bound) from t_lo
static int check_layout(const char* who, int64_t P, int64_t p_pad, int64_t hi_from, int64_t q_rows, double t_lo,
                        double t_h, int nb, const void* order) {
  if (!order || hi_from >= P) {
    REQUIRE(q_rows >= p_pad && q_rows % TILE == 0, "%s: q_rows must be >= p_pad and a multiple of 128", who);
    return FLMM_OK;
  }
  REQUIRE(hi_from % 256 == 0, "%s: a layout needs hi_from a multiple of 256", who);
  REQUIRE(q_rows % TILE == 0 && q_rows >= hi_from + screen_order_slots(P - hi_from) && q_rows >= p_pad,
          "%s: q_rows below flmm_screen_q_rows", who);
  REQUIRE(nb >= 1 && nb <= SO_NB_MAX && isfinite(t_lo) && t_h > 0.0 && t_h <= 0.2,
          "%s: the layout needs 1 <= nb <= %d buckets of width 0 < t_h <= 1/5", who, SO_NB_MAX);
  REQUIRE(P - hi_from < (1LL << 31) - (int64_t)SO_PAD * SO_NB_MAX, "%s: too many screened columns", who);
  return FLMM_OK;
}

int flmm_delta_refine(int prec, int mode, const float* uy, int64_t P, int64_t p_pad, int64_t kpad, const double* s,
                      int64_t neff, const double* s_lam, int64_t nlam, double nco, int n_grid, double lo, double hi,
                      const int32_t* idx, const double* lam64, double* log_delta, double* rWr, double* h2,
                      void* q1_hi, void* q1_lo, void* q2_hi, void* q2_lo, float* q1_scale, float* q2_scale,
                      int64_t hi_from, int64_t q_rows, double t_lo, double t_h, int nb, void* order,
                      void* stream) {
  REQUIRE(hi_from >= 0, "flmm_delta_refine: hi_from must be >= 0");
  const int rl = check_layout("flmm_delta_refine", P, p_pad, hi_from, q_rows, t_lo, t_h, nb, order);
  if (rl) return rl;
  REQUIRE(mode == FLMM_DELTA_GOLDEN || mode == FLMM_DELTA_NEWTON || mode == FLMM_DELTA_FIXED ||
              mode == FLMM_DELTA_GRID,
          "flmm_delta_refine: bad mode %d", mode);
  REQUIRE(idx || mode == FLMM_DELTA_FIXED || mode == FLMM_DELTA_GRID, "flmm_delta_refine: null idx");
  const int rc = check_refine("flmm_delta_refine", prec, uy, P, p_pad, kpad, s, neff, s_lam, nlam, nco, n_grid,
                              log_delta, rWr, h2, q1_hi, q1_lo, q2_hi, q2_lo, q1_scale, q2_scale);
  if (rc) return rc;
  if (p_pad == 0) return FLMM_OK;
  return check_launch(launch_delta_refine(prec, mode, uy, P, p_pad, kpad, s, neff, s_lam, nlam, nco, n_grid, lo, hi,
                                          idx, lam64, log_delta, rWr, h2, q1_hi, q1_lo, q2_hi, q2_lo, q1_scale,
                                          q2_scale, 0.0, 0.0, nullptr, hi_from, q_rows, t_lo, t_h, nb,
                                          hi_from < P ? order : nullptr, (hipStream_t)stream),
                      "flmm_delta_refine");
}

int flmm_gram(int prec, const void* z_hi, const void* z_lo, const float* z_scale, int64_t rows_pad, int64_t kpad,
              int64_t n, int n_split, double* partial, double* out, int64_t ld, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_gram: bad precision %d", prec);
  REQUIRE(rows_pad > 0 && rows_pad % TILE == 0 && kpad > 0 && kpad % KCHUNK == 0 && n > 0 && n <= rows_pad &&
              n_split >= 1 && ld >= n,
          "flmm_gram: need rows_pad %% 128, kpad %% 32, 0 < n <= rows_pad, n_split >= 1, ld >= n");
  const bool two = two_planes(prec);
  REQUIRE(z_hi && (!two || z_lo) && partial && out, "flmm_gram: null pointer");
  REQUIRE(aligned16(z_hi) && (!two || aligned16(z_lo)), "flmm_gram: planes must be 16-byte aligned");
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || z_scale, "flmm_gram: split-FP16 needs the row scales");
  const hipStream_t st = (hipStream_t)stream;
  hipError_t e;
#define FLMM_GRAM(PR)                                                                                                \
  do {                                                                                                               \
    TileArgs<PR, 1> t{};                                                                                             \
    t.a[0][0] = z_hi; t.a[0][1] = z_lo; t.b[0][0] = z_hi; t.b[0][1] = z_lo;                                          \
    t.a_rows_pad = rows_pad; t.b_rows_pad = rows_pad; t.nchunks = (int)(kpad / KCHUNK);                              \
    e = launch_gram<PR>(t, prec == FLMM_PREC_SPLIT_FP16 ? z_scale : nullptr, n_split, partial, n, out, ld, st);      \
  } while (0)
  if (prec == FLMM_PREC_SPLIT_BF16) FLMM_GRAM(PREC_SPLIT_BF16);
  else if (prec == FLMM_PREC_SPLIT_FP16) FLMM_GRAM(PREC_SPLIT_FP16);
  else FLMM_GRAM(PREC_FP32);
#undef FLMM_GRAM
  return check_launch(e, "flmm_gram");
}

int flmm_gram_f64(const double* z, int64_t ldz, int64_t n, int64_t m, int n_split, double* partial, double* out,
                  int64_t ld, void* stream) {
  REQUIRE(n > 0 && m > 0 && ldz >= m && n_split >= 1 && ld >= n,
          "flmm_gram_f64: need n > 0, m > 0, ldz >= m, n_split >= 1, ld >= n");
  REQUIRE(z && partial && out, "flmm_gram_f64: null pointer");
  return check_launch(launch_gram_f64(z, ldz, n, m, n_split, partial, out, ld, (hipStream_t)stream), "flmm_gram_f64");
}

int flmm_rotate_f64(const double* z, int64_t ldz, const int64_t* col_index, int64_t mc, int64_t n, const double* u,
                    int64_t ldu, int64_t neff, double* out, int64_t rows_pad, int64_t kpad, void* stream) {
  REQUIRE(mc >= 0 && n >= 1 && neff >= 1 && neff <= kpad && kpad % TILE == 0 && rows_pad % TILE == 0 &&
              rows_pad >= mc && ldu >= neff && (col_index || ldz >= mc) && ldz >= 1,
          "flmm_rotate_f64: need 0 <= mc <= rows_pad, 1 <= neff <= kpad, rows_pad and kpad %% 128, ldu >= neff, "
          "ldz >= mc without col_index");
  REQUIRE(z && u && out, "flmm_rotate_f64: null pointer");
  if (rows_pad == 0) return FLMM_OK;
  return check_launch(launch_rotate_f64(z, ldz, col_index, mc, n, u, ldu, neff, out, rows_pad, kpad, (hipStream_t)stream),
                      "flmm_rotate_f64");
}

int flmm_gemm_f64(const double* a, int64_t rows, int64_t k, const double* b, int64_t cols, double* c, void* stream) {
  REQUIRE(rows >= 0 && rows % 64 == 0 && k > 0 && k % KCHUNK == 0 && cols > 0 && cols % TILE == 0,
          "flmm_gemm_f64: need rows %% 64, k %% 32 and cols %% 128 == 0");
  REQUIRE(a && b && c, "flmm_gemm_f64: null pointer");
  if (rows == 0) return FLMM_OK;
  return check_launch(launch_gemm_f64(a, rows, k, b, cols, c, (hipStream_t)stream), "flmm_gemm_f64");
}

int flmm_rescore_f64(const double* US, const double* UY, int64_t ld, const double* s, int64_t neff,
                     const double* log_delta, const int64_t* us_row, const int64_t* col, const int64_t* out_row,
                     int64_t out_row_base, int64_t n, const unsigned long long* n_dev, double nu, double* f,
                     double* beta, double* se, double* sfve, int64_t out_ld, void* stream) {
  REQUIRE(n >= 0 && neff >= 1 && neff <= ld && out_ld >= 1 && nu > 0.0 && out_row_base >= 0,
          "flmm_rescore_f64: bad sizes");
  REQUIRE(n == 0 || (US && UY && s && log_delta && us_row && col && f && beta && se && sfve),
          "flmm_rescore_f64: null pointer");
  if (n == 0) return FLMM_OK;
  return check_launch(launch_rescore(US, UY, ld, s, neff, log_delta, us_row, col, out_row, out_row_base, n, n_dev, nu,
                                     f, beta, se, sfve, out_ld, (hipStream_t)stream),
                      "flmm_rescore_f64");
}

int flmm_select_above(const double* f, int64_t rows, int64_t cols, int64_t ld, double thr, int64_t* out_row,
                      int64_t* out_col, int64_t cap, unsigned long long* count, void* stream) {
  REQUIRE(rows >= 0 && cols >= 0 && ld >= cols && cap >= 0, "flmm_select_above: bad sizes");
  REQUIRE(count && (rows * cols == 0 || f) && (cap == 0 || (out_row && out_col)), "flmm_select_above: null pointer");
  if (rows * cols == 0) return FLMM_OK;
  return check_launch(launch_select_above(f, rows, cols, ld, thr, out_row, out_col, cap, count, (hipStream_t)stream),
                      "flmm_select_above");
}

int flmm_factored_weight_planes(int prec, int64_t P, int64_t p_pad, int64_t rpad, const double* t, const double* lam,
                                int64_t R, const double* log_delta, void* q2_hi, void* q2_lo, float* q2_scale,
                                int64_t hi_from, const void* order, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_factored_weight_planes: bad precision %d", prec);
  REQUIRE(hi_from >= 0, "flmm_factored_weight_planes: hi_from must be >= 0");
  REQUIRE(P >= 0 && P <= p_pad && p_pad % TILE == 0, "flmm_factored_weight_planes: bad column extent");
  REQUIRE(rpad > 0 && rpad % TILE == 0 && R >= 1 && R <= rpad,
          "flmm_factored_weight_planes: need 1 <= R <= rpad, rpad a multiple of 128");
  const bool two = two_planes(prec);
  REQUIRE(t && lam && log_delta && q2_hi && (!two || q2_lo), "flmm_factored_weight_planes: null pointer");
  REQUIRE(aligned16(q2_hi) && (!two || aligned16(q2_lo)), "flmm_factored_weight_planes: planes must be 16-byte aligned");
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || q2_scale, "flmm_factored_weight_planes: split-FP16 needs q2_scale");
  if (p_pad == 0) return FLMM_OK;
  return check_launch(launch_factored_planes(prec, P, p_pad, rpad, t, lam, R, log_delta, q2_hi, q2_lo, q2_scale,
                                             hi_from, order != nullptr && hi_from < P, (hipStream_t)stream),
                      "flmm_factored_weight_planes");
}

int flmm_delta_refine_learned(int prec, const float* uy, int64_t P, int64_t p_pad, int64_t kpad, const double* s,
                              int64_t neff, const double* s_lam, int64_t nlam, double nco, int n_grid, double lo,
                              double hi, const int32_t* idx, const double* lam64, double radius, double judge_tol,
                              double* log_delta, int32_t* fallback, double* rWr, double* h2, void* q1_hi,
                              void* q1_lo, void* q2_hi, void* q2_lo, float* q1_scale, float* q2_scale,
                              int64_t hi_from, int64_t q_rows, double t_lo, double t_h, int nb, void* order,
                              void* stream) {
  REQUIRE(radius > 0.0 && judge_tol >= 0.0 && idx, "flmm_delta_refine_learned: bad radius / tol / idx");
  REQUIRE(hi_from >= 0, "flmm_delta_refine_learned: hi_from must be >= 0");
  const int rl = check_layout("flmm_delta_refine_learned", P, p_pad, hi_from, q_rows, t_lo, t_h, nb, order);
  if (rl) return rl;
  const int rc = check_refine("flmm_delta_refine_learned", prec, uy, P, p_pad, kpad, s, neff, s_lam, nlam, nco,
                              n_grid, log_delta, rWr, h2, q1_hi, q1_lo, q2_hi, q2_lo, q1_scale, q2_scale);
  if (rc) return rc;
  if (p_pad == 0) return FLMM_OK;
  return check_launch(launch_delta_refine(prec, 3, uy, P, p_pad, kpad, s, neff, s_lam, nlam, nco, n_grid, lo, hi, idx,
                                          lam64, log_delta, rWr, h2, q1_hi, q1_lo, q2_hi, q2_lo, q1_scale, q2_scale,
                                          radius, judge_tol, fallback, hi_from, q_rows, t_lo, t_h, nb,
                                          hi_from < P ? order : nullptr, (hipStream_t)stream),
                      "flmm_delta_refine_learned");
}

int flmm_mlp_features(const float* uy, int64_t P, int64_t p_pad, int64_t neff, const int32_t* bin_of_k, double n,
                      double c, float* feat, void* stream) {
  REQUIRE(P >= 0 && P <= p_pad && p_pad % TILE == 0 && neff >= 1, "flmm_mlp_features: bad sizes");
  REQUIRE((uy && bin_of_k && feat) || P == 0, "flmm_mlp_features: null pointer");
  if (P == 0) return FLMM_OK;
  return check_launch(launch_mlp_features(uy, P, p_pad, neff, bin_of_k, (float)log(n), (float)c, feat,
                                          (hipStream_t)stream),
                      "flmm_mlp_features");
}

int flmm_mlp_dense_fwd(const float* in, int64_t B, int n_in, const float* W, const float* bias, int n_out, int act,
                       float* out, void* stream) {
  REQUIRE(B >= 0 && n_in > 0 && n_out > 0 && (act == 0 || act == 1), "flmm_mlp_dense_fwd: bad sizes / activation");
  REQUIRE((in && W && bias && out) || B == 0, "flmm_mlp_dense_fwd: null pointer");
  if (B == 0) return FLMM_OK;
  return check_launch(launch_mlp_dense_fwd(in, B, n_in, W, bias, n_out, act, out, (hipStream_t)stream),
                      "flmm_mlp_dense_fwd");
}

int flmm_mlp_dense_bwd(const float* dout, float* dz, const float* out, const float* in, int64_t B, int n_in,
                       int n_out, int act, const float* W, float* din, float* dW, float* db, void* stream) {
  REQUIRE(B > 0 && n_in > 0 && n_out > 0 && (act == 0 || act == 1), "flmm_mlp_dense_bwd: bad sizes / activation");
  REQUIRE(dout && dz && out && in && W && dW && db, "flmm_mlp_dense_bwd: null pointer");
  return check_launch(launch_mlp_dense_bwd(dout, dz, out, in, B, n_in, n_out, act, W, din, dW, db,
                                           (hipStream_t)stream),
                      "flmm_mlp_dense_bwd");
}

int flmm_mlp_mse_grad(const float* pred, const float* target, int64_t B, float* dpred, void* stream) {
  REQUIRE(B > 0 && pred && target && dpred, "flmm_mlp_mse_grad: bad arguments");
  return check_launch(launch_mlp_mse_grad(pred, target, B, dpred, (hipStream_t)stream), "flmm_mlp_mse_grad");
}

int flmm_adam(float* w, const float* g, float* m, float* v, int64_t n, double lr, double beta1, double beta2,
              double eps, int step, void* stream) {
  REQUIRE(n >= 0 && step >= 1 && lr > 0.0, "flmm_adam: bad arguments");
  REQUIRE((w && g && m && v) || n == 0, "flmm_adam: null pointer");
  if (n == 0) return FLMM_OK;
  return check_launch(launch_adam(w, g, m, v, n, (float)lr, (float)beta1, (float)beta2, (float)eps, step,
                                  (hipStream_t)stream),
                      "flmm_adam");
}

int flmm_wald_scan(int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi, const void* s2_lo,
                   const float* s1_scale, const float* s2_scale, int64_t mc_pad, const void* q1_hi,
                   const void* q1_lo, const void* q2_hi, const void* q2_lo, const float* q1_scale,
                   const float* q2_scale, int64_t p_pad, int64_t kpad, int64_t mc, int64_t P, int64_t n_real,
                   const double* rWr, const double* h2, double nu, const int64_t* row_index, double* f, double* beta,
                   double* se, double* sfve, double* nullh2, double* pmax, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_wald_scan: bad precision %d", prec);
  REQUIRE(mc >= 0 && mc <= mc_pad && mc_pad % TILE == 0, "flmm_wald_scan: bad SNP extent");
  REQUIRE(P >= 0 && P <= p_pad && p_pad % TILE == 0, "flmm_wald_scan: bad column extent");
  REQUIRE(kpad > 0 && kpad % TILE == 0, "flmm_wald_scan: kpad must be a positive multiple of 128");
  REQUIRE(n_real >= 0 && n_real <= P, "flmm_wald_scan: n_real must be in [0, P]");
  REQUIRE(nu > 0.0, "flmm_wald_scan: dof must be positive");
  const bool two = two_planes(prec);
  REQUIRE(s1_hi && s2_hi && q1_hi && q2_hi && (!two || (s1_lo && s2_lo && q1_lo && q2_lo)),
          "flmm_wald_scan: null operand plane");
  REQUIRE(aligned16(s1_hi) && aligned16(s2_hi) && aligned16(q1_hi) && aligned16(q2_hi) &&
              (!two || (aligned16(s1_lo) && aligned16(s2_lo) && aligned16(q1_lo) && aligned16(q2_lo))),
          "flmm_wald_scan: operand planes must be 16-byte aligned");
  REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || (s1_scale && s2_scale && q1_scale && q2_scale),
          "flmm_wald_scan: split-FP16 needs the four row-scale arrays");
  REQUIRE(rWr && pmax, "flmm_wald_scan: null rWr / pmax");
  REQUIRE(n_real == 0 || (f && beta && se && sfve && nullh2 && h2 && row_index), "flmm_wald_scan: null output");
  if (mc == 0 || P == 0) return FLMM_OK;
  ScanOut out{f, beta, se, sfve, nullh2, n_real};
  ScanScales sc{nullptr, nullptr, nullptr, nullptr};
  if (prec == FLMM_PREC_SPLIT_FP16) sc = ScanScales{s1_scale, s2_scale, q1_scale, q2_scale};
  // tile variant: "wide" (128 x 256, column operand streamed to registers) needs p_pad % 256 == 0;
  // FLMM_SCAN_VARIANT=square forces the 128 x 128 LDS-staged variant
  const char* var = getenv("FLMM_SCAN_VARIANT");
  const bool wide = (p_pad % 256 == 0) && !(var && strcmp(var, "square") == 0);
  const int nch = (int)(kpad / KCHUNK);
  const hipStream_t st = (hipStream_t)stream;
  hipError_t e;
#define FLMM_SCAN(PR)                                                                                                \
  do {                                                                                                               \
    TileArgs<PR, 2> t{};                                                                                             \
    t.a[0][0] = s1_hi; t.a[0][1] = s1_lo; t.a[1][0] = s2_hi; t.a[1][1] = s2_lo;                                      \
    t.b[0][0] = q1_hi; t.b[0][1] = q1_lo; t.b[1][0] = q2_hi; t.b[1][1] = q2_lo;                                      \
    t.a_rows_pad = mc_pad; t.b_rows_pad = p_pad; t.nchunks = nch;                                                    \
    e = wide ? launch_wald_scan_wide<PR>(t, sc, mc, P, n_real, rWr, h2, nu, row_index, out, pmax, st)                \
             : launch_wald_scan<PR>(t, sc, mc, P, n_real, rWr, h2, nu, row_index, out, pmax, st);                    \
  } while (0)
  if (prec == FLMM_PREC_SPLIT_BF16) FLMM_SCAN(PREC_SPLIT_BF16);
  else if (prec == FLMM_PREC_SPLIT_FP16) FLMM_SCAN(PREC_SPLIT_FP16);
  else FLMM_SCAN(PREC_FP32);
#undef FLMM_SCAN
  return check_launch(e, "flmm_wald_scan");
}

int flmm_wald_scan_lowrank(int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi, const void* s2_lo,
                           const float* s1_scale, const float* s2_scale, int64_t mc_pad, const void* q1_hi,
                           const void* q1_lo, const void* q2_hi, const void* q2_lo, const float* q1_scale,
                           const float* q2_scale, int64_t p_pad, int64_t kpad, int64_t kpad2, int64_t mc, int64_t P,
                           int64_t n_real, const double* rWr, const double* h2, double nu, const int64_t* row_index,
                           double* f, double* beta, double* se, double* sfve, double* nullh2, double* pmax,
                           void* stream) {
  return scan_lowrank("flmm_wald_scan_lowrank", prec, s1_hi, s1_lo, s2_hi, s2_lo, s1_scale, s2_scale, mc_pad, q1_hi,
                      q1_lo, q2_hi, q2_lo, q1_scale, q2_scale, p_pad, kpad, kpad2, mc, P, n_real, rWr, h2, nu,
                      row_index, f, beta, se, sfve, nullh2, pmax, stream);
}

int flmm_wald_scan_lowrank_noalloc(int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi,
                                   const void* s2_lo, const float* s1_scale, const float* s2_scale, int64_t mc_pad,
                                   const void* q1_hi, const void* q1_lo, const void* q2_hi, const void* q2_lo,
                                   const float* q1_scale, const float* q2_scale, int64_t p_pad, int64_t kpad,
                                   int64_t kpad2, int64_t mc, int64_t P, int64_t n_real, const double* rWr,
                                   const double* h2, double nu, const int64_t* row_index, double* f, double* beta,
                                   double* se, double* sfve, double* nullh2, double* max_F, void* stream) {
  const char* who = "flmm_wald_scan_lowrank_noalloc";
  REQUIRE(mc >= 0 && P >= 0 && p_pad >= 0, "%s: bad sizes", who);
  REQUIRE(max_F, "%s: null max_F", who);
  const int64_t n_mt = (mc + TILE - 1) / TILE;
  const int64_t need = n_mt * p_pad * (int64_t)sizeof(double);  // the kernel's per-tile maxima: n_mt x p_pad
  REQUIRE(need <= FLMM_NOALLOC_SCRATCH_BYTES,
          "%s: %lld SNP tiles x p_pad %lld need %lld bytes of scratch, above the maximum of %lld bytes; split the SNP "
          "block (ceil(mc / 128) x p_pad x 8 <= %lld)",
          who, (long long)n_mt, (long long)p_pad, (long long)need, (long long)FLMM_NOALLOC_SCRATCH_BYTES,
          (long long)FLMM_NOALLOC_SCRATCH_BYTES);
  if (mc == 0 || P == 0) return scan_lowrank(who, prec, s1_hi, s1_lo, s2_hi, s2_lo, s1_scale, s2_scale, mc_pad, q1_hi,
                                             q1_lo, q2_hi, q2_lo, q1_scale, q2_scale, p_pad, kpad, kpad2, mc, P,
                                             n_real, rWr, h2, nu, row_index, f, beta, se, sfve, nullh2, max_F,
                                             stream);
  const hipStream_t st = (hipStream_t)stream;
  double* pm = scratch_ptr(st, who);
  if (!pm) return g_err.empty() ? fail(FLMM_ERR_HIP, "%s: no scratch", who) : FLMM_ERR_HIP;
  // every (tile, column < P) entry of the maxima is written by this call's scan before the fold reads it, so the
  // scratch needs no clearing; calls that overlap on one GPU share the array and race
  const int rc = scan_lowrank(who, prec, s1_hi, s1_lo, s2_hi, s2_lo, s1_scale, s2_scale, mc_pad, q1_hi, q1_lo, q2_hi,
                              q2_lo, q1_scale, q2_scale, p_pad, kpad, kpad2, mc, P, n_real, rWr, h2, nu, row_index, f,
                              beta, se, sfve, nullh2, pm, stream);
  if (rc != FLMM_OK) return rc;
  return check_launch(launch_max_reduce(pm, n_mt, p_pad, P, max_F, st), who);
}

// ---- certified screen of the permutation columns (screen.hip) ----
// Error constants of the bound stated in screen.hip.  u': fp16 rounding of an FP32 value within 2^-21 (relative) of
// the FP64 one (the SNP planes round once; the screened columns' hi planes come from FP32 arithmetic, ~5 rounding
// units); gamma(n): FP32
// accumulation of n terms; SUBNORMAL_EPS: the largest absolute error of an fp16 operand below the normal range
// (2^-25: the conversions and v_mfma_f32_32x32x16_f16 keep fp16 subnormals on gfx950; 2^-14 would cover flushing);
// EXPSUM_REL: the factored weights' relative accuracy (1e-9 target, expsum.py); ETA: the subnormal share allowed in
// the second product's relative bound, which sets dmin; SLACK: the FP32 evaluation of x = A^2 g / D (a few ulp).
static constexpr double SC_U = 1.0 / 2048.0 + 1.0 / 2097152.0;
static constexpr double SUBNORMAL_EPS = 1.0 / 33554432.0;
static constexpr double EXPSUM_REL = 1e-8;
static constexpr double SC_ETA = 1.0 / 1048576.0;
static constexpr double SC_SLACK = 1.0 / 524288.0;
static double sc_gamma(double n) { return n / 16777216.0 / (1.0 - n / 16777216.0); }

// phases: 1 = the screen kernel, 2 = the resolution kernel
int64_t flmm_screen_order_bytes(int64_t n_cols) {
  if (n_cols < 0) return -1;
  return 4 * screen_order_ints(n_cols);
}

int64_t flmm_screen_q_rows(int64_t p_pad, int64_t hi_from, int64_t P) {
  if (p_pad < 0 || hi_from < 0 || P < 0 || P > p_pad) return -1;
  if (hi_from >= P) return p_pad;
  const int64_t need = (hi_from + screen_order_slots(P - hi_from) + 255) / 256 * 256;
  return need > p_pad ? need : p_pad;
}

int64_t flmm_screen_work_bytes(int64_t mc_pad, int64_t kpad) {
  if (mc_pad < 0 || kpad < 0) return -1;
  return 8 * (int64_t)SO_NODES * ((kpad + 31) / 32 * 32 + (mc_pad + 31) / 32 * 32);
}

static int screened(const char* who, int phases, const void* s1_hi, const void* s2_hi, const float* s1_scale,
                    const float* s2_scale, int64_t mc_pad, const void* q1_hi, const void* q2_hi, const float* q1_scale,
                    const float* q2_scale, int64_t p_pad, int64_t q_rows, int64_t kpad, int64_t kpad2, int64_t mc,
                    int64_t p_first, int64_t P, const double* us, int64_t ld_us, const float* uy, const double* s,
                    int64_t neff, double s_lo, double s_hi, const double* log_delta, const double* rWr, double nu,
                    double* max_F, void* rec, const void* order, void* work, unsigned long long* n_eval,
                    void* stream) {
  REQUIRE(mc >= 0 && mc <= mc_pad && mc_pad % TILE == 0 && mc <= 64 * TILE, "%s: bad SNP extent (mc <= 8192)", who);
  REQUIRE(P >= 0 && P <= p_pad && p_pad % 256 == 0 && p_first >= 0 && p_first % 256 == 0 && p_first <= P,
          "%s: bad column extent (p_pad and p_first multiples of 256, p_first <= P <= p_pad)", who);
  REQUIRE(q_rows >= p_pad && q_rows % 256 == 0, "%s: q_rows must be >= p_pad and a multiple of 256", who);
  REQUIRE(mc_pad <= (1LL << 24) && q_rows <= (1LL << 24), "%s: rows_pad above 2^24", who);
  REQUIRE(kpad > 0 && kpad % TILE == 0 && kpad2 > 0 && kpad2 % TILE == 0 && kpad2 <= kpad,
          "%s: need 0 < kpad2 <= kpad, both multiples of 128", who);
  REQUIRE(neff >= 0 && neff <= kpad && ld_us >= kpad, "%s: need neff <= kpad <= ld_us", who);
  REQUIRE(nu > 0.0 && s_lo <= s_hi, "%s: need nu > 0 and s_lo <= s_hi", who);
  REQUIRE(s1_hi && s2_hi && q1_hi && q2_hi && s1_scale && s2_scale && q1_scale && q2_scale,
          "%s: null plane or scale", who);
  REQUIRE(aligned16(s1_hi) && aligned16(s2_hi) && aligned16(q1_hi) && aligned16(q2_hi),
          "%s: operand planes must be 16-byte aligned", who);
  REQUIRE(us && uy && s && log_delta && rWr && max_F && rec, "%s: null pointer", who);
  const bool bucketed = order != nullptr;
  REQUIRE(!bucketed || work, "%s: the bucketed layout needs the workspace (flmm_screen_work_bytes)", who);
  REQUIRE(!bucketed || q_rows >= p_first + screen_order_slots(P - p_first),
          "%s: q_rows below flmm_screen_q_rows for the bucketed layout", who);
  if (mc == 0 || P == p_first) return FLMM_OK;
  const int64_t Pl = P - p_first;
  const double u = SC_U, K = (double)kpad, K2 = (double)kpad2, e = SUBNORMAL_EPS;
  const double kA = 2 * u + u * u + sc_gamma(2 * K) * (1 + 2 * u);
  // subnormal part of the first product, relative to s q sqrt(d rWr) >= 2^26 / R (rows scaled to (2^13, 2^14])
  const double kArel = (e * 16384.0 * (sqrt(K) + K) * (1 + u) * (1 + u) + K * e * e) / 67108864.0;
  const double kD = 2 * u + u * u + sc_gamma(2 * K2) * (1 + 2 * u) + EXPSUM_REL + SC_ETA + SC_SLACK;
  // the bucketed tile's sWs: 1 / sWs interpolated within 3.1 e^h 0.0481 h^3 / 6 (h <= 1/5), d~ within that over
  // (1 - it), plus the FP32 evaluation (SC_SLACK): inside kD
  static_assert(3.1 * 1.2215 * 0.0481 * 0.008 / 6.0 * 1.001 + SC_SLACK < 2 * SC_U, "bucketed sWs error within kD");
  // second product: subnormal error <= e (sum Ch + sum Bh) + K2 e^2 with both rows' entries <= 2^14 (sum <= K2 2^14
  // covers the max-scaled exact-weight planes too); a row whose scaled D is below that / ETA is a forced candidate
  const float dmin = (float)((e * 16384.0 * 2.0 * K2 * (1 + u) + K2 * e * e) / SC_ETA);
  const hipStream_t st = (hipStream_t)stream;
  uint64_t* r = (uint64_t*)rec;
  if (phases & 1) {
    if (bucketed) {
      // every screened column on the sWy-only tile, sWs from the node sums of its log-delta bucket
      const ScreenOrder o = screen_order((void*)order, Pl);
      double* wnode = (double*)work;
      double* dnode = wnode + (kpad + 31) / 32 * 32 * SO_NODES;
      int rc = check_launch(launch_screen_nodes(us, ld_us, mc, s, neff, kpad, order, wnode, dnode, st), who);
      if (rc != FLMM_OK) return rc;
      BucketArgs g{};
      g.s1 = (const u16*)s1_hi;
      g.q1 = (const u16*)q1_hi + p_first * 32;
      g.s1s = s1_scale;
      g.q2s = q2_scale + p_first;
      g.log_delta = log_delta + p_first;
      g.dnode = dnode;
      g.mc = mc;
      g.head = o.head;
      g.tb = o.tb;
      g.order = o.order;
      g.gbucket = o.gbucket;
      g.a_rows_pad = mc_pad;
      g.b_rows_pad = q_rows;
      g.nchunks = (int)(kpad / KCHUNK);
      rc = check_launch(launch_screen_bucketed(g, mc, Pl, r, Pl, st), who);
      if (rc != FLMM_OK) return rc;
    } else {
      ScreenArgs g{};
      g.s1 = (const u16*)s1_hi;
      g.s2 = (const u16*)s2_hi;
      g.q1 = (const u16*)q1_hi + p_first * 32;
      g.q2 = (const u16*)q2_hi + p_first * 32;
      g.s1s = s1_scale;
      g.s2s = s2_scale;
      g.a_rows_pad = mc_pad;
      g.b_rows_pad = q_rows;
      g.nchunks = (int)(kpad / KCHUNK);
      g.nchunks2 = (int)(kpad2 / KCHUNK);
      const int rc = check_launch(launch_wald_screen(g, mc, Pl, dmin, r, Pl, st), who);
      if (rc != FLMM_OK) return rc;
    }
  }
  if (phases & 2)
    return check_launch(launch_screen_resolve(r, Pl, mc, Pl, p_first, q1_scale, q2_scale, rWr, log_delta, us, ld_us,
                                              uy, p_pad, s, neff, s_lo, s_hi, nu, kA, kArel, kD, max_F, n_eval, st),
                        who);
  return FLMM_OK;
}

#define FLMM_SCREENED_ARGS                                                                                           \
  s1_hi, s2_hi, s1_scale, s2_scale, mc_pad, q1_hi, q2_hi, q1_scale, q2_scale, p_pad, q_rows, kpad, kpad2, mc,     \
      p_first, P, us, ld_us, uy, s, neff, s_lo, s_hi, log_delta, rWr, nu, max_F, rec, order, work, n_eval, stream
#define FLMM_SCREENED_PARAMS                                                                                         \
  const void *s1_hi, const void *s2_hi, const float *s1_scale, const float *s2_scale, int64_t mc_pad,                \
      const void *q1_hi, const void *q2_hi, const float *q1_scale, const float *q2_scale, int64_t p_pad,             \
      int64_t q_rows, int64_t kpad, int64_t kpad2, int64_t mc, int64_t p_first, int64_t P, const double *us,         \
      int64_t ld_us, const float *uy, const double *s, int64_t neff, double s_lo, double s_hi,                       \
      const double *log_delta, const double *rWr, double nu, double *max_F, void *rec, const void *order,            \
      void *work, unsigned long long *n_eval, void *stream

int flmm_wald_scan_screened(FLMM_SCREENED_PARAMS) {
  return screened("flmm_wald_scan_screened", 3, FLMM_SCREENED_ARGS);
}
int flmm_wald_screen(FLMM_SCREENED_PARAMS) { return screened("flmm_wald_screen", 1, FLMM_SCREENED_ARGS); }
int flmm_screen_resolve(FLMM_SCREENED_PARAMS) { return screened("flmm_screen_resolve", 2, FLMM_SCREENED_ARGS); }
#undef FLMM_SCREENED_ARGS
#undef FLMM_SCREENED_PARAMS

int64_t flmm_scan_table_bytes(int n_groups) {
  if (n_groups < 0) return -1;
  return (int64_t)(4 + n_groups + 1) * 8 + (int64_t)n_groups * (int64_t)sizeof(flmm_scan_group);
}

int flmm_scan_table_fill(int prec, int variant, const flmm_scan_group* groups, int n_groups, void* host_table) {
  REQUIRE(valid_prec(prec), "flmm_scan_table_fill: bad precision %d", prec);
  REQUIRE(variant == FLMM_SCAN_SQUARE || variant == FLMM_SCAN_WIDE || variant == FLMM_SCAN_WIDE_FACTORED,
          "flmm_scan_table_fill: bad variant %d", variant);
  REQUIRE(n_groups >= 0 && n_groups <= 65535, "flmm_scan_table_fill: n_groups must be in [0, 65535]");
  REQUIRE(host_table && (groups || n_groups == 0), "flmm_scan_table_fill: null pointer");
  const bool two = two_planes(prec);
  const int64_t tp = variant == FLMM_SCAN_SQUARE ? TILE : 256;
  int64_t* hdr = (int64_t*)host_table;
  int64_t total = 0;
  for (int g = 0; g < n_groups; ++g) {
    const flmm_scan_group& j = groups[g];
    REQUIRE(j.mc >= 0 && j.mc <= j.mc_pad && j.mc_pad % TILE == 0, "flmm_scan_table_fill: group %d: bad SNP extent", g);
    REQUIRE(j.P >= 0 && j.P <= j.p_pad && j.p_pad % tp == 0,
            "flmm_scan_table_fill: group %d: bad column extent (p_pad %% %lld required)", g, (long long)tp);
    REQUIRE(j.kpad > 0 && j.kpad % TILE == 0, "flmm_scan_table_fill: group %d: kpad must be a positive multiple of 128",
            g);
    if (variant == FLMM_SCAN_WIDE_FACTORED)
      REQUIRE(j.kpad2 > 0 && j.kpad2 % TILE == 0 && j.kpad2 < j.kpad,
              "flmm_scan_table_fill: group %d: the factored variant needs kpad2 a multiple of 128 in (0, kpad)", g);
    else
      REQUIRE(j.kpad2 == 0, "flmm_scan_table_fill: group %d: kpad2 > 0 needs the factored variant", g);
    REQUIRE(j.n_real >= 0 && j.n_real <= j.P, "flmm_scan_table_fill: group %d: n_real must be in [0, P]", g);
    REQUIRE(j.nu > 0.0, "flmm_scan_table_fill: group %d: dof must be positive", g);
    REQUIRE(j.s1_hi && j.s2_hi && j.q1_hi && j.q2_hi && (!two || (j.s1_lo && j.s2_lo && j.q1_lo && j.q2_lo)),
            "flmm_scan_table_fill: group %d: null operand plane", g);
    REQUIRE(aligned16(j.s1_hi) && aligned16(j.s2_hi) && aligned16(j.q1_hi) && aligned16(j.q2_hi) &&
                (!two || (aligned16(j.s1_lo) && aligned16(j.s2_lo) && aligned16(j.q1_lo) && aligned16(j.q2_lo))),
            "flmm_scan_table_fill: group %d: operand planes must be 16-byte aligned", g);
    REQUIRE(prec != FLMM_PREC_SPLIT_FP16 || (j.s1_scale && j.s2_scale && j.q1_scale && j.q2_scale),
            "flmm_scan_table_fill: group %d: split-FP16 needs the four row-scale arrays", g);
    REQUIRE(j.rWr && j.pmax && (j.max_F || j.P == 0), "flmm_scan_table_fill: group %d: null rWr / pmax / max_F", g);
    REQUIRE(j.n_real == 0 || (j.f && j.beta && j.se && j.sfve && j.nullh2 && j.h2 && j.row_index),
            "flmm_scan_table_fill: group %d: null output", g);
    hdr[4 + g] = total;
    total += ((j.mc + TILE - 1) / TILE) * ((j.P + tp - 1) / tp);
  }
  hdr[0] = n_groups;
  hdr[1] = total;
  hdr[2] = variant;
  hdr[3] = 0;
  hdr[4 + n_groups] = total;
  if (n_groups > 0) memcpy(hdr + 4 + n_groups + 1, groups, (size_t)n_groups * sizeof(flmm_scan_group));
  g_err.clear();
  return FLMM_OK;
}

int flmm_wald_scan_grouped(int prec, int variant, const void* table, int blocks, void* stream) {
  REQUIRE(valid_prec(prec), "flmm_wald_scan_grouped: bad precision %d", prec);
  REQUIRE(variant == FLMM_SCAN_SQUARE || variant == FLMM_SCAN_WIDE || variant == FLMM_SCAN_WIDE_FACTORED,
          "flmm_wald_scan_grouped: bad variant %d", variant);
  REQUIRE(table && aligned16(table), "flmm_wald_scan_grouped: table must be a 16-byte aligned device pointer");
  REQUIRE(blocks > 0 && blocks % 8 == 0, "flmm_wald_scan_grouped: blocks must be a positive multiple of 8, got %d",
          blocks);
  return check_launch(launch_wald_scan_grouped(prec, variant, (const int64_t*)table, blocks, (hipStream_t)stream),
                      "flmm_wald_scan_grouped");
}

int flmm_max_reduce_grouped(const void* table, int n_groups, int64_t max_P, void* stream) {
  REQUIRE(table, "flmm_max_reduce_grouped: null table");
  REQUIRE(n_groups >= 0 && n_groups <= 65535 && max_P >= 0, "flmm_max_reduce_grouped: bad sizes");
  if (n_groups == 0 || max_P == 0) return FLMM_OK;
  return check_launch(launch_max_reduce_grouped((const int64_t*)table, n_groups, max_P, (hipStream_t)stream),
                      "flmm_max_reduce_grouped");
}

int flmm_standardise(int dtype, const void* src, int64_t n, int64_t m, int64_t ld_src, double* dst, int64_t ld_dst,
                     double* mean, double* std, void* stream) {
  REQUIRE(dtype == FLMM_DTYPE_F64 || dtype == FLMM_DTYPE_I8, "flmm_standardise: bad dtype %d", dtype);
  REQUIRE(n >= 0 && m >= 0 && ld_src >= m && ld_dst >= m, "flmm_standardise: bad sizes");
  REQUIRE((src && dst) || n == 0 || m == 0, "flmm_standardise: null pointer");
  if (n == 0 || m == 0) return FLMM_OK;
  return check_launch(launch_standardise(dtype, src, n, m, ld_src, dst, ld_dst, mean, std, (hipStream_t)stream),
                      "flmm_standardise");
}

int flmm_max_reduce(const double* pmax, int64_t n_tiles, int64_t ld, int64_t P, double* max_F, void* stream) {
  REQUIRE(n_tiles >= 0 && P >= 0 && ld >= P, "flmm_max_reduce: bad sizes");
  REQUIRE((pmax || n_tiles == 0) && (max_F || P == 0), "flmm_max_reduce: null pointer");
  if (P == 0 || n_tiles == 0) return FLMM_OK;
  return check_launch(launch_max_reduce(pmax, n_tiles, ld, P, max_F, (hipStream_t)stream), "flmm_max_reduce");
}

int flmm_neglog10p_f1(const double* F, int64_t n, double df2, double* out, void* stream) {
  REQUIRE(n >= 0 && df2 > 0.0, "flmm_neglog10p_f1: n >= 0 and df2 > 0 required");
  REQUIRE((F && out) || n == 0, "flmm_neglog10p_f1: null pointer");
  if (n == 0) return FLMM_OK;
  return check_launch(launch_neglog10p(F, n, df2, out, (hipStream_t)stream), "flmm_neglog10p_f1");
}

int flmm_effect_size(const double* beta, int64_t M, int64_t n_real, const double* g_var, const double* p_var,
                     double* out, void* stream) {
  REQUIRE(M >= 0 && n_real >= 0, "flmm_effect_size: bad sizes");
  REQUIRE((beta && g_var && p_var && out) || M * n_real == 0, "flmm_effect_size: null pointer");
  if (M * n_real == 0) return FLMM_OK;
  return check_launch(launch_effect_size(beta, M, n_real, g_var, p_var, out, (hipStream_t)stream),
                      "flmm_effect_size");
}

}  // extern "C"
