// Host-side launch wrappers of the gfx950 kernels (pack.hip, rotate.hip, delta.hip, scan.hip, scan2.hip,
// standardise.hip, pvalue.hip, mlp.hip).
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

// per-row power-of-two scales of the four scan planes (split-FP16; all NULL otherwise)
struct ScanScales {
  const float* s1;
  const float* s2;
  const float* q1;
  const float* q2;
};

hipError_t launch_pack_planes(int prec, const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride,
                              int64_t k_stride, const int64_t* row_index, const double* row_norm, void* out_a,
                              void* out_b, float* row_scale, int64_t rows_pad, int64_t kpad, hipStream_t stream);
hipError_t launch_pack_columns_projected(int prec, const double* y, int64_t N, int64_t P, const double* x, int C,
                                         const double* b, double* norm, void* out_a, void* out_b, float* row_scale,
                                         int64_t rows_pad, int64_t kpad, hipStream_t stream);
hipError_t launch_row_norms(const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                            const int64_t* row_index, double* norm, hipStream_t stream);
template <int PREC>
hipError_t launch_rotate(int mode, const TileArgs<PREC, 1>& args, const float* a_scale, const float* b_scale,
                         void* o0, void* o1, void* o2, void* o3, float* out_scale1, float* out_scale2,
                         hipStream_t stream);
hipError_t launch_delta_weights(const double* s, int64_t neff, const double* s_lam, int64_t nlam, int64_t kpad, int G,
                                double lo, double hi, float* wplane, float* lam, double* lam64, hipStream_t stream);
hipError_t launch_delta_grid_split(const float* wplane, const float* lam, int G, double nco, const float* uy,
                                   const float* col_scale, double s_min, double lo, double hi, int64_t P, int64_t p_pad,
                                   int64_t kpad, int* idx, void* wsplit, hipStream_t stream);
hipError_t launch_delta_grid(const float* wplane, const float* lam, int G, double nco, const float* uy, int64_t P,
                             int64_t p_pad, int64_t kpad, int* idx, hipStream_t stream);
hipError_t launch_delta_refine(int prec, int mode, const float* uy, int64_t P, int64_t p_pad, int64_t kpad,
                               const double* s, int64_t neff, const double* s_lam, int64_t nlam, double nco, int G,
                               double lo, double hi, const int* idx, const double* lam64, double* log_delta,
                               double* rWr, double* h2,
                               void* q1a, void* q1b, void* q2a, void* q2b, float* q1s, float* q2s, double radius,
                               double judge_tol, int* flags, int64_t hi_from, int64_t q_rows, double t_lo,
                               double t_h, int nb, void* order_ws, hipStream_t stream);
hipError_t launch_gemm_f64(const double* x2, int64_t rows, int64_t kpad, const double* A, int64_t rpad, double* C,
                           hipStream_t stream);
hipError_t launch_select_above(const double* f, int64_t rows, int64_t cols, int64_t ld, double thr, int64_t* out_row,
                               int64_t* out_col, int64_t cap, unsigned long long* count, hipStream_t stream);
hipError_t launch_rescore(const double* US, const double* UY, int64_t ld, const double* s, int64_t neff,
                          const double* log_delta, const int64_t* us_row, const int64_t* col, const int64_t* out_row,
                          int64_t out_row_base, int64_t n, const unsigned long long* n_dev, double nu, double* f,
                          double* beta, double* se, double* sfve, int64_t out_ld, hipStream_t stream);
template <int PREC>
hipError_t launch_gram(const TileArgs<PREC, 1>& args, const float* row_scale, int n_split, double* partial, int64_t n,
                       double* out, int64_t ld, hipStream_t stream);
int gram_tile_pairs(int64_t rows_pad);
hipError_t launch_rotate_f64(const double* z, int64_t ldz, const int64_t* col_index, int64_t mc, int64_t n,
                             const double* u, int64_t ldu, int64_t neff, double* out, int64_t rows_pad, int64_t kpad,
                             hipStream_t stream);
hipError_t launch_gram_f64(const double* Z, int64_t ldz, int64_t n, int64_t m, int n_split, double* partial,
                           double* out, int64_t ld, hipStream_t stream);
hipError_t launch_factored_planes(int prec, int64_t P, int64_t p_pad, int64_t rpad, const double* t, const double* lam,
                                  int64_t R, const double* log_delta, void* qa, void* qb, float* qs, int64_t hi_from,
                                  bool layout, hipStream_t stream);
template <int PREC>
hipError_t launch_wald_scan(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P,
                            int64_t n_real, const double* rWr, const double* h2, double nu, const int64_t* row_index,
                            const ScanOut& out, double* pmax, hipStream_t stream);
template <int PREC>
hipError_t launch_wald_scan_wide(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P,
                                 int64_t n_real, const double* rWr, const double* h2, double nu,
                                 const int64_t* row_index, const ScanOut& out, double* pmax, hipStream_t stream);
template <int PREC>
hipError_t launch_wald_scan_lowrank(const TileArgs<PREC, 2>& args, const ScanScales& sc, int64_t Mc, int64_t P,
                                    int64_t n_real, const double* rWr, const double* h2, double nu,
                                    const int64_t* row_index, const ScanOut& out, double* pmax, hipStream_t stream);
// certified screen of the permutation columns (screen.hip)
struct ScreenArgs {
  const u16* s1;       // SNP-side hi planes: S1 = US over nchunks K-chunks, S2 over nchunks2, rows a_rows_pad
  const u16* s2;
  const u16* q1;       // column-side hi planes from the first screened column on (row stride b_rows_pad)
  const u16* q2;
  const float* s1s;    // the SNP rows' power-of-two scales
  const float* s2s;
  int64_t a_rows_pad;
  int64_t b_rows_pad;
  int nchunks;         // K / 32 of the first product (sWy)
  int nchunks2;        // K / 32 of the second (sWs): < nchunks for factored weights, == nchunks for exact ones
};
// The screened columns' layout for the bucketed screen (int32 workspace of screen_order_ints(n) ints, n the
// screened columns): column i's log delta t falls in bucket floor((t - t_lo) / t_h) (clamped to [0, nb)); bucket b's
// columns hold plane slots bstart[b] .. bstart[b] + bcount[b] - 1, bstart[b] a multiple of SO_PAD, the slots up to
// bstart[b + 1] are padding (order = -1); head[0] = the slot count S = bstart[nb], head[1] = nb, tb = {t_lo, t_h};
// order[slot] = column, pos[column] = slot, gbucket[g] = the bucket of slots SO_PAD g .. SO_PAD g + SO_PAD - 1
// (columns and slots counted from the first screened column).  The order within a bucket is not fixed; nothing
// depends on it.
constexpr int SO_NB_MAX = 120;  // buckets (2 SO_NB_MAX + 1 nodes fit SO_NODES)
constexpr int SO_PAD = 32;     // slot alignment of a bucket: one MFMA column block
__host__ __device__ inline int64_t screen_order_up64(int64_t n) { return (n + 63) / 64 * 64; }
__host__ __device__ inline int64_t screen_order_slots(int64_t n) { return n + (int64_t)SO_PAD * SO_NB_MAX; }  // the most slots S
__host__ __device__ inline int64_t screen_order_nblk(int64_t n) { return (n + 1023) / 1024; }
struct ScreenOrder {
  int* head;     // [0] S, [1] nb
  double* tb;    // [0] t_lo, [1] t_h
  int* bcount;   // SO_NB_MAX
  int* bstart;   // SO_NB_MAX + 1
  int* order;    // screen_order_slots(n)
  int* pos;      // n
  int* gbucket;  // screen_order_slots(n) / SO_PAD
  int* blk;      // per 1,024-column block and bucket: counts, then offsets (nblk x SO_NB_MAX)
};
__host__ __device__ inline int64_t screen_order_ints(int64_t n) {
  return 64 + 128 + 128 + screen_order_up64(screen_order_slots(n)) + screen_order_up64(n) +
         screen_order_up64(screen_order_slots(n) / SO_PAD) + screen_order_up64(screen_order_nblk(n) * SO_NB_MAX);
}
__host__ __device__ inline ScreenOrder screen_order(void* ws, int64_t n) {
  int* p = (int*)ws;
  ScreenOrder o;
  o.head = p;
  o.tb = (double*)(p + 8);
  o.bcount = p + 64;
  o.bstart = o.bcount + 128;
  o.order = o.bstart + 128;
  o.pos = o.order + screen_order_up64(screen_order_slots(n));
  o.gbucket = o.pos + screen_order_up64(n);
  o.blk = o.gbucket + screen_order_up64(screen_order_slots(n) / SO_PAD);
  return o;
}
// The bucketed screen tile: every screened column of bucket b has its sWs of SNP row m interpolated from the FP64
// sums dnode[m][2b + j] = sum_k US[m][k]^2 / (s_k + exp(t_lo + (2b + j) t_h / 2)), j = 0, 1, 2 (1 / sWs quadratic
// in log delta), so only the sWy product runs on MFMA.
constexpr int SO_NODES = 256;  // node columns of dnode (2 SO_NB_MAX + 1 used at most)
static_assert(2 * SO_NB_MAX + 1 <= SO_NODES, "the nodes fit dnode's columns");
struct BucketArgs {
  const u16* s1;         // SNP-side hi planes of the first product, rows a_rows_pad
  const u16* q1;         // column-side hi planes, fragment-native layout by slot (frag_off), row stride b_rows_pad
  const float* s1s;      // the SNP rows' power-of-two scales
  const float* q2s;      // the columns' second-product scales (from the first screened column on)
  const double* log_delta;  // the columns' log delta (from the first screened column on)
  const double* dnode;   // (rows, SO_NODES) FP64
  int64_t mc;            // SNP rows
  const int* head;       // ScreenOrder
  const double* tb;
  const int* order;
  const int* gbucket;
  int64_t a_rows_pad;
  int64_t b_rows_pad;
  int nchunks;
};
hipError_t launch_screen_classify(const double* log_delta, int64_t n, double t_lo, double t_h, int nb,
                                  void* order_ws, hipStream_t stream);
// dnode of the bucketed screen for SNP rows m < mc of us (row stride ld_us, k < neff): node weights into
// wnode (kpad x SO_NODES FP64), then dnode (rup(mc, 32) x SO_NODES FP64) from the layout's t_lo, t_h, nb
hipError_t launch_screen_nodes(const double* us, int64_t ld_us, int64_t mc, const double* s, int64_t neff,
                               int64_t kpad, const void* order_ws, double* wnode, double* dnode, hipStream_t stream);
hipError_t launch_screen_bucketed(const BucketArgs& g, int64_t Mc, int64_t P, uint64_t* rec, int64_t rec_ld,
                                  hipStream_t stream);
hipError_t launch_wald_screen(const ScreenArgs& g, int64_t Mc, int64_t P, float dmin, uint64_t* rec, int64_t rec_ld,
                              hipStream_t stream);
hipError_t launch_screen_resolve(const uint64_t* rec, int64_t rec_ld, int64_t Mc, int64_t P, int64_t p_first,
                                 const float* q1s, const float* q2s, const double* rWr, const double* log_delta,
                                 const double* US, int64_t ld_us, const float* uy, int64_t p_pad, const double* s,
                                 int64_t neff, double s_min, double s_max, double nu, double kA, double kArel,
                                 double kD, double* max_F, unsigned long long* n_eval, hipStream_t stream);
hipError_t launch_max_reduce(const double* pmax, int64_t n_mtiles, int64_t ld, int64_t P, double* max_F,
                             hipStream_t stream);
hipError_t launch_standardise(int dtype, const void* src, int64_t n, int64_t m, int64_t lds, double* dst, int64_t ldd,
                              double* mean_out, double* std_out, hipStream_t stream);
hipError_t launch_neglog10p(const double* F, int64_t n, double df2, double* out, hipStream_t stream);
hipError_t launch_effect_size(const double* beta, int64_t M, int64_t n_real, const double* g_var,
                              const double* p_var, double* out, hipStream_t stream);
hipError_t launch_mlp_features(const float* uy, int64_t P, int64_t p_pad, int64_t neff, const int* bin_of_k,
                               float logN, float C, float* feat, hipStream_t stream);
hipError_t launch_mlp_dense_fwd(const float* in, int64_t B, int n_in, const float* W, const float* bias, int n_out,
                                int act, float* out, hipStream_t stream);
hipError_t launch_mlp_dense_bwd(const float* dout_in, float* dz, const float* out, const float* in, int64_t B,
                                int n_in, int n_out, int act, const float* W, float* din, float* dW, float* db,
                                hipStream_t stream);
hipError_t launch_mlp_mse_grad(const float* pred, const float* target, int64_t B, float* dpred, hipStream_t stream);
hipError_t launch_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                       float eps, int t, hipStream_t stream);

// scan epilogue: unscaled (sWy, sWs) of accumulator pair (a, d) for SNP row m and column p
__device__ inline void unscale_sums(const ScanScales& sc, int64_t m, double inv_q1, double inv_q2, double& a,
                                    double& d) {
  if (sc.s1) {
    a *= inv_q1 / (double)gmem(sc.s1)[m];
    d *= inv_q2 / (double)gmem(sc.s2)[m];
  }
}

// grouped scans (scan_grouped.hip); the table layout is described in include/fasterlmm_amd.h
hipError_t launch_wald_scan_grouped(int prec, int variant, const int64_t* table, int blocks, hipStream_t stream);
hipError_t launch_max_reduce_grouped(const int64_t* table, int n_groups, int64_t max_P, hipStream_t stream);

// .bed decoding (bed.hip)
enum BedOut : int { BED_I8 = 0, BED_F64 = 1, BED_F32 = 2 };
hipError_t launch_bed_decode(int out_kind, const uint8_t* bed, int64_t rec_bytes, const int64_t* row_idx, int64_t n,
                             const int64_t* col_idx, int64_t m, bool count_a1, void* out, int64_t ld,
                             double* mean_out, double* std_out, hipStream_t stream);

}  // namespace flmm
