/* fasterlmm_amd.h -- C ABI of the MI355X (gfx950) LMM association scan.
 *
 * Plain pointers and sizes only: every array argument is a device pointer, every stream argument is a
 * hipStream_t passed as void* (NULL = the default stream).  Calls are asynchronous on that stream, allocate
 * nothing, and never synchronise the host.  Each returns FLMM_OK or an error code; flmm_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Operand planes.  MFMA operands are "planes": a rows x K matrix stored in K-chunks of 32, chunk-major,
 * element (row, k) at ((k / 32) * rows_pad + row) * 32 + k % 32.  rows_pad is a multiple of 128 and K a
 * multiple of 128 (the padded extents are zero).  Precisions:
 *   FLMM_PREC_SPLIT_FP16  (hi, lo) pairs of IEEE fp16 bits on rows scaled by a power of two (one float per row,
 *                         "row scale"), x * scale ~= hi + lo to ~22 significant bits;
 *   FLMM_PREC_SPLIT_BF16  (hi, lo) pairs of bf16 bits, x ~= hi + lo to ~16 significant bits, no scales;
 *   FLMM_PREC_FP32        one float plane (pass NULL for lo and for the scales).
 * A product of two split operands is hi*hi + hi*lo + lo*hi on MFMA with FP32 accumulation.
 *
 * The reference (gbrach/FaST-ER-LMM v1.6.2) has no native boundary: its hot path is a set of Python
 * functions on torch tensors.  Each entry point below names the reference function whose arithmetic it
 * replaces; the Python layer in fast-er-lmm_amd/fasterlmm_amd keeps the reference's function names and
 * signatures on top of these.
 */
#ifndef FASTERLMM_AMD_H
#define FASTERLMM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum flmm_status { FLMM_OK = 0, FLMM_ERR_ARG = 1, FLMM_ERR_HIP = 2 };
enum flmm_precision { FLMM_PREC_SPLIT_BF16 = 0, FLMM_PREC_FP32 = 1, FLMM_PREC_SPLIT_FP16 = 2 };
enum flmm_rotate_mode { FLMM_ROTATE_UY = 0, FLMM_ROTATE_US = 1, FLMM_ROTATE_US_X2 = 2 };
enum flmm_dtype { FLMM_DTYPE_F64 = 0, FLMM_DTYPE_I8 = 1 };
enum flmm_delta_mode { FLMM_DELTA_GOLDEN = 0, FLMM_DELTA_NEWTON = 1, FLMM_DELTA_FIXED = 2, FLMM_DELTA_GRID = 4 };
enum flmm_bed_out { FLMM_BED_I8 = 0, FLMM_BED_F64 = 1, FLMM_BED_F32 = 2 };
enum flmm_scan_variant { FLMM_SCAN_SQUARE = 0, FLMM_SCAN_WIDE = 1, FLMM_SCAN_WIDE_FACTORED = 2 };

/* One scan of a grouped launch (flmm_wald_scan_grouped): the arguments of flmm_wald_scan for one SNP block
 * against one column set, plus that column set's per-tile maxima (ceil(mc / 128) x p_pad doubles, scratch of the
 * launch) and its running genome-wide maximum max_F (P doubles, updated by flmm_max_reduce_grouped).  All
 * pointers are device pointers; the layout is fixed (30 eight-byte fields). */
typedef struct flmm_scan_group {
  const void* s1_hi;
  const void* s1_lo;
  const void* s2_hi;
  const void* s2_lo;
  const float* s1_scale;
  const float* s2_scale;
  int64_t mc_pad;
  const void* q1_hi;
  const void* q1_lo;
  const void* q2_hi;
  const void* q2_lo;
  const float* q1_scale;
  const float* q2_scale;
  int64_t p_pad;
  int64_t kpad;
  int64_t kpad2; /* 0; FLMM_SCAN_WIDE_FACTORED: the second product runs over the first kpad2 < kpad */
  int64_t mc;
  int64_t P;
  int64_t n_real;
  const double* rWr;
  const double* h2;
  double nu;
  const int64_t* row_index;
  double* f;
  double* beta;
  double* se;
  double* sfve;
  double* nullh2;
  double* pmax;
  double* max_F;
} flmm_scan_group;

/* ABI version of this header (bumped on any signature change). */
int flmm_abi_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* flmm_last_error(void);

/* Grouped scans in one persistent launch (no counterpart in the reference, which loops over missingness groups
 * with one loco_scan_compat each: fasterlmm/cli.py:337-375, fasterlmm/perms.py:62).  The launch reads a device
 * table: flmm_scan_table_bytes(n) bytes, filled on the host by flmm_scan_table_fill (which validates every
 * group as flmm_wald_scan does, for the tile variant FLMM_SCAN_SQUARE (128 x 128) or FLMM_SCAN_WIDE
 * (128 x 256, needs p_pad % 256 == 0), or as flmm_wald_scan_lowrank does for FLMM_SCAN_WIDE_FACTORED (every group
 * with 0 < kpad2 < kpad: factored weights)) and copied to the device by the caller.  flmm_wald_scan_grouped runs every
 * tile of every group on a persistent grid of `blocks` workgroups (a positive multiple of 8); it takes nothing
 * but the table, so a captured graph of it replays after the table is rewritten with the same variant and
 * precision.  flmm_max_reduce_grouped then folds each group's per-tile maxima into its max_F (grid sized by
 * n_groups and the largest P). */
int64_t flmm_scan_table_bytes(int n_groups);
int flmm_scan_table_fill(int prec, int variant, const flmm_scan_group* groups, int n_groups, void* host_table);
int flmm_wald_scan_grouped(int prec, int variant, const void* table, int blocks, void* stream);
int flmm_max_reduce_grouped(const void* table, int n_groups, int64_t max_P, void* stream);

/* Column standardisation (fasterlmm/io.py:223-254): for each of the m columns of the (n, m) row-major src
 * (FP64 with NaN = missing, or int8 dosages with negative = missing), NaN-aware mean and population std,
 * dst = (x - mean) / std, std == 0 -> column of zeros, missing -> 0 after standardising.  mean / std (m doubles
 * each, may be NULL) receive the column statistics.  src == dst is allowed for FP64. */
int flmm_standardise(int dtype, const void* src, int64_t n, int64_t m, int64_t ld_src, double* dst, int64_t ld_dst,
                     double* mean, double* std, void* stream);

/* Decode a PLINK 1 SNP-major .bed on the device, replacing the pysnptools reads of fasterlmm/io.py:54-67
 * (read_plink) and fasterlmm/io_stream.py:60-71 (read_block, which standardises the block it reads).  bed = the
 * file's bytes after its 3 magic bytes (0x6C 0x1B 0x01): one record of rec_bytes = ceil(n_file / 4) bytes per
 * variant, individual i in byte i / 4 at bit 2 (i % 4).  Output row r < n is individual row_idx[r] (row_idx
 * NULL: r), output column j < m is variant col_idx[j] (col_idx NULL: j); the caller guarantees
 * row_idx[r] < 4 rec_bytes and col_idx[j] < the number of records.  out is (n, m) row-major, leading dimension
 * ld.  count_a1 != 0: dosage = count of the first (.bim A1) allele, codes 00 -> 2, 10 -> 1, 11 -> 0,
 * 01 -> missing; count_a1 == 0: 2 minus that.  out_kind FLMM_BED_I8: int8 dosages, -1 = missing;
 * FLMM_BED_F64 / FLMM_BED_F32: the unit standardisation of flmm_standardise over the selected rows (io.py:247-254),
 * with mean / std (m doubles each, may be NULL) receiving the statistics. */
int flmm_bed_decode(const uint8_t* bed, int64_t rec_bytes, const int64_t* row_idx, int64_t n,
                    const int64_t* col_idx, int64_t m, int count_a1, int out_kind, void* out, int64_t ld,
                    double* mean, double* std, void* stream);

/* L2 norm of each row r < n_rows of the matrix addressed as in flmm_pack_planes (the magnitude bound that sets
 * a split-FP16 row scale). */
int flmm_row_norms(const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                   const int64_t* row_index, double* norm, void* stream);

/* Pack an FP64 matrix into operand planes.  Element (r, k) = src[src_row(r) * row_stride + k * k_stride],
 * src_row(r) = row_index ? row_index[r] : r, for r < n_rows and k < n_k; zeros elsewhere up to
 * (rows_pad, kpad).  Split-FP16 multiplies row r by 2^(14 - ceil(log2(row_norm[r]))) and writes that scale to
 * row_scale[r] (rows_pad floats); other precisions ignore row_norm / row_scale (may be NULL).  Feeds the
 * rotation products of fastlmm_compat_rotate / snp_wald_scan_compat (fasterlmm/core.py:517-518, 628-629): the
 * columns of Y - X X^+ Y, the SNPs of one chromosome (gathered by row_index, core.py:709-716), and U_eff^T. */
int flmm_pack_planes(int prec, const double* src, int64_t n_rows, int64_t n_k, int64_t row_stride, int64_t k_stride,
                     const int64_t* row_index, const double* row_norm, void* plane_hi, void* plane_lo,
                     float* row_scale, int64_t rows_pad, int64_t kpad, void* stream);

/* The regression of the columns on the covariates and their packing in one pass (fastlmm_compat_rotate's
 * Y_r = Y - X (X^+ Y), fasterlmm/core.py:517, and the packing of Y_r for U_eff^T Y_r, core.py:518): planes of the
 * transposed view of Y_r (row p = column p, k = strain n < N, zeros up to (rows_pad, kpad)) from Y (N x P row-major
 * FP64), X (N x C row-major) and B = X^+ Y (C x P row-major), each element formed as y - x_0 b_0 - x_1 b_1 - ...
 * without writing Y_r; C <= 8.  Split-FP16 first writes the rows' L2 norms to norm (P doubles) and scales as
 * flmm_pack_planes does. */
int flmm_pack_columns_projected(int prec, const double* y, int64_t N, int64_t P, const double* x, int C,
                                const double* b, double* norm, void* plane_hi, void* plane_lo, float* row_scale,
                                int64_t rows_pad, int64_t kpad, void* stream);

/* Rotation product x[i][j] = sum_n A[i][n] B[j][n] (A: a_rows_pad x n_pad planes, B: b_rows_pad x n_pad planes
 * whose rows are orthonormal basis vectors).  Replaces U_eff.T @ Y_r (core.py:518) and U_eff.T @ S_r
 * (core.py:629).  a_scale / b_scale are the operands' row scales (split-FP16 only).
 *   FLMM_ROTATE_UY: out0 = one fp32 plane (rows a_rows_pad, K = b_rows_pad) of x.
 *   FLMM_ROTATE_US: the scan's SNP-side planes S1 = x (out0, out1) and S2 = x^2 (out2, out3), with row scales
 *                   out_scale1 / out_scale2 (split-FP16); for FLMM_PREC_FP32 out0 = x, out1 = x^2.
 *   FLMM_ROTATE_US_X2: S1 = x as FLMM_ROTATE_US (out0, out1; out_scale1) and out2 = x^2 as FP64, row-major
 *                   (a_rows_pad x b_rows_pad), from which the caller forms the factored-weight S2 = x^2 A
 *                   (flmm_factored_weight_planes). */
int flmm_rotate(int prec, int mode, const void* a_hi, const void* a_lo, const float* a_scale, int64_t a_rows_pad,
                const void* b_hi, const void* b_lo, const float* b_scale, int64_t b_rows_pad, int64_t n_pad,
                void* out0, void* out1, void* out2, void* out3, float* out_scale1, float* out_scale2, void* stream);

/* Grid weights for the delta search: fp32 plane wplane (128 rows x kpad) with w[g][k] = 1/(s_k + delta_g) for
 * g < n_grid, k < neff (0 elsewhere), and lam[g] = sum_j log(s_lam_j + delta_g) over the nlam-long log-det
 * spectrum (128 floats; lam64, if not NULL, the same in FP64), on the grid linspace(lo, hi, n_grid) in log delta
 * (core.py:238, 569).  n_grid <= 128.
 * The compat loss uses s_lam = s (core.py:539); the GLS path's ML loss uses the full spectrum (core.py:177). */
int flmm_delta_weights(const double* s, int64_t neff, const double* s_lam, int64_t nlam, int64_t kpad, int n_grid,
                       double lo, double hi, float* wplane, float* lam, double* lam64, void* stream);

/* Grid argmin of the profile loss nco * log(sum_k w_gk y_k^2) + lam[g] (core.py:525-542, 569-571; nco = Neff for
 * the compat path) for columns 0..P-1 of the fp32 UY plane (rows p_pad, K = kpad); idx[p] = first grid index of
 * the minimum. */
int flmm_delta_grid(const float* wplane, const float* lam, int n_grid, double nco, const float* uy, int64_t P,
                    int64_t p_pad, int64_t kpad, int32_t* idx, void* stream);

/* flmm_delta_grid on split-FP16 MFMA: the same argmin from the same fp32 wplane / lam (flmm_delta_weights) and UY
 * plane, both operands scaled and split into fp16 (hi, lo) planes as they are staged (three products, FP32
 * accumulation, ~2^-21 relative: the accuracy of the FP32 product).  Needs a power-of-two col_scale[p] with
 * col_scale[p] * |column p| <= 2^14 (the rotated columns' plane scale: an orthonormal basis keeps the norm) and
 * 0 <= s_min <= every s_k (grid row g is scaled by the power of two of 1 / (s_min + delta_g)).  wsplit (may be
 * NULL): 4 * 128 * kpad bytes of device workspace, where the grid rows' split planes are formed once per call
 * instead of in every column tile (the same values, less VALU work per tile). */
int flmm_delta_grid_split(const float* wplane, const float* lam, int n_grid, double nco, const float* uy,
                          const float* col_scale, double s_min, double lo, double hi, int64_t P, int64_t p_pad,
                          int64_t kpad, int32_t* idx, void* wsplit, void* stream);

/* Refine log delta in [grid[i-1], grid[i+1]] (core.py:575-595), i = the FP64 grid minimum reached from the FP32 grid
 * argmin idx (the reference brackets the FP64 argmin; FP32 ties on flat likelihoods are resolved in FP64, each
 * step using lam64 -- the FP64 log-det table of flmm_delta_weights -- or, if NULL, per-column sums;
 * FLMM_DELTA_GOLDEN replays the reference's 50
 * golden-section steps, FLMM_DELTA_NEWTON converges to the same local minimiser by safeguarded Newton;
 * FLMM_DELTA_FIXED reads log_delta[p] as given and skips the search; FLMM_DELTA_GRID returns the grid node of
 * the first minimum of the FP64 loss, as refine=False does (core.py:570-573); idx may be NULL for both), then
 * write log_delta[p], rWr[p] = max(sum w y^2, 1e-300) (core.py:641-643), h2[p] = 1/(1+delta) (core.py:668-669)
 * and the scan's column-side planes Q1 = w o y (q1_hi, q1_lo) and Q2 = w (q2_hi, q2_lo), rows p_pad, with row
 * scales q1_scale / q2_scale (split-FP16); q2_hi == NULL skips Q2 (factored weights).  Columns hi_from <= p < P
 * (those flmm_wald_scan_screened reads: a multiple of 256 at or past the real columns) get the hi planes only,
 * evaluated in FP32; hi_from >= P writes both planes of every column.  The loss is nco * (log 2pi + log(R / nco) + 1)
 * + sum_j log(s_lam_j + delta).
 * The Q planes hold q_rows rows per K-chunk (q_rows >= p_pad, a multiple of 128; uy keeps p_pad).
 * order (may be NULL; flmm_screen_order_bytes(P - hi_from) bytes of device workspace) lays the columns
 * hi_from <= p < P out for the bucketed screen: bucket b of nb holds the columns with
 * t_lo + b t_h <= log_delta < t_lo + (b + 1) t_h (those outside [t_lo, t_lo + nb t_h] join the first or last
 * bucket and are screened as forced), each bucket's plane rows adjacent from hi_from on and padded to 32 rows; only
 * their Q1 hi plane is written (in the layout the screen tile loads), and order records the layout for
 * flmm_factored_weight_planes and flmm_wald_scan_screened.  With order: hi_from a multiple of 256,
 * q_rows >= flmm_screen_q_rows(p_pad, hi_from, P), 1 <= nb <= 120 and 0 < t_h <= 1/5. */
int flmm_delta_refine(int prec, int mode, const float* uy, int64_t P, int64_t p_pad, int64_t kpad, const double* s,
                      int64_t neff, const double* s_lam, int64_t nlam, double nco, int n_grid, double lo, double hi,
                      const int32_t* idx, const double* lam64, double* log_delta, double* rWr, double* h2,
                      void* q1_hi, void* q1_lo, void* q2_hi, void* q2_lo, float* q1_scale, float* q2_scale,
                      int64_t hi_from, int64_t q_rows, double t_lo, double t_h, int nb, void* order, void* stream);
/* Bytes of the layout workspace for n_cols screened columns (-1 on a bad size).  No counterpart in the reference. */
int64_t flmm_screen_order_bytes(int64_t n_cols);
/* Q-plane rows a bucketed layout of the columns hi_from <= p < P needs (p_pad when hi_from >= P; -1 on bad sizes).
 * No counterpart in the reference. */
int64_t flmm_screen_q_rows(int64_t p_pad, int64_t hi_from, int64_t P);

/* Factored weights (no counterpart in the reference, whose sWs = w^T (US o US) runs over all Neff eigenvalues,
 * core.py:640-651).  With a positive exponential sum 1/z = sum_{r<R} lam_r^2 exp(-t_r z) (relative accuracy
 * stated by the caller on the scan's range of z = s_k + delta_p),
 *   sWs[m][p] = sum_k US[m][k]^2 / (s_k + delta_p) = sum_r C[m][r] B[p][r],
 *   C = (US o US) A, A[k][r] = lam_r exp(-t_r s_k),   B[p][r] = lam_r exp(-t_r delta_p),
 * so the scan's second product runs over rpad ~ 128 instead of Neff (flmm_wald_scan_lowrank with kpad2 = rpad;
 * S2 = planes of C, e.g. from FLMM_ROTATE_US_X2 and flmm_pack_planes).  This call writes B (delta_p =
 * exp(log_delta[p]), p < P) as the Q2 planes (rows p_pad, K = rpad, zero for r >= R and p >= P), split-FP16 rows
 * scaled by the power of two of their L2 norm (q2_scale); columns hi_from <= p < P get the hi plane only (as
 * flmm_delta_refine), or, with a bucketed layout order (the one flmm_delta_refine wrote; may be NULL), their scale
 * alone, set to 1 (the bucketed screen reads no Q2 plane; its records and flmm_wald_scan_screened's resolution read
 * the same scale, so any power of two agrees).  p_pad: the Q planes' rows (flmm_delta_refine's q_rows). */
int flmm_factored_weight_planes(int prec, int64_t P, int64_t p_pad, int64_t rpad, const double* t, const double* lam,
                                int64_t R, const double* log_delta, void* q2_hi, void* q2_lo, float* q2_scale,
                                int64_t hi_from, const void* order, void* stream);

/* Kinship Gram G = Z Z^T (n x n FP64, leading dimension ld) for Z (n x m) packed as operand planes with the strains
 * as rows (rows_pad x kpad, flmm_pack_planes; K = the SNPs): the product of io.grm's K = Z Z^T / M and of the LOCO
 * kernels (fasterlmm/io.py:257-265, core.py:708-716) on MFMA in the library's split precisions.  Upper-triangular
 * 128 x 128 tile pairs x n_split K ranges run as workgroups; each folds its FP32 sums into FP64 every 1,024 k and
 * writes its FP64 tile to partial (n_split x rows_pad x rows_pad doubles, caller scratch); the splits are summed
 * in a fixed order into out, mirrored, so G is deterministic and exactly symmetric.  No counterpart in the
 * reference's arithmetic, which forms K in FP64. */
int flmm_gram(int prec, const void* z_hi, const void* z_lo, const float* z_scale, int64_t rows_pad, int64_t kpad,
              int64_t n, int n_split, double* partial, double* out, int64_t ld, void* stream);

/* Kinship Gram G = Z Z^T (n x n FP64, leading dimension ld) of Z (n x m FP64, row-major, row stride ldz >= m) on
 * the FP64 matrix cores: io.grm's K = Z Z^T / M before the division and the LOCO kernels (fasterlmm/io.py:257-265,
 * core.py:708-716), in the reference's FP64 arithmetic.  Upper-triangular 128 x 128 tile pairs x n_split K ranges
 * run as workgroups, each writing its FP64 tile to partial (n_split x rows_pad x rows_pad doubles, rows_pad = n
 * rounded up to 128, caller scratch); the splits are summed in a fixed order into out and mirrored, so G is
 * deterministic and exactly symmetric. */
int flmm_gram_f64(const double* z, int64_t ldz, int64_t n, int64_t m, int n_split, double* partial, double* out,
                  int64_t ld, void* stream);

/* FP64 rotation of SNP columns: out[i][j] = sum_{k<n} z[k * ldz + c(i)] * u[k * ldu + j], c(i) = col_index[i] (or i
 * when col_index is NULL), for i < mc, j < neff; the rest of the rows_pad x kpad output (row-major, multiples of
 * 128) is zero.  col_index (device int64, mc entries) must hold values in [0, ldz); it is not checked on the host.  z = the standardised genotypes (strains x SNPs), u = the basis U_eff (strains x neff, pointer at
 * its first column): US = U_eff^T S_r of core.py:628-629 in the reference's FP64, on the FP64 matrix cores; the
 * caller packs US and its square (or factored products) into the scan's SNP planes. */
int flmm_rotate_f64(const double* z, int64_t ldz, const int64_t* col_index, int64_t mc, int64_t n, const double* u,
                    int64_t ldu, int64_t neff, double* out, int64_t rows_pad, int64_t kpad, void* stream);

/* FP64 row GEMM C = A B (A: rows x k, B: k x cols, C: rows x cols, all row-major; rows % 64 == 0, k % 32 == 0,
 * cols % 128 == 0); every element sums k in order, so results do not depend on the number of rows per call.  Used
 * for the factored weights' SNP side C = (US o US) A (with FLMM_ROTATE_US_X2) and for the FP64 operands of
 * flmm_rescore_f64 (S^T U_eff, Y^T U_eff: U_eff.T @ S_r and U_eff.T @ Y_r of core.py:518, 629). */
int flmm_gemm_f64(const double* a, int64_t rows, int64_t k, const double* b, int64_t cols, double* c, void* stream);

/* FP64 re-evaluation of n selected tests (the scan epilogue, core.py:640-667, on exact FP64 sums): test i pairs
 * row us_row[i] of US (rotated SNPs, leading dimension ld) with column col[i] (row col[i] of UY, leading
 * dimension ld; delta = exp(log_delta[col[i]])): a = sum_k US w UY, d = sum_k US^2 w, rWr = sum_k w UY^2,
 * w_k = 1 / (s_k + delta), k < neff; then f, beta, se, sfve as the scan writes them, at
 * [row * out_ld + col[i]], row = out_row[i], or out_row_base + us_row[i] where out_row is NULL.  n_dev (may be
 * NULL): a device count, read by the kernel; the tests are then i < min(*n_dev, n) (n the lists' capacity), so a
 * caller can launch on flmm_select_above's output without reading its count.  The split-precision scan's FP32
 * sums carry ~1e-7 relative error, which
 * rss = rWr - sWy^2 / sWs amplifies by ~F / nu: re-scoring the large-F tests of the real columns keeps
 * |d log10 p| at FP64 level where p is smallest. */
int flmm_rescore_f64(const double* US, const double* UY, int64_t ld, const double* s, int64_t neff,
                     const double* log_delta, const int64_t* us_row, const int64_t* col, const int64_t* out_row,
                     int64_t out_row_base, int64_t n, const unsigned long long* n_dev, double nu, double* f,
                     double* beta, double* se, double* sfve, int64_t out_ld, void* stream);

/* The tests the FP64 re-scoring takes (core.py:640-667's F of one SNP block, the selection before
 * flmm_rescore_f64): every (row, column) of f (rows x cols, leading dimension ld, FP64) with f > thr, written to
 * out_row / out_col in an arbitrary order; *count (zeroed by the caller) receives the number of such tests, of which
 * the first cap are written (a caller seeing *count > cap selects again with a larger list).  No counterpart in the
 * reference, which re-scores nothing (its scan is FP64 throughout). */
int flmm_select_above(const double* f, int64_t rows, int64_t cols, int64_t ld, double thr, int64_t* out_row,
                      int64_t* out_col, int64_t cap, unsigned long long* count, void* stream);

/* Learned-proposal refinement (no counterpart in the reference: the learned delta proposal).  log_delta[p]
 * holds the proposal t0 on entry.  Newton on the loss slope inside [t0 - radius, t0 + radius]; the result is
 * accepted when it converges there, lies in the grid optimum's bracket [grid[idx[p] - 1], grid[idx[p] + 1]],
 * and its loss is at most judge_tol above the loss at the grid optimum grid[idx[p]]; otherwise fallback[p] = 1
 * and the column is refined in the grid bracket as FLMM_DELTA_NEWTON does.  Outputs as flmm_delta_refine;
 * fallback may be NULL. */
int flmm_delta_refine_learned(int prec, const float* uy, int64_t P, int64_t p_pad, int64_t kpad, const double* s,
                              int64_t neff, const double* s_lam, int64_t nlam, double nco, int n_grid, double lo,
                              double hi, const int32_t* idx, const double* lam64, double radius, double judge_tol,
                              double* log_delta, int32_t* fallback, double* rWr, double* h2, void* q1_hi,
                              void* q1_lo, void* q2_hi, void* q2_lo, float* q1_scale, float* q2_scale,
                              int64_t hi_from, int64_t q_rows, double t_lo, double t_h, int nb, void* order,
                              void* stream);

/* Spectral features of P rotated columns for the learned delta proposal (34 fp32 per column, row-major):
 * feature b < 32 = log(mean y^2 in eigenvalue bin b / mean y^2), bin_of_k[k] in [0, 32) for k < neff;
 * feature 32 = log(n) / 10, feature 33 = c / 10.  No counterpart in the reference. */
int flmm_mlp_features(const float* uy, int64_t P, int64_t p_pad, int64_t neff, const int32_t* bin_of_k, double n,
                      double c, float* feat, void* stream);

/* Dense layer forward: out[b][j] = act(bias[j] + sum_i W[j][i] in[b][i]), act 1 = tanh, 0 = identity. */
int flmm_mlp_dense_fwd(const float* in, int64_t B, int n_in, const float* W, const float* bias, int n_out, int act,
                       float* out, void* stream);

/* Dense layer backward: dz = dout * act'(out) (dz may alias dout), dW = dz^T in, db = sum_b dz,
 * din = dz W (din may be NULL). */
int flmm_mlp_dense_bwd(const float* dout, float* dz, const float* out, const float* in, int64_t B, int n_in,
                       int n_out, int act, const float* W, float* din, float* dW, float* db, void* stream);

/* Gradient of the mean squared error: dpred[b] = 2 (pred[b] - target[b]) / B. */
int flmm_mlp_mse_grad(const float* pred, const float* target, int64_t B, float* dpred, void* stream);

/* One Adam step (Kingma & Ba) on n parameters; step counts from 1. */
int flmm_adam(float* w, const float* g, float* m, float* v, int64_t n, double lr, double beta1, double beta2,
              double eps, int step, void* stream);

/* Fused Wald scan of one SNP block in the low-rank basis (snp_wald_scan_lowrank, fasterlmm/lowrank.py:239-316):
 * the arguments of flmm_wald_scan, with the first product sWy = S1.Q1 over kpad and the second sWs = S2.Q2 over
 * the first kpad2 < kpad entries only (S2 / Q2 planes hold kpad2).  The caller lays out
 *   S1 = [U^T s_r ; s_r]  (K = kpad2 on-rank entries, then the N entries of the regressed-out SNP s_r)
 *   Q1 = [w o U^T y_r ; (y_r - U U^T y_r) / delta]   (same split)
 *   S2 = [(U^T s_r)^2 ; |s_r - U U^T s_r|^2 at entry k]   Q2 = [w ; 1/delta at entry k]   (k = rank)
 * so sWy and sWs carry both the on-rank and the exact off-rank terms (lowrank.py:282-289; the cross term uses
 * (s_r - U U^T s_r)^T (y_r - U U^T y_r) = s_r^T (y_r - U U^T y_r)), with one row / column scale per plane.
 * The same call runs the factored-weight compat scan (flmm_factored_weight_planes): S1 = US, Q1 = w o UY over
 * kpad, S2 = C, Q2 = B over kpad2.  Needs p_pad % 256 == 0 and kpad2 % 128 == 0. */
int flmm_wald_scan_lowrank(int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi, const void* s2_lo,
                           const float* s1_scale, const float* s2_scale, int64_t mc_pad, const void* q1_hi,
                           const void* q1_lo, const void* q2_hi, const void* q2_lo, const float* q1_scale,
                           const float* q2_scale, int64_t p_pad, int64_t kpad, int64_t kpad2, int64_t mc, int64_t P,
                           int64_t n_real, const double* rWr, const double* h2, double nu, const int64_t* row_index,
                           double* f, double* beta, double* se, double* sfve, double* nullh2, double* pmax,
                           void* stream);

/* Allocation-free variant of flmm_wald_scan_lowrank followed by flmm_max_reduce.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Arguments as flmm_wald_scan_lowrank, except that the per-tile maxima
 * live in the library's own device scratch (a file-scope array of FLMM_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends) and are folded into max_F (P doubles, max_F[p] = max(max_F[p], block maximum)) by the same
 * call, so the caller passes only inputs and outputs.  Calls no hipMalloc / hipFree.  Needs
 * ceil(mc / 128) x p_pad x 8 <= FLMM_NOALLOC_SCRATCH_BYTES (FLMM_ERR_ARG otherwise: split the SNP block).  All
 * calls on one GPU share its scratch: calls that overlap on one GPU (different streams, or threads) race. */
#define FLMM_NOALLOC_SCRATCH_BYTES (64LL << 20)
int flmm_wald_scan_lowrank_noalloc(int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi,
                                   const void* s2_lo, const float* s1_scale, const float* s2_scale, int64_t mc_pad,
                                   const void* q1_hi, const void* q1_lo, const void* q2_hi, const void* q2_lo,
                                   const float* q1_scale, const float* q2_scale, int64_t p_pad, int64_t kpad,
                                   int64_t kpad2, int64_t mc, int64_t P, int64_t n_real, const double* rWr,
                                   const double* h2, double nu, const int64_t* row_index, double* f, double* beta,
                                   double* se, double* sfve, double* nullh2, double* max_F, void* stream);

/* Certified screen of the permutation columns' maximum F (perm_threshold's max_F, fasterlmm/perms.py:60-68, over
 * the scan of core.py:640-667), split-FP16 planes only.  Replaces flmm_wald_scan_lowrank + flmm_max_reduce for the
 * columns p_first <= p < P of one SNP block (p_first a multiple of 256, no real column among them): the two scan
 * products run on the hi planes alone (one fp16 MFMA instead of the split's three), each test's sums carry a proven
 * error bound, and only the tests whose bound can reach a column's maximum are re-evaluated in FP64 from
 *   us     the block's rotated SNPs in FP64, (>= mc rows, ld_us >= kpad) row-major, row m = plane row m,
 *   uy     the rotated columns' FP32 plane (p_pad x kpad, plane layout), s the basis spectrum (neff values, all
 *          within [s_lo, s_hi]), log_delta and rWr per column,
 * with the exact weights 1 / (s_k + delta); max_F[p] = max(max_F[p], those values).  The result is the maximum of
 * FP64 evaluations of every test that can hold it.  S1 / Q1 hold kpad, S2 / Q2 kpad2 (== kpad: exact weights);
 * the Q planes have q_rows rows per K-chunk, uy p_pad.
 * rec: caller workspace of ceil(mc / 128) x (P - p_first) uint64; n_eval (may be NULL) accumulates the number of
 * FP64 evaluations.  Needs mc <= 8192 and mc_pad, q_rows <= 2^24.
 * order (may be NULL): the bucketed layout flmm_delta_refine wrote for these columns (its hi_from == p_first):
 * every column then runs the sWy product alone, its 1 / sWs interpolated (quadratic in log delta, sWs within 2.5e-4
 * relative, inside the bound's kD) from FP64 sums of us^2 with the weights at three nodes of its bucket; S2 and Q2
 * are not read.  Same bounds, same result semantics.  work: device workspace of flmm_screen_work_bytes(mc_pad,
 * kpad) bytes, with order. */
int flmm_wald_scan_screened(const void* s1_hi, const void* s2_hi, const float* s1_scale, const float* s2_scale,
                            int64_t mc_pad, const void* q1_hi, const void* q2_hi, const float* q1_scale,
                            const float* q2_scale, int64_t p_pad, int64_t q_rows, int64_t kpad, int64_t kpad2,
                            int64_t mc, int64_t p_first, int64_t P, const double* us, int64_t ld_us, const float* uy,
                            const double* s, int64_t neff, double s_lo, double s_hi, const double* log_delta,
                            const double* rWr, double nu, double* max_F, void* rec, const void* order, void* work,
                            unsigned long long* n_eval, void* stream);
/* The two halves of flmm_wald_scan_screened, same arguments: flmm_wald_screen runs the screen alone (writes rec),
 * flmm_screen_resolve the FP64 re-evaluation from a rec the screen wrote (reads rec, updates max_F); in that order
 * on one stream they equal one flmm_wald_scan_screened call (a caller times them apart). */
int flmm_wald_screen(const void* s1_hi, const void* s2_hi, const float* s1_scale, const float* s2_scale, int64_t mc_pad,
                     const void* q1_hi, const void* q2_hi, const float* q1_scale, const float* q2_scale, int64_t p_pad,
                     int64_t q_rows, int64_t kpad, int64_t kpad2, int64_t mc, int64_t p_first, int64_t P,
                     const double* us, int64_t ld_us, const float* uy, const double* s, int64_t neff, double s_lo,
                     double s_hi, const double* log_delta, const double* rWr, double nu, double* max_F, void* rec,
                     const void* order, void* work, unsigned long long* n_eval, void* stream);
int flmm_screen_resolve(const void* s1_hi, const void* s2_hi, const float* s1_scale, const float* s2_scale,
                        int64_t mc_pad, const void* q1_hi, const void* q2_hi, const float* q1_scale,
                        const float* q2_scale, int64_t p_pad, int64_t q_rows, int64_t kpad, int64_t kpad2, int64_t mc,
                        int64_t p_first, int64_t P, const double* us, int64_t ld_us, const float* uy, const double* s,
                        int64_t neff, double s_lo, double s_hi, const double* log_delta, const double* rWr, double nu,
                        double* max_F, void* rec, const void* order, void* work, unsigned long long* n_eval,
                        void* stream);
/* Bytes of the bucketed screen's workspace for mc_pad SNP rows and kpad (-1 on a bad size).  No counterpart in the
 * reference (the workspace of flmm_wald_scan_screened). */
int64_t flmm_screen_work_bytes(int64_t mc_pad, int64_t kpad);

/* Fused Wald scan of one SNP block (snp_wald_scan_compat's tile loop, core.py:636-667): sWy = S1.Q1, sWs = S2.Q2 on
 * MFMA, FP64 epilogue with dof nu; split-FP16 sums are divided by the planes' row scales.  Real columns
 * p < n_real write f, beta, se, sfve, nullh2 at [row_index[m] * n_real + p] of (M, n_real) row-major FP64 arrays
 * (may be NULL when n_real == 0); every column writes its maximum F over the block's tile rows to pmax[tile][p]
 * (pmax: ceil(mc/128) x p_pad doubles).  With p_pad a multiple of 256 the 128 x 256 tile variant runs
 * (environment FLMM_SCAN_VARIANT=square forces the 128 x 128 one); both give identical results. */
int flmm_wald_scan(int prec, const void* s1_hi, const void* s1_lo, const void* s2_hi, const void* s2_lo,
                   const float* s1_scale, const float* s2_scale, int64_t mc_pad, const void* q1_hi,
                   const void* q1_lo, const void* q2_hi, const void* q2_lo, const float* q1_scale,
                   const float* q2_scale, int64_t p_pad, int64_t kpad, int64_t mc, int64_t P, int64_t n_real,
                   const double* rWr, const double* h2, double nu, const int64_t* row_index, double* f, double* beta,
                   double* se, double* sfve, double* nullh2, double* pmax, void* stream);

/* max_F[p] = max(max_F[p], max_t pmax[t * ld + p]) for t < n_tiles (the genome-wide running max of
 * core.py:657 and core.py:723). */
int flmm_max_reduce(const double* pmax, int64_t n_tiles, int64_t ld, int64_t P, double* max_F, void* stream);

/* -log10 of the F(1, df2) survival function of n values (the caller's F -> p, fasterlmm/cli.py:405-406),
 * evaluated in log space (finite far below the FP64 range of p).  F <= 0 -> 0, F = inf -> inf, NaN -> NaN. */
int flmm_neglog10p_f1(const double* F, int64_t n, double df2, double* out, void* stream);

/* EffectSize = beta^2 * g_var[m] / p_var[j] for an (M, n_real) row-major beta (fasterlmm/cli.py:132). */
int flmm_effect_size(const double* beta, int64_t M, int64_t n_real, const double* g_var, const double* p_var,
                     double* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
