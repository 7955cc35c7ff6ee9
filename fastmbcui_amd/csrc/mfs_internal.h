// mfs_internal.h -- declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mfs {

constexpr int kThreads = 256;     // every kernel here runs 4 wave64 per block
constexpr int kMaxRhs = 16;       // right-hand sides per PCG batch
constexpr int kGramChunk = 128;   // M^T rows are padded with zeros to a multiple of this
constexpr int kMaxLocal = 256;    // largest supernode (local rows); above 128, TH in global scratch

// Per-supernode descriptor (built on the host from the plan).
struct SuperDesc {
  int32_t first;     // opener column (smallest member)
  int32_t ns;        // local size = rows of the opener's column
  int32_t rows_off;  // sptr[first]: offset of the local rows in sind
  int32_t mem_beg;   // members are sup_ind[mem_beg .. mem_end)
  int32_t mem_end;
  int32_t pad;
};

// Device-resident PCG state for a batch of up to kMaxRhs right-hand sides.
struct PcgState {
  double rhs2[kMaxRhs];     // ||r0||^2
  double thresh[kMaxRhs];   // tol^2 ||r0||^2
  double absnew[kMaxRhs];   // r^T z of the current iteration
  double dot_pq[kMaxRhs];
  double dot_rr[kMaxRhs];
  double dot_rz[kMaxRhs];
  int32_t active[kMaxRhs];  // 1 while the column iterates
  int32_t iters[kMaxRhs];
  int32_t n_active;
  int32_t iter;
  int32_t maxit;
  int32_t pad;
  uint32_t counter[8];      // arrival counters of the last-block reductions
  // one-pass normal matvec (k_fused.hip): launch epoch of its tagged records,
  // waves finished in the current launch, set when a bounded wait gave up
  uint32_t fepoch;
  uint32_t fdone;
  uint32_t fabort;
  uint32_t fstale;          // its record polls that found a stale record (all launches)
  double beta[kMaxRhs];     // r^T z (new) / r^T z (old), set by k_dot_rz for k_upd_p
};

enum Counter { C_PQ = 0, C_RR = 1, C_R0 = 4, C_NORM = 5 };

// Packed point records of the matrix-free form (layouts in kernel_eval.h).
// tgt32 == nullptr means the dense form (M is stored).
struct PointsOp {
  const float *tgt32 = nullptr;   // m x 12
  const float *src32 = nullptr;   // n x 8
  const double *tgt64 = nullptr;  // m x 8
  const double *src64 = nullptr;  // n x 4
  int64_t m = 0, n = 0;
  int32_t target_blocks = 0;      // blocks per pass (owners x splits); 0: kPointsBlocks
};

// ----------------------------------------------------------------- plan kernels
// Exhaustive scan for columns j_first .. j_first + ncols - 1, or cols[0..ncols).
hipError_t launch_pattern(hipStream_t st, int32_t n, int32_t j_first, const int32_t *cols,
                          int32_t ncols, const double *px, const double *py, const double *pz,
                          const double *scale, double rho, int32_t *cnt, int32_t *ptr,
                          int32_t *ind, int fill);
// Grid search for the columns j < j_end; the fill also sorts each column with
// at most pattern_sort_max() rows (longer ones must be refilled by the scan).
struct GridBand;
hipError_t launch_grid_pattern(hipStream_t st, int32_t j_end, const GridBand *bands,
                               int32_t nbands, const int32_t *cell_start, const int32_t *cell_pts,
                               const double *px, const double *py, const double *pz,
                               const double *scale, double rho, int32_t *cnt, int32_t *ptr,
                               int32_t *ind, int fill);
int32_t pattern_sort_max();
hipError_t launch_scan(hipStream_t st, const int32_t *cnt, int32_t n, int32_t *ptr,
                       int64_t *total);

// --------------------------------------------------------------- factor kernels
// Columns c_begin .. c_end of M (c_begin a multiple of 64) into rows of M^T.
hipError_t launch_transpose(hipStream_t st, const float *M, int64_t m, int64_t n, float *Mt,
                            int64_t ldmt, int64_t c_begin, int64_t c_end);
// precision: 0 = FP64 compute / FP64 values, 1 = FP32 compute / FP32 values,
// 2 = FP64 compute / FP32 values (retry of FP32 failures).
// One scene's inputs and outputs of the factor (device pointers).  Dense form:
// Mt (n x ldmt) and a row of ldmt zeros Mz (read by the split-BF16 Gram for
// padding columns); points form: tgt (m x 12) and src (n x 8) records.
//
// Band form of the dense FP32 factor (k_band_gram): `gram` holds `parts`
// partial Grams per supernode (part r at gram + r * gram_part), supernode s's
// at offset goff[s] (16 T x 16 T, row-major, both triangles; -1: not in a
// band), which the Cholesky launches sum in part order.  The other supernodes
// read their columns from a compact M^T whose row for original column c is
// colmap[c] (nullptr: Mt holds every column, row c).
struct FactorScene {
  const float *Mt;
  const float *Mz;
  int64_t ldmt;
  const float *tgt, *src;
  int64_t m;
  const SuperDesc *sd;
  const int32_t *sind, *Pinv, *sup_ind, *sptr;
  void *val;        // factor values, TV
  int32_t *status;  // per supernode: 0 done, 1 non-positive pivot
  const int32_t *colmap = nullptr;
  const float *gram = nullptr;
  const int64_t *goff = nullptr;
  int64_t gram_part = 0;
  int32_t parts = 0;
};

// ---- band Gram (dense FP32 factor, BF16 split products) -------------------
// A band is kBandCols consecutive columns of M; its supernodes (all columns
// inside it) are packed into waves of one tile count T each, at most
// band_capacity(T) supernodes per wave (what the registers hold with no spill
// next to the staging: 251 of 256), 8 waves per block.  Block `blk`
// sums its waves' Grams over the rows of part `part` of the row range.
constexpr int kBandCols = 1024;
constexpr int kBandWaves = 8;
__host__ __device__ constexpr int band_capacity(int T) {
  return T == 1 ? 18 : T == 2 ? 8 : T == 3 ? 5 : T == 4 ? 3 : T == 6 ? 1 : 0;
}
struct BandWave {
  int32_t tiles;  // T (0: idle wave)
  int32_t count;  // supernodes, at most band_capacity(T)
  int32_t first;  // their ids are items[first .. first + count)
};
struct BandBlock {
  int32_t c0;    // first column of the block's window, even (-1: empty block)
  int32_t part;  // row part
  int32_t cols;  // window width, even, at most kBandCols
  BandWave w[kBandWaves];
};
// One launch over `count` blocks (n even: 8-byte loads of column pairs);
// rows of part r: [r * part_rows, min(m, (r + 1) * part_rows)).
hipError_t launch_band_gram(hipStream_t st, const float *M, int64_t m, int64_t n,
                            const BandBlock *blocks, int32_t count, const int32_t *items,
                            const SuperDesc *sd, const int32_t *sind, const int32_t *Pinv,
                            const int64_t *goff, float *gram, int64_t gram_part,
                            int64_t part_rows);
// Mt[i][f] = M[f][cols[i]] for f < m, 0 for m <= f < ldmt, i < ncols.
hipError_t launch_gather_transpose(hipStream_t st, const float *M, int64_t m, int64_t n,
                                   const int32_t *cols, int32_t ncols, float *Mt, int64_t ldmt);

// A supernode of a scene: one block of a factor launch.
struct FactorItem {
  int32_t scene, s;
};

// Row splits of a factor launch: each supernode's Gram is summed over
// `splits` blocks of factor_chunk(rows, splits) rows (rows: the scene's M^T
// rows, or targets in the points form) whose partial Grams meet in `partial`
// (tiles <= 8: count * splits * (16 tiles)^2 elements of the compute type;
// larger tiles use th_scratch, sized count * splits blocks); `counters`
// (count int32, zero before the first launch, left zero) counts arrivals.
// splits = 1: one block per supernode, the other fields unused.
struct FactorSplit {
  int32_t splits = 1;
  void *partial = nullptr;
  int32_t *counters = nullptr;
};

// Factors the supernodes items[0..count) (all of `tiles` tiles) of the scenes
// in `scenes` (device table); points = the scenes are in the points form (the
// Gram's operands are evaluated from the records).  tiles in {1, 2, 3, 4, 6,
// 8, 12, 16}; for 12 and 16 th_scratch holds count * splits *
// factor_scratch_elems(tiles) elements of the compute type (8 bytes for
// precision 0 and 2, 4 for 1).  split_bf16 (precision 1, either form): the
// Gram's FP32 products as three BF16 products of a two-part split.
// gram_in (precision 1, dense form): the Grams come summed from the band
// partials (FactorScene::gram); only the Cholesky and the KL columns run.
hipError_t launch_super_factor(hipStream_t st, int tiles, int precision, bool points,
                               const FactorScene *scenes, const FactorItem *items, int32_t count,
                               void *th_scratch, const FactorSplit &sp = FactorSplit{},
                               bool split_bf16 = false, bool gram_in = false);
// The points form's buckets of a batch in one launch (k_factor_queue): bucket
// b's units (count * sp.splits) are units first_unit .. of the queue, its
// items items[offset ..], its scratch and split partials as in its own launch.
struct FactorQueueBucket {
  int32_t tiles, offset, count, first_unit;
  void *th_scratch;
  FactorSplit sp;
};
// LDS bytes of one unit of `tiles` tiles (precision 0: FP64, 1: FP32).
size_t factor_unit_lds(int tiles, int precision);
// Resident blocks of the queue launch with `lds_bytes` of LDS (0: none fit).
int factor_queue_blocks(int precision, bool split_bf16, size_t lds_bytes, int cus);
// `next`: the unit counter, zero before the launch.  `blocks`: at most
// factor_queue_blocks (all resident; none waits for another).  split_bf16 as
// for launch_super_factor.
hipError_t launch_factor_queue(hipStream_t st, int precision, const FactorScene *scenes,
                               const FactorItem *items, const FactorQueueBucket *qb, int32_t nb,
                               int32_t total, int32_t *next, size_t lds_bytes, int blocks,
                               bool split_bf16);
inline int64_t factor_scratch_elems(int tiles) {
  return tiles > 8 ? int64_t(16 * tiles) * (16 * tiles + 1) : 0;
}
// Row splits for a launch of `count` supernodes over `rows` rows (M^T rows,
// or targets in the points form): enough blocks for `target` in all, at most
// 8 splits, at least 256 rows each.
inline int32_t factor_splits(int32_t count, int64_t rows, int64_t target) {
  if (count <= 0 || count >= target) return 1;
  int64_t s = (target + count - 1) / count;
  if (s > 8) s = 8;
  while (s > 1 && rows / s < 256) --s;
  return int32_t(s);
}
__host__ __device__ inline int64_t factor_chunk(int64_t rows, int32_t splits) {
  return ((rows + splits - 1) / splits + 255) / 256 * 256;
}
hipError_t launch_gather_values(hipStream_t st, int precision, int32_t nnz, const int32_t *tmap,
                                const void *val, void *tval);

// Grid of the one-pass normal matvec (k_fused.hip): `slices` column slices
// (a power of two, `width` columns each) times `groups` row groups of
// `panels` panels of 16 rows; slices == 0 when the shape does not qualify.
struct FusedShape {
  int32_t slices = 0, lg_slices = 0, groups = 0, width = 0, panels = 0;
};

// ---------------------------------------------------------------- solve kernels
// TA = double (plain / residual) or float (mixed inner PCG).
template <typename TA>
struct SolveArgs {
  const float *M;
  int64_t m, n;
  int K;                       // batch width (1, 2, 4, 8 or 16)
  const int32_t *P, *Pinv;     // orig -> new, new -> orig
  const int32_t *sptr, *sind;  // L, CSC (= L^T in CSR)
  const int32_t *tptr, *tcol;  // L, CSR
  const TA *lval, *ltval;      // factor values in both layouts
  TA *x, *r, *p, *q, *z, *t;   // n x K, new order
  TA *pv;                      // n x K, orig order (copy of p for the matvec)
  TA *y;                       // m x K
  double *part;                // partial column sums (FP64 for both precisions)
  double *dpart;               // per-block dot partials
  PcgState *st;
  int32_t row_blocks;          // number of row blocks (splits) of the M^T pass
  PointsOp pts;                // points form when pts.tgt32 != nullptr (M unused)
  FusedShape fused;            // grid of the one-pass normal matvec (slices == 0: two passes)
  uint64_t *frec = nullptr;    // its exchange records (fused_records(fused) of them)
};

// One HBM pass over M for q = M^T (M p), dense FP32 batches of 8 (k_fused.hip):
// K == 8, n even, n <= 65,536 (at most 64 slices of 1024 columns), the
// panels of a row group below 65,535; one block per CU.
FusedShape fused_shape(int64_t m, int64_t n, int K, int cus);
size_t fused_records(const FusedShape &f);
// part[g][c][r] (FP64, original column order, g < f.groups) = the row group
// g's share of (M^T M X)[c][r] (X: n x 8 FP32, original order).
hipError_t mv_fused8(hipStream_t st, const float *M, int64_t m, int64_t n, const float *X,
                     double *part, uint64_t *records, PcgState *state, const FusedShape &f);

// Q (n x kval, FP64, original column order) = M^T (M X) for the FP32 batch in
// a.pv (n x a.K, original order, zero beyond kval), through the iteration's
// own matvec (one pass when a.fused > 0, else two); a.st->n_active must be
// non-zero.
hipError_t normal_apply32(hipStream_t st, const SolveArgs<float> &a, int kval, double *Q,
                          void (*track)(void *, int, int), void *track_ctx);

// Blocks of the dense M^T pass (k_mv_cols, TA = double when fp64) resident
// on one CU; vec: n % 4 == 0.
int cols_resident_blocks(bool fp64, int K, bool vec);
// Blocks of the preconditioner products (k_spmv) for n rows and batch width
// K, with or without the dot: the dot's partials (SolveArgs::dpart) take one
// slot of K per block.
int64_t spmv_grid(int64_t n, int K, bool dot);

// thresh_abs (device, K values, may be null) replaces tol^2 ||r0||^2 as the
// stopping threshold of each column (used by the refinement's inner solves).
template <typename TA>
hipError_t pcg_init(hipStream_t st, const SolveArgs<TA> &a, const float *B, int64_t ldb,
                    int64_t col0, int rhs_space, int k_valid, double tol, int32_t maxit,
                    const double *B64, const double *thresh_abs = nullptr);
hipError_t inner_thresholds(hipStream_t st, const double *nb2, double tol_frac, int K, double *thr);
template <typename TA>
hipError_t pcg_iteration(hipStream_t st, const SolveArgs<TA> &a, void (*track)(void *, int, int),
                         void *track_ctx);
template <typename TA>
hipError_t pcg_finish(hipStream_t st, const SolveArgs<TA> &a, float *X, int64_t ldx, int64_t col0,
                      int k_valid, double *X64);
// The FP64 iterate (new order) to X (n x ldx, original order, columns col0 ..).
hipError_t pcg_finish64(hipStream_t st, const SolveArgs<double> &a, double *X, int64_t ldx,
                        int64_t col0, int k_valid);

// FP64 normal-equation residual R = M^T (B - M X) in original order, plus
// per-column ||R||^2 into out2[0..K) and (if with_rhs) ||M^T B||^2 into rhs2.
hipError_t residual64(hipStream_t st, const SolveArgs<double> &a, const float *B, int64_t ldb,
                      int64_t col0, int k_valid, const double *X64, double *R, double *out2,
                      void (*track)(void *, int, int), void *track_ctx);
// Source-space (VpL) residual R = B - M^T M X (B: n x k, original order), or
// R = B when X64 is null, with per-column ||R||^2 into out2.
hipError_t residual64_source(hipStream_t st, const SolveArgs<double> &a, const float *B,
                             int64_t ldb, int64_t col0, int k_valid, const double *X64, double *R,
                             double *out2, void (*track)(void *, int, int), void *track_ctx);

// ---------------------------------------------------------------- points form
// Splits of the other side so that owners x splits fills the chip.
constexpr int32_t kPointsBlocks = 4096;
int32_t points_splits(int64_t owners, int64_t others, int32_t target_blocks);
hipError_t pack_points(hipStream_t st, const double *x, const double *nx, const double *y,
                       int64_t m, int64_t n, float *t32, double *t64, float *s32, double *s64);
// Y (m x K) = M X, or B - M X when Bsub is given; part holds
// points_splits(m, n, op.target_blocks) * m * K doubles.
template <typename TA>
hipError_t points_rows(hipStream_t st, const PointsOp &op, int K, const TA *X, const float *Bsub,
                       int64_t ldb, int64_t col0, int kval, TA *Y, double *part,
                       const PcgState *state);
// part[points_splits(n, m, op.target_blocks)][n][K] = split sums of M^T Y.
template <typename TA>
hipError_t points_cols(hipStream_t st, const PointsOp &op, int K, const TA *Y, double *part,
                       const PcgState *state);
template <typename TA>
hipError_t points_cols_reduce(hipStream_t st, const PointsOp &op, int K, const double *part,
                              TA *out);

// Y (m x K) = M X for FP32 X (n x K, K in {1, 2, 4, 8, 16}) through the row pass.
hipError_t dense_rows32(hipStream_t st, const float *M, int64_t m, int64_t n, const float *X, int K,
                        float *Y);
// G (m x n) = U W^T, U m x K, W n x K, K in {2, 4, 8, 16, 32}, FP32.
hipError_t outer_update32(hipStream_t st, int64_t m, int64_t n, int K, const float *U, const float *W,
                          float *G);

// Kernel assembly of M from FP64 points.
hipError_t launch_assemble(hipStream_t st, const double *x, const double *nx, const double *y,
                           int64_t m, int64_t n, float *M);

// ------------------------------------------------- caller steps (k_caller.hip)
// Face solid angles M (nf x n) of the mesh (V, F) seen from the sources S and
// their backward; scratch holds solid_angle_scratch_bytes(nv, nf, n) bytes and
// *bad (device) is set to 1 when F holds an index outside [0, nv).
size_t solid_angle_scratch_bytes(int64_t nv, int64_t nf, int64_t n);
template <typename T>
hipError_t solid_angle_fwd(hipStream_t st, const T *V, int64_t nv, const int32_t *F, int64_t nf,
                           const T *S, int64_t n, T *M, void *scratch, int32_t *bad);
template <typename T>
hipError_t solid_angle_bwd(hipStream_t st, const T *V, int64_t nv, const int32_t *F, int64_t nf,
                           const T *S, int64_t n, const T *G, T *gV, T *gS, void *scratch,
                           int32_t *bad);
// phi (nf x k) = SL X with SL[f][s] = 1 / |C_f - S_s|, and its backward.
size_t single_layer_scratch_bytes(int64_t nf, int64_t n, int64_t k);
template <typename T>
hipError_t single_layer_fwd(hipStream_t st, const T *S, int64_t n, const T *C, int64_t nf,
                            const T *X, int64_t k, T *phi, void *scratch);
template <typename T>
hipError_t single_layer_bwd(hipStream_t st, const T *S, int64_t n, const T *C, int64_t nf,
                            const T *X, int64_t k, const T *Gp, T *gS, T *gC, T *gX,
                            void *scratch);

}  // namespace mfs
