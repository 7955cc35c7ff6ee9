/* mfs_core.h -- C-ABI of the MI355X (gfx950) preconditioned MFS solve.
 *
 * Plain pointers and sizes only; no torch types.  Device pointers are owned by
 * the caller; every call is asynchronous on the context's HIP stream unless
 * it returns host data.  Entry points return an MFS_* status and never exit
 * the process (the reference exits on any CUDA/cuBLAS/cuSPARSE error,
 * src/macro.h:6-53).  Contexts are independent: there is no process-global
 * factor cache (the reference keeps one, mfs_torch/csrc/mfs_torch_ops.cu:17-19).
 *
 * Layouts:
 *   M  : m x n, row-major FP32 (torch's layout of the op's M argument)
 *   B  : m x k (boundary data) or n x k (source-space data), row-major FP32
 *   X  : n x k, row-major FP32, original source order
 */
#ifndef MFS_CORE_H
#define MFS_CORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfs_ctx mfs_ctx;

enum mfs_status {
    MFS_OK = 0,
    MFS_ERR_ARG = 1,        /* bad shape / pointer / parameter */
    MFS_ERR_HIP = 2,        /* a HIP runtime call failed */
    MFS_ERR_DUPLICATE = 3,  /* coincident source points (zero length scale) */
    MFS_ERR_NOT_SPD = 4,    /* a supernode Gram block lost positive definiteness */
    MFS_ERR_STATE = 5,      /* call order (e.g. factor before plan) */
    MFS_ERR_ALLOC = 6       /* device or host allocation failed */
};

enum mfs_precision {
    MFS_FP64 = 0,           /* plain drop-in: FP64 factor and FP64 PCG */
    MFS_FP32 = 1            /* chosen form: FP32 factor (FP64 retry of failed pivots) */
};

enum mfs_rhs_space {
    MFS_RHS_BOUNDARY = 0,   /* rhs = M^T b, b of length m (pcg_guesss) */
    MFS_RHS_SOURCE = 1      /* rhs given directly, length n (pcg_VpL_guesss) */
};

/* Solve statistics of the last mfs_solve_* call and phase timings (HIP events
 * on the context stream, milliseconds) of the last call of each phase. */
typedef struct mfs_stats {
    int32_t n, n_super, max_local;
    int32_t band_supernodes;  /* last dense FP32 factor: supernodes whose Gram came from the
                                 band kernel (0: none, or another form) */
    int64_t nnz_pattern, nnz_factor, theta_entries;
    double sum_ns2, sum_ns3;
    double plan_ms, factor_ms, solve_ms;
    int32_t iters[64];        /* per right-hand side (first 64) */
    double relres[64];        /* sqrt(RESD2/RHS2) per right-hand side (first 64) */
    int32_t outer_iters;      /* refinement sweeps of mfs_solve_mixed */
    int32_t retried_supernodes; /* FP32 factor: supernodes redone in FP64 */
    double plan_order_ms;       /* host maximin ordering */
    double plan_pattern_ms;     /* neighbour grids + device pattern search, to host */
    int32_t plan_grid_columns;  /* columns found by the grid search (rest: scan) */
    int32_t fp64_sweeps;        /* mixed solves: sweeps run in FP64 after FP32 sweeps stalled */
    int32_t reverted_columns;   /* mixed solves: columns restored after a sweep made them worse */
    int32_t fused_stale_polls;  /* one-pass matvec: exchange polls that found a stale record
                                   (summed over this context's launches) */
} mfs_stats;

/* Context lifecycle.  `stream` is a hipStream_t (NULL = legacy default). */
int mfs_ctx_create(mfs_ctx **out, int device, void *stream);
void mfs_ctx_destroy(mfs_ctx *ctx);
int mfs_ctx_set_stream(mfs_ctx *ctx, void *stream);
const char *mfs_ctx_last_error(const mfs_ctx *ctx);
int mfs_ctx_stats(const mfs_ctx *ctx, mfs_stats *out);

/* Ordering, rho-radius pattern and supernodes from host source points
 * (n x 3 row-major FP32, as the op's Vs_s argument).
 * Replaces mfs_torch_ops.cu:39-54 (Vs_s -> FP64, fps_sampler, fast_fps) and
 * :113-126 (simpl_sparsity, aggregate, super_sparsity, set_supernodes,
 * set_sppatt).  rho is rounded to float for the pattern as at :114. */
int mfs_plan(mfs_ctx *ctx, const float *src_xyz_host, int64_t n, double rho,
             double lambda, int32_t max_super);
/* Same with FP64 host points (used for synthetic scenes). */
int mfs_plan_f64(mfs_ctx *ctx, const double *src_xyz_host, int64_t n, double rho,
                 double lambda, int32_t max_super);

/* The plan's maximin ordering alone, on the host (no device work): xyz is
 * n x 3 row-major FP64; P[orig] = new, Pinv[new] = orig, scale[new] ascending
 * (any output may be NULL).  Replaces fps_sampler::fast_fps
 * (src/preprocess.h:782-831, create_ordering_3d src/maxmin.h:588-665).
 * MFS_ERR_DUPLICATE for coincident points. */
int mfs_order_maximin(const double *xyz, int64_t n, int32_t *P, int32_t *Pinv, double *scale);

/* Host copies of the plan (any pointer may be NULL).  Sizes: P, Pinv, scale,
 * sup_parent: n; ptr, sptr: n+1; ind: nnz_pattern; sind: nnz_factor;
 * sup_ptr: n_super+1; sup_ind: n. */
int mfs_plan_export(const mfs_ctx *ctx, int32_t *P, int32_t *Pinv, double *scale,
                    int32_t *ptr, int32_t *ind, int32_t *sup_ptr, int32_t *sup_ind,
                    int32_t *sup_parent, int32_t *sptr, int32_t *sind);

/* Numeric factor from a dense device M.  Replaces build_cov_mat_LS_float +
 * gpu_super_klchol::compute (mfs_torch_ops.cu:104-134, kl_chol.cu:772-812). */
int mfs_factor_dense(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, int precision);
/* Factor values to host as FP64, CSC layout of the supernodal pattern. */
int mfs_factor_export(const mfs_ctx *ctx, double *val);

/* Plain drop-in solve: one FP64 PCG per right-hand side, in order, with the
 * reference's stopping rule ||r||^2 < tol^2 ||r0||^2 on the normal-equation
 * residual and its zero right-hand-side shortcut.  Replaces the column loop of
 * mfs_torch_ops.cu:159-187 (pcg_guesss) and :332-360 (pcg_VpL_guesss). */
int mfs_solve_dense(mfs_ctx *ctx, const float *M, int64_t m, int64_t n,
                    const float *B, int64_t k, int rhs_space, float *X,
                    int32_t maxit, double tol);
/* mfs_solve_dense with the strengths written in FP64 (X: n x k row-major
 * double, original order): the plain form's own FP64 iterate, without the
 * FP32 rounding of the op's output. */
int mfs_solve_dense64(mfs_ctx *ctx, const float *M, int64_t m, int64_t n,
                      const float *B, int64_t k, int rhs_space, double *X,
                      int32_t maxit, double tol);

/* Chosen (mixed) form: all k right-hand sides batched through one FP32 PCG
 * whose matvec reads M once per pass for the whole batch, wrapped in FP64
 * iterative refinement on the true residual until
 * ||M^T(b - M x)|| <= tol ||M^T b|| for every column (or maxit total inner
 * iterations).  Each inner sweep stops at an absolute residual of
 * inner_tol * tol * ||M^T b|| (inner_tol in (0, 1], default 0.5 when out of
 * range).  X0 (n x k, may be NULL) is a warm start.  rhs_space
 * MFS_RHS_SOURCE takes B as n x k source-space data (the VpL system
 * M^T M x = b) and certifies ||b - M^T M x|| <= tol ||b|| instead.  No
 * counterpart in the reference. */
int mfs_solve_mixed(mfs_ctx *ctx, const float *M, int64_t m, int64_t n,
                    const float *B, int64_t k, int rhs_space, const float *X0,
                    float *X, int32_t maxit, double tol, double inner_tol);

/* FP64 relative normal-equation residual per column,
 * ||M^T (b - M x)|| / ||M^T b|| (rhs_space 0), written to host out[k]. */
int mfs_normal_residual(mfs_ctx *ctx, const float *M, int64_t m, int64_t n,
                        const float *B, int64_t k, const float *X, double *out);

/* ---- Matrix-free (points) form.  The reference op never receives target
 * points and normals (it takes M), so these stand beside the drop-in as
 * siblings; the kernel is the one src/kernels.h:24-40 evaluates (up to the
 * factor -1/(4 pi)), applied without storing M as the point-kernel driver
 * kl_chol.cu:55-126 evaluates it.
 *
 * Packs the targets x, normals nx (m x 3) and the planned sources y (n x 3,
 * original order; all device FP64) and builds the supernodal factor with the
 * Gram evaluated from the points (no m x n matrix).  precision as in
 * mfs_factor_dense; MFS_FP32 retries failed supernodes in FP64.  Replaces
 * build_cov_mat_LS_float + compute (kl_chol.cu:772-812, 867-889) for callers
 * that hold the geometry instead of M. */
int mfs_factor_points(mfs_ctx *ctx, const double *x, const double *nx, int64_t m,
                      const double *y, int64_t n, int precision);
/* mfs_solve_mixed with matrix-free passes: FP32 kernel evaluation (hi/lo
 * split coordinates, hardware rsqrt) with cascaded FP32/FP64 sums inside
 * PCG, FP64 evaluation for the refinement residual.  B: m x k FP32 device
 * boundary data; X0 (may be NULL), X: n x k.  Replaces pcg_guesss
 * (kl_chol.cu:891-1088) for the points form. */
int mfs_solve_points(mfs_ctx *ctx, const float *B, int64_t k, const float *X0, float *X,
                     int32_t maxit, double tol, double inner_tol);
/* mfs_solve_points with FP64 warm start X0 (may be NULL) and FP64 output X,
 * for tolerances below what FP32 strengths can hold (an FP32 copy of X
 * perturbs the residual of an ill-conditioned fit by ~1e-7). */
int mfs_solve_points64(mfs_ctx *ctx, const float *B, int64_t k, const double *X0, double *X,
                       int32_t maxit, double tol, double inner_tol);
/* mfs_normal_residual for the points form (FP64 kernel evaluation). */
int mfs_points_residual(mfs_ctx *ctx, const float *B, int64_t k, const float *X, double *out);
/* Same for FP64 strengths X (n x k). */
int mfs_points_residual64(mfs_ctx *ctx, const float *B, int64_t k, const double *X, double *out);
/* out = M in (trans 0: in n x k, out m x k) or M^T in (trans 1: in m x k,
 * out n x k), row-major device arrays of float (MFS_FP32) or double
 * (MFS_FP64), kernel evaluated in that precision.  The matvec of
 * kl_chol.cu:928-936 / 992-1002 without a stored M. */
int mfs_points_apply(mfs_ctx *ctx, int trans, int precision, const void *in, int64_t k,
                     void *out);

/* Q (n x k, FP64, row-major) = M^T (M X) for the dense FP32 M (m x n,
 * row-major, device) and X (n x k, FP32, device), 1 <= k <= 16, through the
 * PCG iteration's own matvec: one HBM pass over M (k_fused.hip) when
 * MFS_PARAM_FUSED allows it for this shape and k pads to 8, else y = M X then
 * M^T y.  Returns after the work is done.  The matvec pair of
 * src/kl_chol.cu:928-936 and :992-1002 (cublasSgemv with M then M^T) for a
 * batch of right-hand sides. */
int mfs_dense_normal_apply(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *X,
                           int64_t k, double *Q);

/* mfs_dense_normal_apply without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference.  The padded X, y = M X and the split
 * partials live in a static 32 MiB device array per GPU (held from module
 * load to process end; about m x n <= 2^32 / k at most -- larger shapes give
 * MFS_ERR_ARG); the caller passes M, X and Q.  Always the two-pass matvec,
 * with the same sums as mfs_dense_normal_apply.  Asynchronous on the
 * context's stream: it returns once the work is queued.  Every call on one
 * GPU shares that array, so calls that overlap on one GPU (other streams,
 * other contexts, other threads) race. */
int mfs_dense_normal_apply_noalloc(mfs_ctx *ctx, const float *M, int64_t m, int64_t n,
                                   const float *X, int64_t k, double *Q);

/* The backward of MinvF_float_fastGrad without the n x n matrix
 * W = A V^T the reference forms (core/mfs_torch_py/ops.py:42-58):
 *   grad_F = M V,   grad_M = (F - M A) V^T - (M V) A^T
 * (= F V^T - M (W + W^T)), with V = (M^T M)^{-1} G the VpL solve of the
 * incoming gradient and A the forward's output.  One pass over M gives M A and
 * M V together (two when 2k > 8), one pass writes grad_M (a rank-2k outer
 * product).  M: m x n, F: m x k, A and V: n x k, grad_M: m x n, grad_F: m x k;
 * row-major FP32 device arrays, 1 <= k <= 16; grad_M or grad_F may be NULL.
 * Asynchronous on the context's stream. */
int mfs_dense_grad(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *F,
                   const float *A, const float *V, int64_t k, float *grad_M, float *grad_F);

/* Dense point-kernel matrix M[i,j] = -(x_i - y_j).n_i / (4 pi |x_i - y_j|^3),
 * evaluated in FP64 and stored FP32.  x, nx: m x 3, y: n x 3, device FP64.
 * Point form of src/kernels.h:24-40 (laplace_nm_3d_sl, up to the factor
 * -1/(4 pi)) as driven by kl_chol.cu:55-126. */
int mfs_assemble_dense(mfs_ctx *ctx, const double *x, const double *nx, const double *y,
                       int64_t m, int64_t n, float *M);

/* ---- Steps either side of the solve in the caller (device arrays of float
 * for MFS_FP32 or double for MFS_FP64; F is int32, row-major throughout).
 * The two solid-angle calls return after their work is done (they check F's
 * indices on the device and read the verdict back; an index outside [0, nv)
 * gives MFS_ERR_ARG); the single-layer calls are asynchronous. */

/* M (nf x n): M[f][s] = 2 atan2(N, D), the solid angle of face f of the mesh
 * (V: nv x 3, F: nf x 3) seen from source s (S: n x 3), with R_i = V[F[f][i]]
 * - S[s], N = R1.(R2 x R3), D = l1 l2 l3 + (R1.R2) l3 + (R1.R3) l2 +
 * (R2.R3) l1.  Replaces compute_solid_angle (utils/torch_utils.py:306-334) as
 * called by core/KTAB_Mfs_spclf1d.py:114 (one call per body, rows stacked). */
int mfs_solid_angle(mfs_ctx *ctx, int precision, const void *V, int64_t nv, const int32_t *F,
                    int64_t nf, const void *S, int64_t n, void *M);

/* Backward of mfs_solid_angle: given gM = dL/dM (nf x n), writes gV = dL/dV
 * (nv x 3) and gS = dL/dS (n x 3); either may be NULL.  Gradients are formed
 * in FP64.  The autograd of compute_solid_angle (utils/torch_utils.py:306-334)
 * through L.backward() in core/KTAB_Mfs_spclf1d.py:136-139. */
int mfs_solid_angle_backward(mfs_ctx *ctx, int precision, const void *V, int64_t nv,
                             const int32_t *F, int64_t nf, const void *S, int64_t n,
                             const void *gM, void *gV, void *gS);

/* phi (nf x k) = SL X with SL[f][s] = 1 / |C[f] - S[s]| (S: n x 3 sources,
 * C: nf x 3 face centres, X: n x k strengths), without storing SL.  Replaces
 * compute_single_layer (utils/torch_utils.py:254-276) followed by SL @
 * strengths (core/KTAB_Mfs_spclf1d.py:123-124). */
int mfs_single_layer_apply(mfs_ctx *ctx, int precision, const void *S, int64_t n, const void *C,
                           int64_t nf, const void *X, int64_t k, void *phi);

/* Backward of mfs_single_layer_apply: given gphi = dL/dphi (nf x k), writes
 * gS (n x 3), gC (nf x 3) and gX (n x k); any may be NULL. */
int mfs_single_layer_backward(mfs_ctx *ctx, int precision, const void *S, int64_t n,
                              const void *C, int64_t nf, const void *X, int64_t k,
                              const void *gphi, void *gS, void *gC, void *gX);

/* Launch-shape parameters (autotuned by fastmbcui_amd/autotune.py; 0 restores
 * the default).  They change how work is split over blocks, not the results
 * beyond summation order. */
enum mfs_param {
    MFS_PARAM_POINTS_BLOCKS = 1,  /* blocks per points pass, owners x splits (4096) */
    MFS_PARAM_COLS_BLOCKS = 2,    /* blocks of the dense M^T pass (0: one round, CUs x resident) */
    MFS_PARAM_FUSED = 3,          /* one-pass M^T M p for dense FP32 batches of 8 where
                                     the shape qualifies (n even, n <= 65,536): 0 default (when m n >= 2^27, else
                                     two passes), 1 off (two passes), 2 always */
    MFS_PARAM_TILE_SET = 4,       /* factor buckets: bit i enables tile count
                                     {16, 12, 8, 6, 4, 3, 2, 1}[i] (16 always on); a
                                     supernode takes the smallest enabled count covering
                                     its rows.  Applies from the next plan (0: all) */
    MFS_PARAM_FACTOR_PHASES = 5,  /* column ranges of the dense factor's transpose that
                                     the Gram overlaps, 1..8; from the next plan (0: 1) */
    MFS_PARAM_FACTOR_STREAMS = 6, /* streams the factor's buckets run on, 1..4 (0: 4) */
    MFS_PARAM_FACTOR_SPLITS = 7,  /* most row splits of a supernode's Gram, 1..8: launches of
                                     fewer supernodes than 4 x CUs spread each Gram's rows
                                     over several blocks and sum their partial Grams (0: 1,
                                     no splits) */
    MFS_PARAM_GRAM_SPLIT = 8,     /* products of the FP32 factor's Gram, dense or points form
                                     (changes its rounding, not only the summation order;
                                     a batch splits when all its scenes do): 1 FP32 MFMA,
                                     2 three BF16 MFMA products of each value split into
                                     hi + lo (relative error per product about 2^-16, as
                                     FP32 sums over 65,536 rows have); 0: the default, 2 */
    MFS_PARAM_FACTOR_BANDS = 9    /* dense FP32 factor with split products: 0 (default) the
                                     Grams of supernodes inside a band of 1024 columns come
                                     from row-major M in one launch, without M^T of those
                                     columns; 1 every Gram from M^T.  From the next factor */
};
int mfs_ctx_set_param(mfs_ctx *ctx, int param, int64_t value);

/* Kernel timing for the bench: while a class is enabled (bit (1 << class) of
 * `mask`), every launch of that class is bracketed by HIP events on the context
 * stream; mfs_ctx_tracked returns the launch count, the summed event time and
 * the algorithmic work: bytes in the dense form (matvec classes: m*n*4 per
 * launch; Gram: m*4 per local column), FLOPs in the points form (m*n*(17+2k)
 * per matvec launch, 17 per kernel evaluation of the Gram). */
enum mfs_kernel_class {
    MFS_K_NONE = 0,
    MFS_K_MATVEC_ROWS = 1,   /* PCG y = M p     (one HBM pass over M, whole batch) */
    MFS_K_MATVEC_COLS = 2,   /* PCG q = M^T y   (one HBM pass over M, whole batch) */
    MFS_K_GRAM = 3,          /* fused supernode Gram + Cholesky + columns of L */
    MFS_K_RESID_ROWS = 4,    /* FP64 residual y = b - M x */
    MFS_K_RESID_COLS = 5,    /* FP64 residual r = M^T y */
    MFS_K_MATVEC_FUSED = 6   /* PCG q = M^T (M p) in one HBM pass over M (dense, FP32, 8 RHS) */
};
int mfs_ctx_track_kernel(mfs_ctx *ctx, int mask);
int mfs_ctx_tracked(mfs_ctx *ctx, int kernel_class, int64_t *launches, double *total_ms,
                    double *bytes);

/* ---- Batches of independent scenes (points form).  No counterpart in the
 * reference, which solves one scene per call (mfs_torch_ops.cu:21-197); a
 * batch runs the same per-scene algorithm for many scenes at once.  Each
 * scene has its own context and stream (mfs_batch_scene gives the context, for
 * mfs_ctx_stats / mfs_plan_export / mfs_factor_export); the batch's work runs
 * on streams of its own.  Every batch call first waits for the work queued on
 * `stream` (given at creation) and returns when its own work is done. */
typedef struct mfs_batch mfs_batch;
typedef struct mfs_batch_stats {
    double plan_ms, factor_ms, solve_ms;
    int32_t factor_launches;    /* merged bucket launches */
    int32_t factor_items;       /* supernodes of all scenes */
    int32_t graph_nodes;        /* nodes of the largest per-scene graph (use_graph 2) */
    int32_t graph_replays;      /* rounds of `chunk` iterations launched */
    int32_t sweeps;             /* refinement sweeps (max over scenes) */
    int32_t graph_captures;     /* per-scene graphs captured (use_graph 2) */
} mfs_batch_stats;
int mfs_batch_create(mfs_batch **out, int device, void *stream, int32_t count);
void mfs_batch_destroy(mfs_batch *b);
mfs_ctx *mfs_batch_scene(mfs_batch *b, int32_t i);
const char *mfs_batch_last_error(const mfs_batch *b);
int mfs_batch_stats_get(const mfs_batch *b, mfs_batch_stats *out);
/* mfs_plan_f64 for every scene (host FP64 sources xyz[i], n[i] x 3) on up to
 * `threads` host threads. */
int mfs_batch_plan(mfs_batch *b, const double *const *xyz, const int64_t *n, double rho,
                   double lambda, int32_t max_super, int32_t threads);
/* mfs_factor_points for every scene (device FP64 x[i], nx[i]: m[i] x 3, y[i]:
 * the planned sources).  mfs_batch_set_factor_mode: 2 -- all scenes'
 * supernodes in one launch whose resident blocks take them from a device
 * queue, largest tile counts first; 1 -- one launch per tile size (and row
 * split count) on four streams; 0 (default) -- the queue launch for up to
 * 32,768 units (supernodes times row splits), else a launch per bucket.  All
 * give the same factor bit for bit.  Replaces the per-scene factor of
 * mfs_torch_ops.cu:128-131 (one scene per call there); the queue launch has
 * no counterpart in the reference. */
int mfs_batch_set_factor_mode(mfs_batch *b, int32_t mode);
int mfs_batch_factor_points(mfs_batch *b, const double *const *x, const double *const *nx,
                            const int64_t *m, const double *const *y, int precision);
/* mfs_solve_points for every scene (B[i]: m[i] x k[i], X[i]: n[i] x k[i], FP32
 * device, k[i] <= 16), the PCG iterations of the running scenes launched in
 * rounds of `chunk` iterations per scene, each scene on its own stream.
 * use_graph 0: direct launches; 2: each scene's `chunk` iterations captured
 * from its own stream into a graph (once per sweep precision) and replayed
 * every round -- the same kernels in the same order, so the results are
 * identical.  use_graph 1 (one multi-stream capture forked and joined with
 * events) is switched off: MFS_ERR_ARG. */
int mfs_batch_solve_points(mfs_batch *b, const float *const *B, const int64_t *k, float *const *X,
                           int32_t maxit, double tol, double inner_tol, int32_t chunk,
                           int use_graph);

#ifdef __cplusplus
}
#endif
#endif
