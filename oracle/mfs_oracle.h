/* mfs_oracle.h -- CPU FP64 restatement of fastMBCUI's preconditioned MFS solve.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fastmbcui_amd/, include/)
 * includes, links or calls this code; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the fastMBCUI source tree).  None of the reference's code builds here (its
 * vendored Eigen has no Core module), so each function is pinned by an
 * independent check in tests/test_oracle.py:
 *   mo_fast_fps       == mo_maximin_brute on tie-free inputs (sequence and
 *                        distances bit-equal); on tie-heavy inputs the maximin
 *                        invariants hold.  The heap's order among exactly equal
 *                        keys follows src/maxmin.h line by line but is
 *                        parity unpinned (no compiled reference to compare).
 *   mo_maximin_brute  O(n^2) definition of the maximin ordering.
 *   mo_pattern        == a numpy evaluation of the same predicate, all pairs.
 *   mo_aggregate      partition / opener / scale-rule / greedy-maximality
 *                        invariants.
 *   mo_super_pattern  == union of member columns restricted to rows >= j.
 *   mo_factor         each column == Theta_j^{-1} e1 / sqrt(e1^T Theta_j^{-1} e1)
 *                        (numpy solve); dense pattern gives L L^T = (M^T M)^{-1};
 *                        mutants (dropped reversal / normalisation /
 *                        permutation) are rejected.
 *   mo_pcg            converges to the dense solve of M^T M x = M^T b; exact
 *                        preconditioner converges in one iteration; stops at the
 *                        first iteration meeting the threshold.
 */
#ifndef MFS_ORACLE_H
#define MFS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (shared meaning with the product's C-ABI). */
#define MO_OK 0
#define MO_ERR_ARG 1
#define MO_ERR_ALLOC 2
#define MO_ERR_DUPLICATE 3   /* ordering met a zero length scale (coincident points) */
#define MO_ERR_NOT_SPD 4     /* non-positive pivot in a supernode Cholesky */

/* Reverse maximin ordering.  src/preprocess.h:782-831 (fast_fps) driving
 * src/maxmin.h:588-665 (create_ordering_3d) with the heap of maxmin.h:194-293.
 *   xyz    : n x 3 row-major source points (FP64)
 *   P      : out, P[orig] = new index (fine -> coarse: new 0 = finest)
 *   Pinv   : out, Pinv[new] = orig
 *   scale  : out, length scale per NEW index, ascending
 *   seq    : out (may be NULL), seq[k] = k-th pivot (coarse -> fine), seq[0] = root
 *   dist   : out (may be NULL), dist[k] = distances[k] of create_ordering_3d
 */
int mo_fast_fps(int32_t n, const double *xyz, int32_t *P, int32_t *Pinv,
                double *scale, int32_t *seq, double *dist);

/* Brute-force exact maximin from a given root, O(n^2).  Checker for
 * mo_fast_fps (plays the role of src/preprocess.h:684-781 brute_force_fps):
 * picks the unselected point farthest from the selected set, lowest index on
 * ties.  Distances are sqrt(((dx*dx+dy*dy)+dz*dz)). */
int mo_maximin_brute(int32_t n, const double *xyz, int32_t root, int32_t *seq,
                     double *dist);

/* rho-radius lower-triangular CSC pattern, src/preprocess.h:235-316 (MRA
 * overload of simpl_sparsity).  Column j holds rows i >= j with
 * ((dx*dx+dy*dy)+dz*dz) < min((rho*l_i)^2, (rho*l_j)^2), ascending.
 *   xyz_new : n x 3 points in NEW order (row c = point Pinv[c])
 *   rho     : already rounded as the reference does ((double)(float)rho)
 * Call with ind == NULL to fill ptr (n+1) only; then again with ind sized
 * ptr[n]. */
int mo_pattern(int32_t n, const double *xyz_new, const double *scale, double rho,
               int32_t *ptr, int32_t *ind);

/* Greedy supernode aggregation, src/preprocess.h:541-602 (group {0,n}).
 * sup_ptr has room for n+1 entries; *n_super receives the count. */
int mo_aggregate(int32_t n, const int32_t *ptr, const int32_t *ind,
                 const double *scale, double lambda, int32_t max_super,
                 int32_t *sup_ptr, int32_t *sup_ind, int32_t *sup_parent,
                 int32_t *n_super);

/* Supernodal pattern, src/preprocess.h:604-671.  Call with sind == NULL to
 * fill sptr (n+1) only; then again with sind sized sptr[n]. */
int mo_super_pattern(int32_t n, const int32_t *ptr, const int32_t *ind,
                     const int32_t *sup_parent, int32_t n_super,
                     int32_t *sptr, int32_t *sind);

/* Supernodal KL factor in FP64: Gram (src/kl_chol.cu:166-223 LS branch with
 * src/kernels.h:69-84), dense upper Cholesky of the reversed Gram
 * (kl_chol.cu:16-36, 347-360) and per-column back-substitution
 * (kl_chol.cu:38-53, 362-387).
 *   Mp  : m x n column-major FP64, column c = column Pinv[c] of M
 *   val : out, nnz = sptr[n] factor values (CSC, same layout as sind)
 * Returns MO_ERR_NOT_SPD if a pivot is not positive; *bad_super receives the
 * first such supernode (or -1). */
int mo_factor(int64_t m, int32_t n, const double *Mp,
              const int32_t *sup_ptr, const int32_t *sup_ind, int32_t n_super,
              const int32_t *sptr, const int32_t *sind, double *val,
              int32_t *bad_super);

/* t = L^T r (CSC gather).  kl_chol.cu:962-964 (CSR A_ = L^T, NON_TRANSPOSE). */
void mo_spmv_lt(int32_t n, const int32_t *sptr, const int32_t *sind,
                const double *val, const double *r, double *t);
/* z = L t (CSC scatter).  kl_chol.cu:966-970 (TRANSPOSE). */
void mo_spmv_l(int32_t n, const int32_t *sptr, const int32_t *sind,
               const double *val, const double *t, double *z);

/* PCG on M^T M x = rhs, x0 = 0.  rhs_space 0: rhs = M^T b with b of length m
 * (pcg_guesss, kl_chol.cu:891-1088); rhs_space 1: rhs given in source space,
 * length n (pcg_VpL_guesss, kl_chol.cu:1090-1264).  preconditioned 0 uses
 * z = r.  Returns iterations in *iters and sqrt(RESD2/RHS2) in *relres
 * (kl_chol.cu:1085-1087). */
int mo_pcg(int64_t m, int32_t n, const double *Mp, const double *rhs,
           int rhs_space, const int32_t *sptr, const int32_t *sind,
           const double *val, int preconditioned, int32_t maxit, double tol,
           double *x, int32_t *iters, double *relres);

#ifdef __cplusplus
}
#endif
#endif
