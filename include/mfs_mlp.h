/* mfs_mlp.h -- learned warm start for the MFS solve (C ABI).
 *
 * No counterpart in the reference: its PCG starts every solve from zero
 * (src/kl_chol.cu:938-941, pcg_guesss).  This network proposes the initial
 * strengths X0 handed to mfs_solve_points / mfs_solve_mixed; the solve's
 * residual test still decides every answer.
 *
 * Per source point: 20 body-frame features (position, normal, body v and w,
 * boundary datum, max-pool over the 16 nearest sources of the same body) ->
 * MLP 128-128-128 (ReLU, BF16 weights, FP32 accumulation) -> 6 coefficients,
 * strength = coefficients . (body-frame v, w).  Training: MSE on solved
 * strengths, Adam on an FP32 master copy.  All buffers are device pointers on
 * the network's stream; every call returns a status (0 = ok, mfs_core.h). */
#ifndef MFS_MLP_H
#define MFS_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfs_mlp mfs_mlp;

int mfs_mlp_create(mfs_mlp **out, int device, void *stream, uint64_t seed);
void mfs_mlp_destroy(mfs_mlp *net);
const char *mfs_mlp_last_error(const mfs_mlp *net);
int mfs_mlp_set_stream(mfs_mlp *net, void *stream);
int64_t mfs_mlp_param_count(const mfs_mlp *net);
int32_t mfs_mlp_feature_dim(void);
/* FP32 master parameters, layout [W0 W1 W2 W3 b0 b1 b2 b3], W_l row-major
 * (inputs x outputs); setting them refreshes the BF16 copy. */
int mfs_mlp_get_params(mfs_mlp *net, float *host);
int mfs_mlp_set_params(mfs_mlp *net, const float *host);
/* Gradients of the last mfs_mlp_loss_grad, same layout (w.r.t. the BF16
 * weights the forward pass used). */
int mfs_mlp_get_grads(mfs_mlp *net, float *host);

/* Features (n x 20) and head factors q (n x 6) of n source points.  y, nrm:
 * n x 3 sources and unit normals (FP64); bval: boundary data (n); body: body
 * index per point, points of a body contiguous, body_start: n_bodies + 1
 * offsets (int64); rot: n_bodies x 9 row-major (world = rot * body), ctr:
 * n_bodies x 3, vw: n_bodies x 6 (v, w in the world frame). */
int mfs_mlp_features(mfs_mlp *net, const double *y, const double *nrm, const double *bval,
                     const int32_t *body, const int64_t *body_start, const double *rot,
                     const double *ctr, const double *vw, int64_t n, float *feat, float *q);

/* loss = weight * sum_j (x_j - target_j)^2 and its gradients (FP32 feat, q,
 * target of P points). */
int mfs_mlp_loss_grad(mfs_mlp *net, const float *feat, const float *q, const float *target,
                      int64_t P, float weight, double *loss);
/* One Adam step (bias-corrected) on the gradients of the last loss_grad. */
int mfs_mlp_adam_step(mfs_mlp *net, float lr, float beta1, float beta2, float eps);
/* loss_grad + Adam (beta 0.9, 0.999, eps 1e-8). */
int mfs_mlp_train_step(mfs_mlp *net, const float *feat, const float *q, const float *target,
                       int64_t P, float weight, float lr, double *loss);
/* One Adam step on the boundary-fit loss sum_s weight ||M_s x_s - b_s||^2 of
 * `samples` scenes whose points lie back to back in feat / q (sample s:
 * rows offsets[s] .. offsets[s+1]), x_s = scale * (head output), M_s applied
 * matrix-free from its packed FP32 records tgt32[s] (m[s] x 12) and src32[s]
 * (n_s x 8) as in mfs_factor_points, b_s (m[s]) FP32 boundary data.  With
 * normal != 0 the loss is sum_s weight ||M_s^T (M_s x_s - b_s)||^2 instead:
 * the normal-equation residual a PCG started from x_s begins with.  Either
 * loss weighs the strengths by what they do to the data, which is what a
 * PCG start needs, instead of by their size. */
int mfs_mlp_fit_step(mfs_mlp *net, const float *feat, const float *q, int32_t samples,
                     const int64_t *offsets, const float *const *tgt32, const float *const *src32,
                     const int64_t *m, const float *const *b, float scale, float weight, float lr,
                     int normal, double *loss);
/* Strengths x (P, FP32): layer-by-layer GEMMs / the fused inference kernel. */
int mfs_mlp_predict(mfs_mlp *net, const float *feat, const float *q, int64_t P, float *x);
int mfs_mlp_infer(mfs_mlp *net, const float *feat, const float *q, int64_t P, float *x);

#ifdef __cplusplus
}
#endif
#endif
