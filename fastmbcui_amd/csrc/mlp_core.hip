// mlp_core.hip -- C-ABI of the learned warm start (include/mfs_mlp.h).
//
// Fixed architecture: 20 features -> 128 -> 128 -> 128 -> 6 (ReLU), head
// x = out . q (k_mlp.hip).  Parameters, flat FP32 master copy:
//   [W0 (20 x 128) | W1 (128 x 128) | W2 (128 x 128) | W3 (128 x 6) |
//    b0 (128) | b1 (128) | b2 (128) | b3 (6)]
// (W_l row-major, input x output); the weights also live as BF16, the form
// the forward and backward passes read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/mfs_core.h"
#include "../../include/mfs_mlp.h"
#include "mfs_internal.h"

namespace mfs {
int32_t mlp_feature_dim();
hipError_t mlp_features(hipStream_t st, int64_t n, const double *y, const double *nrm,
                        const double *bval, const int32_t *body, const int64_t *body_start,
                        const double *rot, const double *ctr, const double *vw, float *feat,
                        float *q);
hipError_t mlp_gemm(hipStream_t st, int op, int64_t M, int64_t N, int64_t K, const float *A,
                    int64_t lda, const uint16_t *B, int64_t ldb, float *C, int64_t ldc,
                    const float *bias, int relu);
int64_t mlp_wgrad_chunks(int64_t P);
hipError_t mlp_wgrad(hipStream_t st, int64_t P, int64_t K, int64_t N, const float *A, int64_t lda,
                     const float *G, int64_t ldg, float *part, float *dW, float *db);
hipError_t mlp_relu_mask(hipStream_t st, int64_t count, const float *A, float *G);
hipError_t mlp_head(hipStream_t st, int64_t P, const float *out, const float *q,
                    const float *target, float w, float *x, float *dout, double *lpart,
                    double *loss);
hipError_t mlp_adam(hipStream_t st, int64_t count, int64_t count_w, float *w, uint16_t *wb,
                    const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                    int32_t step);
hipError_t mlp_to_bf16(hipStream_t st, int64_t count, const float *w, uint16_t *wb);
hipError_t mlp_infer(hipStream_t st, int64_t P, const float *feat, const float *q,
                     const uint16_t *wb, const float *bias_all, float *x);
hipError_t mlp_scale(hipStream_t st, int64_t P, float s, float *x);
hipError_t mlp_sqnorm(hipStream_t st, int64_t P, const float *r, float w, double *lpart,
                      double *acc);
hipError_t mlp_fit_dout(hipStream_t st, int64_t P, const float *g, const float *q, float c, float *dout);
}  // namespace mfs

using namespace mfs;

namespace {

constexpr int kLayers = 4;
constexpr int kDims[kLayers + 1] = {20, 128, 128, 128, 6};

struct Buf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc(&p, bytes > 256 ? bytes : 256);
    if (e == hipSuccess) cap = bytes > 256 ? bytes : 256;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

}  // namespace

struct mfs_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int64_t w_off[kLayers] = {}, b_off[kLayers] = {};
  int64_t n_w = 0, n_all = 0;
  Buf w, wb, g, m, v, act, gbuf0, gbuf1, out, dout, loss, fit, wpart, lpart;
  int64_t cap_points = 0;
  int32_t step = 0;
  double *hloss = nullptr;  // pinned

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
};

#define MHIPCHK(call, what)                                                              \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return net->fail(MFS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

// Block partials of the loss sums (one double per 256 entries).
hipError_t ensure_partials(mfs_mlp *net, int64_t count) {
  return net->lpart.ensure(size_t((count + kThreads - 1) / kThreads) * sizeof(double));
}

int ensure_points(mfs_mlp *net, int64_t P) {
  if (P <= net->cap_points) return MFS_OK;
  const size_t hid = size_t(P) * 128 * sizeof(float);
  MHIPCHK(net->act.ensure(3 * hid), "alloc activations");
  MHIPCHK(net->gbuf0.ensure(hid), "alloc");
  MHIPCHK(net->gbuf1.ensure(hid), "alloc");
  MHIPCHK(net->out.ensure(size_t(P) * 6 * sizeof(float)), "alloc");
  MHIPCHK(net->dout.ensure(size_t(P) * 6 * sizeof(float)), "alloc");
  // weight-gradient partials: chunks x (largest K N + N) floats
  MHIPCHK(net->wpart.ensure(size_t(mlp_wgrad_chunks(P)) * (128 * 128 + 128) * sizeof(float)), "alloc");
  MHIPCHK(ensure_partials(net, P), "alloc");
  net->cap_points = P;
  return MFS_OK;
}

// A_1..A_3 (ReLU outputs) into act, the network output into net->out.
int forward(mfs_mlp *net, const float *feat, int64_t P) {
  const float *in = feat;
  const float *bias = net->w.as<float>();
  for (int l = 0; l < kLayers; ++l) {
    const int64_t di = kDims[l], dn = kDims[l + 1];
    float *o = (l < kLayers - 1) ? net->act.as<float>() + size_t(l) * P * 128 : net->out.as<float>();
    MHIPCHK(mlp_gemm(net->stream, 0, P, dn, di, in, di, net->wb.as<uint16_t>() + net->w_off[l], dn, o,
                     dn, bias + net->b_off[l], l < kLayers - 1 ? 1 : 0),
            "forward gemm");
    in = o;
  }
  return MFS_OK;
}

// Gradients (written to net->g, every entry) from dout.
int backward(mfs_mlp *net, const float *feat, int64_t P) {
  const float *G = net->dout.as<float>();
  float *bufs[2] = {net->gbuf0.as<float>(), net->gbuf1.as<float>()};
  for (int l = kLayers - 1; l >= 0; --l) {
    const int64_t di = kDims[l], dn = kDims[l + 1];
    const float *A = (l == 0) ? feat : net->act.as<float>() + size_t(l - 1) * P * 128;
    MHIPCHK(mlp_wgrad(net->stream, P, di, dn, A, di, G, dn, net->wpart.as<float>(),
                      net->g.as<float>() + net->w_off[l], net->g.as<float>() + net->b_off[l]),
            "weight gradient");
    if (l > 0) {
      float *Gp = bufs[l % 2];
      MHIPCHK(mlp_gemm(net->stream, 2, P, di, dn, G, dn, net->wb.as<uint16_t>() + net->w_off[l], dn, Gp,
                       di, nullptr, 0),
              "input gradient");
      MHIPCHK(mlp_relu_mask(net->stream, P * di, A, Gp), "relu backward");
      G = Gp;
    }
  }
  return MFS_OK;
}

}  // namespace

extern "C" {

int mfs_mlp_create(mfs_mlp **out, int device, void *stream, uint64_t seed) {
  if (out == nullptr) return MFS_ERR_ARG;
  *out = nullptr;
  mfs_mlp *net = new (std::nothrow) mfs_mlp();
  if (net == nullptr) return MFS_ERR_ALLOC;
  net->device = device;
  net->stream = static_cast<hipStream_t>(stream);
  int64_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    net->w_off[l] = off;
    off += int64_t(kDims[l]) * kDims[l + 1];
  }
  net->n_w = off;
  for (int l = 0; l < kLayers; ++l) {
    net->b_off[l] = off;
    off += kDims[l + 1];
  }
  net->n_all = off;
  // He-normal weights, zero biases (host, seeded)
  std::vector<float> h(size_t(net->n_all), 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < kLayers; ++l) {
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / float(kDims[l])));
    for (int64_t i = 0; i < int64_t(kDims[l]) * kDims[l + 1]; ++i) h[size_t(net->w_off[l] + i)] = nd(rng);
  }
  bool ok = hipSetDevice(device) == hipSuccess && net->w.ensure(size_t(net->n_all) * 4) == hipSuccess &&
            net->wb.ensure(size_t(net->n_w) * 2) == hipSuccess &&
            net->g.ensure(size_t(net->n_all) * 4) == hipSuccess &&
            net->m.ensure(size_t(net->n_all) * 4) == hipSuccess &&
            net->v.ensure(size_t(net->n_all) * 4) == hipSuccess && net->loss.ensure(8) == hipSuccess &&
            hipHostMalloc(reinterpret_cast<void **>(&net->hloss), 8) == hipSuccess;
  ok = ok && hipMemcpy(net->w.p, h.data(), h.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(net->m.p, 0, size_t(net->n_all) * 4) == hipSuccess &&
       hipMemset(net->v.p, 0, size_t(net->n_all) * 4) == hipSuccess &&
       mlp_to_bf16(net->stream, net->n_w, net->w.as<float>(), net->wb.as<uint16_t>()) == hipSuccess &&
       hipStreamSynchronize(net->stream) == hipSuccess;
  if (!ok) {
    mfs_mlp_destroy(net);
    return MFS_ERR_HIP;
  }
  *out = net;
  return MFS_OK;
}

void mfs_mlp_destroy(mfs_mlp *net) {
  if (net == nullptr) return;
  (void)hipSetDevice(net->device);
  (void)hipStreamSynchronize(net->stream);
  for (Buf *b : {&net->w, &net->wb, &net->g, &net->m, &net->v, &net->act, &net->gbuf0, &net->gbuf1,
                 &net->out, &net->dout, &net->loss, &net->fit, &net->wpart, &net->lpart})
    b->release();
  if (net->hloss) (void)hipHostFree(net->hloss);
  delete net;
}

const char *mfs_mlp_last_error(const mfs_mlp *net) { return net ? net->err.c_str() : "null network"; }

int64_t mfs_mlp_param_count(const mfs_mlp *net) { return net ? net->n_all : 0; }

int32_t mfs_mlp_feature_dim(void) { return mlp_feature_dim(); }

int mfs_mlp_set_stream(mfs_mlp *net, void *stream) {
  if (net == nullptr) return MFS_ERR_ARG;
  net->stream = static_cast<hipStream_t>(stream);
  return MFS_OK;
}

int mfs_mlp_get_params(mfs_mlp *net, float *host) {
  if (net == nullptr || host == nullptr) return MFS_ERR_ARG;
  MHIPCHK(hipMemcpyAsync(host, net->w.p, size_t(net->n_all) * 4, hipMemcpyDeviceToHost, net->stream), "copy");
  MHIPCHK(hipStreamSynchronize(net->stream), "sync");
  return MFS_OK;
}

int mfs_mlp_set_params(mfs_mlp *net, const float *host) {
  if (net == nullptr || host == nullptr) return MFS_ERR_ARG;
  MHIPCHK(hipMemcpyAsync(net->w.p, host, size_t(net->n_all) * 4, hipMemcpyHostToDevice, net->stream), "copy");
  MHIPCHK(mlp_to_bf16(net->stream, net->n_w, net->w.as<float>(), net->wb.as<uint16_t>()), "bf16");
  MHIPCHK(hipStreamSynchronize(net->stream), "sync");
  return MFS_OK;
}

int mfs_mlp_get_grads(mfs_mlp *net, float *host) {
  if (net == nullptr || host == nullptr) return MFS_ERR_ARG;
  MHIPCHK(hipMemcpyAsync(host, net->g.p, size_t(net->n_all) * 4, hipMemcpyDeviceToHost, net->stream), "copy");
  MHIPCHK(hipStreamSynchronize(net->stream), "sync");
  return MFS_OK;
}

int mfs_mlp_features(mfs_mlp *net, const double *y, const double *nrm, const double *bval,
                     const int32_t *body, const int64_t *body_start, const double *rot,
                     const double *ctr, const double *vw, int64_t n, float *feat, float *q) {
  if (net == nullptr || y == nullptr || nrm == nullptr || bval == nullptr || body == nullptr ||
      body_start == nullptr || rot == nullptr || ctr == nullptr || vw == nullptr || feat == nullptr ||
      q == nullptr || n <= 0)
    return MFS_ERR_ARG;
  MHIPCHK(hipSetDevice(net->device), "hipSetDevice");
  MHIPCHK(mlp_features(net->stream, n, y, nrm, bval, body, body_start, rot, ctr, vw, feat, q), "features");
  return MFS_OK;
}

int mfs_mlp_loss_grad(mfs_mlp *net, const float *feat, const float *q, const float *target,
                      int64_t P, float weight, double *loss) {
  if (net == nullptr || feat == nullptr || q == nullptr || target == nullptr || P <= 0)
    return MFS_ERR_ARG;
  MHIPCHK(hipSetDevice(net->device), "hipSetDevice");
  int rc = ensure_points(net, P);
  if (rc != MFS_OK) return rc;
  if ((rc = forward(net, feat, P)) != MFS_OK) return rc;
  MHIPCHK(hipMemsetAsync(net->loss.p, 0, 8, net->stream), "clear loss");
  MHIPCHK(mlp_head(net->stream, P, net->out.as<float>(), q, target, weight, nullptr, net->dout.as<float>(),
                   net->lpart.as<double>(), net->loss.as<double>()),
          "head");
  if ((rc = backward(net, feat, P)) != MFS_OK) return rc;
  if (loss != nullptr) {
    MHIPCHK(hipMemcpyAsync(net->hloss, net->loss.p, 8, hipMemcpyDeviceToHost, net->stream), "copy");
    MHIPCHK(hipStreamSynchronize(net->stream), "sync");
    *loss = *net->hloss;
  }
  return MFS_OK;
}

int mfs_mlp_adam_step(mfs_mlp *net, float lr, float beta1, float beta2, float eps) {
  if (net == nullptr) return MFS_ERR_ARG;
  net->step += 1;
  MHIPCHK(mlp_adam(net->stream, net->n_all, net->n_w, net->w.as<float>(), net->wb.as<uint16_t>(),
                   net->g.as<float>(), net->m.as<float>(), net->v.as<float>(), lr, beta1, beta2, eps,
                   net->step),
          "adam");
  return MFS_OK;
}

int mfs_mlp_train_step(mfs_mlp *net, const float *feat, const float *q, const float *target,
                       int64_t P, float weight, float lr, double *loss) {
  int rc = mfs_mlp_loss_grad(net, feat, q, target, P, weight, loss);
  if (rc != MFS_OK) return rc;
  return mfs_mlp_adam_step(net, lr, 0.9f, 0.999f, 1e-8f);
}

// Fit loss: sum over samples s of w ||M_s (scale x_s) - b_s||^2, x_s the head's
// output on sample s's n_s points (samples back to back in feat / q), M_s
// applied matrix-free from the packed records (k_points.hip).  Gradient:
// dL/dx_s = 2 w scale M_s^T (M_s scale x_s - b_s).
int mfs_mlp_fit_step(mfs_mlp *net, const float *feat, const float *q, int32_t samples,
                     const int64_t *offsets, const float *const *tgt32, const float *const *src32,
                     const int64_t *m, const float *const *b, float scale, float weight, float lr,
                     int normal, double *loss) {
  if (net == nullptr || feat == nullptr || q == nullptr || samples <= 0 || offsets == nullptr ||
      tgt32 == nullptr || src32 == nullptr || m == nullptr || b == nullptr)
    return MFS_ERR_ARG;
  MHIPCHK(hipSetDevice(net->device), "hipSetDevice");
  const int64_t P = offsets[samples];
  int rc = ensure_points(net, P);
  if (rc != MFS_OK) return rc;
  if ((rc = forward(net, feat, P)) != MFS_OK) return rc;
  int64_t max_m = 0, max_n = 0;
  for (int32_t s = 0; s < samples; ++s) {
    max_m = std::max<int64_t>(max_m, m[s]);
    max_n = std::max<int64_t>(max_n, offsets[s + 1] - offsets[s]);
  }
  // scratch: x (P), r (max_m), g (P), split partials
  const size_t part = size_t(std::max(points_splits(max_m, max_n, 0) * max_m,
                                      points_splits(max_n, max_m, 0) * max_n));
  MHIPCHK(net->fit.ensure((size_t(P) * 3 + size_t(max_m)) * sizeof(float) + part * sizeof(double)),
          "alloc fit scratch");
  float *x = net->fit.as<float>();
  float *g = x + P;
  float *h = g + P;
  float *r = h + P;
  double *partial = reinterpret_cast<double *>(r + ((max_m + 1) / 2) * 2);
  MHIPCHK(ensure_partials(net, std::max(P, max_m)), "alloc");
  MHIPCHK(mlp_head(net->stream, P, net->out.as<float>(), q, nullptr, 0.f, x, nullptr, nullptr, nullptr),
          "head");
  MHIPCHK(mlp_scale(net->stream, P, scale, x), "scale");
  MHIPCHK(hipMemsetAsync(net->loss.p, 0, 8, net->stream), "clear loss");
  for (int32_t s = 0; s < samples; ++s) {
    PointsOp op;
    op.tgt32 = tgt32[s];
    op.src32 = src32[s];
    op.m = m[s];
    op.n = offsets[s + 1] - offsets[s];
    const float *xs = x + offsets[s];
    // r = b - M x (rows pass with the boundary data subtracted from)
    MHIPCHK(points_rows<float>(net->stream, op, 1, xs, b[s], 1, 0, 1, r, partial, nullptr), "M x");
    if (!normal)
      MHIPCHK(mlp_sqnorm(net->stream, op.m, r, weight, net->lpart.as<double>(), net->loss.as<double>()),
              "loss");
    // g = M^T r
    MHIPCHK(points_cols<float>(net->stream, op, 1, r, partial, nullptr), "M^T r");
    MHIPCHK(points_cols_reduce<float>(net->stream, op, 1, partial, g + offsets[s]), "M^T r");
    if (normal) {
      // loss w |M^T r|^2, dL/dx = -2 w M^T M (M^T r): h = M g, g <- M^T h
      MHIPCHK(mlp_sqnorm(net->stream, op.n, g + offsets[s], weight, net->lpart.as<double>(),
                         net->loss.as<double>()),
              "loss");
      MHIPCHK(points_rows<float>(net->stream, op, 1, g + offsets[s], nullptr, 1, 0, 1, h, partial, nullptr),
              "M g");
      MHIPCHK(points_cols<float>(net->stream, op, 1, h, partial, nullptr), "M^T M g");
      MHIPCHK(points_cols_reduce<float>(net->stream, op, 1, partial, g + offsets[s]), "M^T M g");
    }
  }
  // dL/dx = -2 w scale g ; dout = dL/dx q
  MHIPCHK(mlp_fit_dout(net->stream, P, g, q, -2.f * weight * scale, net->dout.as<float>()), "dout");
  if ((rc = backward(net, feat, P)) != MFS_OK) return rc;
  if (loss != nullptr) {
    MHIPCHK(hipMemcpyAsync(net->hloss, net->loss.p, 8, hipMemcpyDeviceToHost, net->stream), "copy");
    MHIPCHK(hipStreamSynchronize(net->stream), "sync");
    *loss = *net->hloss;
  }
  return mfs_mlp_adam_step(net, lr, 0.9f, 0.999f, 1e-8f);
}

int mfs_mlp_predict(mfs_mlp *net, const float *feat, const float *q, int64_t P, float *x) {
  if (net == nullptr || feat == nullptr || q == nullptr || x == nullptr || P <= 0) return MFS_ERR_ARG;
  MHIPCHK(hipSetDevice(net->device), "hipSetDevice");
  int rc = ensure_points(net, P);
  if (rc != MFS_OK) return rc;
  if ((rc = forward(net, feat, P)) != MFS_OK) return rc;
  MHIPCHK(mlp_head(net->stream, P, net->out.as<float>(), q, nullptr, 0.f, x, nullptr, nullptr, nullptr),
          "head");
  return MFS_OK;
}

int mfs_mlp_infer(mfs_mlp *net, const float *feat, const float *q, int64_t P, float *x) {
  if (net == nullptr || feat == nullptr || q == nullptr || x == nullptr || P <= 0) return MFS_ERR_ARG;
  MHIPCHK(hipSetDevice(net->device), "hipSetDevice");
  MHIPCHK(mlp_infer(net->stream, P, feat, q, net->wb.as<uint16_t>(), net->w.as<float>() + net->b_off[0], x),
          "infer");
  return MFS_OK;
}

}  // extern "C"
