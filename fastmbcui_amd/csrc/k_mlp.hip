// k_mlp.hip -- the learned warm start: a per-point MLP that maps geometry and
// rigid-body velocities to an initial guess of the MFS source strengths.
// No counterpart in the reference (its PCG always starts from zero,
// src/kl_chol.cu:938-941); PCG from the guess remains the judge of every
// answer, the network only changes the iteration count.
//
// Per source point j of body b (R_b, c_b its pose, v_b, w_b its velocities):
//   features (body frame, 20):  r = R^T (y_j - c_b), n = R^T n_j, R^T v_b,
//     R^T w_b, the boundary datum b_j, and the max over the 16 nearest sources
//     of the same body of (r_k - r_j, n_k, b_k)        (kNN max-pool, 7)
//   MLP: 20 -> 128 -> 128 -> 128 -> 6, ReLU; BF16 weights (FP32 master copy
//     for Adam), FP32 activations and accumulation
//   head: x_j = sum_d out_jd q_jd,  q_j = (R^T v_b, R^T w_b)
// The head is linear in the velocities as the strengths are; in the body
// frame every body of one shape looks alike, so the network learns one
// response per shape instead of every orientation.
//
// Kernels: feature assembly, a tiled GEMM with fused bias/ReLU (forward,
// input gradients, weight gradients with split-K), bias gradients, the
// head + loss, Adam, and a fused inference kernel (whole MLP per 64 points,
// weights in LDS).
#include "mfs_internal.h"

namespace mfs {
namespace {

constexpr int kKnn = 16;
constexpr int kFeat = 20;

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Round to nearest even (NaN kept quiet).
__device__ inline uint16_t f32_to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return uint16_t((u >> 16) | 0x40);
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(r >> 16);
}

// ------------------------------------------------------------------ features
__global__ __launch_bounds__(kThreads) void k_features(
    int64_t n, const double *__restrict__ y, const double *__restrict__ nrm,
    const double *__restrict__ bval, const int32_t *__restrict__ body,
    const int64_t *__restrict__ body_start, const double *__restrict__ rot,
    const double *__restrict__ ctr, const double *__restrict__ vw, float *__restrict__ feat,
    float *__restrict__ q) {
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= n) return;
  const int32_t b = body[j];
  const double *R = rot + 9 * b;  // row-major: world = R * local
  auto to_local = [&](const double *w, double *o) {
    for (int d = 0; d < 3; ++d) o[d] = R[0 + d] * w[0] + R[3 + d] * w[1] + R[6 + d] * w[2];
  };
  double d[3] = {y[3 * j] - ctr[3 * b], y[3 * j + 1] - ctr[3 * b + 1], y[3 * j + 2] - ctr[3 * b + 2]};
  double r[3], nl[3], vl[3], wl[3];
  to_local(d, r);
  to_local(nrm + 3 * j, nl);
  to_local(vw + 6 * b, vl);
  to_local(vw + 6 * b + 3, wl);
  float *f = feat + j * kFeat;
  for (int e = 0; e < 3; ++e) {
    f[e] = float(r[e]);
    f[3 + e] = float(nl[e]);
    f[6 + e] = float(vl[e]);
    f[9 + e] = float(wl[e]);
    q[j * 6 + e] = float(vl[e]);
    q[j * 6 + 3 + e] = float(wl[e]);
  }
  f[12] = float(bval[j]);
  // 16 nearest sources of the same body (squared distances, insertion sort)
  float bd[kKnn];
  int64_t bi[kKnn];
  for (int t = 0; t < kKnn; ++t) {
    bd[t] = 3.4e38f;
    bi[t] = -1;
  }
  const double yj0 = y[3 * j], yj1 = y[3 * j + 1], yj2 = y[3 * j + 2];
  for (int64_t k = body_start[b]; k < body_start[b + 1]; ++k) {
    if (k == j) continue;
    const double e0 = y[3 * k] - yj0, e1 = y[3 * k + 1] - yj1, e2 = y[3 * k + 2] - yj2;
    const float dd = float(e0 * e0 + e1 * e1 + e2 * e2);
    if (dd >= bd[kKnn - 1]) continue;
    int t = kKnn - 1;
    while (t > 0 && bd[t - 1] > dd) {
      bd[t] = bd[t - 1];
      bi[t] = bi[t - 1];
      --t;
    }
    bd[t] = dd;
    bi[t] = k;
  }
  float pool[7];
  for (int e = 0; e < 7; ++e) pool[e] = -3.4e38f;
  int found = 0;
  for (int t = 0; t < kKnn; ++t) {
    const int64_t k = bi[t];
    if (k < 0) continue;
    ++found;
    double dk[3] = {y[3 * k] - ctr[3 * b], y[3 * k + 1] - ctr[3 * b + 1], y[3 * k + 2] - ctr[3 * b + 2]};
    double rk[3], nk[3];
    to_local(dk, rk);
    to_local(nrm + 3 * k, nk);
    for (int e = 0; e < 3; ++e) {
      pool[e] = fmaxf(pool[e], float(rk[e] - r[e]));
      pool[3 + e] = fmaxf(pool[3 + e], float(nk[e]));
    }
    pool[6] = fmaxf(pool[6], float(bval[k]));
  }
  for (int e = 0; e < 7; ++e) f[13 + e] = found ? pool[e] : 0.f;
}

// ------------------------------------------------------------------ GEMM
// C[M x N] = op(A) op(B) (+ bias) (ReLU), or C += op(A) op(B) with split-K
// (gridDim.z > 1, atomic adds).  op(A) is M x K: A row-major M x K (lda), or
// its transpose stored K x M when TA.  op(B) is K x N: B row-major K x N, or
// stored N x K when TB.  A is FP32; B is FP32 or BF16 (uint16_t).
template <bool TA, bool TB, typename BT>
__global__ __launch_bounds__(kThreads) void k_gemm(int64_t M, int64_t N, int64_t K,
                                                   const float *__restrict__ A, int64_t lda,
                                                   const BT *__restrict__ B, int64_t ldb,
                                                   float *__restrict__ C, int64_t ldc,
                                                   const float *__restrict__ bias, int relu,
                                                   int64_t kchunk) {
  constexpr int TM = 64, TN = 64, TK = 16;
  __shared__ float As[TK][TM + 4];
  __shared__ float Bs[TK][TN + 4];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;  // 4 x 4 outputs per thread
  const int64_t m0 = int64_t(blockIdx.y) * TM, n0 = int64_t(blockIdx.x) * TN;
  const int64_t kb = int64_t(blockIdx.z) * kchunk;
  const int64_t ke = (kb + kchunk < K) ? kb + kchunk : K;
  float acc[4][4] = {};
  for (int64_t k0 = kb; k0 < ke; k0 += TK) {
    for (int idx = tid; idx < TK * TM; idx += kThreads) {
      const int kk = idx / TM, mm = idx % TM;
      const int64_t gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < ke) v = TA ? A[gk * lda + gm] : A[gm * lda + gk];
      As[kk][mm] = v;
    }
    for (int idx = tid; idx < TK * TN; idx += kThreads) {
      const int kk = idx / TN, nn = idx % TN;
      const int64_t gn = n0 + nn, gk = k0 + kk;
      float v = 0.f;
      if (gn < N && gk < ke) {
        const BT e = TB ? B[gn * ldb + gk] : B[gk * ldb + gn];
        if constexpr (sizeof(BT) == 2)
          v = bf16_to_f32(e);
        else
          v = e;
      }
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; ++kk) {
      float a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) bb[jj] = Bs[kk][tx * 4 + jj];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) acc[i][jj] += a[i] * bb[jj];
    }
    __syncthreads();
  }
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int64_t gn = n0 + tx * 4 + jj;
      if (gn >= N) continue;
      if (split) {
        atomicAdd(C + gm * ldc + gn, acc[i][jj]);
      } else {
        float v = acc[i][jj] + (bias ? bias[gn] : 0.f);
        if (relu) v = fmaxf(v, 0.f);
        C[gm * ldc + gn] = v;
      }
    }
  }
}

// G[i] *= (A[i] > 0): ReLU backward through the stored activation.
__global__ void k_relu_mask(int64_t count, const float *__restrict__ A, float *__restrict__ G) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count && !(A[i] > 0.f)) G[i] = 0.f;
}

// db[c] += sum_p G[p][c] (one block per 256 columns x row slice, atomics).
__global__ void k_bias_grad(int64_t P, int64_t N, const float *__restrict__ G, int64_t rows_per,
                            float *__restrict__ db) {
  const int64_t c = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (c >= N) return;
  const int64_t p0 = int64_t(blockIdx.y) * rows_per;
  const int64_t p1 = (p0 + rows_per < P) ? p0 + rows_per : P;
  float s = 0.f;
  for (int64_t p = p0; p < p1; ++p) s += G[p * N + c];
  atomicAdd(db + c, s);
}

// Head and loss: x_j = out_j . q_j; with a target, loss += (x_j - t_j)^2 * w
// and dout_j = 2 w (x_j - t_j) q_j.
__global__ void k_head(int64_t P, const float *__restrict__ out, const float *__restrict__ q,
                       const float *__restrict__ target, float w, float *__restrict__ x,
                       float *__restrict__ dout, double *__restrict__ loss) {
  __shared__ double red[kThreads];
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double l = 0.0;
  if (j < P) {
    float s = 0.f;
    for (int d = 0; d < 6; ++d) s += out[j * 6 + d] * q[j * 6 + d];
    if (x) x[j] = s;
    if (target) {
      const float e = s - target[j];
      l = double(e) * double(e) * double(w);
      for (int d = 0; d < 6; ++d) dout[j * 6 + d] = 2.f * w * e * q[j * 6 + d];
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

// Adam on the FP32 master weights, refreshing the BF16 copy.
// Parameters [weights (count_w) | biases]; only the weights have a BF16 copy.
__global__ void k_adam(int64_t count, int64_t count_w, float *__restrict__ w,
                       uint16_t *__restrict__ wb, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (i < count_w) wb[i] = f32_to_bf16(wi);
}

__global__ void k_to_bf16(int64_t count, const float *__restrict__ w, uint16_t *__restrict__ wb) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) wb[i] = f32_to_bf16(w[i]);
}

// Fused inference: 64 points per block, the whole MLP with BF16 weights in
// LDS (as FP32 after conversion per use), activations in LDS, and the head.
// Dimensions fixed to the network above (20 -> 128 x 3 -> 6).
constexpr int kIH = 128, kIP = 64;

__global__ __launch_bounds__(kThreads) void k_infer(int64_t P, const float *__restrict__ feat,
                                                    const float *__restrict__ q,
                                                    const uint16_t *__restrict__ wb,
                                                    const float *__restrict__ bias_all,
                                                    float *__restrict__ x) {
  extern __shared__ unsigned char smem[];
  uint16_t *W = reinterpret_cast<uint16_t *>(smem);  // all weights (bf16)
  constexpr int nW = kFeat * kIH + 2 * kIH * kIH + kIH * 6;
  float *H0 = reinterpret_cast<float *>(smem + ((nW * 2 + 15) / 16) * 16);
  float *H1 = H0 + kIP * kIH;
  for (int i = threadIdx.x; i < nW; i += kThreads) W[i] = wb[i];
  const int64_t p0 = int64_t(blockIdx.x) * kIP;
  for (int i = threadIdx.x; i < kIP * kFeat; i += kThreads) {
    const int64_t p = p0 + i / kFeat;
    H1[i] = (p < P) ? feat[p * kFeat + i % kFeat] : 0.f;
  }
  __syncthreads();
  const uint16_t *Wl[4] = {W, W + kFeat * kIH, W + kFeat * kIH + kIH * kIH,
                           W + kFeat * kIH + 2 * kIH * kIH};
  const float *bl[4] = {bias_all, bias_all + kIH, bias_all + 2 * kIH, bias_all + 3 * kIH};
  const int din[4] = {kFeat, kIH, kIH, kIH};
  const int dout[4] = {kIH, kIH, kIH, 6};
  float *in = H1, *out = H0;
  for (int l = 0; l < 4; ++l) {
    for (int idx = threadIdx.x; idx < kIP * dout[l]; idx += kThreads) {
      const int p = idx / dout[l], o = idx % dout[l];
      float s = bl[l][o];
      for (int i = 0; i < din[l]; ++i) s += in[p * din[l] + i] * bf16_to_f32(Wl[l][i * dout[l] + o]);
      out[p * dout[l] + o] = (l < 3) ? fmaxf(s, 0.f) : s;
    }
    __syncthreads();
    float *t = in;
    in = out;
    out = t;
  }
  for (int p = threadIdx.x; p < kIP; p += kThreads) {
    const int64_t gp = p0 + p;
    if (gp >= P) continue;
    float s = 0.f;
    for (int d = 0; d < 6; ++d) s += in[p * 6 + d] * q[gp * 6 + d];
    x[gp] = s;
  }
}

__global__ void k_scale(int64_t P, float s, float *__restrict__ x) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < P) x[i] *= s;
}

__global__ void k_sqnorm(int64_t P, const float *__restrict__ r, float w, double *__restrict__ acc) {
  __shared__ double red[kThreads];
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const double v = (i < P) ? double(r[i]) : 0.0;
  red[threadIdx.x] = v * v * double(w);
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(acc, red[0]);
}

__global__ void k_fit_dout(int64_t P, const float *__restrict__ g, const float *__restrict__ q,
                           float c, float *__restrict__ dout) {
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= P) return;
  const float d = c * g[j];
  for (int e = 0; e < 6; ++e) dout[j * 6 + e] = d * q[j * 6 + e];
}

inline unsigned blocks_for(int64_t count) { return unsigned((count + kThreads - 1) / kThreads); }

}  // namespace

hipError_t mlp_scale(hipStream_t st, int64_t P, float s, float *x) {
  k_scale<<<blocks_for(P), kThreads, 0, st>>>(P, s, x);
  return hipGetLastError();
}

hipError_t mlp_sqnorm(hipStream_t st, int64_t P, const float *r, float w, double *acc) {
  k_sqnorm<<<blocks_for(P), kThreads, 0, st>>>(P, r, w, acc);
  return hipGetLastError();
}

hipError_t mlp_fit_dout(hipStream_t st, int64_t P, const float *g, const float *q, float c,
                        float *dout) {
  k_fit_dout<<<blocks_for(P), kThreads, 0, st>>>(P, g, q, c, dout);
  return hipGetLastError();
}

int32_t mlp_feature_dim() { return kFeat; }

hipError_t mlp_features(hipStream_t st, int64_t n, const double *y, const double *nrm,
                        const double *bval, const int32_t *body, const int64_t *body_start,
                        const double *rot, const double *ctr, const double *vw, float *feat,
                        float *q) {
  k_features<<<blocks_for(n), kThreads, 0, st>>>(n, y, nrm, bval, body, body_start, rot, ctr, vw,
                                                 feat, q);
  return hipGetLastError();
}

// op: 0 = C = A B (+bias, relu), A P x K FP32, B K x N BF16
//     1 = C += A^T G (split-K), A P x M FP32, G P x N FP32, C M x N FP32
//     2 = C = G B^T, G P x N FP32, B K x N BF16 -> C P x K
hipError_t mlp_gemm(hipStream_t st, int op, int64_t M, int64_t N, int64_t K, const float *A,
                    int64_t lda, const void *B, int64_t ldb, float *C, int64_t ldc,
                    const float *bias, int relu) {
  const dim3 tiles(unsigned((N + 63) / 64), unsigned((M + 63) / 64), 1);
  if (op == 0) {
    k_gemm<false, false, uint16_t><<<tiles, kThreads, 0, st>>>(
        M, N, K, A, lda, static_cast<const uint16_t *>(B), ldb, C, ldc, bias, relu, K);
  } else if (op == 1) {
    // K is the point count: split it so the chip is busy (about 1024 blocks)
    const int64_t nt = int64_t(tiles.x) * tiles.y;
    int64_t splits = (1024 + nt - 1) / nt;
    int64_t kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + 15) / 16 * 16;
    splits = (K + kchunk - 1) / kchunk;
    const dim3 g(tiles.x, tiles.y, unsigned(splits < 1 ? 1 : splits));
    k_gemm<true, false, float><<<g, kThreads, 0, st>>>(M, N, K, A, lda,
                                                       static_cast<const float *>(B), ldb, C, ldc,
                                                       nullptr, 0, kchunk);
  } else {
    k_gemm<false, true, uint16_t><<<tiles, kThreads, 0, st>>>(
        M, N, K, A, lda, static_cast<const uint16_t *>(B), ldb, C, ldc, nullptr, 0, K);
  }
  return hipGetLastError();
}

hipError_t mlp_relu_mask(hipStream_t st, int64_t count, const float *A, float *G) {
  k_relu_mask<<<blocks_for(count), kThreads, 0, st>>>(count, A, G);
  return hipGetLastError();
}

hipError_t mlp_bias_grad(hipStream_t st, int64_t P, int64_t N, const float *G, float *db) {
  const int64_t rows = 512;
  const dim3 g(blocks_for(N), unsigned((P + rows - 1) / rows));
  k_bias_grad<<<g, kThreads, 0, st>>>(P, N, G, rows, db);
  return hipGetLastError();
}

hipError_t mlp_head(hipStream_t st, int64_t P, const float *out, const float *q,
                    const float *target, float w, float *x, float *dout, double *loss) {
  k_head<<<blocks_for(P), kThreads, 0, st>>>(P, out, q, target, w, x, dout, loss);
  return hipGetLastError();
}

hipError_t mlp_adam(hipStream_t st, int64_t count, int64_t count_w, float *w, uint16_t *wb,
                    const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                    int32_t step) {
  const float c1 = 1.f - powf(b1, float(step)), c2 = 1.f - powf(b2, float(step));
  k_adam<<<blocks_for(count), kThreads, 0, st>>>(count, count_w, w, wb, g, m, v, lr, b1, b2, eps,
                                                 c1, c2);
  return hipGetLastError();
}

hipError_t mlp_to_bf16(hipStream_t st, int64_t count, const float *w, uint16_t *wb) {
  k_to_bf16<<<blocks_for(count), kThreads, 0, st>>>(count, w, wb);
  return hipGetLastError();
}

size_t mlp_infer_lds() {
  constexpr int nW = kFeat * kIH + 2 * kIH * kIH + kIH * 6;
  return size_t((nW * 2 + 15) / 16 * 16) + 2 * size_t(kIP) * kIH * sizeof(float);
}

hipError_t mlp_infer(hipStream_t st, int64_t P, const float *feat, const float *q,
                     const uint16_t *wb, const float *bias_all, float *x) {
  const size_t lds = mlp_infer_lds();
  static bool attr = false;  // more than 64 KiB of dynamic LDS must be allowed once
  if (!attr) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_infer),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  k_infer<<<unsigned((P + kIP - 1) / kIP), kThreads, lds, st>>>(P, feat, q, wb, bias_all, x);
  return hipGetLastError();
}

}  // namespace mfs
