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
//     for Adam), FP32 activations rounded to BF16 as matrix-product operands,
//     FP32 accumulation
//   head: x_j = sum_d out_jd q_jd,  q_j = (R^T v_b, R^T w_b)
// The head is linear in the velocities as the strengths are; in the body
// frame every body of one shape looks alike, so the network learns one
// response per shape instead of every orientation.
//
// Kernels: feature assembly; BF16 matrix-core GEMMs with fused bias/ReLU
// (forward, input gradients) and the weight and bias gradients as fixed-size
// point-chunk partials summed in chunk order (deterministic: no atomics);
// the head + loss (block partials, summed in a fixed order); Adam; and a
// fused inference kernel (whole MLP per 64 points on the matrix cores).
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

// ------------------------------------------------------------------ GEMMs
// BF16 matrix-core products with FP32 accumulation: both operands are BF16
// (the weights as stored, the FP32 activations and gradients rounded to
// nearest even as they are loaded), so every product is exact in FP32.
//
// v_mfma_f32_16x16x16bf16_1k, K = 16 in four groups of four per lane:
//   A[i][k] in lane i + 16 (k / 4), element k % 4
//   B[k][j] in lane j + 16 (k / 4), element k % 4
//   D[i][j] in lane j + 16 (i / 4), element i % 4
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline v4f mfma_bf16(s4 a, s4 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ inline s4 pack_bf16(float a, float b, float c, float d) {
  return s4{short(f32_to_bf16(a)), short(f32_to_bf16(b)), short(f32_to_bf16(c)), short(f32_to_bf16(d))};
}

constexpr int kGM = 64;     // rows of C per block (16 per wave)
constexpr int kMaxN = 128;  // widest layer
constexpr int kMaxK = 128;
constexpr int kChunk = 256; // points per weight-gradient partial (fixed: the sum order never
                            // depends on the device)
constexpr int kSub = 32;    // points staged in LDS per round

// C[M x N] = bf(A)[M x K] Bop (+ bias) (ReLU), N, K <= 128.  Bop[k][j] is the
// BF16 weight B[k ldb + j] (forward: W is K x N) or, with TB, B[j ldb + k]
// (input gradient G W^T: W is N x K).  One block per 64 rows; Bop sits in
// LDS as Bs[j][k] so that a lane's four k values are one 8-byte read.
template <bool TB>
__global__ __launch_bounds__(kThreads) void k_gemm_bf16(int64_t M, int N, int K,
                                                        const float *__restrict__ A, int64_t lda,
                                                        const uint16_t *__restrict__ B, int64_t ldb,
                                                        float *__restrict__ C, int64_t ldc,
                                                        const float *__restrict__ bias, int relu) {
  constexpr int kLd = kMaxK + 4;  // 264-byte rows
  __shared__ __align__(16) uint16_t Bs[kMaxN][kLd];
  const int Kp = (K + 15) / 16 * 16, Np = (N + 15) / 16 * 16;
  for (int idx = threadIdx.x; idx < Np * Kp; idx += kThreads) {
    const int j = TB ? idx / Kp : idx % Np;  // consecutive threads read consecutive addresses
    const int k = TB ? idx % Kp : idx / Np;
    uint16_t v = 0;
    if (j < N && k < K) v = TB ? B[j * ldb + k] : B[k * ldb + j];
    Bs[j][k] = v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  const int64_t r0 = int64_t(blockIdx.x) * kGM + w * 16;
  const int64_t row = r0 + li;
  const float *arow = A + (row < M ? row : 0) * lda;
  const int NT = Np / 16;
  v4f acc[kMaxN / 16];
#pragma unroll
  for (int t = 0; t < kMaxN / 16; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kp; k0 += 16) {
    float a[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + lk + v;
      a[v] = (row < M && k < K) ? arow[k] : 0.f;
    }
    const s4 af = pack_bf16(a[0], a[1], a[2], a[3]);
#pragma unroll
    for (int t = 0; t < kMaxN / 16; ++t)
      if (t < NT) acc[t] = mfma_bf16(af, *reinterpret_cast<const s4 *>(&Bs[t * 16 + li][k0 + lk]), acc[t]);
  }
#pragma unroll
  for (int t = 0; t < kMaxN / 16; ++t) {
    if (t >= NT) continue;
    const int c = t * 16 + li;
    if (c >= N) continue;
    const float bc = bias ? bias[c] : 0.f;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = r0 + lk + v;
      if (r >= M) continue;
      float x = acc[t][v] + bc;
      if (relu) x = fmaxf(x, 0.f);
      C[r * ldc + c] = x;
    }
  }
}

// Weight and bias gradients of one layer, chunk c = blockIdx.x of kChunk
// points:  part[c][i N + j] = sum_p bf(A[p][i]) bf(G[p][j])  (K x N <= 128 x 128)
//          partb[c][j]     = sum_p G[p][j]                  (FP32, in point order)
// Points are staged kSub at a time in LDS, transposed (As[i][p], Gs[j][p]) so
// that a lane's four points are one 8-byte read; wave w owns output tiles
// w, w + 4, ... of the (K / 16) x (N / 16) tile grid.
__global__ __launch_bounds__(kThreads) void k_wgrad_bf16(int64_t P, int K, int N,
                                                         const float *__restrict__ A, int64_t lda,
                                                         const float *__restrict__ G, int64_t ldg,
                                                         float *__restrict__ part,
                                                         float *__restrict__ partb) {
  constexpr int kLd = kSub + 4;  // 72-byte rows
  __shared__ __align__(16) uint16_t As[kMaxK][kLd];
  __shared__ __align__(16) uint16_t Gs[kMaxN][kLd];
  const int Kp = (K + 15) / 16 * 16, Np = (N + 15) / 16 * 16;
  const int NT = Np / 16, tiles = (Kp / 16) * NT;
  const int64_t p0 = int64_t(blockIdx.x) * kChunk;
  const int64_t p1 = (p0 + kChunk < P) ? p0 + kChunk : P;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  constexpr int kTilesPerWave = (kMaxK / 16) * (kMaxN / 16) / 4;
  v4f acc[kTilesPerWave];
#pragma unroll
  for (int s = 0; s < kTilesPerWave; ++s) acc[s] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int64_t s0 = p0; s0 < p1; s0 += kSub) {
    __syncthreads();  // the previous round's reads are done
    for (int idx = threadIdx.x; idx < kSub * Kp; idx += kThreads) {
      const int pp = idx / Kp, i = idx % Kp;
      const int64_t p = s0 + pp;
      As[i][pp] = (p < p1 && i < K) ? f32_to_bf16(A[p * lda + i]) : uint16_t(0);
    }
    for (int idx = threadIdx.x; idx < kSub * Np; idx += kThreads) {
      const int pp = idx / Np, j = idx % Np;
      const int64_t p = s0 + pp;
      Gs[j][pp] = (p < p1 && j < N) ? f32_to_bf16(G[p * ldg + j]) : uint16_t(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kSub; kk += 16) {
#pragma unroll
      for (int s = 0; s < kTilesPerWave; ++s) {
        const int t = w + 4 * s;
        if (t < tiles) {
          const int ti = t / NT, tj = t - ti * NT;
          const s4 af = *reinterpret_cast<const s4 *>(&As[ti * 16 + li][kk + lk]);
          const s4 bf = *reinterpret_cast<const s4 *>(&Gs[tj * 16 + li][kk + lk]);
          acc[s] = mfma_bf16(af, bf, acc[s]);
        }
      }
    }
  }
  float *pc = part + int64_t(blockIdx.x) * K * N;
#pragma unroll
  for (int s = 0; s < kTilesPerWave; ++s) {
    const int t = w + 4 * s;
    if (t >= tiles) continue;
    const int ti = t / NT, tj = t - ti * NT;
    const int j = tj * 16 + li;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = ti * 16 + lk + v;
      if (i < K && j < N) pc[i * N + j] = acc[s][v];
    }
  }
  if (int(threadIdx.x) < N) {
    float sb = 0.f;
    for (int64_t p = p0; p < p1; ++p) sb += G[p * ldg + threadIdx.x];
    partb[int64_t(blockIdx.x) * N + threadIdx.x] = sb;
  }
}

// out[e] = sum over chunks c (in order) of part[c][e], e < count.
__global__ void k_sum_chunks(int64_t chunks, int64_t count, const float *__restrict__ part,
                             float *__restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (e >= count) return;
  float s = 0.f;
  for (int64_t c = 0; c < chunks; ++c) s += part[c * count + e];
  out[e] = s;
}

// *acc += sum of partial[0 .. count) in a fixed order (one block).
__global__ void k_sum_partials(int64_t count, const double *__restrict__ partial,
                               double *__restrict__ acc) {
  __shared__ double red[kThreads];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += kThreads) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc += red[0];
}

// G[i] *= (A[i] > 0): ReLU backward through the stored activation.
__global__ void k_relu_mask(int64_t count, const float *__restrict__ A, float *__restrict__ G) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count && !(A[i] > 0.f)) G[i] = 0.f;
}

// Head and loss: x_j = out_j . q_j; with a target, block b's share of the
// loss, sum over its points of (x_j - t_j)^2 w, goes to lpart[b] (summed in
// a fixed order afterwards), and dout_j = 2 w (x_j - t_j) q_j.
__global__ void k_head(int64_t P, const float *__restrict__ out, const float *__restrict__ q,
                       const float *__restrict__ target, float w, float *__restrict__ x,
                       float *__restrict__ dout, double *__restrict__ lpart) {
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
  if (threadIdx.x == 0 && lpart) lpart[blockIdx.x] = red[0];
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

// Fused inference: the whole MLP and the head for 64 points per block (16
// per wave) on the matrix cores.  The BF16 weights are staged once per block
// in LDS, transposed (Wt[j][k]) so that a lane's four k values are one 8-byte
// read; each wave keeps its points' layer inputs in LDS as BF16 -- the same
// rounding, products and summation order as the layer-by-layer GEMMs, so
// both paths give the same values.  Dimensions fixed to the network above
// (20 -> 128 x 3 -> 6); layer 0's 20 inputs are padded to 32 with zeros.
constexpr int kIH = 128, kIP = 64;
constexpr int kLdI = kIH + 4;         // activation / weight row, BF16 elements (264 B)
constexpr int kLdI0 = 32 + 4;         // layer-0 weight row (72 B)
constexpr size_t kOffW0 = 0;
constexpr size_t kOffW1 = kOffW0 + size_t(kIH) * kLdI0 * 2;
constexpr size_t kOffW2 = kOffW1 + size_t(kIH) * kLdI * 2;
constexpr size_t kOffW3 = kOffW2 + size_t(kIH) * kLdI * 2;
constexpr size_t kOffH = kOffW3 + size_t(16) * kLdI * 2;
constexpr size_t kOffO = kOffH + size_t(4) * 2 * 16 * kLdI * 2;
constexpr size_t kInferLds = kOffO + size_t(4) * 16 * 8 * sizeof(float);

__global__ __launch_bounds__(kThreads) void k_infer(int64_t P, const float *__restrict__ feat,
                                                    const float *__restrict__ q,
                                                    const uint16_t *__restrict__ wb,
                                                    const float *__restrict__ bias_all,
                                                    float *__restrict__ x) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t *Wt[4] = {reinterpret_cast<uint16_t *>(smem + kOffW0), reinterpret_cast<uint16_t *>(smem + kOffW1),
                     reinterpret_cast<uint16_t *>(smem + kOffW2), reinterpret_cast<uint16_t *>(smem + kOffW3)};
  const int din[4] = {kFeat, kIH, kIH, kIH};
  const int dn[4] = {kIH, kIH, kIH, 6};
  const int kp[4] = {32, kIH, kIH, kIH};
  const int np[4] = {kIH, kIH, kIH, 16};
  const int ldw[4] = {kLdI0, kLdI, kLdI, kLdI};
  {
    int64_t off = 0;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      for (int idx = threadIdx.x; idx < np[l] * kp[l]; idx += kThreads) {
        const int j = idx % np[l], k = idx / np[l];
        Wt[l][j * ldw[l] + k] = (j < dn[l] && k < din[l]) ? wb[off + int64_t(k) * dn[l] + j] : uint16_t(0);
      }
      off += int64_t(din[l]) * dn[l];
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  const int64_t pb = int64_t(blockIdx.x) * kIP + w * 16;
  uint16_t *Hin = reinterpret_cast<uint16_t *>(smem + kOffH) + w * 2 * 16 * kLdI;
  uint16_t *Hout = Hin + 16 * kLdI;
  float *O = reinterpret_cast<float *>(smem + kOffO) + w * 16 * 8;
  for (int e = lane; e < 16 * 32; e += 64) {
    const int pt = e / 32, k = e % 32;
    const int64_t gp = pb + pt;
    Hin[pt * kLdI + k] = (gp < P && k < kFeat) ? f32_to_bf16(feat[gp * kFeat + k]) : uint16_t(0);
  }
  __syncthreads();
  const float *bl = bias_all;
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const int NT = np[l] / 16;
    v4f acc[kIH / 16];
#pragma unroll
    for (int t = 0; t < kIH / 16; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kp[l]; k0 += 16) {
      const s4 af = *reinterpret_cast<const s4 *>(Hin + li * kLdI + k0 + lk);
#pragma unroll
      for (int t = 0; t < kIH / 16; ++t)
        if (t < NT)
          acc[t] = mfma_bf16(af, *reinterpret_cast<const s4 *>(Wt[l] + (t * 16 + li) * ldw[l] + k0 + lk), acc[t]);
    }
#pragma unroll
    for (int t = 0; t < kIH / 16; ++t) {
      if (t >= NT) continue;
      const int j = t * 16 + li;
      if (j >= dn[l]) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float z = acc[t][v] + bl[j];
        if (l < 3)
          Hout[(lk + v) * kLdI + j] = f32_to_bf16(fmaxf(z, 0.f));
        else
          O[(lk + v) * 8 + j] = z;
      }
    }
    bl += dn[l];
    __syncthreads();
    uint16_t *t = Hin;
    Hin = Hout;
    Hout = t;
  }
  if (lane < 16) {
    const int64_t gp = pb + lane;
    if (gp < P) {
      float s = 0.f;
      for (int d = 0; d < 6; ++d) s += O[lane * 8 + d] * q[gp * 6 + d];
      x[gp] = s;
    }
  }
}

__global__ void k_scale(int64_t P, float s, float *__restrict__ x) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < P) x[i] *= s;
}

// lpart[b] = w * sum of r_i^2 over block b's entries.
__global__ void k_sqnorm(int64_t P, const float *__restrict__ r, float w, double *__restrict__ lpart) {
  __shared__ double red[kThreads];
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const double v = (i < P) ? double(r[i]) : 0.0;
  red[threadIdx.x] = v * v * double(w);
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
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

hipError_t mlp_sqnorm(hipStream_t st, int64_t P, const float *r, float w, double *lpart,
                      double *acc) {
  k_sqnorm<<<blocks_for(P), kThreads, 0, st>>>(P, r, w, lpart);
  k_sum_partials<<<1, kThreads, 0, st>>>(int64_t(blocks_for(P)), lpart, acc);
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

// op 0: C = bf(A) B (+bias, ReLU), A M x K FP32 (lda), B K x N BF16 (ldb)
// op 2: C = bf(A) B^T, A M x K FP32, B N x K BF16 (ldb) -> C M x N
hipError_t mlp_gemm(hipStream_t st, int op, int64_t M, int64_t N, int64_t K, const float *A,
                    int64_t lda, const uint16_t *B, int64_t ldb, float *C, int64_t ldc,
                    const float *bias, int relu) {
  if (N < 1 || N > kMaxN || K < 1 || K > kMaxK || M < 1 || (op != 0 && op != 2))
    return hipErrorInvalidValue;
  const unsigned blocks = unsigned((M + kGM - 1) / kGM);
  if (op == 0)
    k_gemm_bf16<false><<<blocks, kThreads, 0, st>>>(M, int(N), int(K), A, lda, B, ldb, C, ldc, bias, relu);
  else
    k_gemm_bf16<true><<<blocks, kThreads, 0, st>>>(M, int(N), int(K), A, lda, B, ldb, C, ldc, bias, relu);
  return hipGetLastError();
}

int64_t mlp_wgrad_chunks(int64_t P) { return (P + kChunk - 1) / kChunk; }

// dW[K x N] = bf(A)^T bf(G) and db[N] = column sums of G over P points, in a
// fixed order: kChunk-point partials (part: chunks x (K N + N) floats), then
// a sum over chunks in chunk order.
hipError_t mlp_wgrad(hipStream_t st, int64_t P, int64_t K, int64_t N, const float *A, int64_t lda,
                     const float *G, int64_t ldg, float *part, float *dW, float *db) {
  if (N < 1 || N > kMaxN || K < 1 || K > kMaxK || P < 1) return hipErrorInvalidValue;
  const int64_t chunks = mlp_wgrad_chunks(P);
  float *partb = part + chunks * K * N;
  k_wgrad_bf16<<<unsigned(chunks), kThreads, 0, st>>>(P, int(K), int(N), A, lda, G, ldg, part, partb);
  k_sum_chunks<<<blocks_for(K * N), kThreads, 0, st>>>(chunks, K * N, part, dW);
  k_sum_chunks<<<blocks_for(N), kThreads, 0, st>>>(chunks, N, partb, db);
  return hipGetLastError();
}

hipError_t mlp_relu_mask(hipStream_t st, int64_t count, const float *A, float *G) {
  k_relu_mask<<<blocks_for(count), kThreads, 0, st>>>(count, A, G);
  return hipGetLastError();
}

hipError_t mlp_head(hipStream_t st, int64_t P, const float *out, const float *q,
                    const float *target, float w, float *x, float *dout, double *lpart,
                    double *loss) {
  k_head<<<blocks_for(P), kThreads, 0, st>>>(P, out, q, target, w, x, dout, loss ? lpart : nullptr);
  if (loss) k_sum_partials<<<1, kThreads, 0, st>>>(int64_t(blocks_for(P)), lpart, loss);
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

size_t mlp_infer_lds() { return kInferLds; }

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
