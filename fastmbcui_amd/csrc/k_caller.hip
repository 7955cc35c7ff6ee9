// k_caller.hip -- the O(faces x sources) steps either side of the solve.
//
// fastMBCUI's Kirchhoff-tensor step (core/KTAB_Mfs_spclf1d.py:97-128) builds
// the MFS matrix from the face solid angles of every body's mesh seen from
// every source (utils/torch_utils.py:306-334, compute_solid_angle) and, after
// the solve, the potential at the face centres as a dense single-layer matrix
// times the strengths (utils/torch_utils.py:254-276 and KTAB:123-124,
// SL @ strengths).  Both run under autograd there and materialise several
// m x n x 3 temporaries.  Here each is a pair of N-body kernels:
//
//   Omega[f][s] = 2 atan2(N, D),   R_i = V[F[f][i]] - S[s],
//     N = R1 . (R2 x R3),  D = l1 l2 l3 + (R1.R2) l3 + (R1.R3) l2 + (R2.R3) l1
//   (Van Oosterom-Strackee);   phi[f][r] = sum_s X[s][r] / |C[f] - S[s]|
//   (applied, the nf x n matrix never exists).
//
// Numerics.  The triple product is evaluated as R1 . ((V2-V1) x (V3-V1)),
// equal to R1 . (R2 x R3) because R1 . (R1 x .) = 0; the face vector
// A = (V2-V1) x (V3-V1) is formed once per face in FP64, so N keeps its
// relative accuracy for sources far from the face (the direct form loses a
// factor |R| / edge to cancellation).  The FP32 forward computes in FP32
// with that form, lengths from the hardware reciprocal square root and atan2
// from an octant-reduced polynomial (relative error 1e-7; the library atan2f
// and correctly rounded sqrtf cost twice the instructions of everything else
// together).  The backward always computes in FP64: dOmega/dR is a
// difference of two terms larger than the result by |R| / edge.
// Sums over the walked side use FP32 chunks of 64 terms added in FP64 (FP32
// inputs) or FP64 throughout (FP64 inputs).
#include "mfs_internal.h"

namespace mfs {
namespace {

constexpr int kChunkTerms = 64;
constexpr int kSaFaceTile = 64;     // faces per block of the forward
constexpr int kSaSrcPerThread = 2;  // sources per thread in the backward
constexpr int kSaGroup = 4;         // faces per LDS reduction group in the backward
constexpr int kFaceRec = 12;        // V1[3] V2[3] V3[3] A[3]

template <int K>
struct IntK {
  static constexpr int value = K;
};

// ------------------------------------------------------------------ helpers

// 1/sqrt(a) in FP64: hardware estimate plus one third-order Householder step.
__device__ inline double rsq64(double a) {
  double y = __builtin_amdgcn_rsq(a);
  const double e = fma(-a * y, y, 1.0);
  return fma(y, e * fma(0.375, e, 0.5), y);
}

template <typename T>
__global__ void k_fill(int64_t count, T *__restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) p[i] = T(0);
}

template <typename T>
__global__ void k_cast(int64_t count, const double *__restrict__ a, T *__restrict__ b) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) b[i] = T(a[i]);
}

// Face records: V1, V2, V3 and A = (V2 - V1) x (V3 - V1), A in FP64.
// An index outside [0, nv) sets *bad and is read as vertex 0.
template <typename T, typename R>
__global__ void k_pack_faces(const T *__restrict__ V, int64_t nv, const int32_t *__restrict__ F,
                             int64_t nf, R *__restrict__ rec, int32_t *__restrict__ bad) {
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (f >= nf) return;
  double v[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int64_t id = F[3 * f + i];
    if (id < 0 || id >= nv) {
      *bad = 1;
      id = 0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) v[i][d] = double(V[3 * id + d]);
  }
  const double e1[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
  const double e2[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
  const double A[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]};
  R *o = rec + f * kFaceRec;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int d = 0; d < 3; ++d) o[3 * i + d] = R(v[i][d]);
#pragma unroll
  for (int d = 0; d < 3; ++d) o[9 + d] = R(A[d]);
}

// ------------------------------------------------------------ solid angle

template <typename T>
struct SaMath;

// atan(t) for t in [0, 1] as t + t s Q(s), s = t^2, Q of degree 7.  The
// coefficients were fitted for this kernel by iteratively reweighted least
// squares on the relative error of atan over [0, 1]; evaluated in FP32 by
// Horner's rule the relative error is below 1e-7 (checked against FP64 atan
// on 2.2e6 points, and through the solid angles against the FP64 oracle).
__device__ inline float atan01(float t) {
  const float s = t * t;
  float q = 0.002920738035124418f;
  q = fmaf(q, s, -0.01636829645957078f);
  q = fmaf(q, s, 0.043212789645338034f);
  q = fmaf(q, s, -0.07552323661080707f);
  q = fmaf(q, s, 0.10666071847234054f);
  q = fmaf(q, s, -0.14211076851855828f);
  q = fmaf(q, s, 0.1999377609739815f);
  q = fmaf(q, s, -0.33333152907277463f);
  return fmaf(t * s, q, t);
}

// atan2(y, x) by octant reduction to atan01 (hardware reciprocal, 1 ulp);
// atan2(0, 0) = 0 and the signs of zero follow atan2.
__device__ inline float atan2_fast(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
  float r = atan01(t);
  r = ay > ax ? 1.5707963267948966f - r : r;
  r = x < 0.0f ? 3.141592653589793f - r : r;
  return copysignf(r, y);
}

// |v| from |v|^2 with the hardware reciprocal square root (0 for 0).
__device__ inline float len_f32(float q) { return q > 0.0f ? q * __builtin_amdgcn_rsqf(q) : 0.0f; }

template <>
struct SaMath<float> {
  static __device__ inline float omega(const float *r, float px, float py, float pz) {
    const float a0 = r[0] - px, a1 = r[1] - py, a2 = r[2] - pz;
    const float b0 = r[3] - px, b1 = r[4] - py, b2 = r[5] - pz;
    const float c0 = r[6] - px, c1 = r[7] - py, c2 = r[8] - pz;
    const float l1 = len_f32(a0 * a0 + a1 * a1 + a2 * a2);
    const float l2 = len_f32(b0 * b0 + b1 * b1 + b2 * b2);
    const float l3 = len_f32(c0 * c0 + c1 * c1 + c2 * c2);
    const float N = a0 * r[9] + a1 * r[10] + a2 * r[11];
    const float d12 = a0 * b0 + a1 * b1 + a2 * b2;
    const float d13 = a0 * c0 + a1 * c1 + a2 * c2;
    const float d23 = b0 * c0 + b1 * c1 + b2 * c2;
    const float D = l1 * l2 * l3 + d12 * l3 + d13 * l2 + d23 * l1;
    return 2.0f * atan2_fast(N, D);
  }
};

template <>
struct SaMath<double> {
  static __device__ inline double omega(const double *r, double px, double py, double pz) {
    const double a0 = r[0] - px, a1 = r[1] - py, a2 = r[2] - pz;
    const double b0 = r[3] - px, b1 = r[4] - py, b2 = r[5] - pz;
    const double c0 = r[6] - px, c1 = r[7] - py, c2 = r[8] - pz;
    const double q1 = a0 * a0 + a1 * a1 + a2 * a2;
    const double q2 = b0 * b0 + b1 * b1 + b2 * b2;
    const double q3 = c0 * c0 + c1 * c1 + c2 * c2;
    const double l1 = q1 * rsq64(q1), l2 = q2 * rsq64(q2), l3 = q3 * rsq64(q3);
    const double N = a0 * r[9] + a1 * r[10] + a2 * r[11];
    const double d12 = a0 * b0 + a1 * b1 + a2 * b2;
    const double d13 = a0 * c0 + a1 * c1 + a2 * c2;
    const double d23 = b0 * c0 + b1 * c1 + b2 * c2;
    const double D = l1 * l2 * l3 + d12 * l3 + d13 * l2 + d23 * l1;
    return 2.0 * atan2(N, D);
  }
};

// M[f][s] for a tile of kSaFaceTile faces (blockIdx / stiles) and 256
// sources (blockIdx % stiles); face records arrive through scalar loads.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_solid_angle(const T *__restrict__ rec,
                                                          const T *__restrict__ S, int64_t nf,
                                                          int64_t n, int64_t stiles,
                                                          T *__restrict__ M) {
  const int64_t st = blockIdx.x % stiles, ft = blockIdx.x / stiles;
  const int64_t s = st * kThreads + threadIdx.x;
  const int64_t f0 = ft * kSaFaceTile;
  const int64_t f1 = (f0 + kSaFaceTile < nf) ? f0 + kSaFaceTile : nf;
  const bool live = s < n;
  const T px = live ? S[3 * s] : T(0), py = live ? S[3 * s + 1] : T(0),
          pz = live ? S[3 * s + 2] : T(0);
  if (f1 - f0 == kSaFaceTile) {
#pragma unroll 2
    for (int t = 0; t < kSaFaceTile; ++t) {
      const T w = SaMath<T>::omega(rec + (f0 + t) * kFaceRec, px, py, pz);
      if (live) M[(f0 + t) * n + s] = w;
    }
  } else {
    for (int64_t f = f0; f < f1; ++f) {
      const T w = SaMath<T>::omega(rec + f * kFaceRec, px, py, pz);
      if (live) M[f * n + s] = w;
    }
  }
}

struct V3 {
  double x, y, z;
};

// Weighted gradients w dOmega/dR_i (i = 1, 2, 3) in FP64.
__device__ inline void omega_grad(const double *r, double px, double py, double pz, double w,
                                  V3 &g1, V3 &g2, V3 &g3) {
  const double a0 = r[0] - px, a1 = r[1] - py, a2 = r[2] - pz;
  const double b0 = r[3] - px, b1 = r[4] - py, b2 = r[5] - pz;
  const double c0 = r[6] - px, c1 = r[7] - py, c2 = r[8] - pz;
  const double q1 = a0 * a0 + a1 * a1 + a2 * a2;
  const double q2 = b0 * b0 + b1 * b1 + b2 * b2;
  const double q3 = c0 * c0 + c1 * c1 + c2 * c2;
  const double i1 = rsq64(q1), i2 = rsq64(q2), i3 = rsq64(q3);
  const double l1 = q1 * i1, l2 = q2 * i2, l3 = q3 * i3;
  const double N = a0 * r[9] + a1 * r[10] + a2 * r[11];
  const double d12 = a0 * b0 + a1 * b1 + a2 * b2;
  const double d13 = a0 * c0 + a1 * c1 + a2 * c2;
  const double d23 = b0 * c0 + b1 * c1 + b2 * c2;
  const double D = l1 * l2 * l3 + d12 * l3 + d13 * l2 + d23 * l1;
  // d(2 atan2(N, D)) = 2 (D dN - N dD) / (N^2 + D^2)
  const double inv = 2.0 * w / (N * N + D * D);
  const double pa = D * inv, pb = -N * inv;
  // dN/dR1 = R2 x R3, dN/dR2 = R3 x R1, dN/dR3 = R1 x R2;
  // dD/dR1 = (l2 l3 + R2.R3) R1 / l1 + l3 R2 + l2 R3, and cyclically.
  const double k1 = pb * (l2 * l3 + d23) * i1, k2 = pb * (l1 * l3 + d13) * i2,
               k3 = pb * (l1 * l2 + d12) * i3;
  const double m12 = pb * l3, m13 = pb * l2, m23 = pb * l1;
  g1.x = pa * (b1 * c2 - b2 * c1) + k1 * a0 + m12 * b0 + m13 * c0;
  g1.y = pa * (b2 * c0 - b0 * c2) + k1 * a1 + m12 * b1 + m13 * c1;
  g1.z = pa * (b0 * c1 - b1 * c0) + k1 * a2 + m12 * b2 + m13 * c2;
  g2.x = pa * (c1 * a2 - c2 * a1) + k2 * b0 + m12 * a0 + m23 * c0;
  g2.y = pa * (c2 * a0 - c0 * a2) + k2 * b1 + m12 * a1 + m23 * c1;
  g2.z = pa * (c0 * a1 - c1 * a0) + k2 * b2 + m12 * a2 + m23 * c2;
  g3.x = pa * (a1 * b2 - a2 * b1) + k3 * c0 + m13 * a0 + m23 * b0;
  g3.y = pa * (a2 * b0 - a0 * b2) + k3 * c1 + m13 * a1 + m23 * b1;
  g3.z = pa * (a0 * b1 - a1 * b0) + k3 * c2 + m13 * a2 + m23 * b2;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Backward for a tile of ftile faces and kThreads * kSaSrcPerThread sources:
//   gfv[f][9] += sum_s G[f][s] dOmega/d(R1, R2, R3),
//   gs[s][3]  -= sum_f G[f][s] (dOmega/dR1 + dOmega/dR2 + dOmega/dR3),
// one FP64 atomic per value and block.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_solid_angle_bwd(
    const double *__restrict__ rec, const T *__restrict__ S, const T *__restrict__ G, int64_t nf,
    int64_t n, int64_t stiles, int64_t ftile, double *__restrict__ gfv, double *__restrict__ gs) {
  __shared__ double red[kThreads / 64][kSaGroup][9];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t st = blockIdx.x % stiles, ft = blockIdx.x / stiles;
  const int64_t f0 = ft * ftile;
  const int64_t f1 = (f0 + ftile < nf) ? f0 + ftile : nf;
  int64_t s[kSaSrcPerThread];
  double p[kSaSrcPerThread][3], acc[kSaSrcPerThread][3];
#pragma unroll
  for (int u = 0; u < kSaSrcPerThread; ++u) {
    s[u] = (st * kSaSrcPerThread + u) * kThreads + threadIdx.x;
    const bool live = s[u] < n;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // Padding lanes: a far point with zero weight, finite and adding 0.
      p[u][d] = live ? double(S[3 * s[u] + d]) : 1e10;
      acc[u][d] = 0.0;
    }
  }
  for (int64_t fg = f0; fg < f1; fg += kSaGroup) {
    const int ng = int((f1 - fg < kSaGroup) ? f1 - fg : kSaGroup);
    for (int t = 0; t < ng; ++t) {
      const int64_t f = fg + t;
      const double *r = rec + f * kFaceRec;
      double e[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < kSaSrcPerThread; ++u) {
        const double w = (s[u] < n) ? double(G[f * n + s[u]]) : 0.0;
        V3 g1, g2, g3;
        omega_grad(r, p[u][0], p[u][1], p[u][2], w, g1, g2, g3);
        e[0] += g1.x; e[1] += g1.y; e[2] += g1.z;
        e[3] += g2.x; e[4] += g2.y; e[5] += g2.z;
        e[6] += g3.x; e[7] += g3.y; e[8] += g3.z;
        acc[u][0] -= g1.x + g2.x + g3.x;
        acc[u][1] -= g1.y + g2.y + g3.y;
        acc[u][2] -= g1.z + g2.z + g3.z;
      }
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const double v = wave_sum(e[c]);
        if (lane == 0) red[wave][t][c] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x < ng * 9) {
      const int t = threadIdx.x / 9, c = threadIdx.x % 9;
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) v += red[w][t][c];
      atomicAdd(gfv + (fg + t) * 9 + c, v);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kSaSrcPerThread; ++u)
    if (s[u] < n)
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(gs + 3 * s[u] + d, acc[u][d]);
}

// gv[F[f][i]] += gfv[f][3 i .. 3 i + 3) (indices outside [0, nv) skipped).
__global__ void k_scatter_vertices(int64_t nf, int64_t nv, const int32_t *__restrict__ F,
                                   const double *__restrict__ gfv, double *__restrict__ gv) {
  const int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (t >= nf * 9) return;
  const int64_t f = t / 9;
  const int c = int(t % 9), i = c / 3, d = c % 3;
  const int64_t id = F[3 * f + i];
  if (id < 0 || id >= nv) return;
  atomicAdd(gv + 3 * id + d, gfv[t]);
}

// ------------------------------------------------------------ single layer

template <typename T>
struct SlMath;

template <>
struct SlMath<float> {
  // 1/|d| from FP32 coordinates (hardware reciprocal square root).
  static __device__ inline float inv_r(float dx, float dy, float dz) {
    return __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
  }
};

template <>
struct SlMath<double> {
  static __device__ inline double inv_r(double dx, double dy, double dz) {
    return rsq64(dx * dx + dy * dy + dz * dz);
  }
};

// part[split][f][c0 + r] = sum_{s in split} X[s][c0 + r] / |C_f - S_s|
// (the chunk accumulators have the input's type, flushed to FP64 every
// kChunkTerms terms).
template <typename T, int K>
__global__ __launch_bounds__(kThreads) void k_sl_rows(const T *__restrict__ C,
                                                      const T *__restrict__ S,
                                                      const T *__restrict__ X, int64_t k,
                                                      int64_t c0, int64_t nf, int64_t n,
                                                      int64_t per_split,
                                                      double *__restrict__ part) {
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t s0 = int64_t(blockIdx.y) * per_split;
  const int64_t s1 = (s0 + per_split < n) ? s0 + per_split : n;
  const bool live = f < nf;
  const T cx = live ? C[3 * f] : T(1e10), cy = live ? C[3 * f + 1] : T(0),
          cz = live ? C[3 * f + 2] : T(0);
  double acc[K];
#pragma unroll
  for (int r = 0; r < K; ++r) acc[r] = 0.0;
  for (int64_t sc = s0; sc < s1; sc += kChunkTerms) {
    const int64_t se = (sc + kChunkTerms < s1) ? sc + kChunkTerms : s1;
    T a[K];
#pragma unroll
    for (int r = 0; r < K; ++r) a[r] = T(0);
    auto term = [&](int64_t j) {
      const T ir = SlMath<T>::inv_r(cx - S[3 * j], cy - S[3 * j + 1], cz - S[3 * j + 2]);
      const T *xj = X + j * k + c0;
#pragma unroll
      for (int r = 0; r < K; ++r) a[r] += ir * xj[r];
    };
    if (se - sc == kChunkTerms) {
#pragma unroll 4
      for (int t = 0; t < kChunkTerms; ++t) term(sc + t);
    } else {
      for (int64_t j = sc; j < se; ++j) term(j);
    }
#pragma unroll
    for (int r = 0; r < K; ++r) acc[r] += double(a[r]);
  }
  if (live) {
    double *dst = part + (int64_t(blockIdx.y) * nf + f) * k + c0;
#pragma unroll
    for (int r = 0; r < K; ++r) dst[r] = acc[r];
  }
}

// Backward, faces own: part3[slot][f] = -sum_{s in split} w (C_f - S_s) / r^3,
// w = sum_r X[s][c0 + r] Gp[f][c0 + r]; slot = chunk * splits + split.
template <typename T, int K>
__global__ __launch_bounds__(kThreads) void k_sl_bwd_faces(
    const T *__restrict__ C, const T *__restrict__ S, const T *__restrict__ X,
    const T *__restrict__ Gp, int64_t k, int64_t c0, int64_t chunk, int64_t nf, int64_t n,
    int64_t per_split, double *__restrict__ part3) {
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t s0 = int64_t(blockIdx.y) * per_split;
  const int64_t s1 = (s0 + per_split < n) ? s0 + per_split : n;
  const bool live = f < nf;
  const T cx = live ? C[3 * f] : T(1e10), cy = live ? C[3 * f + 1] : T(0),
          cz = live ? C[3 * f + 2] : T(0);
  T g[K];
#pragma unroll
  for (int r = 0; r < K; ++r) g[r] = live ? Gp[f * k + c0 + r] : T(0);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t sc = s0; sc < s1; sc += kChunkTerms) {
    const int64_t se = (sc + kChunkTerms < s1) ? sc + kChunkTerms : s1;
    T a[3] = {T(0), T(0), T(0)};
    auto term = [&](int64_t j) {
      const T dx = cx - S[3 * j], dy = cy - S[3 * j + 1], dz = cz - S[3 * j + 2];
      const T ir = SlMath<T>::inv_r(dx, dy, dz);
      const T *xj = X + j * k + c0;
      T w = T(0);
#pragma unroll
      for (int r = 0; r < K; ++r) w += xj[r] * g[r];
      const T c = w * (ir * ir * ir);
      a[0] += c * dx;
      a[1] += c * dy;
      a[2] += c * dz;
    };
    if (se - sc == kChunkTerms) {
#pragma unroll 4
      for (int t = 0; t < kChunkTerms; ++t) term(sc + t);
    } else {
      for (int64_t j = sc; j < se; ++j) term(j);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += double(a[d]);
  }
  if (live) {
    double *dst = part3 + ((chunk * gridDim.y + blockIdx.y) * nf + f) * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) dst[d] = -acc[d];
  }
}

// Backward, sources own: partx[split][s][c0 + r] = sum_{f in split} Gp[f][c0 + r] / r
// and part3[slot][s] = +sum_{f in split} w (C_f - S_s) / r^3.
template <typename T, int K>
__global__ __launch_bounds__(kThreads) void k_sl_bwd_sources(
    const T *__restrict__ C, const T *__restrict__ S, const T *__restrict__ X,
    const T *__restrict__ Gp, int64_t k, int64_t c0, int64_t chunk, int64_t nf, int64_t n,
    int64_t per_split, double *__restrict__ partx, double *__restrict__ part3) {
  const int64_t s = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t f0 = int64_t(blockIdx.y) * per_split;
  const int64_t f1 = (f0 + per_split < nf) ? f0 + per_split : nf;
  const bool live = s < n;
  const T px = live ? S[3 * s] : T(1e10), py = live ? S[3 * s + 1] : T(0),
          pz = live ? S[3 * s + 2] : T(0);
  T x[K];
#pragma unroll
  for (int r = 0; r < K; ++r) x[r] = live ? X[s * k + c0 + r] : T(0);
  double ax[K], acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < K; ++r) ax[r] = 0.0;
  for (int64_t fc = f0; fc < f1; fc += kChunkTerms) {
    const int64_t fe = (fc + kChunkTerms < f1) ? fc + kChunkTerms : f1;
    T a[3] = {T(0), T(0), T(0)};
    T b[K];
#pragma unroll
    for (int r = 0; r < K; ++r) b[r] = T(0);
    auto term = [&](int64_t j) {
      const T dx = C[3 * j] - px, dy = C[3 * j + 1] - py, dz = C[3 * j + 2] - pz;
      const T ir = SlMath<T>::inv_r(dx, dy, dz);
      const T *gj = Gp + j * k + c0;
      T w = T(0);
#pragma unroll
      for (int r = 0; r < K; ++r) {
        b[r] += gj[r] * ir;
        w += x[r] * gj[r];
      }
      const T c = w * (ir * ir * ir);
      a[0] += c * dx;
      a[1] += c * dy;
      a[2] += c * dz;
    };
    if (fe - fc == kChunkTerms) {
#pragma unroll 4
      for (int t = 0; t < kChunkTerms; ++t) term(fc + t);
    } else {
      for (int64_t j = fc; j < fe; ++j) term(j);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[d] += double(a[d]);
#pragma unroll
    for (int r = 0; r < K; ++r) ax[r] += double(b[r]);
  }
  if (live) {
    double *dx = partx + (int64_t(blockIdx.y) * n + s) * k + c0;
#pragma unroll
    for (int r = 0; r < K; ++r) dx[r] = ax[r];
    double *d3 = part3 + ((chunk * gridDim.y + blockIdx.y) * n + s) * 3;
#pragma unroll
    for (int d = 0; d < 3; ++d) d3[d] = acc[d];
  }
}

// out[i] = sum_{p < parts} part[p * count + i]
template <typename T>
__global__ void k_sum_parts(int64_t count, int64_t parts, const double *__restrict__ part,
                            T *__restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= count) return;
  double s = 0.0;
  for (int64_t p = 0; p < parts; ++p) s += part[p * count + i];
  out[i] = T(s);
}

inline unsigned blocks_for(int64_t count) { return unsigned((count + kThreads - 1) / kThreads); }

// Splits of the walked side so that a pass has about kCallerBlocks blocks,
// with at least 256 terms per split and at most 256 splits.
constexpr int64_t kCallerBlocks = 2048;
inline int64_t caller_splits(int64_t owners, int64_t others) {
  const int64_t ob = (owners + kThreads - 1) / kThreads;
  int64_t sp = (kCallerBlocks + ob - 1) / ob;
  const int64_t cap = (others + 255) / 256;
  if (sp > cap) sp = cap;
  if (sp > 256) sp = 256;
  return sp < 1 ? 1 : sp;
}

// f(IntK<K>{}) for K in 1..8 (right-hand sides per pass).
template <typename Fn>
hipError_t with_k(int K, Fn &&f) {
  switch (K) {
    case 1: return f(IntK<1>{});
    case 2: return f(IntK<2>{});
    case 3: return f(IntK<3>{});
    case 4: return f(IntK<4>{});
    case 5: return f(IntK<5>{});
    case 6: return f(IntK<6>{});
    case 7: return f(IntK<7>{});
    default: return f(IntK<8>{});
  }
}

}  // namespace

// ----------------------------------------------------------------- host side

size_t solid_angle_scratch_bytes(int64_t nv, int64_t nf, int64_t n) {
  // face records (FP64 at most) + gfv + gv + gs, each 256-byte aligned
  return size_t(nf) * kFaceRec * 8 + size_t(nf) * 9 * 8 + size_t(nv) * 3 * 8 +
         size_t(n) * 3 * 8 + 4 * 256;
}

template <typename T>
hipError_t solid_angle_fwd(hipStream_t st, const T *V, int64_t nv, const int32_t *F, int64_t nf,
                           const T *S, int64_t n, T *M, void *scratch, int32_t *bad) {
  if (nf == 0 || n == 0) return hipSuccess;
  T *rec = static_cast<T *>(scratch);
  k_pack_faces<T, T><<<blocks_for(nf), kThreads, 0, st>>>(V, nv, F, nf, rec, bad);
  const int64_t stiles = (n + kThreads - 1) / kThreads;
  const int64_t ftiles = (nf + kSaFaceTile - 1) / kSaFaceTile;
  k_solid_angle<T><<<unsigned(stiles * ftiles), kThreads, 0, st>>>(rec, S, nf, n, stiles, M);
  return hipGetLastError();
}

template <typename T>
hipError_t solid_angle_bwd(hipStream_t st, const T *V, int64_t nv, const int32_t *F, int64_t nf,
                           const T *S, int64_t n, const T *G, T *gV, T *gS, void *scratch,
                           int32_t *bad) {
  auto align = [](int64_t elems) { return (elems + 31) / 32 * 32; };
  double *rec = static_cast<double *>(scratch);
  double *gfv = rec + align(nf * kFaceRec);
  double *gv = gfv + align(nf * 9);
  double *gs = gv + align(nv * 3);
  const int64_t zero = (gs + n * 3) - gfv;
  if (zero > 0) k_fill<double><<<blocks_for(zero), kThreads, 0, st>>>(zero, gfv);
  if (nf > 0 && n > 0) {
    k_pack_faces<T, double><<<blocks_for(nf), kThreads, 0, st>>>(V, nv, F, nf, rec, bad);
    const int64_t per_block = int64_t(kThreads) * kSaSrcPerThread;
    const int64_t stiles = (n + per_block - 1) / per_block;
    // Face tile: about 2048 blocks in all, between 16 and 1024 faces.
    int64_t ftile = (nf * stiles + 2047) / 2048;
    ftile = ftile < 16 ? 16 : (ftile > 1024 ? 1024 : ftile);
    const int64_t ftiles = (nf + ftile - 1) / ftile;
    k_solid_angle_bwd<T><<<unsigned(stiles * ftiles), kThreads, 0, st>>>(rec, S, G, nf, n,
                                                                         stiles, ftile, gfv, gs);
    if (gV) k_scatter_vertices<<<blocks_for(nf * 9), kThreads, 0, st>>>(nf, nv, F, gfv, gv);
  }
  if (gV && nv > 0) k_cast<T><<<blocks_for(nv * 3), kThreads, 0, st>>>(nv * 3, gv, gV);
  if (gS && n > 0) k_cast<T><<<blocks_for(n * 3), kThreads, 0, st>>>(n * 3, gs, gS);
  return hipGetLastError();
}

size_t single_layer_scratch_bytes(int64_t nf, int64_t n, int64_t k) {
  const int64_t chunks = (k + 7) / 8;
  const int64_t sf = caller_splits(nf, n), ss = caller_splits(n, nf);
  const size_t fwd = size_t(sf) * nf * k;
  const size_t bwd = size_t(ss) * n * k + size_t(chunks) * ss * n * 3 +
                     size_t(chunks) * sf * nf * 3;
  return (fwd > bwd ? fwd : bwd) * 8 + 256;
}

template <typename T>
hipError_t single_layer_fwd(hipStream_t st, const T *S, int64_t n, const T *C, int64_t nf,
                            const T *X, int64_t k, T *phi, void *scratch) {
  if (nf == 0 || k == 0) return hipSuccess;
  double *part = static_cast<double *>(scratch);
  if (n == 0) {
    k_fill<T><<<blocks_for(nf * k), kThreads, 0, st>>>(nf * k, phi);
    return hipGetLastError();
  }
  const int64_t sp = caller_splits(nf, n);
  const int64_t per = (n + sp - 1) / sp;
  const dim3 grid(blocks_for(nf), unsigned(sp));
  for (int64_t c0 = 0; c0 < k; c0 += 8) {
    const int kk = int((k - c0 < 8) ? k - c0 : 8);
    hipError_t e = with_k(kk, [&](auto KC) {
      k_sl_rows<T, decltype(KC)::value><<<grid, kThreads, 0, st>>>(C, S, X, k, c0, nf, n, per,
                                                                    part);
      return hipGetLastError();
    });
    if (e != hipSuccess) return e;
  }
  k_sum_parts<T><<<blocks_for(nf * k), kThreads, 0, st>>>(nf * k, sp, part, phi);
  return hipGetLastError();
}

template <typename T>
hipError_t single_layer_bwd(hipStream_t st, const T *S, int64_t n, const T *C, int64_t nf,
                            const T *X, int64_t k, const T *Gp, T *gS, T *gC, T *gX,
                            void *scratch) {
  if (nf == 0 || n == 0 || k == 0) {
    if (gS && n > 0) k_fill<T><<<blocks_for(n * 3), kThreads, 0, st>>>(n * 3, gS);
    if (gX && n * k > 0) k_fill<T><<<blocks_for(n * k), kThreads, 0, st>>>(n * k, gX);
    if (gC && nf > 0) k_fill<T><<<blocks_for(nf * 3), kThreads, 0, st>>>(nf * 3, gC);
    return hipGetLastError();
  }
  const int64_t chunks = (k + 7) / 8;
  const int64_t sf = caller_splits(nf, n), ss = caller_splits(n, nf);
  double *partx = static_cast<double *>(scratch);
  double *parts3 = partx + ss * n * k;
  double *partf3 = parts3 + chunks * ss * n * 3;
  const int64_t per_s = (nf + ss - 1) / ss, per_f = (n + sf - 1) / sf;
  const bool src = gS || gX, face = gC != nullptr;
  for (int64_t c0 = 0, ch = 0; c0 < k; c0 += 8, ++ch) {
    const int kk = int((k - c0 < 8) ? k - c0 : 8);
    hipError_t e = with_k(kk, [&](auto KC) {
      constexpr int K = decltype(KC)::value;
      if (src)
        k_sl_bwd_sources<T, K><<<dim3(blocks_for(n), unsigned(ss)), kThreads, 0, st>>>(
            C, S, X, Gp, k, c0, ch, nf, n, per_s, partx, parts3);
      if (face)
        k_sl_bwd_faces<T, K><<<dim3(blocks_for(nf), unsigned(sf)), kThreads, 0, st>>>(
            C, S, X, Gp, k, c0, ch, nf, n, per_f, partf3);
      return hipGetLastError();
    });
    if (e != hipSuccess) return e;
  }
  if (gX) k_sum_parts<T><<<blocks_for(n * k), kThreads, 0, st>>>(n * k, ss, partx, gX);
  if (gS) k_sum_parts<T><<<blocks_for(n * 3), kThreads, 0, st>>>(n * 3, chunks * ss, parts3, gS);
  if (gC)
    k_sum_parts<T><<<blocks_for(nf * 3), kThreads, 0, st>>>(nf * 3, chunks * sf, partf3, gC);
  return hipGetLastError();
}

#define MFS_CALLER_INST(T)                                                                      \
  template hipError_t solid_angle_fwd<T>(hipStream_t, const T *, int64_t, const int32_t *,     \
                                         int64_t, const T *, int64_t, T *, void *, int32_t *);  \
  template hipError_t solid_angle_bwd<T>(hipStream_t, const T *, int64_t, const int32_t *,     \
                                         int64_t, const T *, int64_t, const T *, T *, T *,      \
                                         void *, int32_t *);                                    \
  template hipError_t single_layer_fwd<T>(hipStream_t, const T *, int64_t, const T *, int64_t, \
                                          const T *, int64_t, T *, void *);                     \
  template hipError_t single_layer_bwd<T>(hipStream_t, const T *, int64_t, const T *, int64_t, \
                                          const T *, int64_t, const T *, T *, T *, T *, void *);
MFS_CALLER_INST(float)
MFS_CALLER_INST(double)

}  // namespace mfs
