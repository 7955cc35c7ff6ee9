// k_points.hip -- matrix-free kernel matvecs for the points form.
//
// The reference multiplies by a stored dense M (src/kl_chol.cu:928-936,
// 992-1002), assembled from the point kernel (src/kernels.h:24-40,
// src/kl_chol.cu:55-126).  Here the kernel is evaluated on the fly and no
// m x n matrix exists, so the solve fits where M would not (N = 262,144 would
// need 275 GB in FP32).
//
// Both passes are N-body loops.  A lane owns one target (rows pass, y = M p)
// or one source (cols pass, q = M^T y) and keeps its point in registers; the
// other side is walked with block-uniform indices, so its records and
// right-hand-side values arrive through scalar loads and enter the VALU as
// SGPR operands -- the inner loop is pure arithmetic: ~22 VALU slots for the
// kernel (hi/lo difference, r^2, rsqrt, d.n, r^-3) plus one FMA per
// right-hand side.  The other side is split over gridDim.y so that the chip
// holds >= 4096 blocks; splits write FP64 partials that a second kernel sums.
//
// Sums run in FP32 over 64-term chunks and in FP64 across chunks (cascaded
// summation: the FP32 part never adds more than 64 terms).  The FP64 passes,
// used for the certified residual of the refinement, evaluate and sum in FP64
// (division-free evaluation, kernel_eval.h dlp64).
#include "kernel_eval.h"
#include "mfs_internal.h"

namespace mfs {
namespace {

constexpr int kChunkTerms = 64;

__device__ inline bool batch_done(const PcgState *st) { return st != nullptr && st->n_active == 0; }

// part[s][f][r] = sum_{j in split s} K(f, j) p[j][r]   (FP32 kernel, K rhs)
template <int K>
__global__ __launch_bounds__(kThreads) void k_pts_rows32(
    const float *__restrict__ tgt, const float *__restrict__ src, int64_t m, int64_t n,
    const float *__restrict__ pv, int64_t per_split, double *__restrict__ part,
    const PcgState *__restrict__ st) {
  if (batch_done(st)) return;
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t j0 = int64_t(blockIdx.y) * per_split;
  const int64_t j1 = (j0 + per_split < n) ? j0 + per_split : n;
  float xh0 = 1e15f, xh1 = 0.f, xh2 = 0.f, xl0 = 0.f, xl1 = 0.f, xl2 = 0.f;
  float n0 = 0.f, n1 = 0.f, n2 = 0.f;  // padding lanes: far away, zero normal
  if (f < m) {
    const float4 a = *reinterpret_cast<const float4 *>(tgt + f * kTgt32);
    const float4 b = *reinterpret_cast<const float4 *>(tgt + f * kTgt32 + 4);
    const float4 c = *reinterpret_cast<const float4 *>(tgt + f * kTgt32 + 8);
    xh0 = a.x; xh1 = a.y; xh2 = a.z; xl0 = a.w; xl1 = b.x; xl2 = b.y;
    n0 = b.z; n1 = b.w; n2 = c.x;
  }
  double acc[K];
#pragma unroll
  for (int r = 0; r < K; ++r) acc[r] = 0.0;
  for (int64_t jc = j0; jc < j1; jc += kChunkTerms) {
    const int64_t je = (jc + kChunkTerms < j1) ? jc + kChunkTerms : j1;
    float c[K];
#pragma unroll
    for (int r = 0; r < K; ++r) c[r] = 0.f;
    auto term = [&](int64_t j) {
      const float *s = src + j * kSrc32;
      const float kv = dlp32((xh0 - s[0]) + (xl0 - s[3]), (xh1 - s[1]) + (xl1 - s[4]),
                             (xh2 - s[2]) + (xl2 - s[5]), n0, n1, n2);
      const float *pj = pv + j * K;
#pragma unroll
      for (int r = 0; r < K; ++r) c[r] += kv * pj[r];
    };
    if (je - jc == kChunkTerms) {
#pragma unroll 4
      for (int t = 0; t < kChunkTerms; ++t) term(jc + t);
    } else {
      for (int64_t j = jc; j < je; ++j) term(j);
    }
#pragma unroll
    for (int r = 0; r < K; ++r) acc[r] += double(c[r]);
  }
  if (f < m) {
    double *dst = part + (int64_t(blockIdx.y) * m + f) * K;
#pragma unroll
    for (int r = 0; r < K; ++r) dst[r] = -kInv4Pi * acc[r];
  }
}

// part[s][j][r] = sum_{f in split s} K(f, j) w[f][r]   (FP32 kernel, K rhs)
template <int K>
__global__ __launch_bounds__(kThreads) void k_pts_cols32(
    const float *__restrict__ tgt, const float *__restrict__ src, int64_t m, int64_t n,
    const float *__restrict__ w, int64_t per_split, double *__restrict__ part,
    const PcgState *__restrict__ st) {
  if (batch_done(st)) return;
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t f0 = int64_t(blockIdx.y) * per_split;
  const int64_t f1 = (f0 + per_split < m) ? f0 + per_split : m;
  float yh0 = 1e15f, yh1 = 0.f, yh2 = 0.f, yl0 = 0.f, yl1 = 0.f, yl2 = 0.f;
  if (j < n) {
    const float4 a = *reinterpret_cast<const float4 *>(src + j * kSrc32);
    const float4 b = *reinterpret_cast<const float4 *>(src + j * kSrc32 + 4);
    yh0 = a.x; yh1 = a.y; yh2 = a.z; yl0 = a.w; yl1 = b.x; yl2 = b.y;
  }
  double acc[K];
#pragma unroll
  for (int r = 0; r < K; ++r) acc[r] = 0.0;
  for (int64_t fc = f0; fc < f1; fc += kChunkTerms) {
    const int64_t fe = (fc + kChunkTerms < f1) ? fc + kChunkTerms : f1;
    float c[K];
#pragma unroll
    for (int r = 0; r < K; ++r) c[r] = 0.f;
    auto term = [&](int64_t f) {
      const float *t = tgt + f * kTgt32;
      const float kv = dlp32((t[0] - yh0) + (t[3] - yl0), (t[1] - yh1) + (t[4] - yl1),
                             (t[2] - yh2) + (t[5] - yl2), t[6], t[7], t[8]);
      const float *wf = w + f * K;
#pragma unroll
      for (int r = 0; r < K; ++r) c[r] += kv * wf[r];
    };
    if (fe - fc == kChunkTerms) {
#pragma unroll 4
      for (int t = 0; t < kChunkTerms; ++t) term(fc + t);
    } else {
      for (int64_t f = fc; f < fe; ++f) term(f);
    }
#pragma unroll
    for (int r = 0; r < K; ++r) acc[r] += double(c[r]);
  }
  if (j < n) {
    double *dst = part + (int64_t(blockIdx.y) * n + j) * K;
#pragma unroll
    for (int r = 0; r < K; ++r) dst[r] = -kInv4Pi * acc[r];
  }
}

// FP64 rows pass: part[s][f][r] = sum_j K(f, j) X[j][r].
template <int K>
__global__ __launch_bounds__(kThreads) void k_pts_rows64(
    const double *__restrict__ tgt, const double *__restrict__ src, int64_t m, int64_t n,
    const double *__restrict__ X, int64_t per_split, double *__restrict__ part) {
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t j0 = int64_t(blockIdx.y) * per_split;
  const int64_t j1 = (j0 + per_split < n) ? j0 + per_split : n;
  double x0 = 1e100, x1 = 0, x2 = 0, n0 = 0, n1 = 0, n2 = 0;
  if (f < m) {
    const double *t = tgt + f * kTgt64;
    x0 = t[0]; x1 = t[1]; x2 = t[2]; n0 = t[3]; n1 = t[4]; n2 = t[5];
  }
  double acc[K];
#pragma unroll
  for (int r = 0; r < K; ++r) acc[r] = 0.0;
#pragma unroll 2
  for (int64_t j = j0; j < j1; ++j) {
    const double *s = src + j * kSrc64;
    const double kv = dlp64(x0 - s[0], x1 - s[1], x2 - s[2], n0, n1, n2);
#pragma unroll
    for (int r = 0; r < K; ++r) acc[r] += kv * X[j * K + r];
  }
  if (f < m) {
    double *dst = part + (int64_t(blockIdx.y) * m + f) * K;
#pragma unroll
    for (int r = 0; r < K; ++r) dst[r] = -kInv4Pi * acc[r];
  }
}

// FP64 cols pass: part[s][j][r] = sum_f K(f, j) Y[f][r].
template <int K>
__global__ __launch_bounds__(kThreads) void k_pts_cols64(
    const double *__restrict__ tgt, const double *__restrict__ src, int64_t m, int64_t n,
    const double *__restrict__ Y, int64_t per_split, double *__restrict__ part) {
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t f0 = int64_t(blockIdx.y) * per_split;
  const int64_t f1 = (f0 + per_split < m) ? f0 + per_split : m;
  double y0 = 1e100, y1 = 0, y2 = 0;
  if (j < n) {
    const double *s = src + j * kSrc64;
    y0 = s[0]; y1 = s[1]; y2 = s[2];
  }
  double acc[K];
#pragma unroll
  for (int r = 0; r < K; ++r) acc[r] = 0.0;
#pragma unroll 2
  for (int64_t f = f0; f < f1; ++f) {
    const double *t = tgt + f * kTgt64;
    const double kv = dlp64(t[0] - y0, t[1] - y1, t[2] - y2, t[3], t[4], t[5]);
#pragma unroll
    for (int r = 0; r < K; ++r) acc[r] += kv * Y[f * K + r];
  }
  if (j < n) {
    double *dst = part + (int64_t(blockIdx.y) * n + j) * K;
#pragma unroll
    for (int r = 0; r < K; ++r) dst[r] = -kInv4Pi * acc[r];
  }
}

// Y[i][r] = sum over splits (or b - sum when Bsub is given), stored as TA.
template <typename TA, int K>
__global__ __launch_bounds__(kThreads) void k_pts_reduce(int64_t count, int32_t splits,
                                                         const double *__restrict__ part,
                                                         const float *__restrict__ Bsub,
                                                         int64_t ldb, int64_t col0, int kval,
                                                         TA *__restrict__ Y,
                                                         const PcgState *__restrict__ st) {
  if (batch_done(st)) return;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= count) return;
  double s[K];
#pragma unroll
  for (int r = 0; r < K; ++r) s[r] = 0.0;
  for (int32_t b = 0; b < splits; ++b) {
    const double *p = part + (int64_t(b) * count + i) * K;
#pragma unroll
    for (int r = 0; r < K; ++r) s[r] += p[r];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    if (Bsub != nullptr) {
      const double b = (r < kval) ? double(Bsub[i * ldb + col0 + r]) : 0.0;
      Y[i * K + r] = TA(b - s[r]);
    } else {
      Y[i * K + r] = TA(s[r]);
    }
  }
}

__global__ void k_pack_targets(int64_t m, const double *__restrict__ x,
                               const double *__restrict__ nx, float *__restrict__ t32,
                               double *__restrict__ t64) {
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (f >= m) return;
  for (int d = 0; d < 3; ++d) {
    const double v = x[3 * f + d];
    const float hi = float(v);
    t32[f * kTgt32 + d] = hi;
    t32[f * kTgt32 + 3 + d] = float(v - double(hi));
    t32[f * kTgt32 + 6 + d] = float(nx[3 * f + d]);
    t32[f * kTgt32 + 9 + d] = 0.f;
    t64[f * kTgt64 + d] = v;
    t64[f * kTgt64 + 3 + d] = nx[3 * f + d];
  }
  t64[f * kTgt64 + 6] = 0.0;
  t64[f * kTgt64 + 7] = 0.0;
}

__global__ void k_pack_sources(int64_t n, const double *__restrict__ y, float *__restrict__ s32,
                               double *__restrict__ s64) {
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= n) return;
  for (int d = 0; d < 3; ++d) {
    const double v = y[3 * j + d];
    const float hi = float(v);
    s32[j * kSrc32 + d] = hi;
    s32[j * kSrc32 + 3 + d] = float(v - double(hi));
    s64[j * kSrc64 + d] = v;
  }
  s32[j * kSrc32 + 6] = 0.f;
  s32[j * kSrc32 + 7] = 0.f;
  s64[j * kSrc64 + 3] = 0.0;
}

inline unsigned blocks_for(int64_t count) { return unsigned((count + kThreads - 1) / kThreads); }

#define PTS_DISPATCH(K, CALL)                        \
  switch (K) {                                       \
    case 1: { constexpr int KK = 1; CALL; break; }   \
    case 2: { constexpr int KK = 2; CALL; break; }   \
    case 4: { constexpr int KK = 4; CALL; break; }   \
    case 8: { constexpr int KK = 8; CALL; break; }   \
    case 16: { constexpr int KK = 16; CALL; break; } \
    default: return hipErrorInvalidValue;            \
  }

}  // namespace

int32_t points_splits(int64_t owners, int64_t others, int32_t target_blocks) {
  const int64_t blocks = (owners + kThreads - 1) / kThreads;
  const int64_t target = target_blocks > 0 ? target_blocks : kPointsBlocks;
  int64_t s = (target + blocks - 1) / blocks;
  const int64_t cap = (others + 255) / 256;  // at least 256 terms per split
  if (s > cap) s = cap;
  if (s > 256) s = 256;
  return int32_t(s < 1 ? 1 : s);
}

hipError_t pack_points(hipStream_t st, const double *x, const double *nx, const double *y,
                       int64_t m, int64_t n, float *t32, double *t64, float *s32, double *s64) {
  k_pack_targets<<<blocks_for(m), kThreads, 0, st>>>(m, x, nx, t32, t64);
  k_pack_sources<<<blocks_for(n), kThreads, 0, st>>>(n, y, s32, s64);
  return hipGetLastError();
}

template <typename TA>
hipError_t points_rows(hipStream_t st, const PointsOp &op, int K, const TA *X, const float *Bsub,
                       int64_t ldb, int64_t col0, int kval, TA *Y, double *part,
                       const PcgState *state) {
  const int32_t S = points_splits(op.m, op.n, op.target_blocks);
  const int64_t per = (op.n + S - 1) / S;
  const dim3 grid(blocks_for(op.m), unsigned(S));
  if constexpr (sizeof(TA) == 4) {
    PTS_DISPATCH(K, (k_pts_rows32<KK><<<grid, kThreads, 0, st>>>(op.tgt32, op.src32, op.m, op.n,
                                                                   X, per, part, state)));
  } else {
    PTS_DISPATCH(K, (k_pts_rows64<KK><<<grid, kThreads, 0, st>>>(op.tgt64, op.src64, op.m, op.n,
                                                                   X, per, part)));
  }
  PTS_DISPATCH(K, (k_pts_reduce<TA, KK><<<blocks_for(op.m), kThreads, 0, st>>>(
                      op.m, S, part, Bsub, ldb, col0, kval, Y, state)));
  return hipGetLastError();
}

template <typename TA>
hipError_t points_cols(hipStream_t st, const PointsOp &op, int K, const TA *Y, double *part,
                       const PcgState *state) {
  const int32_t S = points_splits(op.n, op.m, op.target_blocks);
  const int64_t per = (op.m + S - 1) / S;
  const dim3 grid(blocks_for(op.n), unsigned(S));
  if constexpr (sizeof(TA) == 4) {
    PTS_DISPATCH(K, (k_pts_cols32<KK><<<grid, kThreads, 0, st>>>(op.tgt32, op.src32, op.m, op.n,
                                                                   Y, per, part, state)));
  } else {
    PTS_DISPATCH(K, (k_pts_cols64<KK><<<grid, kThreads, 0, st>>>(op.tgt64, op.src64, op.m, op.n,
                                                                   Y, per, part)));
  }
  return hipGetLastError();
}

template <typename TA>
hipError_t points_cols_reduce(hipStream_t st, const PointsOp &op, int K, const double *part,
                              TA *out) {
  const int32_t S = points_splits(op.n, op.m, op.target_blocks);
  PTS_DISPATCH(K, (k_pts_reduce<TA, KK><<<blocks_for(op.n), kThreads, 0, st>>>(
                      op.n, S, part, nullptr, 0, 0, K, out, nullptr)));
  return hipGetLastError();
}

template hipError_t points_rows<float>(hipStream_t, const PointsOp &, int, const float *,
                                       const float *, int64_t, int64_t, int, float *, double *,
                                       const PcgState *);
template hipError_t points_rows<double>(hipStream_t, const PointsOp &, int, const double *,
                                        const float *, int64_t, int64_t, int, double *, double *,
                                        const PcgState *);
template hipError_t points_cols<float>(hipStream_t, const PointsOp &, int, const float *, double *,
                                       const PcgState *);
template hipError_t points_cols<double>(hipStream_t, const PointsOp &, int, const double *,
                                        double *, const PcgState *);
template hipError_t points_cols_reduce<float>(hipStream_t, const PointsOp &, int, const double *,
                                              float *);
template hipError_t points_cols_reduce<double>(hipStream_t, const PointsOp &, int, const double *,
                                               double *);

}  // namespace mfs
