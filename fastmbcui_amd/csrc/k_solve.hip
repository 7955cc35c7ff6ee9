// k_solve.hip -- PCG on M^T M with the supernodal KL preconditioner z = L (L^T r).
//
// Replaces pcg_guesss / pcg_VpL_guesss (src/kl_chol.cu:891-1264), whose every
// step was a separate cuBLAS/cuSPARSE call with a host round trip per dot.
// Here a batch of K right-hand sides moves through seven kernels per
// iteration, the dots are FP64 block partials finished by the last block to
// arrive, and alpha/beta/convergence live in device memory (PcgState), so the
// host only polls a flag every few iterations.
//
//   y = M p          k_mv_rows     one HBM pass over M for the whole batch
//   part = M^T y     k_mv_cols     second HBM pass, row blocks -> partials
//   q = sum(part)    k_mv_reduce   + dot(p, q)
//   x += a p, r -= a q            k_upd_xr   + dot(r, r), convergence
//   t = L^T r        k_spmv (CSC of L read as CSR of L^T)
//   z = L t          k_spmv (CSR of L)  + dot(r, z)
//   p = z + b p      k_upd_p    (+ copy of p in original order for the matvec)
#include <type_traits>

#include "kernel_eval.h"
#include "mfs_internal.h"

namespace mfs {
namespace {

// ------------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum K values over the block into red[0..K) (valid after the call, all threads).
template <int K>
__device__ inline void block_sum(double (&v)[K], double *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const double s = wave_sum(v[r]);
    if (lane == 0) red[wave * K + r] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int r = threadIdx.x;
    red[4 * K + r] = ((red[r] + red[K + r]) + red[2 * K + r]) + red[3 * K + r];
  }
  __syncthreads();
}

// Publish this block's K partials, then report whether it is the last block to
// arrive.  Thread 0 stores the partials and releases them with a device-scope
// fence before the arrival atomic; every thread of the last block fences again
// (acquire) before reading, so partials written on another XCD (own L2) are
// seen.  The last block re-arms the counter, so launches replay without a
// memset.
template <int K>
__device__ inline bool publish_and_arrive(const double *blocksum, double *dpart, uint32_t *counter,
                                          uint32_t nblocks, int *flag) {
  if (threadIdx.x == 0) {
    for (int r = 0; r < K; ++r) dpart[size_t(blockIdx.x) * K + r] = blocksum[r];
    __threadfence();
    const uint32_t prev = atomicAdd(counter, 1u);
    *flag = (prev == nblocks - 1);
    if (*flag) atomicExch(counter, 0u);
  }
  __syncthreads();
  if (*flag) __threadfence();
  return *flag != 0;
}

// Sum the published partials with the whole (last) block: thread t takes
// blocks t, t+256, ..., then a fixed-shape block reduction.  Deterministic;
// totals land in red[4K .. 5K) for every thread.
template <int K>
__device__ inline void gather_partials(const double *dpart, uint32_t nblocks, double *red) {
  double v[K];
#pragma unroll
  for (int r = 0; r < K; ++r) v[r] = 0.0;
  for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int r = 0; r < K; ++r) v[r] += dpart[size_t(b) * K + r];
  block_sum<K>(v, red);
}

__device__ inline bool batch_done(const PcgState *st) { return st != nullptr && st->n_active == 0; }

// ------------------------------------------------------------------ y = M x
// A block of 8 waves owns 8*RW rows; each wave owns RW rows and its lanes
// stream float4 of M along them.  x is staged in LDS in chunks of CH columns,
// right-hand-side major (xs[r][c]) so a lane's four columns for one
// right-hand side are one conflict-free 16-byte read, double-buffered so the
// next chunk is staged while this one is consumed: one barrier per chunk.
// Optional Bsub gives y = b - M x.
constexpr int kRowsThreads = 512;
constexpr int kF32x8Rows = 4;    // k_mv_rows_f32x8: rows per wave
constexpr int kF32x8Depth = 2;   // k_mv_rows_f32x8: x chunks per group, 2 steps each
constexpr int kF64x8Rows = 4;    // k_mv_rows_f64x8: rows per wave
constexpr int kF64x8Depth = 2;   // k_mv_rows_f64x8: x chunks per group, 2 steps each

// M is streamed once per pass: non-temporal loads keep it from displacing the
// vectors (x, y, partials) that are reused from L2.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ inline float4 load_stream4(const float *p) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

template <typename TA, int K>
struct RowsCfg {
  // Rows per wave: each x value read from LDS serves RW rows.  FP32 batches
  // of 8 take 8 rows (240 VGPRs, no spill; 2.59 ms instead of 2.88 ms per
  // pass at m = n = 65,536); wider rows spill once K * sizeof(TA) >= 64.
  static constexpr int RW =
      (K == 1) ? 8 : (K == 8 && sizeof(TA) == 4 ? 8 : (K * int(sizeof(TA)) >= 64 ? 2 : 4));
  // two buffers of K*CH*sizeof(TA) within 64 KB, CH a multiple of 256, at most 2048
  static constexpr int CHB = 32768 / (K * int(sizeof(TA)));
  static constexpr int CH = CHB >= 2048 ? 2048 : (CHB >= 1024 ? 1024 : (CHB >= 512 ? 512 : 256));
};

template <typename TA, int K, int RW, bool VEC>
__global__ __launch_bounds__(kRowsThreads) void k_mv_rows(const float *__restrict__ M, int64_t m,
                                                          int64_t n, const TA *__restrict__ X,
                                                          const float *__restrict__ Bsub,
                                                          int64_t ldb, int64_t col0, int kval,
                                                          TA *__restrict__ Y,
                                                          const PcgState *st) {
  if (batch_done(st)) return;
  constexpr int CH = RowsCfg<TA, K>::CH;
  __shared__ __align__(16) TA xs[2][K * CH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = (int64_t(blockIdx.x) * (kRowsThreads / 64) + wave) * RW;
  TA acc[RW][K];
#pragma unroll
  for (int i = 0; i < RW; ++i)
#pragma unroll
    for (int r = 0; r < K; ++r) acc[i][r] = TA(0);

  // x chunk: each thread owns XPT values; the next chunk is loaded into
  // registers before this chunk's M loads are issued and written to LDS after
  // the chunk is consumed, so the L2 latency of x hides behind HBM traffic.
  constexpr int XPT = CH * K / kRowsThreads;
  TA xr_next[XPT];
  auto fetch = [&](int64_t c0) {
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      const int idx = threadIdx.x + j * kRowsThreads;
      const int cl = idx / K, r = idx % K;
      const int64_t c = c0 + cl;
      xr_next[j] = (c < n) ? X[c * K + r] : TA(0);
    }
  };
  auto store = [&](TA *dst) {
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      const int idx = threadIdx.x + j * kRowsThreads;
      const int cl = idx / K, r = idx % K;
      dst[r * CH + cl] = xr_next[j];
    }
  };
  fetch(0);
  store(xs[0]);
  __syncthreads();
  int buf = 0;
  for (int64_t c0 = 0; c0 < n; c0 += CH) {
    const bool more = c0 + CH < n;
    if (more) fetch(c0 + CH);
    const TA *xb = xs[buf];
#pragma unroll 2
    for (int u = 0; u < CH / 256; ++u) {
      // VEC (n % 4 == 0): a lane owns 4 consecutive columns (one float4).
      // Otherwise rows are not 16-byte aligned and a lane owns columns
      // lane, lane + 64, lane + 128, lane + 192 of the 256-column step, so
      // each scalar load instruction still reads 256 contiguous bytes.
      const int cl = VEC ? u * 256 + lane * 4 : u * 256 + lane;
      constexpr int cs = VEC ? 1 : 64;  // column stride between a lane's 4 values
      const int64_t cc = c0 + cl;
      float4 mv[RW];
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const int64_t f = row0 + i;
        const float *rowp = M + f * n;
        if (f < m && VEC && cc + 3 < n) {
          mv[i] = load_stream4(rowp + cc);
        } else if (f < m) {
          mv[i].x = (cc + 0 * cs < n) ? rowp[cc + 0 * cs] : 0.f;
          mv[i].y = (cc + 1 * cs < n) ? rowp[cc + 1 * cs] : 0.f;
          mv[i].z = (cc + 2 * cs < n) ? rowp[cc + 2 * cs] : 0.f;
          mv[i].w = (cc + 3 * cs < n) ? rowp[cc + 3 * cs] : 0.f;
        } else {
          mv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
#pragma unroll
      for (int r = 0; r < K; ++r) {
        const TA *xr = xb + r * CH + cl;
        const TA x0 = xr[0], x1 = xr[cs], x2 = xr[2 * cs], x3 = xr[3 * cs];
#pragma unroll
        for (int i = 0; i < RW; ++i) {
          TA s = acc[i][r];
          s += TA(mv[i].x) * x0;
          s += TA(mv[i].y) * x1;
          s += TA(mv[i].z) * x2;
          s += TA(mv[i].w) * x3;
          acc[i][r] = s;
        }
      }
    }
    if (more) store(xs[buf ^ 1]);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    const int64_t f = row0 + i;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const TA s = wave_sum(acc[i][r]);
      if (lane == 0 && f < m) {
        if (Bsub != nullptr) {
          const TA b = (r < kval) ? TA(Bsub[f * ldb + col0 + r]) : TA(0);
          Y[f * K + r] = b - s;
        } else {
          Y[f * K + r] = s;
        }
      }
    }
  }
}

// FP64 batches of 8 (the certified residual y = b - M x and its kin).  In
// k_mv_rows these take two rows per wave: each 16-byte x value read from LDS
// serves only two rows (8 bytes of LDS per byte of M, the LDS pipe's limit)
// and a wave has 2 KB of M in flight between barriers (3.4 TB/s at m = n =
// 65,536).  Here a wave takes RW = 4 rows and keeps D = 2 P steps of 256
// columns in flight: the loads of step t + D are issued as soon as step t is
// computed, across the x-chunk barriers too (M loads do not depend on x).
// Loads complete in order, so waiting for an x fetch also waits for every M
// load issued before it: the x of chunk h is fetched P - 1 chunks ahead (P
// register sets), behind M loads that are the next ones needed anyway.  The
// op sequence is the same on every step (rows clamped to m - 1, columns to
// n - 4, x fetches to column n - 1, zeroed where used), so the compiler's
// wait counts are exact.  Steps past the row (the last P chunks may overhang
// n) add zeros.  Needs n % 4 == 0.  RW = 4, P = 2: 244 VGPRs, 2.78 ms at
// 65,536^2 (6.2 TB/s); RW = 2 with P = 2..4 stays at 4.0 TB/s.
template <int RW, int P>
__global__ __launch_bounds__(kRowsThreads) void k_mv_rows_f64x8(const float *__restrict__ M, int64_t m,
                                                                int64_t n, const double *__restrict__ X,
                                                                const float *__restrict__ Bsub,
                                                                int64_t ldb, int64_t col0, int kval,
                                                                double *__restrict__ Y,
                                                                const PcgState *st) {
  if (batch_done(st)) return;
  constexpr int K = 8, D = 2 * P;
  constexpr int CH = RowsCfg<double, K>::CH;
  static_assert(CH == 512, "two 256-column steps per x chunk");
  static_assert(P >= 2, "the x of the next chunk is fetched at least one chunk ahead");
  constexpr int XPT = CH * K / kRowsThreads;
  __shared__ __align__(16) double xs[2][K * CH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = (int64_t(blockIdx.x) * (kRowsThreads / 64) + wave) * RW;
  const float *rowp[RW];
  bool rok[RW];
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    rok[i] = row0 + i < m;
    rowp[i] = M + (rok[i] ? row0 + i : m - 1) * n;
  }
  double acc[RW][K];
#pragma unroll
  for (int i = 0; i < RW; ++i)
#pragma unroll
    for (int r = 0; r < K; ++r) acc[i][r] = 0.0;

  auto loadm = [&](int64_t c, float4 (&mv)[RW]) {  // c: this lane's first column
    const int64_t cc = c < n ? c : n - 4;
#pragma unroll
    for (int i = 0; i < RW; ++i) mv[i] = load_stream4(rowp[i] + cc);
  };
  // x in LDS: column c of row r at r CH + (c & 2) CH / 4 + 2 (c >> 2) + (c & 1),
  // so a lane's pairs (4q, 4q + 1) and (4q + 2, 4q + 3) sit in two 16-byte
  // slots that are consecutive across lanes: bank-conflict-free 128-bit reads
  // (the plain [r][c] layout puts lanes 32 bytes apart, two-way conflicts).
  auto xslot = [](int r, int c) { return r * CH + (c & 2) * (CH / 4) + 2 * (c >> 2) + (c & 1); };
  auto compute = [&](int64_t c, int cl, const float4 (&mv)[RW], const double *xb) {
    const bool cok = c < n;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const double2 xa = *reinterpret_cast<const double2 *>(xb + xslot(r, cl));
      const double2 xc = *reinterpret_cast<const double2 *>(xb + xslot(r, cl + 2));
      const double x0 = xa.x, x1 = xa.y, x2 = xc.x, x3 = xc.y;
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const bool use = rok[i] && cok;
        double s = acc[i][r];
        s += double(use ? mv[i].x : 0.f) * x0;
        s += double(use ? mv[i].y : 0.f) * x1;
        s += double(use ? mv[i].z : 0.f) * x2;
        s += double(use ? mv[i].w : 0.f) * x3;
        acc[i][r] = s;
      }
    }
  };
  auto fetch = [&](int64_t c0, double (&xq)[XPT]) {
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      const int idx = threadIdx.x + j * kRowsThreads;
      const int64_t c = c0 + idx / K;
      xq[j] = X[(c < n ? c : n - 1) * K + idx % K];
    }
  };
  auto store = [&](int64_t c0, const double (&xq)[XPT], double *dst) {
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      const int idx = threadIdx.x + j * kRowsThreads;
      dst[xslot(idx % K, idx / K)] = (c0 + idx / K < n) ? xq[j] : 0.0;
    }
  };
  float4 mv[D][RW];     // step t lives in slot t % D
  double xq[P][XPT];    // x of chunk h lives in set h % P
#pragma unroll
  for (int d = 0; d < D; ++d) loadm(d * 256 + 4 * lane, mv[d]);
#pragma unroll
  for (int h = 0; h < P - 1; ++h) fetch(h * CH, xq[h]);
  store(0, xq[0], xs[0]);
  __syncthreads();
  int buf = 0;
  // scheduling fences keep each phase in place: the machine scheduler would
  // otherwise hoist the FMAs above the loads and wait for them
  for (int64_t c0 = 0; c0 < n; c0 += D * 256) {
#pragma unroll
    for (int h = 0; h < P; ++h) {
      const int64_t cc = c0 + h * CH;      // this x chunk: slots 2h, 2h + 1
      fetch(cc + (P - 1) * CH, xq[(h + P - 1) % P]);   // past the row: clamped, never read
      __builtin_amdgcn_sched_barrier(0);
      compute(cc + 4 * lane, 4 * lane, mv[2 * h], xs[buf]);
      __builtin_amdgcn_sched_barrier(0);
      loadm(cc + D * 256 + 4 * lane, mv[2 * h]);
      __builtin_amdgcn_sched_barrier(0);
      compute(cc + 256 + 4 * lane, 256 + 4 * lane, mv[2 * h + 1], xs[buf]);
      __builtin_amdgcn_sched_barrier(0);
      loadm(cc + 256 + D * 256 + 4 * lane, mv[2 * h + 1]);
      __builtin_amdgcn_sched_barrier(0);
      store(cc + CH, xq[(h + 1) % P], xs[buf ^ 1]);
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    const int64_t f = row0 + i;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const double sum = wave_sum(acc[i][r]);
      if (lane == 0 && f < m) {
        if (Bsub != nullptr) {
          const double bv = (r < kval) ? double(Bsub[f * ldb + col0 + r]) : 0.0;
          Y[f * K + r] = bv - sum;
        } else {
          Y[f * K + r] = sum;
        }
      }
    }
  }
}

// FP32 batches of 8 (the PCG's y = M p), n % 4 == 0: k_mv_rows' sums (each
// lane's columns in the same order, then the wave sum) with the same
// pipelining as k_mv_rows_f64x8 -- 4 rows per wave, D = 2 P steps of 256
// columns in flight across the x-chunk barriers, x fetched P - 1 chunks
// ahead.  FP32 x values sit in LDS as [r][c]: a lane's four columns are one
// 16-byte read and lanes are 16 bytes apart, so no swizzle is needed.  At
// m = n = 65,536: 2.44 ms (7.05 TB/s) against 2.55 ms for k_mv_rows with 8
// rows per wave (which holds 2 KB of M per wave between barriers).
template <int RW, int P>
__global__ __launch_bounds__(kRowsThreads) void k_mv_rows_f32x8(const float *__restrict__ M, int64_t m,
                                                                int64_t n, const float *__restrict__ X,
                                                                const float *__restrict__ Bsub,
                                                                int64_t ldb, int64_t col0, int kval,
                                                                float *__restrict__ Y,
                                                                const PcgState *st) {
  if (batch_done(st)) return;
  constexpr int K = 8, CH = 512, SPC = CH / 256, D = P * SPC;
  static_assert(P >= 2, "the x of the next chunk is fetched at least one chunk ahead");
  constexpr int XPT = CH * K / kRowsThreads;
  __shared__ __align__(16) float xs[2][K * CH];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = (int64_t(blockIdx.x) * (kRowsThreads / 64) + wave) * RW;
  const float *rowp[RW];
  bool rok[RW];
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    rok[i] = row0 + i < m;
    rowp[i] = M + (rok[i] ? row0 + i : m - 1) * n;
  }
  float acc[RW][K];
#pragma unroll
  for (int i = 0; i < RW; ++i)
#pragma unroll
    for (int r = 0; r < K; ++r) acc[i][r] = 0.f;
  auto loadm = [&](int64_t c, float4 (&mv)[RW]) {  // c: this lane's first column
    const int64_t cc = c < n ? c : n - 4;
#pragma unroll
    for (int i = 0; i < RW; ++i) mv[i] = load_stream4(rowp[i] + cc);
  };
  auto compute = [&](int64_t c, int cl, const float4 (&mv)[RW], const float *xb) {
    const bool cok = c < n;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const float4 xv = *reinterpret_cast<const float4 *>(xb + r * CH + cl);
#pragma unroll
      for (int i = 0; i < RW; ++i) {
        const bool use = rok[i] && cok;
        float sum = acc[i][r];
        sum += (use ? mv[i].x : 0.f) * xv.x;
        sum += (use ? mv[i].y : 0.f) * xv.y;
        sum += (use ? mv[i].z : 0.f) * xv.z;
        sum += (use ? mv[i].w : 0.f) * xv.w;
        acc[i][r] = sum;
      }
    }
  };
  auto fetch = [&](int64_t c0, float (&xq)[XPT]) {
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      const int idx = threadIdx.x + j * kRowsThreads;
      const int64_t c = c0 + idx / K;
      xq[j] = X[(c < n ? c : n - 1) * K + idx % K];
    }
  };
  auto store = [&](int64_t c0, const float (&xq)[XPT], float *dst) {
#pragma unroll
    for (int j = 0; j < XPT; ++j) {
      const int idx = threadIdx.x + j * kRowsThreads;
      dst[(idx % K) * CH + idx / K] = (c0 + idx / K < n) ? xq[j] : 0.f;
    }
  };
  float4 mv[D][RW];   // step t lives in slot t % D
  float xq[P][XPT];   // x of chunk h lives in set h % P
#pragma unroll
  for (int d = 0; d < D; ++d) loadm(d * 256 + 4 * lane, mv[d]);
#pragma unroll
  for (int h = 0; h < P - 1; ++h) fetch(h * CH, xq[h]);
  store(0, xq[0], xs[0]);
  __syncthreads();
  int buf = 0;
  // scheduling fences keep each phase in place; the empty asm pins each
  // step's FMAs before the next loads
  for (int64_t c0 = 0; c0 < n; c0 += D * 256) {
#pragma unroll
    for (int h = 0; h < P; ++h) {
      const int64_t cc = c0 + h * CH;
      fetch(cc + (P - 1) * CH, xq[(h + P - 1) % P]);   // past the row: clamped, never read
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < SPC; ++t) {
        compute(cc + t * 256 + 4 * lane, t * 256 + 4 * lane, mv[h * SPC + t], xs[buf]);
#pragma unroll
        for (int i = 0; i < RW; ++i)
#pragma unroll
          for (int r = 0; r < K; ++r) asm volatile("" : "+v"(acc[i][r]));
        __builtin_amdgcn_sched_barrier(0);
        loadm(cc + t * 256 + D * 256 + 4 * lane, mv[h * SPC + t]);
        __builtin_amdgcn_sched_barrier(0);
      }
      store(cc + CH, xq[(h + 1) % P], xs[buf ^ 1]);
      __syncthreads();
      buf ^= 1;
    }
  }
#pragma unroll
  for (int i = 0; i < RW; ++i) {
    const int64_t f = row0 + i;
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const float sum = wave_sum(acc[i][r]);
      if (lane == 0 && f < m) {
        if (Bsub != nullptr) {
          const float bv = (r < kval) ? Bsub[f * ldb + col0 + r] : 0.f;
          Y[f * K + r] = bv - sum;
        } else {
          Y[f * K + r] = sum;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ part = M^T y
// Block (cx, rb): columns [cx*1024, +1024) (4 per thread, float4), rows of row
// block rb.  FP32 products are summed over 32-row chunks in TA and the chunk
// sums are carried in FP64, so the FP32 path keeps a compensated running sum.
constexpr int kRowChunk = 32;
constexpr int kF64x8ColsDepth = 4;   // k_mv_cols_f64x8: steps of 4 rows in flight per thread

// VEC (n % 4 == 0): a thread owns 4 consecutive columns (one float4 per row);
// otherwise it owns columns c, c + 256, c + 512, c + 768 of its block's 1024,
// so each scalar load instruction reads 256 contiguous bytes.  RU rows are in
// flight per thread (16 for small grids, where each thread walks many rows).
template <typename TA, int K, bool VEC, int RU>
__global__ __launch_bounds__(kThreads) void k_mv_cols(const float *__restrict__ M, int64_t m,
                                                      int64_t n, const TA *__restrict__ Y,
                                                      int64_t rows_per_block,
                                                      double *__restrict__ part,
                                                      const PcgState *st) {
  if (batch_done(st)) return;
  // FP32 products are summed over 32-row chunks in FP32 and the chunk sums are
  // carried in FP64; FP64 products go straight into the FP64 sums.
  constexpr bool kTwoLevel = sizeof(TA) == 4;
  __shared__ TA ys[kRowChunk * K];
  const int64_t i0 = VEC ? (int64_t(blockIdx.x) * kThreads + threadIdx.x) * 4
                         : int64_t(blockIdx.x) * 4 * kThreads + threadIdx.x;
  constexpr int64_t cs = VEC ? 1 : kThreads;  // column stride between a thread's 4 columns
  const int64_t fb = int64_t(blockIdx.y) * rows_per_block;
  const int64_t fe = (fb + rows_per_block < m) ? fb + rows_per_block : m;
  double acc[4][K];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < K; ++r) acc[j][r] = 0.0;

  for (int64_t fc = fb; fc < fe; fc += kRowChunk) {
    const int rows = int((fe - fc < kRowChunk) ? fe - fc : kRowChunk);
    __syncthreads();
    for (int idx = threadIdx.x; idx < rows * K; idx += kThreads) ys[idx] = Y[fc * K + idx];
    __syncthreads();
    TA c[kTwoLevel ? 4 : 1][kTwoLevel ? K : 1];
    if (kTwoLevel) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < K; ++r) c[kTwoLevel ? j : 0][kTwoLevel ? r : 0] = TA(0);
    }
    auto add = [&](int j, int r, TA v) {
      if (kTwoLevel)
        c[kTwoLevel ? j : 0][kTwoLevel ? r : 0] += v;
      else
        acc[j][r] += double(v);
    };
    if (i0 < n) {
      int f = 0;
      for (; f + RU <= rows; f += RU) {
        float4 v[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
          const float *rowp = M + (fc + f + u) * n;
          if (VEC && i0 + 3 < n) {
            v[u] = load_stream4(rowp + i0);
          } else {
            v[u].x = rowp[i0];
            v[u].y = (i0 + cs < n) ? rowp[i0 + cs] : 0.f;
            v[u].z = (i0 + 2 * cs < n) ? rowp[i0 + 2 * cs] : 0.f;
            v[u].w = (i0 + 3 * cs < n) ? rowp[i0 + 3 * cs] : 0.f;
          }
        }
#pragma unroll
        for (int u = 0; u < RU; ++u)
#pragma unroll
          for (int r = 0; r < K; ++r) {
            const TA yv = ys[(f + u) * K + r];
            add(0, r, TA(v[u].x) * yv);
            add(1, r, TA(v[u].y) * yv);
            add(2, r, TA(v[u].z) * yv);
            add(3, r, TA(v[u].w) * yv);
          }
      }
      for (; f < rows; ++f) {
        const float *rowp = M + (fc + f) * n;
        float4 v;
        v.x = rowp[i0];
        v.y = (i0 + cs < n) ? rowp[i0 + cs] : 0.f;
        v.z = (i0 + 2 * cs < n) ? rowp[i0 + 2 * cs] : 0.f;
        v.w = (i0 + 3 * cs < n) ? rowp[i0 + 3 * cs] : 0.f;
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const TA yv = ys[f * K + r];
          add(0, r, TA(v.x) * yv);
          add(1, r, TA(v.y) * yv);
          add(2, r, TA(v.z) * yv);
          add(3, r, TA(v.w) * yv);
        }
      }
    }
    if (kTwoLevel) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < K; ++r) acc[j][r] += double(c[kTwoLevel ? j : 0][kTwoLevel ? r : 0]);
    }
  }
  if (i0 < n) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (i0 + j * cs < n) {
        double *dst = part + (int64_t(blockIdx.y) * n + i0 + j * cs) * K;
#pragma unroll
        for (int r = 0; r < K; ++r) dst[r] = acc[j][r];
      }
    }
  }
}

// FP32 batches of 8 (the PCG's M^T y), n % 4 == 0: k_mv_cols' sums (FP32
// over 32-row chunks from the block's first row, carried in FP64) with M kept
// in flight as in k_mv_cols_f64x8 -- D steps of RU rows per thread, y in
// 64-row chunks fetched one chunk ahead.  Launched one block per CU (the
// row split sized for a single resident block): at m = n = 65,536, 64 x 4
// blocks take 2.43 ms (7.06 TB/s); k_mv_cols at its best split 2.55 ms,
// this kernel with 2 resident blocks per CU 2.57-2.61 ms.
template <int RU, int D>
__global__ __launch_bounds__(kThreads) void k_mv_cols_f32x8(const float *__restrict__ M, int64_t m,
                                                            int64_t n, const float *__restrict__ Y,
                                                            int64_t rows_per_block,
                                                            double *__restrict__ part,
                                                            const PcgState *st) {
  if (batch_done(st)) return;
  constexpr int K = 8, CHR = 64, SPC = CHR / RU;
  static_assert(D >= 1 && D < SPC && SPC % D == 0, "ring slots repeat within a chunk");
  static_assert(kRowChunk % RU == 0 && CHR % kRowChunk == 0, "FP32 chunks within y chunks");
  constexpr int YPT = CHR * K / kThreads;  // y values fetched per thread
  __shared__ float ys[2][CHR * K];
  const int64_t i0 = (int64_t(blockIdx.x) * kThreads + threadIdx.x) * 4;
  const bool col_ok = i0 < n;
  const int64_t ic = col_ok ? i0 : n - 4;
  const int64_t fb = int64_t(blockIdx.y) * rows_per_block;
  const int64_t fe = (fb + rows_per_block < m) ? fb + rows_per_block : m;
  double acc[4][K];
  float c[4][K];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < K; ++r) {
      acc[j][r] = 0.0;
      c[j][r] = 0.f;
    }
  if (fb >= fe) {  // an empty row block still writes its (zero) partials
    if (col_ok)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < K; ++r) part[(int64_t(blockIdx.y) * n + i0 + j) * K + r] = 0.0;
    return;
  }
  auto loadm = [&](int64_t f, float4 (&v)[RU]) {  // rows f .. f + RU - 1
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int64_t fr = f + u < fe ? f + u : fe - 1;
      v[u] = load_stream4(M + fr * n + ic);
    }
  };
  float yq[YPT];
  auto fetch = [&](int64_t fc) {
#pragma unroll
    for (int j = 0; j < YPT; ++j) {
      const int idx = threadIdx.x + j * kThreads;
      const int64_t f = fc + idx / K;
      yq[j] = Y[(f < fe ? f : fe - 1) * K + idx % K];
    }
  };
  auto store = [&](int64_t fc, float *dst) {
#pragma unroll
    for (int j = 0; j < YPT; ++j) {
      const int idx = threadIdx.x + j * kThreads;
      dst[idx] = (fc + idx / K < fe) ? yq[j] : 0.f;
    }
  };
  float4 ring[D][RU];  // step t (rows fb + RU t ..) lives in slot t % D
#pragma unroll
  for (int d = 0; d < D; ++d) loadm(fb + d * RU, ring[d]);
  fetch(fb);
  store(fb, ys[0]);
  __syncthreads();
  int buf = 0;
  for (int64_t fc = fb; fc < fe; fc += CHR) {
    fetch(fc + CHR);  // past the range: clamped, stored as zeros
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < SPC; ++t) {
      const float *yb = ys[buf] + t * RU * K;
#pragma unroll
      for (int u = 0; u < RU; ++u)
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const float yv = yb[u * K + r];
          c[0][r] += ring[t % D][u].x * yv;
          c[1][r] += ring[t % D][u].y * yv;
          c[2][r] += ring[t % D][u].z * yv;
          c[3][r] += ring[t % D][u].w * yv;
        }
      if ((t + 1) % (kRowChunk / RU) == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int r = 0; r < K; ++r) {
            acc[j][r] += double(c[j][r]);
            c[j][r] = 0.f;
          }
      }
      // pin the step's FMAs here: without it the instruction selector
      // gathers every step's loads first and the FMAs after them
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < K; ++r) asm volatile("" : "+v"(c[j][r]));
      __builtin_amdgcn_sched_barrier(0);
      loadm(fc + (t + D) * RU, ring[t % D]);
      __builtin_amdgcn_sched_barrier(0);
    }
    store(fc + CHR, ys[buf ^ 1]);
    __syncthreads();
    buf ^= 1;
  }
  if (col_ok) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double *dst = part + (int64_t(blockIdx.y) * n + i0 + j) * K;
#pragma unroll
      for (int r = 0; r < K; ++r) dst[r] = acc[j][r];
    }
  }
}

// FP64 batches of 8 (the certified residual's M^T r), n % 4 == 0: the same
// sums as k_mv_cols<double, 8, true, 4> in the same order, with M kept in
// flight.  k_mv_cols loads 4 rows, computes them and drains at every 32-row
// chunk of y; with FP64 arithmetic that left it at 5.3 TB/s.  Here each
// thread keeps D steps of 4 rows in flight (a ring of 4 D float4), the loads
// of step t + D are issued once step t is computed, and y comes in 64-row
// chunks (4 KB of LDS, two buffers) fetched one chunk ahead, before the M
// loads of the chunk, so that waiting for it never waits for M loads still
// ahead.  Rows past the block's range read its last row and take y = 0;
// columns past n read the row's last float4 and are not stored.  At m = n =
// 65,536: 2.64 ms (6.5 TB/s) against 3.20 ms.
template <int D>
__global__ __launch_bounds__(kThreads) void k_mv_cols_f64x8(const float *__restrict__ M, int64_t m,
                                                            int64_t n, const double *__restrict__ Y,
                                                            int64_t rows_per_block,
                                                            double *__restrict__ part,
                                                            const PcgState *st) {
  if (batch_done(st)) return;
  constexpr int K = 8, RU = 4, CHR = 64, SPC = CHR / RU;  // 16 steps of 4 rows per y chunk
  static_assert(D >= 1 && D < SPC && SPC % D == 0, "ring slots repeat within a chunk");
  constexpr int YPT = CHR * K / kThreads;                 // y values fetched per thread
  __shared__ double ys[2][CHR * K];
  const int64_t i0 = (int64_t(blockIdx.x) * kThreads + threadIdx.x) * 4;
  const bool col_ok = i0 < n;
  const int64_t ic = col_ok ? i0 : n - 4;
  const int64_t fb = int64_t(blockIdx.y) * rows_per_block;
  const int64_t fe = (fb + rows_per_block < m) ? fb + rows_per_block : m;
  double acc[4][K];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < K; ++r) acc[j][r] = 0.0;
  if (fb >= fe) {  // an empty row block still writes its (zero) partials
    if (col_ok)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < K; ++r) part[(int64_t(blockIdx.y) * n + i0 + j) * K + r] = 0.0;
    return;
  }
  auto loadm = [&](int64_t f, float4 (&v)[RU]) {  // rows f .. f + 3
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int64_t fr = f + u < fe ? f + u : fe - 1;
      v[u] = load_stream4(M + fr * n + ic);
    }
  };
  double yq[YPT];
  auto fetch = [&](int64_t fc) {
#pragma unroll
    for (int j = 0; j < YPT; ++j) {
      const int idx = threadIdx.x + j * kThreads;
      const int64_t f = fc + idx / K;
      yq[j] = Y[(f < fe ? f : fe - 1) * K + idx % K];
    }
  };
  auto store = [&](int64_t fc, double *dst) {
#pragma unroll
    for (int j = 0; j < YPT; ++j) {
      const int idx = threadIdx.x + j * kThreads;
      dst[idx] = (fc + idx / K < fe) ? yq[j] : 0.0;
    }
  };
  float4 ring[D][RU];  // step t (rows fb + 4 t ..) lives in slot t % D
#pragma unroll
  for (int d = 0; d < D; ++d) loadm(fb + d * RU, ring[d]);
  fetch(fb);
  store(fb, ys[0]);
  __syncthreads();
  int buf = 0;
  for (int64_t fc = fb; fc < fe; fc += CHR) {
    fetch(fc + CHR);  // past the range: clamped, stored as zeros
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < SPC; ++t) {
      const double *yb = ys[buf] + t * RU * K;
#pragma unroll
      for (int u = 0; u < RU; ++u)
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const double yv = yb[u * K + r];
          acc[0][r] += double(ring[t % D][u].x) * yv;
          acc[1][r] += double(ring[t % D][u].y) * yv;
          acc[2][r] += double(ring[t % D][u].z) * yv;
          acc[3][r] += double(ring[t % D][u].w) * yv;
        }
      // pin the step's FMAs here: without it the instruction selector
      // gathers every step's loads first and the FMAs after them
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < K; ++r) asm volatile("" : "+v"(acc[j][r]));
      __builtin_amdgcn_sched_barrier(0);
      loadm(fc + (t + D) * RU, ring[t % D]);
      __builtin_amdgcn_sched_barrier(0);
    }
    store(fc + CHR, ys[buf ^ 1]);
    __syncthreads();
    buf ^= 1;
  }
  if (col_ok) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double *dst = part + (int64_t(blockIdx.y) * n + i0 + j) * K;
#pragma unroll
      for (int r = 0; r < K; ++r) dst[r] = acc[j][r];
    }
  }
}

// ------------------------------------------------------------------ reductions of part
enum ReduceMode { RED_Q = 0, RED_R0 = 1, RED_ORIG = 2 };

// RED_Q:    q_new[P[i]] = s, dot(pv, s) -> st->dot_pq
// RED_R0:   r_new[P[i]] = s, ||s||^2 -> rhs2/thresh/active (PCG start)
// RED_ORIG: o[i] = s (double, original order), ||s||^2 -> out2
template <typename TA, int K, int MODE>
__global__ __launch_bounds__(kThreads) void k_mv_reduce(
    const double *__restrict__ part, int32_t row_blocks, int64_t n, const int32_t *__restrict__ P,
    const TA *__restrict__ pv, TA *__restrict__ out_new, double *__restrict__ out_orig,
    double *__restrict__ dpart, PcgState *st, double *out2, double tol, int kval, int32_t maxit) {
  if (MODE == RED_Q && batch_done(st)) return;
  __shared__ double red[5 * K];
  __shared__ int last;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double d[K];
#pragma unroll
  for (int r = 0; r < K; ++r) d[r] = 0.0;
  if (i < n) {
    double s[K];
#pragma unroll
    for (int r = 0; r < K; ++r) s[r] = 0.0;
#pragma unroll 8
    for (int32_t b = 0; b < row_blocks; ++b) {
      const double *src = part + (int64_t(b) * n + i) * K;
#pragma unroll
      for (int r = 0; r < K; ++r) s[r] += src[r];
    }
    if (MODE == RED_ORIG) {
#pragma unroll
      for (int r = 0; r < K; ++r) {
        out_orig[i * K + r] = s[r];
        d[r] = s[r] * s[r];
      }
    } else {
      const int64_t dst = int64_t(P[i]) * K;
#pragma unroll
      for (int r = 0; r < K; ++r) {
        const TA v = TA(s[r]);
        out_new[dst + r] = v;
        d[r] = (MODE == RED_Q) ? double(pv[i * K + r]) * double(v) : double(v) * double(v);
      }
    }
  }
  block_sum<K>(d, red);
  const uint32_t ctr = (MODE == RED_Q) ? C_PQ : (MODE == RED_R0 ? C_R0 : C_NORM);
  if (publish_and_arrive<K>(red + 4 * K, dpart, &st->counter[ctr], gridDim.x, &last)) {
    gather_partials<K>(dpart, gridDim.x, red);
    if (threadIdx.x < K) {
      const int r = threadIdx.x;
      const double tot = red[4 * K + r];
      if (MODE == RED_Q) {
        st->dot_pq[r] = tot;
      } else if (MODE == RED_R0) {
        st->rhs2[r] = tot;
        st->thresh[r] = tol * tol * tot;
        st->active[r] = (r < kval && tot >= 1e-16 && maxit > 0) ? 1 : 0;  // kl_chol.cu:950
        st->iters[r] = 0;
        st->dot_rr[r] = tot;
      } else {
        out2[r] = tot;
      }
    }
    if (MODE == RED_R0) {
      __syncthreads();
      if (threadIdx.x == 0) {
        int na = 0;
        for (int r = 0; r < K; ++r) na += st->active[r];
        st->n_active = na;
        st->iter = 0;
        st->maxit = maxit;
      }
    }
  }
}

// r_new[c] = b[Pinv[c]] for source-space right-hand sides (VpL), with ||r||^2.
template <typename TA, int K>
__global__ __launch_bounds__(kThreads) void k_load_source_rhs(
    int64_t n, const int32_t *__restrict__ Pinv, const float *__restrict__ B, int64_t ldb,
    int64_t col0, const double *__restrict__ B64, int kval, TA *__restrict__ r_new,
    double *__restrict__ dpart, PcgState *st, double tol, int32_t maxit,
    const double *__restrict__ thresh_abs) {
  __shared__ double red[5 * K];
  __shared__ int last;
  const int64_t c = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double d[K];
#pragma unroll
  for (int r = 0; r < K; ++r) d[r] = 0.0;
  if (c < n) {
    const int64_t o = Pinv[c];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      TA v = TA(0);
      if (r < kval) v = B64 ? TA(B64[o * K + r]) : TA(B[o * ldb + col0 + r]);
      r_new[c * K + r] = v;
      d[r] = double(v) * double(v);
    }
  }
  block_sum<K>(d, red);
  if (publish_and_arrive<K>(red + 4 * K, dpart, &st->counter[C_R0], gridDim.x, &last)) {
    gather_partials<K>(dpart, gridDim.x, red);
    if (threadIdx.x < K) {
      const int r = threadIdx.x;
      const double tot = red[4 * K + r];
      const double th = thresh_abs ? thresh_abs[r] : tol * tol * tot;
      st->rhs2[r] = tot;
      st->thresh[r] = th;
      st->active[r] = (r < kval && tot >= 1e-16 && maxit > 0 && !(tot < th)) ? 1 : 0;
      st->iters[r] = 0;
      st->dot_rr[r] = tot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int na = 0;
      for (int r = 0; r < K; ++r) na += st->active[r];
      st->n_active = na;
      st->iter = 0;
      st->maxit = maxit;
    }
  }
}

// y = B (FP32 boundary data, strided) as TA, m x K.
template <typename TA, int K>
__global__ __launch_bounds__(kThreads) void k_load_boundary(int64_t m, const float *__restrict__ B,
                                                            int64_t ldb, int64_t col0, int kval,
                                                            TA *__restrict__ y) {
  const int64_t f = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (f >= m) return;
#pragma unroll
  for (int r = 0; r < K; ++r) y[f * K + r] = (r < kval) ? TA(B[f * ldb + col0 + r]) : TA(0);
}

// R = b - R (sub) or R = b, b the source-space FP32 data in original order;
// per-column ||R||^2 into out2 (last-block reduction on counter C_NORM).
template <int K>
__global__ __launch_bounds__(kThreads) void k_source_residual(
    int64_t n, const float *__restrict__ B, int64_t ldb, int64_t col0, int kval, int sub,
    double *__restrict__ R, double *__restrict__ dpart, PcgState *st, double *out2) {
  __shared__ double red[5 * K];
  __shared__ int last;
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double d[K];
#pragma unroll
  for (int r = 0; r < K; ++r) d[r] = 0.0;
  if (i < n) {
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const double b = (r < kval) ? double(B[i * ldb + col0 + r]) : 0.0;
      const double v = sub ? b - R[i * K + r] : b;
      R[i * K + r] = v;
      d[r] = v * v;
    }
  }
  block_sum<K>(d, red);
  if (publish_and_arrive<K>(red + 4 * K, dpart, &st->counter[C_NORM], gridDim.x, &last)) {
    gather_partials<K>(dpart, gridDim.x, red);
    if (threadIdx.x < K) out2[threadIdx.x] = red[4 * K + threadIdx.x];
  }
}

template <typename TA>
__global__ __launch_bounds__(kThreads) void k_zero(int64_t count, TA *__restrict__ a) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) a[i] = TA(0);
}

// ------------------------------------------------------------------ x, r update
template <typename TA, int K>
__global__ __launch_bounds__(kThreads) void k_upd_xr(int64_t n, TA *__restrict__ x,
                                                     TA *__restrict__ r, const TA *__restrict__ p,
                                                     const TA *__restrict__ q,
                                                     double *__restrict__ dpart, PcgState *st) {
  if (batch_done(st)) return;
  __shared__ double red[5 * K];
  __shared__ int last;
  const int64_t c = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  double d[K];
#pragma unroll
  for (int k = 0; k < K; ++k) d[k] = 0.0;
  if (c < n) {
    // every slot's loads first: loads inside the per-slot branches completed
    // one slot at a time
    TA xv[K], pk[K], rk[K], qk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      xv[k] = x[c * K + k];
      pk[k] = p[c * K + k];
      rk[k] = r[c * K + k];
      qk[k] = q[c * K + k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (st->active[k]) {
        const TA alpha = TA(st->absnew[k] / st->dot_pq[k]);
        x[c * K + k] = xv[k] + alpha * pk[k];
        const TA rv = rk[k] + (-alpha) * qk[k];
        r[c * K + k] = rv;
        d[k] = double(rv) * double(rv);
      }
    }
  }
  block_sum<K>(d, red);
  if (publish_and_arrive<K>(red + 4 * K, dpart, &st->counter[C_RR], gridDim.x, &last)) {
    gather_partials<K>(dpart, gridDim.x, red);
    if (threadIdx.x < K) {
      const int k = threadIdx.x;
      const double tot = red[4 * K + k];
      if (st->active[k]) {
        st->iters[k] += 1;
        st->dot_rr[k] = tot;
        if (tot < st->thresh[k]) st->active[k] = 0;   // kl_chol.cu:1038
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int it = st->iter + 1;
      st->iter = it;
      int na = 0;
      for (int k = 0; k < K; ++k) {
        if (it >= st->maxit) st->active[k] = 0;       // while (i < maxits)
        na += st->active[k];
      }
      st->n_active = na;
    }
  }
}

// ------------------------------------------------------------------ sparse L / L^T
// Lanes per row of the preconditioner products.  Each lane walks the entries
// of its row one after another (a chain of dependent index -> value loads), so
// batches of 8 and 16 give every row several lanes per right-hand side, which
// take every 4th (8th) entry and are summed by shuffles.  At rho = 3 a column
// of L (a row of L^T r) has 36 entries on average and at most 86 -- 32 lanes
// per row; a row of L (L t, the product with the dot) 36 on average but up to
// 251, so it takes 64 lanes.  The product with the dot keeps the block count
// of 8 lanes per row (several row groups per block, one after another) to
// bound its per-block partials.  At N = 65,536, K = 8 (average over the
// launches): L^T r 0.035 -> 0.025 ms, L t with the dot 0.114 -> 0.075 ms
// (with the partials summed by k_dot_rz); then, with four entries' loads in
// flight per lane and every pass's row bounds and w loaded up front, 0.018
// and 0.049 ms (the same sums in the same order).
__host__ __device__ constexpr int spmv_lanes(int K, bool dot) { return K >= 8 ? (dot ? 64 : 32) : 8; }
__host__ __device__ constexpr int spmv_passes(int K, bool dot) { return dot ? spmv_lanes(K, dot) / 8 : 1; }

// out[row] = sum_e val[e] * in[idx[e]] for K right-hand sides; spmv_lanes(K,
// DOT) lanes per row.  With DOT, also dot(out, w) per right-hand side ->
// st->dot_rz.
template <typename TA, int K, bool DOT>
__global__ __launch_bounds__(kThreads) void k_spmv(int64_t n, const int32_t *__restrict__ ptr,
                                                   const int32_t *__restrict__ idx,
                                                   const TA *__restrict__ val,
                                                   const TA *__restrict__ in, TA *__restrict__ out,
                                                   const TA *__restrict__ w,
                                                   double *__restrict__ dpart, PcgState *st) {
  if (batch_done(st)) return;
  constexpr int G = spmv_lanes(K, DOT);        // lanes per row
  constexpr int PASSES = spmv_passes(K, DOT);  // row groups per block, one after another
  constexpr int SPLIT = (K >= G) ? 1 : G / K;  // lanes sharing one right-hand side
  constexpr int RPL = (K > G) ? K / G : 1;     // right-hand sides per lane
  __shared__ double red[5 * K];
  const int g = threadIdx.x & (G - 1);
  const int rbase = (K >= G) ? g : (g % K);
  const int phase = (K >= G) ? 0 : (g / K);
  double d[K];
#pragma unroll
  for (int k = 0; k < K; ++k) d[k] = 0.0;
  // every pass's row bounds (and w) loaded up front: the passes' dependent
  // load chains then start from registers
  int32_t rb[PASSES], re[PASSES];
  TA wv[PASSES][RPL];
#pragma unroll
  for (int pass = 0; pass < PASSES; ++pass) {
    const int64_t row = (int64_t(blockIdx.x) * PASSES + pass) * (kThreads / G) + threadIdx.x / G;
    rb[pass] = 0;
    re[pass] = 0;
#pragma unroll
    for (int u = 0; u < RPL; ++u) wv[pass][u] = TA(0);
    if (row < n) {
      rb[pass] = ptr[row];
      re[pass] = ptr[row + 1];
      if (DOT && phase == 0) {
#pragma unroll
        for (int u = 0; u < RPL; ++u) wv[pass][u] = w[row * K + rbase + u * G];
      }
    }
  }
#pragma unroll
  for (int pass = 0; pass < PASSES; ++pass) {
    const int64_t row = (int64_t(blockIdx.x) * PASSES + pass) * (kThreads / G) + threadIdx.x / G;
    TA acc[RPL];
#pragma unroll
    for (int u = 0; u < RPL; ++u) acc[u] = TA(0);
    if (row < n) {
      const int32_t b = rb[pass], e = re[pass];
      int32_t t = b + phase;
      // four entries' loads in flight, added in the same order as one by one
      for (; t + 3 * SPLIT < e; t += 4 * SPLIT) {
        TA v[4], x[4][RPL];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v[j] = val[t + j * SPLIT];
          const int64_t col = idx[t + j * SPLIT];
#pragma unroll
          for (int u = 0; u < RPL; ++u) x[j][u] = in[col * K + rbase + u * G];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int u = 0; u < RPL; ++u) acc[u] += v[j] * x[j][u];
      }
      for (; t < e; t += SPLIT) {
        const TA v = val[t];
        const int64_t col = idx[t];
#pragma unroll
        for (int u = 0; u < RPL; ++u) acc[u] += v * in[col * K + rbase + u * G];
      }
    }
#pragma unroll
    for (int off = K; off < G; off <<= 1)
#pragma unroll
      for (int u = 0; u < RPL; ++u) acc[u] += __shfl_xor(acc[u], off, 64);
    if (row < n && phase == 0) {
#pragma unroll
      for (int u = 0; u < RPL; ++u) {
        const int k = rbase + u * G;
        out[row * K + k] = acc[u];
        if (DOT) {
          // w outside the per-slot selects (a load inside them became one
          // dependent load per slot)
          const double wk = double(wv[pass][u]);
#pragma unroll
          for (int kk = 0; kk < K; ++kk)
            if (kk == k) d[kk] = fma(double(acc[u]), wk, d[kk]);
        }
      }
    }
  }
  if (DOT) {
    // plain stores: k_dot_rz (the next launch) sums the partials, so the
    // blocks skip the device-scope fence and arrival atomic of
    // publish_and_arrive (2048 of them per launch at N = 65,536)
    block_sum<K>(d, red);
    if (threadIdx.x < K) dpart[size_t(blockIdx.x) * K + threadIdx.x] = red[4 * K + threadIdx.x];
  }
}

// st->dot_rz = the sum of the nblocks partials k_spmv<.., true> left in dpart.
// Also beta = dot_rz / absnew and then absnew = dot_rz for the iterating
// columns (INIT: absnew = dot_rz for all), so k_upd_p, the next launch, reads
// beta and needs no last-block step of its own (kl_chol.cu:1056-1077).
template <int K, bool INIT>
__global__ __launch_bounds__(kThreads) void k_dot_rz(const double *__restrict__ dpart,
                                                     uint32_t nblocks, PcgState *st) {
  if (batch_done(st)) return;
  __shared__ double red[5 * K];
  gather_partials<K>(dpart, nblocks, red);
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    const double rz = red[4 * K + k];
    st->dot_rz[k] = rz;
    if (INIT) {
      st->absnew[k] = rz;
    } else if (st->active[k]) {
      st->beta[k] = rz / st->absnew[k];
      st->absnew[k] = rz;
    }
  }
}

// ------------------------------------------------------------------ p update
// INIT: p = z (kl_chol.cu:957-982).  Else p = z + beta p with beta from
// k_dot_rz (kl_chol.cu:1056-1077).  Also scatters p into original order for
// the matvec.
template <typename TA, int K, bool INIT>
__global__ __launch_bounds__(kThreads) void k_upd_p(int64_t n, const int32_t *__restrict__ Pinv,
                                                    TA *__restrict__ p, const TA *__restrict__ z,
                                                    TA *__restrict__ pv, const PcgState *st) {
  if (batch_done(st)) return;
  const int64_t c = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (c < n) {
    const int64_t o = Pinv[c];
    // every slot's loads first (see k_upd_xr)
    TA zk[K], pk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      zk[k] = z[c * K + k];
      pk[k] = INIT ? TA(0) : p[c * K + k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      TA v;
      if (INIT) {
        v = zk[k];
      } else if (st->active[k]) {
        const TA beta = TA(st->beta[k]);
        v = beta * pk[k] + zk[k];
      } else {
        v = pk[k];
      }
      p[c * K + k] = v;
      pv[o * K + k] = v;
    }
  }
}

// X_out[i] = x_new[P[i]] (FP32, strided) and optionally X64[i] += x_new[P[i]].
template <typename TA, int K>
__global__ __launch_bounds__(kThreads) void k_finish(int64_t n, const int32_t *__restrict__ P,
                                                     const TA *__restrict__ x, float *__restrict__ X,
                                                     int64_t ldx, int64_t col0, int kval,
                                                     double *__restrict__ X64) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const int64_t s = int64_t(P[i]) * K;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k >= kval) continue;
    if (X64 != nullptr) {
      const double v = X64[i * K + k] + double(x[s + k]);
      X64[i * K + k] = v;
      if (X != nullptr) X[i * ldx + col0 + k] = float(v);
    } else if (X != nullptr) {
      X[i * ldx + col0 + k] = float(x[s + k]);
    }
  }
}

template <int K>
__global__ __launch_bounds__(kThreads) void k_finish64(int64_t n, const int32_t *__restrict__ P,
                                                       const double *__restrict__ x,
                                                       double *__restrict__ X, int64_t ldx,
                                                       int64_t col0, int kval) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const int64_t s = int64_t(P[i]) * K;
  for (int k = 0; k < kval && k < K; ++k) X[i * ldx + col0 + k] = x[s + k];
}

// ------------------------------------------------------------------ launch helpers
inline unsigned blocks_for(int64_t count) { return unsigned((count + kThreads - 1) / kThreads); }

template <typename TA, int K>
void mv_rows(hipStream_t s, const SolveArgs<TA> &a, const TA *X, const float *Bsub, int64_t ldb,
             int64_t col0, int kval, TA *Y, const PcgState *st) {
  if (a.pts.tgt32 != nullptr) {  // points form (errors surface through hipGetLastError)
    (void)points_rows<TA>(s, a.pts, K, X, Bsub, ldb, col0, kval, Y, a.part, st);
    return;
  }
  // Fewer rows per wave when the grid would not give every CU a few blocks.
  constexpr int RW = RowsCfg<TA, K>::RW;
  constexpr int RWS = RW > 2 ? 2 : RW;
  constexpr int waves = kRowsThreads / 64;
  const bool small = (a.m + waves * RW - 1) / (waves * RW) < 1024;
  const int64_t rpb = int64_t(waves) * (small ? RWS : RW);
  const unsigned grid = unsigned((a.m + rpb - 1) / rpb);
  const bool vec = a.n % 4 == 0;
  // batches of 8 with float4 rows: the pipelined passes, once their grid
  // covers the CUs (fewer rows keep k_mv_rows' shorter blocks)
  if constexpr (sizeof(TA) == 4 && K == 8) {
    const int64_t rows = int64_t(waves) * kF32x8Rows;
    if (vec && (a.m + rows - 1) / rows >= 256) {
      k_mv_rows_f32x8<kF32x8Rows, kF32x8Depth><<<unsigned((a.m + rows - 1) / rows), kRowsThreads, 0, s>>>(
          a.M, a.m, a.n, X, Bsub, ldb, col0, kval, Y, st);
      return;
    }
  }
  if constexpr (sizeof(TA) == 8 && K == 8) {
    if (!small && vec) {
      const int64_t rows = int64_t(waves) * kF64x8Rows;
      k_mv_rows_f64x8<kF64x8Rows, kF64x8Depth><<<unsigned((a.m + rows - 1) / rows), kRowsThreads, 0, s>>>(
          a.M, a.m, a.n, X, Bsub, ldb, col0, kval, Y, st);
      return;
    }
  }
  if (small && vec)
    k_mv_rows<TA, K, RWS, true><<<grid, kRowsThreads, 0, s>>>(a.M, a.m, a.n, X, Bsub, ldb, col0, kval, Y, st);
  else if (small)
    k_mv_rows<TA, K, RWS, false><<<grid, kRowsThreads, 0, s>>>(a.M, a.m, a.n, X, Bsub, ldb, col0, kval, Y, st);
  else if (vec)
    k_mv_rows<TA, K, RW, true><<<grid, kRowsThreads, 0, s>>>(a.M, a.m, a.n, X, Bsub, ldb, col0, kval, Y, st);
  else
    k_mv_rows<TA, K, RW, false><<<grid, kRowsThreads, 0, s>>>(a.M, a.m, a.n, X, Bsub, ldb, col0, kval, Y, st);
}

template <typename TY, int K>
void mv_cols_raw(hipStream_t s, const float *M, int64_t m, int64_t n, int32_t row_blocks,
                 const TY *Y, double *part, const PcgState *st) {
  const int64_t rpb = (m + row_blocks - 1) / row_blocks;
  dim3 grid(unsigned((n + 4 * kThreads - 1) / (4 * kThreads)), unsigned(row_blocks));
  // Small grids: each thread walks many rows, so keep more of them in flight.
  constexpr int RUS = K <= 2 ? 16 : 4;
  const bool small = int64_t(grid.x) * grid.y < 1024;
  if constexpr (std::is_same<TY, double>::value && K == 8) {
    if (n % 4 == 0) {
      k_mv_cols_f64x8<kF64x8ColsDepth><<<grid, kThreads, 0, s>>>(M, m, n, Y, rpb, part, st);
      return;
    }
  }
  if constexpr (std::is_same<TY, float>::value && K == 8) {
    if (n % 4 == 0) {
      k_mv_cols_f32x8<4, 2><<<grid, kThreads, 0, s>>>(M, m, n, Y, rpb, part, st);
      return;
    }
  }
  if (n % 4 == 0) {
    if (small)
      k_mv_cols<TY, K, true, RUS><<<grid, kThreads, 0, s>>>(M, m, n, Y, rpb, part, st);
    else
      k_mv_cols<TY, K, true, 4><<<grid, kThreads, 0, s>>>(M, m, n, Y, rpb, part, st);
  } else {
    if (small)
      k_mv_cols<TY, K, false, RUS><<<grid, kThreads, 0, s>>>(M, m, n, Y, rpb, part, st);
    else
      k_mv_cols<TY, K, false, 4><<<grid, kThreads, 0, s>>>(M, m, n, Y, rpb, part, st);
  }
}

}  // namespace

int64_t spmv_grid(int64_t n, int K, bool dot) {
  const int64_t rows = int64_t(kThreads / spmv_lanes(K, dot)) * spmv_passes(K, dot);  // per block
  return (n + rows - 1) / rows;
}

// Blocks of the dense M^T pass resident per CU (the instantiation launched
// for large grids: 4 rows in flight per thread).
int cols_resident_blocks(bool fp64, int K, bool vec) {
  static int cache[2][5][2] = {};
  const int ki = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : K == 8 ? 3 : 4;
  int &c = cache[fp64][ki][vec];
  if (c > 0) return c;
  const void *f = nullptr;
#define MFS_COLS_FN(T, KK)                                                                         \
  f = vec ? reinterpret_cast<const void *>(k_mv_cols<T, KK, true, 4>)                             \
          : reinterpret_cast<const void *>(k_mv_cols<T, KK, false, 4>)
  if (fp64) {
    switch (K) {
      case 1: MFS_COLS_FN(double, 1); break;
      case 2: MFS_COLS_FN(double, 2); break;
      case 4: MFS_COLS_FN(double, 4); break;
      case 8:
        if (vec) {  // k_mv_cols_f64x8 is sized for one resident block per CU
          c = 1;
          return c;
        }
        MFS_COLS_FN(double, 8);
        break;
      default: MFS_COLS_FN(double, 16); break;
    }
  } else {
    switch (K) {
      case 1: MFS_COLS_FN(float, 1); break;
      case 2: MFS_COLS_FN(float, 2); break;
      case 4: MFS_COLS_FN(float, 4); break;
      case 8:
        if (vec) {  // k_mv_cols_f32x8 is sized for one resident block per CU
          c = 1;
          return c;
        }
        MFS_COLS_FN(float, 8);
        break;
      default: MFS_COLS_FN(float, 16); break;
    }
  }
#undef MFS_COLS_FN
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kThreads, 0) != hipSuccess || nb < 1)
    nb = 2;
  c = nb;
  return c;
}

namespace {

// part = split sums of M^T Y (a.row_blocks splits in both forms).
template <typename TA, typename TY, int K>
void mv_cols(hipStream_t s, const SolveArgs<TA> &a, const TY *Y, const PcgState *st) {
  if (a.pts.tgt32 != nullptr)
    (void)points_cols<TY>(s, a.pts, K, Y, a.part, st);
  else
    mv_cols_raw<TY, K>(s, a.M, a.m, a.n, a.row_blocks, Y, a.part, st);
}

typedef void (*TrackFn)(void *, int, int);
inline void track(TrackFn fn, void *ctx, int cls, int phase) {
  if (fn) fn(ctx, cls, phase);
}

template <typename TA, int K>
hipError_t init_k(hipStream_t s, const SolveArgs<TA> &a, const float *B, int64_t ldb, int64_t col0,
                  int rhs_space, int kval, double tol, int32_t maxit, const double *B64,
                  const double *thresh_abs) {
  const unsigned gn = blocks_for(a.n);
  k_zero<TA><<<blocks_for(a.n * K), kThreads, 0, s>>>(a.n * K, a.x);
  if (rhs_space == 0) {
    k_load_boundary<TA, K><<<blocks_for(a.m), kThreads, 0, s>>>(a.m, B, ldb, col0, kval, a.y);
    mv_cols<TA, TA, K>(s, a, a.y, nullptr);
    k_mv_reduce<TA, K, RED_R0><<<gn, kThreads, 0, s>>>(a.part, a.row_blocks, a.n, a.P, nullptr,
                                                        a.r, nullptr, a.dpart, a.st, nullptr, tol,
                                                        kval, maxit);
  } else {
    k_load_source_rhs<TA, K><<<gn, kThreads, 0, s>>>(a.n, a.Pinv, B, ldb, col0, B64, kval, a.r,
                                                      a.dpart, a.st, tol, maxit, thresh_abs);
  }
  // z = L (L^T r); p = z; absnew = r^T z.
  k_spmv<TA, K, false><<<unsigned(spmv_grid(a.n, K, false)), kThreads, 0, s>>>(
      a.n, a.sptr, a.sind, a.lval, a.r, a.t, nullptr, nullptr, a.st);
  const unsigned gz = unsigned(spmv_grid(a.n, K, true));
  k_spmv<TA, K, true><<<gz, kThreads, 0, s>>>(a.n, a.tptr, a.tcol, a.ltval, a.t, a.z, a.r, a.dpart,
                                              a.st);
  k_dot_rz<K, true><<<1, kThreads, 0, s>>>(a.dpart, gz, a.st);
  k_upd_p<TA, K, true><<<gn, kThreads, 0, s>>>(a.n, a.Pinv, a.p, a.z, a.pv, a.st);
  return hipGetLastError();
}

// q = M^T (M p): one pass over M (dense FP32 batch of 8 with the exchange
// records set up) or two (y = M p, then split partials of M^T y).
template <typename TA, int K>
hipError_t normal_matvec(hipStream_t s, const SolveArgs<TA> &a, TrackFn fn, void *tctx,
                         int32_t *splits) {
  if constexpr (sizeof(TA) == 4 && K == 8) {
    if (a.fused.slices > 0 && a.pts.tgt32 == nullptr) {
      track(fn, tctx, 6, 0);
      const hipError_t e = mv_fused8(s, a.M, a.m, a.n, a.pv, a.part, a.frec, a.st, a.fused);
      track(fn, tctx, 6, 1);
      *splits = a.fused.groups;
      return e;
    }
  }
  track(fn, tctx, 1, 0);
  mv_rows<TA, K>(s, a, a.pv, nullptr, 0, 0, K, a.y, a.st);
  track(fn, tctx, 1, 1);
  track(fn, tctx, 2, 0);
  mv_cols<TA, TA, K>(s, a, a.y, a.st);
  track(fn, tctx, 2, 1);
  *splits = a.row_blocks;
  return hipSuccess;
}

template <typename TA, int K>
hipError_t iter_k(hipStream_t s, const SolveArgs<TA> &a, TrackFn fn, void *tctx) {
  const unsigned gn = blocks_for(a.n);
  int32_t splits = 0;
  const hipError_t e = normal_matvec<TA, K>(s, a, fn, tctx, &splits);
  if (e != hipSuccess) return e;
  k_mv_reduce<TA, K, RED_Q><<<gn, kThreads, 0, s>>>(a.part, splits, a.n, a.P, a.pv, a.q,
                                                     nullptr, a.dpart, a.st, nullptr, 0.0, K, 0);
  k_upd_xr<TA, K><<<gn, kThreads, 0, s>>>(a.n, a.x, a.r, a.p, a.q, a.dpart, a.st);
  k_spmv<TA, K, false><<<unsigned(spmv_grid(a.n, K, false)), kThreads, 0, s>>>(
      a.n, a.sptr, a.sind, a.lval, a.r, a.t, nullptr, nullptr, a.st);
  const unsigned gz = unsigned(spmv_grid(a.n, K, true));
  k_spmv<TA, K, true><<<gz, kThreads, 0, s>>>(a.n, a.tptr, a.tcol, a.ltval, a.t, a.z, a.r, a.dpart,
                                              a.st);
  k_dot_rz<K, false><<<1, kThreads, 0, s>>>(a.dpart, gz, a.st);
  k_upd_p<TA, K, false><<<gn, kThreads, 0, s>>>(a.n, a.Pinv, a.p, a.z, a.pv, a.st);
  return hipGetLastError();
}

template <typename TA, int K>
hipError_t finish_k(hipStream_t s, const SolveArgs<TA> &a, float *X, int64_t ldx, int64_t col0,
                    int kval, double *X64) {
  k_finish<TA, K><<<blocks_for(a.n), kThreads, 0, s>>>(a.n, a.P, a.x, X, ldx, col0, kval, X64);
  return hipGetLastError();
}

// X64 == nullptr: R = M^T B, with FP32 products summed in FP32 over 32-row
// chunks and FP64 across chunks (sets ||M^T b|| and the first right-hand side;
// both only need FP32-level accuracy).  Otherwise R = M^T (B - M X64) fully in
// FP64 -- the residual the stopping test certifies.
template <int K>
hipError_t residual_k(hipStream_t s, const SolveArgs<double> &a, const float *B, int64_t ldb,
                      int64_t col0, int kval, const double *X64, double *R, double *out2,
                      TrackFn fn, void *tctx) {
  if (X64 != nullptr) {
    track(fn, tctx, 4, 0);
    mv_rows<double, K>(s, a, X64, B, ldb, col0, kval, a.y, nullptr);
    track(fn, tctx, 4, 1);
    track(fn, tctx, 5, 0);
    mv_cols<double, double, K>(s, a, a.y, nullptr);
    track(fn, tctx, 5, 1);
  } else {
    float *yf = reinterpret_cast<float *>(a.y);
    k_load_boundary<float, K><<<blocks_for(a.m), kThreads, 0, s>>>(a.m, B, ldb, col0, kval, yf);
    track(fn, tctx, 2, 0);
    mv_cols<double, float, K>(s, a, yf, nullptr);
    track(fn, tctx, 2, 1);
  }
  k_mv_reduce<double, K, RED_ORIG><<<blocks_for(a.n), kThreads, 0, s>>>(
      a.part, a.row_blocks, a.n, a.P, nullptr, nullptr, R, a.dpart, a.st, out2, 0.0, kval, 0);
  return hipGetLastError();
}

// Source-space FP64 residual R = B - M^T M X in original order (R = B when
// X64 is null), per-column ||R||^2 into out2.
template <int K>
hipError_t residual_src_k(hipStream_t s, const SolveArgs<double> &a, const float *B, int64_t ldb,
                          int64_t col0, int kval, const double *X64, double *R, double *out2,
                          TrackFn fn, void *tctx) {
  if (X64 != nullptr) {
    track(fn, tctx, 4, 0);
    mv_rows<double, K>(s, a, X64, nullptr, 0, 0, kval, a.y, nullptr);
    track(fn, tctx, 4, 1);
    track(fn, tctx, 5, 0);
    mv_cols<double, double, K>(s, a, a.y, nullptr);
    track(fn, tctx, 5, 1);
    k_mv_reduce<double, K, RED_ORIG><<<blocks_for(a.n), kThreads, 0, s>>>(
        a.part, a.row_blocks, a.n, a.P, nullptr, nullptr, R, a.dpart, a.st, out2, 0.0, kval, 0);
  }
  k_source_residual<K><<<blocks_for(a.n), kThreads, 0, s>>>(a.n, B, ldb, col0, kval,
                                                           X64 != nullptr ? 1 : 0, R, a.dpart,
                                                           a.st, out2);
  return hipGetLastError();
}

}  // namespace

#define MFS_DISPATCH_K(K, CALL)                  \
  switch (K) {                                   \
    case 1: { constexpr int KK = 1; return CALL; } \
    case 2: { constexpr int KK = 2; return CALL; } \
    case 4: { constexpr int KK = 4; return CALL; } \
    case 8: { constexpr int KK = 8; return CALL; } \
    case 16: { constexpr int KK = 16; return CALL; } \
    default: return hipErrorInvalidValue;        \
  }

template <typename TA>
hipError_t pcg_init(hipStream_t s, const SolveArgs<TA> &a, const float *B, int64_t ldb,
                    int64_t col0, int rhs_space, int kval, double tol, int32_t maxit,
                    const double *B64, const double *thresh_abs) {
  MFS_DISPATCH_K(a.K, (init_k<TA, KK>(s, a, B, ldb, col0, rhs_space, kval, tol, maxit, B64, thresh_abs)));
}

template <typename TA>
hipError_t pcg_iteration(hipStream_t s, const SolveArgs<TA> &a, TrackFn fn, void *tctx) {
  MFS_DISPATCH_K(a.K, (iter_k<TA, KK>(s, a, fn, tctx)));
}

template <typename TA>
hipError_t pcg_finish(hipStream_t s, const SolveArgs<TA> &a, float *X, int64_t ldx, int64_t col0,
                      int kval, double *X64) {
  MFS_DISPATCH_K(a.K, (finish_k<TA, KK>(s, a, X, ldx, col0, kval, X64)));
}

// Q (n x kval, FP64, original order) = the sum of `splits` partials.
__global__ void k_sum_splits(const double *__restrict__ part, int32_t splits, int64_t n, int K,
                             int kval, double *__restrict__ Q) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < kval; ++r) {
    double s = 0.0;
    for (int32_t b = 0; b < splits; ++b) s += part[(int64_t(b) * n + i) * K + r];
    Q[i * kval + r] = s;
  }
}

hipError_t normal_apply32(hipStream_t s, const SolveArgs<float> &a, int kval, double *Q,
                          TrackFn fn, void *tctx) {
  int32_t splits = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (a.K) {
    case 1: e = normal_matvec<float, 1>(s, a, fn, tctx, &splits); break;
    case 2: e = normal_matvec<float, 2>(s, a, fn, tctx, &splits); break;
    case 4: e = normal_matvec<float, 4>(s, a, fn, tctx, &splits); break;
    case 8: e = normal_matvec<float, 8>(s, a, fn, tctx, &splits); break;
    case 16: e = normal_matvec<float, 16>(s, a, fn, tctx, &splits); break;
    default: break;
  }
  if (e != hipSuccess) return e;
  k_sum_splits<<<blocks_for(a.n), kThreads, 0, s>>>(a.part, splits, a.n, a.K, kval, Q);
  return hipGetLastError();
}

hipError_t residual64(hipStream_t s, const SolveArgs<double> &a, const float *B, int64_t ldb,
                      int64_t col0, int kval, const double *X64, double *R, double *out2,
                      TrackFn fn, void *tctx) {
  MFS_DISPATCH_K(a.K, (residual_k<KK>(s, a, B, ldb, col0, kval, X64, R, out2, fn, tctx)));
}

hipError_t residual64_source(hipStream_t s, const SolveArgs<double> &a, const float *B, int64_t ldb,
                             int64_t col0, int kval, const double *X64, double *R, double *out2,
                             TrackFn fn, void *tctx) {
  MFS_DISPATCH_K(a.K, (residual_src_k<KK>(s, a, B, ldb, col0, kval, X64, R, out2, fn, tctx)));
}

template hipError_t pcg_init<double>(hipStream_t, const SolveArgs<double> &, const float *, int64_t,
                                     int64_t, int, int, double, int32_t, const double *,
                                     const double *);
template hipError_t pcg_init<float>(hipStream_t, const SolveArgs<float> &, const float *, int64_t,
                                    int64_t, int, int, double, int32_t, const double *,
                                    const double *);

// thr[r] = (frac * tol)^2 * nb2[r]: absolute inner thresholds from ||M^T b||^2.
__global__ void k_inner_thresholds(const double *nb2, double tol_frac, int K, double *thr) {
  const int r = threadIdx.x;
  if (r < K) thr[r] = tol_frac * tol_frac * nb2[r];
}

hipError_t inner_thresholds(hipStream_t s, const double *nb2, double tol_frac, int K, double *thr) {
  k_inner_thresholds<<<1, 64, 0, s>>>(nb2, tol_frac, K, thr);
  return hipGetLastError();
}
hipError_t pcg_finish64(hipStream_t s, const SolveArgs<double> &a, double *X, int64_t ldx,
                        int64_t col0, int kval) {
  MFS_DISPATCH_K(a.K, (k_finish64<KK><<<blocks_for(a.n), kThreads, 0, s>>>(a.n, a.P, a.x, X, ldx, col0,
                                                                          kval),
                       hipGetLastError()));
}

template hipError_t pcg_iteration<double>(hipStream_t, const SolveArgs<double> &, TrackFn, void *);
template hipError_t pcg_iteration<float>(hipStream_t, const SolveArgs<float> &, TrackFn, void *);
template hipError_t pcg_finish<double>(hipStream_t, const SolveArgs<double> &, float *, int64_t,
                                       int64_t, int, double *);
template hipError_t pcg_finish<float>(hipStream_t, const SolveArgs<float> &, float *, int64_t,
                                      int64_t, int, double *);

// ------------------------------------------------------------------ backward products
// Y (m x K, FP32) = M X (X: n x K FP32, K padded) through the PCG's row pass.
hipError_t dense_rows32(hipStream_t s, const float *M, int64_t m, int64_t n, const float *X, int K,
                        float *Y) {
  SolveArgs<float> a{};
  a.M = M;
  a.m = m;
  a.n = n;
  a.K = K;
  a.pts = PointsOp{};
  switch (K) {
    case 1: mv_rows<float, 1>(s, a, X, nullptr, 0, 0, K, Y, nullptr); break;
    case 2: mv_rows<float, 2>(s, a, X, nullptr, 0, 0, K, Y, nullptr); break;
    case 4: mv_rows<float, 4>(s, a, X, nullptr, 0, 0, K, Y, nullptr); break;
    case 8: mv_rows<float, 8>(s, a, X, nullptr, 0, 0, K, Y, nullptr); break;
    case 16: mv_rows<float, 16>(s, a, X, nullptr, 0, 0, K, Y, nullptr); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

namespace {

// G (m x n, FP32) = U W^T, U: m x K, W: n x K (FP32, row-major, K padded).
// Block: 32 rows x 1024 columns; a thread owns 4 adjacent columns and keeps
// their K values of W in registers; the block walks its rows together, so a
// row of U is the same for every lane and arrives by scalar loads as SGPR
// operands of the FMAs; each row of G is written as whole float4 lines.  FP32
// products summed over r in order.
constexpr int kOuterRows = 32;
template <int K>
__global__ __launch_bounds__(kThreads) void k_outer(int64_t m, int64_t n, const float *__restrict__ U,
                                                    const float *__restrict__ W, float *__restrict__ G) {
  const int64_t c0 = (int64_t(blockIdx.x) * kThreads + threadIdx.x) * 4;
  const int64_t f0 = int64_t(blockIdx.y) * kOuterRows;
  if (c0 >= n) return;
  float w[4][K];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int r = 0; r < K; ++r) w[j][r] = (c0 + j < n) ? W[(c0 + j) * K + r] : 0.f;
  const bool vec = (n % 4 == 0);
  const int rows = int(m - f0 < kOuterRows ? m - f0 : kOuterRows);
  // rows in groups of RG: the group's U rows are loaded together (one wait per
  // group instead of one per row); rows past the block's end read its first
  // row and are not stored
  constexpr int RG = K <= 8 ? 8 : 4;
  for (int i0 = 0; i0 < rows; i0 += RG) {
    float u[RG][K];
#pragma unroll
    for (int g = 0; g < RG; ++g) {
      const float *__restrict__ up = U + (f0 + (i0 + g < rows ? i0 + g : 0)) * K;  // uniform
#pragma unroll
      for (int r = 0; r < K; ++r) u[g][r] = up[r];
    }
#pragma unroll
    for (int g = 0; g < RG; ++g) {
      if (i0 + g >= rows) break;
      float o[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < K; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] += u[g][r] * w[j][r];
      float *dst = G + (f0 + i0 + g) * n + c0;
      if (vec) {
        *reinterpret_cast<float4 *>(dst) = make_float4(o[0], o[1], o[2], o[3]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (c0 + j < n) dst[j] = o[j];
      }
    }
  }
}

}  // namespace

hipError_t outer_update32(hipStream_t s, int64_t m, int64_t n, int K, const float *U, const float *W,
                          float *G) {
  const dim3 grid(unsigned((n + 4 * kThreads - 1) / (4 * kThreads)),
                  unsigned((m + kOuterRows - 1) / kOuterRows));
  switch (K) {
    case 2: k_outer<2><<<grid, kThreads, 0, s>>>(m, n, U, W, G); break;
    case 4: k_outer<4><<<grid, kThreads, 0, s>>>(m, n, U, W, G); break;
    case 8: k_outer<8><<<grid, kThreads, 0, s>>>(m, n, U, W, G); break;
    case 16: k_outer<16><<<grid, kThreads, 0, s>>>(m, n, U, W, G); break;
    case 32: k_outer<32><<<grid, kThreads, 0, s>>>(m, n, U, W, G); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ kernel assembly
// M[i,j] = -(x_i - y_j).n_i / (4 pi r^3) in FP64, stored FP32; thread per
// (row, column), rows strided over gridDim.y (at most 65,535 row blocks).
__global__ __launch_bounds__(kThreads) void k_assemble(const double *__restrict__ x,
                                                       const double *__restrict__ nx,
                                                       const double *__restrict__ y, int64_t m,
                                                       int64_t n, float *__restrict__ M) {
  const int64_t j = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (j >= n) return;
  for (int64_t i = blockIdx.y; i < m; i += gridDim.y)
    M[i * n + j] = float(kernel64(x[3 * i] - y[3 * j], x[3 * i + 1] - y[3 * j + 1],
                                  x[3 * i + 2] - y[3 * j + 2], nx[3 * i], nx[3 * i + 1],
                                  nx[3 * i + 2]));
}

hipError_t launch_assemble(hipStream_t s, const double *x, const double *nx, const double *y,
                           int64_t m, int64_t n, float *M) {
  dim3 grid(blocks_for(n), unsigned(m < 65535 ? m : 65535));
  k_assemble<<<grid, kThreads, 0, s>>>(x, nx, y, m, n, M);
  return hipGetLastError();
}

}  // namespace mfs
