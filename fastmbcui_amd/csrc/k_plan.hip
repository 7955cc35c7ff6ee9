// k_plan.hip -- rho-radius sparsity pattern of the KL factor on the GPU.
//
// Replaces simpl_sparsity (src/preprocess.h:235-316), which runs a nanoflann
// radius search per column on the host.  Column j of the lower-triangular
// pattern holds the rows i >= j with
//     ((dx*dx + dy*dy) + dz*dz) < min((rho*l_i)^2, (rho*l_j)^2),
// evaluated here in FP64 with contraction disabled, so each product and sum is
// rounded exactly as the reference's host code rounds it; the set, and hence
// the pattern, is bit-identical.
//
// One wave per column scans the rows i >= j in 64-wide strides; ballot +
// mbcnt place the hits in ascending order without a sort.
#include "mfs_internal.h"

namespace mfs {
namespace {

__device__ inline double sq_dist(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  const double a = dx * dx;
  const double b = dy * dy;
  const double c = dz * dz;
  return (a + b) + c;
}

template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_pattern(int32_t n, const double *__restrict__ px,
                                                      const double *__restrict__ py,
                                                      const double *__restrict__ pz,
                                                      const double *__restrict__ scale, double rho,
                                                      int32_t *__restrict__ cnt,
                                                      const int32_t *__restrict__ ptr,
                                                      int32_t *__restrict__ ind) {
  const int lane = threadIdx.x & 63;
  const int32_t j = int32_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double xj = px[j], yj = py[j], zj = pz[j];
  const double lj = rho * scale[j];
  const double lj2 = lj * lj;
  int32_t base = FILL ? ptr[j] : 0;
  for (int32_t i0 = j; i0 < n; i0 += 64) {
    const int32_t i = i0 + lane;
    bool hit = false;
    if (i < n) {
      const double d2 = sq_dist(px[i] - xj, py[i] - yj, pz[i] - zj);
      const double li = rho * scale[i];
      const double li2 = li * li;
      const double lim = (li2 < lj2) ? li2 : lj2;
      hit = d2 < lim;
    }
    const uint64_t mask = __ballot(hit);
    if (FILL && hit) {
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
      ind[base + int32_t(below)] = i;
    }
    base += int32_t(__popcll(mask));
  }
  if (!FILL && lane == 0) cnt[j] = base;
}

// Exclusive scan of cnt[0..n) into ptr[0..n], one block of 1024 threads; each
// thread owns a contiguous segment.  Sums in int64 to detect int32 overflow.
__global__ __launch_bounds__(1024) void k_scan(const int32_t *__restrict__ cnt, int32_t n,
                                               int32_t *__restrict__ ptr, int64_t *total) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t seg = (int64_t(n) + 1023) / 1024;
  const int64_t b = t * seg;
  const int64_t e = (b + seg < n) ? b + seg : n;
  int64_t s = 0;
  for (int64_t i = b; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int64_t v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int64_t run = part[t] - s;  // exclusive prefix of this segment
  for (int64_t i = b; i < e; ++i) {
    ptr[i] = int32_t(run);
    run += cnt[i];
  }
  if (t == 1023) {
    ptr[n] = int32_t(part[1023]);
    *total = part[1023];
  }
}

}  // namespace

hipError_t launch_pattern(hipStream_t st, int32_t n, const double *px, const double *py,
                          const double *pz, const double *scale, double rho, int32_t *cnt,
                          int32_t *ptr, int32_t *ind, int fill) {
  const unsigned grid = unsigned((n + 3) / 4);
  if (fill)
    k_pattern<true><<<grid, kThreads, 0, st>>>(n, px, py, pz, scale, rho, nullptr, ptr, ind);
  else
    k_pattern<false><<<grid, kThreads, 0, st>>>(n, px, py, pz, scale, rho, cnt, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_scan(hipStream_t st, const int32_t *cnt, int32_t n, int32_t *ptr,
                       int64_t *total) {
  k_scan<<<1, 1024, 0, st>>>(cnt, n, ptr, total);
  return hipGetLastError();
}

}  // namespace mfs
