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
// Neighbour search: a uniform grid per band of columns (plan_host.h,
// build_grid_bands) whose cells are at least as wide as the band's largest
// search radius, so one wave per column visits the 27 cells around the
// column's point and tests only the points found there; the hits are written
// unordered and then sorted per column.  The few coarsest columns, whose
// radii span the scene, and any column with more than kSortMax rows use the
// exhaustive scan: one wave per column over the rows i >= j in 64-wide
// strides, where ballot + mbcnt place the hits in ascending order.
#include "mfs_internal.h"
#include "plan_host.h"

namespace mfs {
namespace {

__device__ inline double sq_dist(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  const double a = dx * dx;
  const double b = dy * dy;
  const double c = dz * dz;
  return (a + b) + c;
}

// The pattern predicate for row i of column j (both in the new order).
__device__ inline bool in_pattern(const double *__restrict__ px, const double *__restrict__ py,
                                  const double *__restrict__ pz, const double *__restrict__ scale,
                                  double rho, int32_t i, double xj, double yj, double zj,
                                  double lj2) {
  const double d2 = sq_dist(px[i] - xj, py[i] - yj, pz[i] - zj);
  const double li = rho * scale[i];
  const double li2 = li * li;
  const double lim = (li2 < lj2) ? li2 : lj2;
  return d2 < lim;
}

// Exhaustive scan for columns j_first + w (or cols[w] when cols is given).
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_pattern(int32_t n, int32_t j_first,
                                                      const int32_t *__restrict__ cols,
                                                      int32_t ncols, const double *__restrict__ px,
                                                      const double *__restrict__ py,
                                                      const double *__restrict__ pz,
                                                      const double *__restrict__ scale, double rho,
                                                      int32_t *__restrict__ cnt,
                                                      const int32_t *__restrict__ ptr,
                                                      int32_t *__restrict__ ind) {
  const int lane = threadIdx.x & 63;
  const int32_t w = int32_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (w >= ncols) return;
  const int32_t j = cols ? cols[w] : j_first + w;
  const double xj = px[j], yj = py[j], zj = pz[j];
  const double lj = rho * scale[j];
  const double lj2 = lj * lj;
  int32_t base = FILL ? ptr[j] : 0;
  for (int32_t i0 = j; i0 < n; i0 += 64) {
    const int32_t i = i0 + lane;
    const bool hit = (i < n) && in_pattern(px, py, pz, scale, rho, i, xj, yj, zj, lj2);
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

// Grid search for the columns j < j_end (all of them in some band).  Hits
// are written in cell order; k_sort_columns orders them afterwards.
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_grid_pattern(
    int32_t j_end, const GridBand *__restrict__ bands, int32_t nbands,
    const int32_t *__restrict__ cell_start, const int32_t *__restrict__ cell_pts,
    const double *__restrict__ px, const double *__restrict__ py, const double *__restrict__ pz,
    const double *__restrict__ scale, double rho, int32_t *__restrict__ cnt,
    const int32_t *__restrict__ ptr, int32_t *__restrict__ ind) {
  const int lane = threadIdx.x & 63;
  const int32_t j = int32_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= j_end) return;
  int32_t b = 0;
  while (b + 1 < nbands && j >= bands[b].j1) ++b;
  const GridBand B = bands[b];
  const double xj = px[j], yj = py[j], zj = pz[j];
  const double lj = rho * scale[j];
  const double lj2 = lj * lj;
  const double q[3] = {xj, yj, zj};
  int64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = int64_t((q[d] - B.lo[d]) * B.inv_h);
    c[d] = c[d] < 0 ? 0 : (c[d] > B.dims[d] - 1 ? B.dims[d] - 1 : c[d]);
  }
  const int32_t *starts = cell_start + B.cell_off;
  const int32_t *pts = cell_pts + B.pt_off;
  int32_t base = FILL ? ptr[j] : 0;
  for (int64_t z = c[2] - 1; z <= c[2] + 1; ++z) {
    if (z < 0 || z >= B.dims[2]) continue;
    for (int64_t y = c[1] - 1; y <= c[1] + 1; ++y) {
      if (y < 0 || y >= B.dims[1]) continue;
      for (int64_t x = c[0] - 1; x <= c[0] + 1; ++x) {
        if (x < 0 || x >= B.dims[0]) continue;
        const int64_t id = (z * B.dims[1] + y) * B.dims[0] + x;
        const int32_t s = starts[id], e = starts[id + 1];
        for (int32_t k0 = s; k0 < e; k0 += 64) {
          const int32_t k = k0 + lane;
          int32_t i = -1;
          if (k < e) i = pts[k];
          const bool hit = (i >= j) && in_pattern(px, py, pz, scale, rho, i, xj, yj, zj, lj2);
          const uint64_t mask = __ballot(hit);
          if (FILL && hit) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
            ind[base + int32_t(below)] = i;
          }
          base += int32_t(__popcll(mask));
        }
      }
    }
  }
  if (!FILL && lane == 0) cnt[j] = base;
}

// Ascending order of each column j < j_end with 2..kSortMax rows: bitonic sort
// in LDS, one block per column.
constexpr int kSortMax = 1024;

__global__ __launch_bounds__(kThreads) void k_sort_columns(const int32_t *__restrict__ ptr,
                                                           int32_t *__restrict__ ind) {
  __shared__ int32_t v[kSortMax];
  const int32_t j = blockIdx.x;
  const int32_t s = ptr[j], c = ptr[j + 1] - s;
  if (c < 2 || c > kSortMax) return;
  int P = 2;
  while (P < c) P <<= 1;
  for (int t = threadIdx.x; t < P; t += kThreads) v[t] = (t < c) ? ind[s + t] : INT32_MAX;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int h = k >> 1; h > 0; h >>= 1) {
      for (int t = threadIdx.x; t < P; t += kThreads) {
        const int u = t ^ h;
        if (u > t) {
          const bool up = (t & k) == 0;
          const int32_t a = v[t], bb = v[u];
          if ((a > bb) == up) {
            v[t] = bb;
            v[u] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < c; t += kThreads) ind[s + t] = v[t];
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

hipError_t launch_pattern(hipStream_t st, int32_t n, int32_t j_first, const int32_t *cols,
                          int32_t ncols, const double *px, const double *py, const double *pz,
                          const double *scale, double rho, int32_t *cnt, int32_t *ptr,
                          int32_t *ind, int fill) {
  if (ncols <= 0) return hipSuccess;
  const unsigned grid = unsigned((ncols + 3) / 4);
  if (fill)
    k_pattern<true><<<grid, kThreads, 0, st>>>(n, j_first, cols, ncols, px, py, pz, scale, rho,
                                               nullptr, ptr, ind);
  else
    k_pattern<false><<<grid, kThreads, 0, st>>>(n, j_first, cols, ncols, px, py, pz, scale, rho,
                                                cnt, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_grid_pattern(hipStream_t st, int32_t j_end, const GridBand *bands,
                               int32_t nbands, const int32_t *cell_start, const int32_t *cell_pts,
                               const double *px, const double *py, const double *pz,
                               const double *scale, double rho, int32_t *cnt, int32_t *ptr,
                               int32_t *ind, int fill) {
  if (j_end <= 0) return hipSuccess;
  const unsigned grid = unsigned((j_end + 3) / 4);
  if (fill) {
    k_grid_pattern<true><<<grid, kThreads, 0, st>>>(j_end, bands, nbands, cell_start, cell_pts,
                                                    px, py, pz, scale, rho, nullptr, ptr, ind);
    k_sort_columns<<<unsigned(j_end), kThreads, 0, st>>>(ptr, ind);
  } else {
    k_grid_pattern<false><<<grid, kThreads, 0, st>>>(j_end, bands, nbands, cell_start, cell_pts,
                                                     px, py, pz, scale, rho, cnt, nullptr,
                                                     nullptr);
  }
  return hipGetLastError();
}

int32_t pattern_sort_max() { return kSortMax; }

hipError_t launch_scan(hipStream_t st, const int32_t *cnt, int32_t n, int32_t *ptr,
                       int64_t *total) {
  k_scan<<<1, 1024, 0, st>>>(cnt, n, ptr, total);
  return hipGetLastError();
}

}  // namespace mfs
