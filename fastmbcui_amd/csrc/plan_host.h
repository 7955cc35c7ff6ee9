// plan_host.h -- host side of the plan: maximin ordering, supernode aggregation,
// supernodal pattern and the row-major (transposed) view of the factor pattern.
//
// The ordering is a serial priority-queue walk; it stays on the host, in native
// code, because every step depends on the previous one (each pivot's
// candidates come from the heap the previous pivot left).  The rho-radius
// pattern itself is computed on the GPU (k_plan.hip).
#pragma once

#include <cstdint>
#include <vector>

namespace mfs {

enum PlanStatus { PLAN_OK = 0, PLAN_ERR_ARG = 1, PLAN_ERR_DUPLICATE = 3 };

// Reverse maximin ordering of n points (xyz row-major FP64).
// Out: P[orig] = new, Pinv[new] = orig, scale[new] ascending.
// Same result as src/preprocess.h:782-831 + src/maxmin.h:588-665.
int order_maximin(int32_t n, const double *xyz, std::vector<int32_t> &P,
                  std::vector<int32_t> &Pinv, std::vector<double> &scale);

// Uniform grids for the rho-radius neighbour search (k_plan.hip).  Columns are
// cut into bands over which the length scale at most doubles; band b's grid
// holds the points i >= j0 (the only possible rows of its columns) in cubic
// cells no smaller than the band's largest search radius rho * l_j, so every
// row of a column lies in the 27 cells around the column's own cell.
struct GridBand {
  int32_t j0, j1;      // columns [j0, j1)
  int32_t dims[3];
  int32_t pad;
  double lo[3];        // grid origin
  double inv_h;        // 1 / cell size
  int64_t cell_off;    // cells start at cell_start[cell_off], dims product + 1 entries
  int64_t pt_off;      // point indices of the band at cell_pts[pt_off + cell_start[...]]
};

// Builds the grid bands for points (new order, FP64 coordinate arrays) with
// ascending scale.  Columns from *j_brute on are left to the exhaustive scan
// (their grids would not prune: a few coarse points spread over the scene).
void build_grid_bands(int32_t n, const double *px, const double *py, const double *pz,
                      const double *scale, double rho, std::vector<GridBand> &bands,
                      std::vector<int32_t> &cell_start, std::vector<int32_t> &cell_pts,
                      int32_t *j_brute);

// Greedy supernodes, src/preprocess.h:541-602 with group {0, n}.
int aggregate(int32_t n, const int32_t *ptr, const int32_t *ind, const double *scale,
              double lambda, int32_t max_super, std::vector<int32_t> &sup_ptr,
              std::vector<int32_t> &sup_ind, std::vector<int32_t> &sup_parent);

// Supernodal pattern, src/preprocess.h:604-671.
void super_pattern(int32_t n, const int32_t *ptr, const int32_t *ind,
                   const std::vector<int32_t> &sup_parent, int32_t n_super,
                   std::vector<int32_t> &sptr, std::vector<int32_t> &sind);

// Row-major view of the CSC factor pattern: row i lists the columns j that
// hold it (ascending) and, in tmap, the CSC position of each entry.
void transpose_pattern(int32_t n, const std::vector<int32_t> &sptr,
                       const std::vector<int32_t> &sind, std::vector<int32_t> &tptr,
                       std::vector<int32_t> &tcol, std::vector<int32_t> &tmap);

}  // namespace mfs
