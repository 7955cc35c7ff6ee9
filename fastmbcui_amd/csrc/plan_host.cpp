// plan_host.cpp -- host side of the plan (see plan_host.h).
//
// Built with -ffp-contract=off: distances are ((dx*dx + dy*dy) + dz*dz) with
// each product rounded, the order the reference's host build evaluates them in
// (src/maxmin.h:370-374, Eigen squaredNorm, nanoflann L2_Adaptor).
#include "plan_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace mfs {
namespace {

inline double sqdist(const double *a, const double *b) {
  const double dx = a[0] - b[0];
  const double dy = a[1] - b[1];
  const double dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

// Binary max-heap of (key, point) slots addressed by position, with the
// reference's shape and tie rules (src/maxmin.h:194-293):
//  * position k has children 2k+1 and 2k+2;
//  * keys start at +inf where a position has fewer than two children and at
//    10000 where it has two;
//  * a key only ever decreases, followed by a sift-down that takes the right
//    child unless the left one is strictly larger, and swaps only when the
//    parent is strictly smaller.
class MaxHeap {
 public:
  explicit MaxHeap(int32_t n) : n_(n), slot_(n), at_(n) {
    for (int32_t k = 0; k < n; ++k) {
      const bool two = int64_t(2) * k + 2 < n;
      slot_[k].key = two ? 10000.0 : std::numeric_limits<double>::infinity();
      slot_[k].pt = k;
      at_[k] = k;
    }
  }
  int32_t top() const { return slot_[0].pt; }
  double top_key() const { return slot_[0].key; }
  double key_of(int32_t pt) const { return slot_[at_[pt]].key; }

  void lower(int32_t pt, double key) {
    int32_t k = at_[pt];
    slot_[k].key = key;
    for (;;) {
      const int32_t l = 2 * k + 1;
      if (l >= n_) return;
      const int32_t r = l + 1;
      const int32_t c = (r >= n_ || slot_[l].key > slot_[r].key) ? l : r;
      if (!(slot_[k].key < slot_[c].key)) return;
      std::swap(slot_[k], slot_[c]);
      at_[slot_[k].pt] = k;
      at_[slot_[c].pt] = c;
      k = c;
    }
  }

 private:
  struct Slot {
    double key;
    int32_t pt;
  };
  int32_t n_;
  std::vector<Slot> slot_;
  std::vector<int32_t> at_;
};

}  // namespace

int order_maximin(int32_t n, const double *xyz, std::vector<int32_t> &P,
                  std::vector<int32_t> &Pinv, std::vector<double> &scale) {
  if (n <= 0 || xyz == nullptr) return PLAN_ERR_ARG;

  // Root: closest to the bounding-box minimum corner, first index on ties
  // (src/preprocess.h:785, 815-823).
  double lo[3] = {xyz[0], xyz[1], xyz[2]};
  for (int32_t i = 1; i < n; ++i)
    for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], xyz[3 * int64_t(i) + d]);
  int32_t root = -1;
  double best = 1e10;
  for (int32_t i = 0; i < n; ++i) {
    const double d = sqdist(xyz + 3 * int64_t(i), lo);
    if (d < best) {
      best = d;
      root = i;
    }
  }
  if (root < 0) return PLAN_ERR_ARG;

  // Pivot k's candidate list is kids[head[k] .. head[k+1]); a point's list
  // index is the pivot whose list it must be searched from (src/maxmin.h:295-336,
  // 440-464).  Distances to a boundary are not used: the reference passes no
  // boundary tree, so its per-point bound stays 1e16 (maxmin.h:619).
  const double kBound = 1e16;
  MaxHeap heap(n);
  std::vector<int32_t> head(size_t(n) + 1, 0);
  std::vector<int32_t> kids;
  kids.reserve(size_t(n) * 8);
  std::vector<int32_t> owner(n, 0);
  std::vector<int32_t> seq(n);
  std::vector<double> dist(n);

  seq[0] = root;
  dist[0] = 0.0;
  const double *xr = xyz + 3 * int64_t(root);
  for (int32_t k = 0; k < n; ++k) {
    kids.push_back(k);
    const double d = std::sqrt(sqdist(xr, xyz + 3 * int64_t(k)));
    if (d > dist[0]) dist[0] = d;
    heap.lower(k, std::min(d, kBound));
  }
  head[1] = n;

  for (int32_t k = 1; k < n; ++k) {
    const int32_t piv = heap.top();
    const double pd = heap.top_key();
    seq[k] = piv;
    dist[k] = pd;
    const int32_t from = head[owner[piv]], to = head[owner[piv] + 1];
    head[k + 1] = head[k];
    const double *xp = xyz + 3 * int64_t(piv);
    const double pd2 = pd * pd;
    for (int32_t t = from; t < to; ++t) {
      const int32_t c = kids[t];
      const double d2 = sqdist(xp, xyz + 3 * int64_t(c));
      if (!(d2 < pd2)) continue;
      double cd = heap.key_of(c);
      if (d2 < cd * cd) {
        cd = std::min(std::sqrt(d2), kBound);
        heap.lower(c, cd);
      }
      kids.push_back(c);
      head[k + 1] += 1;
      if (std::sqrt(d2) + cd < pd) owner[c] = k;
    }
  }

  // Permutation (fine -> coarse) and ascending scale (preprocess.h:822-830).
  P.assign(n, -1);
  Pinv.resize(n);
  scale.resize(n);
  for (int32_t k = 0; k < n; ++k) {
    if (P[seq[k]] != -1) return PLAN_ERR_DUPLICATE;
    P[seq[k]] = n - 1 - k;
    Pinv[n - 1 - k] = seq[k];
    scale[n - 1 - k] = dist[k];
  }
  for (int32_t i = 0; i + 1 < n; ++i)
    if (!(scale[i] <= scale[i + 1])) return PLAN_ERR_DUPLICATE;
  if (n > 1 && !(scale[0] > 0.0)) return PLAN_ERR_DUPLICATE;
  return PLAN_OK;
}

int aggregate(int32_t n, const int32_t *ptr, const int32_t *ind, const double *scale,
              double lambda, int32_t max_super, std::vector<int32_t> &sup_ptr,
              std::vector<int32_t> &sup_ind, std::vector<int32_t> &sup_parent) {
  std::vector<uint8_t> taken(n, 0);
  sup_ptr.clear();
  sup_ptr.push_back(0);
  sup_ind.clear();
  sup_ind.reserve(n);
  sup_parent.assign(n, -1);
  int32_t s = 0;
  for (int32_t j = 0; j < n; ++j) {
    if (taken[j]) continue;
    const double lim = lambda * scale[j];
    int32_t size = 0;
    for (int32_t e = ptr[j]; e < ptr[j + 1]; ++e) {
      const int32_t i = ind[e];
      if (taken[i] || !(scale[i] <= lim)) continue;
      taken[i] = 1;
      sup_ind.push_back(i);
      sup_parent[i] = s;
      if (++size >= max_super) break;
    }
    sup_ptr.push_back(int32_t(sup_ind.size()));
    ++s;
  }
  return int32_t(sup_ind.size()) == n ? PLAN_OK : PLAN_ERR_DUPLICATE;
}

void super_pattern(int32_t n, const int32_t *ptr, const int32_t *ind,
                   const std::vector<int32_t> &sup_parent, int32_t n_super,
                   std::vector<int32_t> &sptr, std::vector<int32_t> &sind) {
  // Each supernode's row set is the sorted union of its columns' rows.  Walk
  // the columns of each supernode, mark rows with a per-supernode stamp, then
  // sort the collected rows.
  std::vector<int32_t> first_of(n_super, -1);
  std::vector<std::vector<int32_t>> members(n_super);
  for (int32_t j = 0; j < n; ++j) members[sup_parent[j]].push_back(j);
  std::vector<int32_t> stamp(n, -1);
  std::vector<int32_t> set_ptr(size_t(n_super) + 1, 0);
  std::vector<int32_t> set_val;
  set_val.reserve(size_t(ptr[n]) * 2);
  for (int32_t s = 0; s < n_super; ++s) {
    const size_t base = set_val.size();
    for (int32_t j : members[s])
      for (int32_t e = ptr[j]; e < ptr[j + 1]; ++e) {
        const int32_t i = ind[e];
        if (stamp[i] != s) {
          stamp[i] = s;
          set_val.push_back(i);
        }
      }
    std::sort(set_val.begin() + base, set_val.end());
    set_ptr[s + 1] = int32_t(set_val.size());
  }
  sptr.assign(size_t(n) + 1, 0);
  std::vector<int32_t> lo(n);
  for (int32_t j = 0; j < n; ++j) {
    const int32_t s = sup_parent[j];
    const int32_t *b = set_val.data() + set_ptr[s];
    const int32_t *e = set_val.data() + set_ptr[s + 1];
    lo[j] = int32_t(std::lower_bound(b, e, j) - set_val.data());
    sptr[j + 1] = sptr[j] + (set_ptr[s + 1] - lo[j]);
  }
  sind.resize(sptr[n]);
  for (int32_t j = 0; j < n; ++j) {
    const int32_t s = sup_parent[j];
    std::memcpy(sind.data() + sptr[j], set_val.data() + lo[j],
                sizeof(int32_t) * size_t(set_ptr[s + 1] - lo[j]));
  }
}

void transpose_pattern(int32_t n, const std::vector<int32_t> &sptr,
                       const std::vector<int32_t> &sind, std::vector<int32_t> &tptr,
                       std::vector<int32_t> &tcol, std::vector<int32_t> &tmap) {
  const int32_t nnz = sptr[n];
  tptr.assign(size_t(n) + 1, 0);
  for (int32_t e = 0; e < nnz; ++e) tptr[sind[e] + 1]++;
  for (int32_t i = 0; i < n; ++i) tptr[i + 1] += tptr[i];
  tcol.resize(nnz);
  tmap.resize(nnz);
  std::vector<int32_t> fill(tptr.begin(), tptr.end() - 1);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t e = sptr[j]; e < sptr[j + 1]; ++e) {
      const int32_t t = fill[sind[e]]++;
      tcol[t] = j;
      tmap[t] = e;
    }
}

void build_grid_bands(int32_t n, const double *px, const double *py, const double *pz,
                      const double *scale, double rho, std::vector<GridBand> &bands,
                      std::vector<int32_t> &cell_start, std::vector<int32_t> &cell_pts,
                      int32_t *j_brute) {
  // Cells per band at most twice the band's points (the clouds are surfaces,
  // most cells of a box are empty), larger cells beyond; a band whose grid
  // would have fewer than kMinCells cells cannot prune and is left to the scan.
  constexpr int64_t kMinCells = 64;
  bands.clear();
  cell_start.clear();
  cell_pts.clear();
  *j_brute = n;
  const double *p[3] = {px, py, pz};
  int32_t j0 = 0;
  while (j0 < n) {
    // columns whose length scale stays within twice the band's first one
    int32_t j1 = j0 + 1;
    while (j1 < n && scale[j1] <= 2.0 * scale[j0]) ++j1;
    // every row of column j lies within rho * l_j (the predicate is strict);
    // the margin keeps |cell(p) - cell(q)| <= 1 clear of rounding
    const double radius = rho * scale[j1 - 1];
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = p[d][j0];
    for (int32_t i = j0 + 1; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], p[d][i]);
        hi[d] = std::max(hi[d], p[d][i]);
      }
    const int64_t kMaxCells = std::max<int64_t>(2 * int64_t(n - j0), 4096);
    double h = radius * (1.0 + 1e-6);
    if (!(h > 0)) break;  // degenerate scale: leave the rest to the scan
    int64_t dims[3], cells = 1;
    for (int pass = 0; pass < 16; ++pass) {
      cells = 1;
      for (int d = 0; d < 3; ++d) {
        dims[d] = std::max<int64_t>(1, int64_t((hi[d] - lo[d]) / h) + 1);
        cells *= dims[d];
      }
      if (cells <= kMaxCells) break;
      h *= std::max(std::cbrt(double(cells) / double(kMaxCells)), 1.05);
    }
    if (cells < kMinCells || cells > kMaxCells) break;
    GridBand b{};
    b.j0 = j0;
    b.j1 = j1;
    for (int d = 0; d < 3; ++d) {
      b.dims[d] = int32_t(dims[d]);
      b.lo[d] = lo[d];
    }
    b.inv_h = 1.0 / h;
    b.cell_off = int64_t(cell_start.size());
    b.pt_off = int64_t(cell_pts.size());
    // counting sort of the points i >= j0 by cell (same arithmetic as the query)
    std::vector<int32_t> cell_of(size_t(n - j0));
    std::vector<int32_t> cnt(size_t(cells) + 1, 0);
    for (int32_t i = j0; i < n; ++i) {
      int64_t c[3];
      for (int d = 0; d < 3; ++d) {
        c[d] = int64_t((p[d][i] - lo[d]) * b.inv_h);
        c[d] = std::min<int64_t>(std::max<int64_t>(c[d], 0), dims[d] - 1);
      }
      const int32_t id = int32_t((c[2] * dims[1] + c[1]) * dims[0] + c[0]);
      cell_of[size_t(i - j0)] = id;
      ++cnt[size_t(id) + 1];
    }
    for (int64_t c = 0; c < cells; ++c) cnt[size_t(c) + 1] += cnt[size_t(c)];
    cell_start.insert(cell_start.end(), cnt.begin(), cnt.end());
    cell_pts.resize(cell_pts.size() + size_t(n - j0));
    for (int32_t i = j0; i < n; ++i)
      cell_pts[size_t(b.pt_off) + size_t(cnt[size_t(cell_of[size_t(i - j0)])]++)] = i;
    bands.push_back(b);
    j0 = j1;
  }
  *j_brute = j0;
}

}  // namespace mfs
