// mfs_core.hip -- C-ABI entry points (include/mfs_core.h) and orchestration.
//
// One mfs_ctx holds one plan (ordering, patterns, supernodes) and one factor on
// one device, plus grow-only device workspaces reused across calls.  It owns
// no process-global state: the reference's static factor cache
// (mfs_torch/csrc/mfs_torch_ops.cu:17-19, 136-157) becomes an explicit context
// that the caller keys by geometry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mfs_core.h"
#include "kernel_eval.h"
#include "mfs_internal.h"
#include "plan_host.h"

namespace {

using namespace mfs;

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p != nullptr) return hipSuccess;
    if (p != nullptr) {
      hipError_t e = hipFree(p);
      p = nullptr;
      cap = 0;
      if (e != hipSuccess) return e;
    }
    const size_t nb = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, nb);
    if (e == hipSuccess) cap = nb;
    return e;
  }
  void release() {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// The dense factor transposes M in `phases` column ranges [phase_col(p),
// phase_col(p + 1)) (at most kMaxPhases) and starts the supernodes whose
// columns (original order) all lie below phase_col(p + 1) once range p is
// done, so the transpose of the later ranges can overlap the Gram of the
// earlier ones.  On surface scenes a supernode's columns nearly always lie in
// one body, whose sources are contiguous, so the Gram work splits about
// evenly over the phases.  The default is one range: with the Gram's products
// on BF16 matrix cores both the transpose and the Gram stream HBM, and at
// 65,536^2 the overlap made the factor slower (21.5 ms with 4 ranges against
// 20.4 ms with 1); with FP32 products (compute-bound) 4 ranges were best.
constexpr int kMaxPhases = 8;
constexpr int kDefaultPhases = 1;
int64_t phase_col(int p, int64_t n, int phases) {
  if (p >= phases) return n;
  return std::min<int64_t>(n, (int64_t(p) * n / phases + 63) / 64 * 64);
}

struct Bucket {
  int tiles;
  int32_t offset, count;
  int32_t phase_off[kMaxPhases + 1];  // phase p: items offset + [phase_off[p], phase_off[p + 1])
};

// Tile counts of the factor buckets, launch order (largest first); bit i of
// a tile-set mask enables kTileSizes[i].  A supernode of local size ns takes
// the smallest enabled count T with 16 T >= ns (zero-padded rows); 16 is
// always enabled, so every supernode up to kMaxLocal rows has one.
constexpr int kTileSizes[] = {16, 12, 8, 6, 4, 3, 2, 1};
constexpr uint32_t kAllTiles = 0xffu;

}  // namespace

// ---- scratch of mfs_dense_normal_apply_noalloc.  The only device memory that
// entry point uses beyond its inputs and outputs: the padded copy of X, y = M X
// and the split partials of M^T y.  Each GPU holds its own copy from module
// load until the process ends.  In host code `scratch` names a host-side
// stand-in, so kernels are given the device address that hipGetSymbolAddress
// returns (once per device, when the first context on it is created).
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
struct ScratchSlot {
  std::once_flag once;
  void *ptr = nullptr;
  hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[kScratchDevices];

// Called with `device` current.
void scratch_lookup(int device) {
  if (device < 0 || device >= kScratchDevices) return;
  ScratchSlot &s = g_scratch[device];
  std::call_once(s.once, [&s]() {
    void *p = nullptr;
    s.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (s.err == hipSuccess) s.ptr = p;
  });
}

}  // namespace

struct mfs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // plan (host)
  bool has_plan = false;
  int32_t n = 0;
  double rho32 = 0;
  std::vector<int32_t> P, Pinv, ptr, ind, sup_ptr, sup_ind, sup_parent, sptr, sind, tptr, tcol, tmap;
  std::vector<double> scale;
  std::vector<SuperDesc> sdesc;
  std::vector<int32_t> list;         // supernodes in launch order, bucket by bucket
  std::vector<FactorItem> items;     // the same as (scene 0, s) items
  std::vector<Bucket> buckets;
  int32_t max_local = 0;

  // plan (device)
  DevBuf dP, dPinv, dsptr, dsind, dtptr, dtcol, dtmap, dsupind, dsd, dlist, dstatus;
  DevBuf dpx, dpy, dpz, dscale, dcnt, dptr, dind, dtotal, dbands, dcells, dcpts, dlong;

  // factor
  bool has_factor = false;
  int prec = MFS_FP64;
  DevBuf dval, dtval, dMt, dval_alt, dtval_alt, dthg, dctr, dfscene;
  FactorScene fscene{};  // what the factor kernels read for this context
  int alt_prec = -1;

  // points form: packed target/source records (kernel_eval.h layouts)
  bool has_points = false;
  PointsOp pts{};
  int32_t points_blocks = 0;  // tuning: blocks per points pass (0 = default)
  int32_t cols_blocks = 0;    // tuning: blocks of the dense M^T pass (0 = default)
  int32_t fused_mode = 0;     // one-pass normal matvec: 0 for M of 2^27 entries up, 1 off, 2 always
  FusedShape frec_shape;      // the grid the exchange records were last set up for
  int cus = 0;                // compute units of the device
  DevBuf dfrec;               // exchange records of the one-pass matvec
  DevBuf dt32, dt64, ds32, ds64, dapin, dapout;

  // caller steps (solid angles, single layer): scratch and the index flag
  DevBuf dcaller, dflag;
  DevBuf dgrad;  // the backward's products (mfs_dense_grad)

  // solve workspaces
  DevBuf dvec, dpart, ddpart, dstate, dx64, dr64, dnorm;
  PcgState *hstate = nullptr;  // pinned
  double *hnorm = nullptr;     // pinned, 2 * kMaxRhs

  // statistics and timing
  mfs_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Side streams (default priority) that run the factor's size buckets side by
  // side, forked from and joined back into `stream` with events.
  static constexpr int kSide = 4;
  hipStream_t side[kSide] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kSide] = {}, tr_ev[kMaxPhases] = {};
  // factor launch shape (autotuned; the plan's buckets depend on tile_mask
  // and phases, so changing them takes effect at the next plan)
  uint32_t tile_mask = kAllTiles;
  int phases = kDefaultPhases;
  int side_streams = kSide;
  uint32_t plan_mask = kAllTiles;  // what the current plan's buckets were built with
  int plan_phases = kDefaultPhases;
  // Row splits of the factor's Gram per supernode, 1..8.  Off by default:
  // the buckets already run side by side on several streams, and at
  // m = n = 65,536 splitting the small launches made the factor slower
  // (26.7 ms against 23.1 ms); the autotuner tries it.
  int max_splits = 1;
  // Products of the dense FP32 factor's Gram: false FP32 MFMA, true (the
  // default) three BF16 MFMA products of a two-part split of each value --
  // 20.4 ms against 23.0 ms for the factor at 65,536^2, with column errors
  // against the FP64 factor no larger (max 9.9e-5 against 1.4e-4).
  bool gram_split = true;
  // Band form of the dense FP32 factor with split products (k_band_gram):
  // the Grams of the supernodes inside a band of kBandCols columns come
  // straight from row-major M, the rest from a compact M^T of their columns.
  // Packed at plan time (plan_bands); MFS_PARAM_FACTOR_BANDS switches it off.
  bool band_mode = true;
  std::vector<BandBlock> band_groups;  // blocks of 8 waves (part 0)
  std::vector<int8_t> band_xcd;        // XCD slot of each block (its band's index mod 8)
  std::vector<int32_t> band_items;     // supernode ids of the waves
  std::vector<int64_t> goff;           // per supernode: its Gram's offset in a part (-1: no band)
  int64_t gram_elems = 0;              // one part: sum over band supernodes of (16 T)^2
  int32_t band_supernodes = 0;
  std::vector<FactorItem> gin_items, rest_items;
  std::vector<Bucket> gin_buckets, rest_buckets;
  std::vector<int32_t> rest_cols;      // original columns the rest read (ascending)
  size_t rest_scratch = 8;             // th_scratch bytes of the rest's large tiles
  DevBuf dbitems, dgoff, dgram, dblocks, dginlist, drestlist, drestcols, dcolmap, dMtc;
  int track_mask = 0;
  struct Track {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs;
    size_t used = 0;
    bool open = false;
    int64_t launches = 0;
    double ms = 0.0, bytes = 0.0;
  } tr[8];
  double bytes_per_launch[8] = {0, 0, 0, 0, 0, 0, 0, 0};

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
  int hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return MFS_ERR_HIP;
  }
};

#define HIPCHK(call, what)                                      \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return ctx->hip_fail(e_, what);       \
  } while (0)

namespace {

template <class T>
hipError_t upload(mfs_ctx *ctx, DevBuf &b, const std::vector<T> &v) {
  hipError_t e = b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (e != hipSuccess) return e;
  if (v.empty()) return hipSuccess;
  return hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
}

int tiles_for(int ns, uint32_t mask) {
  const int need = std::max((ns + 15) / 16, 1);
  int best = -1;
  for (int i = 0; i < 8; ++i)
    if (((mask | 1u) >> i) & 1u)   // bit 0 (16 tiles) is always on
      if (kTileSizes[i] >= need) best = kTileSizes[i];
  return best;
}

// ---- kernel-class timing (events bracket each launch of an enabled class)
void track_cb(void *p, int cls, int phase) {
  mfs_ctx *ctx = static_cast<mfs_ctx *>(p);
  if (cls <= 0 || cls >= 8 || !(ctx->track_mask & (1 << cls))) return;
  mfs_ctx::Track &t = ctx->tr[cls];
  if (phase == 0) {
    if (t.used == t.pairs.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess) return;
      if (hipEventCreate(&b) != hipSuccess) return;
      t.pairs.push_back({a, b});
    }
    (void)hipEventRecord(t.pairs[t.used].first, ctx->stream);
    t.open = true;
  } else if (t.open) {
    (void)hipEventRecord(t.pairs[t.used].second, ctx->stream);
    t.used++;
    t.launches++;
    t.bytes += ctx->bytes_per_launch[cls];
    t.open = false;
  }
}

// Forget the last `count` launches of a class (iterations launched after the
// batch converged: they exit at once and move no data).
void track_discard(mfs_ctx *ctx, int cls, int count) {
  if (count <= 0 || cls <= 0 || cls >= 8 || !(ctx->track_mask & (1 << cls))) return;
  mfs_ctx::Track &t = ctx->tr[cls];
  const int k = std::min<int>(count, int(t.used));
  t.used -= size_t(k);
  t.launches -= k;
  t.bytes -= k * ctx->bytes_per_launch[cls];
}

void track_fold(mfs_ctx *ctx) {
  bool any = false;
  for (auto &t : ctx->tr) any = any || t.used > 0;
  if (!any) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &t : ctx->tr) {
    for (size_t i = 0; i < t.used; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, t.pairs[i].first, t.pairs[i].second) == hipSuccess) t.ms += ms;
    }
    t.used = 0;
  }
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void phase_begin(mfs_ctx *ctx) { (void)hipEventRecord(ctx->ev0, ctx->stream); }
double phase_end(mfs_ctx *ctx) {
  (void)hipEventRecord(ctx->ev1, ctx->stream);
  (void)hipEventSynchronize(ctx->ev1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  return ms;
}

// The band form's packing.  A supernode goes into a band when its tile count
// has a band capacity and its columns (original order) lie within kBandCols
// of its first column rounded down to even; bands are runs of kBandCols
// columns, each opened at the (even-rounded) first column of the supernodes
// not yet placed.  A band's supernodes are
// packed into waves of one tile count (band_capacity(T) each) and the waves
// into blocks of 8, heaviest first onto the least loaded block, the two waves
// sharing a SIMD (w and w + 4) one heavy, one light.  The other supernodes
// (spanning bands, or larger tiles) read a compact M^T of their columns.
int plan_bands(mfs_ctx *ctx, int32_t n) {
  const int32_t n_super = int32_t(ctx->sdesc.size());
  ctx->band_groups.clear();
  ctx->band_xcd.clear();
  ctx->band_items.clear();
  ctx->gin_items.clear();
  ctx->rest_items.clear();
  ctx->gin_buckets.clear();
  ctx->rest_buckets.clear();
  ctx->rest_cols.clear();
  ctx->goff.assign(size_t(n_super), -1);
  ctx->gram_elems = 0;
  ctx->band_supernodes = 0;
  ctx->rest_scratch = 8;
  std::vector<int32_t> cmin(n_super), cmax(n_super), tt(n_super), elig;
  for (int32_t s = 0; s < n_super; ++s) {
    const SuperDesc &d = ctx->sdesc[s];
    int32_t lo = n, hi = -1;
    for (int32_t p = 0; p < d.ns; ++p) {
      const int32_t c = ctx->Pinv[ctx->sind[d.rows_off + p]];
      lo = std::min(lo, c);
      hi = std::max(hi, c);
    }
    cmin[s] = lo;
    cmax[s] = hi;
    tt[s] = tiles_for(d.ns, ctx->plan_mask);
    // bands start at even columns (the band Gram loads column pairs, 8-byte aligned)
    if (band_capacity(tt[s]) > 0 && hi - (lo & ~1) < kBandCols) elig.push_back(s);
  }
  std::stable_sort(elig.begin(), elig.end(), [&](int32_t a, int32_t b) { return cmin[a] < cmin[b]; });
  std::vector<std::pair<int32_t, std::vector<int32_t>>> bands;
  for (int32_t s : elig) {
    if (bands.empty() || cmax[s] >= bands.back().first + kBandCols) bands.push_back({cmin[s] & ~1, {}});
    bands.back().second.push_back(s);
  }
  static constexpr int kBandTiles[] = {6, 4, 3, 2, 1};
  std::vector<std::vector<FactorItem>> gin_by_t(17);
  for (size_t bi = 0; bi < bands.size(); ++bi) {
    struct Wv {
      BandWave w;
      double work;
      int32_t lo, hi;  // its supernodes' columns: lo (even) .. hi
    };
    std::vector<Wv> waves;
    for (int T : kBandTiles) {
      std::vector<int32_t> v;
      for (int32_t s : bands[bi].second)
        if (tt[s] == T) v.push_back(s);
      const int C = band_capacity(T), P = T * (T + 1) / 2;
      for (size_t k = 0; k < v.size(); k += size_t(C)) {
        const int32_t cnt = int32_t(std::min(v.size() - k, size_t(C)));
        Wv wv{{T, cnt, int32_t(ctx->band_items.size())}, 0.0, n, -1};
        for (int32_t q = 0; q < cnt; ++q) {
          const int32_t sid = v[k + size_t(q)];
          wv.lo = std::min(wv.lo, cmin[sid] & ~1);
          wv.hi = std::max(wv.hi, cmax[sid]);
          ctx->band_items.push_back(sid);
          ctx->goff[sid] = ctx->gram_elems;
          ctx->gram_elems += int64_t(16 * T) * (16 * T);
          gin_by_t[T].push_back({0, sid});
        }
        // MFMA cycles (3 products of 16 per tile pair) + operand split per tile
        wv.work = double(cnt) * (48.0 * P + 32.0 * T);
        waves.push_back(wv);
      }
    }
    // Blocks: the waves in column order, kBandWaves consecutive ones each, so
    // that a block stages only the window of columns its supernodes read.  At
    // 64 ellipsoids x 1024 points this cut the columns staged from 82 to 61
    // GB against blocks balanced by work over the whole band, with no heavier
    // waves per block (the factor 17.4 -> 16.3 ms).
    std::stable_sort(waves.begin(), waves.end(), [](const Wv &a, const Wv &b) { return a.lo < b.lo; });
    for (size_t b0 = 0; b0 < waves.size(); b0 += size_t(kBandWaves)) {
      std::vector<Wv> blk(waves.begin() + std::ptrdiff_t(b0),
                          waves.begin() + std::ptrdiff_t(std::min(b0 + size_t(kBandWaves), waves.size())));
      int32_t lo = n, hi = -1;
      for (const Wv &w : blk) {
        lo = std::min(lo, w.lo);
        hi = std::max(hi, w.hi);
      }
      std::stable_sort(blk.begin(), blk.end(), [](const Wv &a, const Wv &b) { return a.work > b.work; });
      BandBlock B{};
      B.c0 = lo;
      B.part = 0;
      // hi - lo < kBandCols: every supernode of the band ends before the band's first column +
      // kBandCols, and lo is at or after it
      B.cols = std::min(kBandCols, (hi + 2 - lo) & ~1);
      for (int i = 0; i < kBandWaves; ++i) B.w[i] = BandWave{0, 0, 0};
      // heaviest first; SIMD i holds the i-th heaviest and the i-th lightest
      const int nw = int(blk.size());
      for (int i = 0; i < nw; ++i) {
        const int pos = i < kBandWaves / 2 ? i : kBandWaves - 1 - (i - kBandWaves / 2);
        B.w[pos] = blk[size_t(i)].w;
      }
      ctx->band_groups.push_back(B);
      ctx->band_xcd.push_back(int8_t(bi % 8));
    }
  }
  for (int T : kBandTiles) {
    if (gin_by_t[T].empty()) continue;
    ctx->gin_buckets.push_back({T, int32_t(ctx->gin_items.size()), int32_t(gin_by_t[T].size()), {}});
    ctx->gin_items.insert(ctx->gin_items.end(), gin_by_t[T].begin(), gin_by_t[T].end());
  }
  ctx->band_supernodes = int32_t(ctx->gin_items.size());
  // the rest, bucketed by tile count, and the columns they read
  std::vector<std::vector<FactorItem>> rest_by_t(17);
  std::vector<char> used(size_t(n), 0);
  for (int32_t s = 0; s < n_super; ++s) {
    if (ctx->goff[s] >= 0) continue;
    rest_by_t[tt[s]].push_back({0, s});
    const SuperDesc &d = ctx->sdesc[s];
    for (int32_t p = 0; p < d.ns; ++p) used[size_t(ctx->Pinv[ctx->sind[d.rows_off + p]])] = 1;
  }
  for (int t : kTileSizes) {
    if (rest_by_t[t].empty()) continue;
    ctx->rest_buckets.push_back({t, int32_t(ctx->rest_items.size()), int32_t(rest_by_t[t].size()), {}});
    ctx->rest_items.insert(ctx->rest_items.end(), rest_by_t[t].begin(), rest_by_t[t].end());
    ctx->rest_scratch = std::max(ctx->rest_scratch, rest_by_t[t].size() * size_t(factor_scratch_elems(t)) * 4);
  }
  std::vector<int32_t> colmap(size_t(n), -1);
  for (int32_t c = 0; c < n; ++c)
    if (used[size_t(c)]) {
      colmap[size_t(c)] = int32_t(ctx->rest_cols.size());
      ctx->rest_cols.push_back(c);
    }
  HIPCHK(upload(ctx, ctx->dbitems, ctx->band_items), "upload");
  HIPCHK(upload(ctx, ctx->dgoff, ctx->goff), "upload");
  HIPCHK(upload(ctx, ctx->dginlist, ctx->gin_items), "upload");
  HIPCHK(upload(ctx, ctx->drestlist, ctx->rest_items), "upload");
  HIPCHK(upload(ctx, ctx->drestcols, ctx->rest_cols), "upload");
  HIPCHK(upload(ctx, ctx->dcolmap, colmap), "upload");
  return MFS_OK;
}

int plan_impl(mfs_ctx *ctx, std::vector<double> &&xyz, int64_t n64, double rho, double lambda,
              int32_t max_super) {
  if (n64 <= 0 || n64 > (int64_t(1) << 30)) return ctx->fail(MFS_ERR_ARG, "plan: bad n");
  if (!(rho > 0) || !(lambda > 0) || max_super <= 0)
    return ctx->fail(MFS_ERR_ARG, "plan: rho, lambda and max_super must be positive");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  const int32_t n = int32_t(n64);
  ctx->has_plan = false;
  ctx->has_factor = false;
  ctx->has_points = false;
  phase_begin(ctx);

  // 1. ordering (host)
  const auto t_order = std::chrono::steady_clock::now();
  int rc = order_maximin(n, xyz.data(), ctx->P, ctx->Pinv, ctx->scale);
  ctx->stats.plan_order_ms = ms_since(t_order);
  const auto t_pattern = std::chrono::steady_clock::now();
  if (rc == PLAN_ERR_DUPLICATE)
    return ctx->fail(MFS_ERR_DUPLICATE, "plan: coincident source points (zero length scale)");
  if (rc != PLAN_OK) return ctx->fail(MFS_ERR_ARG, "plan: ordering failed");

  // 2. rho-radius pattern (device); rho rounded through float as the op does.
  ctx->rho32 = double(float(rho));
  std::vector<double> px(n), py(n), pz(n);
  for (int32_t c = 0; c < n; ++c) {
    const int64_t o = ctx->Pinv[c];
    px[c] = xyz[3 * o];
    py[c] = xyz[3 * o + 1];
    pz[c] = xyz[3 * o + 2];
  }
  HIPCHK(upload(ctx, ctx->dpx, px), "upload points");
  HIPCHK(upload(ctx, ctx->dpy, py), "upload points");
  HIPCHK(upload(ctx, ctx->dpz, pz), "upload points");
  HIPCHK(upload(ctx, ctx->dscale, ctx->scale), "upload scale");
  HIPCHK(ctx->dcnt.ensure(size_t(n) * 4), "alloc");
  HIPCHK(ctx->dptr.ensure(size_t(n + 1) * 4), "alloc");
  HIPCHK(ctx->dtotal.ensure(8), "alloc");
  // grids for the neighbour search (host), then count / scan / fill (device)
  std::vector<GridBand> bands;
  std::vector<int32_t> cell_start, cell_pts;
  int32_t j_brute = n;
  build_grid_bands(n, px.data(), py.data(), pz.data(), ctx->scale.data(), ctx->rho32, bands,
                   cell_start, cell_pts, &j_brute);
  HIPCHK(upload(ctx, ctx->dbands, bands), "upload grid");
  HIPCHK(upload(ctx, ctx->dcells, cell_start), "upload grid");
  HIPCHK(upload(ctx, ctx->dcpts, cell_pts), "upload grid");
  const double *dx_ = ctx->dpx.as<double>(), *dy_ = ctx->dpy.as<double>(), *dz_ = ctx->dpz.as<double>();
  const double *dsc = ctx->dscale.as<double>();
  HIPCHK(launch_grid_pattern(ctx->stream, j_brute, ctx->dbands.as<GridBand>(), int32_t(bands.size()),
                             ctx->dcells.as<int32_t>(), ctx->dcpts.as<int32_t>(), dx_, dy_, dz_, dsc,
                             ctx->rho32, ctx->dcnt.as<int32_t>(), nullptr, nullptr, 0),
         "pattern count (grid)");
  HIPCHK(launch_pattern(ctx->stream, n, j_brute, nullptr, n - j_brute, dx_, dy_, dz_, dsc, ctx->rho32,
                        ctx->dcnt.as<int32_t>(), nullptr, nullptr, 0),
         "pattern count (scan)");
  HIPCHK(launch_scan(ctx->stream, ctx->dcnt.as<int32_t>(), n, ctx->dptr.as<int32_t>(),
                     ctx->dtotal.as<int64_t>()),
         "pattern scan");
  int64_t total = 0;
  ctx->ptr.resize(size_t(n) + 1);
  HIPCHK(hipMemcpyAsync(&total, ctx->dtotal.p, 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
  HIPCHK(hipMemcpyAsync(ctx->ptr.data(), ctx->dptr.p, size_t(n + 1) * 4, hipMemcpyDeviceToHost,
                        ctx->stream),
         "copy pattern");
  HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
  if (total > INT32_MAX) return ctx->fail(MFS_ERR_ARG, "plan: pattern exceeds 2^31 entries");
  // grid columns too long for the in-LDS sort are refilled, in order, by the scan
  std::vector<int32_t> long_cols;
  for (int32_t j = 0; j < j_brute; ++j)
    if (ctx->ptr[j + 1] - ctx->ptr[j] > pattern_sort_max()) long_cols.push_back(j);
  HIPCHK(upload(ctx, ctx->dlong, long_cols), "upload");
  HIPCHK(ctx->dind.ensure(size_t(std::max<int64_t>(total, 1)) * 4), "alloc pattern");
  HIPCHK(launch_grid_pattern(ctx->stream, j_brute, ctx->dbands.as<GridBand>(), int32_t(bands.size()),
                             ctx->dcells.as<int32_t>(), ctx->dcpts.as<int32_t>(), dx_, dy_, dz_, dsc,
                             ctx->rho32, nullptr, ctx->dptr.as<int32_t>(), ctx->dind.as<int32_t>(), 1),
         "pattern fill (grid)");
  HIPCHK(launch_pattern(ctx->stream, n, j_brute, nullptr, n - j_brute, dx_, dy_, dz_, dsc, ctx->rho32,
                        nullptr, ctx->dptr.as<int32_t>(), ctx->dind.as<int32_t>(), 1),
         "pattern fill (scan)");
  HIPCHK(launch_pattern(ctx->stream, n, 0, ctx->dlong.as<int32_t>(), int32_t(long_cols.size()), dx_,
                        dy_, dz_, dsc, ctx->rho32, nullptr, ctx->dptr.as<int32_t>(),
                        ctx->dind.as<int32_t>(), 1),
         "pattern fill (long columns)");
  ctx->ind.resize(size_t(total));
  if (total > 0)
    HIPCHK(hipMemcpyAsync(ctx->ind.data(), ctx->dind.p, size_t(total) * 4, hipMemcpyDeviceToHost,
                          ctx->stream),
           "copy pattern");
  HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
  ctx->stats.plan_grid_columns = j_brute;
  ctx->stats.plan_pattern_ms = ms_since(t_pattern);

  // 3. supernodes and supernodal pattern (host)
  rc = aggregate(n, ctx->ptr.data(), ctx->ind.data(), ctx->scale.data(), lambda, max_super,
                 ctx->sup_ptr, ctx->sup_ind, ctx->sup_parent);
  if (rc != PLAN_OK)
    return ctx->fail(MFS_ERR_DUPLICATE, "plan: a column lacks its diagonal (zero length scale)");
  const int32_t n_super = int32_t(ctx->sup_ptr.size()) - 1;
  super_pattern(n, ctx->ptr.data(), ctx->ind.data(), ctx->sup_parent, n_super, ctx->sptr, ctx->sind);
  transpose_pattern(n, ctx->sptr, ctx->sind, ctx->tptr, ctx->tcol, ctx->tmap);

  // 4. supernode descriptors, bucketed by tile count (largest first in each)
  ctx->plan_mask = ctx->tile_mask;
  ctx->plan_phases = ctx->phases;
  ctx->sdesc.resize(n_super);
  ctx->max_local = 0;
  double s2 = 0, s3 = 0;
  int64_t theta = 0;
  std::vector<std::vector<int32_t>> by_t(17);
  for (int32_t s = 0; s < n_super; ++s) {
    SuperDesc d{};
    d.first = ctx->sup_ind[ctx->sup_ptr[s]];
    d.ns = ctx->sptr[d.first + 1] - ctx->sptr[d.first];
    d.rows_off = ctx->sptr[d.first];
    d.mem_beg = ctx->sup_ptr[s];
    d.mem_end = ctx->sup_ptr[s + 1];
    ctx->sdesc[s] = d;
    ctx->max_local = std::max(ctx->max_local, d.ns);
    s2 += double(d.ns) * d.ns;
    s3 += double(d.ns) * d.ns * d.ns;
    theta += int64_t(d.ns) * d.ns;
    const int t = tiles_for(d.ns, ctx->plan_mask);
    if (t < 0)
      return ctx->fail(MFS_ERR_ARG, "plan: supernode local size " + std::to_string(d.ns) +
                                        " exceeds the supported maximum of " +
                                        std::to_string(kMaxLocal));
    by_t[t].push_back(s);
  }
  // Launch order within a bucket: supernodes sorted along a Morton curve of
  // their opener's position, then dealt so that block b runs the (b / 8)-th
  // supernode of the b % 8-th contiguous slice.  Workgroups are dispatched to
  // the 8 XCDs round-robin, so each XCD walks one spatially compact slice and
  // blocks running together on it gather overlapping rows of M^T through the
  // same L2.
  {
    double lo[3] = {px[0], py[0], pz[0]}, hi[3] = {px[0], py[0], pz[0]};
    for (int32_t c = 1; c < n; ++c) {
      lo[0] = std::min(lo[0], px[c]); hi[0] = std::max(hi[0], px[c]);
      lo[1] = std::min(lo[1], py[c]); hi[1] = std::max(hi[1], py[c]);
      lo[2] = std::min(lo[2], pz[c]); hi[2] = std::max(hi[2], pz[c]);
    }
    auto spread = [](uint64_t v) {  // 21 bits -> every third bit
      v &= 0x1fffff;
      v = (v | v << 32) & 0x1f00000000ffffULL;
      v = (v | v << 16) & 0x1f0000ff0000ffULL;
      v = (v | v << 8) & 0x100f00f00f00f00fULL;
      v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
      v = (v | v << 2) & 0x1249249249249249ULL;
      return v;
    };
    std::vector<uint64_t> key(n_super);
    for (int32_t s = 0; s < n_super; ++s) {
      const int32_t c = ctx->sdesc[s].first;
      const double v[3] = {px[c], py[c], pz[c]};
      uint64_t k = 0;
      for (int d = 0; d < 3; ++d) {
        const double w = hi[d] > lo[d] ? (v[d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
        k |= spread(uint64_t(std::min(w, 1.0) * 2097151.0)) << d;
      }
      key[s] = k;
    }
    // column phase of each supernode (see phase_col)
    std::vector<int8_t> phase(n_super);
    for (int32_t s = 0; s < n_super; ++s) {
      const SuperDesc &d = ctx->sdesc[s];
      int32_t mc = 0;
      for (int32_t p = 0; p < d.ns; ++p) mc = std::max(mc, ctx->Pinv[ctx->sind[d.rows_off + p]]);
      int ph = 0;
      while (ph + 1 < ctx->plan_phases && mc >= phase_col(ph + 1, n, ctx->plan_phases)) ++ph;
      phase[s] = int8_t(ph);
    }
    ctx->list.clear();
    ctx->buckets.clear();
    constexpr int kXcd = 8;
    for (int t : kTileSizes) {
      std::vector<int32_t> &v = by_t[t];
      if (v.empty()) continue;
      std::stable_sort(v.begin(), v.end(), [&](int32_t a, int32_t b) {
        return phase[a] != phase[b] ? phase[a] < phase[b] : key[a] < key[b];
      });
      Bucket b{t, int32_t(ctx->list.size()), int32_t(v.size()), {}};
      int32_t lo = 0;
      for (int ph = 0; ph < ctx->plan_phases; ++ph) {
        int32_t hi = lo;
        while (hi < int32_t(v.size()) && phase[v[hi]] == ph) ++hi;
        b.phase_off[ph] = lo;
        const int32_t cnt = hi - lo;
        const int32_t per = (cnt + kXcd - 1) / kXcd;
        for (int32_t i = 0; i < per; ++i)
          for (int x = 0; x < kXcd; ++x) {
            const int32_t src = x * per + i;
            if (src < cnt) ctx->list.push_back(v[lo + src]);
          }
        lo = hi;
      }
      for (int ph = ctx->plan_phases; ph <= kMaxPhases; ++ph) b.phase_off[ph] = lo;
      ctx->buckets.push_back(b);
    }
  }

  // 5. upload
  HIPCHK(upload(ctx, ctx->dP, ctx->P), "upload");
  HIPCHK(upload(ctx, ctx->dPinv, ctx->Pinv), "upload");
  HIPCHK(upload(ctx, ctx->dsptr, ctx->sptr), "upload");
  HIPCHK(upload(ctx, ctx->dsind, ctx->sind), "upload");
  HIPCHK(upload(ctx, ctx->dtptr, ctx->tptr), "upload");
  HIPCHK(upload(ctx, ctx->dtcol, ctx->tcol), "upload");
  HIPCHK(upload(ctx, ctx->dtmap, ctx->tmap), "upload");
  HIPCHK(upload(ctx, ctx->dsupind, ctx->sup_ind), "upload");
  HIPCHK(upload(ctx, ctx->dsd, ctx->sdesc), "upload");
  ctx->items.resize(ctx->list.size());
  for (size_t i = 0; i < ctx->items.size(); ++i) ctx->items[i] = {0, ctx->list[i]};
  HIPCHK(upload(ctx, ctx->dlist, ctx->items), "upload");
  HIPCHK(ctx->dstatus.ensure(size_t(n_super) * 4), "alloc");
  {
    const int rc = plan_bands(ctx, n);
    if (rc != MFS_OK) return rc;
  }

  ctx->n = n;
  ctx->has_plan = true;
  ctx->stats.n = n;
  ctx->stats.n_super = n_super;
  ctx->stats.max_local = ctx->max_local;
  ctx->stats.nnz_pattern = int64_t(ctx->ptr[n]);
  ctx->stats.nnz_factor = int64_t(ctx->sptr[n]);
  ctx->stats.theta_entries = theta;
  ctx->stats.sum_ns2 = s2;
  ctx->stats.sum_ns3 = s3;
  ctx->stats.plan_ms = phase_end(ctx);
  return MFS_OK;
}

// Factor values in the precision a solver needs (converted copy if required).
__global__ void k_convert_d2f(int64_t count, const double *a, float *b) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) b[i] = float(a[i]);
}
__global__ void k_convert_f2d(int64_t count, const float *a, double *b) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) b[i] = double(a[i]);
}

int values_for(mfs_ctx *ctx, int want, const void **lval, const void **ltval) {
  if (ctx->prec == want) {
    *lval = ctx->dval.p;
    *ltval = ctx->dtval.p;
    return MFS_OK;
  }
  const int64_t nnz = ctx->sptr[ctx->n];
  const size_t w = (want == MFS_FP64) ? 8 : 4;
  if (ctx->alt_prec != want) {
    HIPCHK(ctx->dval_alt.ensure(size_t(std::max<int64_t>(nnz, 1)) * w), "alloc");
    HIPCHK(ctx->dtval_alt.ensure(size_t(std::max<int64_t>(nnz, 1)) * w), "alloc");
    const unsigned g = unsigned((nnz + kThreads - 1) / kThreads);
    if (nnz > 0) {
      if (want == MFS_FP32) {
        k_convert_d2f<<<g, kThreads, 0, ctx->stream>>>(nnz, ctx->dval.as<double>(), ctx->dval_alt.as<float>());
        k_convert_d2f<<<g, kThreads, 0, ctx->stream>>>(nnz, ctx->dtval.as<double>(), ctx->dtval_alt.as<float>());
      } else {
        k_convert_f2d<<<g, kThreads, 0, ctx->stream>>>(nnz, ctx->dval.as<float>(), ctx->dval_alt.as<double>());
        k_convert_f2d<<<g, kThreads, 0, ctx->stream>>>(nnz, ctx->dtval.as<float>(), ctx->dtval_alt.as<double>());
      }
      HIPCHK(hipGetLastError(), "convert factor");
    }
    ctx->alt_prec = want;
  }
  *lval = ctx->dval_alt.p;
  *ltval = ctx->dtval_alt.p;
  return MFS_OK;
}

// Row blocks of the M^T pass: about `target` blocks in all (default: one
// full round, every CU holding as many as fit), at least 64 rows each.  At
// m = n = 65,536 (FP32, 8 columns, 3 resident blocks per CU) one round of
// 768 blocks reads M at 6.74 TB/s where 2048 (2.7 rounds) reach 6.31 TB/s and
// 1024 (1.3 rounds) 5.6 TB/s: a partial last round leaves CUs idle.
int32_t row_blocks_for(int64_t m, int64_t n, int32_t target, int64_t per_round) {
  const int64_t cols = (n + 1023) / 1024;
  const int64_t want = target > 0 ? target : std::max<int64_t>(per_round, 1);
  int64_t rb = (want + cols - 1) / cols;
  rb = std::min<int64_t>(rb, std::max<int64_t>(m / 64, 1));
  return int32_t(std::max<int64_t>(rb, 1));
}

int batch_width(int64_t k) {
  if (k <= 1) return 1;
  if (k <= 2) return 2;
  if (k <= 4) return 4;
  if (k <= 8) return 8;
  return 16;
}

// One pass over M per FP32 iteration for dense batches of 8 when the shape
// qualifies: by default once M holds kFusedMinElems entries or more (below
// that the two passes are faster: at 8192^2 0.113 ms against 0.118 ms, at
// 16384^2 0.361 ms against 0.287 ms), always with MFS_PARAM_FUSED = 2, never
// with 1.  Sets up the exchange records and the partials of every row group
// of the grid (part is re-pointed if the buffer grows).
constexpr double kFusedMinElems = double(int64_t(1) << 27);

int setup_fused(mfs_ctx *ctx, const PointsOp *pts, int64_t m, int64_t n, int K, FusedShape &shape,
                uint64_t *&records, double *&part) {
  shape = FusedShape{};
  records = nullptr;
  if (pts != nullptr || ctx->fused_mode == 1) return MFS_OK;
  if (ctx->fused_mode == 0 && double(m) * double(n) < kFusedMinElems) return MFS_OK;
  const FusedShape f = fused_shape(m, n, K, ctx->cus);
  if (f.slices <= 0) return MFS_OK;
  const size_t bytes = fused_records(f) * sizeof(uint64_t);
  const FusedShape &o = ctx->frec_shape;
  const bool same = o.slices == f.slices && o.groups == f.groups && o.width == f.width &&
                    o.panels == f.panels;
  if (ctx->dfrec.cap < bytes || ctx->dfrec.p == nullptr || !same) {
    HIPCHK(ctx->dfrec.ensure(bytes), "alloc exchange records");
    // A record's tag holds 16 bits of the launch count: launches of one grid
    // rewrite every record they read, but a record left by another grid (or an
    // earlier owner of this memory) could match a tag 65,536 launches later.
    HIPCHK(hipMemsetAsync(ctx->dfrec.p, 0, ctx->dfrec.cap, ctx->stream), "clear records");
    ctx->frec_shape = f;
  }
  HIPCHK(ctx->dpart.ensure(size_t(f.groups) * size_t(n) * size_t(K) * sizeof(double)),
         "alloc partials");
  part = ctx->dpart.as<double>();
  shape = f;
  records = ctx->dfrec.as<uint64_t>();
  return MFS_OK;
}

// Lay out a SolveArgs workspace for batch width K.
template <typename TA>
int workspace(mfs_ctx *ctx, const float *M, const PointsOp *pts, int64_t m, int64_t n, int K,
              SolveArgs<TA> &a) {
  const int want = sizeof(TA) == 8 ? MFS_FP64 : MFS_FP32;
  const void *lv = nullptr, *ltv = nullptr;
  int rc = values_for(ctx, want, &lv, &ltv);
  if (rc != MFS_OK) return rc;
  a.M = M;
  a.m = m;
  a.n = n;
  a.K = K;
  a.P = ctx->dP.as<int32_t>();
  a.Pinv = ctx->dPinv.as<int32_t>();
  a.sptr = ctx->dsptr.as<int32_t>();
  a.sind = ctx->dsind.as<int32_t>();
  a.tptr = ctx->dtptr.as<int32_t>();
  a.tcol = ctx->dtcol.as<int32_t>();
  a.lval = static_cast<const TA *>(lv);
  a.ltval = static_cast<const TA *>(ltv);
  a.row_blocks = pts ? points_splits(n, m, pts->target_blocks)
                     : row_blocks_for(m, n, ctx->cols_blocks,
                                      int64_t(ctx->cus) * cols_resident_blocks(sizeof(TA) == 8, K, n % 4 == 0));
  a.pts = pts ? *pts : PointsOp{};
  const size_t nv = size_t(n) * K, mv = size_t(m) * K;
  HIPCHK(ctx->dvec.ensure((7 * nv + mv) * sizeof(TA) + 64), "alloc vectors");
  TA *base = ctx->dvec.as<TA>();
  a.x = base;
  a.r = base + nv;
  a.p = base + 2 * nv;
  a.q = base + 3 * nv;
  a.z = base + 4 * nv;
  a.t = base + 5 * nv;
  a.pv = base + 6 * nv;
  a.y = base + 7 * nv;
  size_t part_elems = size_t(a.row_blocks) * nv;
  if (pts) part_elems = std::max(part_elems, size_t(points_splits(m, n, pts->target_blocks)) * mv);
  HIPCHK(ctx->dpart.ensure(part_elems * sizeof(double)), "alloc partials");
  a.part = ctx->dpart.as<double>();
  const size_t maxblocks = size_t(spmv_grid(n, K, true)) + size_t((m + kThreads - 1) / kThreads) + 16;
  HIPCHK(ctx->ddpart.ensure(maxblocks * K * sizeof(double)), "alloc dot partials");
  a.dpart = ctx->ddpart.as<double>();
  if (ctx->dstate.p == nullptr) {
    HIPCHK(ctx->dstate.ensure(sizeof(PcgState)), "alloc state");
    HIPCHK(hipMemsetAsync(ctx->dstate.p, 0, sizeof(PcgState), ctx->stream), "memset state");
  }
  a.st = ctx->dstate.as<PcgState>();
  a.fused = FusedShape{};
  a.frec = nullptr;
  if (sizeof(TA) == 4) return setup_fused(ctx, pts, m, n, K, a.fused, a.frec, a.part);
  return MFS_OK;
}

// Run PCG iterations until every column is inactive (polls the device flag).
template <typename TA>
int iterate(mfs_ctx *ctx, const SolveArgs<TA> &a, int32_t maxit) {
  int chunk = 2;
  int32_t launched = 0;
  for (;;) {
    HIPCHK(hipMemcpyAsync(ctx->hstate, a.st, sizeof(PcgState), hipMemcpyDeviceToHost, ctx->stream),
           "poll state");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    if (ctx->hstate->n_active == 0) break;
    if (launched > maxit + 64) return ctx->fail(MFS_ERR_STATE, "pcg: iteration bookkeeping diverged");
    for (int i = 0; i < chunk; ++i) {
      HIPCHK(pcg_iteration<TA>(ctx->stream, a, track_cb, ctx), "pcg iteration");
      ++launched;
    }
    chunk = std::min(chunk * 2, 16);
  }
  const int dead = launched - ctx->hstate->iter;
  ctx->stats.fused_stale_polls = int32_t(ctx->hstate->fstale);
  track_discard(ctx, MFS_K_MATVEC_ROWS, dead);
  track_discard(ctx, MFS_K_MATVEC_COLS, dead);
  track_discard(ctx, MFS_K_MATVEC_FUSED, dead);
  if (ctx->hstate->fabort != 0u) {
    const uint32_t zero = 0;
    (void)hipMemcpyAsync(&a.st->fabort, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream);
    (void)hipStreamSynchronize(ctx->stream);
    return ctx->fail(MFS_ERR_HIP, "one-pass matvec: a wait for another block timed out");
  }
  return MFS_OK;
}

void record_batch_stats(mfs_ctx *ctx, int64_t col0, int kval) {
  for (int r = 0; r < kval; ++r) {
    const int64_t c = col0 + r;
    if (c >= 64) break;
    ctx->stats.iters[c] = ctx->hstate->iters[r];
    const double rhs2 = ctx->hstate->rhs2[r];
    ctx->stats.relres[c] = rhs2 > 0 ? std::sqrt(ctx->hstate->dot_rr[r] / rhs2) : 0.0;
  }
}

__global__ void k_x64_from_f32(int64_t n, const float *X, int64_t ldx, int64_t col0, int kval,
                               int K, double *X64) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < K; ++r) X64[i * K + r] = (r < kval && X) ? double(X[i * ldx + col0 + r]) : 0.0;
}

// pv (n x K, zero beyond k) = X (n x k, row-major), FP32.
__global__ void k_pack_rhs(int64_t n, const float *X, int64_t k, int K, float *pv) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n * K) return;
  const int64_t row = i / K, r = i % K;
  pv[i] = r < k ? X[row * k + r] : 0.f;
}

// Columns r with bit r of `mask` set: X64 = backup (a sweep that made them worse).
__global__ void k_restore_cols(int64_t n, int K, uint32_t mask, const double *backup, double *X64) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < K; ++r)
    if (mask & (1u << r)) X64[i * K + r] = backup[i * K + r];
}

// X (n x Kx) = [A | V] padded with zeros (A, V: n x k); with `which` = 1 only
// A, 2 only V (each padded to Kx).
__global__ void k_pack_av(int64_t n, const float *A, const float *V, int64_t k, int Kx, int which,
                          float *X) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n * Kx) return;
  const int64_t j = i / Kx, r = i % Kx;
  float v = 0.f;
  if (which == 0) v = r < k ? A[j * k + r] : (r < 2 * k ? V[j * k + r - k] : 0.f);
  else v = r < k ? (which == 1 ? A : V)[j * k + r] : 0.f;
  X[i] = v;
}

// The backward's factors: U (m x Ku) = [F - MA | -MV], W (n x Ku) = [V | A]
// (zero padded), grad_F = MV when asked.  MA, MV: columns ca / cv of Y
// (m x Ky).
__global__ void k_grad_uw(int64_t m, int64_t n, int64_t k, const float *F, const float *Y, int Ky,
                          int ca, int cv, const float *YV, int Kyv, const float *A, const float *V,
                          int Ku, float *U, float *W, float *grad_F) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < m) {
    for (int r = 0; r < Ku; ++r) {
      float u = 0.f;
      if (r < k) u = F[i * k + r] - Y[i * Ky + ca + r];
      else if (r < 2 * k) u = -YV[i * Kyv + cv + r - k];
      U[i * Ku + r] = u;
    }
    if (grad_F != nullptr)
      for (int r = 0; r < k; ++r) grad_F[i * k + r] = YV[i * Kyv + cv + r];
  }
  if (i < n) {
    for (int r = 0; r < Ku; ++r)
      W[i * Ku + r] = r < k ? V[i * k + r] : (r < 2 * k ? A[i * k + r - k] : 0.f);
  }
}

// FP64 strided columns <-> the K-padded FP64 iterate (FP64 warm start / output).
__global__ void k_x64_from_f64(int64_t n, const double *X, int64_t ldx, int64_t col0, int kval,
                               int K, double *X64) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < K; ++r) X64[i * K + r] = (r < kval && X) ? X[i * ldx + col0 + r] : 0.0;
}

__global__ void k_x64_to_f64(int64_t n, const double *X64, int K, double *X, int64_t ldx,
                             int64_t col0, int kval) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < kval; ++r) X[i * ldx + col0 + r] = X64[i * K + r];
}

__global__ void k_x64_to_f32(int64_t n, const double *X64, int K, float *X, int64_t ldx,
                             int64_t col0, int kval) {
  const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  for (int r = 0; r < kval; ++r) X[i * ldx + col0 + r] = float(X64[i * K + r]);
}

int check_dense_args(mfs_ctx *ctx, const float *M, int64_t m, int64_t n) {
  if (!ctx->has_plan) return ctx->fail(MFS_ERR_STATE, "no plan: call mfs_plan first");
  if (M == nullptr || m <= 0 || n != ctx->n)
    return ctx->fail(MFS_ERR_ARG, "M must be m x n with n equal to the planned source count");
  return MFS_OK;
}

constexpr int kEvalFlops = 17;  // one kernel evaluation, rsqrt counted as 1

int ensure_side_streams(mfs_ctx *ctx) {
  if (ctx->fork_ev != nullptr) return MFS_OK;
  for (int i = 0; i < mfs_ctx::kSide; ++i) {
    HIPCHK(hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking), "create stream");
    HIPCHK(hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming), "create event");
  }
  for (int i = 0; i < kMaxPhases; ++i)
    HIPCHK(hipEventCreateWithFlags(&ctx->tr_ev[i], hipEventDisableTiming), "create event");
  HIPCHK(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming), "create event");
  return MFS_OK;
}

int check_points(mfs_ctx *ctx) {
  if (!ctx->has_plan) return ctx->fail(MFS_ERR_STATE, "no plan: call mfs_plan first");
  if (!ctx->has_points)
    return ctx->fail(MFS_ERR_STATE, "no points: call mfs_factor_points first");
  return MFS_OK;
}

// Algorithmic work per tracked launch of the solve passes: bytes of M read
// (dense form) or kernel FLOPs, 17 per evaluation plus 2 per right-hand side
// (points form).
void set_pass_work(mfs_ctx *ctx, const PointsOp *pts, int64_t m, int64_t n, int K = 1) {
  for (int c = 1; c < 8; ++c) {
    if (pts == nullptr) {
      ctx->bytes_per_launch[c] = double(m) * double(n) * 4.0;
    } else {
      ctx->bytes_per_launch[c] = double(m) * double(n) * double(kEvalFlops + 2 * K);
    }
  }
}

// Factor phase 1: the M^T buffer (dense form), value buffers, cleared
// statuses, and the scene record the factor kernels read (ctx->fscene,
// ctx->dfscene).
int factor_prepare(mfs_ctx *ctx, const float *M, const PointsOp *pts, int64_t m, int64_t n,
                   int precision, int parts = 0) {
  if (precision != MFS_FP64 && precision != MFS_FP32)
    return ctx->fail(MFS_ERR_ARG, "factor: precision must be MFS_FP64 or MFS_FP32");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  ctx->has_factor = false;
  ctx->alt_prec = -1;
  const int64_t ldmt = (m + kGramChunk - 1) / kGramChunk * kGramChunk;
  // Dense form: M^T of every column (filled by factor_impl, range by range),
  // or in the band form (parts > 0) a compact M^T of the rest's columns, and
  // the band partial Grams.  The row after the last stays zero.
  float *mt = nullptr;
  int64_t mt_rows = 0;
  if (pts == nullptr) {
    DevBuf &buf = parts > 0 ? ctx->dMtc : ctx->dMt;
    mt_rows = parts > 0 ? int64_t(ctx->rest_cols.size()) : n;
    HIPCHK(buf.ensure(size_t(mt_rows + 1) * size_t(ldmt) * sizeof(float)), "alloc M^T");
    mt = buf.as<float>();
    HIPCHK(hipMemsetAsync(mt + size_t(mt_rows) * size_t(ldmt), 0, size_t(ldmt) * sizeof(float),
                          ctx->stream),
           "clear zero row");
    if (parts > 0)
      HIPCHK(ctx->dgram.ensure(size_t(std::max<int64_t>(ctx->gram_elems, 1)) * size_t(parts) * sizeof(float)),
             "alloc band Grams");
  }
  const int64_t nnz = ctx->sptr[ctx->n];
  const size_t w = (precision == MFS_FP64) ? 8 : 4;
  HIPCHK(ctx->dval.ensure(size_t(std::max<int64_t>(nnz, 1)) * w), "alloc factor");
  HIPCHK(ctx->dtval.ensure(size_t(std::max<int64_t>(nnz, 1)) * w), "alloc factor");
  const int32_t n_super = int32_t(ctx->sdesc.size());
  HIPCHK(hipMemsetAsync(ctx->dstatus.p, 0, size_t(n_super) * 4, ctx->stream), "memset");
  FactorScene &F = ctx->fscene;
  F.Mt = mt;
  F.Mz = pts ? nullptr : mt + size_t(mt_rows) * size_t(ldmt);
  F.ldmt = ldmt;
  F.colmap = parts > 0 ? ctx->dcolmap.as<int32_t>() : nullptr;
  F.gram = parts > 0 ? ctx->dgram.as<float>() : nullptr;
  F.goff = parts > 0 ? ctx->dgoff.as<int64_t>() : nullptr;
  F.gram_part = parts > 0 ? ctx->gram_elems : 0;
  F.parts = parts;
  F.tgt = pts ? pts->tgt32 : nullptr;
  F.src = pts ? pts->src32 : nullptr;
  F.m = m;
  F.sd = ctx->dsd.as<SuperDesc>();
  F.sind = ctx->dsind.as<int32_t>();
  F.Pinv = ctx->dPinv.as<int32_t>();
  F.sup_ind = ctx->dsupind.as<int32_t>();
  F.sptr = ctx->dsptr.as<int32_t>();
  F.val = ctx->dval.p;
  F.status = ctx->dstatus.as<int32_t>();
  HIPCHK(ctx->dfscene.ensure(sizeof(FactorScene)), "alloc");
  HIPCHK(hipMemcpyAsync(ctx->dfscene.p, &ctx->fscene, sizeof(FactorScene), hipMemcpyHostToDevice,
                        ctx->stream),
         "upload scene");
  double sum_ns = 0;
  for (auto &d : ctx->sdesc) sum_ns += d.ns;
  // Work per launch: bytes of the gathered columns (dense), kernel FLOPs (points).
  ctx->bytes_per_launch[MFS_K_GRAM] = double(m) * sum_ns * (pts ? double(kEvalFlops) : 4.0);
  return MFS_OK;
}

// Factor phase 3 (after the bucket launches): statuses, FP64 retry of failed
// FP32 pivots (values stay FP32), transposed copy of the values.
int factor_finish(mfs_ctx *ctx, bool points, int precision, const float *M = nullptr,
                  int64_t m = 0, int64_t n = 0) {
  const int32_t n_super = int32_t(ctx->sdesc.size());
  const int64_t nnz = ctx->sptr[ctx->n];
  std::vector<int32_t> status(n_super);
  HIPCHK(hipMemcpyAsync(status.data(), ctx->dstatus.p, size_t(n_super) * 4, hipMemcpyDeviceToHost,
                        ctx->stream),
         "copy status");
  HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
  std::vector<int32_t> failed;
  for (int32_t s = 0; s < n_super; ++s)
    if (status[s] != 0) failed.push_back(s);
  ctx->stats.retried_supernodes = 0;
  if (!failed.empty() && precision == MFS_FP32) {
    ctx->stats.retried_supernodes = int32_t(failed.size());
    if (!points && ctx->fscene.parts > 0) {
      // the band form built no M^T of the band columns: the FP64 retry reads
      // M^T of every column
      FactorScene &F = ctx->fscene;
      const int64_t ldmt = F.ldmt;
      HIPCHK(ctx->dMt.ensure(size_t(n + 1) * size_t(ldmt) * sizeof(float)), "alloc M^T");
      HIPCHK(hipMemsetAsync(ctx->dMt.as<float>() + size_t(n) * size_t(ldmt), 0,
                            size_t(ldmt) * sizeof(float), ctx->stream),
             "clear zero row");
      HIPCHK(launch_transpose(ctx->stream, M, m, n, ctx->dMt.as<float>(), ldmt, 0, n), "transpose");
      F.Mt = ctx->dMt.as<float>();
      F.Mz = F.Mt + size_t(n) * size_t(ldmt);
      F.colmap = nullptr;
      F.gram = nullptr;
      F.goff = nullptr;
      F.gram_part = 0;
      F.parts = 0;
      HIPCHK(hipMemcpyAsync(ctx->dfscene.p, &F, sizeof(FactorScene), hipMemcpyHostToDevice, ctx->stream),
             "upload scene");
    }
    std::vector<std::vector<FactorItem>> by_t(17);
    for (int32_t s : failed) by_t[tiles_for(ctx->sdesc[s].ns, ctx->plan_mask)].push_back({0, s});
    std::vector<FactorItem> flist;
    std::vector<Bucket> fb;
    size_t scratch_bytes = 8;
    for (int t : kTileSizes) {
      if (by_t[t].empty()) continue;
      fb.push_back({t, int32_t(flist.size()), int32_t(by_t[t].size())});
      flist.insert(flist.end(), by_t[t].begin(), by_t[t].end());
      scratch_bytes = std::max(scratch_bytes, by_t[t].size() * size_t(factor_scratch_elems(t)) * 8);
    }
    DevBuf dfl;
    HIPCHK(upload(ctx, dfl, flist), "upload retry list");
    HIPCHK(ctx->dthg.ensure(scratch_bytes), "alloc factor scratch");
    for (const Bucket &b : fb)
      HIPCHK(launch_super_factor(ctx->stream, b.tiles, 2, points, ctx->dfscene.as<FactorScene>(),
                                 dfl.as<FactorItem>() + b.offset, b.count, ctx->dthg.p),
             "supernode factor retry");
    HIPCHK(hipMemcpyAsync(status.data(), ctx->dstatus.p, size_t(n_super) * 4,
                          hipMemcpyDeviceToHost, ctx->stream),
           "copy status");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    dfl.release();
    failed.clear();
    for (int32_t s = 0; s < n_super; ++s)
      if (status[s] != 0) failed.push_back(s);
  }
  if (!failed.empty())
    return ctx->fail(MFS_ERR_NOT_SPD, "factor: supernode " + std::to_string(failed[0]) +
                                          " has a non-positive pivot (" +
                                          std::to_string(failed.size()) + " supernodes failed)");
  HIPCHK(launch_gather_values(ctx->stream, precision == MFS_FP64 ? 0 : 1, int32_t(nnz),
                              ctx->dtmap.as<int32_t>(), ctx->dval.p, ctx->dtval.p),
         "gather values");
  ctx->prec = precision;
  ctx->has_factor = true;
  return MFS_OK;
}

// Row parts of the band Gram: as many as 8, at least 512 rows each.  Shorter
// blocks balance better over the CUs (a band's blocks carry different
// loads): at m = n = 65,536 (306 blocks per part) the band factor took
// 18.3 ms with 5 parts and 17.3-17.5 ms with 8, 12 or 16.
int band_parts(const mfs_ctx *ctx, int64_t m) {
  (void)ctx;
  int R = 1;
  while (R < 8 && m / (R + 1) >= 512) ++R;
  return R;
}

// Dense FP32 factor in the band form: one band-Gram launch over all bands
// (straight from M), then on the side streams the rest (compact M^T of their
// columns, then their buckets) and the Cholesky and KL columns of the band
// supernodes from the summed partials, bucket by bucket.
int factor_bands(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, int parts) {
  const int64_t part_rows = ((m + parts - 1) / parts + 31) / 32 * 32;
  // Launch order: blocks are dealt to the XCDs round-robin (block i to XCD
  // i mod 8), so XCD x gets the blocks of the bands with index x mod 8, part
  // by part, a band's blocks side by side: blocks that share a band's rows
  // run together on one L2.
  std::vector<std::vector<BandBlock>> seq(8);
  for (int r = 0; r < parts; ++r)
    for (size_t g = 0; g < ctx->band_groups.size(); ++g) {
      BandBlock b = ctx->band_groups[g];
      b.part = r;
      seq[size_t(ctx->band_xcd[g])].push_back(b);
    }
  size_t len = 0;
  for (auto &q : seq) len = std::max(len, q.size());
  BandBlock empty{};
  empty.c0 = -1;
  std::vector<BandBlock> blocks(8 * len, empty);
  for (size_t x = 0; x < 8; ++x)
    for (size_t i = 0; i < seq[x].size(); ++i) blocks[8 * i + x] = seq[x][i];
  HIPCHK(upload(ctx, ctx->dblocks, blocks), "upload band blocks");
  HIPCHK(ctx->dthg.ensure(ctx->rest_scratch), "alloc factor scratch");
  const int nside = std::max(1, std::min<int>(ctx->side_streams, mfs_ctx::kSide));
  // The rest (compact M^T of their columns, then their buckets) on side
  // stream 0, queued before the band Gram so that it can run beside it.
  HIPCHK(hipEventRecord(ctx->fork_ev, ctx->stream), "fork");
  HIPCHK(hipStreamWaitEvent(ctx->side[0], ctx->fork_ev, 0), "fork");
  if (!ctx->rest_items.empty()) {
    HIPCHK(launch_gather_transpose(ctx->side[0], M, m, n, ctx->drestcols.as<int32_t>(),
                                   int32_t(ctx->rest_cols.size()), ctx->dMtc.as<float>(), ctx->fscene.ldmt),
           "gather M^T");
    for (const Bucket &b : ctx->rest_buckets)
      HIPCHK(launch_super_factor(ctx->side[0], b.tiles, 1, false, ctx->dfscene.as<FactorScene>(),
                                 ctx->drestlist.as<FactorItem>() + b.offset, b.count, ctx->dthg.p,
                                 FactorSplit{}, true),
             "supernode factor");
  }
  HIPCHK(launch_band_gram(ctx->stream, M, m, n, ctx->dblocks.as<BandBlock>(), int32_t(blocks.size()),
                          ctx->dbitems.as<int32_t>(), ctx->dsd.as<SuperDesc>(), ctx->dsind.as<int32_t>(),
                          ctx->dPinv.as<int32_t>(), ctx->dgoff.as<int64_t>(), ctx->dgram.as<float>(),
                          ctx->gram_elems, part_rows),
         "band Gram");
  HIPCHK(hipEventRecord(ctx->tr_ev[0], ctx->stream), "fork");
  for (int i = 1; i < nside; ++i) HIPCHK(hipStreamWaitEvent(ctx->side[i], ctx->tr_ev[0], 0), "fork");
  if (nside == 1) HIPCHK(hipStreamWaitEvent(ctx->side[0], ctx->tr_ev[0], 0), "fork");
  // Cholesky and KL columns of the band supernodes
  for (size_t i = 0; i < ctx->gin_buckets.size(); ++i) {
    const Bucket &b = ctx->gin_buckets[i];
    hipStream_t st = nside > 1 ? ctx->side[1 + i % size_t(nside - 1)] : ctx->side[0];
    HIPCHK(launch_super_factor(st, b.tiles, 1, false, ctx->dfscene.as<FactorScene>(),
                               ctx->dginlist.as<FactorItem>() + b.offset, b.count, nullptr,
                               FactorSplit{}, false, true),
           "supernode factor");
  }
  for (int i = 0; i < nside; ++i) {
    HIPCHK(hipEventRecord(ctx->join_ev[i], ctx->side[i]), "join");
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->join_ev[i], 0), "join");
  }
  return MFS_OK;
}

// Supernodal factor from M (pts == nullptr) or from the packed points.
int factor_impl(mfs_ctx *ctx, const float *M, const PointsOp *pts, int64_t m, int64_t n,
                int precision) {
  if (precision != MFS_FP64 && precision != MFS_FP32)
    return ctx->fail(MFS_ERR_ARG, "factor: precision must be MFS_FP64 or MFS_FP32");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  phase_begin(ctx);
  const bool bands = pts == nullptr && precision == MFS_FP32 && ctx->gram_split && ctx->band_mode &&
                     !ctx->band_groups.empty() && n % 2 == 0;
  const int parts = bands ? band_parts(ctx, m) : 0;
  int rc = factor_prepare(ctx, M, pts, m, n, precision, parts);
  if (rc != MFS_OK) return rc;
  ctx->stats.band_supernodes = bands ? ctx->band_supernodes : 0;
  if (bands) {
    rc = ensure_side_streams(ctx);
    if (rc != MFS_OK) return rc;
    track_cb(ctx, MFS_K_GRAM, 0);
    rc = factor_bands(ctx, M, m, n, parts);
    if (rc != MFS_OK) return rc;
    track_cb(ctx, MFS_K_GRAM, 1);
    rc = factor_finish(ctx, false, precision, M, m, n);
    if (rc != MFS_OK) return rc;
    ctx->stats.factor_ms = phase_end(ctx);
    track_fold(ctx);
    return MFS_OK;
  }
  track_cb(ctx, MFS_K_GRAM, 0);
  // Buckets (largest tiles first) run on side streams at once, so that the
  // few large supernodes and each launch's tail overlap the bulk of the
  // smaller ones instead of leaving CUs idle between launches.
  rc = ensure_side_streams(ctx);
  if (rc != MFS_OK) return rc;
  // The launches: one per (bucket, column range) in the dense form, one per
  // bucket from the points.  Launches of few supernodes split each Gram's
  // rows over several blocks (factor_splits) so the chip stays busy; their
  // partial Grams, the large buckets' TH and the arrival counters get
  // regions of their own (sized for FP64, the widest compute type), since
  // launches on different streams run at once.
  const int prec_code = precision == MFS_FP64 ? 0 : 1;
  const int nside = std::min<int>(ctx->side_streams, int(ctx->buckets.size()));
  struct FLaunch {
    size_t bucket;
    int phase;
    int32_t lo, cnt;
    FactorSplit sp;
    size_t off;      // bytes into dthg: TH (large tiles) or partial Grams
    size_t ctr;      // counters offset (int32 elements)
  };
  const int64_t rows = pts ? m : ctx->fscene.ldmt;
  const int64_t target = int64_t(ctx->cus) * 4;
  const int nph = pts ? 1 : ctx->plan_phases;
  std::vector<FLaunch> fl;
  size_t bytes = 8, ctrs = 1;
  for (int ph = 0; ph < nph; ++ph)
    for (size_t bi = 0; bi < ctx->buckets.size(); ++bi) {
      const Bucket &b = ctx->buckets[bi];
      const int32_t lo = pts ? 0 : b.phase_off[ph];
      const int32_t cnt = pts ? b.count : b.phase_off[ph + 1] - lo;
      if (cnt <= 0) continue;
      FLaunch L{bi, ph, lo, cnt, FactorSplit{}, bytes, ctrs};
      L.sp.splits = std::min(ctx->max_splits, factor_splits(cnt, rows, target));
      const size_t ns = size_t(16) * b.tiles;
      const size_t blocks = size_t(cnt) * L.sp.splits;
      if (b.tiles > 8) bytes += blocks * size_t(factor_scratch_elems(b.tiles)) * 8;
      else if (L.sp.splits > 1) bytes += blocks * ns * ns * 8;
      bytes = (bytes + 255) / 256 * 256;
      if (L.sp.splits > 1) ctrs += size_t(cnt);
      fl.push_back(L);
    }
  HIPCHK(ctx->dthg.ensure(bytes), "alloc factor scratch");
  HIPCHK(ctx->dctr.ensure(ctrs * 4), "alloc factor counters");
  HIPCHK(hipMemsetAsync(ctx->dctr.p, 0, ctrs * 4, ctx->stream), "clear factor counters");
  for (FLaunch &L : fl) {
    const Bucket &b = ctx->buckets[L.bucket];
    if (L.sp.splits > 1) {
      L.sp.counters = ctx->dctr.as<int32_t>() + L.ctr;
      if (b.tiles <= 8) L.sp.partial = ctx->dthg.as<char>() + L.off;
    }
  }
  auto launch = [&](const FLaunch &L) -> int {
    const Bucket &b = ctx->buckets[L.bucket];
    HIPCHK(launch_super_factor(ctx->side[L.bucket % nside], b.tiles, prec_code, pts != nullptr,
                               ctx->dfscene.as<FactorScene>(),
                               ctx->dlist.as<FactorItem>() + b.offset + L.lo, L.cnt,
                               ctx->dthg.as<char>() + L.off, L.sp,
                               ctx->gram_split && prec_code == 1),
           "supernode factor");
    return MFS_OK;
  };
  if (pts == nullptr) {
    // M^T range by range on the context stream; a bucket's side stream runs
    // its phase-p supernodes once range p is done, phases in turn, while the
    // later ranges are transposed.
    const int64_t ldmt = ctx->fscene.ldmt;
    for (int ph = 0; ph < nph; ++ph) {
      HIPCHK(launch_transpose(ctx->stream, M, m, n, ctx->dMt.as<float>(), ldmt, phase_col(ph, n, nph),
                              phase_col(ph + 1, n, nph)),
             "transpose");
      HIPCHK(hipEventRecord(ctx->tr_ev[ph], ctx->stream), "fork");
    }
    size_t li = 0;
    for (int ph = 0; ph < nph; ++ph) {
      for (int i = 0; i < nside; ++i) HIPCHK(hipStreamWaitEvent(ctx->side[i], ctx->tr_ev[ph], 0), "fork");
      // this range's launches (fl holds them range by range, in bucket order)
      for (; li < fl.size() && fl[li].phase == ph; ++li)
        if ((rc = launch(fl[li])) != MFS_OK) return rc;
    }
  } else {
    HIPCHK(hipEventRecord(ctx->fork_ev, ctx->stream), "fork");
    for (int i = 0; i < nside; ++i) HIPCHK(hipStreamWaitEvent(ctx->side[i], ctx->fork_ev, 0), "fork");
    for (const FLaunch &L : fl)
      if ((rc = launch(L)) != MFS_OK) return rc;
  }
  for (int i = 0; i < nside; ++i) {
    HIPCHK(hipEventRecord(ctx->join_ev[i], ctx->side[i]), "join");
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->join_ev[i], 0), "join");
  }
  track_cb(ctx, MFS_K_GRAM, 1);
  rc = factor_finish(ctx, pts != nullptr, precision, M, m, n);
  if (rc != MFS_OK) return rc;
  ctx->stats.factor_ms = phase_end(ctx);
  track_fold(ctx);
  return MFS_OK;
}

// Mixed-precision solve with M (pts == nullptr) or matrix-free from the points.
// One column batch (col0 .. col0 + kval) of the mixed-precision solve, as a
// sequence of phases so that many scenes can run them side by side:
//   begin   -- X64 from X0, ||M^T b||^2 and the first residual, norms to host
//   decide  -- (host, after a sync) relative residuals; is another sweep due?
//   start   -- inner FP32 PCG on M^T M d = R64 with absolute thresholds
//   (PCG iterations: polled by `iterate`, or launched in rounds by a batch)
//   end     -- X64 += d, FP64 residual, norms to host
//   finish  -- X64 to the FP64 and/or FP32 outputs
struct MixedRun {
  mfs_ctx *ctx = nullptr;
  const float *B = nullptr;
  int rhs_space = 0;  // MFS_RHS_BOUNDARY: B m x k; MFS_RHS_SOURCE: B n x k
  int64_t ldb = 0, col0 = 0, n = 0;
  int kval = 0, K = 1;
  int32_t maxit = 0;
  double tol = 0, inner_tol = 0.5;
  SolveArgs<double> a64{};
  SolveArgs<float> a32{};
  double *X64 = nullptr, *R64 = nullptr, *nb2 = nullptr, *res2 = nullptr, *thr = nullptr;
  double *X64b = nullptr;       // the iterate before the running sweep
  int32_t inner_used = 0;
  int outer = 0;
  std::vector<int32_t> iters;
  // Safeguard of the FP32 sweeps: a column whose FP64 residual a sweep made
  // worse is restored from the backup, and once a sweep gains less than 4x on
  // a column that is not yet done the remaining sweeps run in FP64 (the plain
  // form's PCG on the same residual system).  FP32 sweeps stall like this
  // only on blocks near singularity, where the FP32 recurrences lose the
  // directions the preconditioner amplifies.
  bool fp64_inner = false;
  std::vector<double> prev_rel;
  int reverted = 0;             // columns restored (all sweeps)
  int fp64_sweeps = 0;          // sweeps run in FP64
  int fault = MFS_OK;           // set when decide() could not restore columns
  bool swept = false;           // a sweep ended since the last decide()

  // FP64 residual of the right-hand-side space: M^T (b - M x) or b - M^T M x.
  hipError_t residual(const double *X, double *out2) {
    return rhs_space == MFS_RHS_SOURCE
               ? residual64_source(ctx->stream, a64, B, ldb, col0, kval, X, R64, out2, track_cb, ctx)
               : residual64(ctx->stream, a64, B, ldb, col0, kval, X, R64, out2, track_cb, ctx);
  }

  int begin(mfs_ctx *c, const float *M, const PointsOp *pts, int64_t m, int64_t n_, const float *B_,
            int64_t ldb_, int64_t col0_, int kval_, int32_t maxit_, double tol_, double inner_tol_,
            const float *X0, const double *X0d, int rhs_space_ = MFS_RHS_BOUNDARY) {
    ctx = c; B = B_; ldb = ldb_; col0 = col0_; n = n_; kval = kval_; K = batch_width(kval);
    rhs_space = rhs_space_;
    maxit = maxit_; tol = tol_; inner_tol = inner_tol_;
    iters.assign(kval, 0);
    int rc = workspace<double>(ctx, M, pts, m, n, K, a64);
    if (rc != MFS_OK) return rc;
    if (pts) set_pass_work(ctx, pts, m, n, K);
    HIPCHK(ctx->dx64.ensure(size_t(n) * K * 8 * 2), "alloc x64");
    HIPCHK(ctx->dr64.ensure(size_t(n) * K * 8), "alloc r64");
    HIPCHK(ctx->dnorm.ensure(4 * kMaxRhs * 8), "alloc norms");
    X64 = ctx->dx64.as<double>();
    X64b = X64 + size_t(n) * K;
    fp64_inner = false;
    prev_rel.assign(kval, -1.0);
    reverted = 0;
    fp64_sweeps = 0;
    swept = false;
    fault = MFS_OK;
    R64 = ctx->dr64.as<double>();
    nb2 = ctx->dnorm.as<double>();
    res2 = nb2 + kMaxRhs;
    thr = nb2 + 2 * kMaxRhs;
    // FP32 inner workspace; its vectors live in the FP64 vector buffer
    const void *lv = nullptr, *ltv = nullptr;
    rc = values_for(ctx, MFS_FP32, &lv, &ltv);
    if (rc != MFS_OK) return rc;
    a32 = SolveArgs<float>{};
    a32.M = M; a32.m = m; a32.n = n; a32.K = K;
    a32.P = a64.P; a32.Pinv = a64.Pinv; a32.sptr = a64.sptr; a32.sind = a64.sind;
    a32.tptr = a64.tptr; a32.tcol = a64.tcol;
    a32.lval = static_cast<const float *>(lv);
    a32.ltval = static_cast<const float *>(ltv);
    a32.row_blocks = a64.row_blocks;
    if (!pts) {
      // the FP32 passes get their own split count (a different kernel with
      // its own occupancy) over the same partials buffer
      a32.row_blocks = row_blocks_for(m, n, ctx->cols_blocks,
                                      int64_t(ctx->cus) * cols_resident_blocks(false, K, n % 4 == 0));
      if (a32.row_blocks > a64.row_blocks) {
        HIPCHK(ctx->dpart.ensure(size_t(a32.row_blocks) * size_t(n) * K * sizeof(double)),
               "alloc partials");
        a64.part = ctx->dpart.as<double>();
      }
    }
    a32.pts = a64.pts;
    float *base = reinterpret_cast<float *>(a64.x);
    const size_t nv = size_t(n) * K;
    a32.x = base; a32.r = base + nv; a32.p = base + 2 * nv; a32.q = base + 3 * nv;
    a32.z = base + 4 * nv; a32.t = base + 5 * nv; a32.pv = base + 6 * nv; a32.y = base + 7 * nv;
    a32.part = a64.part; a32.dpart = a64.dpart; a32.st = a64.st;
    rc = setup_fused(ctx, pts, m, n, K, a32.fused, a32.frec, a32.part);
    if (rc != MFS_OK) return rc;
    a64.part = a32.part;

    const unsigned gn = unsigned((n + kThreads - 1) / kThreads);
    if (X0d != nullptr)
      k_x64_from_f64<<<gn, kThreads, 0, ctx->stream>>>(n, X0d, ldb, col0, kval, K, X64);
    else
      k_x64_from_f32<<<gn, kThreads, 0, ctx->stream>>>(n, X0, ldb, col0, kval, K, X64);
    // ||M^T b||^2 (||b||^2 for source-space data) and, for a zero start, the
    // first residual.
    HIPCHK(residual(nullptr, nb2), "residual");
    if (X0 != nullptr || X0d != nullptr)
      HIPCHK(residual(X64, res2), "residual");
    else
      HIPCHK(hipMemcpyAsync(res2, nb2, kMaxRhs * 8, hipMemcpyDeviceToDevice, ctx->stream), "copy");
    HIPCHK(hipMemcpyAsync(ctx->hnorm, nb2, 2 * kMaxRhs * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
    return MFS_OK;
  }

  double rel_of(int r) const {
    const double b2 = ctx->hnorm[r], e2 = ctx->hnorm[kMaxRhs + r];
    return b2 > 0 ? std::sqrt(e2 / b2) : 0.0;
  }

  // After ctx->stream has been synchronized.  Returns whether another sweep
  // is due; may restore columns (and then recompute and re-read the residual).
  bool decide() {
    if (swept) {
      swept = false;
      uint32_t worse = 0;
      bool stalled = false;
      for (int r = 0; r < kval; ++r) {
        if (!(ctx->hnorm[r] >= 1e-16) || !(prev_rel[r] >= 0)) continue;
        const double rel = rel_of(r);
        if (!(rel <= prev_rel[r])) worse |= 1u << r;   // also NaN
        if (!(rel <= tol) && !(rel <= 0.25 * prev_rel[r])) stalled = true;
      }
      if (stalled) fp64_inner = true;
      if (worse != 0u) {
        const unsigned gn = unsigned((n + kThreads - 1) / kThreads);
        k_restore_cols<<<gn, kThreads, 0, ctx->stream>>>(n, K, worse, X64b, X64);
        for (int r = 0; r < kval; ++r) reverted += (worse >> r) & 1u;
        hipError_t e = residual(X64, res2);
        if (e == hipSuccess)
          e = hipMemcpyAsync(ctx->hnorm, nb2, 2 * kMaxRhs * 8, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
          fault = ctx->hip_fail(e, "mixed solve: restore columns");
          return false;
        }
      }
    }
    bool done = true;
    for (int r = 0; r < kval; ++r) {
      const double rel = rel_of(r);
      prev_rel[r] = rel;
      if (col0 + r < 64) ctx->stats.relres[col0 + r] = rel;
      if (ctx->hnorm[r] >= 1e-16 && !(rel <= tol)) done = false;
    }
    return !(done || inner_used >= maxit || outer >= 30);
  }

  int start() {
    // Each sweep aims straight at inner_tol * tol * ||M^T b|| (absolute), so
    // one sweep normally meets the tolerance and the FP64 residual check
    // only confirms it; FP32 round-off left above it is removed by the next.
    HIPCHK(inner_thresholds(ctx->stream, nb2, inner_tol * tol, K, thr), "inner thresholds");
    HIPCHK(hipMemcpyAsync(X64b, X64, size_t(n) * K * 8, hipMemcpyDeviceToDevice, ctx->stream),
           "backup");
    if (fp64_inner) {
      HIPCHK(pcg_init<double>(ctx->stream, a64, nullptr, 0, 0, 1, K, 0.0, maxit - inner_used, R64, thr),
             "inner init");
    } else {
      // An FP32 sweep leaves a quarter of the remaining budget (it gets at
      // least 64), so a stall still leaves iterations for FP64 sweeps; a
      // sweep cut short restarts the Krylov space, so the cap is kept high.
      const int32_t left = maxit - inner_used;
      const int32_t cap = std::min(left, std::max<int32_t>(64, left - left / 4));
      HIPCHK(pcg_init<float>(ctx->stream, a32, nullptr, 0, 0, 1, K, 0.0, cap, R64, thr), "inner init");
    }
    return MFS_OK;
  }

  // One PCG iteration of the running sweep (FP32, or FP64 after a stall).
  hipError_t step(hipStream_t st, void (*fn)(void *, int, int), void *tctx) {
    return fp64_inner ? pcg_iteration<double>(st, a64, fn, tctx) : pcg_iteration<float>(st, a32, fn, tctx);
  }
  // Runs the sweep's iterations to the end (polling the state).
  int run_inner() {
    return fp64_inner ? iterate<double>(ctx, a64, maxit - inner_used)
                      : iterate<float>(ctx, a32, maxit - inner_used);
  }

  // After ctx->hstate holds the finished inner PCG's state.
  int end() {
    if (fp64_inner)
      HIPCHK(pcg_finish<double>(ctx->stream, a64, nullptr, 0, 0, K, X64), "inner finish");
    else
      HIPCHK(pcg_finish<float>(ctx->stream, a32, nullptr, 0, 0, K, X64), "inner finish");
    int32_t used = 0;
    for (int r = 0; r < kval; ++r) {
      iters[r] += ctx->hstate->iters[r];
      used = std::max(used, ctx->hstate->iters[r]);
    }
    inner_used += std::max(used, 1);
    ++outer;
    if (fp64_inner) ++fp64_sweeps;
    swept = true;
    HIPCHK(residual(X64, res2), "residual");
    HIPCHK(hipMemcpyAsync(ctx->hnorm, nb2, 2 * kMaxRhs * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
    return MFS_OK;
  }

  int finish(float *X, double *Xd) {
    const unsigned gn = unsigned((n + kThreads - 1) / kThreads);
    if (Xd != nullptr)
      k_x64_to_f64<<<gn, kThreads, 0, ctx->stream>>>(n, X64, K, Xd, ldb, col0, kval);
    if (X != nullptr)
      k_x64_to_f32<<<gn, kThreads, 0, ctx->stream>>>(n, X64, K, X, ldb, col0, kval);
    HIPCHK(hipGetLastError(), "output");
    for (int r = 0; r < kval; ++r)
      if (col0 + r < 64) ctx->stats.iters[col0 + r] = iters[r];
    return MFS_OK;
  }
};

// X0 / X are FP32; X0d / Xd (used instead when given) are FP64 -- for
// tolerances below what an FP32 copy of the strengths can hold.
int solve_mixed_impl(mfs_ctx *ctx, const float *M, const PointsOp *pts, int64_t m, int64_t n,
                     const float *B, int64_t k, int rhs_space, const float *X0, float *X,
                     int32_t maxit, double tol, double inner_tol, const double *X0d = nullptr,
                     double *Xd = nullptr) {
  if (!ctx->has_factor) return ctx->fail(MFS_ERR_STATE, "no factor: factor before solving");
  if (B == nullptr || (X == nullptr && Xd == nullptr) || k <= 0 ||
      (rhs_space != MFS_RHS_BOUNDARY && rhs_space != MFS_RHS_SOURCE))
    return ctx->fail(MFS_ERR_ARG, "solve_mixed: bad arguments");
  if (!(inner_tol > 0) || !(inner_tol <= 1)) inner_tol = 0.5;
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  phase_begin(ctx);
  set_pass_work(ctx, pts, m, n);
  int outer_total = 0;
  ctx->stats.fp64_sweeps = 0;
  ctx->stats.reverted_columns = 0;
  for (int64_t col0 = 0; col0 < k; col0 += kMaxRhs) {
    const int kval = int(std::min<int64_t>(kMaxRhs, k - col0));
    MixedRun run;
    int rc = run.begin(ctx, M, pts, m, n, B, k, col0, kval, maxit, tol, inner_tol, X0, X0d, rhs_space);
    if (rc != MFS_OK) return rc;
    for (;;) {
      HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
      if (!run.decide()) {
        if (run.fault != MFS_OK) return run.fault;
        break;
      }
      if ((rc = run.start()) != MFS_OK) return rc;
      if ((rc = run.run_inner()) != MFS_OK) return rc;
      if ((rc = run.end()) != MFS_OK) return rc;
    }
    if ((rc = run.finish(X, Xd)) != MFS_OK) return rc;
    outer_total = std::max(outer_total, run.outer);
    ctx->stats.fp64_sweeps = std::max(ctx->stats.fp64_sweeps, run.fp64_sweeps);
    ctx->stats.reverted_columns += run.reverted;
  }
  ctx->stats.outer_iters = outer_total;
  ctx->stats.solve_ms = phase_end(ctx);
  track_fold(ctx);
  return MFS_OK;
}

// FP64 normal-equation residual per column, with M or from the points.
int residual_impl(mfs_ctx *ctx, const float *M, const PointsOp *pts, int64_t m, int64_t n,
                  const float *B, int64_t k, const float *X, double *out,
                  const double *Xd = nullptr) {
  int rc = MFS_OK;
  if (!ctx->has_factor) return ctx->fail(MFS_ERR_STATE, "no factor: factor before solving");
  if (B == nullptr || (X == nullptr && Xd == nullptr) || k <= 0)
    return ctx->fail(MFS_ERR_ARG, "residual: bad arguments");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  for (int64_t col0 = 0; col0 < k; col0 += kMaxRhs) {
    const int kval = int(std::min<int64_t>(kMaxRhs, k - col0));
    const int K = batch_width(kval);
    SolveArgs<double> a{};
    rc = workspace<double>(ctx, M, pts, m, n, K, a);
    if (rc != MFS_OK) return rc;
    HIPCHK(ctx->dx64.ensure(size_t(n) * K * 8), "alloc");
    HIPCHK(ctx->dr64.ensure(size_t(n) * K * 8), "alloc");
    HIPCHK(ctx->dnorm.ensure(4 * kMaxRhs * 8), "alloc");
    double *nb2 = ctx->dnorm.as<double>();
    const unsigned gn = unsigned((n + kThreads - 1) / kThreads);
    if (Xd != nullptr)
      k_x64_from_f64<<<gn, kThreads, 0, ctx->stream>>>(n, Xd, k, col0, kval, K, ctx->dx64.as<double>());
    else
      k_x64_from_f32<<<gn, kThreads, 0, ctx->stream>>>(n, X, k, col0, kval, K, ctx->dx64.as<double>());
    HIPCHK(residual64(ctx->stream, a, B, k, col0, kval, nullptr, ctx->dr64.as<double>(), nb2, nullptr, nullptr), "residual");
    HIPCHK(residual64(ctx->stream, a, B, k, col0, kval, ctx->dx64.as<double>(), ctx->dr64.as<double>(),
                      nb2 + kMaxRhs, nullptr, nullptr), "residual");
    HIPCHK(hipMemcpyAsync(ctx->hnorm, nb2, 2 * kMaxRhs * 8, hipMemcpyDeviceToHost, ctx->stream), "copy");
    HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
    for (int r = 0; r < kval; ++r) {
      const double b2 = ctx->hnorm[r];
      out[col0 + r] = b2 > 0 ? std::sqrt(ctx->hnorm[kMaxRhs + r] / b2) : 0.0;
    }
  }
  return MFS_OK;
}

}  // namespace

extern "C" {

int mfs_ctx_create(mfs_ctx **out, int device, void *stream) {
  if (out == nullptr) return MFS_ERR_ARG;
  *out = nullptr;
  mfs_ctx *ctx = new (std::nothrow) mfs_ctx();
  if (ctx == nullptr) return MFS_ERR_ALLOC;
  ctx->device = device;
  ctx->stream = static_cast<hipStream_t>(stream);
  if (hipSetDevice(device) != hipSuccess || hipEventCreate(&ctx->ev0) != hipSuccess ||
      hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&ctx->hstate), sizeof(PcgState)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&ctx->hnorm), 4 * kMaxRhs * sizeof(double)) != hipSuccess ||
      hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    mfs_ctx_destroy(ctx);
    return MFS_ERR_HIP;
  }
  scratch_lookup(device);  // the device is current (set above)
  *out = ctx;
  return MFS_OK;
}

void mfs_ctx_destroy(mfs_ctx *ctx) {
  if (ctx == nullptr) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  for (DevBuf *b : {&ctx->dP, &ctx->dPinv, &ctx->dsptr, &ctx->dsind, &ctx->dtptr, &ctx->dtcol,
                    &ctx->dtmap, &ctx->dsupind, &ctx->dsd, &ctx->dlist, &ctx->dstatus, &ctx->dpx,
                    &ctx->dpy, &ctx->dpz, &ctx->dscale, &ctx->dcnt, &ctx->dptr, &ctx->dind,
                    &ctx->dtotal, &ctx->dbands, &ctx->dcells, &ctx->dcpts, &ctx->dlong, &ctx->dval, &ctx->dtval, &ctx->dMt, &ctx->dval_alt,
                    &ctx->dtval_alt, &ctx->dthg, &ctx->dctr, &ctx->dfscene, &ctx->dvec, &ctx->dpart, &ctx->ddpart, &ctx->dstate,
                    &ctx->dx64, &ctx->dr64, &ctx->dnorm, &ctx->dt32, &ctx->dt64, &ctx->ds32,
                    &ctx->ds64, &ctx->dapin, &ctx->dapout, &ctx->dcaller, &ctx->dflag, &ctx->dfrec,
                    &ctx->dgrad})
    b->release();
  if (ctx->hstate) (void)hipHostFree(ctx->hstate);
  if (ctx->hnorm) (void)hipHostFree(ctx->hnorm);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (int i = 0; i < mfs_ctx::kSide; ++i) {
    if (ctx->side[i]) (void)hipStreamDestroy(ctx->side[i]);
    if (ctx->join_ev[i]) (void)hipEventDestroy(ctx->join_ev[i]);
  }
  for (int i = 0; i < kMaxPhases; ++i)
    if (ctx->tr_ev[i]) (void)hipEventDestroy(ctx->tr_ev[i]);
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  for (auto &t : ctx->tr)
    for (auto &pr : t.pairs) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  delete ctx;
}

int mfs_ctx_set_stream(mfs_ctx *ctx, void *stream) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  ctx->stream = static_cast<hipStream_t>(stream);
  return MFS_OK;
}

const char *mfs_ctx_last_error(const mfs_ctx *ctx) {
  return ctx == nullptr ? "null context" : ctx->err.c_str();
}

int mfs_ctx_stats(const mfs_ctx *ctx, mfs_stats *out) {
  if (ctx == nullptr || out == nullptr) return MFS_ERR_ARG;
  *out = ctx->stats;
  return MFS_OK;
}

int mfs_plan(mfs_ctx *ctx, const float *src_xyz_host, int64_t n, double rho, double lambda,
             int32_t max_super) {
  if (ctx == nullptr || src_xyz_host == nullptr) return MFS_ERR_ARG;
  std::vector<double> xyz(size_t(std::max<int64_t>(n, 0)) * 3);
  for (size_t i = 0; i < xyz.size(); ++i) xyz[i] = double(src_xyz_host[i]);
  return plan_impl(ctx, std::move(xyz), n, rho, lambda, max_super);
}

int mfs_plan_f64(mfs_ctx *ctx, const double *src_xyz_host, int64_t n, double rho, double lambda,
                 int32_t max_super) {
  if (ctx == nullptr || src_xyz_host == nullptr) return MFS_ERR_ARG;
  std::vector<double> xyz(src_xyz_host, src_xyz_host + size_t(std::max<int64_t>(n, 0)) * 3);
  return plan_impl(ctx, std::move(xyz), n, rho, lambda, max_super);
}

int mfs_order_maximin(const double *xyz, int64_t n, int32_t *P, int32_t *Pinv, double *scale) {
  if (xyz == nullptr || n <= 0 || n > (int64_t(1) << 30)) return MFS_ERR_ARG;
  std::vector<int32_t> p, pinv;
  std::vector<double> sc;
  const int rc = order_maximin(int32_t(n), xyz, p, pinv, sc);
  if (rc == PLAN_ERR_DUPLICATE) return MFS_ERR_DUPLICATE;
  if (rc != PLAN_OK) return MFS_ERR_ARG;
  if (P != nullptr) std::memcpy(P, p.data(), size_t(n) * 4);
  if (Pinv != nullptr) std::memcpy(Pinv, pinv.data(), size_t(n) * 4);
  if (scale != nullptr) std::memcpy(scale, sc.data(), size_t(n) * 8);
  return MFS_OK;
}

int mfs_plan_export(const mfs_ctx *ctx, int32_t *P, int32_t *Pinv, double *scale, int32_t *ptr,
                    int32_t *ind, int32_t *sup_ptr, int32_t *sup_ind, int32_t *sup_parent,
                    int32_t *sptr, int32_t *sind) {
  if (ctx == nullptr || !ctx->has_plan) return MFS_ERR_STATE;
  auto cp = [](auto *dst, const auto &v) {
    if (dst != nullptr && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(P, ctx->P);
  cp(Pinv, ctx->Pinv);
  cp(scale, ctx->scale);
  cp(ptr, ctx->ptr);
  cp(ind, ctx->ind);
  cp(sup_ptr, ctx->sup_ptr);
  cp(sup_ind, ctx->sup_ind);
  cp(sup_parent, ctx->sup_parent);
  cp(sptr, ctx->sptr);
  cp(sind, ctx->sind);
  return MFS_OK;
}

int mfs_factor_export(const mfs_ctx *ctx_c, double *val) {
  mfs_ctx *ctx = const_cast<mfs_ctx *>(ctx_c);
  if (ctx == nullptr || val == nullptr) return MFS_ERR_ARG;
  if (!ctx->has_factor) return ctx->fail(MFS_ERR_STATE, "no factor");
  const int64_t nnz = ctx->sptr[ctx->n];
  if (ctx->prec == MFS_FP64) {
    HIPCHK(hipMemcpy(val, ctx->dval.p, size_t(nnz) * 8, hipMemcpyDeviceToHost), "copy factor");
  } else {
    std::vector<float> tmp(nnz);
    HIPCHK(hipMemcpy(tmp.data(), ctx->dval.p, size_t(nnz) * 4, hipMemcpyDeviceToHost), "copy factor");
    for (int64_t i = 0; i < nnz; ++i) val[i] = tmp[i];
  }
  return MFS_OK;
}

}  // extern "C"

namespace {

// The plain form's column loop; the strengths go to X (FP32) or Xd (FP64).
int solve_dense_impl(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *B, int64_t k,
                     int rhs_space, float *X, double *Xd, int32_t maxit, double tol) {
  int rc = check_dense_args(ctx, M, m, n);
  if (rc != MFS_OK) return rc;
  if (!ctx->has_factor) return ctx->fail(MFS_ERR_STATE, "no factor: call mfs_factor_dense first");
  if (B == nullptr || (X == nullptr && Xd == nullptr) || k <= 0 || (rhs_space != 0 && rhs_space != 1))
    return ctx->fail(MFS_ERR_ARG, "solve: bad right-hand side arguments");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  SolveArgs<double> a{};
  rc = workspace<double>(ctx, M, nullptr, m, n, 1, a);
  if (rc != MFS_OK) return rc;
  phase_begin(ctx);
  for (int c = 1; c < 8; ++c) ctx->bytes_per_launch[c] = double(m) * double(n) * 4.0;
  for (int64_t c = 0; c < k; ++c) {
    HIPCHK(pcg_init<double>(ctx->stream, a, B, k, c, rhs_space, 1, tol, maxit, nullptr), "pcg init");
    rc = iterate<double>(ctx, a, maxit);
    if (rc != MFS_OK) return rc;
    if (Xd != nullptr)
      HIPCHK(pcg_finish64(ctx->stream, a, Xd, k, c, 1), "pcg finish");
    else
      HIPCHK(pcg_finish<double>(ctx->stream, a, X, k, c, 1, nullptr), "pcg finish");
    record_batch_stats(ctx, c, 1);
  }
  ctx->stats.outer_iters = 0;
  ctx->stats.solve_ms = phase_end(ctx);
  track_fold(ctx);
  return MFS_OK;
}

}  // namespace

extern "C" {

int mfs_solve_dense(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *B, int64_t k,
                    int rhs_space, float *X, int32_t maxit, double tol) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if (X == nullptr) return ctx->fail(MFS_ERR_ARG, "solve: X is null");
  return solve_dense_impl(ctx, M, m, n, B, k, rhs_space, X, nullptr, maxit, tol);
}

int mfs_solve_dense64(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *B, int64_t k,
                      int rhs_space, double *X, int32_t maxit, double tol) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if (X == nullptr) return ctx->fail(MFS_ERR_ARG, "solve: X is null");
  return solve_dense_impl(ctx, M, m, n, B, k, rhs_space, nullptr, X, maxit, tol);
}

int mfs_factor_dense(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, int precision) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  int rc = check_dense_args(ctx, M, m, n);
  if (rc != MFS_OK) return rc;
  return factor_impl(ctx, M, nullptr, m, n, precision);
}

int mfs_solve_mixed(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *B, int64_t k,
                    int rhs_space, const float *X0, float *X, int32_t maxit, double tol,
                    double inner_tol) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  int rc = check_dense_args(ctx, M, m, n);
  if (rc != MFS_OK) return rc;
  return solve_mixed_impl(ctx, M, nullptr, m, n, B, k, rhs_space, X0, X, maxit, tol, inner_tol);
}

int mfs_normal_residual(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *B,
                        int64_t k, const float *X, double *out) {
  if (ctx == nullptr || out == nullptr) return MFS_ERR_ARG;
  int rc = check_dense_args(ctx, M, m, n);
  if (rc != MFS_OK) return rc;
  return residual_impl(ctx, M, nullptr, m, n, B, k, X, out);
}

int mfs_factor_points(mfs_ctx *ctx, const double *x, const double *nx, int64_t m, const double *y,
                      int64_t n, int precision) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if (!ctx->has_plan) return ctx->fail(MFS_ERR_STATE, "no plan: call mfs_plan first");
  if (x == nullptr || nx == nullptr || y == nullptr || m <= 0 || m > INT32_MAX || n != ctx->n)
    return ctx->fail(MFS_ERR_ARG, "points: need m > 0 targets and the n planned sources");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  ctx->has_points = false;
  HIPCHK(ctx->dt32.ensure(size_t(m) * kTgt32 * 4), "alloc points");
  HIPCHK(ctx->dt64.ensure(size_t(m) * kTgt64 * 8), "alloc points");
  HIPCHK(ctx->ds32.ensure(size_t(n) * kSrc32 * 4), "alloc points");
  HIPCHK(ctx->ds64.ensure(size_t(n) * kSrc64 * 8), "alloc points");
  HIPCHK(pack_points(ctx->stream, x, nx, y, m, n, ctx->dt32.as<float>(), ctx->dt64.as<double>(),
                     ctx->ds32.as<float>(), ctx->ds64.as<double>()),
         "pack points");
  ctx->pts.tgt32 = ctx->dt32.as<float>();
  ctx->pts.src32 = ctx->ds32.as<float>();
  ctx->pts.tgt64 = ctx->dt64.as<double>();
  ctx->pts.src64 = ctx->ds64.as<double>();
  ctx->pts.m = m;
  ctx->pts.n = n;
  ctx->pts.target_blocks = ctx->points_blocks;
  ctx->has_points = true;
  return factor_impl(ctx, nullptr, &ctx->pts, m, n, precision);
}

int mfs_solve_points(mfs_ctx *ctx, const float *B, int64_t k, const float *X0, float *X,
                     int32_t maxit, double tol, double inner_tol) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  int rc = check_points(ctx);
  if (rc != MFS_OK) return rc;
  return solve_mixed_impl(ctx, nullptr, &ctx->pts, ctx->pts.m, ctx->n, B, k, 0, X0, X, maxit, tol,
                          inner_tol);
}

int mfs_solve_points64(mfs_ctx *ctx, const float *B, int64_t k, const double *X0, double *X,
                       int32_t maxit, double tol, double inner_tol) {
  if (ctx == nullptr || X == nullptr) return MFS_ERR_ARG;
  int rc = check_points(ctx);
  if (rc != MFS_OK) return rc;
  return solve_mixed_impl(ctx, nullptr, &ctx->pts, ctx->pts.m, ctx->n, B, k, 0, nullptr, nullptr,
                          maxit, tol, inner_tol, X0, X);
}

int mfs_points_residual64(mfs_ctx *ctx, const float *B, int64_t k, const double *X, double *out) {
  if (ctx == nullptr || out == nullptr || X == nullptr) return MFS_ERR_ARG;
  int rc = check_points(ctx);
  if (rc != MFS_OK) return rc;
  return residual_impl(ctx, nullptr, &ctx->pts, ctx->pts.m, ctx->n, B, k, nullptr, out, X);
}

int mfs_points_residual(mfs_ctx *ctx, const float *B, int64_t k, const float *X, double *out) {
  if (ctx == nullptr || out == nullptr) return MFS_ERR_ARG;
  int rc = check_points(ctx);
  if (rc != MFS_OK) return rc;
  return residual_impl(ctx, nullptr, &ctx->pts, ctx->pts.m, ctx->n, B, k, X, out);
}

int mfs_dense_normal_apply(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *X,
                           int64_t k, double *Q) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if (M == nullptr || X == nullptr || Q == nullptr || m <= 0 || n <= 0 || k <= 0 || k > kMaxRhs)
    return ctx->fail(MFS_ERR_ARG, "dense_normal_apply: bad arguments (1 <= k <= 16)");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  const int K = batch_width(k);
  SolveArgs<float> a{};
  a.M = M;
  a.m = m;
  a.n = n;
  a.K = K;
  a.row_blocks = row_blocks_for(m, n, ctx->cols_blocks,
                                int64_t(ctx->cus) * cols_resident_blocks(false, K, n % 4 == 0));
  a.pts = PointsOp{};
  const size_t nv = size_t(n) * K, mv = size_t(m) * K;
  HIPCHK(ctx->dvec.ensure((nv + mv) * sizeof(float) + 64), "alloc vectors");
  a.pv = ctx->dvec.as<float>();
  a.y = a.pv + nv;
  HIPCHK(ctx->dpart.ensure(size_t(a.row_blocks) * nv * sizeof(double)), "alloc partials");
  a.part = ctx->dpart.as<double>();
  if (ctx->dstate.p == nullptr) {
    HIPCHK(ctx->dstate.ensure(sizeof(PcgState)), "alloc state");
    HIPCHK(hipMemsetAsync(ctx->dstate.p, 0, sizeof(PcgState), ctx->stream), "memset state");
  }
  a.st = ctx->dstate.as<PcgState>();
  int rc = setup_fused(ctx, nullptr, m, n, K, a.fused, a.frec, a.part);
  if (rc != MFS_OK) return rc;
  const int32_t one = 1;
  HIPCHK(hipMemcpyAsync(&a.st->n_active, &one, sizeof(one), hipMemcpyHostToDevice, ctx->stream),
         "arm state");
  if (k < K) HIPCHK(hipMemsetAsync(a.pv, 0, nv * sizeof(float), ctx->stream), "memset");
  HIPCHK(hipMemcpy2DAsync(a.pv, K * sizeof(float), X, k * sizeof(float), k * sizeof(float), n,
                          hipMemcpyDeviceToDevice, ctx->stream), "pack input");
  for (int c = 1; c < 8; ++c) ctx->bytes_per_launch[c] = double(m) * double(n) * 4.0;
  HIPCHK(normal_apply32(ctx->stream, a, int(k), Q, track_cb, ctx), "normal apply");
  HIPCHK(hipStreamSynchronize(ctx->stream), "sync");
  HIPCHK(hipMemcpy(ctx->hstate, a.st, sizeof(PcgState), hipMemcpyDeviceToHost), "read state");
  ctx->stats.fused_stale_polls = int32_t(ctx->hstate->fstale);
  if (ctx->hstate->fabort != 0u) {
    const uint32_t zero = 0;
    (void)hipMemcpy(&a.st->fabort, &zero, sizeof(zero), hipMemcpyHostToDevice);
    return ctx->fail(MFS_ERR_HIP, "one-pass matvec: a wait for another block timed out");
  }
  return MFS_OK;
}

int mfs_dense_normal_apply_noalloc(mfs_ctx *ctx, const float *M, int64_t m, int64_t n,
                                   const float *X, int64_t k, double *Q) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if (M == nullptr || X == nullptr || Q == nullptr || m <= 0 || n <= 0 || k <= 0 || k > kMaxRhs)
    return ctx->fail(MFS_ERR_ARG, "dense_normal_apply_noalloc: bad arguments (1 <= k <= 16)");
  if (ctx->device < 0 || ctx->device >= kScratchDevices)
    return ctx->fail(MFS_ERR_ARG, "dense_normal_apply_noalloc: device index above 63");
  const ScratchSlot &slot = g_scratch[ctx->device];
  if (slot.ptr == nullptr)
    return ctx->fail(MFS_ERR_HIP, std::string("dense_normal_apply_noalloc: no scratch address: ") +
                                      hipGetErrorString(slot.err));
  const int K = batch_width(k);
  SolveArgs<float> a{};
  a.M = M;
  a.m = m;
  a.n = n;
  a.K = K;
  // the split count of mfs_dense_normal_apply, so both give the same sums
  a.row_blocks = row_blocks_for(m, n, ctx->cols_blocks,
                                int64_t(ctx->cus) * cols_resident_blocks(false, K, n % 4 == 0));
  auto up256 = [](int64_t b) { return (b + 255) / 256 * 256; };
  const int64_t pv_bytes = up256(n * K * 4), y_bytes = up256(m * K * 4);
  const int64_t part_bytes = int64_t(a.row_blocks) * n * K * 8;
  const int64_t need = pv_bytes + y_bytes + part_bytes;
  if (need > int64_t(kScratchBytes) || (n * K + kThreads - 1) / kThreads > INT32_MAX)
    return ctx->fail(MFS_ERR_ARG, "dense_normal_apply_noalloc: needs " + std::to_string(need) +
                                      " bytes of scratch; the static scratch holds " +
                                      std::to_string(kScratchBytes) + " (m x n about 2^32 / k at most)");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  char *base = static_cast<char *>(slot.ptr);
  a.pv = reinterpret_cast<float *>(base);
  a.y = reinterpret_cast<float *>(base + pv_bytes);
  a.part = reinterpret_cast<double *>(base + pv_bytes + y_bytes);
  a.pts = PointsOp{};
  a.st = nullptr;  // no batch state: the passes always run
  a.fused = FusedShape{};  // the two-pass matvec (the one-pass kernel needs exchange records)
  k_pack_rhs<<<unsigned((n * K + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(n, X, k, K,
                                                                                       a.pv);
  HIPCHK(hipGetLastError(), "pack input");
  HIPCHK(normal_apply32(ctx->stream, a, int(k), Q, nullptr, nullptr), "normal apply");
  return MFS_OK;
}

int mfs_dense_grad(mfs_ctx *ctx, const float *M, int64_t m, int64_t n, const float *F,
                   const float *A, const float *V, int64_t k, float *grad_M, float *grad_F) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if (M == nullptr || F == nullptr || A == nullptr || V == nullptr || m <= 0 || n <= 0 || k <= 0 ||
      k > kMaxRhs || m > int64_t(65535) * 32 || n > INT32_MAX)
    return ctx->fail(MFS_ERR_ARG, "dense_grad: bad arguments (1 <= k <= 16, m <= 2,097,120)");
  if (grad_M == nullptr && grad_F == nullptr) return MFS_OK;
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  // One row pass over M for [MA | MV] when 2k <= 8, else one each (the row
  // pass of batches of 8 runs at 7 TB/s, the one of batches of 16 at a third
  // of that: two passes of 8 are faster than one of 16).
  const bool one = 2 * k <= 8;
  const int Kx = one ? batch_width(2 * k) : batch_width(k);
  const int Ku = 2 * k <= 2 ? 2 : 2 * k <= 4 ? 4 : 2 * k <= 8 ? 8 : 2 * k <= 16 ? 16 : 32;
  const size_t xs = size_t(n) * Kx, ys = size_t(m) * Kx;
  const size_t floats = (one ? xs + ys : 2 * (xs + ys)) + size_t(m) * Ku + size_t(n) * Ku;
  HIPCHK(ctx->dgrad.ensure(floats * sizeof(float) + 256), "alloc gradient workspace");
  float *X = ctx->dgrad.as<float>(), *Y = X + xs;
  float *X2 = one ? X : Y + ys, *Y2 = one ? Y : X2 + xs;
  float *U = Y2 + ys, *W = U + size_t(m) * Ku;
  const unsigned gx = unsigned((n * Kx + kThreads - 1) / kThreads);
  if (one) {
    k_pack_av<<<gx, kThreads, 0, ctx->stream>>>(n, A, V, k, Kx, 0, X);
    HIPCHK(dense_rows32(ctx->stream, M, m, n, X, Kx, Y), "row pass");
  } else {
    k_pack_av<<<gx, kThreads, 0, ctx->stream>>>(n, A, V, k, Kx, 1, X);
    HIPCHK(dense_rows32(ctx->stream, M, m, n, X, Kx, Y), "row pass");
    k_pack_av<<<gx, kThreads, 0, ctx->stream>>>(n, A, V, k, Kx, 2, X2);
    HIPCHK(dense_rows32(ctx->stream, M, m, n, X2, Kx, Y2), "row pass");
  }
  const int64_t rows = std::max(m, n);
  k_grad_uw<<<unsigned((rows + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(
      m, n, k, F, Y, Kx, 0, one ? int(k) : 0, Y2, Kx, A, V, Ku, U, W, grad_F);
  HIPCHK(hipGetLastError(), "gradient factors");
  if (grad_M != nullptr) HIPCHK(outer_update32(ctx->stream, m, n, Ku, U, W, grad_M), "outer product");
  return MFS_OK;
}

int mfs_points_apply(mfs_ctx *ctx, int trans, int precision, const void *in, int64_t k,
                     void *out) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  int rc = check_points(ctx);
  if (rc != MFS_OK) return rc;
  if (in == nullptr || out == nullptr || k <= 0 || (trans != 0 && trans != 1) ||
      (precision != MFS_FP32 && precision != MFS_FP64))
    return ctx->fail(MFS_ERR_ARG, "points_apply: bad arguments");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  const PointsOp &op = ctx->pts;
  const size_t w = precision == MFS_FP64 ? 8 : 4;
  const int64_t rows_in = trans ? op.m : op.n, rows_out = trans ? op.n : op.m;
  const size_t part_rows = std::max(size_t(points_splits(op.m, op.n, op.target_blocks)) * size_t(op.m),
                                    size_t(points_splits(op.n, op.m, op.target_blocks)) * size_t(op.n));
  for (int64_t col0 = 0; col0 < k; col0 += kMaxRhs) {
    const int kval = int(std::min<int64_t>(kMaxRhs, k - col0));
    const int K = batch_width(kval);
    HIPCHK(ctx->dapin.ensure(size_t(rows_in) * K * w), "alloc");
    HIPCHK(ctx->dapout.ensure(size_t(rows_out) * K * w), "alloc");
    HIPCHK(ctx->dpart.ensure(part_rows * K * sizeof(double)), "alloc partials");
    if (kval < K) HIPCHK(hipMemsetAsync(ctx->dapin.p, 0, size_t(rows_in) * K * w, ctx->stream), "memset");
    HIPCHK(hipMemcpy2DAsync(ctx->dapin.p, K * w, static_cast<const char *>(in) + col0 * w, k * w,
                            kval * w, rows_in, hipMemcpyDeviceToDevice, ctx->stream),
           "pack input");
    double *part = ctx->dpart.as<double>();
    if (precision == MFS_FP32) {
      const float *xi = ctx->dapin.as<float>();
      float *yo = ctx->dapout.as<float>();
      if (trans == 0) {
        HIPCHK(points_rows<float>(ctx->stream, op, K, xi, nullptr, 0, 0, K, yo, part, nullptr), "rows");
      } else {
        HIPCHK(points_cols<float>(ctx->stream, op, K, xi, part, nullptr), "cols");
        HIPCHK(points_cols_reduce<float>(ctx->stream, op, K, part, yo), "cols reduce");
      }
    } else {
      const double *xi = ctx->dapin.as<double>();
      double *yo = ctx->dapout.as<double>();
      if (trans == 0) {
        HIPCHK(points_rows<double>(ctx->stream, op, K, xi, nullptr, 0, 0, K, yo, part, nullptr), "rows");
      } else {
        HIPCHK(points_cols<double>(ctx->stream, op, K, xi, part, nullptr), "cols");
        HIPCHK(points_cols_reduce<double>(ctx->stream, op, K, part, yo), "cols reduce");
      }
    }
    HIPCHK(hipMemcpy2DAsync(static_cast<char *>(out) + col0 * w, k * w, ctx->dapout.p, K * w,
                            kval * w, rows_out, hipMemcpyDeviceToDevice, ctx->stream),
           "unpack output");
  }
  return MFS_OK;
}

int mfs_assemble_dense(mfs_ctx *ctx, const double *x, const double *nx, const double *y, int64_t m,
                       int64_t n, float *M) {
  if (ctx == nullptr || x == nullptr || nx == nullptr || y == nullptr || M == nullptr || m <= 0 ||
      n <= 0 || m > INT32_MAX)
    return MFS_ERR_ARG;
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(launch_assemble(ctx->stream, x, nx, y, m, n, M), "assemble");
  return MFS_OK;
}

// ------------------------------------------------------------ caller steps

namespace {

bool caller_dims_ok(int64_t a, int64_t b) {
  return a >= 0 && b >= 0 && a <= INT32_MAX && b <= INT32_MAX;
}

// Reads the device index flag back (synchronises the context stream).
int caller_check_flag(mfs_ctx *ctx) {
  int32_t bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, ctx->dflag.p, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream),
         "read index flag");
  HIPCHK(hipStreamSynchronize(ctx->stream), "synchronize");
  if (bad) return ctx->fail(MFS_ERR_ARG, "solid_angle: face index outside [0, nv)");
  return MFS_OK;
}

}  // namespace

int mfs_solid_angle(mfs_ctx *ctx, int precision, const void *V, int64_t nv, const int32_t *F,
                    int64_t nf, const void *S, int64_t n, void *M) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if ((precision != MFS_FP32 && precision != MFS_FP64) || !caller_dims_ok(nv, nf) ||
      !caller_dims_ok(n, 0) || (nf > 0 && (F == nullptr || V == nullptr)) ||
      (n > 0 && S == nullptr) || (nf > 0 && n > 0 && M == nullptr))
    return ctx->fail(MFS_ERR_ARG, "solid_angle: bad arguments");
  if (nf > 0 && nv == 0) return ctx->fail(MFS_ERR_ARG, "solid_angle: faces without vertices");
  const int64_t blocks = (n + kThreads - 1) / kThreads * ((nf + 63) / 64);
  if (blocks > INT32_MAX) return ctx->fail(MFS_ERR_ARG, "solid_angle: too large");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(ctx->dcaller.ensure(solid_angle_scratch_bytes(nv, nf, n)), "alloc caller scratch");
  HIPCHK(ctx->dflag.ensure(256), "alloc");
  HIPCHK(hipMemsetAsync(ctx->dflag.p, 0, sizeof(int32_t), ctx->stream), "clear flag");
  int32_t *bad = ctx->dflag.as<int32_t>();
  if (precision == MFS_FP32)
    HIPCHK(solid_angle_fwd<float>(ctx->stream, static_cast<const float *>(V), nv, F, nf,
                                  static_cast<const float *>(S), n, static_cast<float *>(M),
                                  ctx->dcaller.p, bad), "solid_angle");
  else
    HIPCHK(solid_angle_fwd<double>(ctx->stream, static_cast<const double *>(V), nv, F, nf,
                                   static_cast<const double *>(S), n, static_cast<double *>(M),
                                   ctx->dcaller.p, bad), "solid_angle");
  return caller_check_flag(ctx);
}

int mfs_solid_angle_backward(mfs_ctx *ctx, int precision, const void *V, int64_t nv,
                             const int32_t *F, int64_t nf, const void *S, int64_t n,
                             const void *gM, void *gV, void *gS) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if ((precision != MFS_FP32 && precision != MFS_FP64) || !caller_dims_ok(nv, nf) ||
      !caller_dims_ok(n, 0) || (nf > 0 && (F == nullptr || V == nullptr)) ||
      (n > 0 && S == nullptr) || (nf > 0 && n > 0 && gM == nullptr))
    return ctx->fail(MFS_ERR_ARG, "solid_angle_backward: bad arguments");
  if (nf > 0 && nv == 0)
    return ctx->fail(MFS_ERR_ARG, "solid_angle_backward: faces without vertices");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(ctx->dcaller.ensure(solid_angle_scratch_bytes(nv, nf, n)), "alloc caller scratch");
  HIPCHK(ctx->dflag.ensure(256), "alloc");
  HIPCHK(hipMemsetAsync(ctx->dflag.p, 0, sizeof(int32_t), ctx->stream), "clear flag");
  int32_t *bad = ctx->dflag.as<int32_t>();
  if (precision == MFS_FP32)
    HIPCHK(solid_angle_bwd<float>(ctx->stream, static_cast<const float *>(V), nv, F, nf,
                                  static_cast<const float *>(S), n,
                                  static_cast<const float *>(gM), static_cast<float *>(gV),
                                  static_cast<float *>(gS), ctx->dcaller.p, bad),
           "solid_angle_backward");
  else
    HIPCHK(solid_angle_bwd<double>(ctx->stream, static_cast<const double *>(V), nv, F, nf,
                                   static_cast<const double *>(S), n,
                                   static_cast<const double *>(gM), static_cast<double *>(gV),
                                   static_cast<double *>(gS), ctx->dcaller.p, bad),
           "solid_angle_backward");
  return caller_check_flag(ctx);
}

int mfs_single_layer_apply(mfs_ctx *ctx, int precision, const void *S, int64_t n, const void *C,
                           int64_t nf, const void *X, int64_t k, void *phi) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if ((precision != MFS_FP32 && precision != MFS_FP64) || !caller_dims_ok(n, nf) ||
      !caller_dims_ok(k, 0) || (n > 0 && S == nullptr) || (nf > 0 && C == nullptr) ||
      (n > 0 && k > 0 && X == nullptr) || (nf > 0 && k > 0 && phi == nullptr))
    return ctx->fail(MFS_ERR_ARG, "single_layer_apply: bad arguments");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(ctx->dcaller.ensure(single_layer_scratch_bytes(nf, n, k)), "alloc caller scratch");
  if (precision == MFS_FP32)
    HIPCHK(single_layer_fwd<float>(ctx->stream, static_cast<const float *>(S), n,
                                   static_cast<const float *>(C), nf,
                                   static_cast<const float *>(X), k, static_cast<float *>(phi),
                                   ctx->dcaller.p), "single_layer_apply");
  else
    HIPCHK(single_layer_fwd<double>(ctx->stream, static_cast<const double *>(S), n,
                                    static_cast<const double *>(C), nf,
                                    static_cast<const double *>(X), k, static_cast<double *>(phi),
                                    ctx->dcaller.p), "single_layer_apply");
  return MFS_OK;
}

int mfs_single_layer_backward(mfs_ctx *ctx, int precision, const void *S, int64_t n,
                              const void *C, int64_t nf, const void *X, int64_t k,
                              const void *gphi, void *gS, void *gC, void *gX) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  if ((precision != MFS_FP32 && precision != MFS_FP64) || !caller_dims_ok(n, nf) ||
      !caller_dims_ok(k, 0) || (n > 0 && S == nullptr) || (nf > 0 && C == nullptr) ||
      (n > 0 && k > 0 && X == nullptr) || (nf > 0 && k > 0 && gphi == nullptr))
    return ctx->fail(MFS_ERR_ARG, "single_layer_backward: bad arguments");
  HIPCHK(hipSetDevice(ctx->device), "hipSetDevice");
  HIPCHK(ctx->dcaller.ensure(single_layer_scratch_bytes(nf, n, k)), "alloc caller scratch");
  if (precision == MFS_FP32)
    HIPCHK(single_layer_bwd<float>(ctx->stream, static_cast<const float *>(S), n,
                                   static_cast<const float *>(C), nf,
                                   static_cast<const float *>(X), k,
                                   static_cast<const float *>(gphi), static_cast<float *>(gS),
                                   static_cast<float *>(gC), static_cast<float *>(gX),
                                   ctx->dcaller.p), "single_layer_backward");
  else
    HIPCHK(single_layer_bwd<double>(ctx->stream, static_cast<const double *>(S), n,
                                    static_cast<const double *>(C), nf,
                                    static_cast<const double *>(X), k,
                                    static_cast<const double *>(gphi), static_cast<double *>(gS),
                                    static_cast<double *>(gC), static_cast<double *>(gX),
                                    ctx->dcaller.p), "single_layer_backward");
  return MFS_OK;
}

int mfs_ctx_set_param(mfs_ctx *ctx, int param, int64_t value) {
  if (ctx == nullptr || value < 0 || value > (int64_t(1) << 24)) return MFS_ERR_ARG;
  switch (param) {
    case MFS_PARAM_POINTS_BLOCKS:
      ctx->points_blocks = int32_t(value);
      ctx->pts.target_blocks = int32_t(value);
      return MFS_OK;
    case MFS_PARAM_COLS_BLOCKS:
      ctx->cols_blocks = int32_t(value);
      return MFS_OK;
    case MFS_PARAM_FUSED:
      if (value > 2) return ctx->fail(MFS_ERR_ARG, "set_param: fused mode is 0, 1 or 2");
      ctx->fused_mode = int32_t(value);
      return MFS_OK;
    case MFS_PARAM_TILE_SET:
      if (value > 0xff) return ctx->fail(MFS_ERR_ARG, "set_param: tile set is an 8-bit mask");
      ctx->tile_mask = value == 0 ? kAllTiles : uint32_t(value) | 1u;
      return MFS_OK;
    case MFS_PARAM_FACTOR_PHASES:
      if (value > kMaxPhases) return ctx->fail(MFS_ERR_ARG, "set_param: 1 to 8 factor column ranges");
      ctx->phases = value == 0 ? kDefaultPhases : int(value);
      return MFS_OK;
    case MFS_PARAM_FACTOR_SPLITS:
      if (value > 8) return ctx->fail(MFS_ERR_ARG, "set_param: 1 to 8 row splits");
      ctx->max_splits = value == 0 ? 1 : int(value);
      return MFS_OK;
    case MFS_PARAM_GRAM_SPLIT:
      if (value > 2) return ctx->fail(MFS_ERR_ARG, "set_param: Gram products are 0, 1 or 2");
      ctx->gram_split = value != 1;
      return MFS_OK;
    case MFS_PARAM_FACTOR_BANDS:
      if (value > 1) return ctx->fail(MFS_ERR_ARG, "set_param: factor bands are 0 (on) or 1 (off)");
      ctx->band_mode = value == 0;
      return MFS_OK;
    case MFS_PARAM_FACTOR_STREAMS:
      if (value > mfs_ctx::kSide) return ctx->fail(MFS_ERR_ARG, "set_param: 1 to 4 factor streams");
      ctx->side_streams = value == 0 ? mfs_ctx::kSide : int(value);
      return MFS_OK;
    default:
      return ctx->fail(MFS_ERR_ARG, "set_param: unknown parameter");
  }
}

int mfs_ctx_track_kernel(mfs_ctx *ctx, int mask) {
  if (ctx == nullptr) return MFS_ERR_ARG;
  track_fold(ctx);
  ctx->track_mask = mask;
  for (auto &t : ctx->tr) {
    t.launches = 0;
    t.ms = 0.0;
    t.bytes = 0.0;
  }
  return MFS_OK;
}

int mfs_ctx_tracked(mfs_ctx *ctx, int kernel_class, int64_t *launches, double *total_ms,
                    double *bytes) {
  if (ctx == nullptr || kernel_class <= 0 || kernel_class >= 8) return MFS_ERR_ARG;
  track_fold(ctx);
  const mfs_ctx::Track &t = ctx->tr[kernel_class];
  if (launches) *launches = t.launches;
  if (total_ms) *total_ms = t.ms;
  if (bytes) *bytes = t.bytes;
  return MFS_OK;
}

}  // extern "C"

// ============================================================================
// Batches of independent scenes (points form).
//
// The reference solves one scene per call.  A batch keeps one context per
// scene, each on a stream of its own, and
//   * plans the scenes on parallel host threads (the ordering is serial per
//     scene, the scenes are independent);
//   * factors the supernodes of all scenes together: one launch per tile-size
//     bucket over the (scene, supernode) items of every scene;
//   * runs the PCG iterations of all scenes in rounds (`chunk` iterations per
//     running scene on the scene's stream, ending in a gather of every
//     scene's active count on the batch's stream); converged scenes drop out
//     of the rounds.  Replaying the rounds from a captured hipGraph is
//     switched off (see mfs_batch_solve_points).
// ============================================================================

namespace {

__global__ void k_gather_active(PcgState *const *states, int32_t count, int32_t *out) {
  const int32_t i = int32_t(blockIdx.x) * kThreads + threadIdx.x;
  if (i < count) out[i] = states[i]->n_active;
}

}  // namespace

struct mfs_batch {
  int device = 0;
  hipStream_t caller = nullptr;  // work queued there before a call is waited for
  hipStream_t root = nullptr;    // the batch's own stream (fork, join, polls)
  std::string err;
  std::vector<mfs_ctx *> scenes;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> ev;     // one per scene (fork/join)
  hipEvent_t root_ev = nullptr;
  hipStream_t side[mfs_ctx::kSide] = {};
  hipEvent_t side_ev[mfs_ctx::kSide] = {};
  DevBuf dtable, ditems, dthg, dctr, dstates, dactive, dqb, dqnext;
  // 0: automatic (the queue launch for up to kQueueUnits units); 1: a launch
  // per bucket; 2: one launch over a device queue of all buckets' units
  int32_t factor_mode = 0;
  std::vector<FactorScene> table;
  std::vector<FactorItem> items;
  std::vector<PcgState *> states;
  int32_t *hactive = nullptr;  // pinned
  mfs_batch_stats stats{};

  int fail(int code, const std::string &msg) {
    err = msg;
    return code;
  }
};

namespace {

#define BHIPCHK(call, what)                                                  \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) return b->fail(MFS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_)); \
  } while (0)

// root waits for every scene stream
int join_scenes(mfs_batch *b) {
  for (size_t i = 0; i < b->scenes.size(); ++i) {
    BHIPCHK(hipEventRecord(b->ev[i], b->streams[i]), "record");
    BHIPCHK(hipStreamWaitEvent(b->root, b->ev[i], 0), "wait");
  }
  return MFS_OK;
}

// every scene stream waits for root
int fork_scenes(mfs_batch *b) {
  BHIPCHK(hipEventRecord(b->root_ev, b->root), "record");
  for (size_t i = 0; i < b->scenes.size(); ++i)
    BHIPCHK(hipStreamWaitEvent(b->streams[i], b->root_ev, 0), "wait");
  return MFS_OK;
}

int scene_error(mfs_batch *b, size_t i, int rc) {
  return b->fail(rc, "scene " + std::to_string(i) + ": " + b->scenes[i]->err);
}

}  // namespace

extern "C" {

int mfs_batch_create(mfs_batch **out, int device, void *stream, int32_t count) {
  if (out == nullptr || count <= 0) return MFS_ERR_ARG;
  *out = nullptr;
  mfs_batch *b = new (std::nothrow) mfs_batch();
  if (b == nullptr) return MFS_ERR_ALLOC;
  b->device = device;
  b->caller = static_cast<hipStream_t>(stream);
  bool ok = hipSetDevice(device) == hipSuccess &&
            hipStreamCreateWithFlags(&b->root, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->root_ev, hipEventDisableTiming) == hipSuccess &&
            hipHostMalloc(reinterpret_cast<void **>(&b->hactive), size_t(count) * 4) == hipSuccess;
  for (int i = 0; ok && i < mfs_ctx::kSide; ++i)
    ok = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->side_ev[i], hipEventDisableTiming) == hipSuccess;
  for (int32_t i = 0; ok && i < count; ++i) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    mfs_ctx *c = nullptr;
    ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
    if (ok) b->streams.push_back(s);
    ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (ok) b->ev.push_back(e);
    ok = ok && mfs_ctx_create(&c, device, s) == MFS_OK;
    if (ok) b->scenes.push_back(c);
  }
  if (!ok) {
    mfs_batch_destroy(b);
    return MFS_ERR_HIP;
  }
  *out = b;
  return MFS_OK;
}

void mfs_batch_destroy(mfs_batch *b) {
  if (b == nullptr) return;
  (void)hipSetDevice(b->device);
  for (hipStream_t s : b->streams) (void)hipStreamSynchronize(s);
  for (mfs_ctx *c : b->scenes) mfs_ctx_destroy(c);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  for (int i = 0; i < mfs_ctx::kSide; ++i) {
    if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
    if (b->side_ev[i]) (void)hipEventDestroy(b->side_ev[i]);
  }
  if (b->root_ev) (void)hipEventDestroy(b->root_ev);
  if (b->root) (void)hipStreamDestroy(b->root);
  if (b->hactive) (void)hipHostFree(b->hactive);
  for (DevBuf *d : {&b->dtable, &b->ditems, &b->dthg, &b->dctr, &b->dstates, &b->dactive}) d->release();
  delete b;
}

mfs_ctx *mfs_batch_scene(mfs_batch *b, int32_t i) {
  if (b == nullptr || i < 0 || size_t(i) >= b->scenes.size()) return nullptr;
  return b->scenes[size_t(i)];
}

const char *mfs_batch_last_error(const mfs_batch *b) {
  return b == nullptr ? "null batch" : b->err.c_str();
}

int mfs_batch_stats_get(const mfs_batch *b, mfs_batch_stats *out) {
  if (b == nullptr || out == nullptr) return MFS_ERR_ARG;
  *out = b->stats;
  return MFS_OK;
}

int mfs_batch_plan(mfs_batch *b, const double *const *xyz, const int64_t *n, double rho,
                   double lambda, int32_t max_super, int32_t threads) {
  if (b == nullptr || xyz == nullptr || n == nullptr) return MFS_ERR_ARG;
  BHIPCHK(hipStreamSynchronize(b->caller), "sync caller");
  const auto t0 = std::chrono::steady_clock::now();
  const size_t count = b->scenes.size();
  const int nt = std::max<int>(1, std::min<int>(threads > 0 ? threads : 1, int(count)));
  std::atomic<size_t> next{0};
  std::vector<int> rcs(count, MFS_OK);
  auto work = [&]() {
    for (;;) {
      const size_t i = next.fetch_add(1);
      if (i >= count) return;
      mfs_ctx *c = b->scenes[i];
      if (xyz[i] == nullptr) {
        rcs[i] = c->fail(MFS_ERR_ARG, "plan: null points");
        continue;
      }
      std::vector<double> pts(xyz[i], xyz[i] + size_t(std::max<int64_t>(n[i], 0)) * 3);
      rcs[i] = plan_impl(c, std::move(pts), n[i], rho, lambda, max_super);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  for (size_t i = 0; i < count; ++i)
    if (rcs[i] != MFS_OK) return scene_error(b, i, rcs[i]);
  b->stats.plan_ms = ms_since(t0);
  return MFS_OK;
}

int mfs_batch_set_factor_mode(mfs_batch *b, int32_t mode) {
  if (b == nullptr) return MFS_ERR_ARG;
  if (mode < 0 || mode > 2)
    return b->fail(MFS_ERR_ARG, "factor mode is 0 (automatic), 1 (a launch per bucket) or 2 (one launch)");
  b->factor_mode = mode;
  return MFS_OK;
}

int mfs_batch_factor_points(mfs_batch *b, const double *const *x, const double *const *nx,
                            const int64_t *m, const double *const *y, int precision) {
  if (b == nullptr || x == nullptr || nx == nullptr || m == nullptr || y == nullptr)
    return MFS_ERR_ARG;
  if (precision != MFS_FP64 && precision != MFS_FP32)
    return b->fail(MFS_ERR_ARG, "factor: precision must be MFS_FP64 or MFS_FP32");
  BHIPCHK(hipSetDevice(b->device), "hipSetDevice");
  BHIPCHK(hipStreamSynchronize(b->caller), "sync caller");
  const auto t0 = std::chrono::steady_clock::now();
  const size_t count = b->scenes.size();
  // 1. per scene: pack the points, clear statuses, scene record
  for (size_t i = 0; i < count; ++i) {
    mfs_ctx *c = b->scenes[i];
    if (!c->has_plan) return b->fail(MFS_ERR_STATE, "scene " + std::to_string(i) + ": no plan");
    if (x[i] == nullptr || nx[i] == nullptr || y[i] == nullptr || m[i] <= 0 || m[i] > INT32_MAX)
      return b->fail(MFS_ERR_ARG, "scene " + std::to_string(i) + ": bad points");
    const int64_t ni = c->n;
    c->has_points = false;
    BHIPCHK(c->dt32.ensure(size_t(m[i]) * kTgt32 * 4), "alloc points");
    BHIPCHK(c->dt64.ensure(size_t(m[i]) * kTgt64 * 8), "alloc points");
    BHIPCHK(c->ds32.ensure(size_t(ni) * kSrc32 * 4), "alloc points");
    BHIPCHK(c->ds64.ensure(size_t(ni) * kSrc64 * 8), "alloc points");
    BHIPCHK(pack_points(c->stream, x[i], nx[i], y[i], m[i], ni, c->dt32.as<float>(),
                        c->dt64.as<double>(), c->ds32.as<float>(), c->ds64.as<double>()),
            "pack points");
    c->pts = PointsOp{c->dt32.as<float>(), c->ds32.as<float>(), c->dt64.as<double>(),
                      c->ds64.as<double>(), m[i], ni, c->points_blocks};
    c->has_points = true;
    const int rc = factor_prepare(c, nullptr, &c->pts, m[i], ni, precision);
    if (rc != MFS_OK) return scene_error(b, i, rc);
  }
  // 2. merged buckets: every scene's supernodes of one tile size in one launch
  // per row-split count.  A scene's supernodes get the split count they get
  // when the scene is factored alone (factor_splits of its bucket size and
  // target rows), so the batch's factor equals the single-scene one bit for
  // bit.
  b->table.resize(count);
  for (size_t i = 0; i < count; ++i) b->table[i] = b->scenes[i]->fscene;
  b->items.clear();
  struct BLaunch {
    int tiles;
    int32_t offset, count;
    FactorSplit sp;
    size_t off, ctr;
  };
  std::vector<BLaunch> buckets;
  size_t scratch_bytes = 8, ctrs = 1;
  const int64_t target = int64_t(b->scenes[0]->cus) * 4;
  for (int t : kTileSizes) {
    for (int32_t S = 1; S <= 8; ++S) {
      const int32_t off = int32_t(b->items.size());
      for (size_t i = 0; i < count; ++i)
        for (const Bucket &bk : b->scenes[i]->buckets) {
          if (bk.tiles != t) continue;
          const int32_t si = std::min(b->scenes[i]->max_splits, factor_splits(bk.count, m[i], target));
          if (si != S) continue;
          for (int32_t q = 0; q < bk.count; ++q)
            b->items.push_back({int32_t(i), b->scenes[i]->list[size_t(bk.offset + q)]});
        }
      const int32_t cnt = int32_t(b->items.size()) - off;
      if (cnt == 0) continue;
      BLaunch L{t, off, cnt, FactorSplit{}, scratch_bytes, ctrs};
      L.sp.splits = S;
      const size_t ns = size_t(16) * t, blocks = size_t(cnt) * S;
      if (t > 8) scratch_bytes += blocks * size_t(factor_scratch_elems(t)) * 8;
      else if (S > 1) scratch_bytes += blocks * ns * ns * 8;
      scratch_bytes = (scratch_bytes + 255) / 256 * 256;
      if (S > 1) ctrs += size_t(cnt);
      buckets.push_back(L);
    }
  }
  BHIPCHK(b->dtable.ensure(count * sizeof(FactorScene)), "alloc");
  BHIPCHK(b->ditems.ensure(std::max<size_t>(b->items.size(), 1) * sizeof(FactorItem)), "alloc");
  BHIPCHK(b->dthg.ensure(scratch_bytes), "alloc factor scratch");
  BHIPCHK(b->dctr.ensure(ctrs * 4), "alloc factor counters");
  if (join_scenes(b) != MFS_OK) return MFS_ERR_HIP;
  BHIPCHK(hipMemsetAsync(b->dctr.p, 0, ctrs * 4, b->root), "clear factor counters");
  BHIPCHK(hipMemcpyAsync(b->dtable.p, b->table.data(), count * sizeof(FactorScene),
                         hipMemcpyHostToDevice, b->root),
          "upload");
  if (!b->items.empty())
    BHIPCHK(hipMemcpyAsync(b->ditems.p, b->items.data(), b->items.size() * sizeof(FactorItem),
                           hipMemcpyHostToDevice, b->root),
            "upload");
  for (BLaunch &L : buckets)
    if (L.sp.splits > 1) {
      L.sp.counters = b->dctr.as<int32_t>() + L.ctr;
      if (L.tiles <= 8) L.sp.partial = b->dthg.as<char>() + L.off;
    }
  const int prec_code = precision == MFS_FP64 ? 0 : 1;
  // Split products in FP32 when every scene asks for them (one launch serves
  // them all).
  bool split = prec_code == 1;
  for (size_t i = 0; i < count; ++i) split = split && b->scenes[i]->gram_split;
  // One launch: every bucket's units (largest tiles first) from a device
  // queue, as many blocks as stay resident with the LDS of the largest tile
  // count present.  Its kernel carries every tile count's code, so it has the
  // registers of the widest (all 256 and accumulators besides, against
  // 32-215 for the tiles' own kernels) and fewer blocks per CU: measured on
  // the uneven batch's scenes, 12.2 ms against 15.2 ms for 24 scenes (17,416
  // supernodes, 4 row splits) but 817 ms against 428 ms for 256 scenes
  // (697,733).  Automatic mode takes it up to kQueueUnits units; a launch per
  // bucket otherwise, or when no block fits.
  constexpr int64_t kQueueUnits = 32768;
  int queue_blocks = 0;
  size_t queue_lds = 0;
  std::vector<FactorQueueBucket> qb;
  int64_t units = 0;
  for (const BLaunch &L : buckets) {
    qb.push_back({L.tiles, L.offset, L.count, int32_t(units), b->dthg.as<char>() + L.off, L.sp});
    units += int64_t(L.count) * L.sp.splits;
    queue_lds = std::max(queue_lds, factor_unit_lds(L.tiles, prec_code));
  }
  const bool use_queue = !buckets.empty() && units < INT32_MAX &&
                         (b->factor_mode == 2 || (b->factor_mode == 0 && units <= kQueueUnits));
  if (use_queue) queue_blocks = factor_queue_blocks(prec_code, split, queue_lds, b->scenes[0]->cus);
  if (queue_blocks > 0) {
    BHIPCHK(b->dqb.ensure(qb.size() * sizeof(FactorQueueBucket)), "alloc");
    BHIPCHK(b->dqnext.ensure(4), "alloc");
    BHIPCHK(hipMemcpyAsync(b->dqb.p, qb.data(), qb.size() * sizeof(FactorQueueBucket),
                           hipMemcpyHostToDevice, b->root),
            "upload");
    BHIPCHK(hipMemsetAsync(b->dqnext.p, 0, 4, b->root), "clear queue");
    BHIPCHK(launch_factor_queue(b->root, prec_code, b->dtable.as<FactorScene>(), b->ditems.as<FactorItem>(),
                                b->dqb.as<FactorQueueBucket>(), int32_t(qb.size()), int32_t(units),
                                b->dqnext.as<int32_t>(), queue_lds, queue_blocks, split),
            "supernode factor queue");
  } else {
    BHIPCHK(hipEventRecord(b->root_ev, b->root), "fork");
    const int nside = std::min<int>(mfs_ctx::kSide, int(buckets.size()));
    for (int i = 0; i < nside; ++i) BHIPCHK(hipStreamWaitEvent(b->side[i], b->root_ev, 0), "fork");
    for (size_t bi = 0; bi < buckets.size(); ++bi) {
      const BLaunch &L = buckets[bi];
      BHIPCHK(launch_super_factor(b->side[bi % nside], L.tiles, prec_code, true, b->dtable.as<FactorScene>(),
                                  b->ditems.as<FactorItem>() + L.offset, L.count, b->dthg.as<char>() + L.off,
                                  L.sp, split),
              "supernode factor");
    }
    for (int i = 0; i < nside; ++i) {
      BHIPCHK(hipEventRecord(b->side_ev[i], b->side[i]), "join");
      BHIPCHK(hipStreamWaitEvent(b->root, b->side_ev[i], 0), "join");
    }
  }
  if (fork_scenes(b) != MFS_OK) return MFS_ERR_HIP;
  // 3. per scene: statuses, FP64 retries, transposed values
  for (size_t i = 0; i < count; ++i) {
    const int rc = factor_finish(b->scenes[i], true, precision);
    if (rc != MFS_OK) return scene_error(b, i, rc);
  }
  b->stats.factor_ms = ms_since(t0);
  b->stats.factor_launches = queue_blocks > 0 ? 1 : int32_t(buckets.size());
  b->stats.factor_items = int32_t(b->items.size());
  return MFS_OK;
}

int mfs_batch_solve_points(mfs_batch *b, const float *const *B, const int64_t *k, float *const *X,
                           int32_t maxit, double tol, double inner_tol, int32_t chunk,
                           int use_graph) {
  if (b == nullptr || B == nullptr || k == nullptr || X == nullptr || chunk <= 0) return MFS_ERR_ARG;
  // use_graph 1 -- one multi-stream capture of all scenes' rounds, forked and
  // joined with events and re-captured as scenes converged -- is switched
  // off: a full test run once ended in a host segmentation fault inside this
  // call in that mode, with no cause found in this code.  use_graph 2 replays
  // per-scene graphs: each scene's `chunk` iterations are captured from its
  // own stream alone (no event is recorded or waited on inside a capture),
  // instantiated once per sweep precision, launched on that stream every
  // round and destroyed only after every stream has been synchronized.
  if (use_graph == 1)
    return b->fail(MFS_ERR_ARG, "batch solve: graph replay is switched off; pass use_graph = 0 or 2");
  if (use_graph != 0 && use_graph != 2) return b->fail(MFS_ERR_ARG, "batch solve: use_graph is 0 or 2");
  if (!(inner_tol > 0) || !(inner_tol <= 1)) inner_tol = 0.5;
  BHIPCHK(hipSetDevice(b->device), "hipSetDevice");
  BHIPCHK(hipStreamSynchronize(b->caller), "sync caller");
  const auto t0 = std::chrono::steady_clock::now();
  const size_t count = b->scenes.size();
  std::vector<MixedRun> runs(count);
  b->states.resize(count);
  for (size_t i = 0; i < count; ++i) {
    mfs_ctx *c = b->scenes[i];
    if (!c->has_points || !c->has_factor)
      return b->fail(MFS_ERR_STATE, "scene " + std::to_string(i) + ": factor the points first");
    if (B[i] == nullptr || X[i] == nullptr || k[i] <= 0 || k[i] > kMaxRhs)
      return b->fail(MFS_ERR_ARG, "scene " + std::to_string(i) + ": 1..16 right-hand sides");
    c->track_mask = 0;
    const int rc = runs[i].begin(c, nullptr, &c->pts, c->pts.m, c->n, B[i], k[i], 0, int(k[i]),
                                 maxit, tol, inner_tol, nullptr, nullptr);
    if (rc != MFS_OK) return scene_error(b, i, rc);
    b->states[i] = runs[i].a32.st;
  }
  BHIPCHK(b->dstates.ensure(count * sizeof(PcgState *)), "alloc");
  BHIPCHK(b->dactive.ensure(count * 4), "alloc");
  BHIPCHK(hipMemcpyAsync(b->dstates.p, b->states.data(), count * sizeof(PcgState *),
                         hipMemcpyHostToDevice, b->root),
          "upload");
  // Per-scene graphs of `chunk` iterations (use_graph 2): [scene][precision],
  // precision 0 for the FP32 sweeps, 1 for the FP64 ones.  Kernel arguments
  // are fixed for the whole call (device state carries the iteration).
  std::vector<std::array<hipGraphExec_t, 2>> execs(count, {nullptr, nullptr});
  auto destroy_graphs = [&]() {
    for (size_t i = 0; i < count; ++i) (void)hipStreamSynchronize(b->streams[i]);
    for (auto &pair : execs)
      for (hipGraphExec_t &e : pair)
        if (e != nullptr) {
          (void)hipGraphExecDestroy(e);
          e = nullptr;
        }
  };
  auto scene_graph = [&](size_t i, hipGraphExec_t *out) -> int {
    hipStream_t s = b->streams[i];
    BHIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    hipError_t e = hipSuccess;
    for (int32_t c = 0; c < chunk && e == hipSuccess; ++c) e = runs[i].step(s, nullptr, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (e == hipSuccess) e = ec;
    size_t nodes = 0;
    if (e == hipSuccess) e = hipGraphGetNodes(g, nullptr, &nodes);
    if (e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    if (g != nullptr) (void)hipGraphDestroy(g);
    BHIPCHK(e, "capture scene iterations");
    ++b->stats.graph_captures;
    b->stats.graph_nodes = std::max<int32_t>(b->stats.graph_nodes, int32_t(nodes));
    return MFS_OK;
  };
  // Iterations: `chunk` per scene on the scene's stream (direct launches or
  // the scene's graph), for the scenes in `set`, then a gather of every
  // scene's active count on root.
  auto launch_chunk = [&](const std::vector<size_t> &set) -> int {
    if (fork_scenes(b) != MFS_OK) return MFS_ERR_HIP;
    for (size_t i : set) {
      if (use_graph == 2) {
        hipGraphExec_t &ex = execs[i][runs[i].fp64_inner ? 1 : 0];
        if (ex == nullptr) {
          const int rc = scene_graph(i, &ex);
          if (rc != MFS_OK) return rc;
        }
        BHIPCHK(hipGraphLaunch(ex, b->streams[i]), "graph launch");
      } else {
        for (int32_t c = 0; c < chunk; ++c)
          BHIPCHK(runs[i].step(b->streams[i], nullptr, nullptr), "pcg iteration");
      }
    }
    if (join_scenes(b) != MFS_OK) return MFS_ERR_HIP;
    k_gather_active<<<unsigned((count + kThreads - 1) / kThreads), kThreads, 0, b->root>>>(
        b->dstates.as<PcgState *>(), int32_t(count), b->dactive.as<int32_t>());
    BHIPCHK(hipGetLastError(), "gather");
    return MFS_OK;
  };
  b->stats.graph_nodes = 0;
  b->stats.graph_captures = 0;
  int32_t replays = 0, sweeps = 0;
  int rc = MFS_OK;
  for (;;) {
    for (size_t i = 0; i < count; ++i) BHIPCHK(hipStreamSynchronize(b->streams[i]), "sync");
    std::vector<size_t> active;
    for (size_t i = 0; i < count; ++i) {
      if (runs[i].decide()) {
        if ((rc = runs[i].start()) != MFS_OK) break;
        active.push_back(i);
      } else if (runs[i].fault != MFS_OK) {
        rc = scene_error(b, i, runs[i].fault);
        break;
      }
    }
    if (rc != MFS_OK || active.empty()) break;
    ++sweeps;
    int32_t budget = 0;
    for (size_t i : active) budget = std::max(budget, maxit - runs[i].inner_used);
    std::vector<size_t> running = active;
    for (int32_t it = 0;; it += chunk) {
      if (it > budget + 64) {
        rc = b->fail(MFS_ERR_STATE, "batch pcg: iteration bookkeeping diverged");
        break;
      }
      if ((rc = launch_chunk(running)) != MFS_OK) break;
      ++replays;
      hipError_t e = hipMemcpyAsync(b->hactive, b->dactive.p, count * 4, hipMemcpyDeviceToHost, b->root);
      if (e == hipSuccess) e = hipStreamSynchronize(b->root);
      if (e != hipSuccess) { rc = b->fail(MFS_ERR_HIP, std::string("poll: ") + hipGetErrorString(e)); break; }
      std::vector<size_t> still;
      for (size_t i : running)
        if (b->hactive[i] != 0) still.push_back(i);
      running.swap(still);
      if (running.empty()) break;
    }
    if (rc != MFS_OK) break;
    if (fork_scenes(b) != MFS_OK) { rc = MFS_ERR_HIP; break; }
    for (size_t i : active) {
      mfs_ctx *c = b->scenes[i];
      hipError_t e = hipMemcpyAsync(c->hstate, runs[i].a32.st, sizeof(PcgState), hipMemcpyDeviceToHost,
                                    c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) { rc = b->fail(MFS_ERR_HIP, std::string("state: ") + hipGetErrorString(e)); break; }
      if ((rc = runs[i].end()) != MFS_OK) { rc = scene_error(b, i, rc); break; }
    }
    if (rc != MFS_OK) break;
  }
  destroy_graphs();
  if (rc != MFS_OK) return rc;
  for (size_t i = 0; i < count; ++i) {
    if ((rc = runs[i].finish(X[i], nullptr)) != MFS_OK) return scene_error(b, i, rc);
    b->scenes[i]->stats.outer_iters = runs[i].outer;
    b->scenes[i]->stats.fp64_sweeps = runs[i].fp64_sweeps;
    b->scenes[i]->stats.reverted_columns = runs[i].reverted;
  }
  for (size_t i = 0; i < count; ++i) BHIPCHK(hipStreamSynchronize(b->streams[i]), "sync");
  b->stats.solve_ms = ms_since(t0);
  b->stats.graph_replays = replays;
  b->stats.sweeps = sweeps;
  return MFS_OK;
}

}  // extern "C"
