// k_fused.hip -- q = M^T (M p) in one HBM pass over M (dense form, FP32 batch of 8).
//
// The reference applies the normal operator as two cuBLAS GEMVs per right-hand
// side (src/kl_chol.cu:928-936 then :992-1002), so M (m x n, 17 GB at
// m = n = 65,536 in FP32) crosses HBM twice per PCG iteration.  Here every
// element of M is loaded once: the lane that loads it keeps it in registers
// until y = M p of its row is complete over all columns, then multiplies it by
// y for its column of M^T y.
//
// Grid.  The columns are cut into S slices (S a power of two, at least 2, at
// most 1024 columns each); the rows into NG groups of NPg panels of 16 rows.
// Block (g, s) = blockIdx.x (s = blockIdx.x mod S) streams the panels of
// group g over slice s.  A block is 8 waves, two per SIMD; wave w owns 128
// columns of the slice.  One 16-byte load per lane covers two rows of the
// wave's panel (lanes 0-31 the even row, 32-63 the odd one, four columns
// each), so a lane holds four columns of 8 rows of every panel.
//
// Step j of a wave (register ring of kTiles panels: j - kDefer .. j + 1):
//   (1) its exchange duty for step j, if any, issues its record loads;
//   (2) the loads of panel j + 1 are issued (non-temporal);
//   (3) panel j arrives; the duty consumes its records;
//   (4) phase A: the 16 x 128 tile goes, one 64-column half at a time,
//       through a wave-private LDS transpose (rows padded to 68 floats:
//       conflict-free both ways) into 16 v_mfma_f32_16x16x4f32 per half with
//       p as the B operand (rows x 8 right-hand sides, padded to 16); the
//       wave's partial y is staged in LDS and the last wave of the block to
//       stage sums the 8 partials in wave order and publishes them;
//   (5) phase B on panel j - kDefer: y of its rows from the LDS y ring
//       (broadcast reads) times the wave's registers, 256 VALU FMAs per lane
//       into FP32 sums that are added into FP64 (LDS) every 8 panels.
// Loads of a wave complete in order, so every record load is issued before
// the prefetch of panel j + 1: waiting for the records never waits for M.
// FP32 MFMA and VALU work of one SIMD do not overlap on this GPU.
//
// Cross-block exchange (tagged 8-byte records, device-coherent stores and
// loads; tag = (launch epoch << 16) | (panel + 1)):
//   publish  panel j, at step j: block partials of the 128 values
//            (value v = 16 rhs + row) -> prec[g][j mod kRing][v][s];
//   reduce   panel j, from step j + kRedLag on (a block's panels in order,
//            up to two per step; a panel with partials still missing waits
//            for a later step, until the step before its fetch, which polls
//            for them): block s sums the S partials of its 128 / S
//            values (two slices per lane, then a fixed butterfly over the
//            value's lanes: deterministic, the same bits in every block) and
//            writes y[g][j mod kRing][v];
//   fetch    panel j, at step j + kDefer: the 128 y records -> LDS y ring
//            (zero for rows past m), flag for the waves.
// A step is 64 KB of M per block (about 4 us at 65,536^2).  A fetch whose
// records are still stale polls again (that wait also drains the wave's
// prefetch); the state counts these polls and the reduce attempts left to a
// later step.  Every wait is bounded: after kSpinLimit polls the wave sets
// PcgState::fabort and stops waiting, the host reports the failure.  Panel
// j's records are rewritten at panel j + kRing; by then every block has
// reduced and fetched panel j (a block at step j + kRing used
// y(j + kRing - kDefer - 1), which needed every block's partials of that
// panel, so every block is past step j + kDefer).
//
// Residency.  The blocks wait for each other, so all S x NG of them must be
// on the GPU at once: the grid is at most one block per CU (the LDS holds one
// block per CU) and fused_shape checks the occupancy.  The launch is an
// ordinary one, and the one-pass launches of this process on a device are
// ordered one after another whatever stream they are queued on, so two such
// grids never share the CUs while each waits for blocks of its own.
// Measured at m = n = 65,536: 4.19-4.21 ms per matvec (4.1 TB/s of M), the
// two passes 2 x 2.45 ms; streaming M alone takes about 2.5 ms, the
// exchange's waits (y of a panel needs the slowest block of the group) most
// of the rest.
#include <map>
#include <mutex>

#include "mfs_internal.h"

namespace mfs {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 8;                        // waves per block (two per SIMD)
constexpr int kThreadsB = kWaves * 64;
constexpr int kWaveCols = 128;                   // columns per wave (four per lane)
constexpr int kSliceCols = kWaves * kWaveCols;   // at most 1024 columns per block
constexpr int kPanel = 16;                       // rows per panel (the MFMA's 16 rows)
constexpr int kLoads = kPanel / 2;               // 16-byte loads per lane per panel
constexpr int kDefer = 3;                        // panel j's y is used at step j + kDefer
constexpr int kPF = 1;                           // panels in flight ahead of panel j
constexpr int kTiles = kDefer + kPF + 1;         // register tiles: j - kDefer .. j + kPF
constexpr int kRedLag = 2;                       // panel j reduced from step j + kRedLag on
constexpr int kRing = 8;                         // record rings (panels)
constexpr int kYSlots = 8;                       // LDS y ring (panels)
constexpr int kStSlots = kDefer + 1;             // LDS staging ring (panels)
constexpr int kFlush = 8;                        // panels between FP64 flushes of q
constexpr int kMaxSlices = 64;
constexpr uint32_t kSpinLimit = 1u << 22;        // polls before giving up

// LDS (bytes): wave-private transposes, p of the slice, FP64 sums of q,
// staging of the waves' partials, y ring, sync words.
constexpr int kTrRow = 68;                                    // floats per transposed row (64 + 4)
constexpr int kTrBytes = kWaves * kPanel * kTrRow * 4;        // 34 KB (one 64-column half at a time)
constexpr int kPBytes = kSliceCols * 8 * 4;                   // 32 KB
constexpr int kQLanes = kWaves * 32;                          // lanes holding q after the halves combine
constexpr int kQBytes = 32 * kQLanes * 8;                     // 64 KB
constexpr int kStBytes = kStSlots * kWaves * 8 * kPanel * 4;  // 16 KB
constexpr int kYBytes = kYSlots * kPanel * 8 * 4;             // 4 KB
constexpr int kPOff = kTrBytes, kQOff = kPOff + kPBytes, kStOff = kQOff + kQBytes,
              kYOff = kStOff + kStBytes, kSyncOff = kYOff + kYBytes;
constexpr int kSyncWords = kStSlots + kYSlots + 2;            // arrivals, y flags, abort, reduced panels
constexpr int kLdsBytes = kSyncOff + kSyncWords * 4;
static_assert(kLdsBytes <= 160 * 1024, "LDS");
static_assert(kRing >= kDefer + 2 && kYSlots > kDefer + 1 && kRedLag < kDefer, "rings");
static_assert((kRing & (kRing - 1)) == 0 && (kYSlots & (kYSlots - 1)) == 0, "rings");

__device__ inline v4f mfma16(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline uint64_t ld_rec(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_rec(uint64_t *p, float v, uint32_t tag) {
  const uint64_t r = uint64_t(__float_as_uint(v)) | (uint64_t(tag) << 32);
  __hip_atomic_store(p, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float rec_val(uint64_t r) { return __uint_as_float(uint32_t(r)); }
__device__ inline uint32_t rec_tag(uint64_t r) { return uint32_t(r >> 32); }
__device__ inline uint32_t lds_get(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_set(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline v4f load_stream4(const float *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
}

__global__ __launch_bounds__(kThreadsB, 1) void k_mv_fused8(
    const float *__restrict__ M, int64_t m, int64_t n, const float *__restrict__ X,
    double *__restrict__ part, uint64_t *__restrict__ prec, uint64_t *__restrict__ yrec,
    PcgState *st, int32_t lgS, int32_t width, int32_t NPg) {
  if (st->n_active == 0) return;  // the batch has converged: uniform early exit
  __shared__ __align__(16) char lds[kLdsBytes];
  float *tr_all = reinterpret_cast<float *>(lds);
  float *pl = reinterpret_cast<float *>(lds + kPOff);
  double *q64 = reinterpret_cast<double *>(lds + kQOff);
  float *stg = reinterpret_cast<float *>(lds + kStOff);
  float *yr = reinterpret_cast<float *>(lds + kYOff);
  uint32_t *sync = reinterpret_cast<uint32_t *>(lds + kSyncOff);
  uint32_t *cnt = sync, *yflag = sync + kStSlots, *labort = sync + kStSlots + kYSlots;
  uint32_t *red_done = labort + 1;  // panels this block has reduced (a prefix)

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int S = 1 << lgS;
  const int s = int(blockIdx.x) & (S - 1), g = int(blockIdx.x) >> lgS;
  const int64_t c0 = int64_t(s) * width;
  const int64_t cend = c0 + width < n ? c0 + width : n;    // one past the slice's last column
  const int64_t c0w = c0 + kWaveCols * w;
  const int64_t row0 = int64_t(g) * NPg * kPanel;
  const uint32_t epoch = st->fepoch;
  auto tag_of = [&](int32_t j) { return (epoch << 16) | uint32_t(j + 1); };

  // ------------------------------------------------------------ set-up
  if (tid < kSyncWords) sync[tid] = 0u;
  for (int i = tid; i < 32 * kQLanes; i += kThreadsB) q64[i] = 0.0;
  // p of the slice as [local column / 4][rhs][local column % 4] (phase A's
  // B operands: four columns of one right-hand side per 16-byte read), zero
  // past the slice (those columns belong to the next block or lie past n)
  for (int i = tid; i < kSliceCols * 8; i += kThreadsB) {
    const int cl = i >> 3, r = i & 7;
    const int64_t col = c0 + cl;
    pl[((cl >> 2) * 8 + r) * 4 + (cl & 3)] = col < cend ? X[col * 8 + r] : 0.f;
  }
  __syncthreads();

  bool dead = false;
  uint32_t stale = 0;
  auto give_up = [&](uint32_t &spins) {
    if (++spins > kSpinLimit || ((spins & 255u) == 0u &&
                                 (lds_get(labort) != 0u ||
                                  __hip_atomic_load(&st->fabort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u))) {
      __hip_atomic_store(&st->fabort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      lds_set(labort, 1u);
      dead = true;
      return true;
    }
    __builtin_amdgcn_s_sleep(1);
    return false;
  };

  // ------------------------------------------------------------ per-lane columns and rows
  const int half = lane >> 5;                                  // the lane's row of a pair
  const int64_t ccol = c0w + 4 * (lane & 31);                  // the lane's four columns
  // columns past n are clamped (their sums are not written); a load takes a
  // scalar row base and a 32-bit lane offset (odd rows: one row further)
  const uint32_t cbyte = uint32_t(ccol + 4 <= n ? ccol : n - 4) * 4u;
  const uint32_t obyte = cbyte + uint32_t(half) * uint32_t(n) * 4u;
  auto load_panel = [&](v4f (&T)[kLoads], int32_t j) {
    const int32_t jc = j < NPg ? j : NPg - 1;
    const int64_t rb = row0 + int64_t(jc) * kPanel;
    if (rb + kPanel <= m) {
#pragma unroll
      for (int u = 0; u < kLoads; ++u)
        T[u] = load_stream4(reinterpret_cast<const float *>(
            reinterpret_cast<const char *>(M + (rb + 2 * u) * n) + obyte));
    } else {  // the last panel: rows past m are clamped (their y is zero)
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const int64_t r = rb + 2 * u + half < m ? rb + 2 * u + half : m - 1;
        T[u] = load_stream4(reinterpret_cast<const float *>(
            reinterpret_cast<const char *>(M + r * n) + cbyte));
      }
    }
  };

  // exchange addressing
  const int VPB = 128 >> lgS;                                   // values reduced per block
  // record addresses as 32-bit byte offsets from the (uniform) ring bases:
  // the rings hold at most 256 x 8 x 128 x 65 records
  auto rec_at = [](uint64_t *base, uint32_t idx) {
    return reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(base) + idx * 8u);
  };
  auto prec_at = [&](int32_t j, int v, int slice) {
    return rec_at(prec, ((uint32_t(g) * kRing + uint32_t(j & (kRing - 1))) * 128u + uint32_t(v)) *
                            uint32_t(S) + uint32_t(slice));
  };
  auto yrec_at = [&](int32_t j, int v) {
    return rec_at(yrec, (uint32_t(g) * kRing + uint32_t(j & (kRing - 1))) * 128u + uint32_t(v));
  };
  // reduce duty: lane covers value s * VPB + (2 lane >> lgS), slices 2 lane mod S and + 1
  const int red_v = s * VPB + ((2 * lane) >> lgS);
  const int red_slice = (2 * lane) & (S - 1);
  auto red_request = [&](int32_t j, uint64_t (&rr)[2]) {
    rr[0] = ld_rec(prec_at(j, red_v, red_slice));
    rr[1] = ld_rec(prec_at(j, red_v, red_slice + 1));
  };
  // sum a panel's partials of this block's values, write y
  auto red_finish = [&](int32_t jr, const uint64_t (&rr)[2]) {
    double acc = double(rec_val(rr[0])) + double(rec_val(rr[1]));
    for (int off = 1; off < (S >> 1); off <<= 1) acc += __shfl_xor(acc, off, 64);
    if ((lane & ((S >> 1) - 1)) == 0) st_rec(yrec_at(jr, red_v), float(acc), tag_of(jr));
  };
  auto fetch_request = [&](int32_t j, uint64_t (&fr)[2]) {
    fr[0] = ld_rec(yrec_at(j, 2 * lane));
    fr[1] = ld_rec(yrec_at(j, 2 * lane + 1));
  };

  v4f T[kTiles][kLoads];  // register ring: panel j in T[j mod kTiles]
  float q[4][8];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 8; ++r) q[c][r] = 0.f;
  float *tr = tr_all + w * (kPanel * kTrRow);
  const int li = lane & 15, lk = lane >> 4;
  const int32_t steps = NPg + kDefer;

#pragma unroll
  for (int i = 0; i < kPF; ++i) load_panel(T[i], i);

  auto step = [&](int32_t j, v4f (&Tj)[kLoads], v4f (&Tn)[kLoads], v4f (&Tb)[kLoads]) {
    const int32_t jf = j - kDefer, jb = j - kDefer;
    // The reduce wave of step j takes the block's two oldest unreduced panels
    // up to j - kRedLag; a panel whose partials are not all published yet is
    // left to the next step's reduce wave (no polling: this block is ahead
    // of another).  The prefix counter only grows; a wave that reads it
    // before an earlier update may sum a panel again, storing the same record.
    const bool red_duty = w == (j & 7) && j - kRedLag >= 0;
    int32_t jr0 = 0, jr1 = 0;
    bool ok0 = false, ok1 = false;
    const bool fetch_duty = w == ((j + 4) & 7) && jf >= 0 && jf < NPg;
    uint64_t rr0[2], rr1[2], fr[2];
    // (1) record loads of this step's duties, before the prefetch
    if (red_duty) {
      const int32_t lim = (j - kRedLag < NPg - 1 ? j - kRedLag : NPg - 1);
      jr0 = int32_t(__builtin_amdgcn_readfirstlane(lds_get(red_done)));
      jr1 = jr0 + 1;
      ok0 = jr0 <= lim;
      ok1 = jr1 <= lim;
      red_request(ok0 ? jr0 : 0, rr0);
      red_request(ok1 ? jr1 : 0, rr1);
    }
    if (fetch_duty) fetch_request(jf, fr);
    // (2) prefetch of panel j + kPF
    load_panel(Tn, j + kPF);
    __builtin_amdgcn_sched_barrier(0);
    // (3) duties.  The first check reads records loaded before the prefetch;
    // only a stale record takes the polling loop, whose loads wait behind the
    // prefetch.
    if (red_duty && ok0) {
      const uint32_t t0 = tag_of(jr0), t1 = tag_of(jr1);
      bool fresh0 = __all(rec_tag(rr0[0]) == t0 && rec_tag(rr0[1]) == t0);
      if (!fresh0 && !dead && jr0 + kDefer - 1 <= j) {
        // its fetch is next: this panel cannot wait for a later step
        uint32_t spins = 0;
        do {
          ++stale;
          if (give_up(spins)) break;
          red_request(jr0, rr0);
          fresh0 = __all(rec_tag(rr0[0]) == t0 && rec_tag(rr0[1]) == t0);
        } while (!fresh0);
      }
      if (fresh0 || dead) {
        red_finish(jr0, rr0);
        uint32_t done = uint32_t(jr0 + 1);
        if (ok1 && __all(rec_tag(rr1[0]) == t1 && rec_tag(rr1[1]) == t1)) {
          red_finish(jr1, rr1);
          done = uint32_t(jr1 + 1);
        }
        if (lane == 0) __hip_atomic_fetch_max(red_done, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else {
        ++stale;
      }
    }
    if (fetch_duty) {
      const uint32_t t = tag_of(jf);
      if (!dead && __any(rec_tag(fr[0]) != t || rec_tag(fr[1]) != t)) {
        uint32_t spins = 0;
        do {
          ++stale;
          if (give_up(spins)) break;
          fetch_request(jf, fr);
        } while (__any(rec_tag(fr[0]) != t || rec_tag(fr[1]) != t));
      }
      float *ys = yr + (jf & (kYSlots - 1)) * (kPanel * 8);
      const int64_t rb = row0 + int64_t(jf) * kPanel;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int v = 2 * lane + u, rhs = v >> 4, row = v & 15;
        ys[row * 8 + rhs] = rb + row < m ? rec_val(u == 0 ? fr[0] : fr[1]) : 0.f;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) lds_set(yflag + (jf & (kYSlots - 1)), uint32_t(jf + 1));
    }
    // (4) phase A on panel j
    if (j < NPg) {
      // The lanes pass the tile to each other through LDS: compiler barriers
      // keep the reads after all of the writes (and the next panel's writes
      // after the reads), with all lanes active.
      v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // 64-column halves: lanes (lane & 31) in [16 h, 16 h + 16) hold them
        asm volatile("" ::: "memory");
        if (((lane >> 4) & 1) == h) {
#pragma unroll
          for (int u = 0; u < kLoads; ++u)
            *reinterpret_cast<v4f *>(tr + (2 * u + half) * kTrRow + 4 * (lane & 15)) = Tj[u];
        }
        asm volatile("" ::: "memory");
        // row li, columns 16 lk + 4 v + e of the half (MFMA t = 4 v + e); p of
        // the same columns: slice column 128 w + 64 h + 16 lk + 4 v + e
        const float *pb = pl + ((32 * w + 16 * h + 4 * lk) * 8 + (lane & 7)) * 4;
        const float *pa = tr + li * kTrRow + 16 * lk;
        // operands of step v + 1 are read while step v's MFMAs run
        v4f an = *reinterpret_cast<const v4f *>(pa);
        v4f bn = *reinterpret_cast<const v4f *>(pb);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const v4f a = an, b = bn;
          if (v + 1 < 4) {
            an = *reinterpret_cast<const v4f *>(pa + 4 * (v + 1));
            bn = *reinterpret_cast<const v4f *>(pb + 32 * (v + 1));
          }
          __builtin_amdgcn_sched_barrier(0);
          acc0 = mfma16(a[0], b[0], acc0);
          acc1 = mfma16(a[1], b[1], acc1);
          acc0 = mfma16(a[2], b[2], acc0);
          acc1 = mfma16(a[3], b[3], acc1);
        }
      }
      const v4f d = acc0 + acc1;  // D[row 4 lk + e][rhs li]
      const int slot = j % kStSlots;
      float *sw = stg + (slot * kWaves + w) * (8 * kPanel);
      if (li < 8) *reinterpret_cast<v4f *>(sw + li * kPanel + 4 * lk) = d;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      uint32_t old = 0;
      if (lane == 0)
        old = __hip_atomic_fetch_add(cnt + slot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      old = __builtin_amdgcn_readfirstlane(old);
      if ((old & (kWaves - 1)) == kWaves - 1) {
        // last wave to stage panel j: sum the 8 partials in wave order, publish
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const float *sb = stg + slot * kWaves * (8 * kPanel);
        int ln = lane;
        asm volatile("" : "+v"(ln));  // formed here, not held across the loop
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int v = ln + 64 * u;   // v = 16 rhs + row
          float sum = sb[v];
#pragma unroll
          for (int w2 = 1; w2 < kWaves; ++w2) sum += sb[w2 * (8 * kPanel) + v];
          st_rec(prec_at(j, v, s), sum, tag_of(j));
        }
      }
    }
    // (5) phase B on panel j - kDefer
    if (jb >= 0) {
      const int ysl = jb & (kYSlots - 1);
      uint32_t spins = 0;
      while (!dead && lds_get(yflag + ysl) != uint32_t(jb + 1))
        if (give_up(spins)) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const float *ys = yr + ysl * (kPanel * 8) + half * 8;
      // y of row pair u + 1 is read while row pair u is multiplied (the LDS
      // latency is not left to the other wave of the SIMD to cover)
      v4f yn0 = *reinterpret_cast<const v4f *>(ys);
      v4f yn1 = *reinterpret_cast<const v4f *>(ys + 4);
#pragma unroll
      for (int u = 0; u < kLoads; ++u) {
        const v4f y0 = yn0, y1 = yn1;
        if (u + 1 < kLoads) {
          yn0 = *reinterpret_cast<const v4f *>(ys + (u + 1) * 16);
          yn1 = *reinterpret_cast<const v4f *>(ys + (u + 1) * 16 + 4);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float mv = Tb[u][c];
          q[c][0] += mv * y0[0];
          q[c][1] += mv * y0[1];
          q[c][2] += mv * y0[2];
          q[c][3] += mv * y0[3];
          q[c][4] += mv * y1[0];
          q[c][5] += mv * y1[1];
          q[c][6] += mv * y1[2];
          q[c][7] += mv * y1[3];
        }
      }
      if ((jb + 1) % kFlush == 0 || jb == NPg - 1) {
        // the two halves of the wave hold the same columns (even / odd rows)
        int qi = w * 32 + (lane & 31);
        asm volatile("" : "+v"(qi));
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            const float both = q[c][r] + __shfl_xor(q[c][r], 32, 64);
            if (half == 0) q64[(c * 8 + r) * kQLanes + qi] += double(both);
            q[c][r] = 0.f;
          }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };

  // Unrolled over the ring, so every tile index is a constant and the ring
  // stays in registers: panel j + kPF goes kPF tiles ahead, panel j - kDefer
  // is kDefer tiles back.
  for (int32_t j = 0; j < steps; j += kTiles) {
#pragma unroll
    for (int sl = 0; sl < kTiles; ++sl) {
      if (j + sl >= steps) break;
      step(j + sl, T[sl], T[(sl + kPF) % kTiles], T[(sl + kTiles - kDefer) % kTiles]);
    }
  }

  // ------------------------------------------------------------ output
  // The lane's columns of M^T y over this group's rows, FP64 (even-row lanes).
  int lo = tid;
  asm volatile("" : "+v"(lo));  // the output addresses are formed here, not held across the loop
  if (((lo >> 5) & 1) == 0) {
    const int64_t ocol = c0 + kWaveCols * (lo >> 6) + 4 * (lo & 31);
    const int qo = (lo >> 6) * 32 + (lo & 31);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t col = ocol + c;
      if (col < cend) {
#pragma unroll
        for (int r = 0; r < 8; ++r) part[(int64_t(g) * n + col) * 8 + r] = q64[(c * 8 + r) * kQLanes + qo];
      }
    }
  }
  if (stale != 0u && lane == 0)
    __hip_atomic_fetch_add(&st->fstale, stale, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // The last block to finish advances the epoch: records of this launch never
  // match a later launch's tags.
  __syncthreads();
  if (tid == 0) {
    const uint32_t done =
        __hip_atomic_fetch_add(&st->fdone, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (done == gridDim.x - 1) {
      __hip_atomic_store(&st->fdone, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&st->fepoch, (epoch + 1u) & 0xffffu, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace

FusedShape fused_shape(int64_t m, int64_t n, int K, int cus) {
  FusedShape f;
  if (K != 8 || n % 4 != 0 || n < 4 || m < 2 || n > int64_t(kMaxSlices) * kSliceCols) return f;
  int64_t S = 2;
  while (S * kSliceCols < n) S *= 2;
  int lg = 0;
  while ((int64_t(1) << lg) < S) ++lg;
  const int64_t NP = (m + kPanel - 1) / kPanel;
  int64_t NG = cus / S;
  if (NG > NP) NG = NP;
  if (NG < 1) return f;
  const int64_t NPg = (NP + NG - 1) / NG;
  if (NPg >= 65535) return f;
  // every block of the grid must be resident at once (cooperative launch)
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mv_fused8, kThreadsB, 0) != hipSuccess ||
      int64_t(per_cu) * cus < S * NG)
    return f;
  // width: columns per slice, a multiple of 4 (16-byte loads stay aligned)
  const int64_t width = ((n + S - 1) / S + 3) / 4 * 4;
  f.slices = int32_t(S);
  f.lg_slices = lg;
  f.groups = int32_t(NG);
  f.width = int32_t(width);
  f.panels = int32_t(NPg);
  return f;
}

size_t fused_records(const FusedShape &f) {
  return size_t(f.groups) * kRing * 128 * (size_t(f.slices) + 1);
}

// One-pass launches of this process on one device, in launch order across
// streams: each waits on the previous one's completion event.
struct FusedOrder {
  std::mutex mu;
  std::map<int, hipEvent_t> last;  // per device; kept until the process ends
};
FusedOrder &fused_order() {
  static FusedOrder *o = new FusedOrder();
  return *o;
}

hipError_t mv_fused8(hipStream_t s, const float *M, int64_t m, int64_t n, const float *X,
                     double *part, uint64_t *records, PcgState *st, const FusedShape &f) {
  uint64_t *prec = records;
  uint64_t *yrec = records + size_t(f.groups) * kRing * 128 * size_t(f.slices);
  int32_t lg = f.lg_slices, width = f.width, NPg = f.panels;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  FusedOrder &o = fused_order();
  std::lock_guard<std::mutex> lock(o.mu);
  auto it = o.last.find(dev);
  if (it == o.last.end()) {
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    it = o.last.emplace(dev, ev).first;
  } else {
    e = hipStreamWaitEvent(s, it->second, 0);
    if (e != hipSuccess) return e;
  }
  k_mv_fused8<<<dim3(unsigned(f.slices * f.groups)), dim3(kThreadsB), 0, s>>>(
      M, m, n, X, part, prec, yrec, st, lg, width, NPg);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipEventRecord(it->second, s);
}

}  // namespace mfs
