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
// columns of the slice, and each of its lanes two adjacent columns (one
// 8-byte load per row).
//
// Step j of a wave (static ring of 5 register tiles: panels j - 3 .. j + 1):
//   (1) its exchange duty for step j, if any, issues its record loads;
//   (2) the loads of panel j + 1 are issued (non-temporal);
//   (3) panel j arrives; the exchange duty consumes its records;
//   (4) phase A: the 16 x 128 tile goes through a wave-private LDS transpose
//       (two 64-column halves, XOR-swizzled: conflict-free both ways) into
//       32 v_mfma_f32_16x16x4f32 with p as the B operand (rows x 8 right-hand
//       sides, padded to 16); the wave's partial y is staged in LDS and the
//       last wave of the block to stage sums the 8 partials in wave order
//       and publishes them;
//   (5) phase B for panel j - 3: y of its 16 rows from the LDS y ring
//       (broadcast reads) times the wave's registers, 256 VALU FMAs per lane
//       into FP32 sums that are added into FP64 (LDS) every 8 panels.
// Loads of a wave complete in order, so every record load is issued before
// the prefetch of panel j + 1: waiting for the records never waits for M.
// FP32 MFMA and VALU work of one SIMD do not overlap on this GPU.
//
// Cross-block exchange (tagged 8-byte records, device-coherent stores and
// loads; tag = (launch epoch << 16) | (panel + 1)):
//   publish  panel j, at step j: block partials of the 128 values
//            (value v = 16 rhs + row) -> prec[g][j mod 8][v][s];
//   reduce   panel j, at step j + 2, by wave (j + 2) mod 8: block s sums the
//            S partials of its 128 / S values (two slices per lane, then a
//            fixed butterfly over the value's lanes: deterministic, the same
//            bits in every block) and writes y[g][j mod 8][v];
//   fetch    panel j, at step j + 3, by wave (j + 7) mod 8: the 128 y
//            records -> LDS y ring (zero for rows past m), flag for the waves.
// A duty whose records are stale polls again (that wait also drains the
// wave's prefetch; the state counts these polls).  Every wait is bounded:
// after kSpinLimit polls the wave sets PcgState::fabort and stops waiting,
// the host reports the failure.  Panel j's records are rewritten at panel
// j + 8; by then every block has reduced and fetched panel j (a block at step
// j + 8 used y(j + 4), which needed every block's partials of panel j + 4).
//
// Residency.  The blocks wait for each other, so all S x NG of them must be
// on the GPU at once: the grid is at most one block per CU (the LDS holds one
// block per CU) and fused_shape checks the occupancy.  The launch is an
// ordinary one, and the one-pass launches of this process on a device are
// ordered one after another whatever stream they are queued on, so two such
// grids never share the CUs while each waits for blocks of its own.
// Measured at m = n = 65,536: 3.89 ms per matvec (4.4 TB/s of M), the two
// passes 2 x 2.45 ms; streaming M alone takes about 2.9 ms, the exchange's
// waits (y of a panel needs the slowest block of the group) the rest.
#include <map>
#include <mutex>
#include <type_traits>

#include "mfs_internal.h"

namespace mfs {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));

constexpr int kWaves = 8;                        // waves per block (two per SIMD)
constexpr int kThreadsB = kWaves * 64;
constexpr int kWaveCols = 128;                   // columns per wave (two per lane)
constexpr int kSliceCols = kWaves * kWaveCols;   // at most 1024 columns per block
constexpr int kPanel = 16;                       // rows per panel (the MFMA's 16 rows)
constexpr int kDefer = 3;                        // panel j's y is used at step j + kDefer
constexpr int kTiles = kDefer + 2;               // register tiles: j - kDefer .. j + 1
constexpr int kRedLag = 2;                       // panel j reduced at step j + kRedLag
constexpr int kRing = 8;                         // record rings (panels)
constexpr int kYSlots = 8;                       // LDS y ring (panels)
constexpr int kStSlots = 4;                      // LDS staging ring (panels)
constexpr int kFlush = 8;                        // panels between FP64 flushes of q
constexpr int kMaxSlices = 64;
constexpr uint32_t kSpinLimit = 1u << 22;        // polls before giving up

// LDS (bytes): wave-private transposes, p of the slice, FP64 sums of q,
// staging of the waves' partials, y ring, sync words.
constexpr int kTrBytes = kWaves * kPanel * 64 * 4;           // 32 KB
constexpr int kPStride = 20;                                  // floats per (16-column group, rhs)
constexpr int kPBytes = (kSliceCols / 16) * 8 * kPStride * 4; // 40 KB
constexpr int kQBytes = 16 * kThreadsB * 8;                   // 64 KB
constexpr int kStBytes = kStSlots * kWaves * 8 * kPanel * 4;  // 16 KB
constexpr int kYBytes = kYSlots * kPanel * 8 * 4;             // 4 KB
constexpr int kPOff = kTrBytes, kQOff = kPOff + kPBytes, kStOff = kQOff + kQBytes,
              kYOff = kStOff + kStBytes, kSyncOff = kYOff + kYBytes;
constexpr int kSyncWords = kStSlots + kYSlots + 1;            // arrivals, y flags, abort
constexpr int kLdsBytes = kSyncOff + kSyncWords * 4;
static_assert(kLdsBytes <= 160 * 1024, "LDS");
static_assert(kRing >= kDefer + 2 && kYSlots > kDefer + 1 && kStSlots > kDefer, "rings");
static_assert(kTiles == 5, "the step loop is unrolled over 5 register tiles");

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
__device__ inline v2f load_stream2(const float *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p));
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
  for (int i = tid; i < 16 * kThreadsB; i += kThreadsB) q64[i] = 0.0;
  // p of the slice: pl[(local column / 16) * 8 * kPStride + rhs * kPStride + column % 16],
  // zero past the slice (those columns belong to the next block or lie past n)
  for (int i = tid; i < kSliceCols * 8; i += kThreadsB) {
    const int cl = i >> 3, r = i & 7;
    const int64_t col = c0 + cl;
    pl[(cl >> 4) * 8 * kPStride + r * kPStride + (cl & 15)] = col < cend ? X[col * 8 + r] : 0.f;
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
  const int64_t ccol = c0w + 2 * lane;                         // the lane's two columns
  const float *mcol = M + (ccol + 2 <= n ? ccol : n - 2);      // clamped (columns past n are not written)
  auto load_panel = [&](v2f (&T)[kPanel], int32_t j) {
    const int32_t jc = j < NPg ? j : NPg - 1;
    const int64_t rb = row0 + int64_t(jc) * kPanel;
#pragma unroll
    for (int u = 0; u < kPanel; ++u) {
      const int64_t r = rb + u < m ? rb + u : m - 1;
      T[u] = load_stream2(mcol + r * n);
    }
  };

  // exchange addressing
  const int VPB = 128 >> lgS;                                   // values reduced per block
  auto prec_at = [&](int32_t j, int v, int slice) {
    return prec + ((size_t(g) * kRing + size_t(j & (kRing - 1))) * 128 + size_t(v)) * size_t(S) +
           size_t(slice);
  };
  auto yrec_at = [&](int32_t j, int v) {
    return yrec + (size_t(g) * kRing + size_t(j & (kRing - 1))) * 128 + size_t(v);
  };
  // reduce duty: lane covers value s * VPB + (2 lane >> lgS), slices 2 lane mod S and + 1
  const int red_v = s * VPB + ((2 * lane) >> lgS);
  const int red_slice = (2 * lane) & (S - 1);
  auto red_request = [&](int32_t j, uint64_t (&rr)[2]) {
    const int32_t jc = j < 0 ? 0 : j;
    rr[0] = ld_rec(prec_at(jc, red_v, red_slice));
    rr[1] = ld_rec(prec_at(jc, red_v, red_slice + 1));
  };
  auto fetch_request = [&](int32_t j, uint64_t (&fr)[2]) {
    const int32_t jc = j < 0 ? 0 : j;
    fr[0] = ld_rec(yrec_at(jc, 2 * lane));
    fr[1] = ld_rec(yrec_at(jc, 2 * lane + 1));
  };

  v2f T0[kPanel], T1[kPanel], T2[kPanel], T3[kPanel], T4[kPanel];
  float q[2][8];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int r = 0; r < 8; ++r) q[c][r] = 0.f;
  float *tr = tr_all + w * (kPanel * 64);
  const int li = lane & 15, lk = lane >> 4;
  const int32_t steps = NPg + kDefer;

  load_panel(T0, 0);

  auto step = [&](int32_t j, v2f (&Tj)[kPanel], v2f (&Tn)[kPanel], v2f (&Tb)[kPanel]) {
    const int32_t jr = j - kRedLag, jf = j - kDefer, jb = j - kDefer;
    const bool red_duty = w == (j & 7) && jr >= 0 && jr < NPg;
    const bool fetch_duty = w == ((j + 7) & 7) && jf >= 0 && jf < NPg;
    uint64_t rr[2], fr[2];
    // (1) record loads of this step's duties, before the prefetch
    if (red_duty) red_request(jr, rr);
    if (fetch_duty) fetch_request(jf, fr);
    // (2) prefetch of panel j + 1
    load_panel(Tn, j + 1);
    __builtin_amdgcn_sched_barrier(0);
    // (3) duties
    // The first check reads records loaded before the prefetch; only a stale
    // record takes the polling loop, whose loads wait behind the prefetch.
    if (red_duty) {
      const uint32_t t = tag_of(jr);
      if (!dead && __any(rec_tag(rr[0]) != t || rec_tag(rr[1]) != t)) {
        uint32_t spins = 0;
        do {
          ++stale;
          if (give_up(spins)) break;
          red_request(jr, rr);
        } while (__any(rec_tag(rr[0]) != t || rec_tag(rr[1]) != t));
      }
      double acc = double(rec_val(rr[0])) + double(rec_val(rr[1]));
      for (int off = 1; off < (S >> 1); off <<= 1) acc += __shfl_xor(acc, off, 64);
      if ((lane & ((S >> 1) - 1)) == 0) st_rec(yrec_at(jr, red_v), float(acc), t);
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
      v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // The lanes pass the tile to each other through LDS: compiler
        // barriers keep every read of a half after all of its writes (and
        // the next half's writes after them), with all lanes active.
        asm volatile("" ::: "memory");
        if ((lane >> 5) == h) {
          const int qf = (lane & 31) >> 1, half = 2 * (lane & 1);
#pragma unroll
          for (int u = 0; u < kPanel; ++u)
            *reinterpret_cast<v2f *>(tr + u * 64 + 4 * (qf ^ u) + half) = Tj[u];
        }
        asm volatile("" ::: "memory");
        const float *pb = pl + ((8 * w + 4 * h + lk) * 8 + (lane & 7)) * kPStride;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const v4f a = *reinterpret_cast<const v4f *>(tr + li * 64 + 4 * ((4 * lk + v) ^ li));
          const v4f b = *reinterpret_cast<const v4f *>(pb + 4 * v);
          acc0 = mfma16(a[0], b[0], acc0);
          acc1 = mfma16(a[1], b[1], acc1);
          acc0 = mfma16(a[2], b[2], acc0);
          acc1 = mfma16(a[3], b[3], acc1);
        }
      }
      const v4f d = acc0 + acc1;  // D[row 4 lk + e][rhs li]
      const int slot = j & (kStSlots - 1);
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
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int v = lane + 64 * u;   // v = 16 rhs + row
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
      const float *ys = yr + ysl * (kPanel * 8);
#pragma unroll
      for (int u = 0; u < kPanel; ++u) {
        const v4f y0 = *reinterpret_cast<const v4f *>(ys + u * 8);
        const v4f y1 = *reinterpret_cast<const v4f *>(ys + u * 8 + 4);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
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
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            q64[(c * 8 + r) * kThreadsB + tid] += double(q[c][r]);
            q[c][r] = 0.f;
          }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };

  for (int32_t j = 0; j < steps; j += kTiles) {
    // panel j in tile j mod 5, panel j + 1 in the next, panel j - 3 two ahead
    step(j, T0, T1, T2);
    if (j + 1 >= steps) break;
    step(j + 1, T1, T2, T3);
    if (j + 2 >= steps) break;
    step(j + 2, T2, T3, T4);
    if (j + 3 >= steps) break;
    step(j + 3, T3, T4, T0);
    if (j + 4 >= steps) break;
    step(j + 4, T4, T0, T1);
  }

  // ------------------------------------------------------------ output
  // The lane's columns of M^T y over this group's rows, FP64.
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int64_t col = ccol + c;
    if (col < cend) {
#pragma unroll
      for (int r = 0; r < 8; ++r) part[(int64_t(g) * n + col) * 8 + r] = q64[(c * 8 + r) * kThreadsB + tid];
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
  if (K != 8 || n % 2 != 0 || n < 2 || m < 1 || n > int64_t(kMaxSlices) * kSliceCols) return f;
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
  // every block of the grid must be resident at once
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mv_fused8, kThreadsB, 0) != hipSuccess ||
      int64_t(per_cu) * cus < S * NG)
    return f;
  // width: columns per slice, even (8-byte loads stay aligned)
  const int64_t width = ((n + S - 1) / S + 1) / 2 * 2;
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
