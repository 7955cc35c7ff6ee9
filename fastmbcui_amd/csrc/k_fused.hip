// k_fused.hip -- q = M^T (M p) in one HBM pass over M (dense form, FP32 batch of 8).
//
// The reference applies the normal operator as two cuBLAS GEMVs per right-hand
// side (src/kl_chol.cu:928-936 then :992-1002), so M (m x n, 17 GB at
// m = n = 65,536 in FP32) crosses HBM twice per PCG iteration.  Here every row
// panel of M is read once: the wave that owns a 64-column piece of a panel
// keeps it in LDS until the panel's y = M p is complete over all columns, then
// multiplies the same piece by y for its columns of M^T y.
//
// Layout.  One block per 256-column slice (G = ceil(n / 256) blocks, all
// resident: one block per CU, launched cooperatively) of 4 compute waves and
// 3 exchange waves.  Compute wave w owns columns c0 + 64 w .. + 63; panels
// are 16 rows.  Both products run on the full-rate FP32 matrix instruction
// v_mfma_f32_16x16x4f32 (the 4x4x1 form issues a quarter of the FLOPs per
// cycle):
//   phase A  D[16 rows][16] += A[16 rows][4 columns] B[4 columns][16], B = p
//            (8 right-hand sides in j < 8, zero above): 16 MFMAs per wave
//            give the wave's partial y of the panel;
//   phase B  D[16 columns][16] += A[16 columns][4 rows] B[4 rows][16], B = y
//            of the panel (zero above j = 8): 16 MFMAs (4 column groups x 4
//            row steps) add the panel's M^T y to the wave's columns.
// Lane layout of the 16x16x4 FP32 MFMA: A[i][k] from lane i + 16 k, B[k][j]
// from lane j + 16 k, D[i][j] in lane j + 16 (i / 4), element i % 4.
//
// A compute wave's step i:
//   (a)  tile i (16 rows x 64 columns, in registers since step i - 4) -> LDS
//        slot i % (D + 1) (wave-private ring, row-major with an XOR swizzle of
//        16-column groups by row % 4, so that both the row-order stores and
//        phase B's column-order reads are free of bank conflicts);
//   (b)  prefetch tile i + 4 (non-temporal float4 loads) into the registers
//        just freed;
//   (c)  phase A from LDS; the wave's 16 x 8 partials go to the block's
//        staging ring (LDS) and the wave counts itself in;
//   (d)  phase B on tile i - D with y of panel i - D from the block's y ring
//        (LDS): FP32 sums per 8 panels, carried in FP64.
// Compute waves touch no cross-block data and wait only on LDS words: a wait
// on a global load would also wait for every load issued before it (vector
// loads complete in order) and drain the wave's prefetch of M.
//
// Exchange waves (each waits only on LDS words and on other blocks' records):
//   publisher  panel j: waits for the four compute waves' staged partials,
//              sums them in wave order and writes one tagged record per value;
//   reducer    panel j, once this block has published panel j + 1: requests
//              the G partials of its values of panel j, and one panel later
//              sums them and writes y;
//   fetcher    panel j, once this block has published panel j + 4: requests
//              y(j), and two panels later (polling while stale) copies it
//              into the y ring and flags it for the compute waves.
// Value v (rhs c, row 4 t + q at v = 32 q + 4 c + t) of panel j is reduced by
// block (128 j + v) mod G: lanes sum blocks lane + 64 u in FP64 in order,
// then a fixed butterfly over the lanes (deterministic).  A panel's y is used
// D = 8 steps after its partials.
//
// Measured at m = n = 65,536 (one apply, 8 right-hand sides, MI355X): 5.3 ms,
// against 4.9 ms for the two passes (k_mv_rows_f32x8 + k_mv_cols_f32x8), so
// it stays opt-in.  The path there: the first form (4x4x1 MFMAs, every wave
// polling y itself) took 8.5 ms; full-rate MFMAs alone made it 10.0 ms,
// because y was late at 87% of the steps and each wait drained the wave's
// prefetch; one exchange wave 7.8 ms, limited by one load round trip
// (~1.2 us under the stream) per two panels; three exchange waves 5.3 ms.
// What remains: the MFMA unit is busy under half of a compute step (one
// compute wave per SIMD cannot hide its own LDS and dependency latencies),
// and the two cross-block hops of y (about 1.2 us each) leave compute waves
// waiting for y about a third of the time.  Eight 32-column compute waves,
// a deeper prefetch (8 tiles) or issuing the step's LDS reads ahead of the
// MFMAs each measured 5.4-5.8 ms.
//
// Cross-block data (partials, y) is exchanged as 8-byte records (value, tag)
// written and read with device-coherent accesses; tag = (launch epoch << 16)
// | (panel + 1).  A reader that sees a stale tag reads again.  The epoch
// advances at the end of every launch (last reducer wave), so records of
// earlier launches never match.  Every wait is bounded: after kSpinLimit
// polls the wave sets PcgState::fabort and stops waiting (the host reports an
// error); the exchange waves pass an abort on to the compute waves in LDS.
#include <type_traits>

#include "mfs_internal.h"

namespace mfs {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kCW = 4;                       // compute waves; waves kCW .. kCW + 2 exchange
constexpr int kBlockThreads = (kCW + 3) * 64;
constexpr int kPanel = 16;                   // rows per panel (the MFMA's 16 rows)
constexpr int kWaveCols = 64;                // columns per compute wave
constexpr int kQ = kWaveCols / 16;           // 16-column groups per compute wave
constexpr int kSlice = kCW * kWaveCols;      // 256 columns per block
constexpr int kDefer = kFusedDefer;          // phase B trails phase A by this many panels
constexpr int kSlots = kFusedSlots;          // record rings (panels)
constexpr int kTiles = kDefer + 1;           // LDS tile slots per compute wave
constexpr int kTileFloats = kPanel * kWaveCols;           // 4 KB
constexpr int kWaveBytes = kTiles * kTileFloats * 4;
constexpr int kStage = kPanel * 8;           // one panel's 16 x 8 values
constexpr int kSP = 4;                       // staging ring (panels)
constexpr int kSY = 8;                       // y ring (panels)
constexpr int kStageOff = kCW * kWaveBytes;                // 147,456 B of tiles
constexpr int kYOff = kStageOff + kSP * kCW * kStage * 4;  // + 8 KB staging
constexpr int kSyncOff = kYOff + kSY * kStage * 4;         // + 4 KB y ring
constexpr int kLdsBytes = kSyncOff + 32 * 4;               // + sync words
// sync words: arrivals per staging slot, rounds consumed per staging slot,
// panel + 1 held per y slot, reads per y slot, abort
constexpr int kCnt = 0, kDone = kCnt + kSP, kYTag = kDone + kSP, kYCnt = kYTag + kSY,
              kAbort = kYCnt + kSY, kPub = kAbort + 1;
constexpr int kPF = 4;                       // tiles in flight (register buffers)
constexpr int kFetchLag = 5;                 // y(j) requested once panel j + 4 is published
constexpr int kRedAhead = 1;                 // the reducer sums panel j one panel after requesting it
constexpr int kBatch = 2;                    // reduction duties per panel in flight (G >= 64)
constexpr int kFlushSteps = 8;               // FP32 phase-B sums between FP64 flushes
constexpr int kMinSlices = 8;                // narrower matrices take the two passes
constexpr uint32_t kSpinLimit = 1u << 20;    // polls before giving up

static_assert(kLdsBytes <= 160 * 1024, "LDS");
// Record rings: panel j's partials are rewritten by panel j + kSlots, whose
// publication needs y(j + kSlots - D), hence every reduction of panel
// j + kSlots - D, hence (rounds run in order) every reduction of panel j; y
// of panel j is rewritten by the reduction of panel j + kSlots, which needs
// every block's partials of it, hence (their compute waves used
// y(j + kSlots - D)) every block's fetch of y(j).
static_assert(kSlots > kDefer && (kSlots & (kSlots - 1)) == 0, "record ring");
static_assert(kFetchLag + 2 < kDefer, "y fetched too late for its use");

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

__device__ inline bool aborted(PcgState *st) {
  return __hip_atomic_load(&st->fabort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}
__device__ inline void set_abort(PcgState *st) {
  __hip_atomic_store(&st->fabort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS words shared by the block's waves.  LDS operations of one wave are
// performed in issue order, so a compiler barrier between a wave's data
// accesses and its flag access is all the ordering needed.
__device__ inline uint32_t lds_get(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_set(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void lds_inc(uint32_t *p) {
  __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ inline v4f load_stream(const float *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
}

// Tile element (row r, wave column c < 64) sits at r * 64 + (c ^ 16 (r % 4)):
// phase B's reads (lane li + 16 lk: row 4 t + lk, column 16 q + li) then
// fall in 64 distinct banks.
static_assert(kWaveCols == 64, "tile swizzle");
__device__ inline int tile_at(int r, int c) { return r * kWaveCols + (c ^ (16 * (r & 3))); }

__global__ __launch_bounds__(kBlockThreads, 1) void k_mv_fused8(
    const float *__restrict__ M, int64_t m, int64_t n, const float *__restrict__ X,
    double *__restrict__ part, uint64_t *__restrict__ prec, uint64_t *__restrict__ yrec,
    PcgState *st, int32_t G, int32_t NP) {
  if (st->n_active == 0) return;  // the batch has converged: uniform early exit
  __shared__ __align__(16) char lds[kLdsBytes];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = blockIdx.x;
  float *stage = reinterpret_cast<float *>(lds + kStageOff);  // [kSP][kCW][16 rows][8]
  float *ybuf = reinterpret_cast<float *>(lds + kYOff);       // [kSY][16 rows][8]
  uint32_t *sync = reinterpret_cast<uint32_t *>(lds + kSyncOff);
  if (threadIdx.x < 32) sync[threadIdx.x] = 0u;
  __syncthreads();
  auto lds_abort = [&]() { return lds_get(sync + kAbort) != 0u; };

  const uint32_t uG = uint32_t(G);
  // partial of value v (= 32 q + i) of panel j from block b; y of value v
  // (32-bit indices: the rings hold kSlots x 128 x (G + 1) records)
  auto part_at = [&](int32_t j, uint32_t v, uint32_t b) {
    return prec + (((uint32_t(j) & (kSlots - 1)) * 4u + (v >> 5)) * uG + b) * 32u + (v & 31u);
  };
  auto y_at = [&](int32_t j, uint32_t v) { return yrec + ((uint32_t(j) & (kSlots - 1)) * 128u + v); };
  if (w >= kCW) {
    // ================================================================ exchange waves
    const uint32_t epoch = st->fepoch;
    auto tag_of = [&](int32_t j) { return (epoch << 16) | uint32_t(j + 1); };
    bool stop = false;
    auto give_up = [&](uint32_t &spins) {
      if (++spins > kSpinLimit || ((spins & 255u) == 0u && aborted(st))) {
        set_abort(st);
        lds_set(sync + kAbort, 1u);
        stop = true;
        return true;
      }
      __builtin_amdgcn_s_sleep(1);
      return false;
    };
    if (w == kCW) {
      // ---------------------------------------------- publisher and y fetcher
      // Iteration j publishes panel j, then moves y of panel j - kYLag into
      // the y ring (requested at iteration j - kYLag + kYAhead).
      // Iteration j: publish panel j (waits only on this block's compute waves).
      for (int32_t j = 0; j < NP; ++j) {
        const int s = j % kSP;
        const uint32_t want = uint32_t(kCW) * uint32_t(j / kSP + 1);
        uint32_t spins = 0;
        while (!stop && lds_get(sync + kCnt + s) < want)
          if (give_up(spins)) break;
        asm volatile("" ::: "memory");
        const float *sg = stage + s * kCW * kStage;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int v = lane + 64 * u, q = v >> 5, c = (v >> 2) & 7, t = v & 3;
          const int e = (4 * t + q) * 8 + c;
          float sum = sg[e];
#pragma unroll
          for (int cw2 = 1; cw2 < kCW; ++cw2) sum += sg[cw2 * kStage + e];  // wave order
          st_rec(part_at(j, uint32_t(v), uint32_t(g)), sum, tag_of(j));
        }
        asm volatile("" ::: "memory");
        if (lane == 0) {
          lds_set(sync + kDone + s, uint32_t(j / kSP + 1));
          lds_set(sync + kPub, uint32_t(j + 1));
        }
      }
    } else if (w == kCW + 2) {
      // ---------------------------------------------- y fetcher
      // Iteration k requests y(k) once this block has published panel
      // k + kFetchLag - 1 (by then its reducers have summed panel k), and
      // moves y(k - 2), requested two iterations before, into the y ring.
      uint64_t yr0[2], yr1[2], yr2[2];  // requests of panels k % 3 == 0, 1, 2
      auto y_request = [&](int32_t j, uint64_t (&dst)[2]) {
        const int32_t jc = j < 0 ? 0 : (j >= NP ? NP - 1 : j);
#pragma unroll
        for (int u = 0; u < 2; ++u) dst[u] = ld_rec(y_at(jc, uint32_t(lane + 64 * u)));
      };
      y_request(0, yr1);  // placeholders for the first two iterations
      y_request(0, yr2);
      auto iter = [&](int32_t k, uint64_t (&mine)[2], uint64_t (&old)[2]) {
        if (k < NP) {
          const uint32_t want = uint32_t(k + kFetchLag < NP ? k + kFetchLag : NP);
          uint32_t spins = 0;
          while (!stop && lds_get(sync + kPub) < want)
            if (give_up(spins)) break;
          y_request(k, mine);
        }
        const int32_t jy = k - 2;
        if (jy >= 0 && jy < NP) {
          const uint32_t t = tag_of(jy);
          uint32_t spins = 0;
          while (!stop && __any(rec_tag(old[0]) != t || rec_tag(old[1]) != t)) {
            if (give_up(spins)) break;
            y_request(jy, old);
          }
          const int s = jy % kSY;
          const uint32_t want = uint32_t(kCW) * uint32_t(jy / kSY);  // readers of the slot's previous panel
          spins = 0;
          while (!stop && lds_get(sync + kYCnt + s) < want)
            if (give_up(spins)) break;
          asm volatile("" ::: "memory");
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int v = lane + 64 * u, q = v >> 5, c = (v >> 2) & 7, t2 = v & 3;
            ybuf[s * kStage + (4 * t2 + q) * 8 + c] = rec_val(old[u]);
          }
          asm volatile("" ::: "memory");
          if (lane == 0) lds_set(sync + kYTag + s, uint32_t(jy + 1));
        }
      };
      for (int32_t k = 0; k < NP + 2; k += 3) {
        iter(k, yr0, yr1);
        iter(k + 1, yr1, yr2);
        iter(k + 2, yr2, yr0);
      }
    } else {
      // ---------------------------------------------- reducer
      // Duties: value v of panel j is reduced by block (128 j + v) mod G, so
      // this block's values of panel j are v0(j) + k G with
      // v0(j) = (g - 128 j) mod G, stepped from panel to panel without
      // division.  Panel j's partials are requested once this block has
      // published panel j + 1 and summed kRedAhead panels later.
      const uint32_t c128 = 128u % uG;
      auto next_v0 = [&](uint32_t v) { return v >= c128 ? v - c128 : v + uG - c128; };
      auto count = [&](uint32_t v0) { return v0 < 128u ? int((127u - v0) / uG) + 1 : 0; };
      uint64_t rr0[kBatch][4], rr1[kBatch][4];  // requests of even / odd panels
      static_assert(kRedAhead == 1, "request ring rotation");
      auto request = [&](int32_t j, uint32_t v0, uint64_t (&dst)[kBatch][4]) {
        const int cnt = count(v0);
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
          if (b >= cnt) continue;
          const uint32_t v = v0 + uint32_t(b) * uG;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t src = uint32_t(lane + 64 * u);
            dst[b][u] = ld_rec(part_at(j, v, src < uG ? src : uG - 1));
          }
        }
      };
      // sum duty (j, v) from its records (polling while stale), write y
      auto finish = [&](int32_t j, uint32_t v, uint64_t (&rec)[4]) {
        const uint32_t t = tag_of(j);
        auto stale = [&]() {
          bool bad = false;
#pragma unroll
          for (int u = 0; u < 4; ++u) bad = bad || (uint32_t(lane + 64 * u) < uG && rec_tag(rec[u]) != t);
          return bad;
        };
        uint32_t spins = 0;
        while (!stop && __any(stale())) {
          if (give_up(spins)) break;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint32_t src = uint32_t(lane + 64 * u);
            rec[u] = ld_rec(part_at(j, v, src < uG ? src : uG - 1));
          }
        }
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (uint32_t(lane + 64 * u) < uG) acc += double(rec_val(rec[u]));
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) st_rec(y_at(j, v), float(acc), t);
      };
      // Iteration j: request panel j's duties into mine, sum panel j - 1's
      // (requested one iteration ago into prev).
      uint32_t v0 = uint32_t(g) % uG;  // v0(0)
      uint32_t v0_prev = 0;
      auto iter = [&](int32_t j, uint64_t (&mine)[kBatch][4], uint64_t (&prev)[kBatch][4]) {
        uint32_t v0_now = v0;
        if (j < NP) {
          // pace: this block has published panel j + 1 (or the last panel)
          const uint32_t want = uint32_t(j + 2 < NP ? j + 2 : NP);
          uint32_t spins = 0;
          while (!stop && lds_get(sync + kPub) < want)
            if (give_up(spins)) break;
          request(j, v0_now, mine);
          v0 = next_v0(v0);
        }
        const int32_t jf = j - 1;
        if (jf >= 0 && jf < NP) {
          const int cnt = count(v0_prev);
#pragma unroll
          for (int b = 0; b < kBatch; ++b)
            if (b < cnt) finish(jf, v0_prev + uint32_t(b) * uG, prev[b]);
          // narrow matrices (G < 64): the duties beyond kBatch, one by one
          for (int b = kBatch; b < cnt; ++b) {
            uint64_t one[4];
            const uint32_t v = v0_prev + uint32_t(b) * uG;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const uint32_t src = uint32_t(lane + 64 * u);
              one[u] = ld_rec(part_at(jf, v, src < uG ? src : uG - 1));
            }
            finish(jf, v, one);
          }
        }
        v0_prev = v0_now;
      };
      const int32_t iters = NP + 1;
      for (int32_t j = 0; j < iters; j += 2) {
        iter(j, rr0, rr1);
        iter(j + 1, rr1, rr0);
      }
      if (lane == 0) {
        const uint32_t prev =
            __hip_atomic_fetch_add(&st->fdone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == uint32_t(G) - 1) {
          __hip_atomic_store(&st->fdone, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&st->fepoch, epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    return;
  }

  // ================================================================ compute waves
  const int li = lane & 15, lk = lane >> 4;
  const int64_t cw = int64_t(g) * kSlice + w * kWaveCols;  // the wave's first column
  float *tiles = reinterpret_cast<float *>(lds + w * kWaveBytes);
  bool stop = false;
  // Wait until the LDS word reaches want (at least, or exactly when EQ).
  auto wait_lds = [&](const uint32_t *word, uint32_t want, bool eq) {
    uint32_t spins = 0;
    for (;;) {
      const uint32_t v = lds_get(word);
      if (eq ? v == want : v >= want) break;
      if (stop) break;
      if (lds_abort() || ++spins > kSpinLimit) {
        set_abort(st);
        stop = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    asm volatile("" ::: "memory");
  };

  // Phase A's B operands: lane (li, lk) of MFMA (q, e) holds p[cw + 16 q +
  // 4 lk + e][li] (zero for li >= 8 and for columns past n).
  float pB[kQ][4];
#pragma unroll
  for (int q = 0; q < kQ; ++q)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t col = cw + 16 * q + 4 * lk + e;
      pB[q][e] = (li < 8 && col < n) ? X[col * 8 + li] : 0.f;
    }

  // Row-order loads: lane (li, lk) reads row li of the panel, columns
  // cw + 16 q + 4 lk .. + 3.  Rows past m read row m - 1 and columns past n
  // read columns n - 4 .. n - 1, so every wave issues the same loads and no
  // value is inspected at load time.  Such columns meet p = 0 in phase A and
  // are never written in phase B; such rows get y = 0 in phase B.
  int64_t lcol[kQ];
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    const int64_t c = cw + 16 * q + 4 * lk;
    lcol[q] = c < n ? c : n - 4;
  }
  auto clampp = [&](int32_t j) { return j < 0 ? 0 : (j >= NP ? NP - 1 : j); };
  auto load_tile = [&](int32_t i, v4f (&buf)[kQ]) {
    const int64_t f = int64_t(i) * kPanel + li;
    const float *row = M + (f < m ? f : m - 1) * n;
#pragma unroll
    for (int q = 0; q < kQ; ++q) buf[q] = load_stream(row + lcol[q]);
  };

  v4f QB[kQ];  // phase B sums: column group q, D[col 16 q + 4 lk + e][rhs li] in element e
  double QD[kQ][4];
#pragma unroll
  for (int q = 0; q < kQ; ++q) {
    QB[q] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int e = 0; e < 4; ++e) QD[q][e] = 0.0;
  }

  // ---- one step; buf holds tile i (loaded at step i - 4) and receives
  // tile i + 4
  auto step = [&](int32_t i, v4f (&buf)[kQ]) {
    // (a) tile i -> LDS slot i % kTiles (row li, 16-column groups swizzled)
    float *slot = tiles + (i % kTiles) * kTileFloats;
#pragma unroll
    for (int q = 0; q < kQ; ++q) *reinterpret_cast<v4f *>(slot + tile_at(li, 16 * q + 4 * lk)) = buf[q];
    // (b) prefetch tile i + 4 into the registers just freed (pinned ahead of
    // phase A's MFMAs, which the scheduler would otherwise hoist above it)
    __builtin_amdgcn_sched_barrier(0);
    load_tile(clampp(i + kPF), buf);
    __builtin_amdgcn_sched_barrier(0);
    // (c) phase A: MFMA (q, e) takes columns 16 q + 4 k + e, k = lane group
    if (i < NP) {
      v4f CA[2] = {v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int q = 0; q < kQ; ++q) {
        const v4f a = *reinterpret_cast<const v4f *>(slot + tile_at(li, 16 * q + 4 * lk));
#pragma unroll
        for (int e = 0; e < 4; ++e) CA[e & 1] = mfma16(a[e], pB[q][e], CA[e & 1]);
      }
      // staging slot i % kSP, once the publisher has taken its previous panel
      const int s = i % kSP;
      wait_lds(sync + kDone + s, uint32_t(i / kSP), false);
      float *sg = stage + (s * kCW + w) * kStage;
      if (li < 8) {
#pragma unroll
        for (int e = 0; e < 4; ++e) sg[(4 * lk + e) * 8 + li] = CA[0][e] + CA[1][e];
      }
      asm volatile("" ::: "memory");
      if (lane == 0) lds_inc(sync + kCnt + s);
    }
    // (d) phase B on tile jb = i - D with y(jb) from the y ring
    const int32_t jb = i - kDefer;
    if (jb >= 0 && jb < NP) {
      const int s = jb % kSY;
      wait_lds(sync + kYTag + s, uint32_t(jb + 1), true);
      float yb[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = 4 * t + lk;
        const bool ok = li < 8 && int64_t(jb) * kPanel + r < m;
        const float v = ybuf[s * kStage + r * 8 + (li & 7)];
        yb[t] = ok ? v : 0.f;
      }
      asm volatile("" ::: "memory");
      if (lane == 0) lds_inc(sync + kYCnt + s);
      const float *tb = tiles + (jb % kTiles) * kTileFloats;
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int q = 0; q < kQ; ++q) QB[q] = mfma16(tb[tile_at(4 * t + lk, 16 * q + li)], yb[t], QB[q]);
      if (jb % kFlushSteps == kFlushSteps - 1 || jb == NP - 1) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
#pragma unroll
          for (int e = 0; e < 4; ++e) QD[q][e] += double(QB[q][e]);
          QB[q] = v4f{0.f, 0.f, 0.f, 0.f};
        }
      }
    }
  };

  // The first four tiles, in order (the scheduler would otherwise reorder
  // them, and the wait counts at the loop header assume tile 0 is oldest).
  static_assert(kPF == 4, "four-buffer rotation");
  const int32_t steps = (NP + kDefer + kPF - 1) / kPF * kPF;
  v4f b0[kQ], b1[kQ], b2[kQ], b3[kQ];
  load_tile(clampp(0), b0);
  __builtin_amdgcn_sched_barrier(0);
  load_tile(clampp(1), b1);
  __builtin_amdgcn_sched_barrier(0);
  load_tile(clampp(2), b2);
  __builtin_amdgcn_sched_barrier(0);
  load_tile(clampp(3), b3);
  __builtin_amdgcn_sched_barrier(0);
  for (int32_t i = 0; i < steps; i += 4) {
    step(i, b0);
    step(i + 1, b1);
    step(i + 2, b2);
    step(i + 3, b3);
  }

  // Each wave owns its columns: write q[c][r] (FP64) for c < n, r < 8.
  if (li < 8) {
#pragma unroll
    for (int q = 0; q < kQ; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t col = cw + 16 * q + 4 * lk + e;
        if (col < n) part[col * 8 + li] = QD[q][e];
      }
  }
}

}  // namespace

int32_t fused_blocks(int64_t m, int64_t n, int K, int cus) {
  if (K != 8 || n % 4 != 0 || n < 4 || m < 1) return 0;
  const int64_t G = (n + kSlice - 1) / kSlice;
  const int64_t NP = (m + kPanel - 1) / kPanel;
  if (G < kMinSlices || G > cus || NP >= 65535) return 0;
  return int32_t(G);
}

size_t fused_records(int32_t G) { return size_t(kSlots) * 128 * size_t(G) + size_t(kSlots) * 128; }

hipError_t mv_fused8(hipStream_t s, const float *M, int64_t m, int64_t n, const float *X,
                     double *part, uint64_t *records, PcgState *st, int32_t G) {
  const int32_t NP = int32_t((m + kPanel - 1) / kPanel);
  uint64_t *prec = records;
  uint64_t *yrec = records + size_t(kSlots) * 128 * size_t(G);
  void *args[] = {(void *)&M, (void *)&m, (void *)&n, (void *)&X, (void *)&part,
                  (void *)&prec, (void *)&yrec, (void *)&st, (void *)&G, (void *)&NP};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_mv_fused8), dim3(G),
                                    dim3(kBlockThreads), args, 0, s);
}

}  // namespace mfs
