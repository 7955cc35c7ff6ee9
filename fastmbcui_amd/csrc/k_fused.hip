// k_fused.hip -- q = M^T (M p) in one HBM pass over M (dense form, FP32 batch of 8).
//
// The reference applies the normal operator as two cuBLAS GEMVs per right-hand
// side (src/kl_chol.cu:928-936 then :992-1002), so M (m x n, 17 GB at
// m = n = 65,536 in FP32) crosses HBM twice per PCG iteration.  Here every row
// panel of M is read once: the block that owns a 256-column slice keeps the
// panel in LDS until the panel's y = M p is complete over all slices, then
// multiplies the same tile by y for its slice of M^T y.
//
// Layout.  One block of 4 waves per column slice c0 = 256 g (G = ceil(n/256)
// blocks, all resident: one block per CU, launched cooperatively).  Panels
// are 16 rows; wave w owns rows 4w .. 4w+3 of every panel, one row quad.  A
// wave's step i:
//   (r)  reducer duty (below): request partial records of panel i - 4;
//   (y)  request y of panel i + 2 - D (used two steps later);
//   (a)  tile i (4 rows x 256 columns, in registers since step i-4) -> LDS;
//   (b)  prefetch tile i+4 (non-temporal float4 per lane, 1 KB per row);
//   (c)  phase A: partial y of its 4 rows over the slice, v_mfma_f32_4x4x1
//        on LDS reads in "4 rows of one column per MFMA block" order, a
//        reduce-scatter over the 16 blocks (permlane swaps and DPP), one
//        tagged record per lane;
//   (d)  phase B on tile i - D: v_mfma_f32_4x4x1 with the tile in row order
//        ("4 columns of one row per block") times y(i - D): q partial of the
//        slice, FP32 per 8 panels, carried in FP64;
//   (s)  reducer duty: sum the records requested in (r), write y.
// D = 8 panels of deferral hide the all-slices reduction of y.
//
// Reduction of y.  The 32 values (4 rows x 8 right-hand sides) of a wave's
// quarter of panel j are summed over the G slices by 8 designated waves (wave
// w of block (8 j + k) mod G sums values 4k .. 4k+3), which request the G
// partial records at the start of step j + 4 and write y at its end; every
// wave of the quarter requests y(j) at step j + D - 2 and uses it at step
// j + D.
//
// Cross-block data (partials, y) is exchanged as 8-byte records (value, tag)
// written and read with device-coherent (sc1) accesses; tag = (launch epoch
// << 16) | (panel + 1).  A reader that sees a stale tag reads again, so no
// wave ever waits for its own stores -- a wait would also drain the HBM
// prefetch of M that is in flight behind them.  The epoch advances at the end
// of every launch (last wave), so records of earlier launches never match.
// Every wait is bounded: after kSpinLimit polls the wave sets
// PcgState::fabort and stops waiting (the host reports an error).
//
// Vector memory operations complete in issue order, so a wait on one load
// also waits on every load issued before it.  Each step therefore issues the
// same sequence (1 record load, 4 tile loads, 1 record store; panel indices
// clamped, surplus results unused) and consumes loads issued 2-4 steps
// earlier, so the compiler's wait counts are exact and never drain the
// prefetch.  Only reducer steps (1 in 32) wait for more; their loads are
// issued and consumed within the step, so no pending load crosses steps in
// registers other than the tile and y rings.
//
// The 4x4x1 MFMA (16 blocks of C[4x4] += A[4x1] B[1x4]) was measured on
// gfx950 with an in-bounds probe: block b = lanes 4b..4b+3, A[i] from lane
// 4b+i, B[n] from lane 4b+n, C[i][n] in lane 4b+n, register i.
#include <type_traits>

#include "mfs_internal.h"

namespace mfs {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kW = 4;                        // waves per block
constexpr int kRows = 4;                     // rows per wave per panel (one MFMA row quad)
constexpr int kPanel = kW * kRows;           // 16 rows
constexpr int kDefer = kFusedDefer;          // phase B trails phase A by this many panels
constexpr int kSlots = kFusedSlots;          // exchange ring (panels)
constexpr int kTiles = kDefer + 1;           // LDS tile slots per wave
constexpr int kRowBytes = 1040;              // 256 floats + 16 B pad
constexpr int kWaveBytes = kTiles * kRows * kRowBytes;
constexpr int kTileLds = kW * kWaveBytes;    // 149,760 B
constexpr int kLdsBytes = kTileLds + kW * 32 * 4;  // + y staging: one block per CU
constexpr int kPF = 4;                       // tiles in flight (register buffers)
constexpr int kYLag = 2;                     // y requested this many steps before use
constexpr int kReducers = 8;                 // waves that sum a panel quarter (4 values each)
constexpr int kRedLag = 4;                   // reducers sum panel j's partials at step j + 4
constexpr int kFlushSteps = 8;               // FP32 phase-B sums between FP64 flushes
constexpr uint32_t kSpinLimit = 1u << 20;    // polls (each a coherent load) before giving up

static_assert(kLdsBytes <= 160 * 1024, "LDS");
// Waves stay within kDefer steps of each other (a wave at step t has used
// y(t - D), so every wave has published panel t - D).  Hence a partial slot
// is rewritten only after its reducers (step j + kRedLag) are done, and a y
// slot only after its readers (step j + D) are done:
static_assert(kSlots > kDefer + kRedLag && kSlots > 2 * kDefer - kRedLag, "exchange ring too short");
static_assert(kDefer - kYLag > kRedLag, "y requested before its reducers have written it");
static_assert((kSlots & (kSlots - 1)) == 0, "ring size");

__device__ inline v4f mfma4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
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

__device__ inline v4f load_stream(const float *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
}

// Lane pairs (a, b) across lane bit 5 / 4: the lane with the bit clear ends
// with a + partner's a, the other with b + partner's b.  v_permlane32_swap
// exchanges the upper half of a with the lower half of b; v_permlane16_swap
// the odd 16-lane rows of a with the even rows of b.
__device__ inline float pair_add32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ inline float pair_add16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false,
                                                  false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// Across lane bit 3 (lane ^ 8 = rotation by 8 inside a 16-lane row).
__device__ inline float pair_add8(float a, float b, int lane) {
  const bool hi = (lane >> 3) & 1;
  const float send = hi ? a : b;
  const float keep = hi ? b : a;
  const int moved = __builtin_amdgcn_update_dpp(0, __float_as_int(send), 0x128, 0xF, 0xF, false);
  return keep + __int_as_float(moved);
}
// x + x of lane ^ 4 (row_shl:4 / row_shr:4 stay inside the 16-lane row).
__device__ inline float sum_xor4(float x, int lane) {
  const int up = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x104, 0xF, 0xF, false);
  const int dn = __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xF, 0xF, false);
  return x + __int_as_float(((lane >> 2) & 1) ? dn : up);
}

__global__ __launch_bounds__(256, 1) void k_mv_fused8(const float *__restrict__ M, int64_t m,
                                                      int64_t n, const float *__restrict__ X,
                                                      double *__restrict__ part,
                                                      uint64_t *__restrict__ prec,
                                                      uint64_t *__restrict__ yrec, PcgState *st,
                                                      int32_t G, int32_t NP) {
  if (st->n_active == 0) return;  // the batch has converged: uniform early exit
  __shared__ __align__(16) char lds[kLdsBytes];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = blockIdx.x;
  const int64_t c0 = int64_t(g) * 256;
  const uint32_t epoch = st->fepoch;
  char *wbase = lds + w * kWaveBytes;
  float *ylds = reinterpret_cast<float *>(lds + kTileLds) + w * 32;
  const int lb = lane >> 2, ln = lane & 3;

  // B operands of phase A: p[c0 + 16 lb + 4 i + j][4 h + ln].
  float pB[4][4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t col = c0 + 16 * lb + 4 * i + j;
#pragma unroll
      for (int h = 0; h < 2; ++h) pB[i][j][h] = (col < n) ? X[col * 8 + 4 * h + ln] : 0.f;
    }

  // Row-order loads: lane owns columns c0 + 4 lane .. +3 of each row.  Rows
  // past m read row m - 1 and columns past n read columns n - 4 .. n - 1, so
  // every wave issues the same loads and no value is inspected at load time
  // (a select on the loaded value would wait for it at once).  Such columns
  // meet p = 0 in phase A and are never written in phase B; such rows get
  // y = 0 in phase B.
  const int64_t lcol = c0 + 4 * lane;
  const int64_t lcol_c = lcol < n ? lcol : n - 4;
  auto load_tile = [&](int32_t i, v4f (&buf)[kRows]) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int64_t f = int64_t(i) * kPanel + kRows * w + r;
      const int64_t fc = f < m ? f : m - 1;
      buf[r] = load_stream(M + fc * n + lcol_c);
    }
  };

  v4f CB[4][2];
  double QD[4][2][4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      CB[j][h] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) QD[j][h][e] = 0.0;
    }

  bool stop = false;  // a wait gave up: finish without waiting
  const int32_t steps = (NP + kDefer + kPF - 1) / kPF * kPF;
  // low half j + 1 is never 0, so zeroed records never match
  auto tag_of = [&](int32_t j) { return (epoch << 16) | uint32_t(j + 1); };
  auto clampp = [&](int32_t j) { return j < 0 ? 0 : (j >= NP ? NP - 1 : j); };
  // Partial record of this wave: lane l at index l of its 64 (lanes l and
  // l ^ 4 hold the same value after the reduce-scatter).  The 32 distinct
  // values are the lanes with bit 2 clear; value 4k + v sits at index 8k + v
  // and is row e = 2 bit 4 + bit 3 of k's bits (bit 3 = k & 1, bit 4 =
  // (k >> 1) & 1), right-hand side 4 h + v with h = k >> 2.
  auto part_rec = [&](int32_t j, int gg) {
    return prec + ((int64_t(j & (kSlots - 1)) * kW + w) * G + gg) * 64;
  };
  // y records of this wave's quarter: row e, right-hand side 4 h + n at
  // index 8 n + 2 e + h, so lane ln's eight values are contiguous.
  auto y_rec = [&](int32_t j) { return yrec + (int64_t(j & (kSlots - 1)) * kW + w) * 32; };

  // ---- reducer duty: the records of the designated panel quarter
  auto red_load = [&](int32_t j, int k, uint64_t (&dst)[4][4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int gg = lane + 64 * t;
      const uint64_t *src = part_rec(j, gg < G ? gg : G - 1) + 8 * k;
#pragma unroll
      for (int v = 0; v < 4; ++v) dst[t][v] = ld_rec(src + v);
    }
  };
  auto red_finish = [&](int32_t rj, int rk, uint64_t (&rr)[4][4]) {
    const uint32_t t = tag_of(rj);
    auto stale = [&](const uint64_t (&a)[4][4]) {
      bool bad = false;
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) bad = bad || (lane + 64 * c < G && rec_tag(a[c][v]) != t);
      return bad;
    };
    uint32_t spins = 0;
    while (__any(stale(rr)) && !stop) {
      if (++spins > kSpinLimit || aborted(st)) {
        set_abort(st);
        stop = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      red_load(rj, rk, rr);
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        if (lane + 64 * c < G) acc[v] += double(rec_val(rr[c][v]));
    // fixed-order butterfly over the 64 lanes (deterministic)
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) acc[v] += __shfl_xor(acc[v], off, 64);
    if (lane < 4) {
      const int e = 2 * ((rk >> 1) & 1) + (rk & 1), h = rk >> 2;
      const double a = lane == 0 ? acc[0] : lane == 1 ? acc[1] : lane == 2 ? acc[2] : acc[3];
      st_rec(y_rec(rj) + 8 * lane + 2 * e + h, float(a), t);
    }
  };

  // ---- phase B on tile jb with the y values of this wave's quarter (lane
  // record yv = value of index lane & 31).  The slow copy (SLOW) issues its
  // MFMAs in another order so that the compiler cannot merge the two copies
  // (a merged copy would see y as loaded just now on every path).
  auto phase_b = [&](int32_t jb, uint64_t yv, auto slow) {
    if (lane < 32) ylds[lane] = rec_val(yv);
    const v4f ya = *reinterpret_cast<const v4f *>(ylds + 8 * ln);
    const v4f yb = *reinterpret_cast<const v4f *>(ylds + 8 * ln + 4);
    const float yy[8] = {ya[0], ya[1], ya[2], ya[3], yb[0], yb[1], yb[2], yb[3]};
    const char *slot = wbase + (jb % kTiles) * (kRows * kRowBytes);
    v4f Wv[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) Wv[r] = *reinterpret_cast<const v4f *>(slot + r * kRowBytes + 16 * lane);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const bool row_ok = int64_t(jb) * kPanel + kRows * w + r < m;
      const float y0 = row_ok ? yy[2 * r] : 0.f;
      const float y1 = row_ok ? yy[2 * r + 1] : 0.f;
      if constexpr (decltype(slow)::value) {
#pragma unroll
        for (int j = 0; j < 4; ++j) CB[j][1] = mfma4(Wv[r][j], y1, CB[j][1]);
#pragma unroll
        for (int j = 0; j < 4; ++j) CB[j][0] = mfma4(Wv[r][j], y0, CB[j][0]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          CB[j][0] = mfma4(Wv[r][j], y0, CB[j][0]);
          CB[j][1] = mfma4(Wv[r][j], y1, CB[j][1]);
        }
      }
    }
  };

  // ---- one step; buf holds tile i (loaded at step i - 4), yuse the y
  // record requested at step i - 2, yreq receives this step's request.
  auto step = [&](int32_t i, v4f (&buf)[kRows], uint64_t &yuse, uint64_t &yreq) {
    // (r) reducer duty: request the records of panel jr's quarter
    const int32_t jr = i - kRedLag;
    const int rk = int((int64_t(g) - int64_t(jr) * kReducers) % G + G) % G;
    const bool reducer = jr >= 0 && jr < NP && rk < kReducers;
    uint64_t rr[4][4];
    if (reducer) red_load(jr, rk, rr);
    asm volatile("" ::: "memory");
    // (y) request y(i + kYLag - D)
    yreq = ld_rec(y_rec(clampp(i + kYLag - kDefer)) + (lane & 31));
    asm volatile("" ::: "memory");  // keep the tile prefetch behind this load
    // (a) tile i -> LDS slot i % kTiles (row order, 16 B per lane)
    char *slot = wbase + (i % kTiles) * (kRows * kRowBytes);
#pragma unroll
    for (int r = 0; r < kRows; ++r) *reinterpret_cast<v4f *>(slot + r * kRowBytes + 16 * lane) = buf[r];
    // (b) prefetch tile i + 4 into the registers just freed
    load_tile(clampp(i + kPF), buf);
    // (c) phase A: 4 rows of one column per MFMA block; two chains per h
    v4f V[4];
    {
      const char *rowp = slot + ln * kRowBytes + 64 * lb;
#pragma unroll
      for (int ii = 0; ii < 4; ++ii) V[ii] = *reinterpret_cast<const v4f *>(rowp + 16 * ii);
    }
    v4f CA[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      CA[h][0] = v4f{0.f, 0.f, 0.f, 0.f};
      CA[h][1] = v4f{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int h = 0; h < 2; ++h) CA[h][ii & 1] = mfma4(V[ii][j], pB[ii][j][h], CA[h][ii & 1]);
    // reduce-scatter over the 16 blocks (lane bits 5..2): the lane ends with
    // row e = 2 bit 4 + bit 3, right-hand side 4 h + ln, h = bit 5.
    float v[8];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < 4; ++e) v[4 * h + e] = CA[h][0][e] + CA[h][1][e];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = pair_add32(v[k], v[k + 4]);
#pragma unroll
    for (int k = 0; k < 2; ++k) v[k] = pair_add16(v[k], v[k + 2]);
    v[0] = pair_add8(v[0], v[1], lane);
    v[0] = sum_xor4(v[0], lane);
    st_rec(part_rec(i, g) + lane, v[0], tag_of(i));
    // (d) phase B on tile jb = i - D with y requested at step i - 2
    const int32_t jb = i - kDefer;
    if (jb >= 0 && jb < NP) {
      const uint32_t t = tag_of(jb);
      if (!__any(rec_tag(yuse) != t) || stop) {
        phase_b(jb, yuse, std::false_type{});
      } else {
        // y(jb) was not complete when requested: poll into a register of
        // its own (reloading yuse would make the compiler treat it as
        // freshly loaded on every path and wait for the prefetch behind it).
        uint64_t yt = yuse;
        uint32_t spins = 0;
        for (;;) {
          if (++spins > kSpinLimit || aborted(st)) {
            set_abort(st);
            stop = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          yt = ld_rec(y_rec(jb) + (lane & 31));
          if (!__any(rec_tag(yt) != t)) break;
        }
        phase_b(jb, yt, std::true_type{});
      }
      if (jb % kFlushSteps == kFlushSteps - 1 || jb == NP - 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int e = 0; e < 4; ++e) QD[j][h][e] += double(CB[j][h][e]);
            CB[j][h] = v4f{0.f, 0.f, 0.f, 0.f};
          }
      }
    }
    // (s) reducer duty: sum and publish y of panel jr's quarter
    if (reducer) red_finish(jr, rk, rr);
  };

  // Prologue: the memory operations of the four steps before the loop's
  // first one (record load, tile, record store each), in the same order, so
  // the waits at the loop header are exact.  The stores write tag-0 records
  // (never accepted by a reader) into this wave's partial record of panel 0
  // ahead of its real one.
  v4f b0[kRows], b1[kRows], b2[kRows], b3[kRows];
  uint64_t y0 = 0, y1 = 0, y2 = 0, y3 = 0;
  uint64_t *own0 = part_rec(0, g) + lane;
  auto pro = [&](int32_t i, v4f (&buf)[kRows], uint64_t &yr) {
    yr = ld_rec(y_rec(0) + (lane & 31));
    asm volatile("" ::: "memory");
    load_tile(clampp(i), buf);
    asm volatile("" ::: "memory");
    st_rec(own0, 0.f, 0u);
    asm volatile("" ::: "memory");
  };
  pro(0, b0, y0);
  pro(1, b1, y1);
  pro(2, b2, y2);
  pro(3, b3, y3);
  // Step i requests into y[i % 4] and uses the request of step i - 2, held in
  // y[(i + 2) % 4].  A multiple of four steps: no path runs part of an
  // iteration (the extra steps reduce and consume nothing).
  static_assert(kYLag == 2, "y ring rotation");
  for (int32_t i = 0; i < steps; i += 4) {
    step(i, b0, y2, y0);
    step(i + 1, b1, y3, y1);
    step(i + 2, b2, y0, y2);
    step(i + 3, b3, y1, y3);
  }

  // Sum the four waves' slices (same columns, different rows) in LDS, in
  // wave order, and write q[c][r] (FP64) for the slice.
  __syncthreads();
  // 32 values per lane; waves 1..3 park theirs (3 x 32 x 64 doubles = 48 KB).
  double *red = reinterpret_cast<double *>(lds);
  if (w > 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          red[((w - 1) * 32 + (8 * j + 4 * h + e)) * 64 + lane] = QD[j][h][e];
  }
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int idx = 8 * j + 4 * h + e;
          double s = QD[j][h][e];
          s += red[(0 * 32 + idx) * 64 + lane];
          s += red[(1 * 32 + idx) * 64 + lane];
          s += red[(2 * 32 + idx) * 64 + lane];
          const int64_t col = c0 + 16 * lb + 4 * e + j;
          if (col < n) part[col * 8 + 4 * h + ln] = s;
        }
  }
  // The last wave of the launch advances the epoch (records of this launch
  // then never match the next one's tags).
  if (lane == 0) {
    const uint32_t prev =
        __hip_atomic_fetch_add(&st->fdone, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == uint32_t(G) * kW - 1) {
      __hip_atomic_store(&st->fdone, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&st->fepoch, epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace

int32_t fused_blocks(int64_t m, int64_t n, int K, int cus) {
  if (K != 8 || n % 4 != 0 || n < 4 || m < 1) return 0;
  const int64_t G = (n + 255) / 256;
  const int64_t NP = (m + kPanel - 1) / kPanel;
  if (G < kReducers || G > cus || NP >= 65535) return 0;
  return int32_t(G);
}

size_t fused_records(int32_t G) {
  return size_t(kSlots) * kW * size_t(G) * 64 + size_t(kSlots) * kW * 32;
}

hipError_t mv_fused8(hipStream_t s, const float *M, int64_t m, int64_t n, const float *X,
                     double *part, uint64_t *records, PcgState *st, int32_t G) {
  const int32_t NP = int32_t((m + kPanel - 1) / kPanel);
  uint64_t *prec = records;
  uint64_t *yrec = records + size_t(kSlots) * kW * size_t(G) * 64;
  void *args[] = {(void *)&M, (void *)&m, (void *)&n, (void *)&X, (void *)&part,
                  (void *)&prec, (void *)&yrec, (void *)&st, (void *)&G, (void *)&NP};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_mv_fused8), dim3(G),
                                    dim3(kW * 64), args, 0, s);
}

}  // namespace mfs
