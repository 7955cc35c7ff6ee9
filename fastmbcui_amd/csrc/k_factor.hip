// k_factor.hip -- supernodal KL factor on gfx950.
//
// The reference builds the factor with three one-thread-per-item kernels
// (src/kl_chol.cu:166-223 Gram, :347-360 Cholesky, :362-387 back-substitution):
// each thread walks whole dot products and triangular solves serially.  Here
// one 256-thread block owns one supernode and does all three steps in LDS:
//
//  1. Gram  TH = A^T A, A = the supernode's columns of M (gathered rows of M^T)
//     or, in the points form, kernel values evaluated on the fly; on MFMA
//     16x16x4 tiles (FP64 or FP32), lower tile triangle only.  Every lane
//     holds exactly the operand entries it feeds to the MFMAs (loaded as
//     float4 rows of M^T, or evaluated), the four waves split the f-range and
//     add their accumulators at the end; stored reversed, TH[ns-1-p][ns-1-q],
//     as the reference stores it (kl_chol.cu:216-220).
//  2. Upper Cholesky TH = U^T U, right-looking, one barrier per column.  A
//     non-positive pivot (FP32: one within the Gram's rounding noise) marks
//     the supernode (status = 1) instead of producing NaN or noise (the
//     reference does not check, kl_chol.cu:33).
//  3. For each member column j with nb rows: x = U[0:nb,0:nb]^{-1} e_{nb-1},
//     written in ascending row order -- the KL column
//     Theta_j^{-1} e_1 / sqrt(e_1^T Theta_j^{-1} e_1) (kl_chol.cu:378-386).
//     Members are solved in batches of 16 with one barrier per row.
#include "kernel_eval.h"
#include "mfs_internal.h"

namespace mfs {
namespace {

template <typename T>
struct Mfma;

// A (16x4) and B (4x16) operands: lane l holds A[l%16][l/16] and B[l/16][l%16].
// The 16x16 result sits in 4 registers per lane with column l%16; the row of
// register v differs by precision (measured on gfx950): FP64 puts rows
// l/16 + 4v in a lane, FP32 puts rows 4(l/16) + v.
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 run(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 run(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <int A, int B>
struct CMax {
  static constexpr int v = A > B ? A : B;
};

// Lower-triangle tile pair pi -> (a, b), a >= b, row-major over a.
__host__ __device__ constexpr int pair_a(int pi) {
  int a = 0;
  while ((a + 1) * (a + 2) / 2 <= pi) ++a;
  return a;
}
__host__ __device__ constexpr int pair_b(int pi) { return pi - pair_a(pi) * (pair_a(pi) + 1) / 2; }

template <int I>
struct IntC {
  static constexpr int value = I;
};
// f(IntC<0>{}), ..., f(IntC<N-1>{}) in order.
template <int N, int I = 0, typename F>
__device__ inline void static_for(F &&f) {
  if constexpr (I < N) {
    f(IntC<I>{});
    static_for<N, I + 1>(f);
  }
}

// One k-step of the Gram for the tile pairs GRP + NG (U0 + u) (all
// compile-time, so each MFMA reads its operands from fixed registers).
template <int T, int NG, int PPW, int GRP, int U0, typename TC, typename acc_t>
__device__ inline void gram_step(const TC (&fr)[T], acc_t (&acc)[PPW]) {
#pragma unroll
  for (int u = 0; u < PPW; ++u) {
    constexpr int kPairs = T * (T + 1) / 2;
    const int pi = GRP + NG * (U0 + u);
    if (pi < kPairs) acc[u] = Mfma<TC>::run(fr[pair_a(pi)], fr[pair_b(pi)], acc[u]);
  }
}

// Split FP32 Gram products (GB): x = hi + lo + r with hi the BF16 rounding of
// x, lo that of x - hi (exact in FP32) and |r| <= 2^-16 |x|; a b is taken as
// hi_a hi_b + hi_a lo_b + lo_a hi_b, summed in FP32 -- the three-product
// scheme of Henry, Tang and Heinecke, "Leveraging the bfloat16 artificial
// intelligence datatype for higher-precision computations" (ARITH 2019).  The
// dropped terms (lo_a lo_b and those with r) total at most about 3 2^-16 |a b|,
// and round to nearest leaves them of either sign but lo_a lo_b.  The BF16
// 16x16x32 MFMA does 8x the k-extent of the FP32 16x16x4 one in half its
// cycles, so the three products cost 3/16 of the FP32 MFMA time.
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

// hi goes back to FP32 from its packed words (element 2i is the low half of
// word i): a shift and a mask per pair, not a conversion per value.
__device__ inline void split_bf16(f4_t x0, f4_t x1, bf8_t &hi, bf8_t &lo) {
  const f8_t x = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
  hi = __builtin_convertvector(x, bf8_t);
  const u4_t h = __builtin_bit_cast(u4_t, hi);
  f8_t hf;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    hf[2 * i] = __uint_as_float(h[i] << 16);
    hf[2 * i + 1] = __uint_as_float(h[i] & 0xffff0000u);
  }
  lo = __builtin_convertvector(x - hf, bf8_t);
}

// split_bf16 for two values (x0 the lower half of each word), as its words.
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));
__device__ inline void split_pair(float x0, float x1, uint32_t &hw, uint32_t &lw) {
  hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2_t{x0, x1}, bf2_t));
  const float h0 = __uint_as_float(hw << 16), h1 = __uint_as_float(hw & 0xffff0000u);
  lw = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2_t{x0 - h0, x1 - h1}, bf2_t));
}

// One 32-row k-step of the split Gram for the tile pairs GRP + NG (U0 + u):
// the large products of all pairs first, then the two small ones, so that no
// MFMA waits for the one before it on the same accumulator.
template <int T, int NG, int PPW, int GRP, int U0>
__device__ inline void gram_step3(const bf8_t (&hi)[T], const bf8_t (&lo)[T], f4_t (&acc)[PPW]) {
  constexpr int kPairs = T * (T + 1) / 2;
#pragma unroll
  for (int u = 0; u < PPW; ++u) {
    const int pi = GRP + NG * (U0 + u);
    if (pi < kPairs)
      acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[pair_a(pi)], hi[pair_b(pi)], acc[u], 0, 0, 0);
  }
#pragma unroll
  for (int u = 0; u < PPW; ++u) {
    const int pi = GRP + NG * (U0 + u);
    if (pi < kPairs)
      acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[pair_a(pi)], lo[pair_b(pi)], acc[u], 0, 0, 0);
  }
#pragma unroll
  for (int u = 0; u < PPW; ++u) {
    const int pi = GRP + NG * (U0 + u);
    if (pi < kPairs)
      acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[pair_a(pi)], hi[pair_b(pi)], acc[u], 0, 0, 0);
  }
}

// ------------------------------------------------------------------ transpose
// Mt[c][f] = M[f][c] for f < m, 0 for m <= f < ldmt, for the columns c from cb
// (a multiple of 64) on.  64x64 tiles through LDS, float4 in and out when the
// row length allows it.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void k_transpose(const float *__restrict__ M, int64_t m,
                                                        int64_t n, float *__restrict__ Mt,
                                                        int64_t ldmt, int64_t cb) {
  __shared__ float tile[64][65];
  const int t = threadIdx.x;
  const int64_t c0 = cb + int64_t(blockIdx.x) * 64, f0 = int64_t(blockIdx.y) * 64;
  const int q = t & 15, r16 = t >> 4;  // 16 float4 per 64-wide row, 16 rows per pass
  for (int r = r16; r < 64; r += 16) {
    const int64_t f = f0 + r, c = c0 + 4 * q;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (f < m) {
      if (VEC && c + 3 < n) {
        // M is read once: non-temporal, so it does not evict what L2 reuses
        typedef float f4v __attribute__((ext_vector_type(4)));
        const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(M + f * n + c));
        v = make_float4(t.x, t.y, t.z, t.w);
      } else {
        if (c + 0 < n) v.x = M[f * n + c + 0];
        if (c + 1 < n) v.y = M[f * n + c + 1];
        if (c + 2 < n) v.z = M[f * n + c + 2];
        if (c + 3 < n) v.w = M[f * n + c + 3];
      }
    }
    tile[r][4 * q + 0] = v.x;
    tile[r][4 * q + 1] = v.y;
    tile[r][4 * q + 2] = v.z;
    tile[r][4 * q + 3] = v.w;
  }
  __syncthreads();
  for (int r = r16; r < 64; r += 16) {
    const int64_t c = c0 + r;
    if (c >= n) continue;
    const float4 v = make_float4(tile[4 * q + 0][r], tile[4 * q + 1][r], tile[4 * q + 2][r],
                                 tile[4 * q + 3][r]);
    *reinterpret_cast<float4 *>(Mt + c * ldmt + f0 + 4 * q) = v;
  }
}

// Mt[i][f] = M[f][cols[i]]: the compact M^T of the columns that supernodes
// outside the bands read (scattered reads along a row of M, 16-byte writes
// along a row of Mt), zero for m <= f < ldmt.
__global__ __launch_bounds__(kThreads) void k_gather_transpose(const float *__restrict__ M, int64_t m,
                                                               int64_t n, const int32_t *__restrict__ cols,
                                                               int32_t ncols, float *__restrict__ Mt,
                                                               int64_t ldmt) {
  __shared__ float tile[64][65];
  const int t = threadIdx.x;
  const int64_t i0 = int64_t(blockIdx.x) * 64, f0 = int64_t(blockIdx.y) * 64;
  const int cc = t & 63;
  const bool have = i0 + cc < ncols;
  const int64_t col = have ? cols[i0 + cc] : 0;
  for (int r = t >> 6; r < 64; r += kThreads / 64) {
    const int64_t f = f0 + r;
    tile[r][cc] = (have && f < m) ? M[f * n + col] : 0.f;
  }
  __syncthreads();
  const int q = t & 15;
  for (int ci = t >> 4; ci < 64; ci += kThreads / 16) {
    const int64_t i = i0 + ci;
    if (i >= ncols) continue;
    *reinterpret_cast<float4 *>(Mt + i * ldmt + f0 + 4 * q) =
        make_float4(tile[4 * q][ci], tile[4 * q + 1][ci], tile[4 * q + 2][ci], tile[4 * q + 3][ci]);
  }
}

// ------------------------------------------------------------------ band Gram
// The Grams of a band's supernodes straight from row-major M, with no M^T: a
// block stages 32 rows x kBandCols columns of M in LDS, column-major with 36
// floats per column (a column's 32 rows contiguous), while the next chunk's
// loads are in flight in registers: 8 bytes (two columns) per lane and row,
// 32 loads per thread, within the wave's count of loads in flight.  Each wave owns whole supernodes of one tile count
// T: lane l supplies, for tile a, chunk rows 8 (l >> 4) .. + 7 of column
// p = 16 a + (l & 15) (two 16-byte LDS reads) as the eight k-values of its
// slot in both operands of the BF16 MFMA, split into hi + lo (three products
// per FP32 product, as in the Gram from M^T), and keeps its supernodes'
// accumulators in registers over all chunks of the block's row part.  LDS
// column kBandCols stays zero: padding columns of a last tile read it.  The
// partial Grams go to gram (16 T x 16 T per supernode, both triangles).
constexpr int kBandRows = 32;
constexpr int kBandStride = kBandRows + 4;
constexpr int kBandThreads = kBandWaves * 64;
constexpr int kBandLds = (kBandCols + 1) * kBandStride;  // floats
static_assert(2 * kBandThreads == kBandCols, "the loads cover a band in two columns per thread");
static_assert(kBandLds * 4 <= 160 * 1024, "LDS");

template <int T>
__device__ inline void band_wave(const float *__restrict__ M, int64_t m, int64_t n, int64_t c0, int cols,
                                 int64_t r0, int64_t r1, float *chunk, const BandWave W,
                                 const int32_t *__restrict__ items, const SuperDesc *__restrict__ sd,
                                 const int32_t *__restrict__ sind, const int32_t *__restrict__ Pinv,
                                 const int64_t *__restrict__ goff, float *__restrict__ gp) {
  constexpr int C = band_capacity(T);
  constexpr int P = T * (T + 1) / 2;
  // supernodes whose MFMAs interleave (an accumulator's next product is at
  // least P * GS instructions later)
  constexpr int GS = T == 1 ? 6 : (T == 2 ? 2 : 1);
  static_assert(C % GS == 0, "groups");
  const int tid = threadIdx.x, lane = tid & 63;
  // LDS offsets (floats) of the lane's column of every (supernode, tile);
  // columns outside the band (none, by construction) read the zero column
  int off[C][T];
#pragma unroll
  for (int j = 0; j < C; ++j) {
    int32_t ns = 0, ro = 0;
    if (j < W.count) {
      const SuperDesc D = sd[items[W.first + j]];
      ns = D.ns;
      ro = D.rows_off;
    }
#pragma unroll
    for (int a = 0; a < T; ++a) {
      const int p = 16 * a + (lane & 15);
      int64_t c = p < ns ? int64_t(Pinv[sind[ro + p]]) - c0 : int64_t(kBandCols);
      if (c < 0 || c > kBandCols) c = kBandCols;
      off[j][a] = int(c) * kBandStride + 8 * (lane >> 4);
    }
  }
  f4_t acc[C][P];
#pragma unroll
  for (int j = 0; j < C; ++j)
#pragma unroll
    for (int u = 0; u < P; ++u) acc[j][u] = f4_t{0.f, 0.f, 0.f, 0.f};

  // Loads: columns 2 tid and 2 tid + 1 of the window of all 32 rows of a
  // chunk (8 bytes per lane, a 512-byte row segment per wave; n is even, so
  // rows and pairs start 8-byte aligned).  A pair past the end of a row is
  // moved back to the row's last pair; those columns lie outside the block's
  // supernodes.  Lanes past the window's width neither load nor store: no
  // supernode of the block reads those columns.
  typedef float f2_t __attribute__((ext_vector_type(2)));
  const bool stage = 2 * tid < cols;
  const int64_t cp = c0 + 2 * tid;
  const uint32_t oc = uint32_t(cp + 1 < n ? cp : n - 2) * 4u;
  f2_t rv[kBandRows];
  auto load = [&](int64_t f0) {
    if (!stage) return;
#pragma unroll
    for (int r = 0; r < kBandRows; ++r) {
      const int64_t f = f0 + r < m ? f0 + r : m - 1;
      const char *row = reinterpret_cast<const char *>(M + f * n);
      rv[r] = *reinterpret_cast<const f2_t *>(row + oc);
    }
  };
  // rows past the part (the next part's, or past m) enter as zeros
  auto store = [&](int64_t f0) {
    if (!stage) return;
#pragma unroll
    for (int q = 0; q < kBandRows / 4; ++q) {
      f4_t va, vb;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool live = f0 + 4 * q + i < r1;
        va[i] = live ? rv[4 * q + i].x : 0.f;
        vb[i] = live ? rv[4 * q + i].y : 0.f;
      }
      *reinterpret_cast<f4_t *>(chunk + (2 * tid) * kBandStride + 4 * q) = va;
      *reinterpret_cast<f4_t *>(chunk + (2 * tid + 1) * kBandStride + 4 * q) = vb;
    }
  };

  const int64_t nch = r1 > r0 ? (r1 - r0 + kBandRows - 1) / kBandRows : 0;
  if (nch > 0) load(r0);
  for (int64_t k = 0; k < nch; ++k) {
    __syncthreads();  // every wave is done with the previous chunk
    store(r0 + k * kBandRows);
    __syncthreads();
    if (k + 1 < nch) load(r0 + (k + 1) * kBandRows);
#pragma unroll
    for (int j0 = 0; j0 < C; j0 += GS) {
      if (j0 >= W.count) break;
      bf8_t hi[GS][T], lo[GS][T];
#pragma unroll
      for (int g = 0; g < GS; ++g)
#pragma unroll
        for (int a = 0; a < T; ++a) {
          const f4_t x0 = *reinterpret_cast<const f4_t *>(chunk + off[j0 + g][a]);
          const f4_t x1 = *reinterpret_cast<const f4_t *>(chunk + off[j0 + g][a] + 4);
          split_bf16(x0, x1, hi[g][a], lo[g][a]);
        }
#pragma unroll
      for (int g = 0; g < GS; ++g)
#pragma unroll
        for (int u = 0; u < P; ++u)
          acc[j0 + g][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[g][pair_a(u)], hi[g][pair_b(u)],
                                                                   acc[j0 + g][u], 0, 0, 0);
#pragma unroll
      for (int g = 0; g < GS; ++g)
#pragma unroll
        for (int u = 0; u < P; ++u)
          acc[j0 + g][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(hi[g][pair_a(u)], lo[g][pair_b(u)],
                                                                   acc[j0 + g][u], 0, 0, 0);
#pragma unroll
      for (int g = 0; g < GS; ++g)
#pragma unroll
        for (int u = 0; u < P; ++u)
          acc[j0 + g][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[g][pair_a(u)], hi[g][pair_b(u)],
                                                                   acc[j0 + g][u], 0, 0, 0);
    }
  }

  // Partial Grams: lane l holds rows 4 (l >> 4) + v, column l & 15 of each
  // 16 x 16 tile product.
  constexpr int NS = 16 * T;
#pragma unroll
  for (int j = 0; j < C; ++j) {
    if (j >= W.count) break;
    float *g = gp + goff[items[W.first + j]];
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const int a = pair_a(u), b = pair_b(u);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int p = 16 * a + 4 * (lane >> 4) + v, q = 16 * b + (lane & 15);
        g[p * NS + q] = acc[j][u][v];
        if (a != b) g[q * NS + p] = acc[j][u][v];
      }
    }
  }
}

__global__ __launch_bounds__(kBandThreads, 1) void k_band_gram(
    const float *__restrict__ M, int64_t m, int64_t n, const BandBlock *__restrict__ blocks,
    const int32_t *__restrict__ items, const SuperDesc *__restrict__ sd,
    const int32_t *__restrict__ sind, const int32_t *__restrict__ Pinv,
    const int64_t *__restrict__ goff, float *__restrict__ gram, int64_t gram_part, int64_t part_rows) {
  __shared__ __align__(16) float chunk[kBandLds];
  const BandBlock &B = blocks[blockIdx.x];
  const int32_t c0 = B.c0;
  if (c0 < 0) return;  // padding block (uniform)
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const BandWave W = B.w[wave];
  if (threadIdx.x < kBandStride) chunk[kBandCols * kBandStride + threadIdx.x] = 0.f;
  const int64_t r0 = int64_t(B.part) * part_rows;
  const int64_t r1 = r0 + part_rows < m ? r0 + part_rows : m;
  float *gp = gram + int64_t(B.part) * gram_part;
  // Every wave runs the same chunk loop (loads, stores, barriers); idle waves
  // (count 0) take the T = 1 path and compute nothing.
  switch (W.tiles) {
    case 2: band_wave<2>(M, m, n, c0, B.cols, r0, r1, chunk, W, items, sd, sind, Pinv, goff, gp); break;
    case 3: band_wave<3>(M, m, n, c0, B.cols, r0, r1, chunk, W, items, sd, sind, Pinv, goff, gp); break;
    case 4: band_wave<4>(M, m, n, c0, B.cols, r0, r1, chunk, W, items, sd, sind, Pinv, goff, gp); break;
    case 6: band_wave<6>(M, m, n, c0, B.cols, r0, r1, chunk, W, items, sd, sind, Pinv, goff, gp); break;
    default: band_wave<1>(M, m, n, c0, B.cols, r0, r1, chunk, W, items, sd, sind, Pinv, goff, gp); break;
  }
}

// ------------------------------------------------------------------ fused supernode factor
// PTS: the points form -- the Gram's operands are evaluated from the packed
// point records (tgt: m x 12, srcp: n x 8) instead of read from M^T.  Each
// lane evaluates exactly the MFMA operand entries it supplies, A[f][p] with
// p = 16 a + (lane & 15) for every tile a and f = f0 + (lane >> 4) (FP32
// products; f = f0 + 8 (lane >> 4) + i, i < 8, with split products), keeping
// its T source points in registers: no LDS staging and no barriers.
// GIN: the Gram comes summed from the band partials (FS.gram); only the
// Cholesky and the KL columns run.
//
// The LDS of one supernode factor, carved from one buffer (a launch per
// bucket declares it per tile count; the queue launch sizes it for its
// largest tile count): TH (NS x RS, none for the large tiles: global scratch),
// the local-to-original column map, the pivots and their inverses, the
// member right-hand sides W, the member batch's bookkeeping, the split flag,
// the FP32 diagonals.
template <int T, typename TC>
struct FactorLds {
  static constexpr int NS = 16 * T, RS = NS + 1, MB = 16;
  static constexpr size_t al(size_t x) { return (x + 15) / 16 * 16; }
  static constexpr size_t o_src = al(T > 8 ? 16 : size_t(NS) * RS * sizeof(TC));
  static constexpr size_t o_dg = al(o_src + size_t(NS) * 4);
  static constexpr size_t o_dinv = al(o_dg + size_t(NS) * sizeof(TC));
  static constexpr size_t o_w = al(o_dinv + size_t(NS) * sizeof(TC));
  static constexpr size_t o_mcol = al(o_w + size_t(NS) * MB * sizeof(TC));
  static constexpr size_t o_mc = o_mcol + MB * 4, o_moff = o_mc + MB * 4, o_last = o_moff + MB * 4;
  static constexpr size_t o_d0 = al(o_last + 4);
  static constexpr size_t bytes = al(o_d0 + (sizeof(TC) == 4 ? size_t(NS) : 1) * sizeof(TC));
};

// One unit of a factor launch: a supernode, or one row split of it.  `unit`
// numbers the units of the launch's items (item * S + split); it also picks
// the unit's regions of th_scratch and of the split partials.  Forced inline:
// with two callers (a bucket's kernel and the batch queue) the compiler would
// keep it out of line, a call per unit with its LDS behind flat pointers and
// its registers spilled.
template <int T, typename TC, typename TV, bool PTS, bool GB, bool GIN>
__device__ __forceinline__ void factor_unit(const FactorScene *__restrict__ scenes,
                                   const FactorItem *__restrict__ items, TC *__restrict__ th_scratch,
                                   FactorSplit sp, int32_t unit, unsigned char *lds) {
  // This unit's supernode and its scene (one launch may mix scenes); with
  // row splits, units item * S .. item * S + S - 1 share one supernode.
  const int32_t S = sp.splits;
  const int32_t bi = S > 1 ? unit / S : unit;
  const int32_t split = S > 1 ? unit % S : 0;
  const FactorItem item = items[bi];
  const FactorScene &FS = scenes[item.scene];
  const float *__restrict__ Mt = FS.Mt;
  const int64_t ldmt = FS.ldmt;
  const float *__restrict__ tgt = FS.tgt;
  const float *__restrict__ srcp = FS.src;
  const int64_t m = FS.m;
  const SuperDesc *__restrict__ sd = FS.sd;
  const int32_t *__restrict__ sind = FS.sind;
  const int32_t *__restrict__ Pinv = FS.Pinv;
  const int32_t *__restrict__ sup_ind = FS.sup_ind;
  const int32_t *__restrict__ sptr = FS.sptr;
  TV *__restrict__ val = static_cast<TV *>(FS.val);
  int32_t *__restrict__ status = FS.status;
  constexpr int NS = 16 * T;                  // padded local size
  // Supernodes of more than 128 local rows (T = 12, 16) keep TH in global
  // scratch (NS x RS per block; it stays in L2) since it would not fit in LDS.
  constexpr bool LARGE = T > 8;
  constexpr int RS = NS + 1;                  // TH row stride
  constexpr int MB = 16;                      // member columns per batch
  constexpr int PAIRS = T * (T + 1) / 2;      // lower-triangle tile pairs
  // Waves split the rows (KS) and/or the tile pairs (NG groups); the larger
  // the tiles (and FP64 accumulators), the more the pairs are split, to bound
  // the accumulator registers.
  constexpr int KS = LARGE ? 1
                     : T <= 4 ? 4
                     : sizeof(TC) == 4 ? (T <= 6 ? 4 : 2)
                     : (T <= 6 ? 2 : 1);
  constexpr int NG = 4 / KS;                  // groups splitting the tile pairs
  // The large tiles also split the pairs over sequential passes over the rows
  // (operands re-read or re-evaluated) to bound the accumulator registers.
  constexpr int NPASS = LARGE ? (sizeof(TC) == 8 ? 3 : 2) : 1;
  constexpr int PPW = (PAIRS + NG * NPASS - 1) / (NG * NPASS);  // tile pairs per wave and pass
  typedef typename Mfma<TC>::v4 acc_t;

  typedef FactorLds<T, TC> LY;
  static_assert(LY::MB == MB, "member batch");
  unsigned char *big_raw = lds;
  int32_t *src = reinterpret_cast<int32_t *>(lds + LY::o_src);
  TC *dg = reinterpret_cast<TC *>(lds + LY::o_dg);
  TC *dinv = reinterpret_cast<TC *>(lds + LY::o_dinv);
  TC *W = reinterpret_cast<TC *>(lds + LY::o_w);
  int32_t *mcol = reinterpret_cast<int32_t *>(lds + LY::o_mcol);
  int32_t *mc = reinterpret_cast<int32_t *>(lds + LY::o_mc);
  int32_t *moff = reinterpret_cast<int32_t *>(lds + LY::o_moff);
  int &last_split = *reinterpret_cast<int *>(lds + LY::o_last);

  // wave index in a scalar register: loop bounds derived from it stay
  // uniform branches (and exact wait counts) instead of exec-masked regions
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int32_t s = item.s;
  const SuperDesc D = sd[s];
  const int ns = D.ns;

  const int32_t *__restrict__ colmap = FS.colmap;
  for (int p = tid; p < NS; p += kThreads) {
    int32_t c = -1;
    if (p < ns) {
      c = Pinv[sind[D.rows_off + p]];
      if (colmap != nullptr) c = colmap[c];  // row of the compact M^T
    }
    src[p] = c;
  }
  __syncthreads();

  // ---- 1. Gram on MFMA ------------------------------------------------------
  // Rows [r0, r1) of this block's split (all rows without splits); chunks are
  // multiples of 256 rows, so every wave's slabs and row quads stay inside.
  const int64_t rows_total = PTS ? m : ldmt;
  const int64_t chunk = S > 1 ? factor_chunk(rows_total, S) : rows_total;
  const int64_t r0 = int64_t(split) * chunk;
  const int64_t r1 = (S > 1 && r0 + chunk < rows_total) ? r0 + chunk : rows_total;
  const int grp = wave % NG, kph = wave / NG;
  TC *TH = LARGE ? th_scratch + size_t(unit) * NS * RS : reinterpret_cast<TC *>(big_raw);
  // Pass P accumulates the pairs grp + NG (P PPW + u), u < PPW, over all rows.
  auto gram_pass = [&](auto pass_c) {
  constexpr int U0 = decltype(pass_c)::value * PPW;
  int pa[PPW], pb[PPW];
#pragma unroll
  for (int u = 0; u < PPW; ++u) {
    int pi = grp + NG * (U0 + u), a = 0;
    if (pi >= PAIRS) pi = -1;
    if (pi >= 0) {
      while ((a + 1) * (a + 2) / 2 <= pi) ++a;
      pa[u] = a;
      pb[u] = pi - a * (a + 1) / 2;
    } else {
      pa[u] = -1;
      pb[u] = -1;
    }
  }
  acc_t acc[PPW];
#pragma unroll
  for (int u = 0; u < PPW; ++u) acc[u] = acc_t{TC(0), TC(0), TC(0), TC(0)};

  auto step = [&](const TC(&fr)[T]) {
    if (NG == 1) {
      gram_step<T, NG, PPW, 0, U0>(fr, acc);
    } else if (NG == 2) {
      if (grp == 0) gram_step<T, NG, PPW, 0, U0>(fr, acc);
      else gram_step<T, NG, PPW, 1 % NG, U0>(fr, acc);
    } else {
      switch (grp) {
        case 0: gram_step<T, NG, PPW, 0, U0>(fr, acc); break;
        case 1: gram_step<T, NG, PPW, 1 % NG, U0>(fr, acc); break;
        case 2: gram_step<T, NG, PPW, 2 % NG, U0>(fr, acc); break;
        default: gram_step<T, NG, PPW, 3 % NG, U0>(fr, acc); break;
      }
    }
  };

  auto step3 = [&](const bf8_t(&hi)[T], const bf8_t(&lo)[T]) {
    if constexpr (GB) {
      if (NG == 1) {
        gram_step3<T, NG, PPW, 0, U0>(hi, lo, acc);
      } else if (NG == 2) {
        if (grp == 0) gram_step3<T, NG, PPW, 0, U0>(hi, lo, acc);
        else gram_step3<T, NG, PPW, 1 % NG, U0>(hi, lo, acc);
      } else {
        switch (grp) {
          case 0: gram_step3<T, NG, PPW, 0, U0>(hi, lo, acc); break;
          case 1: gram_step3<T, NG, PPW, 1 % NG, U0>(hi, lo, acc); break;
          case 2: gram_step3<T, NG, PPW, 2 % NG, U0>(hi, lo, acc); break;
          default: gram_step3<T, NG, PPW, 3 % NG, U0>(hi, lo, acc); break;
        }
      }
    }
  };

  if constexpr (PTS) {
    float yh[T][3], yl[T][3];
    bool ok[T];
#pragma unroll
    for (int a = 0; a < T; ++a) {
      const int p = a * 16 + (lane & 15);
      ok[a] = p < ns;
      const float *y = srcp + int64_t(ok[a] ? src[p] : src[0]) * 8;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        yh[a][d] = y[d];
        yl[a][d] = y[3 + d];
      }
    }
    // Row records, branch-free: a row past the end reads the last row with a
    // zero scale, so its values are exactly zero (a target never coincides
    // with a source, so the distance stays finite).  A branch around the
    // loads would end in a wait for them, exposing every row's latency.
    typedef const __attribute__((address_space(1))) float gfloat;
    typedef const __attribute__((address_space(1))) f4_t gf4;
    gfloat *const tg = (gfloat *)tgt;
    const float scale = float(-kInv4Pi);
    struct Rec {
      f4_t a, b;   // x_hi[3] x_lo[0] | x_lo[1..2] n[0..1]
      float c, s;  // n[2], the scale (0 past the end)
    };
    auto load_rec = [&](int64_t f, Rec &r) {
      const bool in = f < m;
      gfloat *p = tg + (in ? f : m - 1) * 12;
      r.a = *(gf4 *)p;
      r.b = *(gf4 *)(p + 4);
      r.c = p[8];
      r.s = in ? scale : 0.f;
    };
    auto eval = [&](const Rec &t, int a) {
      return dlp32((t.a.x - yh[a][0]) + (t.a.w - yl[a][0]), (t.a.y - yh[a][1]) + (t.b.x - yl[a][1]),
                   (t.a.z - yh[a][2]) + (t.b.y - yl[a][2]), t.s * t.b.z, t.s * t.b.w, t.s * t.c);
    };
    if constexpr (GB) {
      // Split products (FP32 values, FP32 sums, as for the Gram from M^T):
      // lane l supplies the k-slots 8 (l >> 4) .. + 7 of both operands, rows
      // f0 + 8 (l >> 4) + i, i < 8, of each 32-row step.  It evaluates them a
      // row at a time, the next row's record in flight (the next step's first
      // row after the eighth), and splits each tile's eight values once they
      // are complete: the same evaluations as the FP32 products, 3/16 of their
      // MFMA time.  FSTEP <= 128 divides the row splits' chunks.
      constexpr int FSTEP = 32 * KS;
      const int64_t fs = r0 + kph * 32 + 8 * (lane >> 4);  // this lane's first row
      Rec cur, nxt;
      load_rec(fs, cur);
      for (int64_t f0 = r0 + kph * 32, fb = fs; f0 < r1; f0 += FSTEP, fb += FSTEP) {
        // Rows are split in pairs as soon as both are evaluated (hi and lo
        // words of two BF16 each), so a step holds the words and one row's
        // values, not all eight rows' values beside their split.
        u4_t hw[T], lw[T];
        float pend[T];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          load_rec(i < 7 ? fb + i + 1 : fb + FSTEP, nxt);
#pragma unroll
          for (int a = 0; a < T; ++a) {
            const float e = ok[a] ? eval(cur, a) : 0.f;
            if (i % 2 == 0) {
              pend[a] = e;
            } else {
              uint32_t h, l;
              split_pair(pend[a], e, h, l);
              hw[a][i / 2] = h;
              lw[a][i / 2] = l;
            }
          }
          cur = nxt;
        }
        bf8_t hi[T], lo[T];
#pragma unroll
        for (int a = 0; a < T; ++a) {
          hi[a] = __builtin_bit_cast(bf8_t, hw[a]);
          lo[a] = __builtin_bit_cast(bf8_t, lw[a]);
        }
        step3(hi, lo);
      }
    } else {
      // FP32 products: one row per lane and k-step.  This loop keeps the
      // branch around its loads (padding rows far away with a zero normal):
      // at 1-6 tiles it measured faster than the branch-free records above
      // (30.8 against 34.3 ms for the factor of 64 ellipsoids x 1024 points).
      auto load_target = [&](int64_t f, float4 (&t)[3]) {
        if (f < m) {
          t[0] = *reinterpret_cast<const float4 *>(tgt + f * 12);
          t[1] = *reinterpret_cast<const float4 *>(tgt + f * 12 + 4);
          t[2] = *reinterpret_cast<const float4 *>(tgt + f * 12 + 8);
        } else {
          t[0] = make_float4(1e15f, 0.f, 0.f, 0.f);
          t[1] = make_float4(0.f, 0.f, 0.f, 0.f);
          t[2] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
      };
      constexpr int FSTEP = 4 * KS;  // rows per k-step over the waves sharing a pair group
      float4 cur[3], nxt[3];
      int64_t f = r0 + kph * 4 + (lane >> 4);
      load_target(f, cur);
      for (int64_t f0 = r0 + kph * 4; f0 < r1; f0 += FSTEP, f += FSTEP) {
        load_target(f + FSTEP, nxt);
        const float sn0 = scale * cur[1].z, sn1 = scale * cur[1].w, sn2 = scale * cur[2].x;
        TC fr[T];
#pragma unroll
        for (int a = 0; a < T; ++a) {
          const float v = dlp32((cur[0].x - yh[a][0]) + (cur[0].w - yl[a][0]),
                                (cur[0].y - yh[a][1]) + (cur[1].x - yl[a][1]),
                                (cur[0].z - yh[a][2]) + (cur[1].y - yl[a][2]), sn0, sn1, sn2);
          fr[a] = ok[a] ? TC(v) : TC(0);
        }
        step(fr);
#pragma unroll
        for (int d = 0; d < 3; ++d) cur[d] = nxt[d];
      }
    }
  } else if constexpr (GB) {
    // Split products (FP32 values, FP32 sums): lane l loads, for every tile a
    // and 32-row group f0, rows f0 + 4 (l >> 4) .. + 3 and f0 + 16 + 4 (l >> 4)
    // .. + 3 of column p = 16 a + (l & 15) (two float4 of a row of M^T; each
    // load instruction reads 64 consecutive bytes per column) and supplies
    // them as the eight k-values of its slot in both operands of the BF16
    // MFMA: any placement of the rows on k serves a Gram as long as both
    // operands share it.  Padding columns, and slabs past the end of the rows
    // (re-read so the loop has no branch), load the zero row Mz instead of
    // being masked.
    typedef const __attribute__((address_space(1))) float gfloat;
    typedef const __attribute__((address_space(1))) f4_t gf4;
    gfloat *const zrow = (gfloat *)FS.Mz + 4 * (lane >> 4);
    gfloat *col[T];
#pragma unroll
    for (int a = 0; a < T; ++a) {
      const int p = a * 16 + (lane & 15);
      col[a] = p < ns ? (gfloat *)Mt + int64_t(src[p]) * ldmt + 4 * (lane >> 4) : zrow;
    }
    // W 32-row groups per slab and wave (64 rows for the small tiles, as the
    // FP32 products read); FSTEP <= 256 divides the row splits' chunks.
    constexpr int W = LARGE ? 1 : (T <= 4 ? 2 : 1);
    constexpr int FSTEP = 32 * W * KS;
    auto load = [&](int64_t f0, bool live, f4_t (&v)[T][2 * W]) {
#pragma unroll
      for (int a = 0; a < T; ++a) {
        gfloat *b = live ? col[a] : zrow;
#pragma unroll
        for (int u = 0; u < W; ++u) {
          v[a][2 * u] = *(gf4 *)(b + f0 + 32 * u);
          v[a][2 * u + 1] = *(gf4 *)(b + f0 + 32 * u + 16);
        }
      }
    };
    auto compute = [&](const f4_t (&v)[T][2 * W]) {
#pragma unroll
      for (int u = 0; u < W; ++u) {
        bf8_t hi[T], lo[T];
#pragma unroll
        for (int a = 0; a < T; ++a) split_bf16(v[a][2 * u], v[a][2 * u + 1], hi[a], lo[a]);
        step3(hi, lo);
      }
    };
    const int64_t fs = r0 + kph * 32 * W;  // this wave's first slab
    if constexpr (!LARGE) {
      // Two register buffers in turn, each slab requested one slab ahead
      // (as in the FP32 path below).
      const int64_t last = fs < r1 ? fs + (r1 - 1 - fs) / FSTEP * FSTEP : fs;
      auto clampf = [&](int64_t f) { return f <= last ? f : last; };
      f4_t b0[T][2 * W], b1[T][2 * W];
      if (fs < r1) {
        load(fs, true, b0);
        for (int64_t f0 = fs; f0 < r1; f0 += 2 * FSTEP) {
          load(clampf(f0 + FSTEP), f0 + FSTEP < r1, b1);
          __builtin_amdgcn_sched_barrier(0);
          compute(b0);
          __builtin_amdgcn_sched_barrier(0);
          load(clampf(f0 + 2 * FSTEP), f0 + 2 * FSTEP < r1, b0);
          __builtin_amdgcn_sched_barrier(0);
          compute(b1);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    } else {
      f4_t cur[T][2 * W];
      for (int64_t f0 = fs; f0 < r1; f0 += FSTEP) {
        load(f0, true, cur);
        compute(cur);
      }
    }
  } else {
    // Lane l loads, for every tile a, the float4 of rows f0 + 4 (l >> 4) .. +3
    // of column p = 16 a + (l & 15) of A (a row of M^T); MFMA i of the group
    // takes component i.  Slot (l >> 4) of MFMA i then holds row
    // f0 + 4 (l >> 4) + i in both operands, so the four MFMAs sum rows
    // f0 .. f0 + 15 exactly once.  Straight from global memory into registers,
    // with the next group's loads in flight: no LDS staging, no barriers.
    // M^T comes from a scene record in memory, so the compiler cannot tell
    // it is global memory: the address-space cast of the base makes these
    // global loads (counted by vmcnt only) instead of flat ones (vmcnt and
    // lgkmcnt, waited for conservatively).
    typedef const __attribute__((address_space(1))) float gfloat;
    typedef float fv4 __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(1))) fv4 gfv4;
    gfloat *col[T];
    bool ok[T];
#pragma unroll
    for (int a = 0; a < T; ++a) {
      const int p = a * 16 + (lane & 15);
      ok[a] = p < ns;
      col[a] = (gfloat *)Mt + int64_t(ok[a] ? src[p] : src[0]) * ldmt + 4 * (lane >> 4);
    }
    // W float4 per tile and step: each wave reads 16 W consecutive rows (64 W
    // bytes) of every column -- whole 128-byte lines from W = 2 on; the small
    // tiles afford W = 4 registers.
    constexpr int W = LARGE ? 1 : (T <= 4 ? 4 : 2);
    // A wave reads slabs [f0, f0 + 16 W) with f0 = kph 16 W + j FSTEP.  16 W
    // divides 128 and ldmt is a multiple of 128, so a slab that starts below
    // ldmt ends within the row; FSTEP (up to 256) need not divide ldmt, so no
    // slab at or past ldmt is ever read: the first load is issued only when
    // the wave's first slab lies below ldmt, and later ones are clamped to
    // its last such slab.
    constexpr int FSTEP = 16 * W * KS;
    // Every lane loads (padding columns read a real row, src[0]'s, and are
    // zeroed when used): a select on a loaded value, or loads under a
    // lane-dependent branch, make the compiler wait for the loads at once,
    // prefetch included.
    auto load = [&](int64_t f0, fv4 (&v)[T][W]) {
#pragma unroll
      for (int a = 0; a < T; ++a)
#pragma unroll
        for (int u = 0; u < W; ++u) v[a][u] = *(gfv4 *)(col[a] + f0 + 16 * u);
    };
    // use[a]: column a's values enter the product (false for padding
    // columns and for a slab re-read past the end of the row)
    auto compute = [&](const fv4 (&v)[T][W], bool live) {
      bool use[T];
#pragma unroll
      for (int a = 0; a < T; ++a) use[a] = ok[a] && live;
#pragma unroll
      for (int u = 0; u < W; ++u) {
        TC fr[T];
#pragma unroll
        for (int a = 0; a < T; ++a) fr[a] = use[a] ? TC(v[a][u].x) : TC(0);
        step(fr);
#pragma unroll
        for (int a = 0; a < T; ++a) fr[a] = use[a] ? TC(v[a][u].y) : TC(0);
        step(fr);
#pragma unroll
        for (int a = 0; a < T; ++a) fr[a] = use[a] ? TC(v[a][u].z) : TC(0);
        step(fr);
#pragma unroll
        for (int a = 0; a < T; ++a) fr[a] = use[a] ? TC(v[a][u].w) : TC(0);
        step(fr);
      }
    };
    const int64_t fs = r0 + kph * 16 * W;  // this wave's first slab
    if constexpr (!LARGE) {
      // Two register buffers in turn, each slab requested one slab ahead.
      // Slabs past the range re-read the wave's last slab (in bounds) and are
      // multiplied in as zeros, so the loop body has no branch: a buffer
      // used under a branch has its loads sunk into that branch and waited
      // for there.
      const int64_t last = fs < r1 ? fs + (r1 - 1 - fs) / FSTEP * FSTEP : fs;
      auto clampf = [&](int64_t f) { return f <= last ? f : last; };
      fv4 b0[T][W], b1[T][W];
      if (fs < r1) {
        load(fs, b0);
        for (int64_t f0 = fs; f0 < r1; f0 += 2 * FSTEP) {
          // scheduling fences: the machine scheduler would otherwise pull
          // the register-only MFMAs above the prefetch and wait for it
          load(clampf(f0 + FSTEP), b1);
          __builtin_amdgcn_sched_barrier(0);
          compute(b0, true);
          __builtin_amdgcn_sched_barrier(0);
          load(clampf(f0 + 2 * FSTEP), b0);
          __builtin_amdgcn_sched_barrier(0);
          compute(b1, f0 + FSTEP < r1);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    } else {  // the large tiles need the registers: no prefetch
      fv4 cur[T][W];
      for (int64_t f0 = fs; f0 < r1; f0 += FSTEP) {
        load(f0, cur);
        compute(cur, true);
      }
    }
  }
  // Accumulators -> TH (reversed order, both triangles); k-split phases add in turn.
#pragma unroll
  for (int ph = 0; ph < KS; ++ph) {
    __syncthreads();
    if (kph == ph) {
#pragma unroll
      for (int u = 0; u < PPW; ++u) {
        if (pa[u] < 0) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int p = pa[u] * 16 + Mfma<TC>::row(lane, v);
          const int q = pb[u] * 16 + (lane & 15);
          if (p < ns && q < ns) {
            const int rp = ns - 1 - p, rq = ns - 1 - q;
            const TC g = acc[u][v];
            if (ph == 0) {
              TH[rp * RS + rq] = g;
              if (pa[u] != pb[u]) TH[rq * RS + rp] = g;
            } else {
              TH[rp * RS + rq] += g;
              if (pa[u] != pb[u]) TH[rq * RS + rp] += g;
            }
          }
        }
      }
    }
  }
  __syncthreads();
  };
  if constexpr (GIN) {
    // TH (reversed order) = the sum of the band partials, in part order
    const float *__restrict__ g = FS.gram + FS.goff[s];
    const int32_t parts = FS.parts;
    const int64_t gpart = FS.gram_part;
    for (int idx = tid; idx < ns * ns; idx += kThreads) {
      const int p = idx / ns, q = idx - p * ns;
      TC sum = TC(0);
      for (int r = 0; r < parts; ++r) sum += TC(g[r * gpart + p * NS + q]);
      TH[(ns - 1 - p) * RS + (ns - 1 - q)] = sum;
    }
    __syncthreads();
  } else {
    static_for<NPASS>(gram_pass);
  }

  // Row splits: every block publishes its partial Gram (LARGE: its TH in
  // global scratch already); the last of the S blocks to arrive sums all S
  // in split order (deterministic) into its TH and carries on alone.  It
  // re-arms the counter for the next launch.
  if (S > 1) {
    TC *part = static_cast<TC *>(sp.partial);
    if constexpr (!LARGE) {
      for (int idx = tid; idx < ns * ns; idx += kThreads) {
        const int r = idx / ns, c = idx % ns;
        part[size_t(unit) * NS * NS + r * NS + c] = TH[r * RS + c];
      }
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) {
      const int32_t prev = atomicAdd(sp.counters + bi, 1);
      last_split = (prev == S - 1);
      if (last_split) atomicExch(sp.counters + bi, 0);
    }
    __syncthreads();
    if (!last_split) return;
    __threadfence();
    for (int idx = tid; idx < ns * ns; idx += kThreads) {
      const int r = idx / ns, c = idx % ns;
      TC sum = TC(0);
      for (int q = 0; q < S; ++q) {
        const size_t blk = size_t(bi) * S + q;
        sum += LARGE ? th_scratch[blk * NS * RS + r * RS + c] : part[blk * NS * NS + r * NS + c];
      }
      TH[r * RS + c] = sum;
    }
    __syncthreads();
  }

  // ---- 2. Upper Cholesky TH = U^T U ----------------------------------------
  // FP32: a pivot below the FP32 Gram's rounding noise relative to its
  // diagonal (8 sqrt(m) eps, kept within [2^-20, 2^-10]) has no correct digits
  // left and is marked like a non-positive one, for the FP64 retry.  At rho = 3
  // on surface scenes every pivot keeps over a fifth of its diagonal.
  TC *d0 = reinterpret_cast<TC *>(lds + LY::o_d0);
  TC floor_rel = TC(0);
  if constexpr (sizeof(TC) == 4) {
    for (int p = tid; p < ns; p += kThreads) d0[p] = TH[p * RS + p];
    floor_rel = TC(8.0 * sqrt(double(m)) * 0x1p-24);
    floor_rel = floor_rel < TC(0x1p-20) ? TC(0x1p-20) : (floor_rel > TC(0x1p-10) ? TC(0x1p-10) : floor_rel);
    __syncthreads();
  }
  for (int j = 0; j < ns; ++j) {
    if (j > 0) {  // finish row j-1 (disjoint from the update of step j)
      const int jr = j - 1;
      for (int t = jr + tid; t < ns; t += kThreads)
        TH[jr * RS + t] = (t == jr) ? dg[jr] : TH[jr * RS + t] * dinv[jr];
    }
    const TC d = TH[j * RS + j];
    bool bad = !(d > TC(0));
    if constexpr (sizeof(TC) == 4) bad = bad || !(d > floor_rel * d0[j]);
    if (bad) {
      if (tid == 0) status[s] = 1;
      return;  // uniform: every thread read the same pivot
    }
    const TC invd = TC(1) / d;
    if (tid == 0) {
      const TC u = sqrt(d);
      dg[j] = u;
      dinv[j] = TC(1) / u;
    }
    for (int r = j + 1 + (tid >> 4); r < ns; r += 16) {
      const TC f = TH[j * RS + r] * invd;
      for (int t = r + (tid & 15); t < ns; t += 16) TH[r * RS + t] -= f * TH[j * RS + t];
    }
    __syncthreads();
  }
  {
    const int jr = ns - 1;
    if (tid == 0) TH[jr * RS + jr] = dg[jr];
  }
  __syncthreads();

  // ---- 3. KL columns of the members ------------------------------------------
  for (int mb0 = D.mem_beg; mb0 < D.mem_end; mb0 += MB) {
    const int nmb = (D.mem_end - mb0 < MB) ? D.mem_end - mb0 : MB;
    if (tid < MB) {
      if (tid < nmb) {
        const int32_t j = sup_ind[mb0 + tid];
        mcol[tid] = j;
        moff[tid] = sptr[j];
        mc[tid] = sptr[j + 1] - sptr[j] - 1;
      } else {
        mcol[tid] = -1;
        moff[tid] = 0;
        mc[tid] = -1;
      }
    }
    __syncthreads();
    int cmax = 0;
    for (int t = 0; t < nmb; ++t) cmax = (mc[t] > cmax) ? mc[t] : cmax;
    for (int idx = tid; idx < (cmax + 1) * MB; idx += kThreads) {
      const int r = idx / MB, t = idx % MB;
      W[idx] = (t < nmb && r == mc[t]) ? TC(1) : TC(0);
    }
    __syncthreads();
    for (int tt = cmax; tt >= 0; --tt) {
      for (int idx = tid; idx < (tt + 1) * MB; idx += kThreads) {
        const int r = idx / MB, t = idx % MB;
        if (t < nmb && tt <= mc[t]) {
          const TC w = W[tt * MB + t] * dinv[tt];
          if (r == tt)
            val[moff[t] + (mc[t] - tt)] = TV(w);
          else
            W[r * MB + t] -= TH[r * RS + tt] * w;
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) status[s] = 0;
}

template <int T, typename TC, typename TV, bool PTS, bool GB = false, bool GIN = false>
__global__ __launch_bounds__(kThreads) void k_super_factor(const FactorScene *__restrict__ scenes,
                                                           const FactorItem *__restrict__ items,
                                                           TC *__restrict__ th_scratch,
                                                           FactorSplit sp) {
  __shared__ __align__(16) unsigned char lds[FactorLds<T, TC>::bytes];
  factor_unit<T, TC, TV, PTS, GB, GIN>(scenes, items, th_scratch, sp, int32_t(blockIdx.x), lds);
}

// The points form's buckets of a batch in one launch: resident blocks take
// the units of all buckets in turn from a device counter (bucket by bucket,
// largest tiles first, as the host lists them), so the large supernodes start
// first and the small ones fill in behind them.  The LDS is sized for the
// largest tile count present.  Each unit runs exactly the per-bucket launch's
// code on the same regions of the scratch and split partials, so the factor
// is the same bit for bit.
template <typename TC, typename TV, bool GB>
__global__ __launch_bounds__(kThreads) void k_factor_queue(const FactorScene *__restrict__ scenes,
                                                           const FactorItem *__restrict__ items,
                                                           const FactorQueueBucket *__restrict__ qb,
                                                           int32_t nb, int32_t total,
                                                           int32_t *__restrict__ next) {
  extern __shared__ __align__(16) unsigned char dyn[];
  __shared__ int32_t unit_s;
  for (;;) {
    if (threadIdx.x == 0) unit_s = atomicAdd(next, 1);
    __syncthreads();
    const int32_t u = unit_s;
    __syncthreads();  // every thread has read the unit before it is overwritten
    if (u >= total) break;
    int32_t b = 0;
    while (b + 1 < nb && u >= qb[b + 1].first_unit) ++b;
    const FactorQueueBucket Q = qb[b];
    const int32_t local = u - Q.first_unit;
    const FactorItem *it = items + Q.offset;
    TC *ths = static_cast<TC *>(Q.th_scratch);
    switch (Q.tiles) {
      case 1: factor_unit<1, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      case 2: factor_unit<2, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      case 3: factor_unit<3, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      case 4: factor_unit<4, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      case 6: factor_unit<6, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      case 8: factor_unit<8, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      case 12: factor_unit<12, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
      default: factor_unit<16, TC, TV, true, GB, false>(scenes, it, ths, Q.sp, local, dyn); break;
    }
    __syncthreads();  // the unit's LDS is free for the next one
  }
}

template <typename TV>
__global__ __launch_bounds__(kThreads) void k_gather_values(int32_t nnz,
                                                            const int32_t *__restrict__ tmap,
                                                            const TV *__restrict__ val,
                                                            TV *__restrict__ tval) {
  const int32_t t = int32_t(blockIdx.x) * kThreads + threadIdx.x;
  if (t < nnz) tval[t] = val[tmap[t]];
}

struct FactorLaunch {
  const FactorScene *scenes;
  const FactorItem *items;
  int32_t count;
  bool points;
  void *th_scratch;
  FactorSplit sp;
  bool split_bf16;
  bool gram_in;
};

template <int T, typename TC, typename TV>
hipError_t launch_t(hipStream_t st, const FactorLaunch &L) {
  if (L.count <= 0) return hipSuccess;
  const unsigned grid = unsigned(int64_t(L.count) * (L.sp.splits > 1 ? L.sp.splits : 1));
  if (L.points) {
    if constexpr (sizeof(TC) == 4 && sizeof(TV) == 4) {
      if (L.split_bf16) {
        k_super_factor<T, TC, TV, true, true><<<grid, kThreads, 0, st>>>(
            L.scenes, L.items, static_cast<TC *>(L.th_scratch), L.sp);
        return hipGetLastError();
      }
    }
    k_super_factor<T, TC, TV, true><<<grid, kThreads, 0, st>>>(
        L.scenes, L.items, static_cast<TC *>(L.th_scratch), L.sp);
  } else if constexpr (sizeof(TC) == 4 && sizeof(TV) == 4) {
    if constexpr (T <= 6) {
      if (L.gram_in) {
        k_super_factor<T, TC, TV, false, false, true><<<grid, kThreads, 0, st>>>(
            L.scenes, L.items, static_cast<TC *>(L.th_scratch), L.sp);
        return hipGetLastError();
      }
    }
    if (L.split_bf16)
      k_super_factor<T, TC, TV, false, true><<<grid, kThreads, 0, st>>>(
          L.scenes, L.items, static_cast<TC *>(L.th_scratch), L.sp);
    else
      k_super_factor<T, TC, TV, false><<<grid, kThreads, 0, st>>>(
          L.scenes, L.items, static_cast<TC *>(L.th_scratch), L.sp);
  } else
    k_super_factor<T, TC, TV, false><<<grid, kThreads, 0, st>>>(
        L.scenes, L.items, static_cast<TC *>(L.th_scratch), L.sp);
  return hipGetLastError();
}

template <typename TC, typename TV>
hipError_t launch_prec(hipStream_t st, int tiles, const FactorLaunch &L) {
  switch (tiles) {
    case 1: return launch_t<1, TC, TV>(st, L);
    case 2: return launch_t<2, TC, TV>(st, L);
    case 3: return launch_t<3, TC, TV>(st, L);
    case 4: return launch_t<4, TC, TV>(st, L);
    case 6: return launch_t<6, TC, TV>(st, L);
    case 8: return launch_t<8, TC, TV>(st, L);
    case 12: return launch_t<12, TC, TV>(st, L);
    case 16: return launch_t<16, TC, TV>(st, L);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_transpose(hipStream_t st, const float *M, int64_t m, int64_t n, float *Mt,
                            int64_t ldmt, int64_t c_begin, int64_t c_end) {
  if (c_end <= c_begin) return hipSuccess;
  dim3 grid(unsigned((c_end - c_begin + 63) / 64), unsigned(ldmt / 64));
  if (n % 4 == 0)
    k_transpose<true><<<grid, kThreads, 0, st>>>(M, m, n, Mt, ldmt, c_begin);
  else
    k_transpose<false><<<grid, kThreads, 0, st>>>(M, m, n, Mt, ldmt, c_begin);
  return hipGetLastError();
}

hipError_t launch_super_factor(hipStream_t st, int tiles, int precision, bool points,
                               const FactorScene *scenes, const FactorItem *items, int32_t count,
                               void *th_scratch, const FactorSplit &sp, bool split_bf16,
                               bool gram_in) {
  if (sp.splits > 1 && ((sp.partial == nullptr && tiles <= 8) || sp.counters == nullptr))
    return hipErrorInvalidValue;
  if (split_bf16 && precision != 1) return hipErrorInvalidValue;
  if (gram_in && (precision != 1 || points || sp.splits > 1 || tiles > 6)) return hipErrorInvalidValue;
  const FactorLaunch L{scenes, items, count, points, th_scratch, sp, split_bf16, gram_in};
  if (precision == 0) return launch_prec<double, double>(st, tiles, L);
  if (precision == 1) return launch_prec<float, float>(st, tiles, L);
  return launch_prec<double, float>(st, tiles, L);
}

size_t factor_unit_lds(int tiles, int precision) {
  auto pick = [&](auto tc) -> size_t {
    typedef decltype(tc) TC;
    switch (tiles) {
      case 1: return FactorLds<1, TC>::bytes;
      case 2: return FactorLds<2, TC>::bytes;
      case 3: return FactorLds<3, TC>::bytes;
      case 4: return FactorLds<4, TC>::bytes;
      case 6: return FactorLds<6, TC>::bytes;
      case 8: return FactorLds<8, TC>::bytes;
      case 12: return FactorLds<12, TC>::bytes;
      case 16: return FactorLds<16, TC>::bytes;
      default: return 0;
    }
  };
  return precision == 1 ? pick(float{}) : pick(double{});
}

int factor_queue_blocks(int precision, bool split_bf16, size_t lds_bytes, int cus) {
  int per_cu = 0;
  const hipError_t e =
      precision != 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_factor_queue<double, double, false>,
                                                                    kThreads, lds_bytes)
      : split_bf16   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_factor_queue<float, float, true>,
                                                                    kThreads, lds_bytes)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_factor_queue<float, float, false>,
                                                                    kThreads, lds_bytes);
  if (e != hipSuccess || per_cu < 1) return 0;
  return per_cu * cus;
}

hipError_t launch_factor_queue(hipStream_t st, int precision, const FactorScene *scenes,
                               const FactorItem *items, const FactorQueueBucket *qb, int32_t nb,
                               int32_t total, int32_t *next, size_t lds_bytes, int blocks, bool split_bf16) {
  if (total <= 0) return hipSuccess;
  if (blocks < 1 || (precision != 0 && precision != 1) || (split_bf16 && precision != 1))
    return hipErrorInvalidValue;
  if (precision != 1)
    k_factor_queue<double, double, false><<<unsigned(blocks), kThreads, lds_bytes, st>>>(scenes, items, qb, nb,
                                                                                         total, next);
  else if (split_bf16)
    k_factor_queue<float, float, true><<<unsigned(blocks), kThreads, lds_bytes, st>>>(scenes, items, qb, nb, total,
                                                                                      next);
  else
    k_factor_queue<float, float, false><<<unsigned(blocks), kThreads, lds_bytes, st>>>(scenes, items, qb, nb,
                                                                                       total, next);
  return hipGetLastError();
}

hipError_t launch_band_gram(hipStream_t st, const float *M, int64_t m, int64_t n,
                            const BandBlock *blocks, int32_t count, const int32_t *items,
                            const SuperDesc *sd, const int32_t *sind, const int32_t *Pinv,
                            const int64_t *goff, float *gram, int64_t gram_part,
                            int64_t part_rows) {
  if (count <= 0) return hipSuccess;
  k_band_gram<<<unsigned(count), kBandThreads, 0, st>>>(M, m, n, blocks, items, sd, sind, Pinv, goff,
                                                        gram, gram_part, part_rows);
  return hipGetLastError();
}

hipError_t launch_gather_transpose(hipStream_t st, const float *M, int64_t m, int64_t n,
                                   const int32_t *cols, int32_t ncols, float *Mt, int64_t ldmt) {
  if (ncols <= 0) return hipSuccess;
  dim3 grid(unsigned((ncols + 63) / 64), unsigned(ldmt / 64));
  k_gather_transpose<<<grid, kThreads, 0, st>>>(M, m, n, cols, ncols, Mt, ldmt);
  return hipGetLastError();
}

hipError_t launch_gather_values(hipStream_t st, int precision, int32_t nnz, const int32_t *tmap,
                                const void *val, void *tval) {
  const unsigned grid = unsigned((nnz + kThreads - 1) / kThreads);
  if (nnz <= 0) return hipSuccess;
  if (precision == 0)
    k_gather_values<double><<<grid, kThreads, 0, st>>>(nnz, tmap, static_cast<const double *>(val),
                                                        static_cast<double *>(tval));
  else
    k_gather_values<float><<<grid, kThreads, 0, st>>>(nnz, tmap, static_cast<const float *>(val),
                                                       static_cast<float *>(tval));
  return hipGetLastError();
}

}  // namespace mfs
