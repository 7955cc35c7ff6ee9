// kernel_eval.h -- the MFS point kernel, evaluated on the device.
//
//   K(x, n_x; y) = -(x - y) . n_x / (4 pi |x - y|^3)
//
// This is the configs' normal derivative of 1/(4 pi r).  The reference's
// laplace_nm_3d_sl (src/kernels.h:24-40) evaluates (x - y) . n / |x - y|^3,
// i.e. -4 pi times this; a constant factor scales M, its Gram and the
// factor, and PCG's iterates and solution are invariant to it.
#pragma once

#include <hip/hip_runtime.h>

namespace mfs {

constexpr double kInv4Pi = 0.07957747154594767;  // 1 / (4 pi)

// Packed point records (original source order, as the columns of M).
//   targets FP32: x_hi[3] x_lo[3] n[3] pad[3]   (12 floats, 48 B)
//   sources FP32: y_hi[3] y_lo[3] pad[2]        (8 floats, 32 B)
//   targets FP64: x[3] n[3] pad[2]              (8 doubles)
//   sources FP64: y[3] pad                      (4 doubles)
// x_hi = fl32(x), x_lo = fl32(x - x_hi): the FP32 difference
// (x_hi - y_hi) + (x_lo - y_lo) keeps nearby pairs accurate to FP32 of the
// difference itself, not of the coordinates.
constexpr int kTgt32 = 12, kSrc32 = 8, kTgt64 = 8, kSrc64 = 4;

// (d . n) / |d|^3 in FP32 with the hardware reciprocal square root.
__device__ inline float dlp32(float dx, float dy, float dz, float n0, float n1, float n2) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  const float ri = __builtin_amdgcn_rsqf(r2);
  return (dx * n0 + dy * n1 + dz * n2) * (ri * ri * ri);
}

// (d . n) / |d|^3 in FP64 without a division or a square root: the hardware
// reciprocal square root (relative error about 2^-23) refined by one
// third-order Householder step for 1/sqrt(a),
//   e = 1 - a y^2,   y <- y (1 + e/2 + 3 e^2 / 8),
// whose error is O(e^3) (below FP64 rounding for |e| <= 2^-20).
__device__ inline double dlp64(double dx, double dy, double dz, double n0, double n1, double n2) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  double ri = __builtin_amdgcn_rsq(r2);
  const double e = fma(-r2 * ri, ri, 1.0);
  ri = fma(ri, e * fma(0.375, e, 0.5), ri);
  return (dx * n0 + dy * n1 + dz * n2) * (ri * ri * ri);
}

// Full FP64 value, the same expression the dense assembly rounds to FP32.
__device__ inline double kernel64(double dx, double dy, double dz, double n0, double n1,
                                  double n2) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double dn = dx * n0 + dy * n1 + dz * n2;
  const double r = sqrt(r2);
  return -dn / (4.0 * 3.14159265358979323846 * r2 * r);
}

}  // namespace mfs
