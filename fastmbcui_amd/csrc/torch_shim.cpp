// torch_shim.cpp -- libmfs_torch.so: the reference's operator library, over the C-ABI.
//
// fastMBCUI's op package loads a shared library that defines the mfs_torch
// operators (core/mfs_torch_py/__init__.py:7-8 loads
// ./build/mfs_torch/csrc/libmfs_torch.so; schemas mfs_torch_ops.cpp:8-9,
// CUDA-key implementations mfs_torch_ops.cu:372-375) and then registers their
// fake kernels and autograd in Python (core/mfs_torch_py/ops.py).  This file
// builds a library with the same two schemas whose CUDA-key implementations
// (the key HIP tensors carry under PyTorch-ROCm) call libmfs_core.so on the
// current HIP stream, so that package loads it unchanged.  Host code only: no
// device code is compiled here.
//
// Same argument checks and messages as mfs_torch_ops.cu:24-33 and :202-215.
// The solve is the plain FP64 form (mfs_solve_dense), as the Python operators
// run by default.  The reference keeps the last factor in process-global
// tensors that are not keyed by geometry (mfs_torch_ops.cu:17-19, 136-157);
// here one context per (device, stream) keeps its plan and factor keyed by
// the source points, rho, M's identity and a fingerprint of M's contents.
#include <ATen/ATen.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>

#include "../../include/mfs_core.h"

namespace {

constexpr double kLambda = 1.5;     // supernode aggregation ratio (mfs_torch_ops.cu:115)
constexpr int32_t kMaxSuper = 2048; // supernode member cap (mfs_torch_ops.cu:88)

struct Slot {
  mfs_ctx *ctx = nullptr;
  std::string plan_key, factor_key;
};

std::mutex g_mu;
std::map<std::pair<int, void *>, Slot> g_slots;
bool g_forward_seen = false;

// 128-bit FNV-1a style hash of a byte range (two independent 64-bit lanes).
void hash_bytes(const void *p, size_t n, uint64_t &h1, uint64_t &h2) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; ++i) {
    h1 = (h1 ^ b[i]) * 0x100000001b3ULL;
    h2 = (h2 ^ b[i]) * 0x00000100000001b3ULL + 0x9e3779b97f4a7c15ULL;
  }
}

std::string hex_key(uint64_t a, uint64_t b) {
  char buf[40];
  std::snprintf(buf, sizeof(buf), "%016llx%016llx", static_cast<unsigned long long>(a),
                static_cast<unsigned long long>(b));
  return buf;
}

std::string geometry_key(const at::Tensor &Vs32, double rho) {
  uint64_t h1 = 0xcbf29ce484222325ULL, h2 = 0x84222325cbf29ce4ULL;
  hash_bytes(Vs32.data_ptr(), size_t(Vs32.numel()) * sizeof(float), h1, h2);
  hash_bytes(&rho, sizeof(rho), h1, h2);
  const int64_t n = Vs32.size(0);
  hash_bytes(&n, sizeof(n), h1, h2);
  return hex_key(h1, h2);
}

// Weighted FP64 sum of up to 65,536 evenly spaced entries of M (synchronises).
double content_fingerprint(const at::Tensor &M) {
  const int64_t count = M.numel();
  if (count == 0) return 0.0;
  const int64_t take = count < 65536 ? count : 65536;
  const auto opts = M.options().dtype(at::kDouble);
  at::Tensor idx = at::linspace(0, double(count - 1), take, opts).round().to(at::kLong);
  at::Tensor w = at::arange(1, take + 1, opts).mul(0.6180339887498949).frac().add(0.5);
  return M.reshape({-1}).index_select(0, idx).to(at::kDouble).mul(w).sum().item<double>();
}

[[noreturn]] void fail(mfs_ctx *ctx, int code, const char *what) {
  TORCH_CHECK(false, what, " failed (status ", code, "): ", mfs_ctx_last_error(ctx));
}

at::Tensor solve(const at::Tensor &M, const at::Tensor &B, const at::Tensor &Vs_s, double rho,
                 int64_t pcg_max_iter, double tol_r, bool timing_verbose, bool solve_verbose,
                 bool forward) {
  // mfs_torch_ops.cu:24-33 (MinvF) / :202-211 (VpL)
  TORCH_CHECK(B.dim() == 2, forward ? "F must be a 2D tensor" : "pL must be a 2D tensor");
  if (forward) {
    TORCH_CHECK(M.size(0) == B.size(0), "M and F must have the same first dimension");
  } else {
    TORCH_CHECK(M.size(1) == B.size(0), "M and pL must have compatible dimensions");
  }
  TORCH_CHECK(M.size(1) == Vs_s.size(0), "M and Vs_s must have compatible dimensions");
  TORCH_CHECK(Vs_s.size(1) == 3, "Vs_s must be 3d points");
  TORCH_CHECK(M.dtype() == at::kFloat, "M must be float32");
  TORCH_CHECK(B.dtype() == at::kFloat, forward ? "F must be float32" : "pL must be float32");
  TORCH_CHECK(Vs_s.dtype() == at::kFloat, "Vs_s must be float32");
  TORCH_INTERNAL_ASSERT(M.device().is_cuda());
  TORCH_INTERNAL_ASSERT(B.device().is_cuda());
  TORCH_INTERNAL_ASSERT(Vs_s.device().is_cpu());
  if (!forward) {  // mfs_torch_ops.cu:213-215
    TORCH_CHECK(g_forward_seen, "Float cache tensors are not defined. Please run MinvF_float first.");
  }

  const auto t0 = std::chrono::steady_clock::now();
  auto since = [](std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
  };
  const at::Tensor Mc = M.contiguous(), Bc = B.contiguous(), Vs = Vs_s.contiguous();
  const double prep_ms = since(t0);
  const int64_t m = Mc.size(0), n = Mc.size(1), k = Bc.size(1);
  const int device = Mc.get_device();
  void *stream = c10::hip::getCurrentHIPStream(device).stream();

  std::lock_guard<std::mutex> lock(g_mu);
  Slot &s = g_slots[{device, stream}];
  if (s.ctx == nullptr) {
    const int rc = mfs_ctx_create(&s.ctx, device, stream);
    TORCH_CHECK(rc == MFS_OK, "mfs_ctx_create failed (status ", rc, ")");
  }
  mfs_ctx_set_stream(s.ctx, stream);
  const std::string gk = geometry_key(Vs, rho);
  double plan_ms = 0.0, order_ms = 0.0, pattern_ms = 0.0, factor_ms = 0.0;
  if (s.plan_key != gk) {
    s.plan_key.clear();
    s.factor_key.clear();
    const int rc = mfs_plan(s.ctx, Vs.data_ptr<float>(), n, rho, kLambda, kMaxSuper);
    if (rc != MFS_OK) fail(s.ctx, rc, "mfs_plan");
    s.plan_key = gk;
    mfs_stats st{};
    mfs_ctx_stats(s.ctx, &st);
    plan_ms = st.plan_ms;
    order_ms = st.plan_order_ms;
    pattern_ms = st.plan_pattern_ms;
  }
  char fk[160];
  std::snprintf(fk, sizeof(fk), "%p|%lld|%lld|%lld|%u|%.17g", Mc.data_ptr(), static_cast<long long>(m),
                static_cast<long long>(n), static_cast<long long>(Mc.stride(0)),
                static_cast<unsigned>(Mc._version()), content_fingerprint(Mc));
  const std::string fkey = gk + fk;
  if (s.factor_key != fkey) {
    s.factor_key.clear();
    const int rc = mfs_factor_dense(s.ctx, Mc.data_ptr<float>(), m, n, MFS_FP64);
    if (rc != MFS_OK) fail(s.ctx, rc, "mfs_factor_dense");
    s.factor_key = fkey;
    mfs_stats st{};
    mfs_ctx_stats(s.ctx, &st);
    factor_ms = st.factor_ms;
  }
  at::Tensor out = at::empty({n, k}, Mc.options());
  if (k > 0) {
    const int rc = mfs_solve_dense(s.ctx, Mc.data_ptr<float>(), m, n, Bc.data_ptr<float>(), k,
                                   forward ? MFS_RHS_BOUNDARY : MFS_RHS_SOURCE, out.data_ptr<float>(),
                                   int32_t(pcg_max_iter), tol_r);
    if (rc != MFS_OK) fail(s.ctx, rc, "mfs_solve_dense");
  }
  mfs_stats st{};
  mfs_ctx_stats(s.ctx, &st);
  if (timing_verbose) {
    // the reference's phases (mfs_torch_ops.cu:48-193); permutation, the
    // covariance copy, the cache copy and the back-permutation are fused into
    // the kernels here, and a reused plan or factor costs nothing
    (void)c10::hip::getCurrentHIPStream(device).synchronize();
    const double setup_ms = plan_ms - order_ms - pattern_ms > 0 ? plan_ms - order_ms - pattern_ms : 0.0;
    const std::pair<const char *, double> phases[] = {
        {"prep", prep_ms},     {"fps_sampler", order_ms}, {"permutation", 0.0},
        {"build_cov_mat", 0.0}, {"pattern", pattern_ms},   {"setup", setup_ms},
        {"compute", factor_ms}, {"cache", 0.0},            {"pcg", st.solve_ms},
        {"perm_back", 0.0},     {"total", since(t0)}};
    for (const auto &p : phases) std::printf("TIME(%s): %g ms\n", p.first, p.second);
  }
  if (solve_verbose) {
    for (int64_t c = 0; c < k && c < 64; ++c)
      std::printf("pcg column %lld: iterations %d, relative residual %.3e\n", static_cast<long long>(c),
                  st.iters[c], st.relres[c]);
  }
  if (forward) g_forward_seen = true;
  return out;
}

at::Tensor minvf_cuda(const at::Tensor &M, const at::Tensor &F, const at::Tensor &Vs_s, double rho,
                      int64_t pcg_max_iter, double tol_r, bool timing_verbose, bool solve_verbose) {
  return solve(M, F, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose, true);
}

at::Tensor vpl_cuda(const at::Tensor &M, const at::Tensor &pL, const at::Tensor &Vs_s, double rho,
                    int64_t pcg_max_iter, double tol_r, bool timing_verbose, bool solve_verbose) {
  return solve(M, pL, Vs_s, rho, pcg_max_iter, tol_r, timing_verbose, solve_verbose, false);
}

}  // namespace

TORCH_LIBRARY(mfs_torch, m) {
  m.def("VpL_fast_float(Tensor M, Tensor pL, Tensor Vs_s, float rho, int pcg_max_iter, float tol_r, "
        "bool timing_verbose=False, bool solve_verbose=False) -> Tensor");
  m.def("MinvF_float_fastGrad(Tensor M, Tensor F, Tensor Vs_s, float rho, int pcg_max_iter, "
        "float tol_r, bool timing_verbose=False, bool solve_verbose=False) -> Tensor");
}

TORCH_LIBRARY_IMPL(mfs_torch, CUDA, m) {
  m.impl("MinvF_float_fastGrad", &minvf_cuda);
  m.impl("VpL_fast_float", &vpl_cuda);
}
