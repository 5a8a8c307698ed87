// FP64 vector reductions and updates: dot, infinity norm, axpy.
//
// Reference semantics (CPUBackend, core/execution/gpu_backend.cpp:124-136):
// dot and axpy run to min(len); norm_inf is max |v| where NaN entries never
// win (std::max(best, |v|) keeps best), so the maximum is order-free and
// bit-exact.  The dot product is a fixed-shape two-pass tree: deterministic
// for a given length, but summed in a different order from the sequential
// CPU loop.
#include "hn_device.hpp"
#include "hn_internal.hpp"

#include <algorithm>
#include <cmath>

namespace hn {

constexpr int kRedNT = 256;
constexpr int kRedBlocks = 1024;

__global__ __launch_bounds__(kRedNT) void k_dot_partial(const double* __restrict__ a,
                                                        const double* __restrict__ b, long long n,
                                                        double* __restrict__ part) {
  __shared__ double red[kRedNT / kWave];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)kRedNT + threadIdx.x; i < n; i += (long long)gridDim.x * kRedNT)
    s += a[i] * b[i];
  s = block_sum<double, kRedNT>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedNT) void k_normi_partial(const double* __restrict__ v, long long n,
                                                          double* __restrict__ part) {
  __shared__ double red[kRedNT / kWave];
  double m = 0.0;
  for (long long i = blockIdx.x * (long long)kRedNT + threadIdx.x; i < n; i += (long long)gridDim.x * kRedNT)
    m = nan_ignoring_max(m, fabs(v[i]));
  m = block_max<double, kRedNT>(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

template <bool IsMax>
__global__ __launch_bounds__(kRedNT) void k_finish(const double* __restrict__ part, int np,
                                                   double* __restrict__ out) {
  __shared__ double red[kRedNT / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedNT) s = IsMax ? nan_ignoring_max(s, part[i]) : s + part[i];
  s = IsMax ? block_max<double, kRedNT>(s, red) : block_sum<double, kRedNT>(s, red);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_axpy(double alpha, const double* __restrict__ x, double* __restrict__ y, long long n) {
  // Rounded like the reference's CPU loop: a product, then a sum (no FMA).
#pragma clang fp contract(off)
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    y[i] = y[i] + alpha * x[i];
}

// Partial sums of the two-pass reductions.  One buffer per process (never
// freed before exit); reductions enqueued on different streams must not run
// concurrently.
static DevBuf& partials() {
  static DevBuf* b = new DevBuf();
  return *b;
}

static int partial_blocks(int64_t n) {
  return (int)std::min<int64_t>(kRedBlocks, std::max<int64_t>(1, (n + kRedNT * 4 - 1) / (kRedNT * 4)));
}

int dot_launch(const double* a, const double* b, int64_t n, double* out, hipStream_t s) {
  const int nb = partial_blocks(n);
  HN_TRY(partials().reserve(sizeof(double) * kRedBlocks));
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kRedNT), 0, s, a, b, (long long)n, partials().as<double>());
  hipLaunchKernelGGL((k_finish<false>), dim3(1), dim3(kRedNT), 0, s, partials().as<double>(), nb, out);
  count_launch(2);
  return check_launch("dot");
}

int norm_inf_launch(const double* v, int64_t n, double* out, hipStream_t s) {
  const int nb = partial_blocks(n);
  HN_TRY(partials().reserve(sizeof(double) * kRedBlocks));
  hipLaunchKernelGGL(k_normi_partial, dim3(nb), dim3(kRedNT), 0, s, v, (long long)n, partials().as<double>());
  hipLaunchKernelGGL((k_finish<true>), dim3(1), dim3(kRedNT), 0, s, partials().as<double>(), nb, out);
  count_launch(2);
  return check_launch("norm_inf");
}

// ---- allocation-free dot product (hn_dot_dev_noalloc) -----------------------
// The same two kernels as hn_dot_dev, with the partial sums in a file-scope
// device array instead of a buffer reserved on first use, so the call never
// allocates.  Each device holds its own copy of the array until the process
// ends.  In host code the array's name is not a device address: the address
// comes from hipGetSymbolAddress, taken once per device with that device
// current (vec_noalloc_setup, called by hn_init) and kept in g_scratch.
static __device__ __align__(256) double scratch[kRedBlocks];
constexpr int kNoAllocDevices = 64;
static double* g_scratch[kNoAllocDevices] = {};
// Largest n: the grid is capped at kRedBlocks blocks, so the array serves
// any length the kernels' signed 64-bit index can reach.
constexpr uint64_t kNoAllocMaxN = uint64_t(1) << 62;

int vec_noalloc_setup(int device) {
  HN_CHECK(device >= 0 && device < kNoAllocDevices, HN_ERR_UNSUPPORTED,
           "allocation-free scratch: device ordinal above 63");
  if (g_scratch[device]) return HN_OK;
  void* p = nullptr;
  HN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[device] = static_cast<double*>(p);
  return HN_OK;
}

int axpy_launch(double alpha, const double* x, double* y, int64_t n, hipStream_t s) {
  if (n <= 0) return HN_OK;
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, alpha, x, y, (long long)n);
  count_launch();
  return check_launch("axpy");
}

}  // namespace hn

using namespace hn;

extern "C" int hn_dot_dev(const double* a, const double* b, uint64_t n, double* out, void* stream) {
  HN_CHECK(out && (n == 0 || (a && b)), HN_ERR_INVALID, "hn_dot_dev: null argument");
  HN_TRY(enter());
  return dot_launch(a, b, (int64_t)n, out, pick_stream(stream));
}

extern "C" int hn_dot_dev_noalloc(const double* a, const double* b, uint64_t n, double* out, void* stream) {
  HN_CHECK(out && (n == 0 || (a && b)), HN_ERR_INVALID, "hn_dot_dev_noalloc: null argument");
  HN_CHECK(n <= kNoAllocMaxN, HN_ERR_INVALID, "hn_dot_dev_noalloc: n above 2^62");
  HN_TRY(enter());
  double* part = g_scratch[ctx().device];
  HN_CHECK(part, HN_ERR_NOT_INIT, "hn_dot_dev_noalloc: scratch address not set up by hn_init");
  const int nb = partial_blocks((int64_t)n);
  hipStream_t s = pick_stream(stream);
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kRedNT), 0, s, a, b, (long long)n, part);
  hipLaunchKernelGGL((k_finish<false>), dim3(1), dim3(kRedNT), 0, s, (const double*)part, nb, out);
  count_launch(2);
  return check_launch("dot (allocation-free)");
}

extern "C" int hn_norm_inf_dev(const double* v, uint64_t n, double* out, void* stream) {
  HN_CHECK(out && (n == 0 || v), HN_ERR_INVALID, "hn_norm_inf_dev: null argument");
  HN_TRY(enter());
  return norm_inf_launch(v, (int64_t)n, out, pick_stream(stream));
}

extern "C" int hn_axpy_dev(double alpha, const double* x, double* y, uint64_t n, void* stream) {
  HN_CHECK(n == 0 || (x && y), HN_ERR_INVALID, "hn_axpy_dev: null argument");
  HN_TRY(enter());
  return axpy_launch(alpha, x, y, (int64_t)n, pick_stream(stream));
}

extern "C" int hn_dot(const double* a, const double* b, uint64_t n, double* out) {
  HN_CHECK(out && (n == 0 || (a && b)), HN_ERR_INVALID, "hn_dot: null argument");
  if (n == 0) {
    *out = 0.0;
    return HN_OK;
  }
  HN_TRY(enter());
  HostOpScratch& sc = host_scratch();
  std::lock_guard<std::mutex> lk(sc.mtx);
  hipStream_t s = ctx().stream;
  HN_TRY(sc.a.reserve(sizeof(double) * n));
  HN_TRY(sc.b.reserve(sizeof(double) * n));
  HN_TRY(sc.red.reserve(sizeof(double)));
  HN_HIP(hipMemcpyAsync(sc.a.p, a, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HN_HIP(hipMemcpyAsync(sc.b.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HN_TRY(dot_launch(sc.a.as<double>(), sc.b.as<double>(), (int64_t)n, sc.red.as<double>(), s));
  HN_HIP(hipMemcpyAsync(out, sc.red.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}

extern "C" int hn_norm_inf(const double* v, uint64_t n, double* out) {
  HN_CHECK(out && (n == 0 || v), HN_ERR_INVALID, "hn_norm_inf: null argument");
  if (n == 0) {
    *out = 0.0;
    return HN_OK;
  }
  HN_TRY(enter());
  HostOpScratch& sc = host_scratch();
  std::lock_guard<std::mutex> lk(sc.mtx);
  hipStream_t s = ctx().stream;
  HN_TRY(sc.a.reserve(sizeof(double) * n));
  HN_TRY(sc.red.reserve(sizeof(double)));
  HN_HIP(hipMemcpyAsync(sc.a.p, v, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HN_TRY(norm_inf_launch(sc.a.as<double>(), (int64_t)n, sc.red.as<double>(), s));
  HN_HIP(hipMemcpyAsync(out, sc.red.p, sizeof(double), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}

extern "C" int hn_axpy(double alpha, const double* x, double* y, uint64_t n) {
  HN_CHECK(n == 0 || (x && y), HN_ERR_INVALID, "hn_axpy: null argument");
  if (n == 0) return HN_OK;
  HN_TRY(enter());
  HostOpScratch& sc = host_scratch();
  std::lock_guard<std::mutex> lk(sc.mtx);
  hipStream_t s = ctx().stream;
  HN_TRY(sc.a.reserve(sizeof(double) * n));
  HN_TRY(sc.b.reserve(sizeof(double) * n));
  HN_HIP(hipMemcpyAsync(sc.a.p, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HN_HIP(hipMemcpyAsync(sc.b.p, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
  HN_TRY(axpy_launch(alpha, sc.a.as<double>(), sc.b.as<double>(), (int64_t)n, s));
  HN_HIP(hipMemcpyAsync(y, sc.b.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}
