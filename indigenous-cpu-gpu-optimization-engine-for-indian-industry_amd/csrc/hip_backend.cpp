// HipBackend: IComputeBackend over the dlopen'ed libhn_kernels.so.
#include "hip_backend.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace hn::backend {

// Entry points bound from libhn_kernels.so (declared in include/hn_kernels.h).
#define HN_BACKEND_FUNCS(X)                                                                          \
  X(int, hn_init, (int, void*))                                                                      \
  X(int, hn_finalize, (void))                                                                        \
  X(int, hn_last_error, (char*, std::size_t))                                                        \
  X(int, hn_mem_info, (std::size_t*, std::size_t*))                                                  \
  X(uint64_t, hn_kernel_executions, (void))                                                          \
  X(void, hn_reset_execution_stats, (void))                                                          \
  X(int, hn_csr_upload, (const double*, const uint64_t*, const uint64_t*, uint64_t, uint64_t, uint64_t, int, void**)) \
  X(int, hn_csr_free, (void*))                                                                       \
  X(int, hn_spmv, (const void*, const double*, double*))                                             \
  X(int, hn_spmm, (const void*, const double*, double*, uint64_t))                                   \
  X(int, hn_dot, (const double*, const double*, uint64_t, double*))                                  \
  X(int, hn_axpy, (double, const double*, double*, uint64_t))                                        \
  X(int, hn_norm_inf, (const double*, uint64_t, double*))                                            \
  X(int, hn_inject_failures, (uint32_t, int))                                                        \
  X(int, hn_dev_alloc, (std::size_t, void**))                                                        \
  X(int, hn_dev_free, (void*))                                                                       \
  X(int, hn_memcpy_h2d, (void*, const void*, std::size_t))                                           \
  X(int, hn_memcpy_d2h, (void*, const void*, std::size_t))                                           \
  X(int, hn_spmv_dev, (const void*, int, const void*, void*, void*))                                 \
  X(int, hn_stream_sync, (void*))

struct KernelApi {
#define HN_DECL(ret, name, args) ret(*name) args = nullptr;
  HN_BACKEND_FUNCS(HN_DECL)
#undef HN_DECL
};

namespace {

// hn_device_info from include/hn_kernels.h (layout-compatible copy).
struct DeviceInfo {
  char name[256];
  char arch[64];
  uint64_t total_memory;
  int compute_units;
  int device;
};

std::string default_library() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_library), &info) && info.dli_fname) {
    std::string self = info.dli_fname;
    const auto slash = self.find_last_of('/');
    if (slash != std::string::npos) return self.substr(0, slash + 1) + "libhn_kernels.so";
  }
  return "libhn_kernels.so";
}

uint64_t fingerprint(const std::vector<double>& v, const std::vector<std::size_t>& c,
                     const std::vector<std::size_t>& r) {
  // FNV-1a over every byte of the three arrays.
  uint64_t h = 1469598103934665603ULL;
  auto eat = [&](const void* p, std::size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (std::size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ULL;
    }
  };
  eat(v.data(), v.size() * sizeof(double));
  eat(c.data(), c.size() * sizeof(std::size_t));
  eat(r.data(), r.size() * sizeof(std::size_t));
  return h;
}

}  // namespace

struct HipBackend::Resident {
  const void* keys[3];
  std::size_t nnz = 0, nrows = 0, ncols = 0;
  uint64_t print = 0;
  void* handle = nullptr;
};

HipBackend::HipBackend(std::string library) : library_(std::move(library)), api_(new KernelApi()) {}

HipBackend::~HipBackend() {
  std::lock_guard<std::mutex> lk(mutex_);
  release_all();
  if (handle_) dlclose(handle_);
}

std::string HipBackend::name() const {
  return device_name_.empty() ? std::string("HIP") : "HIP(" + device_name_ + ")";
}

void HipBackend::initialize() {
  std::lock_guard<std::mutex> lk(mutex_);
  if (initialized_) return;
  initialized_ = true;
  const std::string path = library_.empty() ? default_library() : library_;
  handle_ = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
  if (!handle_) {
    const char* e = dlerror();
    last_error_ = std::string("cannot load ") + path + ": " + (e ? e : "?");
    return;
  }
#define HN_BIND(ret, name, args)                                                    \
  api_->name = reinterpret_cast<ret(*) args>(dlsym(handle_, #name));               \
  if (!api_->name) {                                                                \
    last_error_ = "missing symbol " #name " in " + path;                            \
    return;                                                                         \
  }
  HN_BACKEND_FUNCS(HN_BIND)
#undef HN_BIND
  DeviceInfo info{};
  if (api_->hn_init(-1, &info) != 0) {
    char buf[512];
    api_->hn_last_error(buf, sizeof buf);
    last_error_ = buf;
    return;
  }
  device_name_ = info.name;
  available_ = true;
}

void HipBackend::finalize() {
  std::lock_guard<std::mutex> lk(mutex_);
  release_all();
}

void HipBackend::release_all() {
  if (available_)
    for (auto& r : cache_) api_->hn_csr_free(r->handle);
  cache_.clear();
}

void HipBackend::require(const char* op) const {
  if (!available_) throw std::runtime_error(std::string("HIP backend not available for ") + op + ": " + last_error_);
}

void HipBackend::raise(const char* op) const {
  char buf[1024];
  api_->hn_last_error(buf, sizeof buf);
  throw std::runtime_error(std::string("HIP ") + op + " failed: " + buf);
}

HipBackend::Resident& HipBackend::resident(const std::vector<double>& values,
                                           const std::vector<std::size_t>& col_indices,
                                           const std::vector<std::size_t>& row_ptr, std::size_t ncols) {
  const std::size_t nrows = row_ptr.size() - 1;
  const uint64_t print = fingerprint(values, col_indices, row_ptr);
  for (auto& r : cache_) {
    if (r->keys[0] == &values && r->keys[1] == &col_indices && r->keys[2] == &row_ptr) {
      if (r->nnz == values.size() && r->nrows == nrows && r->ncols == ncols && r->print == print) return *r;
      api_->hn_csr_free(r->handle);
      r->handle = nullptr;
      if (api_->hn_csr_upload(values.data(), reinterpret_cast<const uint64_t*>(col_indices.data()),
                              reinterpret_cast<const uint64_t*>(row_ptr.data()), nrows, ncols, values.size(), 1,
                              &r->handle) != 0)
        raise("matrix upload");
      r->nnz = values.size();
      r->nrows = nrows;
      r->ncols = ncols;
      r->print = print;
      return *r;
    }
  }
  auto r = std::make_unique<Resident>();
  r->keys[0] = &values;
  r->keys[1] = &col_indices;
  r->keys[2] = &row_ptr;
  static_assert(sizeof(std::size_t) == sizeof(uint64_t), "64-bit size_t expected");
  if (api_->hn_csr_upload(values.data(), reinterpret_cast<const uint64_t*>(col_indices.data()),
                          reinterpret_cast<const uint64_t*>(row_ptr.data()), nrows, ncols, values.size(), 1,
                          &r->handle) != 0)
    raise("matrix upload");
  r->nnz = values.size();
  r->nrows = nrows;
  r->ncols = ncols;
  r->print = print;
  cache_.push_back(std::move(r));
  return *cache_.back();
}

void HipBackend::spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                      const std::vector<std::size_t>& row_ptr, const std::vector<double>& x,
                      std::vector<double>& y) {
  std::lock_guard<std::mutex> lk(mutex_);
  require("spmv");
  // gpu_backend.cpp:207-210: empty matrix -> zero vector of nrows
  if (values.empty() || row_ptr.size() < 2) {
    y.assign(row_ptr.size() >= 1 ? row_ptr.size() - 1 : 0, 0.0);
    return;
  }
  Resident& r = resident(values, col_indices, row_ptr, x.size());
  y.resize(r.nrows);
  if (api_->hn_spmv(r.handle, x.data(), y.data()) != 0) raise("spmv");
}

void HipBackend::spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                      const std::vector<std::size_t>& row_ptr, const std::vector<double>& B,
                      std::vector<double>& C, std::size_t ncols_B) {
  std::lock_guard<std::mutex> lk(mutex_);
  require("spmm");
  const std::size_t nrows = row_ptr.size() >= 1 ? row_ptr.size() - 1 : 0;
  if (values.empty() || row_ptr.size() < 2 || ncols_B == 0) {
    C.assign(nrows * ncols_B, 0.0);
    return;
  }
  Resident& r = resident(values, col_indices, row_ptr, B.size() / ncols_B);
  C.resize(nrows * ncols_B);
  if (api_->hn_spmm(r.handle, B.data(), C.data(), ncols_B) != 0) raise("spmm");
}

double HipBackend::dot(const std::vector<double>& a, const std::vector<double>& b) {
  std::lock_guard<std::mutex> lk(mutex_);
  require("dot");
  double out = 0.0;
  if (api_->hn_dot(a.data(), b.data(), std::min(a.size(), b.size()), &out) != 0) raise("dot");
  return out;
}

void HipBackend::axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) {
  std::lock_guard<std::mutex> lk(mutex_);
  require("axpy");
  if (api_->hn_axpy(alpha, x.data(), y.data(), std::min(x.size(), y.size())) != 0) raise("axpy");
}

double HipBackend::norm_inf(const std::vector<double>& v) {
  std::lock_guard<std::mutex> lk(mutex_);
  require("norm_inf");
  double out = 0.0;
  if (api_->hn_norm_inf(v.data(), v.size(), &out) != 0) raise("norm_inf");
  return out;
}

std::size_t HipBackend::device_memory() const {
  if (!available_) return 0;
  std::size_t f = 0, t = 0;
  return api_->hn_mem_info(&f, &t) == 0 ? t : 0;
}

std::size_t HipBackend::free_memory() const {
  if (!available_) return 0;
  std::size_t f = 0, t = 0;
  return api_->hn_mem_info(&f, &t) == 0 ? f : 0;
}

bool HipBackend::was_gpu_kernel_executed() const { return available_ && api_->hn_kernel_executions() > 0; }
std::size_t HipBackend::gpu_kernel_executions() const { return available_ ? api_->hn_kernel_executions() : 0; }
void HipBackend::reset_execution_stats() {
  if (available_) api_->hn_reset_execution_stats();
}

void HipBackend::inject_failures(unsigned count) {
  if (available_) api_->hn_inject_failures(count, 2 /* HN_ERR_HIP */);
}

std::unique_ptr<IComputeBackend> create_hip_backend(std::string library) {
  auto b = std::make_unique<HipBackend>(std::move(library));
  b->initialize();
  if (!b->is_available()) return nullptr;
  return b;
}

std::string hip_backend_probe_error(std::string library) {
  HipBackend b(std::move(library));
  b.initialize();
  return b.is_available() ? std::string() : b.last_error();
}

// ---- calibration ---------------------------------------------------------------
void HipBackend::calibrate() {
  using clock = std::chrono::steady_clock;
  auto ms_since = [](clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  };
  measured_spmv_gbps_ = measured_h2d_gbps_ = measured_d2h_gbps_ = 0.0;
  if (available_) {
    // the reference's synthetic problem: 4096 rows x 32 nonzeros, column
    // (i + j) mod 4096, values 1, x = 1
    const std::size_t nrows = 4096, per_row = 32, nnz = nrows * per_row;
    std::vector<std::size_t> rp(nrows + 1), ci(nnz);
    std::vector<double> val(nnz, 1.0), x(nrows, 1.0);
    for (std::size_t i = 0; i < nrows; ++i) {
      rp[i] = i * per_row;
      for (std::size_t j = 0; j < per_row; ++j) ci[i * per_row + j] = (i + j) % nrows;
    }
    rp[nrows] = nnz;
    // device-only rate (the model charges the vector transfers separately):
    // matrix and vectors resident, the product timed with the stream drained
    void* a = nullptr;
    void *dx = nullptr, *dy = nullptr;
    if (api_->hn_csr_upload(val.data(), reinterpret_cast<const uint64_t*>(ci.data()),
                            reinterpret_cast<const uint64_t*>(rp.data()), nrows, nrows, nnz, 1, &a) == 0 &&
        api_->hn_dev_alloc(nrows * sizeof(double), &dx) == 0 && api_->hn_dev_alloc(nrows * sizeof(double), &dy) == 0 &&
        api_->hn_memcpy_h2d(dx, x.data(), nrows * sizeof(double)) == 0 &&
        api_->hn_spmv_dev(a, 1, dx, dy, nullptr) == 0 && api_->hn_stream_sync(nullptr) == 0) {
      constexpr int kReps = 5;
      const auto t0 = clock::now();
      bool ok = true;
      for (int r = 0; r < kReps && ok; ++r) ok = api_->hn_spmv_dev(a, 1, dx, dy, nullptr) == 0;
      ok = ok && api_->hn_stream_sync(nullptr) == 0;
      const double ms = ms_since(t0);
      if (ok && ms > 0.0)
        measured_spmv_gbps_ = (double)nnz * (sizeof(double) + sizeof(std::size_t)) * kReps / (ms / 1e3) / 1e9;
    } else {
      char buf[512];
      api_->hn_last_error(buf, sizeof buf);
      last_error_ = std::string("calibration SpMV failed: ") + buf;
    }
    if (dx) api_->hn_dev_free(dx);
    if (dy) api_->hn_dev_free(dy);
    if (a) api_->hn_csr_free(a);
    void* d = nullptr;
    const std::size_t bytes = nnz * sizeof(double);
    if (api_->hn_dev_alloc(bytes, &d) == 0) {
      auto t0 = clock::now();
      if (api_->hn_memcpy_h2d(d, val.data(), bytes) == 0) {
        const double ms = ms_since(t0);
        if (ms > 0.0) measured_h2d_gbps_ = (double)bytes / (ms / 1e3) / 1e9;
      }
      std::vector<double> back(nnz);
      t0 = clock::now();
      if (api_->hn_memcpy_d2h(back.data(), d, bytes) == 0) {
        const double ms = ms_since(t0);
        if (ms > 0.0) measured_d2h_gbps_ = (double)bytes / (ms / 1e3) / 1e9;
      }
      api_->hn_dev_free(d);
    }
  }
  if (measured_h2d_gbps_ <= 0.0) measured_h2d_gbps_ = 6.0;
  if (measured_d2h_gbps_ <= 0.0) measured_d2h_gbps_ = 6.0;
  if (measured_spmv_gbps_ <= 0.0) measured_spmv_gbps_ = 50.0;
}

// ---- CPU backend ------------------------------------------------------------------
void CpuBackend::spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                      const std::vector<std::size_t>& row_ptr, const std::vector<double>& x, std::vector<double>& y) {
  const std::size_t nrows = row_ptr.size() >= 1 ? row_ptr.size() - 1 : 0;
  y.assign(nrows, 0.0);
  for (std::size_t i = 0; i < nrows; ++i) {
    double sum = 0.0;
    for (std::size_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) sum += values[k] * x[col_indices[k]];
    y[i] = sum;
  }
}

void CpuBackend::spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                      const std::vector<std::size_t>& row_ptr, const std::vector<double>& B, std::vector<double>& C,
                      std::size_t ncols_B) {
  const std::size_t nrows = row_ptr.size() >= 1 ? row_ptr.size() - 1 : 0;
  C.assign(nrows * ncols_B, 0.0);
  for (std::size_t i = 0; i < nrows; ++i)
    for (std::size_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
      const std::size_t j = col_indices[k];
      const double a = values[k];
      for (std::size_t c = 0; c < ncols_B; ++c) C[i * ncols_B + c] += a * B[j * ncols_B + c];
    }
}

double CpuBackend::dot(const std::vector<double>& a, const std::vector<double>& b) {
  double sum = 0.0;
  for (std::size_t i = 0; i < a.size() && i < b.size(); ++i) sum += a[i] * b[i];
  return sum;
}

void CpuBackend::axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) {
  for (std::size_t i = 0; i < x.size() && i < y.size(); ++i) y[i] += alpha * x[i];
}

double CpuBackend::norm_inf(const std::vector<double>& v) {
  double mx = 0.0;
  for (const double e : v) mx = std::max(mx, std::abs(e));
  return mx;
}

// ---- cost model ---------------------------------------------------------------------
namespace {
struct ModelParams {
  double h2d_gbps = 6.0;
  double d2h_gbps = 6.0;
  double device_spmv_gbps = 0.0;  // set by a calibrated backend
  double cpu_gbps = 8.0;          // sustained single-thread SpMV stream rate
  double launch_overhead_ms = 0.01;
  double per_call_overhead_ms = 0.02;
};
ModelParams& model() {
  static ModelParams p;
  return p;
}
std::mutex& model_mutex() {
  static std::mutex m;
  return m;
}
}  // namespace

void HipCostModel::set_device_params(double device_spmv_gbps, double h2d_gbps, double d2h_gbps) {
  std::lock_guard<std::mutex> lk(model_mutex());
  if (device_spmv_gbps > 0.0) model().device_spmv_gbps = device_spmv_gbps;
  if (h2d_gbps > 0.0) model().h2d_gbps = h2d_gbps;
  if (d2h_gbps > 0.0) model().d2h_gbps = d2h_gbps;
}

void HipCostModel::reset_device_params() {
  std::lock_guard<std::mutex> lk(model_mutex());
  model() = ModelParams();
}

double HipCostModel::estimate_cpu_spmv(std::size_t nnz, std::size_t nrows) {
  const double bytes = (double)nnz * (sizeof(double) + sizeof(std::size_t) + sizeof(double)) +
                       (double)nrows * (sizeof(double) + sizeof(std::size_t));
  std::lock_guard<std::mutex> lk(model_mutex());
  return bytes / (model().cpu_gbps * 1e9) * 1e3;
}

double HipCostModel::estimate_gpu_spmv(std::size_t nnz, std::size_t nrows, std::size_t row_ptr_size,
                                       int batch_size) {
  std::lock_guard<std::mutex> lk(model_mutex());
  const ModelParams& p = model();
  const double matrix_bytes = (double)nnz * (sizeof(double) + sizeof(std::size_t)) +
                              (double)row_ptr_size * sizeof(std::size_t);
  const double vector_bytes = (double)nrows * (sizeof(double) + sizeof(std::size_t));
  const double upload_ms = matrix_bytes / (p.h2d_gbps * 1e9) * 1e3 +
                           (vector_bytes / (p.h2d_gbps * 1e9) + vector_bytes / (p.d2h_gbps * 1e9)) * 1e3 * batch_size;
  const double kernel_ms =
      (double)nnz * (2.0 * sizeof(double) + sizeof(std::size_t)) / (p.device_spmv_gbps * 1e9) * 1e3 * batch_size;
  return upload_ms + kernel_ms + p.launch_overhead_ms + p.per_call_overhead_ms;
}

HipCostModel::Decision HipCostModel::decide(const std::vector<double>& values,
                                            const std::vector<std::size_t>& col_indices,
                                            const std::vector<std::size_t>& row_ptr, std::size_t nrows,
                                            std::size_t ncols, int batch_size, bool force_gpu) {
  (void)col_indices;
  (void)ncols;
  Decision d;
  const std::size_t nnz = values.size();
  double rate;
  {
    std::lock_guard<std::mutex> lk(model_mutex());
    rate = model().device_spmv_gbps;
  }
  if (rate <= 0.0 && !force_gpu) {
    d.reason = "no calibrated GPU device; CPU dispatch";
    d.estimated_cpu_time_ms = estimate_cpu_spmv(nnz, nrows) * batch_size;
    return d;
  }
  d.estimated_cpu_time_ms = estimate_cpu_spmv(nnz, nrows) * batch_size;
  d.estimated_gpu_time_ms = estimate_gpu_spmv(nnz, nrows, row_ptr.size(), batch_size);
  const bool big_enough = nnz >= 50000 && nrows >= 2000;
  const double speedup = d.estimated_gpu_time_ms > 0.0 ? d.estimated_cpu_time_ms / d.estimated_gpu_time_ms : 0.0;
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.2f", speedup);
  if (force_gpu || (big_enough && speedup >= 1.5)) {
    d.backend = ComputeBackendType::HIP;
    d.reason = force_gpu ? "GPU execution forced by options"
                         : std::string("GPU estimated ~") + buf + "x faster (nnz=" + std::to_string(nnz) +
                               ", matrix resident, batch=" + std::to_string(batch_size) + ")";
  } else if (!big_enough) {
    d.reason = "problem too small to amortize GPU launch/transfer overhead";
  } else {
    d.reason = std::string("GPU estimate not faster enough than CPU (~") + buf + "x)";
  }
  return d;
}

// ---- AUTO backend -------------------------------------------------------------------
AutoHipBackend::AutoHipBackend(bool force_gpu, std::string library)
    : force_gpu_(force_gpu), library_(std::move(library)) {}

std::string AutoHipBackend::name() const {
  std::lock_guard<std::mutex> lk(mutex_);
  return name_;
}

void AutoHipBackend::initialize() {
  std::lock_guard<std::mutex> lk(mutex_);
  if (initialized_) return;
  initialized_ = true;
  cpu_ = std::make_unique<CpuBackend>();
  auto h = std::make_unique<HipBackend>(library_);
  h->initialize();
  if (h->is_available()) {
    h->calibrate();
    HipCostModel::set_device_params(h->measured_spmv_gbps(), h->measured_h2d_gbps(), h->measured_d2h_gbps());
    hip_ = std::move(h);
  }
  name_ = hip_ ? "Auto(" + hip_->name() + " | CPU fallback)" : "Auto(CPU)";
}

void AutoHipBackend::finalize() {
  std::lock_guard<std::mutex> lk(mutex_);
  if (hip_) hip_->finalize();
}

void AutoHipBackend::spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                          const std::vector<std::size_t>& row_ptr, const std::vector<double>& x,
                          std::vector<double>& y) {
  const std::size_t nrows = row_ptr.size() >= 1 ? row_ptr.size() - 1 : 0;
  const auto d = HipCostModel::decide(values, col_indices, row_ptr, nrows, 0, 1, force_gpu_);
  if (d.backend == ComputeBackendType::HIP && hip_ && hip_->is_available()) {
    try {
      hip_->spmv(values, col_indices, row_ptr, x, y);
      return;
    } catch (const std::exception&) {
      // a device failure: this call runs on the CPU
    }
  }
  cpu_->spmv(values, col_indices, row_ptr, x, y);
}

void AutoHipBackend::spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                          const std::vector<std::size_t>& row_ptr, const std::vector<double>& B,
                          std::vector<double>& C, std::size_t ncols_B) {
  const std::size_t nrows = row_ptr.size() >= 1 ? row_ptr.size() - 1 : 0;
  const auto d = HipCostModel::decide(values, col_indices, row_ptr, nrows, ncols_B, (int)ncols_B, force_gpu_);
  if (d.backend == ComputeBackendType::HIP && hip_ && hip_->is_available()) {
    try {
      hip_->spmm(values, col_indices, row_ptr, B, C, ncols_B);
      return;
    } catch (const std::exception&) {
      // a device failure: this call runs on the CPU
    }
  }
  cpu_->spmm(values, col_indices, row_ptr, B, C, ncols_B);
}

double AutoHipBackend::dot(const std::vector<double>& a, const std::vector<double>& b) {
  if (hip_ && a.size() >= 10000) return hip_->dot(a, b);
  return cpu_->dot(a, b);
}

void AutoHipBackend::axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) {
  if (hip_ && x.size() >= 10000)
    hip_->axpy(alpha, x, y);
  else
    cpu_->axpy(alpha, x, y);
}

double AutoHipBackend::norm_inf(const std::vector<double>& v) {
  if (hip_ && v.size() >= 10000) return hip_->norm_inf(v);
  return cpu_->norm_inf(v);
}

std::unique_ptr<IComputeBackend> create_auto_backend(bool force_gpu, std::string library) {
  auto b = std::make_unique<AutoHipBackend>(force_gpu, std::move(library));
  b->initialize();
  return b;
}

}  // namespace hn::backend
