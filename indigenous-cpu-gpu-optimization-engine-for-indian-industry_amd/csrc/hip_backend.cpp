// HipBackend: IComputeBackend over the dlopen'ed libhn_kernels.so.
#include "hip_backend.hpp"

#include <dlfcn.h>

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
  X(int, hn_inject_failures, (uint32_t, int))

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

}  // namespace hn::backend
