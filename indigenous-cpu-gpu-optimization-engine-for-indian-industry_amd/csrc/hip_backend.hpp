// C++ host side of the HIP compute backend.
//
// hn::backend::IComputeBackend has exactly the virtual interface of the
// reference's hypernova::execution::IComputeBackend
// (core/execution/gpu_backend.hpp:18-51): same names, argument types and
// meaning, so HipBackend can be registered for ComputeBackendType::HIP with
// a one-line adapter in the reference's factory.  HipBackend loads libhn_kernels.so
// at run time with dlopen and binds its C entry points through one table,
// the way the reference binds its driver (core/execution/cuda_driver.cpp:24-48,
// 112-145), so a host without a GPU or without the library still links and
// starts; after a successful initialize(), failures throw std::runtime_error
// like CudaDriverBackend (gpu_backend.cpp:12-14).
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace hn::backend {

enum class ComputeBackendType { CPU, CUDA, HIP, SYCL, AUTO };  // gpu_backend.hpp:10-16

class IComputeBackend {
 public:
  virtual ~IComputeBackend() = default;
  virtual ComputeBackendType type() const = 0;
  virtual std::string name() const = 0;
  virtual bool is_available() const = 0;
  virtual void initialize() = 0;
  virtual void finalize() = 0;
  virtual void spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                    const std::vector<std::size_t>& row_ptr, const std::vector<double>& x,
                    std::vector<double>& y) = 0;
  virtual void spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                    const std::vector<std::size_t>& row_ptr, const std::vector<double>& B,
                    std::vector<double>& C, std::size_t ncols_B) = 0;
  virtual double dot(const std::vector<double>& a, const std::vector<double>& b) = 0;
  virtual void axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) = 0;
  virtual double norm_inf(const std::vector<double>& v) = 0;
  virtual std::size_t device_memory() const = 0;
  virtual std::size_t free_memory() const = 0;
  virtual bool was_gpu_kernel_executed() const { return false; }
  virtual std::size_t gpu_kernel_executions() const { return 0; }
  virtual void reset_execution_stats() {}
};

struct KernelApi;  // the dlopen'ed entry points of libhn_kernels.so

class HipBackend final : public IComputeBackend {
 public:
  // library: path of libhn_kernels.so ("" = next to this library, then the
  // dynamic loader's search path).
  explicit HipBackend(std::string library = "");
  ~HipBackend() override;
  HipBackend(const HipBackend&) = delete;
  HipBackend& operator=(const HipBackend&) = delete;

  ComputeBackendType type() const override { return ComputeBackendType::HIP; }
  std::string name() const override;
  bool is_available() const override { return available_; }
  void initialize() override;
  void finalize() override;

  // Matrix residency: one device copy per (values, col_indices, row_ptr)
  // container triple, re-uploaded when a size or a content fingerprint
  // changes (the reference keys only by values.data(), gpu_backend.cpp:287-306,
  // which returns stale results when a new matrix reuses the buffer).
  void spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
            const std::vector<std::size_t>& row_ptr, const std::vector<double>& x,
            std::vector<double>& y) override;
  void spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
            const std::vector<std::size_t>& row_ptr, const std::vector<double>& B, std::vector<double>& C,
            std::size_t ncols_B) override;
  double dot(const std::vector<double>& a, const std::vector<double>& b) override;
  void axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) override;
  double norm_inf(const std::vector<double>& v) override;
  std::size_t device_memory() const override;
  std::size_t free_memory() const override;
  bool was_gpu_kernel_executed() const override;
  std::size_t gpu_kernel_executions() const override;
  void reset_execution_stats() override;

  const std::string& last_error() const { return last_error_; }
  // Test hook (hn_inject_failures): the next `count` library calls fail as a
  // device failure would, so a caller's CPU fallback can be exercised.
  void inject_failures(unsigned count);

 private:
  struct Resident;
  Resident& resident(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                     const std::vector<std::size_t>& row_ptr, std::size_t ncols);
  void require(const char* op) const;
  [[noreturn]] void raise(const char* op) const;
  void release_all();

  std::string library_;
  void* handle_ = nullptr;
  std::unique_ptr<KernelApi> api_;
  bool initialized_ = false;
  bool available_ = false;
  std::string device_name_;
  std::string last_error_;
  std::vector<std::unique_ptr<Resident>> cache_;
  mutable std::mutex mutex_;
};

// Factory mirroring ComputeBackendFactory::create_backend for the HIP slot
// (gpu_backend.cpp:911-935): never throws; returns nullptr when no usable
// device or library is present (the caller then keeps its CPU backend).
std::unique_ptr<IComputeBackend> create_hip_backend(std::string library = "");
// Empty when a GPU is usable, else the reason (cuda_backend_probe_error,
// gpu_backend.cpp:948-952).
std::string hip_backend_probe_error(std::string library = "");

}  // namespace hn::backend
