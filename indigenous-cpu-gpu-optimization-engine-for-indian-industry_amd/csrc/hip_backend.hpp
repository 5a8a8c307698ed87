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

  // CudaDriverBackend::calibrate (gpu_backend.cpp:489-578): the device-only
  // SpMV rate on a 4096 x 32-per-row synthetic matrix (matrix and vectors
  // resident), and pageable H2D / D2H rates of 1 MiB.  Rates that cannot be
  // measured default to 50, 6 and 6 GB/s.
  void calibrate();
  double measured_spmv_gbps() const { return measured_spmv_gbps_; }
  double measured_h2d_gbps() const { return measured_h2d_gbps_; }
  double measured_d2h_gbps() const { return measured_d2h_gbps_; }

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
  double measured_spmv_gbps_ = 0.0, measured_h2d_gbps_ = 0.0, measured_d2h_gbps_ = 0.0;
  mutable std::mutex mutex_;
};

// CPUBackend (gpu_backend.cpp:80-140): the reference CPU kernels, the
// fallback of the AUTO backend.
class CpuBackend final : public IComputeBackend {
 public:
  ComputeBackendType type() const override { return ComputeBackendType::CPU; }
  std::string name() const override { return "CPU"; }
  bool is_available() const override { return true; }
  void initialize() override {}
  void finalize() override {}
  void spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
            const std::vector<std::size_t>& row_ptr, const std::vector<double>& x, std::vector<double>& y) override;
  void spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
            const std::vector<std::size_t>& row_ptr, const std::vector<double>& B, std::vector<double>& C,
            std::size_t ncols_B) override;
  double dot(const std::vector<double>& a, const std::vector<double>& b) override;
  void axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) override;
  double norm_inf(const std::vector<double>& v) override;
  std::size_t device_memory() const override { return 0; }
  std::size_t free_memory() const override { return 0; }
};

// GPUCostModel (gpu_backend.hpp:64-100, gpu_backend.cpp:756-883) with the
// HIP backend in the GPU slot: the same byte estimates, the same gate
// (nnz >= 50,000 and rows >= 2,000, or forced) and the same 1.5x rule; until
// a calibrated backend sets the device rates, every decision is CPU.
class HipCostModel {
 public:
  struct Decision {
    ComputeBackendType backend = ComputeBackendType::CPU;
    double estimated_cpu_time_ms = 0.0;
    double estimated_gpu_time_ms = 0.0;
    std::string reason;
  };
  static Decision decide(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
                         const std::vector<std::size_t>& row_ptr, std::size_t nrows, std::size_t ncols,
                         int batch_size = 1, bool force_gpu = false);
  static void set_device_params(double device_spmv_gbps, double h2d_gbps, double d2h_gbps);
  // Back to the uncalibrated state (tests).
  static void reset_device_params();
  static double estimate_cpu_spmv(std::size_t nnz, std::size_t nrows);
  static double estimate_gpu_spmv(std::size_t nnz, std::size_t nrows, std::size_t row_ptr_size, int batch_size);
};

// AutoBackend (gpu_backend.cpp:623-746) over HipBackend: calibrates the HIP
// backend at initialize() and feeds its rates to HipCostModel; each SpMV or
// SpMM goes to the device when the model picks it, and a device call that
// throws is retried on the CPU for that call only; dot, axpy and norm_inf go
// to the device from 10,000 entries.  Always available (CPU without a GPU).
class AutoHipBackend final : public IComputeBackend {
 public:
  explicit AutoHipBackend(bool force_gpu = false, std::string library = "");
  ComputeBackendType type() const override { return ComputeBackendType::AUTO; }
  std::string name() const override;
  bool is_available() const override { return true; }
  void initialize() override;
  void finalize() override;
  void spmv(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
            const std::vector<std::size_t>& row_ptr, const std::vector<double>& x, std::vector<double>& y) override;
  void spmm(const std::vector<double>& values, const std::vector<std::size_t>& col_indices,
            const std::vector<std::size_t>& row_ptr, const std::vector<double>& B, std::vector<double>& C,
            std::size_t ncols_B) override;
  double dot(const std::vector<double>& a, const std::vector<double>& b) override;
  void axpy(double alpha, const std::vector<double>& x, std::vector<double>& y) override;
  double norm_inf(const std::vector<double>& v) override;
  std::size_t device_memory() const override { return hip_ ? hip_->device_memory() : 0; }
  std::size_t free_memory() const override { return hip_ ? hip_->free_memory() : 0; }
  bool was_gpu_kernel_executed() const override { return hip_ && hip_->was_gpu_kernel_executed(); }
  std::size_t gpu_kernel_executions() const override { return hip_ ? hip_->gpu_kernel_executions() : 0; }
  void reset_execution_stats() override {
    if (hip_) hip_->reset_execution_stats();
  }
  void set_force_gpu(bool force) { force_gpu_ = force; }
  HipBackend* hip() { return hip_.get(); }

 private:
  bool force_gpu_ = false;
  std::string library_;
  std::unique_ptr<CpuBackend> cpu_;
  std::unique_ptr<HipBackend> hip_;
  std::string name_ = "Auto(CPU)";
  bool initialized_ = false;
  mutable std::mutex mutex_;
};

// Factory mirroring ComputeBackendFactory::create_backend for the HIP slot
// (gpu_backend.cpp:911-935): never throws; returns nullptr when no usable
// device or library is present (the caller then keeps its CPU backend).
std::unique_ptr<IComputeBackend> create_hip_backend(std::string library = "");
// Empty when a GPU is usable, else the reason (cuda_backend_probe_error,
// gpu_backend.cpp:948-952).
std::string hip_backend_probe_error(std::string library = "");
// ComputeBackendFactory::create_backend(AUTO): an initialised
// AutoHipBackend (never null).
std::unique_ptr<IComputeBackend> create_auto_backend(bool force_gpu = false, std::string library = "");

}  // namespace hn::backend
