// C++ tests of hn::backend::HipBackend, written after the reference's own
// backend tests (tests/unit/test_gpu_backend.cpp:26-292): the same random
// matrices (std::mt19937 recipe of make_problem, :26-49), the same
// tolerance rule |got - ref| <= tol * max(1, |ref|) (:67-74), the same cases.
// Usage: test_hip_backend [--no-gpu]   (--no-gpu runs only the fallback cases)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "hip_backend.hpp"

using namespace hn::backend;

namespace {

int g_failures = 0;
#define EXPECT(cond)                                                                   \
  do {                                                                                 \
    if (!(cond)) {                                                                     \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                  \
      ++g_failures;                                                                    \
    }                                                                                  \
  } while (0)

struct Problem {
  std::vector<double> values;
  std::vector<std::size_t> col_indices, row_ptr;
  std::vector<double> x;
  std::size_t nrows = 0, ncols = 0;
};

Problem make_problem(std::size_t nrows, std::size_t ncols, double density, std::uint32_t seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> val(-4.0, 4.0);
  Problem p;
  p.nrows = nrows;
  p.ncols = ncols;
  p.row_ptr.push_back(0);
  for (std::size_t i = 0; i < nrows; ++i) {
    for (std::size_t j = 0; j < ncols; ++j) {
      if (std::uniform_real_distribution<double>(0.0, 1.0)(rng) < density) {
        p.values.push_back(val(rng));
        p.col_indices.push_back(j);
      }
    }
    p.row_ptr.push_back(p.values.size());
  }
  p.x.resize(ncols);
  for (auto& v : p.x) v = val(rng);
  return p;
}

std::vector<double> cpu_spmv(const Problem& p) {
  std::vector<double> y(p.nrows, 0.0);
  for (std::size_t i = 0; i < p.nrows; ++i) {
    double s = 0.0;
    for (std::size_t k = p.row_ptr[i]; k < p.row_ptr[i + 1]; ++k) s += p.values[k] * p.x[p.col_indices[k]];
    y[i] = s;
  }
  return y;
}

void expect_close(const std::vector<double>& got, const std::vector<double>& ref, double tol) {
  EXPECT(got.size() == ref.size());
  if (got.size() != ref.size()) return;
  for (std::size_t i = 0; i < ref.size(); ++i)
    if (std::abs(got[i] - ref[i]) > tol * std::max(1.0, std::abs(ref[i]))) {
      std::printf("  row %zu: %.17g vs %.17g\n", i, got[i], ref[i]);
      ++g_failures;
      return;
    }
}

void run(const char* name, const std::function<void()>& fn) {
  const int before = g_failures;
  std::printf("[ RUN  ] %s\n", name);
  std::fflush(stdout);
  try {
    fn();
  } catch (const std::exception& e) {
    std::printf("  exception: %s\n", e.what());
    ++g_failures;
  }
  std::printf("[ %s ] %s\n", g_failures == before ? " OK " : "FAIL", name);
  std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
  const bool gpu = !(argc > 1 && std::strcmp(argv[1], "--no-gpu") == 0);

  run("HipBackendTest.FactoryReturnsNullWithoutLibrary", [] {
    // gpu_backend.cpp:911-935: the factory never throws; no device -> no backend
    EXPECT(create_hip_backend("/nonexistent/libhn_kernels.so") == nullptr);
    EXPECT(!hip_backend_probe_error("/nonexistent/libhn_kernels.so").empty());
    HipBackend b("/nonexistent/libhn_kernels.so");
    b.initialize();
    EXPECT(!b.is_available());
    bool threw = false;
    try {
      std::vector<double> y;
      b.spmv({1.0}, {0}, {0, 1}, {1.0}, y);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    EXPECT(threw);
  });
  run("HipCostModelTest.UnsetDeviceMeansCPU", [] {  // test_gpu_backend.cpp:133-141
    HipCostModel::reset_device_params();
    const Problem p = make_problem(5000, 5000, 0.01, 123);
    const auto d = HipCostModel::decide(p.values, p.col_indices, p.row_ptr, p.nrows, p.ncols);
    EXPECT(d.backend == ComputeBackendType::CPU);
    EXPECT(!d.reason.empty());
    EXPECT(d.estimated_cpu_time_ms >= 0.0);
  });
  run("HipCostModelTest.LargeBatchFavorsGPUWhenCalibrated", [] {  // :143-153
    const Problem p = make_problem(6000, 4000, 0.05, 321);
    HipCostModel::set_device_params(700.0, 8.0, 8.0);
    const auto d = HipCostModel::decide(p.values, p.col_indices, p.row_ptr, p.nrows, p.ncols, 64);
    EXPECT(d.backend == ComputeBackendType::HIP);
    HipCostModel::reset_device_params();
  });
  run("HipCostModelTest.SmallProblemStaysOnCPU", [] {  // :155-161
    Problem p;  // make_diagonal(100)
    p.nrows = p.ncols = 100;
    p.row_ptr.push_back(0);
    for (std::size_t i = 0; i < 100; ++i) {
      p.values.push_back(1.0 + i);
      p.col_indices.push_back(i);
      p.row_ptr.push_back(i + 1);
    }
    HipCostModel::set_device_params(700.0, 8.0, 8.0);
    const auto d = HipCostModel::decide(p.values, p.col_indices, p.row_ptr, p.nrows, p.ncols);
    EXPECT(d.backend == ComputeBackendType::CPU);
    HipCostModel::reset_device_params();
  });
  run("AutoHipBackendTest.FallsBackToCPUWithoutGPU", [] {  // :163-175
    auto b = create_auto_backend(false, "/nonexistent/libhn_kernels.so");
    EXPECT(b && b->is_available());
    EXPECT(b->name() == "Auto(CPU)");
    const Problem p = make_problem(64, 48, 0.3, 5);
    std::vector<double> y;
    b->spmv(p.values, p.col_indices, p.row_ptr, p.x, y);
    expect_close(y, cpu_spmv(p), 1e-12);
    b->finalize();
  });
  if (!gpu) return g_failures ? 1 : 0;

  auto be = create_hip_backend();
  if (!be) {
    std::printf("no usable HIP device: %s\n", hip_backend_probe_error().c_str());
    return 2;
  }
  std::printf("backend: %s\n", be->name().c_str());

  run("HipBackendTest.SpmvShortRowsMatchesCpu", [&] {  // test_gpu_backend.cpp:211-222
    const Problem p = make_problem(1500, 1200, 0.003, 77);
    std::vector<double> y;
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y);
    expect_close(y, cpu_spmv(p), 1e-10);
  });
  run("HipBackendTest.SpmvLongRowsMatchesCpu", [&] {  // :224-234
    const Problem p = make_problem(300, 800, 0.08, 4242);
    std::vector<double> y;
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y);
    expect_close(y, cpu_spmv(p), 1e-10);
  });
  run("HipBackendTest.SpmvEmptyRows", [&] {  // :236-258
    std::vector<double> vals;
    std::vector<std::size_t> cols, rp{0};
    for (std::size_t i = 0; i < 256; ++i) {
      if (i % 4 != 0) {
        vals.push_back(2.0);
        cols.push_back(i);
      }
      rp.push_back(vals.size());
    }
    std::vector<double> x(256, 1.5), y;
    be->spmv(vals, cols, rp, x, y);
    EXPECT(y.size() == 256);
    for (std::size_t i = 0; i < y.size(); ++i) EXPECT(std::abs(y[i] - ((i % 4 == 0) ? 0.0 : 3.0)) <= 1e-12);
  });
  run("HipBackendTest.EmptyMatrixIsSafe", [&] {  // :126-131
    std::vector<double> y;
    be->spmv({}, {}, {0}, {}, y);
    EXPECT(y.empty());
  });
  run("HipBackendTest.SpmmMatchesCpu", [&] {  // :260-276
    const Problem p = make_problem(400, 300, 0.05, 2024);
    const std::size_t k = 5;
    std::vector<double> B(p.ncols * k);
    std::mt19937 rng(8);
    std::uniform_real_distribution<double> val(-1.0, 1.0);
    for (auto& v : B) v = val(rng);
    std::vector<double> C, ref(p.nrows * k, 0.0);
    be->spmm(p.values, p.col_indices, p.row_ptr, B, C, k);
    for (std::size_t i = 0; i < p.nrows; ++i)
      for (std::size_t e = p.row_ptr[i]; e < p.row_ptr[i + 1]; ++e)
        for (std::size_t c = 0; c < k; ++c) ref[i * k + c] += p.values[e] * B[p.col_indices[e] * k + c];
    expect_close(C, ref, 1e-10);
  });
  run("HipBackendTest.MatrixResidencyIsReused", [&] {  // :278-292
    const Problem p = make_problem(2000, 1500, 0.03, 555);
    std::vector<double> y1, y2;
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y1);
    const std::size_t free1 = be->free_memory();
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y2);
    EXPECT(be->free_memory() <= free1 + (1u << 20));
    expect_close(y2, y1, 1e-12);
  });
  run("HipBackendTest.NewContentsSameBuffersNotStale", [&] {
    Problem p = make_problem(500, 400, 0.02, 9);
    std::vector<double> y1, y2;
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y1);
    for (auto& v : p.values) v *= 2.0;  // same containers, new matrix
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y2);
    expect_close(y2, cpu_spmv(p), 1e-10);
  });
  run("HipBackendTest.InjectedFailureFallsBackToCpu", [&] {
    // AutoBackend::spmv (gpu_backend.cpp:688-694): a device call that throws
    // is retried on the CPU for that call only.  The failure is forced from
    // the host; the next call runs on the device again.
    auto* hb = dynamic_cast<HipBackend*>(be.get());
    EXPECT(hb != nullptr);
    const Problem p = make_problem(800, 600, 0.01, 31);
    auto auto_spmv = [&](std::vector<double>& y) -> bool {  // true: the device answered
      try {
        be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y);
        return true;
      } catch (const std::exception&) {
        y = cpu_spmv(p);
        return false;
      }
    };
    std::vector<double> y1, y2;
    hb->inject_failures(1);
    EXPECT(!auto_spmv(y1));
    expect_close(y1, cpu_spmv(p), 0.0);
    EXPECT(auto_spmv(y2));
    expect_close(y2, cpu_spmv(p), 1e-10);
    bool threw = false;
    hb->inject_failures(1);
    try {
      std::vector<double> y;
      be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y);
    } catch (const std::runtime_error& e) {
      threw = std::strstr(e.what(), "injected failure") != nullptr;
    }
    EXPECT(threw);
  });
  run("AutoHipBackendTest.CalibratedDispatch", [&] {
    // AutoBackend over the HIP backend: calibration feeds the cost model; a
    // large SpMV runs on the device, a small one on the CPU, both equal to
    // the CPU product; an injected device failure falls back to the CPU for
    // that call (gpu_backend.cpp:623-746)
    HipCostModel::reset_device_params();
    auto b = create_auto_backend();
    auto* ab = dynamic_cast<AutoHipBackend*>(b.get());
    EXPECT(ab && ab->hip());
    if (!ab || !ab->hip()) return;
    std::printf("  %s: spmv %.1f GB/s, h2d %.1f GB/s, d2h %.1f GB/s\n", b->name().c_str(),
                ab->hip()->measured_spmv_gbps(), ab->hip()->measured_h2d_gbps(), ab->hip()->measured_d2h_gbps());
    EXPECT(b->name().rfind("Auto(", 0) == 0 && b->name().find("CPU fallback") != std::string::npos);
    EXPECT(ab->hip()->measured_spmv_gbps() > 0.0);
    const Problem big = make_problem(20000, 20000, 0.002, 99);  // ~800k nonzeros
    const Problem small = make_problem(300, 300, 0.02, 98);
    const auto db = HipCostModel::decide(big.values, big.col_indices, big.row_ptr, big.nrows, big.ncols);
    const auto ds = HipCostModel::decide(small.values, small.col_indices, small.row_ptr, small.nrows, small.ncols);
    std::printf("  big: %s\n  small: %s\n", db.reason.c_str(), ds.reason.c_str());
    EXPECT(ds.backend == ComputeBackendType::CPU);
    std::vector<double> y;
    b->reset_execution_stats();
    b->spmv(small.values, small.col_indices, small.row_ptr, small.x, y);
    expect_close(y, cpu_spmv(small), 0.0);
    EXPECT(b->gpu_kernel_executions() == 0);
    b->spmv(big.values, big.col_indices, big.row_ptr, big.x, y);
    expect_close(y, cpu_spmv(big), 1e-10);
    EXPECT((db.backend == ComputeBackendType::HIP) == (b->gpu_kernel_executions() > 0));
    ab->set_force_gpu(true);
    ab->hip()->inject_failures(1);
    b->spmv(big.values, big.col_indices, big.row_ptr, big.x, y);
    expect_close(y, cpu_spmv(big), 0.0);  // the CPU answered
    b->finalize();
    HipCostModel::reset_device_params();
  });
  run("HipBackendTest.VectorOps", [&] {  // CPUBackend semantics, gpu_backend.cpp:124-136
    std::vector<double> a(20000), b(20000);
    for (std::size_t i = 0; i < a.size(); ++i) {
      a[i] = std::sin(0.1 * i);
      b[i] = std::cos(0.3 * i);
    }
    double ref = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) ref += a[i] * b[i];
    EXPECT(std::abs(be->dot(a, b) - ref) <= 1e-10 * std::max(1.0, std::abs(ref)));
    std::vector<double> y = b, yr = b;
    be->axpy(0.5, a, y);
    for (std::size_t i = 0; i < yr.size(); ++i) yr[i] += 0.5 * a[i];
    expect_close(y, yr, 0.0);
    a[7] = -3.5;
    EXPECT(be->norm_inf(a) == 3.5);
  });
  run("HipBackendTest.ExecutionStats", [&] {  // gpu_backend.hpp:48-50
    be->reset_execution_stats();
    EXPECT(!be->was_gpu_kernel_executed());
    const Problem p = make_problem(64, 48, 0.3, 5);
    std::vector<double> y;
    be->spmv(p.values, p.col_indices, p.row_ptr, p.x, y);
    EXPECT(be->was_gpu_kernel_executed() && be->gpu_kernel_executions() > 0);
    EXPECT(be->type() == ComputeBackendType::HIP && be->is_available());
  });
  be->finalize();
  std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "PASSED", g_failures);
  return g_failures ? 1 : 0;
}
