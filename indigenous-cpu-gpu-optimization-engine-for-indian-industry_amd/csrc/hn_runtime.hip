// Lifecycle, error slot, device buffers and events of libhn_kernels.so.
//
// Mirrors the responsibilities of the reference's dynamically loaded driver
// layer (core/execution/cuda_driver.cpp:65-145,204-236): device selection by
// largest memory, one execution context per process, a per-failure message.
// Unlike the reference, the HIP runtime is linked directly (this library is
// itself the dlopen'ed unit) and every entry point re-establishes the device,
// so any host thread may call.
#include "hn_internal.hpp"

#include <cstring>
#include <sstream>

namespace hn {

namespace {
thread_local std::string t_last_error;
std::mutex g_init_mtx;
}  // namespace

void set_error(const std::string& msg) { t_last_error = msg; }
const std::string& get_error() { return t_last_error; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  std::ostringstream os;
  os << what << " failed: " << hipGetErrorName(e) << " (" << hipGetErrorString(e) << ") at "
     << file << ":" << line;
  int code = (e == hipErrorOutOfMemory) ? HN_ERR_NOMEM : HN_ERR_HIP;
  return fail(code, os.str());
}

Context& ctx() {
  static Context c;
  return c;
}

// Failures forced from the host (hn_inject_failures): the next g_inject
// calls of enter() fail with g_inject_code before touching the device.
namespace {
std::atomic<int64_t> g_inject{0};
std::atomic<int> g_inject_code{HN_ERR_HIP};
}  // namespace

int enter() {
  if (g_inject.load(std::memory_order_relaxed) > 0 && g_inject.fetch_sub(1) > 0)
    return fail(g_inject_code.load(), "injected failure (hn_inject_failures)");
  Context& c = ctx();
  if (!c.initialized) return fail(HN_ERR_NOT_INIT, "hn_init() has not been called");
  int cur = -1;
  HN_HIP(hipGetDevice(&cur));
  if (cur != c.device) HN_HIP(hipSetDevice(c.device));
  return HN_OK;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what, __FILE__, __LINE__);
  return HN_OK;
}

int DevBuf::reserve(size_t want) {
  if (want <= bytes && p) return HN_OK;
  release();
  if (want == 0) want = 256;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  }
  bytes = want;
  return HN_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

HostOpScratch& host_scratch() {
  static HostOpScratch* s = new HostOpScratch();  // released in hn_finalize
  return *s;
}

}  // namespace hn

using namespace hn;

extern "C" int hn_init(int device, hn_device_info* info) {
  std::lock_guard<std::mutex> lk(g_init_mtx);
  Context& c = ctx();
  if (c.initialized) {
    if (info) *info = c.info;
    return HN_OK;
  }
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count <= 0)
    return fail(HN_ERR_NO_DEVICE, "no HIP device visible");
  int best = device;
  if (best < 0) {
    size_t best_mem = 0;
    for (int d = 0; d < count; ++d) {
      hipDeviceProp_t p;
      if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
      if (p.totalGlobalMem > best_mem) {
        best_mem = p.totalGlobalMem;
        best = d;
      }
    }
    if (best < 0) best = 0;
  }
  HN_CHECK(best < count, HN_ERR_INVALID, "device ordinal out of range");
  hipDeviceProp_t p;
  HN_HIP(hipGetDeviceProperties(&p, best));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
    return fail(HN_ERR_NO_DEVICE, std::string("device is ") + p.gcnArchName +
                                      ", this library is built for gfx950 only");
  HN_HIP(hipSetDevice(best));
  HN_TRY(vec_noalloc_setup(best));
  HN_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  std::memset(&c.info, 0, sizeof(c.info));
  std::snprintf(c.info.name, sizeof(c.info.name), "%s", p.name);
  std::snprintf(c.info.arch, sizeof(c.info.arch), "%s", p.gcnArchName);
  c.info.total_memory = p.totalGlobalMem;
  c.info.compute_units = p.multiProcessorCount;
  c.info.device = best;
  c.device = best;
  c.initialized = true;
  if (info) *info = c.info;
  return HN_OK;
}

extern "C" int hn_finalize(void) {
  std::lock_guard<std::mutex> lk(g_init_mtx);
  Context& c = ctx();
  if (!c.initialized) return HN_OK;
  (void)hipSetDevice(c.device);
  (void)hipStreamSynchronize(c.stream);
  {
    HostOpScratch& s = host_scratch();
    std::lock_guard<std::mutex> lk2(s.mtx);
    s.a.release();
    s.b.release();
    s.c.release();
    s.red.release();
  }
  release_verify_scratch();
  reset_batch_tuning();
  (void)hipStreamDestroy(c.stream);
  c.stream = nullptr;
  c.initialized = false;
  return HN_OK;
}

extern "C" int hn_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return HN_ERR_INVALID;
  std::snprintf(buf, len, "%s", get_error().c_str());
  return HN_OK;
}

extern "C" int hn_mem_info(size_t* free_bytes, size_t* total_bytes) {
  HN_TRY(enter());
  size_t f = 0, t = 0;
  HN_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return HN_OK;
}

extern "C" uint64_t hn_kernel_executions(void) { return ctx().launches.load(); }
extern "C" void hn_reset_execution_stats(void) { ctx().launches.store(0); }

extern "C" int hn_inject_failures(uint32_t count, int code) {
  HN_CHECK(code != HN_OK, HN_ERR_INVALID, "hn_inject_failures: the injected code must be an error");
  g_inject_code.store(code);
  g_inject.store((int64_t)count);
  return HN_OK;
}

extern "C" int hn_dev_alloc(size_t bytes, void** ptr) {
  HN_CHECK(ptr, HN_ERR_INVALID, "hn_dev_alloc: null out pointer");
  HN_TRY(enter());
  HN_HIP(hipMalloc(ptr, bytes ? bytes : 256));
  return HN_OK;
}

extern "C" int hn_dev_free(void* ptr) {
  HN_TRY(enter());
  if (ptr) HN_HIP(hipFree(ptr));
  return HN_OK;
}

extern "C" int hn_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HN_TRY(enter());
  if (bytes == 0) return HN_OK;
  HN_CHECK(dst && src, HN_ERR_INVALID, "hn_memcpy_h2d: null pointer");
  HN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().stream));
  HN_HIP(hipStreamSynchronize(ctx().stream));
  return HN_OK;
}

extern "C" int hn_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HN_TRY(enter());
  if (bytes == 0) return HN_OK;
  HN_CHECK(dst && src, HN_ERR_INVALID, "hn_memcpy_d2h: null pointer");
  HN_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream));
  HN_HIP(hipStreamSynchronize(ctx().stream));
  return HN_OK;
}

extern "C" int hn_stream_sync(void* stream) {
  HN_TRY(enter());
  HN_HIP(hipStreamSynchronize(pick_stream(stream)));
  return HN_OK;
}

namespace {
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

extern "C" int hn_copy_dev(void* dst, const void* src, size_t bytes, void* stream) {
  HN_CHECK(bytes % 16 == 0, HN_ERR_INVALID, "hn_copy_dev: bytes must be a multiple of 16");
  HN_CHECK(bytes == 0 || (dst && src), HN_ERR_INVALID, "hn_copy_dev: null pointer");
  HN_TRY(enter());
  const long long n = (long long)(bytes / 16);
  if (n == 0) return HN_OK;
  hipLaunchKernelGGL(k_copy16, dim3(grid_for(n, 256, 65536)), dim3(256), 0, pick_stream(stream),
                     (const uint4*)src, (uint4*)dst, n);
  return check_launch("k_copy16");
}

struct hn_event {
  hipEvent_t ev = nullptr;
};

extern "C" int hn_event_create(hn_event** ev) {
  HN_CHECK(ev, HN_ERR_INVALID, "hn_event_create: null out pointer");
  HN_TRY(enter());
  hn_event* e = new hn_event();
  hipError_t r = hipEventCreate(&e->ev);
  if (r != hipSuccess) {
    delete e;
    return hip_fail(r, "hipEventCreate", __FILE__, __LINE__);
  }
  *ev = e;
  return HN_OK;
}

extern "C" int hn_event_destroy(hn_event* ev) {
  if (!ev) return HN_OK;
  HN_TRY(enter());
  (void)hipEventDestroy(ev->ev);
  delete ev;
  return HN_OK;
}

extern "C" int hn_event_record(hn_event* ev, void* stream) {
  HN_CHECK(ev, HN_ERR_INVALID, "hn_event_record: null event");
  HN_TRY(enter());
  HN_HIP(hipEventRecord(ev->ev, pick_stream(stream)));
  return HN_OK;
}

extern "C" int hn_event_elapsed_ms(const hn_event* start, const hn_event* stop, float* ms) {
  HN_CHECK(start && stop && ms, HN_ERR_INVALID, "hn_event_elapsed_ms: null argument");
  HN_TRY(enter());
  HN_HIP(hipEventSynchronize(stop->ev));
  HN_HIP(hipEventElapsedTime(ms, start->ev, stop->ev));
  return HN_OK;
}
