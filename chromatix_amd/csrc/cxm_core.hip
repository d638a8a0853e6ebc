// Library state of libcxm: thread-local error text, per-device twiddle
// tables, version.  The tables live in a file-scope __device__ array (one
// copy per GPU) filled once per device by a kernel that evaluates the
// twiddles in FP64, so no allocation is ever made for them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "cxm_internal.h"

namespace cxm {

namespace {
thread_local std::string g_last_error;
constexpr int kTwTotal = tw_total();
constexpr int kMaxDevices = 64;
std::mutex g_init_mu;
bool g_ready[kMaxDevices] = {};
float2* g_tw_ptr[kMaxDevices] = {};
float2* g_scratch_ptr[kMaxDevices] = {};
int g_cus[kMaxDevices] = {};
}  // namespace

__device__ __align__(256) float2 g_twiddles[kTwTotal];

// Workspace of the allocation-free entry points (cxm_prop2d_c64_noalloc):
// kNoallocScratchBytes per GPU, held for the life of the process.  In host
// code `scratch` names a host-side stand-in, not device memory: kernels get
// the per-device address taken by hipGetSymbolAddress in init_device.
__device__ __align__(256) float2 scratch[kNoallocScratchBytes / sizeof(float2)];

__global__ void init_twiddles_kernel(float2* tw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTwTotal) return;
  int off = 0, L = 1;
  for (int j = 0; j < kTwCount; ++j) {
    if (i < off + kTwLengths[j]) { L = kTwLengths[j]; break; }
    off += kTwLengths[j];
  }
  const int k = i - off;
  // exp(-2 pi i k / L) in FP64, rounded once to f32
  double s, c;
  sincospi(-2.0 * (double)k / (double)L, &s, &c);
  tw[i] = make_float2((float)c, (float)s);
}

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

static int init_device(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail(CXM_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ready[dev]) return CXM_OK;
  int prev = 0;
  CXM_HIP_CHECK(hipGetDevice(&prev));
  if (prev != dev) CXM_HIP_CHECK(hipSetDevice(dev));
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_twiddles));
  if (e != hipSuccess) {
    if (prev != dev) (void)hipSetDevice(prev);
    return fail(CXM_ERR_HIP, std::string("hipGetSymbolAddress(twiddles): ") + hipGetErrorString(e));
  }
  void* sp = nullptr;
  e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    if (prev != dev) (void)hipSetDevice(prev);
    return fail(CXM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(init_twiddles_kernel, dim3((kTwTotal + 255) / 256), dim3(256), 0, 0, (float2*)p);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (prev != dev) (void)hipSetDevice(prev);
  if (e != hipSuccess) return fail(CXM_ERR_HIP, std::string("twiddle init: ") + hipGetErrorString(e));
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  g_cus[dev] = cus;
  g_tw_ptr[dev] = (float2*)p;
  g_scratch_ptr[dev] = (float2*)sp;
  g_ready[dev] = true;
  return CXM_OK;
}

int device_cus() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  return (dev >= 0 && dev < kMaxDevices && g_cus[dev] > 0) ? g_cus[dev] : 256;
}

int ensure_init() {
  int dev = 0;
  CXM_HIP_CHECK(hipGetDevice(&dev));
  if (dev >= 0 && dev < kMaxDevices && g_ready[dev]) return CXM_OK;
  return init_device(dev);
}

float2* noalloc_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  return g_scratch_ptr[dev];
}

const float2* twiddles_for(int n) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int off = tw_offset(n);
  return off < 0 ? nullptr : g_tw_ptr[dev] + off;
}

}  // namespace cxm

extern "C" int cxm_init(int device) { return cxm::init_device(device); }
extern "C" const char* cxm_last_error(void) { return cxm::g_last_error.c_str(); }
extern "C" int cxm_version(void) { return 100; }
extern "C" size_t cxm_noalloc_scratch_bytes(void) { return cxm::kNoallocScratchBytes; }
extern "C" int cxm_fft_supported(int64_t n) { return cxm::fft_len_supported(n) ? 1 : 0; }

// ---------------------------------------------------------------------------
// Optional per-kernel timing with HIP events recorded on the launch stream
// around every kernel (off by default; used by the benchmark).
// ---------------------------------------------------------------------------
#include <map>
#include <vector>

namespace cxm {
namespace {
struct TimingRec { std::string name; hipEvent_t a, b; };
std::mutex g_tmu;
bool g_timing = false;
std::vector<TimingRec> g_recs;
std::vector<std::pair<std::string, std::pair<double, long long>>> g_summary;
}  // namespace

bool timing_on() { return g_timing; }

void* timing_start(hipStream_t s) {
  if (!g_timing) return nullptr;
  hipEvent_t a;
  if (hipEventCreate(&a) != hipSuccess) return nullptr;
  (void)hipEventRecord(a, s);
  return (void*)a;
}

void timing_stop(void* start, hipStream_t s, const std::string& name) {
  if (!start) return;
  hipEvent_t b;
  if (hipEventCreate(&b) != hipSuccess) return;
  (void)hipEventRecord(b, s);
  std::lock_guard<std::mutex> lk(g_tmu);
  g_recs.push_back({name, (hipEvent_t)start, b});
}
}  // namespace cxm

extern "C" int cxm_timing_enable(int on) {
  using namespace cxm;
  std::lock_guard<std::mutex> lk(g_tmu);
  for (auto& r : g_recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
  g_recs.clear();
  g_summary.clear();
  g_timing = on != 0;
  return CXM_OK;
}

// Waits for the recorded events and aggregates per kernel name; returns the
// number of distinct kernels (then read with cxm_timing_get).
extern "C" int cxm_timing_collect(void) {
  using namespace cxm;
  std::lock_guard<std::mutex> lk(g_tmu);
  std::map<std::string, std::pair<double, long long>> agg;
  for (auto& r : g_recs) {
    if (hipEventSynchronize(r.b) != hipSuccess) return -fail(CXM_ERR_HIP, "event sync failed");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return -fail(CXM_ERR_HIP, "elapsed failed");
    auto& e = agg[r.name];
    e.first += ms;
    e.second += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  g_recs.clear();
  g_summary.assign(agg.begin(), agg.end());
  return (int)g_summary.size();
}

extern "C" int cxm_timing_get(int i, char* name, int name_len, double* total_ms, int64_t* launches) {
  using namespace cxm;
  std::lock_guard<std::mutex> lk(g_tmu);
  if (i < 0 || i >= (int)g_summary.size()) return fail(CXM_ERR_INVALID, "timing index out of range");
  const auto& e = g_summary[i];
  if (name && name_len > 0) {
    strncpy(name, e.first.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_ms) *total_ms = e.second.first;
  if (launches) *launches = e.second.second;
  return CXM_OK;
}

// ---------------------------------------------------------------------------
// Tuning options (launch-shape variants of the kernels; see cxm.h).
// ---------------------------------------------------------------------------
namespace cxm {
constexpr int kNumOpts = 32;
namespace { int g_opts[kNumOpts] = {0}; }
int option(int key) { return (key >= 0 && key < kNumOpts) ? g_opts[key] : 0; }

}  // namespace cxm

extern "C" int cxm_set_option(int key, int value) {
  if (key < 0 || key >= cxm::kNumOpts) return cxm::fail(CXM_ERR_INVALID, "unknown option");
  cxm::g_opts[key] = value;
  return CXM_OK;
}
