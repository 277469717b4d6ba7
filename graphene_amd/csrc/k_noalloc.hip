// Allocation-free ||x||_2 for callers that cannot allocate device memory.
//
// The grid reduction of the norm kernel needs scratch memory: the arrival
// counters, one partial sum per block and the state block that receives the
// result. gph_vector_norm takes it from the matrix (allocated when the matrix
// is created); this variant takes it from `scratch`, a static device array of
// fixed size that each GPU holds for the life of the process. Host code never
// uses the array's name as a device address: the address is looked up once
// per device with hipGetSymbolAddress and kept.
//
// Every call on one GPU shares that GPU's array. A call clears the counters
// it needs before its reduction, so the contents left by earlier calls do not
// matter; two calls that overlap on one GPU (different threads or streams)
// race on the array and must not be issued.
#include <mutex>

#include "kernels.hpp"

namespace gph {

namespace {

// Largest grid of the vector kernels (vecGrid) and one double per block;
// 16 KiB of arrival counters before the partials; the state block after.
constexpr int kNoAllocMaxBlocks = 1024;
constexpr size_t kNoAllocPartials = (size_t)kNoAllocMaxBlocks * sizeof(double);
constexpr size_t kNoAllocStateOff = kArriveBytes + kNoAllocPartials;
constexpr size_t kNoAllocBytes = kNoAllocStateOff + 256;
static_assert(sizeof(DevState) <= 256, "state block fits its slot");
static_assert(kNoAllocStateOff % 256 == 0, "state block aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per
// device, with that device current).
void* scratchAddress() {
  int dev = 0;
  GPH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    fail(kRuntime, "device id " + std::to_string(dev) + " beyond the supported 64");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    GPH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(kHip, "hipGetSymbolAddress returned no address for the scratch array");
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace

size_t noAllocScratchBytes() { return kNoAllocBytes; }

double norm2NoAlloc(const void* x, DType t, int64_t n, hipStream_t s) {
  if (t != DType::F32 && t != DType::F64)
    fail(kUnsupported, "norm: type must be float32 or float64");
  if (n < 0 || n > 0x7fffffffLL)
    fail(kInvalid, "allocation-free norm: n = " + std::to_string(n) +
                       " is beyond the maximum of 2^31 - 1 elements");
  if ((size_t)vecGrid(n) * sizeof(double) > kNoAllocPartials)
    fail(kInvalid, "allocation-free norm: the reduction grid exceeds the scratch array");
  char* base = static_cast<char*>(scratchAddress());
  DevState* st = reinterpret_cast<DevState*>(base + kNoAllocStateOff);
  // counters to zero: the first-level arrival counters and the state's
  // top-level counter (a fill kernel, so the call stays capturable)
  launchFill(base, DType::F32, (int64_t)(kNoAllocBytes / 4), 0.0, s);
  VecArgs v;
  v.n = (int32_t)n;
  v.a = x;
  v.state = st;
  v.partials = base;
  v.fin = kFinNormAux;
  launchNorm2(v, t, s);
  double out = 0;
  GPH_HIP(hipMemcpyAsync(&out, &st->aux, sizeof(double), hipMemcpyDeviceToHost, s));
  GPH_HIP(hipStreamSynchronize(s));
  return out;
}

}  // namespace gph
