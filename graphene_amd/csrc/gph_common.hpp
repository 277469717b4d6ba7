// Shared definitions of the MI355X sparse-solver library: element types,
// error handling and an owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>

namespace gph {

// Element types. Names follow lg/common/Type.cpp:69-88 ("float16",
// "float32", "float64", "twofloat32"). TF32 is the double-word type
// two<float> (a float pair {hi, lo}, 8 bytes; dword.hpp), used as the
// extended precision of mixed-precision refinement.
enum class DType : int32_t { F16 = 1, F32 = 2, F64 = 3, TF32 = 4 };

inline size_t dtypeSize(DType t) {
  switch (t) {
    case DType::F16:
      return 2;
    case DType::F32:
      return 4;
    case DType::F64:
    case DType::TF32:
      return 8;
  }
  throw std::invalid_argument("unknown element type");
}

inline const char* dtypeName(DType t) {
  switch (t) {
    case DType::F16:
      return "float16";
    case DType::F32:
      return "float32";
    case DType::F64:
      return "float64";
    case DType::TF32:
      return "twofloat32";
  }
  return "?";
}

// Status codes of the C-ABI (mirrored in include/graphene_amd.h).
enum Status : int32_t {
  kOk = 0,
  kInvalid = 1,
  kRuntime = 2,
  kHip = 3,
  kUnsupported = 4,
};

struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int32_t code, const std::string& msg) {
  throw Error(code, msg);
}

inline void hipCheck(hipError_t e, const char* what, const char* file,
                     int line) {
  if (e != hipSuccess) {
    fail(kHip, std::string(what) + " failed: " + hipGetErrorString(e) + " (" +
                   file + ":" + std::to_string(line) + ")");
  }
}
#define GPH_HIP(x) ::gph::hipCheck((x), #x, __FILE__, __LINE__)

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }

  void alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 256;
    GPH_HIP(hipMalloc(&p_, bytes));
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T>
  void upload(const T* host, size_t count, hipStream_t s = nullptr) {
    alloc(count * sizeof(T));
    if (count) {
      GPH_HIP(hipMemcpyAsync(p_, host, count * sizeof(T),
                             hipMemcpyHostToDevice, s));
      GPH_HIP(hipStreamSynchronize(s));
    }
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p_);
  }
  void* get() const { return p_; }
  size_t bytes() const { return n_; }
  explicit operator bool() const { return p_ != nullptr; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace gph
