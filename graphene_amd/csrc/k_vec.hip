// Fused vector kernels of PBiCGStab and MPIR (gfx950).
//
// The reference expresses these as TensorDSL expressions that are fused per
// tile (lg/dsl/tensor/Expression.cpp:375-542) followed by 4-stage reductions
// (:597-769). Here each step of pbicgstab/Solver.cpp:97-180 is one kernel
// that streams its vectors once and, where the step ends in a dot product,
// reduces in the same pass (wave shuffles, LDS, last-block grid sum).
#include "kernels.hpp"

namespace gph {

int vecGrid(int64_t n) {
  int64_t g = (n + kVecThreads * 4 - 1) / (kVecThreads * 4);
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return (int)g;
}

// x = (x + alpha*y) + omega*z ; r = s - omega*t ; then r.r and r0.r
// (pbicgstab/Solver.cpp:162-180, next rho :91).
template <typename X>
__global__ __launch_bounds__(kVecThreads) void updateXRKernel(VecArgs v,
                                                             const X* r0) {
  __shared__ X lds[2 * kVecThreads / 64];
  DevState* st = v.state;
  if (st->done) return;
  const X alpha = (X)st->alpha, omega = (X)st->omega;
  const X* y = static_cast<const X*>(v.a);
  const X* z = static_cast<const X*>(v.b);
  const X* s = static_cast<const X*>(v.c);
  const X* t = static_cast<const X*>(v.d);
  X* x = static_cast<X*>(v.o1);
  X* r = static_cast<X*>(v.o2);
  X acc[2] = {0, 0};
  // kU grid-stride elements per pass: all their loads are issued before
  // the first store (the stores may alias them as far as the compiler
  // knows); the sums take the elements in the same order as a plain
  // grid-stride loop
  constexpr int kU = 8;
  const int64_t stride = (int64_t)gridDim.x * kVecThreads;
  for (int64_t i0 = blockIdx.x * kVecThreads + threadIdx.x; i0 < v.n; i0 += kU * stride) {
    X xv[kU], yv[kU], zv[kU], sv[kU], tv[kU], wv[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int64_t i = i0 + u * stride;
      if (i < v.n) {
        xv[u] = x[i];
        yv[u] = y[i];
        zv[u] = z[i];
        sv[u] = s[i];
        tv[u] = t[i];
        wv[u] = r0[i];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int64_t i = i0 + u * stride;
      if (i < v.n) {
        x[i] = (xv[u] + alpha * yv[u]) + omega * zv[u];
        const X ri = sv[u] - omega * tv[u];
        r[i] = ri;
        acc[0] = acc[0] + ri * ri;
        acc[1] = acc[1] + wv[u] * ri;
      }
    }
  }
  X tot[2];
  if (gridReduce<X, 2, kVecThreads>(acc, v.partials,
                                    &st->counter[0], tot, lds)) {
    if (threadIdx.x == 0)
      finalize<X>(st, v.fin ? v.fin : kFinUpdate, (double)tot[0], (double)tot[1]);
  }
}

// p = r + beta*(p - omega*v) with beta = (rho/rhoOld)*(alpha/omega) when
// iter != 0 (pbicgstab/Solver.cpp:98-101); optional copy to o2 (no
// preconditioner: y = p, :111-113).
template <typename X>
__global__ __launch_bounds__(kVecThreads) void updatePKernel(VecArgs v) {
  DevState* st = v.state;
  if (st->done) return;
  const X* r = static_cast<const X*>(v.a);
  const X* vv = static_cast<const X*>(v.c);
  X* p = static_cast<X*>(v.o1);
  X* y = static_cast<X*>(v.o2);
  const bool upd = st->iter != 0;
  const X beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
  const X omega = (X)st->omega;
  for (int i = blockIdx.x * kVecThreads + threadIdx.x; i < v.n;
       i += gridDim.x * kVecThreads) {
    X pi = p[i];
    if (upd) {
      pi = r[i] + beta * (pi - omega * vv[i]);
      p[i] = pi;
    }
    if (y) y[i] = pi;
  }
}

// s = r - alpha*v (pbicgstab/Solver.cpp:128); optional copy to o2.
template <typename X>
__global__ __launch_bounds__(kVecThreads) void updateSKernel(VecArgs v) {
  DevState* st = v.state;
  if (st->done) return;
  const X alpha = (X)st->alpha;
  const X* r = static_cast<const X*>(v.a);
  const X* vv = static_cast<const X*>(v.c);
  X* s = static_cast<X*>(v.o1);
  X* z = static_cast<X*>(v.o2);
  for (int i = blockIdx.x * kVecThreads + threadIdx.x; i < v.n;
       i += gridDim.x * kVecThreads) {
    const X si = r[i] - alpha * vv[i];
    s[i] = si;
    if (z) z[i] = si;
  }
}

// Sum of squares in the operand type (Tensor::norm L2,
// lg/dsl/tensor/Tensor.cpp:425-439), result handed to finalize().
template <typename T>
__global__ __launch_bounds__(kVecThreads) void norm2Kernel(VecArgs v) {
  __shared__ T lds[kVecThreads / 64];
  DevState* st = v.state;
  if (v.checkDone && st->done) return;
  const T* a = static_cast<const T*>(v.a);
  T acc[1] = {0};
  for (int i = blockIdx.x * kVecThreads + threadIdx.x; i < v.n;
       i += gridDim.x * kVecThreads)
    acc[0] = acc[0] + a[i] * a[i];
  T tot[1];
  if (gridReduce<T, 1, kVecThreads>(acc, v.partials,
                                    &st->counter[0], tot, lds)) {
    if (threadIdx.x == 0) finalize<T>(st, v.fin, (double)tot[0], 0.0);
  }
}

// sum |x| (L1, OP = 0) or max |x| (LINF, OP = 1) in the operand type,
// handed to finalize() as the finished norm. The reference's L1 is the
// signed sum (Tensor.cpp:430-431) and its LINF reduction is not available
// (Expression.cpp:502-521); these are the textbook norms.
template <typename T, int OP>
__global__ __launch_bounds__(kVecThreads) void normKindKernel(VecArgs v) {
  __shared__ T lds[kVecThreads / 64];
  DevState* st = v.state;
  if (v.checkDone && st->done) return;
  const T* a = static_cast<const T*>(v.a);
  T acc[1] = {0};
  for (int i = blockIdx.x * kVecThreads + threadIdx.x; i < v.n;
       i += gridDim.x * kVecThreads)
    acc[0] = combine<OP>(acc[0], a[i] < (T)0 ? -a[i] : a[i]);
  T tot[1];
  if (gridReduce<T, 1, kVecThreads, OP>(acc, v.partials, &st->counter[0], tot, lds)) {
    if (threadIdx.x == 0) finalize<T>(st, v.fin, (double)tot[0], 0.0);
  }
}

template <typename T>
__global__ void fillKernel(T* x, int64_t n, T value) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = value;
}

template <typename I, typename O>
__global__ void castKernel(const I* in, O* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (O)in[i];
}
template <>
__global__ void castKernel<float, __half>(const float* in, __half* out,
                                          int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = __float2half_rn(in[i]);
}
template <>
__global__ void castKernel<double, __half>(const double* in, __half* out,
                                           int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = __float2half_rn((float)in[i]);
}

// phi = phi + ext(d) (iterative-refinement/Solver.cpp:103)
template <typename T>
__global__ void accumulateKernel(double* phi, const T* d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    phi[i] = phi[i] + (double)d[i];
}

static int elemGrid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (int)g;
}

void launchUpdateXR(const VecArgs& v, const void* r0, DType x, hipStream_t s) {
  const int g = vecGrid(v.n);
  if (x == DType::F32)
    hipLaunchKernelGGL(updateXRKernel<float>, dim3(g), dim3(kVecThreads), 0,
                       s, v, static_cast<const float*>(r0));
  else if (x == DType::F64)
    hipLaunchKernelGGL(updateXRKernel<double>, dim3(g), dim3(kVecThreads), 0,
                       s, v, static_cast<const double*>(r0));
  else
    fail(kUnsupported, "working type must be float32 or float64");
  GPH_HIP(hipGetLastError());
}

void launchUpdateP(const VecArgs& v, DType x, hipStream_t s) {
  const int g = vecGrid(v.n);
  if (x == DType::F32)
    hipLaunchKernelGGL(updatePKernel<float>, dim3(g), dim3(kVecThreads), 0, s,
                       v);
  else if (x == DType::F64)
    hipLaunchKernelGGL(updatePKernel<double>, dim3(g), dim3(kVecThreads), 0,
                       s, v);
  else
    fail(kUnsupported, "working type must be float32 or float64");
  GPH_HIP(hipGetLastError());
}

void launchUpdateS(const VecArgs& v, DType x, hipStream_t s) {
  const int g = vecGrid(v.n);
  if (x == DType::F32)
    hipLaunchKernelGGL(updateSKernel<float>, dim3(g), dim3(kVecThreads), 0, s,
                       v);
  else if (x == DType::F64)
    hipLaunchKernelGGL(updateSKernel<double>, dim3(g), dim3(kVecThreads), 0,
                       s, v);
  else
    fail(kUnsupported, "working type must be float32 or float64");
  GPH_HIP(hipGetLastError());
}

void launchNorm2(const VecArgs& v, DType t, hipStream_t s) {
  const int g = vecGrid(v.n);
  if (t == DType::F32)
    hipLaunchKernelGGL(norm2Kernel<float>, dim3(g), dim3(kVecThreads), 0, s,
                       v);
  else if (t == DType::F64)
    hipLaunchKernelGGL(norm2Kernel<double>, dim3(g), dim3(kVecThreads), 0, s,
                       v);
  else
    fail(kUnsupported, "norm: type must be float32 or float64");
  GPH_HIP(hipGetLastError());
}

void launchNormKind(const VecArgs& v, DType t, int kind, hipStream_t s) {
  if (kind == kNormL2) {
    fail(kInvalid, "launchNormKind: the L2 norm has its own kernel");
  }
  const int g = vecGrid(v.n);
  const bool mx = kind == kNormLinf;
  if (t == DType::F32) {
    if (mx)
      hipLaunchKernelGGL((normKindKernel<float, 1>), dim3(g), dim3(kVecThreads), 0, s, v);
    else
      hipLaunchKernelGGL((normKindKernel<float, 0>), dim3(g), dim3(kVecThreads), 0, s, v);
  } else if (t == DType::F64) {
    if (mx)
      hipLaunchKernelGGL((normKindKernel<double, 1>), dim3(g), dim3(kVecThreads), 0, s, v);
    else
      hipLaunchKernelGGL((normKindKernel<double, 0>), dim3(g), dim3(kVecThreads), 0, s, v);
  } else {
    fail(kUnsupported, "norm: type must be float32 or float64");
  }
  GPH_HIP(hipGetLastError());
}

void launchFill(void* x, DType t, int64_t n, double value, hipStream_t s) {
  if (n <= 0) return;
  const int g = elemGrid(n);
  if (t == DType::F32)
    hipLaunchKernelGGL(fillKernel<float>, dim3(g), dim3(256), 0, s,
                       static_cast<float*>(x), n, (float)value);
  else if (t == DType::F64)
    hipLaunchKernelGGL(fillKernel<double>, dim3(g), dim3(256), 0, s,
                       static_cast<double*>(x), n, value);
  else
    hipLaunchKernelGGL(fillKernel<__half>, dim3(g), dim3(256), 0, s,
                       static_cast<__half*>(x), n,
                       __float2half_rn((float)value));
  GPH_HIP(hipGetLastError());
}

// x_i = U[-1,1) from a counter-based hash of (seed, i): splitmix64 of
// seed + (i+1)*golden, mapped as ((u >> 40) * 2^-24) * 2 - 1.
__global__ void uniformKernel(double* x, int64_t n, uint64_t seed) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    x[i] = (double)((float)(z >> 40) * 5.9604644775390625e-8f) * 2.0 - 1.0;
  }
}

void launchUniform(double* x, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(uniformKernel, dim3(elemGrid(n)), dim3(256), 0, s, x, n, seed);
  GPH_HIP(hipGetLastError());
}

void launchCast(const void* in, DType it, void* out, DType ot, int64_t n,
                hipStream_t s) {
  if (launchCastDW(in, it, out, ot, n, s)) return;
  if (n <= 0) return;
  const int g = elemGrid(n);
  const DType H = DType::F16, F = DType::F32, D = DType::F64;
  if (it == F && ot == D)
    hipLaunchKernelGGL((castKernel<float, double>), dim3(g), dim3(256), 0, s,
                       static_cast<const float*>(in), static_cast<double*>(out), n);
  else if (it == D && ot == F)
    hipLaunchKernelGGL((castKernel<double, float>), dim3(g), dim3(256), 0, s,
                       static_cast<const double*>(in), static_cast<float*>(out), n);
  else if (it == F && ot == F)
    hipLaunchKernelGGL((castKernel<float, float>), dim3(g), dim3(256), 0, s,
                       static_cast<const float*>(in), static_cast<float*>(out), n);
  else if (it == D && ot == D)
    hipLaunchKernelGGL((castKernel<double, double>), dim3(g), dim3(256), 0, s,
                       static_cast<const double*>(in), static_cast<double*>(out), n);
  else if (it == F && ot == H)
    hipLaunchKernelGGL((castKernel<float, __half>), dim3(g), dim3(256), 0, s,
                       static_cast<const float*>(in), static_cast<__half*>(out), n);
  else if (it == D && ot == H)
    hipLaunchKernelGGL((castKernel<double, __half>), dim3(g), dim3(256), 0, s,
                       static_cast<const double*>(in), static_cast<__half*>(out), n);
  else
    fail(kUnsupported, "unsupported cast");
  GPH_HIP(hipGetLastError());
}

void launchAccumulate(double* phi, const void* d, DType dt, int64_t n,
                      hipStream_t s) {
  if (n <= 0) return;
  const int g = elemGrid(n);
  if (dt == DType::F32)
    hipLaunchKernelGGL(accumulateKernel<float>, dim3(g), dim3(256), 0, s, phi,
                       static_cast<const float*>(d), n);
  else if (dt == DType::F64)
    hipLaunchKernelGGL(accumulateKernel<double>, dim3(g), dim3(256), 0, s,
                       phi, static_cast<const double*>(d), n);
  else
    fail(kUnsupported, "accumulate: correction must be float32 or float64");
  GPH_HIP(hipGetLastError());
}

}  // namespace gph
