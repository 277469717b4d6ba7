// Double-word ("twofloat32") kernels of mixed-precision refinement (gfx950).
//
// With extendedPrecisionType "twofloat32" the refinement keeps the solution
// phi as double words and evaluates the residual in double-word arithmetic
// (iterative-refinement/Solver.cpp:79, :103, :107 with the residual of
// CRSMatrix.cpp:107-125 / LDUMatrix.cpp:130-159 instantiated for
// two<float>): per row sum = dw(diag) * phi_i, then sum += dw(a_ij) * phi_j
// for every stored entry in stored order (CRS: local column order; LDU: owned
// faces, then losort faces), r_i = fl32(dw(b_i) - sum). Arithmetic in
// dword.hpp. One thread per row, entries walked in the same streams as the
// SpMV kernels (sliced ELL or row blocks); the residual is evaluated once per
// outer step, a small part of a solve.
#include "dword.hpp"
#include "kernels.hpp"

namespace gph {

static constexpr int kDwThreads = 256;

// sum of one row in the sliced-ELL streams (slot = slice * 64 + lane)
__device__ __forceinline__ DWord dwRowSliced(const SpmvArgs& a, const DWord* __restrict__ x,
                                             int slice, int lane, int len, DWord sum) {
  const float* __restrict__ val = static_cast<const float*>(a.sval) + a.sliceOff[slice] + lane;
  const int cbase = a.sliceBase[slice];
  const uint16_t* __restrict__ cw = a.scol + a.sliceColW[slice];
  if (cbase >= 0) {
    const uint16_t* __restrict__ col = cw + lane;
    for (int k = 0; k < len; k++)
      sum = dwAdd(sum, dwMul(dwFrom(val[k * 64]), x[cbase + (int)col[k * 64]]));
  } else {
    const int32_t* __restrict__ col = reinterpret_cast<const int32_t*>(cw) + lane;
    for (int k = 0; k < len; k++) sum = dwAdd(sum, dwMul(dwFrom(val[k * 64]), x[col[k * 64]]));
  }
  return sum;
}

// sum of one row in the row streams (CRS: one stream; LDU: owned faces,
// then losort faces)
__device__ __forceinline__ DWord dwRowStreams(const SpmvArgs& a, const DWord* __restrict__ x,
                                              int row, DWord sum) {
  for (int s = 0; s < a.nstreams; s++) {
    const int32_t* __restrict__ col = a.col[s];
    const float* __restrict__ val = static_cast<const float*>(a.val[s]);
    for (int e = a.ptr[s][row]; e < a.ptr[s][row + 1]; e++)
      sum = dwAdd(sum, dwMul(dwFrom(val[e]), x[col[e]]));
  }
  return sum;
}

// r = fl32(b - A phi); out64 (optional) = b - A phi rounded to float64;
// RED: sum of the squared float64 residuals, finalised with a.fin
template <bool RED>
__global__ __launch_bounds__(kDwThreads) void dwResidualKernel(SpmvArgs a) {
  __shared__ double redLds[kDwThreads / 64];
  const DWord* __restrict__ x = static_cast<const DWord*>(a.x);
  const float* __restrict__ b = static_cast<const float*>(a.b);
  const float* __restrict__ diag = static_cast<const float*>(a.diag);
  float* __restrict__ r = static_cast<float*>(a.out);
  const bool sliced = a.numSlices > 0;
  const int64_t slots = sliced ? (int64_t)a.numSlices * 64 : (int64_t)a.n;
  double acc[1] = {0.0};
  for (int64_t q = blockIdx.x * (int64_t)kDwThreads + threadIdx.x; q < slots;
       q += (int64_t)gridDim.x * kDwThreads) {
    int row = (int)q, len = 0;
    if (sliced) {
      row = a.perm ? a.perm[q] : (int)q;
      len = a.rowLen[q];
      if (row >= a.n) continue;
    }
    DWord sum = dwMul(dwFrom(diag[row]), x[row]);
    sum = sliced ? dwRowSliced(a, x, (int)(q >> 6), (int)(q & 63), len, sum)
                 : dwRowStreams(a, x, row, sum);
    const DWord res = dwSub(dwFrom(b[row]), sum);
    r[row] = dwEval(res);
    const double rd = dwToDouble(res);
    if (a.out64) a.out64[row] = rd;
    if (RED) acc[0] = acc[0] + rd * rd;
  }
  if constexpr (RED) {
    double tot[1];
    if (gridReduce<double, 1, kDwThreads>(acc, a.partials, &a.state->counter[0], tot, redLds)) {
      if (threadIdx.x == 0) finalize<double>(a.state, a.fin, tot[0], 0.0);
    }
  }
}

static int dwGrid(int64_t work) {
  int64_t g = (work + kDwThreads - 1) / kDwThreads;
  if (g < 1) g = 1;
  if (g > 2048) g = 2048;
  return (int)g;
}

void launchResidualDW(const SpmvArgs& a, hipStream_t s) {
  if (a.n <= 0) return;
  const int64_t work = a.numSlices > 0 ? (int64_t)a.numSlices * 64 : (int64_t)a.n;
  const int g = dwGrid(work);
  if (a.state && a.fin != kFinNone)
    hipLaunchKernelGGL(dwResidualKernel<true>, dim3(g), dim3(kDwThreads), 0, s, a);
  else
    hipLaunchKernelGGL(dwResidualKernel<false>, dim3(g), dim3(kDwThreads), 0, s, a);
  GPH_HIP(hipGetLastError());
}

// phi = phi + dw(d) (iterative-refinement/Solver.cpp:103: the correction is
// cast to the extended type, then added in it)
__global__ void dwAccumulateKernel(DWord* phi, const float* d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    phi[i] = dwAdd(phi[i], dwFrom(d[i]));
}

void launchAccumulateDW(void* phi, const void* d, DType dt, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  if (dt != DType::F32) fail(kUnsupported, "twofloat32 refinement: the correction must be float32");
  hipLaunchKernelGGL(dwAccumulateKernel, dim3(dwGrid(n)), dim3(kDwThreads), 0, s,
                     static_cast<DWord*>(phi), static_cast<const float*>(d), n);
  GPH_HIP(hipGetLastError());
}

template <typename I>
__global__ void toDWKernel(const I* in, DWord* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = dwFrom(in[i]);
}
__global__ void fromDWKernel(const DWord* in, float* out32, double* out64, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (out32) out32[i] = dwEval(in[i]);
    if (out64) out64[i] = dwToDouble(in[i]);
  }
}

// casts to and from double words (Tensor::cast with twofloat32)
bool launchCastDW(const void* in, DType it, void* out, DType ot, int64_t n, hipStream_t s) {
  if (it != DType::TF32 && ot != DType::TF32) return false;
  if (n <= 0) return true;
  const int g = dwGrid(n);
  if (it == DType::F32 && ot == DType::TF32)
    hipLaunchKernelGGL(toDWKernel<float>, dim3(g), dim3(kDwThreads), 0, s,
                       static_cast<const float*>(in), static_cast<DWord*>(out), n);
  else if (it == DType::F64 && ot == DType::TF32)
    hipLaunchKernelGGL(toDWKernel<double>, dim3(g), dim3(kDwThreads), 0, s,
                       static_cast<const double*>(in), static_cast<DWord*>(out), n);
  else if (it == DType::TF32 && ot == DType::TF32)
    GPH_HIP(hipMemcpyAsync(out, in, (size_t)n * sizeof(DWord), hipMemcpyDeviceToDevice, s));
  else if (it == DType::TF32 && ot == DType::F32)
    hipLaunchKernelGGL(fromDWKernel, dim3(g), dim3(kDwThreads), 0, s,
                       static_cast<const DWord*>(in), static_cast<float*>(out),
                       static_cast<double*>(nullptr), n);
  else if (it == DType::TF32 && ot == DType::F64)
    hipLaunchKernelGGL(fromDWKernel, dim3(g), dim3(kDwThreads), 0, s,
                       static_cast<const DWord*>(in), static_cast<float*>(nullptr),
                       static_cast<double*>(out), n);
  else
    fail(kUnsupported, std::string("unsupported cast: ") + dtypeName(it) + " to " + dtypeName(ot));
  GPH_HIP(hipGetLastError());
  return true;
}

}  // namespace gph
