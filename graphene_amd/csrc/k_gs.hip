// Gauss-Seidel sweep over the jagged-diagonal level stream (gfx950).
//
// Same arithmetic as the reference's sweep (lg/matrix/solver/gauss-seidel/
// Solver.cpp:45-86, codelet lg/codelet/ipu/matrix/solvers/gauss-seidel.cpp:41-61):
// per row x_i = (b_i - sum_j a_ij x_j) / a_ii over the row's entries in
// stored order, x_j the value this sweep has already written for in-tile
// j < i (forward) or j > i (backward), the previous x otherwise (halo
// included). One workgroup per tile, the values this sweep writes in LDS.
// Per level (stream format in gs_pack.hpp) two phases: every thread takes
// consecutive entries of the level's stream (in gather order: previous-x
// columns ascending) and stores its product a_ij * x_j at the entry's
// position in LDS (x_j from LDS or the previous x; the stream loads are
// coalesced, the gathers of neighbouring lanes share cache lines, and no
// lane idles on a short row); then one thread per row
// subtracts its products from b_i in the row's stored order and writes
// x_i. The result is written to xout at the end (xout != xin).
#include "kernels.hpp"

namespace gph {

static size_t gsMaxShared() { return 160 * 1024 - 4096; }

// entries per thread and step of a level's product phase
constexpr int kGsStep = 4;

// LDS: the x values this sweep writes, the level's products, the level's
// diagonal starts (kGsMaxWidth + 1 ints)
size_t gsStreamSharedBytes(const GsArgs& a, DType xt) {
  return (size_t)(a.maxRows + a.maxLevelEntries) * dtypeSize(xt) + (kGsMaxWidth + 1) * 4;
}

template <typename X>
__global__ __launch_bounds__(kGsThreads) void gsChunkKernel(GsArgs a,
                                                            const float* __restrict__ diag,
                                                            const X* __restrict__ b,
                                                            const X* __restrict__ xin,
                                                            X* __restrict__ xout, DevState* st,
                                                            int checkDone) {
  extern __shared__ __align__(16) unsigned char smemRaw[];
  if (checkDone && st->done) return;
  X* xs = reinterpret_cast<X*>(smemRaw);
  X* prod = xs + a.maxRows;
  int* doff = reinterpret_cast<int*>(prod + a.maxLevelEntries);
  // consecutive tiles on one XCD: their x windows overlap in its L2
  const int tile = xcdRemap(blockIdx.x, gridDim.x);
  const int t0 = a.tileStart[tile], n = a.tileStart[tile + 1] - t0;
  const int tid = threadIdx.x;
  const float* __restrict__ val = a.val;
  const uint32_t* __restrict__ code = a.code;
  const uint16_t* __restrict__ pos = a.pos;
  const int32_t* __restrict__ dg = a.diagStart;
  const int l0 = a.lvlBase[tile], nL = a.lvlBase[tile + 1] - l0;
  for (int L = 0; L < nL; L++) {
    const int4 lv = reinterpret_cast<const int4*>(a.lvl)[l0 + L];
    const int e0 = dg[lv.z], e1 = dg[lv.z + lv.w];
    // the first row of this thread (loads independent of the products)
    uint32_t sl = 0;
    X bi = 0, di = 1;
    if (tid < lv.y) {
      sl = a.slot[lv.x + tid];
      const int i = t0 + (int)(sl >> 8);
      bi = b[i];
      di = cvt<X>(diag[i]);
    }
    if (tid <= lv.w) doff[tid] = dg[lv.z + tid] - e0;
    // kGsStep entries per thread and step: the code and value loads, the x
    // reads, then the products (indices past the level read its last entry)
    for (int e = e0 + tid; e < e1; e += kGsStep * kGsThreads) {
      uint32_t c[kGsStep];
      float v[kGsStep];
      int ps[kGsStep];
#pragma unroll
      for (int k = 0; k < kGsStep; k++) {
        const int ek = min(e + k * kGsThreads, e1 - 1);
        c[k] = code[ek];
        v[k] = val[ek];
        ps[k] = pos[ek];
      }
      X xv[kGsStep];
#pragma unroll
      for (int k = 0; k < kGsStep; k++)
        xv[k] = (c[k] & kGsFreshBit) ? xs[c[k] & ~kGsFreshBit] : xin[c[k]];
#pragma unroll
      for (int k = 0; k < kGsStep; k++)
        if (e + k * kGsThreads < e1) prod[ps[k]] = cvt<X>(v[k]) * xv[k];
    }
    __syncthreads();
    for (int q = tid; q < lv.y; q += kGsThreads) {
      if (q != tid) {
        sl = a.slot[lv.x + q];
        const int i = t0 + (int)(sl >> 8);
        bi = b[i];
        di = cvt<X>(diag[i]);
      }
      const int len = (int)(sl & 255u);
      X s = bi;
      // 8 products per step, read together, subtracted in order
      for (int j = 0; j < len; j += 8) {
        X p[8];
#pragma unroll
        for (int k = 0; k < 8; k++) p[k] = prod[doff[min(j + k, len - 1)] + q];
#pragma unroll
        for (int k = 0; k < 8; k++) s = j + k < len ? s - p[k] : s;
      }
      xs[sl >> 8] = s / di;
    }
    __syncthreads();
  }
  for (int q = tid; q < n; q += kGsThreads) xout[t0 + q] = xs[q];
}

void launchGsStream(const GsArgs& a, const void* diag, const void* b, const void* xin,
                    void* xout, DType xt, DevState* st, int checkDone, hipStream_t s) {
  if (a.T <= 0) return;
  const size_t shm = gsStreamSharedBytes(a, xt);
  if (shm > gsMaxShared()) fail(kUnsupported, "Gauss-Seidel stream kernel: tile too large for LDS");
  auto launch = [&](auto kern, auto* xtyped) {
    using XT = std::remove_pointer_t<decltype(xtyped)>;
    static bool attr = false;
    if (!attr) {
      GPH_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)gsMaxShared()));
      attr = true;
    }
    hipLaunchKernelGGL(kern, dim3(a.T), dim3(kGsThreads), shm, s, a,
                       static_cast<const float*>(diag),
                       static_cast<const XT*>(b), static_cast<const XT*>(xin),
                       static_cast<XT*>(xout), st, checkDone);
  };
  if (xt == DType::F32)
    launch(gsChunkKernel<float>, (float*)nullptr);
  else if (xt == DType::F64)
    launch(gsChunkKernel<double>, (double*)nullptr);
  else
    fail(kUnsupported, "Gauss-Seidel vectors must be float32 or float64");
  GPH_HIP(hipGetLastError());
}

}  // namespace gph
