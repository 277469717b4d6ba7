// Gauss-Seidel sweep over the jagged-diagonal level stream (gfx950).
//
// Same arithmetic as the reference's sweep (lg/matrix/solver/gauss-seidel/
// Solver.cpp:45-86, codelet lg/codelet/ipu/matrix/solvers/gauss-seidel.cpp:41-61):
// per row x_i = (b_i - sum_j a_ij x_j) / a_ii over the row's entries in
// stored order, x_j the value this sweep has already written for in-tile
// j < i (forward) or j > i (backward), the previous x otherwise (halo
// included). One workgroup per tile walks the tile's levels with a barrier
// per level; its waves take groups of 64 of the level's rows (sorted by
// length; stream format in gs_pack.hpp), lane = row, so the entries a wave
// loads for one entry index are adjacent, and a group runs as many entry
// steps as its longest row. A row runs in steps of 16 entries whose codes
// and values are loaded one step ahead, so a step waits on one round of x
// reads. The sweep's own x values sit in LDS; the result is written to xout
// at the end (xout != xin).
#include "kernels.hpp"

namespace gph {

static size_t gsMaxShared() { return 160 * 1024 - 4096; }

size_t gsStreamSharedBytes(const GsArgs& a, DType xt) {
  return (size_t)(a.maxRows + 1) * dtypeSize(xt);
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
  // consecutive tiles on one XCD: their x windows overlap in its L2
  const int tile = xcdRemap(blockIdx.x, gridDim.x);
  const int t0 = a.tileStart[tile], n = a.tileStart[tile + 1] - t0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int nw = kGsThreads / 64;
  const float* __restrict__ val = a.val;
  const uint32_t* __restrict__ code = a.code;
  const int32_t* __restrict__ dg = a.diagStart;
  const int l0 = a.lvlBase[tile], nL = a.lvlBase[tile + 1] - l0;
  auto xOf = [&](uint32_t c) { return (c & kGsFreshBit) ? xs[c & ~kGsFreshBit] : xin[c]; };
  for (int L = 0; L < nL; L++) {
    const int4 lv = reinterpret_cast<const int4*>(a.lvl)[l0 + L];
    for (int g0 = 64 * wave; g0 < lv.y; g0 += 64 * nw) {
      const bool own = g0 + lane < lv.y;
      const uint32_t sl = a.slot[lv.x + g0 + (own ? lane : 0)];
      const int len = own ? (int)(sl & 255u) : 0;
      const int r = (int)(sl >> 8);
      const int W = __builtin_amdgcn_readfirstlane(len);  // the group's longest row
      const int i = t0 + r;
      X s = b[i];
      // entry j of the lane's row: diagonal j's start + the row's slot;
      // lanes whose row is shorter (or past the level) read the diagonal's
      // first entry and do not subtract it
      constexpr int S = 16;
      uint32_t cA[S], cB[S];
      float vA[S], vB[S];
      auto ld = [&](int j0, uint32_t (&cd)[S], float (&v)[S]) {
#pragma unroll
        for (int k = 0; k < S; k++) {
          const int j = j0 + k;
          const uint32_t d = (uint32_t)dg[lv.z + min(j, W - 1)];
          const uint32_t e = j < len ? d + (uint32_t)(g0 + lane) : d;
          cd[k] = code[e];
          v[k] = val[e];
        }
      };
      auto step = [&](int j0, const uint32_t (&cd)[S], const float (&v)[S]) {
        X xv[S];
#pragma unroll
        for (int k = 0; k < S; k++) xv[k] = xOf(cd[k]);
#pragma unroll
        for (int k = 0; k < S; k++) {
          const X p = cvt<X>(v[k]) * xv[k];
          s = j0 + k < len ? s - p : s;
        }
      };
      if (W > 0) {
        ld(0, cA, vA);
        for (int j = 0; j < W; j += 2 * S) {
          ld(j + S, cB, vB);
          step(j, cA, vA);
          if (j + S >= W) break;
          ld(j + 2 * S, cA, vA);
          step(j + S, cB, vB);
        }
      }
      if (own) xs[r] = s / cvt<X>(diag[i]);
    }
    __syncthreads();
  }
  for (int q = threadIdx.x; q < n; q += kGsThreads) xout[t0 + q] = xs[q];
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
