// Sliced-ELL SpMV for K right-hand sides at once (gfx950): y_j = A x_j,
// j < K, with the matrix streams read once for all K vectors.
//
// Same layout and arithmetic per vector as the single-vector sliced kernel
// (k_spmv.hip, sellKernel; CRSMatrix::spmv, lg/matrix/details/crs/CRSMatrix.cpp:62-80):
// one wave per 64-row slice, lane = row; per row sum = diag*x, then the
// stored entries in order, no FMA contraction. Each entry's column and
// value are loaded once and multiplied with the K gathered x values, so a
// vector's result is bit-identical to the single-vector kernel's.
// Vectors are K blocks of `stride` elements (x_j = x + j*stride). A vector
// whose solver state in `states` is finished is skipped (its output is left
// as it is).
#include "kernels.hpp"

namespace gph {

template <int K>
__global__ __launch_bounds__(kSpmvThreads) void sellMultiKernel(SpmvArgs a, int64_t stride,
                                                                const DevState* states,
                                                                const DevState* glob) {
  if (glob && glob->done) return;
  const int lb = xcdRemap(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int W = kSpmvThreads / 64;
  const int groups = (a.numSlices + W - 1) / W;
  const int per = (groups + gridDim.x - 1) / gridDim.x;
  const int g0 = lb * per, g1 = min(groups, g0 + per);
  const float* __restrict__ x = static_cast<const float*>(a.x);
  float* __restrict__ y = static_cast<float*>(a.out);
  bool act[K];
#pragma unroll
  for (int j = 0; j < K; j++) act[j] = !(states && states[j].done);
  for (int g = g0; g < g1; g++) {
    const int slice = g * W + wave;
    if (slice >= a.numSlices) break;
    const int slot = slice * 64 + lane;
    const int row = a.perm ? a.perm[slot] : slot;
    const int len = a.rowLen[slot];
    if (row >= a.n) continue;
    const float d = static_cast<const float*>(a.diag)[row];
    float sum[K];
#pragma unroll
    for (int j = 0; j < K; j++) sum[j] = act[j] ? d * x[j * stride + row] : 0.0f;
    const float* __restrict__ val = static_cast<const float*>(a.sval) + a.sliceOff[slice] + lane;
    const int cbase = a.sliceBase[slice];
    const uint16_t* __restrict__ cw = a.scol + a.sliceColW[slice];
    auto rowSum = [&](auto* __restrict__ col, int cb) {
      int k = 0;
      for (; k + 4 <= len; k += 4) {
        int c[4];
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
          c[e] = cb + (int)streamLoad(&col[(k + e) * 64]);
          v[e] = streamLoad(&val[(k + e) * 64]);
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
          if (!act[j]) continue;
          float xv[4];
#pragma unroll
          for (int e = 0; e < 4; e++) xv[e] = x[j * stride + c[e]];
#pragma unroll
          for (int e = 0; e < 4; e++) sum[j] = sum[j] + v[e] * xv[e];
        }
      }
      for (; k < len; k++) {
        const int c = cb + (int)streamLoad(&col[k * 64]);
        const float v = streamLoad(&val[k * 64]);
#pragma unroll
        for (int j = 0; j < K; j++)
          if (act[j]) sum[j] = sum[j] + v * x[j * stride + c];
      }
    };
    if (cbase >= 0)
      rowSum(cw + lane, cbase);  // 16-bit offsets from the slice base
    else
      rowSum(reinterpret_cast<const int32_t*>(cw) + lane, 0);
#pragma unroll
    for (int j = 0; j < K; j++)
      if (act[j]) __builtin_nontemporal_store(sum[j], y + j * stride + row);
  }
}

// blocks of the kernel resident at once on the current device
template <int K>
static int multiResident() {
  int dev = 0, cus = 0, per = 0;
  GPH_HIP(hipGetDevice(&dev));
  GPH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  GPH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, sellMultiKernel<K>, kSpmvThreads,
                                                       0));
  return cus * (per < 1 ? 1 : per);
}

template <int K>
static void launchMultiK(const SpmvArgs& a, int64_t stride, const DevState* states,
                         const DevState* glob, hipStream_t s) {
  static const int resident = multiResident<K>();
  int grid = resident;
  const int groups = (a.numSlices + kSpmvThreads / 64 - 1) / (kSpmvThreads / 64);
  if (grid > groups) grid = groups;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(sellMultiKernel<K>, dim3(grid), dim3(kSpmvThreads), 0, s, a, stride,
                     states, glob);
  GPH_HIP(hipGetLastError());
}

void launchSpmvMulti(const SpmvArgs& a, int k, int64_t stride, const DevState* states,
                     const DevState* glob, hipStream_t s) {
  if (a.numSlices <= 0) fail(kInvalid, "multi-vector SpMV needs the sliced layout");
  switch (k) {
    case 1: return launchMultiK<1>(a, stride, states, glob, s);
    case 2: return launchMultiK<2>(a, stride, states, glob, s);
    case 3: return launchMultiK<3>(a, stride, states, glob, s);
    case 4: return launchMultiK<4>(a, stride, states, glob, s);
    case 5: return launchMultiK<5>(a, stride, states, glob, s);
    case 6: return launchMultiK<6>(a, stride, states, glob, s);
    case 7: return launchMultiK<7>(a, stride, states, glob, s);
    case 8: return launchMultiK<8>(a, stride, states, glob, s);
    default:
      fail(kInvalid, "multi-vector SpMV: 1 to 8 vectors, not " + std::to_string(k));
  }
}

}  // namespace gph
