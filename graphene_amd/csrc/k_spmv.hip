// Row-gather kernels for gfx950: CRS and LDU SpMV and residual.
//
// Reference kernels: CRSMatrix::spmv / residual (lg/matrix/details/crs/CRSMatrix.cpp:62-80,
// :107-125) and LDUMatrix::spmv / residual (lg/matrix/details/ldu/LDUMatrix.cpp:62-99,
// :126-165). Per row: sum = diag*x, then every stored entry in order
// (CRS: local column order; LDU: owned faces, then losort faces), no FMA
// contraction (the library is compiled with -ffp-contract=off), so results
// match a sequential evaluation bit for bit.
//
// Layout per block (256 threads, one row per thread, <= 256 rows): the
// block's contiguous range of stream entries is processed in chunks of
// kSpmvCap; threads stream column indices and values with coalesced loads,
// gather x, and stage the products in LDS; each thread then adds the
// products of its own row in stored order. Blocks are remapped so that each
// XCD sweeps one contiguous row range (x window stays in that XCD's L2).
// Optional fused epilogues: residual, extra output copies, dot products
// with a last-block grid reduction that updates the solver state.
#include "kernels.hpp"

namespace gph {

template <typename AT, int RED>
struct RedAcc {
  static constexpr int K = (RED == kRedSelfW) ? 2 : 1;
  AT v[K];
  __device__ RedAcc() {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = (AT)0;
  }
};

// Row inputs of the epilogue: b[row] (residual) and w[row] (dot partner).
template <typename XT, typename BT, int MODE, int RED>
struct RowIn {
  BT b{};
  XT w{};
  __device__ __forceinline__ void load(const SpmvArgs& a, int row, bool has) {
    if constexpr (MODE == kModeResidual) {
      if (has) b = static_cast<const BT*>(a.b)[row];
    }
    if constexpr (RED == kRedW || RED == kRedSelfW) {
      if (has) w = static_cast<const XT*>(a.w)[row];
    }
  }
};

template <typename T, bool NT>
__device__ __forceinline__ void storeOut(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// Row result: residual, stores (non-temporal with NT), dot-product
// contributions.
template <typename XT, typename BT, typename OT, typename AT, int MODE, int RED,
          bool NT = false>
__device__ __forceinline__ void spmvRowOut(const SpmvArgs& a, int row, bool has,
                                           AT sum, const RowIn<XT, BT, MODE, RED>& in,
                                           RedAcc<AT, RED>& acc) {
  AT o = sum;
  if constexpr (MODE == kModeResidual) {
    if (has) o = (AT)in.b - sum;
  }
  if (has) {
    const OT ov = (OT)o;
    storeOut<OT, NT>(static_cast<OT*>(a.out) + row, ov);
    if (a.out2) storeOut<OT, NT>(static_cast<OT*>(a.out2) + row, ov);
    if (a.out3) storeOut<OT, NT>(static_cast<OT*>(a.out3) + row, ov);
    if (a.out64) storeOut<double, NT>(a.out64 + row, (double)o);
    if constexpr (RED != kRedNone) {
      AT wv = 0;
      if constexpr (RED == kRedW || RED == kRedSelfW) wv = (AT)in.w;
      if constexpr (RED == kRedSelf) acc.v[0] = acc.v[0] + o * o;
      if constexpr (RED == kRedW) acc.v[0] = acc.v[0] + o * wv;
      if constexpr (RED == kRedSelfW) {
        acc.v[0] = acc.v[0] + o * o;
        acc.v[1] = acc.v[1] + o * wv;
      }
    }
  }
}

// Block reduction of the dot contributions + last-block finalisation.
template <typename AT, int RED, int NT>
__device__ __forceinline__ void spmvFinish(const SpmvArgs& a,
                                           RedAcc<AT, RED>& acc, AT* redLds) {
  if constexpr (RED != kRedNone) {
    constexpr int K = RedAcc<AT, RED>::K;
    AT tot[K];
    if (gridReduce<AT, K, NT>(acc.v, a.partials,
                              &a.state->counter[0], tot, redLds)) {
      if (threadIdx.x == 0)
        finalize<AT>(a.state, a.fin, (double)tot[0],
                     K > 1 ? (double)tot[1] : 0.0);
    }
  }
}

template <int S, typename VT, typename XT, typename BT, typename OT,
          typename AT, int MODE, int RED>
__global__ __launch_bounds__(kSpmvThreads) void spmvKernel(SpmvArgs a) {
  __shared__ AT prod[kSpmvCap];
  __shared__ AT redLds[2 * kSpmvThreads / 64];
  if (a.checkDone && a.state->done) return;
  const int lb = a.remap ? xcdRemap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int r0 = a.blockRow[lb], r1 = a.blockRow[lb + 1];
  const int tid = threadIdx.x;
  const int row = r0 + tid;
  const bool has = row < r1;
  const XT* __restrict__ x = static_cast<const XT*>(a.x);

  AT sum = 0;
  if (has) sum = cvt<AT>(static_cast<const VT*>(a.diag)[row]) * (AT)x[row];

#pragma unroll
  for (int s = 0; s < S; s++) {
    const int32_t* __restrict__ ptr = a.ptr[s];
    // column-ordered chunks (one stream): products go to their entry's
    // position in the chunk
    const bool srt = S == 1 && a.gcol != nullptr;
    const int32_t* __restrict__ col = srt ? a.gcol : a.col[s];
    const VT* __restrict__ val = static_cast<const VT*>(srt ? a.gval : a.val[s]);
    const uint16_t* __restrict__ pos = a.gpos;
    const int e0 = ptr[r0], e1 = ptr[r1];
    const int mb = has ? ptr[row] : 0;
    const int me = has ? ptr[row + 1] : 0;
    for (int c0 = e0; c0 < e1; c0 += kSpmvCap) {
      const int cn = min(kSpmvCap, e1 - c0);
      constexpr int U = kSpmvCap / kSpmvThreads;
      int ci[U], pk[U];
      VT vv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = tid + u * kSpmvThreads;
        pk[u] = k;
        if (k < cn) {
          ci[u] = col[c0 + k];
          vv[u] = val[c0 + k];
          if (srt) pk[u] = pos[c0 + k];
        }
      }
      XT xv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = tid + u * kSpmvThreads;
        if (k < cn) xv[u] = x[ci[u]];
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = tid + u * kSpmvThreads;
        if (k < cn) prod[pk[u]] = cvt<AT>(vv[u]) * (AT)xv[u];
      }
      __syncthreads();
      const int lo = max(mb, c0), hi = min(me, c0 + cn);
      for (int k = lo; k < hi; k++) sum = sum + prod[k - c0];
      __syncthreads();
    }
  }

  RedAcc<AT, RED> acc;
  RowIn<XT, BT, MODE, RED> in;
  in.load(a, row, has);
  spmvRowOut<XT, BT, OT, AT, MODE, RED>(a, row, has, sum, in, acc);
  spmvFinish<AT, RED, kSpmvThreads>(a, acc, redLds);
}

// Sliced ELL: one wave per 64-row slice; lane l owns row perm[s*64+l]
// (or s*64+l) and walks its entries k = 0..len-1 at slot
// sliceOff[s] + k*64 + l, so every wave-level value load is one contiguous
// 256-byte (4-byte types) segment. Column indices are 16-bit offsets from a
// per-slice base where the slice's columns span less than 2^16 (2 bytes per
// entry instead of 4), 32-bit columns otherwise. Per row the order is the
// stored order, as in the row-block kernel.
//
// A wave takes its slices in chunks of CH and stores a chunk's results
// after the chunk (rows, sums and epilogue inputs stay in
// registers): a store counts against the wave's memory counter until it is
// acknowledged, so a store issued at the end of every slice would delay
// the next slice's first waits. Results are stored non-temporally (the
// output is not read again by this kernel).
// Chunk length: 4 slices, 2 where the held epilogue state (FP64 sums,
// residual and dot inputs) would cost occupancy.
template <typename AT, int MODE, int RED>
constexpr int sellChunk() {
  return (sizeof(AT) == 8 || MODE == kModeResidual || RED != kRedNone) ? 2 : 4;
}

// SKIP: batched solves, slices whose systems have all finished are skipped
// (a separate instantiation: the check's loads in the slice loop would make
// the compiler's load waits conservative for every launch)
template <typename VT, typename XT, typename BT, typename OT, typename AT,
          int MODE, int RED, bool SKIP>
__global__ __launch_bounds__(kSpmvThreads) void sellKernel(SpmvArgs a) {
  __shared__ AT redLds[2 * kSpmvThreads / 64];
  if (a.checkDone && a.state->done) return;
  // batched refinement: the residual of the systems whose outer step is due
  // (phase 1) is skipped at once when none is (the batch state's iter
  // counts them)
  if (SKIP && a.runPhase == 1 && a.state && a.state->iter == 0) return;
  // each block owns one contiguous range of slice groups (4 slices = 4
  // waves per group); blocks are remapped so an XCD sweeps adjacent ranges
  const int lb = a.remap ? xcdRemap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int W = kSpmvThreads / 64;
  const int groups = (a.numSlices + W - 1) / W;
  const int per = (groups + gridDim.x - 1) / gridDim.x;
  const int g0 = lb * per, g1 = min(groups, g0 + per);
  const XT* __restrict__ x = static_cast<const XT*>(a.x);
  using In = RowIn<XT, BT, MODE, RED>;
  RedAcc<AT, RED> acc;
  constexpr int CH = sellChunk<AT, MODE, RED>();
  // SKIP: the block's groups in rounds of 64 (256 slices); each thread
  // first tests one slice of the round (are any of its systems in the run
  // phase), all at once, and a round without such a slice is passed over
  __shared__ uint8_t runLds[SKIP ? kSpmvThreads : 1];
  const int rg = SKIP ? kSpmvThreads / W : g1 - g0;
  for (int r0 = g0; r0 < g1; r0 += rg) {
  const int r1 = min(g1, r0 + rg);
  if constexpr (SKIP) {
    const int sl = r0 * W + (int)threadIdx.x;
    bool run = false;
    if (sl < r1 * W && sl < a.numSlices)
      for (int k = a.sliceSys[2 * sl]; k <= a.sliceSys[2 * sl + 1]; k++)
        run = run || a.sysPhase[k] == a.runPhase;
    runLds[threadIdx.x] = run ? 1 : 0;
    if (!__syncthreads_or(run ? 1 : 0)) continue;
  }
  for (int c0 = r0; c0 < r1; c0 += CH) {
    int rows[CH];
    bool hass[CH];
    AT sums[CH];
    In ins[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) {
      const int g = c0 + i;
      const int slice = g * W + wave;
      bool sliceOk = g < r1 && slice < a.numSlices;
      if (SKIP && sliceOk) sliceOk = runLds[slice - r0 * W] != 0;
      const int slot = slice * 64 + lane;
      int row = 0, len = 0;
      if (sliceOk) {
        row = a.perm ? a.perm[slot] : slot;
        len = a.rowLen[slot];
      }
      const bool has = sliceOk && row < a.n;
      ins[i].load(a, row, has);
      AT sum = 0;
      if (has) {
        sum = cvt<AT>(static_cast<const VT*>(a.diag)[row]) * (AT)x[row];
        const VT* __restrict__ val = static_cast<const VT*>(a.sval) + a.sliceOff[slice] + lane;
        const int cbase = a.sliceBase[slice];
        const uint16_t* __restrict__ cw = a.scol + a.sliceColW[slice];
        // 8 entries per step: 16 independent stream loads, then 8 gathers,
        // then the adds in stored order; the rest of the row in steps of 4,
        // 2 and 1 (one round of loads each); stream loads are non-temporal
        auto rowSum = [&](auto* __restrict__ col, int cb) {
          int k = 0;
          auto batch = [&](auto nTag) {
            constexpr int B = decltype(nTag)::value;
            int c[B];
            VT v[B];
#pragma unroll
            for (int j = 0; j < B; j++) {
              c[j] = cb + (int)streamLoad(&col[(k + j) * 64]);
              v[j] = streamLoad(&val[(k + j) * 64]);
            }
            XT xv[B];
#pragma unroll
            for (int j = 0; j < B; j++) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < B; j++) sum = sum + cvt<AT>(v[j]) * (AT)xv[j];
            k += B;
          };
          while (k + 8 <= len) batch(std::integral_constant<int, 8>{});
          if (k + 4 <= len) batch(std::integral_constant<int, 4>{});
          if (k + 2 <= len) batch(std::integral_constant<int, 2>{});
          if (k < len) batch(std::integral_constant<int, 1>{});
        };
        if (cbase >= 0)
          rowSum(cw + lane, cbase);  // 16-bit offsets from the slice base
        else
          rowSum(reinterpret_cast<const int32_t*>(cw) + lane, 0);
      }
      rows[i] = row;
      hass[i] = has;
      sums[i] = sum;
    }
#pragma unroll
    for (int i = 0; i < CH; i++)
      spmvRowOut<XT, BT, OT, AT, MODE, RED, true>(a, rows[i], hass[i], sums[i], ins[i], acc);
  }
  if constexpr (SKIP) __syncthreads();  // the round's flags are read
  }
  spmvFinish<AT, RED, kSpmvThreads>(a, acc, redLds);
}

// Blocks of a kernel that are resident at once on the current device (the
// sliced kernel's grid: one round of blocks, each with an equal share).
template <typename K>
static int residentBlocks(K kernel) {
  int dev = 0, cus = 0, per = 0;
  GPH_HIP(hipGetDevice(&dev));
  GPH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  GPH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kSpmvThreads, 0));
  const int n = cus * per;
  return n < 1 ? 1 : (n > kSellMaxGrid ? kSellMaxGrid : n);
}

template <int S, typename VT, typename XT, typename BT, typename OT,
          typename AT, int MODE, int RED>
static void launchOne(const SpmvArgs& a, hipStream_t s) {
  if (S == 0) {
    if (a.numSlices <= 0) return;
    const int groups = (a.numSlices + kSpmvThreads / 64 - 1) / (kSpmvThreads / 64);
    if (a.sliceSys) {
      static const int grid = residentBlocks(sellKernel<VT, XT, BT, OT, AT, MODE, RED, true>);
      hipLaunchKernelGGL((sellKernel<VT, XT, BT, OT, AT, MODE, RED, true>),
                         dim3(groups < grid ? groups : grid), dim3(kSpmvThreads), 0, s, a);
    } else {
      static const int grid = residentBlocks(sellKernel<VT, XT, BT, OT, AT, MODE, RED, false>);
      hipLaunchKernelGGL((sellKernel<VT, XT, BT, OT, AT, MODE, RED, false>),
                         dim3(groups < grid ? groups : grid), dim3(kSpmvThreads), 0, s, a);
    }
  } else {
    if (a.numBlocks <= 0) return;
    hipLaunchKernelGGL((spmvKernel<S == 0 ? 1 : S, VT, XT, BT, OT, AT, MODE, RED>),
                       dim3(a.numBlocks), dim3(kSpmvThreads), 0, s, a);
  }
  GPH_HIP(hipGetLastError());
}

template <typename VT, typename XT, typename BT, typename OT, typename AT,
          int MODE>
static bool dispatchRed(const SpmvArgs& a, int red, hipStream_t s) {
  auto go = [&](auto sTag) {
    constexpr int S = decltype(sTag)::value;
    switch (red) {
      case kRedNone:
        launchOne<S, VT, XT, BT, OT, AT, MODE, kRedNone>(a, s);
        return true;
      case kRedSelf:
        launchOne<S, VT, XT, BT, OT, AT, MODE, kRedSelf>(a, s);
        return true;
      case kRedW:
        launchOne<S, VT, XT, BT, OT, AT, MODE, kRedW>(a, s);
        return true;
      case kRedSelfW:
        launchOne<S, VT, XT, BT, OT, AT, MODE, kRedSelfW>(a, s);
        return true;
    }
    return false;
  };
  if (a.numSlices > 0) return go(std::integral_constant<int, 0>{});
  if (a.nstreams == 1) return go(std::integral_constant<int, 1>{});
  if (a.nstreams == 2) return go(std::integral_constant<int, 2>{});
  return false;
}

// Supported type combinations (value, x, b, out, accumulate).
void launchSpmv(const SpmvArgs& a, DType vt, DType xt, DType bt, DType ot,
                DType at, int mode, int red, hipStream_t s) {
  using F16 = __half;
  const DType H = DType::F16, F = DType::F32, D = DType::F64;
  bool ok = false;
  if (mode == kModeSpmv) {
    if (vt == F && xt == F && ot == F && at == F)
      ok = dispatchRed<float, float, float, float, float, kModeSpmv>(a, red, s);
    else if (vt == H && xt == F && ot == F && at == F)
      ok = dispatchRed<F16, float, float, float, float, kModeSpmv>(a, red, s);
    else if (vt == D && xt == D && ot == D && at == D)
      ok = dispatchRed<double, double, double, double, double, kModeSpmv>(a, red, s);
    else if (vt == F && xt == D && ot == D && at == D)
      ok = dispatchRed<float, double, double, double, double, kModeSpmv>(a, red, s);
    else if (vt == F && xt == F && ot == F && at == D)
      ok = dispatchRed<float, float, float, float, double, kModeSpmv>(a, red, s);
  } else if (mode == kModeResidual) {
    if (vt == F && xt == F && bt == F && ot == F && at == F)
      ok = dispatchRed<float, float, float, float, float, kModeResidual>(a, red, s);
    else if (vt == H && xt == F && bt == F && ot == F && at == F)
      ok = dispatchRed<F16, float, float, float, float, kModeResidual>(a, red, s);
    else if (vt == D && xt == D && bt == D && ot == D && at == D)
      ok = dispatchRed<double, double, double, double, double, kModeResidual>(a, red, s);
    else if (vt == F && xt == D && bt == D && ot == D && at == D)
      ok = dispatchRed<float, double, double, double, double, kModeResidual>(a, red, s);
    else if (vt == F && xt == D && bt == F && ot == F && at == D)
      ok = dispatchRed<float, double, float, float, double, kModeResidual>(a, red, s);
    else if (vt == F && xt == D && bt == D && ot == F && at == D)
      ok = dispatchRed<float, double, double, float, double, kModeResidual>(a, red, s);
    else if (vt == F && xt == D && bt == F && ot == D && at == D)
      ok = dispatchRed<float, double, float, double, double, kModeResidual>(a, red, s);
    else if (vt == F && xt == F && bt == F && ot == F && at == D)
      ok = dispatchRed<float, float, float, float, double, kModeResidual>(a, red, s);
  }
  if (!ok)
    fail(kUnsupported, std::string("unsupported SpMV type combination: value ") +
                           dtypeName(vt) + ", x " + dtypeName(xt) + ", b " +
                           dtypeName(bt) + ", out " + dtypeName(ot) +
                           ", accumulate " + dtypeName(at));
}

}  // namespace gph
