// ILU(0) substitution with one workgroup per tile, for tiles of any size
// (gfx950).
//
// Same arithmetic as ilu/Solver.cpp:380-442 / :526-548 (block-Jacobi over
// tiles; forward t = b_i, t -= L_ij x_j over the in-tile lower entries in
// stored order, x_i = t; backward t = 0, t -= U_ij x_j over the in-tile
// upper entries descending, x_i = (x_i + t) / d_i; DILU: x_i = t / d_i and
// x_i += t / d_i). Per row the products are subtracted one after the other
// in that order, so x equals the sequential substitution bit for bit.
//
// 1024 threads walk the tile's levels (the rows of a level are
// independent) with one barrier per level; x sits in LDS when the tile
// fits (global memory otherwise). The right-hand side, with the PBiCGStab
// p- or s-update fused, is formed in one coalesced pass first. The factor
// stream (level-major, entry j of level slot s at j*R + s, so a level's
// loads are coalesced) is prefetched D levels ahead into a register ring:
// the loads for level L+D are issued when level L is done, so the chain of
// dependent levels waits on LDS and the barrier, not on memory. Thread t
// takes slot t of every level (the stream splits levels of more than 1024
// rows into pieces).
#include "kernels.hpp"

namespace gph {

// LDS: the tile's level tables (both directions), then x and its zero slot
// when they fit.
size_t blockSharedBytes(const BlockArgs& a, DType xt) {
  return blockInfoBytes(a.maxLevels) + (size_t)(a.maxRows + 1) * dtypeSize(xt);
}

static size_t blockMaxShared() { return 160 * 1024 - 4096; }

// Records of the stream: (row, 1/d) per slot and (column, value) per entry,
// 8 bytes, or 16 for float64 values (value in the upper half).
template <typename T>
struct BlockRec {
  using Raw = uint2;
  __device__ static uint32_t id(Raw q) { return q.x; }
  __device__ static T value(Raw q) {
    if constexpr (sizeof(T) == 2)
      return __ushort_as_half((unsigned short)(q.y & 0xffffu));
    else
      return __uint_as_float(q.y);
  }
};
template <>
struct BlockRec<double> {
  using Raw = uint4;
  __device__ static uint32_t id(Raw q) { return q.x; }
  __device__ static double value(Raw q) {
    return __longlong_as_double((long long)(((uint64_t)q.w << 32) | q.z));
  }
};

// x of the tile: LDS with a zero slot at n, or global memory read past the
// vector L1 (rows written by other waves of the workgroup in earlier
// levels), the zero slot read as +0.
template <typename X, bool LDSX>
__device__ __forceinline__ X readX(const X* xs, uint32_t c, uint32_t n) {
  if constexpr (LDSX) {
    return xs[c];
  } else {
    const X v = __hip_atomic_load(xs + (c < n ? c : 0u), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    return c < n ? v : (X)0;
  }
}

template <typename FT, typename X, int E, int D, bool LDSX, bool BWD>
__device__ __forceinline__ void blockSweep(const BlockDirArgs& d, const int4* info, int nL, X* xs,
                                           uint32_t n, int dilu) {
  if (nL <= 0) return;
  using RR = BlockRec<InvDiag<FT>>;
  using ER = BlockRec<FT>;
  const typename RR::Raw* __restrict__ rowA = static_cast<const typename RR::Raw*>(d.rows);
  const typename ER::Raw* __restrict__ entA = static_cast<const typename ER::Raw*>(d.ents);
  const int tid = threadIdx.x;
  const int wave0 = __builtin_amdgcn_readfirstlane(tid & ~63);  // first thread of the wave

  int4 rinfo[D];
  typename RR::Raw rrow[D];
  typename ER::Raw rent[D][E];
  // Loads of level L into ring slot k, D levels before their use (L past
  // the last level: the last level again, not used). Threads past the
  // level's rows read its last slot; waves past them load nothing: the one
  // CU's vector memory path bounds this kernel, and it spends the same time
  // on a wave's load whether or not its data are used. No load depends on
  // another, and the level table is in LDS.
  auto load = [&](int k, int L) {
    const int4 f = info[min(L, nL - 1)];
    rinfo[k] = f;
    if (wave0 >= f.y) return;
    const int s = min(tid, f.y - 1);
    rrow[k] = rowA[(uint32_t)(f.x + s)];
#pragma unroll
    for (int j = 0; j < E; j++) rent[k][j] = entA[(uint32_t)(f.z + j * f.y + s)];
  };

#pragma unroll
  for (int k = 0; k < D; k++) load(k, k);
  // U = 4 D levels per loop iteration (the compiler's waits for ring slots
  // across the loop's back edge are full ones)
  constexpr int U = 4 * D;
  for (int L0 = 0; L0 < nL; L0 += U) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = u % D;
      if (L0 + u < nL && wave0 < rinfo[k].y) {
        // every thread of the wave computes (a thread past the level's rows
        // on the level's last slot), only the level's rows store; waves
        // past the level's rows skip the level
        const uint32_t r = RR::id(rrow[k]);
        X xv[E];
#pragma unroll
        for (int j = 0; j < E; j++) xv[j] = readX<X, LDSX>(xs, ER::id(rent[k][j]), n);
        const X xr = readX<X, LDSX>(xs, r, n);
        X acc = BWD ? (X)0 : xr;
#pragma unroll
        for (int j = 0; j < E; j++) acc = acc - cvt<X>(ER::value(rent[k][j])) * xv[j];
        const X inv = cvt<X>(RR::value(rrow[k]));
        X res;
        if constexpr (!BWD)
          res = dilu ? acc * inv : acc;
        else
          res = dilu ? xr + acc * inv : (xr + acc) * inv;
        if (tid < rinfo[k].y) xs[r] = res;
      }
      load(k, L0 + u + D);
      __syncthreads();
    }
  }
}

template <typename FT, typename X, int E, int D, bool LDSX>
__global__ __launch_bounds__(kBlockThreads) void iluBlockKernel(
    BlockArgs a, const X* __restrict__ rhs, X* __restrict__ out, PreOp pre, DevState* st,
    int checkDone) {
  extern __shared__ __align__(16) unsigned char smemRaw[];
  if (checkDone && st->done) return;
  const int tile = blockIdx.x;
  const int t0 = a.tileStart[tile], n = a.tileStart[tile + 1] - t0;
  // level tables of both directions
  int4* info = reinterpret_cast<int4*>(smemRaw);
  const int lbF = a.fwd.lvlBase[tile], nF = a.fwd.lvlBase[tile + 1] - lbF;
  const int lbB = a.bwd.lvlBase[tile], nB = a.bwd.lvlBase[tile + 1] - lbB;
  for (int q = threadIdx.x; q < nF + nB; q += kBlockThreads)
    info[q] = q < nF ? reinterpret_cast<const int4*>(a.fwd.info)[lbF + q]
                     : reinterpret_cast<const int4*>(a.bwd.info)[lbB + q - nF];
  X* xs = LDSX ? reinterpret_cast<X*>(smemRaw + blockInfoBytes(a.maxLevels)) : out + t0;
  if (LDSX && threadIdx.x == 0) xs[n] = (X)0;  // the zero slot

  // right-hand side into x
  X beta = 0, omega = 0, alpha = 0;
  bool pupd = false;
  if (pre.kind == kPreP && st) {
    pupd = st->iter != 0;
    beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
    omega = (X)st->omega;
  } else if (pre.kind == kPreS && st) {
    alpha = (X)st->alpha;
  }
  const X* r = static_cast<const X*>(pre.r);
  X* pv = static_cast<X*>(pre.p);
  const X* vv = static_cast<const X*>(pre.v);
  for (int li = threadIdx.x; li < n; li += kBlockThreads) {
    const int i = t0 + li;
    X s;
    if (pre.kind == kPreP) {
      X pi = pv[i];
      if (pupd) {
        pi = r[i] + beta * (pi - omega * vv[i]);
        pv[i] = pi;
      }
      s = pi;
    } else if (pre.kind == kPreS) {
      s = r[i] - alpha * vv[i];
      pv[i] = s;
    } else {
      s = rhs[i];
    }
    xs[li] = s;
  }
  __syncthreads();
  blockSweep<FT, X, E, D, LDSX, false>(a.fwd, info, nF, xs, (uint32_t)n, a.dilu);
  blockSweep<FT, X, E, D, LDSX, true>(a.bwd, info + nF, nB, xs, (uint32_t)n, a.dilu);
  if (LDSX)
    for (int li = threadIdx.x; li < n; li += kBlockThreads) out[t0 + li] = xs[li];
}

template <typename FT, typename X, int E, int D, bool LDSX>
static void launchBlockKernel(const BlockArgs& a, size_t shm, const void* rhs, void* out,
                              const PreOp& pre, DevState* st, int checkDone, hipStream_t s) {
  auto k = iluBlockKernel<FT, X, E, D, LDSX>;
  static bool attr = false;
  if (!attr) {
    GPH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)blockMaxShared()));
    attr = true;
  }
  hipLaunchKernelGGL(k, dim3(a.T), dim3(kBlockThreads), shm, s, a, static_cast<const X*>(rhs),
                     static_cast<X*>(out), pre, st, checkDone);
  GPH_HIP(hipGetLastError());
}

template <typename FT, typename X, int E, int D>
static void launchBlockED(const BlockArgs& a, const void* rhs, void* out, const PreOp& pre,
                          DevState* st, int checkDone, hipStream_t s) {
  const size_t shm = blockSharedBytes(a, sizeof(X) == 8 ? DType::F64 : DType::F32);
  if (shm <= blockMaxShared())
    launchBlockKernel<FT, X, E, D, true>(a, shm, rhs, out, pre, st, checkDone, s);
  else if (blockInfoBytes(a.maxLevels) <= blockMaxShared())
    launchBlockKernel<FT, X, E, D, false>(a, blockInfoBytes(a.maxLevels), rhs, out, pre, st,
                                          checkDone, s);
  else
    fail(kUnsupported, "ILU block substitution: too many levels in a tile");
}

// Ring depth by width: about 80 registers of prefetched stream per thread.
template <typename FT, typename X>
static void launchBlockTyped(const BlockArgs& a, const void* rhs, void* out, const PreOp& pre,
                             DevState* st, int checkDone, hipStream_t s) {
  constexpr bool wide = sizeof(FT) == 8;
  if (a.width == 4)
    return launchBlockED<FT, X, 4, wide ? 4 : 6>(a, rhs, out, pre, st, checkDone, s);
  if (a.width == 8)
    return launchBlockED<FT, X, 8, wide ? 2 : 4>(a, rhs, out, pre, st, checkDone, s);
  if (a.width == 16)
    return launchBlockED<FT, X, 16, wide ? 1 : 2>(a, rhs, out, pre, st, checkDone, s);
  fail(kUnsupported, "ILU block substitution: stream width must be 4, 8 or 16");
}

void launchIluBlock(const BlockArgs& a, DType ft, const void* rhs, void* out,
                    DType xt, const PreOp& pre, DevState* st, int checkDone, hipStream_t s) {
  if (a.T <= 0) return;
  if (ft == DType::F32 && xt == DType::F32)
    launchBlockTyped<float, float>(a, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F16 && xt == DType::F32)
    launchBlockTyped<__half, float>(a, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F64 && xt == DType::F64)
    launchBlockTyped<double, double>(a, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F32 && xt == DType::F64)
    launchBlockTyped<float, double>(a, rhs, out, pre, st, checkDone, s);
  else
    fail(kUnsupported, std::string("unsupported ILU apply types: factors ") + dtypeName(ft) +
                           ", vectors " + dtypeName(xt));
}

}  // namespace gph
