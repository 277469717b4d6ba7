// ILU(0) substitution over the level-major packed factors (gfx950).
//
// Same arithmetic as ilu/Solver.cpp:380-442 / :526-548 (block-Jacobi over
// tiles, forward then backward substitution per tile), re-organised for one
// wave per tile:
//   * the tile's right-hand side (with the PBiCGStab p- or s-update fused),
//     the inverted diagonal, row descriptors and level tables are loaded
//     once into LDS with coalesced loads;
//   * the factor entries stream through a ring of kPackDepth slabs held in
//     registers, so the loads of later levels are in flight while the wave
//     works on the current level (the sweep is a chain of dependent levels,
//     latency- rather than bandwidth-limited without this);
//   * per level, lane k multiplies entry k with the x value it references
//     (LDS) into a product buffer, then one lane per row subtracts its
//     products in stored order - exactly the reference's sequential row sum.
#include <climits>

#include "kernels.hpp"

namespace gph {

static constexpr int kDepth = 16;  // must equal kPackDepth of ilu_pack.hpp

struct PackedLds {
  size_t xs, invd, prod, rows, slabEnd, rowStart, total;
};

__host__ __device__ inline PackedLds packedLayout(const PackedArgs& p,
                                                  size_t xsize) {
  PackedLds L;
  size_t o = 0;
  L.xs = o;
  o += (size_t)(p.maxRows + 1) * xsize;
  L.invd = o;
  o += (size_t)p.maxRows * xsize;
  L.prod = o;
  o += (size_t)(p.maxLevelEntries + 64) * xsize;
  o = (o + 7) & ~(size_t)7;
  L.rows = o;
  o += (size_t)p.maxRows * 8;
  L.slabEnd = o;
  o += (size_t)(p.maxLevels + 1) * 4;
  L.rowStart = o;
  o += (size_t)(p.maxLevels + 1) * 4;
  L.total = (o + 15) & ~(size_t)15;
  return L;
}

size_t packedSharedBytes(const PackedArgs& p, DType xt) {
  return packedLayout(p, dtypeSize(xt)).total;
}

template <typename FT, typename X, bool BACKWARD>
__device__ __forceinline__ void packedSweep(const PackedDirArgs& d, int tile,
                                            int t0, int nt, X* xs,
                                            const X* invd, X* prod,
                                            uint2* rows, int* slabEnd,
                                            int* rowStart, int dilu) {
  const int lane = threadIdx.x;
  const int lo = d.lvlOff[tile];
  const int nL = d.lvlOff[tile + 1] - lo - 1;
  for (int q = lane; q <= nL; q += kPackedThreads) {
    slabEnd[q] = d.slabEnd[lo + q];
    rowStart[q] = d.rowStart[lo + q];
  }
  const uint2* ri = reinterpret_cast<const uint2*>(d.rowInfo);
  for (int q = lane; q < nt; q += kPackedThreads) rows[q] = ri[t0 + q];
  __syncthreads();
  const int S = nL > 0 ? slabEnd[nL - 1] : 0;
  const int Stot = ((S + kDepth - 1) / kDepth) * kDepth;
  // 16-bit slots are stored slab-pair interleaved (the stream starts on an
  // even slab): lane l finds slots (2p, l) and (2p+1, l) in the low and high
  // half of 32-bit word p*64 + l. The ring holds these words as loaded;
  // narrowing them in the loop would make every iteration wait on the loads.
  constexpr bool kHalfVals = sizeof(FT) == 2;
  const size_t wbase = (size_t)d.slabOff[tile] * (kPackedThreads / 2) + lane;
  const uint32_t* __restrict__ cp = reinterpret_cast<const uint32_t*>(d.col) + wbase;
  const uint32_t* __restrict__ hp = reinterpret_cast<const uint32_t*>(d.val) + wbase;
  const FT* __restrict__ vp =
      static_cast<const FT*>(d.val) + (size_t)d.slabOff[tile] * kPackedThreads + lane;

  // The ring starts empty (zeros) and the loop starts one depth early: the
  // first pass only issues the loads, in the same order as every later
  // pass, so the load waits the compiler places hold for the steady state.
  uint32_t cw[kDepth / 2] = {};
  uint32_t hw[kHalfVals ? kDepth / 2 : 1] = {};
  FT v[kHalfVals ? 1 : kDepth] = {};
  int lvl = 0, lvlFirst = 0, lvlEnd = slabEnd[0];
  for (int sb = -kDepth; sb < Stot; sb += kDepth) {
#pragma unroll
    for (int u = 0; u < kDepth; u++) {
      const int s = sb + u;
      const uint32_t c = (u & 1) ? cw[u >> 1] >> 16 : cw[u >> 1] & 0xffffu;
      X a;
      if constexpr (kHalfVals)
        a = cvt<X>(__ushort_as_half(
            (unsigned short)((u & 1) ? hw[u >> 1] >> 16 : hw[u >> 1] & 0xffffu)));
      else
        a = cvt<X>(v[u]);
      // slabs before 0 (empty ring) and past S (ring overrun into the next
      // group's stream or the padding, after the last level's row sums)
      // write their product to a dead slot; every ring word is consumed on
      // every path, so no refill has to wait for an unused load
      const int pi = (s >= 0 && s < S) ? (s - lvlFirst) * kPackedThreads + lane : lane;
      prod[pi] = a * xs[c];
      if (u & 1) {  // both slots of the word consumed: refill it
        const size_t nw = (size_t)((s + kDepth) >> 1) * kPackedThreads;
        cw[u >> 1] = cp[nw];
        if constexpr (kHalfVals) hw[u >> 1] = hp[nw];
      }
      if constexpr (!kHalfVals) v[u] = vp[(size_t)(s + kDepth) * kPackedThreads];
      if (s + 1 == lvlEnd) {
        __syncthreads();
        const int r1 = rowStart[lvl + 1];
        for (int r = rowStart[lvl] + lane; r < r1; r += kPackedThreads) {
          const uint2 info = rows[r];
          const int loc = (int)(info.x & 0xffffu);
          const int len = (int)(info.x >> 16);
          const X* pp = prod + info.y;
          X acc = BACKWARD ? (X)0 : xs[loc];
          // issue the LDS reads of 8 products at once, then subtract in order
          for (int a0 = 0; a0 < len; a0 += 8) {
            X pv[8];
#pragma unroll
            for (int j = 0; j < 8; j++) pv[j] = (a0 + j < len) ? pp[a0 + j] : (X)0;
#pragma unroll
            for (int j = 0; j < 8; j++)
              if (a0 + j < len) acc = acc - pv[j];
          }
          if (BACKWARD)
            xs[loc] = dilu ? xs[loc] + acc * invd[loc] : (xs[loc] + acc) * invd[loc];
          else
            xs[loc] = dilu ? acc * invd[loc] : acc;
        }
        __syncthreads();
        lvl++;
        lvlFirst = lvlEnd;
        lvlEnd = slabEnd[lvl];
      }
    }
  }
  __syncthreads();
}

template <typename FT, typename X>
__global__ __launch_bounds__(kPackedThreads) void iluPackedKernel(
    PackedArgs p, const FT* __restrict__ invD, const X* __restrict__ rhs,
    X* __restrict__ out, PreOp pre, DevState* st, int checkDone) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (checkDone && st->done) return;
  const PackedLds L = packedLayout(p, sizeof(X));
  X* xs = reinterpret_cast<X*>(smem + L.xs);
  X* invd = reinterpret_cast<X*>(smem + L.invd);
  X* prod = reinterpret_cast<X*>(smem + L.prod);
  uint2* rows = reinterpret_cast<uint2*>(smem + L.rows);
  int* slabEnd = reinterpret_cast<int*>(smem + L.slabEnd);
  int* rowStart = reinterpret_cast<int*>(smem + L.rowStart);
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int t0 = p.tileStart[tile], nt = p.tileStart[tile + 1] - t0;

  // right-hand side (with the fused PBiCGStab update) and 1/diag into LDS
  X beta = 0, omega = 0, alpha = 0;
  bool pupd = false;
  if (pre.kind == kPreP) {
    pupd = st->iter != 0;
    beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
    omega = (X)st->omega;
  } else if (pre.kind == kPreS) {
    alpha = (X)st->alpha;
  }
  const X* r = static_cast<const X*>(pre.r);
  X* pv = static_cast<X*>(pre.p);
  const X* vv = static_cast<const X*>(pre.v);
  for (int q = lane; q < nt; q += kPackedThreads) {
    const int i = t0 + q;
    X s;
    if (pre.kind == kPreP) {
      s = pv[i];
      if (pupd) {
        s = r[i] + beta * (s - omega * vv[i]);
        pv[i] = s;
      }
    } else if (pre.kind == kPreS) {
      s = r[i] - alpha * vv[i];
      pv[i] = s;
    } else {
      s = rhs[i];
    }
    xs[q] = s;
    invd[q] = cvt<X>(invD[i]);
  }
  if (lane == 0) xs[nt] = (X)0;
  __syncthreads();
  packedSweep<FT, X, false>(p.fwd, tile, t0, nt, xs, invd, prod, rows, slabEnd,
                            rowStart, p.dilu);
  packedSweep<FT, X, true>(p.bwd, tile, t0, nt, xs, invd, prod, rows, slabEnd,
                           rowStart, p.dilu);
  for (int q = lane; q < nt; q += kPackedThreads) out[t0 + q] = xs[q];
}

template <typename FT, typename X>
static void packedTyped(const PackedArgs& p, const void* invD, const void* rhs,
                        void* out, const PreOp& pre, DevState* st,
                        int checkDone, hipStream_t s) {
  const size_t shm = packedSharedBytes(p, sizeof(X) == 8 ? DType::F64 : DType::F32);
  auto k = iluPackedKernel<FT, X>;
  if (shm > 64 * 1024)
    GPH_HIP(hipFuncSetAttribute((const void*)k,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
  hipLaunchKernelGGL(k, dim3(p.T), dim3(kPackedThreads), shm, s, p,
                     static_cast<const FT*>(invD), static_cast<const X*>(rhs),
                     static_cast<X*>(out), pre, st, checkDone);
  GPH_HIP(hipGetLastError());
}

void launchIluPacked(const PackedArgs& p, const void* invD, DType ft,
                     const void* rhs, void* out, DType xt, const PreOp& pre,
                     DevState* st, int checkDone, hipStream_t s) {
  if (p.T <= 0) return;
  if (ft == DType::F32 && xt == DType::F32)
    packedTyped<float, float>(p, invD, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F16 && xt == DType::F32)
    packedTyped<__half, float>(p, invD, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F64 && xt == DType::F64)
    packedTyped<double, double>(p, invD, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F32 && xt == DType::F64)
    packedTyped<float, double>(p, invD, rhs, out, pre, st, checkDone, s);
  else
    fail(kUnsupported, std::string("unsupported packed ILU types: factors ") +
                           dtypeName(ft) + ", vectors " + dtypeName(xt));
}

template <typename T>
__global__ void gatherKernel(const T* src, const int32_t* idx, T* dst,
                             int64_t n) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = idx[k];
    dst[k] = i < 0 ? (T)0 : src[i];
  }
}

void launchGather(const void* src, DType t, const int32_t* idx, void* dst,
                  int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  if (t == DType::F32)
    hipLaunchKernelGGL(gatherKernel<float>, dim3((int)g), dim3(256), 0, s,
                       static_cast<const float*>(src), idx,
                       static_cast<float*>(dst), n);
  else if (t == DType::F64)
    hipLaunchKernelGGL(gatherKernel<double>, dim3((int)g), dim3(256), 0, s,
                       static_cast<const double*>(src), idx,
                       static_cast<double*>(dst), n);
  else
    hipLaunchKernelGGL(gatherKernel<__half>, dim3((int)g), dim3(256), 0, s,
                       static_cast<const __half*>(src), idx,
                       static_cast<__half*>(dst), n);
  GPH_HIP(hipGetLastError());
}

}  // namespace gph
