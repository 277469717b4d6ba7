// ILU(0) substitution over level records, one lane per row (gfx950).
//
// Same arithmetic as ilu/Solver.cpp:380-442 / :526-548 (block-Jacobi over
// tiles, forward then backward substitution per tile). One wave runs a pair
// of tile groups: lanes 0-31 the first, 32-63 the second, each group's x in
// LDS with a zero slot that stays +0. A step holds up to 32 rows of one
// level per group: lane r reads its row's columns and factor values (two
// per read, row-major) and subtracts the products from its partial sum in
// the reference's order; entries past a row's end refer to the zero slot
// with value 0 and subtract +0, which leaves the sum unchanged. Steps of a
// level are independent; a step's LDS writes precede the next step's reads
// in the wave's program order.
//
// The record stream (forward then backward steps, contiguous per pair) is
// copied into an LDS ring with LDS-DMA (global_load_lds_dwordx4: 1 KiB per
// wave instruction, no registers). The wave keeps kLvlLag chunks in flight
// beyond the current step and waits on the memory counter only for the
// chunks that hold the current step and the next step's header, so the
// record loads of later steps overlap the dependent chain of levels. No
// record crosses the ring end (the packer splits or pads), so a record is
// read at fixed offsets from its ring position.
#include "kernels.hpp"

namespace gph {

using LdsPtr = __attribute__((address_space(3))) void*;

// One LDS-DMA wave load: 16 bytes per lane from `src` to LDS byte address
// lds + 16 * lane. Written as inline assembly so that the compiler does not
// treat every later LDS read as a possible reader of the DMA data (it would
// wait for all outstanding DMA before each one); the kernel waits for the
// chunks it reads itself.
__device__ __forceinline__ void ldsDma16(const void* src, uint32_t lds) {
  asm volatile("global_load_lds_dwordx4 %0, off" ::"v"(src), "{m0}"(lds) : "memory");
}

// Waits until at most kLvlLag, or no, vector memory operations are
// outstanding (in issue order, so all older ones have completed).
__device__ __forceinline__ void waitVmem(bool lag) {
  if (lag) {
    static_assert(kLvlLag == 10, "wait count below");
    asm volatile("s_waitcnt vmcnt(10)" ::: "memory");
  } else {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  }
}

__host__ __device__ inline size_t levelXsOffset() { return kLvlRing; }

// x of both halves, each with its zero slot at index maxRows.
size_t levelSharedBytes(const LevelArgs& p, DType xt) {
  return ((size_t)kLvlRing + 2 * (size_t)(p.maxRows + 1) * dtypeSize(xt) + 15) & ~(size_t)15;
}

template <typename T>
__device__ __forceinline__ T ldsAt(const unsigned char* base, uint32_t off) {
  return *reinterpret_cast<const T*>(base + off);
}

// Two consecutive factor values of a row (one 4-, 8- or 16-byte LDS read).
template <typename FT, typename X>
__device__ __forceinline__ void ldsPair(const unsigned char* p, X& a, X& b) {
  if constexpr (sizeof(FT) == 2) {
    const uint32_t u = *reinterpret_cast<const uint32_t*>(p);
    a = cvt<X>(__ushort_as_half((unsigned short)(u & 0xffffu)));
    b = cvt<X>(__ushort_as_half((unsigned short)(u >> 16)));
  } else if constexpr (sizeof(FT) == 4) {
    const float2 v = *reinterpret_cast<const float2*>(p);
    a = cvt<X>(v.x);
    b = cvt<X>(v.y);
  } else {
    const double2 v = *reinterpret_cast<const double2*>(p);
    a = cvt<X>(v.x);
    b = cvt<X>(v.y);
  }
}

template <typename FT, typename X>
__global__ __launch_bounds__(64) void iluLevelKernel(LevelArgs p, const X* __restrict__ rhs,
                                                     X* __restrict__ out, PreOp pre,
                                                     DevState* st, int checkDone) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (checkDone && st->done) return;
  unsigned char* ring = smem;
  const uint32_t ringLds = (uint32_t)(uintptr_t)(LdsPtr)smem;
  const int pair = blockIdx.x, lane = threadIdx.x;
  const int half = lane >> 5, r = lane & 31;
  // this lane's half: group 2*pair + half (absent for the last odd group)
  const int gi = 2 * pair + half;
  const int t0 = gi < p.G ? p.groupStart[gi] : 0;
  const int nt = gi < p.G ? p.groupStart[gi + 1] - t0 : 0;
  const int zs = p.maxRows;  // zero slot
  X* xs = reinterpret_cast<X*>(smem + levelXsOffset()) + half * (zs + 1);
  const unsigned char* blob = p.bytes + p.pairOff[pair] + lane * 16;
  const int nF = p.steps[2 * pair], nS = nF + p.steps[2 * pair + 1];

  // chunk c of the stream -> ring bytes (c * 1 KiB) mod ring size
  int issued = 0;
  auto issue = [&](int upTo) {
    for (; issued < upTo; issued++)
      ldsDma16(blob + (size_t)issued * kLvlChunk,
               ringLds + ((uint32_t)(issued * kLvlChunk) & (kLvlRing - 1)));
  };
  issue(kLvlLag);

  // right-hand side (with the fused PBiCGStab update) into x, 32 lanes per
  // half; the zero slot stays +0
  const X* rv = static_cast<const X*>(pre.r);
  X* pv = static_cast<X*>(pre.p);
  const X* vv = static_cast<const X*>(pre.v);
  auto stage = [&](auto f, X* wb) {
    for (int q = r; q < nt; q += 32) {
      const X a = f(t0 + q);
      if (wb) wb[t0 + q] = a;
      xs[q] = a;
    }
  };
  if (pre.kind == kPreP) {
    if (st->iter != 0) {
      const X beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
      const X omega = (X)st->omega;
      stage([&](int i) { return rv[i] + beta * (pv[i] - omega * vv[i]); }, pv);
    } else {
      stage([&](int i) { return pv[i]; }, (X*)nullptr);
    }
  } else if (pre.kind == kPreS) {
    const X alpha = (X)st->alpha;
    stage([&](int i) { return rv[i] - alpha * vv[i]; }, pv);
  } else {
    stage([&](int i) { return rhs[i]; }, (X*)nullptr);
  }
  if (r == 0) xs[zs] = (X)0;

  constexpr uint32_t es = sizeof(FT);
  auto pad16 = [](uint32_t b) { return (b + 15u) & ~15u; };
  waitVmem(false);
  uint4 hdr = ldsAt<uint4>(ring, 0);
  uint32_t cur = 0;
  for (int s = 0; s < nS; s++) {
    const uint32_t Rm = __builtin_amdgcn_readfirstlane(hdr.x) & 0xffffu;
    const uint32_t W = (uint32_t)__builtin_amdgcn_readfirstlane(hdr.x) >> 16;
    const uint32_t bytes = __builtin_amdgcn_readfirstlane(hdr.y);
    const uint32_t flags = __builtin_amdgcn_readfirstlane(hdr.z);
    const uint32_t next = cur + bytes;
    // chunks holding this step and the next header; the ring may run up to
    // one ring length past the chunk the step starts in
    const int need = (int)((next + 16 + kLvlChunk - 1) / kLvlChunk);
    const int limit = (int)(cur / kLvlChunk) + kLvlRing / kLvlChunk;
    const int upTo = min(need + kLvlLag, limit);
    issue(upTo);
    waitVmem(upTo == need + kLvlLag);
    const uint4 hdrNext = ldsAt<uint4>(ring, next & (kLvlRing - 1));
    const bool bwd = s >= nF;
    // the record does not cross the ring end
    const unsigned char* rec = ring + (cur & (kLvlRing - 1));
    const uint32_t qInv = 16 + pad16(4 * Rm);
    const uint32_t qc = qInv + ((flags & 1u) ? pad16(es * 2 * Rm) : 0u);
    const uint32_t qv = qc + pad16(4 * Rm * W);
    if ((uint32_t)r < Rm) {
      const uint32_t q = half * Rm + r;  // row slot
      const int loc = ldsAt<uint16_t>(rec, 16 + 2 * q);
      X d = (X)0;
      if (flags & 1u) d = cvt<X>(ldsAt<FT>(rec, qInv + es * q));
      const X px = xs[loc];
      X acc = bwd ? (X)0 : px;
      // row q: W/2 column words (two 16-bit columns each) and W values
      const unsigned char* cp = rec + qc + 2 * q * W;
      const unsigned char* vp = rec + qv + es * q * W;
      const uint32_t P = W >> 1;
      uint32_t k = 0;
      // 4 pairs per step: 8 reads, 8 gathers, then the subtractions in order
      for (; k + 4 <= P; k += 4) {
        uint32_t cw[4];
        X v[8];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          cw[u] = ldsAt<uint32_t>(cp, 4 * (k + u));
          ldsPair<FT>(vp + 2 * es * (k + u), v[2 * u], v[2 * u + 1]);
        }
        X xv[8];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          xv[2 * u] = xs[cw[u] & 0xffffu];
          xv[2 * u + 1] = xs[cw[u] >> 16];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc = acc - v[u] * xv[u];
      }
      // remaining 1-3 pairs: clamped reads, the extra ones not subtracted
      if (k < P) {
        uint32_t cw[3];
        X v[6];
#pragma unroll
        for (int u = 0; u < 3; u++) {
          const uint32_t kk = k + u < P ? k + u : P - 1;
          cw[u] = ldsAt<uint32_t>(cp, 4 * kk);
          ldsPair<FT>(vp + 2 * es * kk, v[2 * u], v[2 * u + 1]);
        }
        X xv[6];
#pragma unroll
        for (int u = 0; u < 3; u++) {
          xv[2 * u] = xs[cw[u] & 0xffffu];
          xv[2 * u + 1] = xs[cw[u] >> 16];
        }
#pragma unroll
        for (int u = 0; u < 3; u++)
          if (k + u < P) {
            acc = acc - v[2 * u] * xv[2 * u];
            acc = acc - v[2 * u + 1] * xv[2 * u + 1];
          }
      }
      if (bwd)
        xs[loc] = p.dilu ? px + acc * d : (px + acc) * d;
      else
        xs[loc] = p.dilu ? acc * d : acc;
    }
    cur = next;
    hdr = hdrNext;
  }
  for (int q = r; q < nt; q += 32) __builtin_nontemporal_store(xs[q], out + t0 + q);
}

template <typename FT, typename X>
static void levelTyped(const LevelArgs& p, const void* rhs, void* out, const PreOp& pre,
                       DevState* st, int checkDone, hipStream_t s) {
  const size_t shm = levelSharedBytes(p, sizeof(X) == 8 ? DType::F64 : DType::F32);
  auto k = iluLevelKernel<FT, X>;
  if (shm > 64 * 1024)
    GPH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
  hipLaunchKernelGGL(k, dim3(p.P), dim3(64), shm, s, p, static_cast<const X*>(rhs),
                     static_cast<X*>(out), pre, st, checkDone);
  GPH_HIP(hipGetLastError());
}

void launchIluLevel(const LevelArgs& p, DType ft, const void* rhs, void* out, DType xt,
                    const PreOp& pre, DevState* st, int checkDone, hipStream_t s) {
  if (p.P <= 0) return;
  if (ft == DType::F32 && xt == DType::F32)
    levelTyped<float, float>(p, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F16 && xt == DType::F32)
    levelTyped<__half, float>(p, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F64 && xt == DType::F64)
    levelTyped<double, double>(p, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F32 && xt == DType::F64)
    levelTyped<float, double>(p, rhs, out, pre, st, checkDone, s);
  else
    fail(kUnsupported, std::string("unsupported level ILU types: factors ") +
                           dtypeName(ft) + ", vectors " + dtypeName(xt));
}

}  // namespace gph
