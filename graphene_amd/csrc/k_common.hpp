// Device-side helpers shared by the gfx950 kernels: the solver state block
// that lives in HBM, deterministic block / grid reductions, and the
// XCD-aware block remap.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gph {

// Scalar state of one Krylov solve. All scalars are stored as double; a
// value computed in float working precision is exactly representable, so
// kernels re-derive the working-precision value by a cast.
struct DevState {
  double rho, rhoOld, alpha, omega;
  double r0n, rF, bNorm, aux;
  int32_t iter, converged, singular, done;
  float absTol, relTol, relRes;
  int32_t minIt, maxIt;
  int32_t innerTotal;  // batched refinement: inner iterations of the finished inner solves
  int32_t pad[2];
  uint32_t counter[8];
};
static_assert(sizeof(DevState) == 144, "DevState layout");

// What the last block of a fused reduction does with the totals.
enum FinOp : int32_t {
  kFinNone = 0,
  kFinStore = 1,      // aux = sum0 (raw sum)
  kFinNormAux = 2,    // aux = sqrt(sum0)
  kFinBNorm = 3,      // bNorm = sqrt(sum0)
  kFinInit = 4,       // PBiCGStab start: r0n = rF = ||r||, rho = r.r
  kFinAlpha = 5,      // alpha = rho / sum0
  kFinOmega = 6,      // omega = sum1 / sum0, singularity check
  kFinUpdate = 7,     // end of iteration: ||r||, convergence, next rho
  // The same steps for a norm other than L2, in two kernels: the vector
  // kernel's reduction updates rho (kFinUpdateRho), then a norm kernel
  // hands over the finished norm value (the ...V ops take it as sum0).
  kFinUpdateRho = 8,    // rhoOld = rho, iteration + 1, rho = sum1
  kFinUpdateNormV = 9,  // rF = norm, convergence, singularity, done
  kFinInitNormV = 10,   // r0n = rF = norm (after kFinInit)
  kFinBNormV = 11,      // bNorm = norm
  kFinNormAuxV = 12,    // aux = norm
};

// Vector norms (lg/matrix/Norm.hpp): L2 = sqrt(sum x^2), L1 = sum |x|,
// LINF = max |x|.
enum NormKind : int32_t { kNormL2 = 0, kNormL1 = 1, kNormLinf = 2 };

// Combination of partial results: sums, or maxima (LINF).
template <int OP, typename R>
__device__ __forceinline__ R combine(R a, R b) {
  if constexpr (OP == 1)
    return a > b ? a : b;
  else
    return a + b;
}

// Singularity threshold Traits<float>::vsmall() (lg/common/Traits.hpp:38-46).
constexpr double kVSmall = 1e-37;

template <typename T>
__device__ __forceinline__ float toF(T v) {
  return (float)v;
}
template <>
__device__ __forceinline__ float toF<__half>(__half v) {
  return __half2float(v);
}
template <typename A, typename V>
__device__ __forceinline__ A cvt(V v) {
  return (A)v;
}
template <>
__device__ __forceinline__ float cvt<float, __half>(__half v) {
  return __half2float(v);
}
template <>
__device__ __forceinline__ double cvt<double, __half>(__half v) {
  return (double)__half2float(v);
}

// Element type of the inverted diagonal stored with ILU(0) factors of type
// FT: float16 off-diagonal factors keep a float32 inverse diagonal (one
// value per row, so the narrow storage saves little there, while rounding
// it to float16 perturbs every row of the substitution).
template <typename FT>
struct InvDiagOf {
  using type = FT;
};
template <>
struct InvDiagOf<__half> {
  using type = float;
};
template <typename FT>
using InvDiag = typename InvDiagOf<FT>::type;

// Load of a matrix stream element that this launch reads exactly once:
// non-temporal, so the streams do not displace the gathered vector from the
// caches (sliced kernel, 128^3 27-point SpMV: 83.1 -> 79.5 us on MI355X; the
// row-block kernel of the power-law matrix keeps ordinary loads, which were
// faster there: 236 against 249 us).
template <typename T>
__device__ __forceinline__ T streamLoad(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ __half streamLoad<__half>(const __half* p) {
  return __ushort_as_half(__builtin_nontemporal_load(reinterpret_cast<const unsigned short*>(p)));
}

// Blocks are dispatched to the 8 XCDs round robin (block b -> XCD b % 8).
// Remap so that each XCD works on one contiguous range of logical blocks,
// keeping the gathered x window of neighbouring rows in one L2.
__device__ __forceinline__ int xcdRemap(int b, int nb) {
  const int q = nb >> 3, rem = nb & 7;
  const int x = b & 7, i = b >> 3;
  return x * q + min(x, rem) + i;
}

// Deterministic block sum of K values over a 1-D block of NT threads
// (NT multiple of 64). Result valid in thread 0.
// OP = 1: maxima instead of sums.
template <typename R, int K, int NT, int OP = 0>
__device__ __forceinline__ void blockSum(R (&v)[K], R* lds /* K*NT/64 */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) {
    R s = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s = combine<OP>(s, __shfl_down(s, off, 64));
    if (lane == 0) lds[k * (NT / 64) + wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      R s = lds[k * (NT / 64)];
      for (int w = 1; w < NT / 64; w++) s = combine<OP>(s, lds[k * (NT / 64) + w]);
      v[k] = s;
    }
  }
  __syncthreads();
}

// Layout of a grid reduction's scratch buffer: first-level arrival counters
// (kArriveSlots of them, one per 128-byte line, zero at allocation), then
// the partial sums.
constexpr int kArriveGroup = 32;    // blocks per first-level counter (at least)
constexpr int kArriveSlots = 128;   // first-level counters
constexpr int kArriveStride = 32;   // u32 per counter line
constexpr size_t kArriveBytes = (size_t)kArriveSlots * kArriveStride * 4;

// Grid-wide reduction epilogue: every block publishes its K partial sums;
// the last block to arrive adds all partials in a fixed order and returns
// true (in all of its threads) with the totals in `tot`. The arrival
// counters are reset by the blocks that finish them, so no memset is needed
// between uses.
//
// Arrival is counted in two levels: blocks of a group of consecutive block
// ids increment their group's counter, and the last block of each group
// increments the top counter. Atomics on one address are executed one after
// the other at the device's coherence point (about 12 ns each here), so a
// single counter for a grid of 2048 blocks serialises for about 25 us at the
// end of a launch; the groups' counters sit on separate lines and proceed
// in parallel.
//
// The partials are written and read with device-scope atomic stores/loads,
// which are coherent across the XCDs' separate L2 caches, and the writer
// waits for its stores to complete before it increments a counter. This
// avoids a device-scope release fence per block (an L2 write-back on this
// hardware, which costs far more than the reduction itself). The last block
// issues all of its loads before the first add (a device-scope load takes
// about a microsecond); the adds take the partials in the same fixed order
// as a plain strided loop.
template <typename R, int K, int NT, int OP = 0>
__device__ bool gridReduce(R (&v)[K], void* scratch, uint32_t* counter,
                           R (&tot)[K], R* lds) {
  __shared__ int isLast;
  blockSum<R, K, NT, OP>(v, lds);
  const int nb = gridDim.x;
  R* partials = reinterpret_cast<R*>(static_cast<char*>(scratch) + kArriveBytes);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++)
      __hip_atomic_store(&partials[k * nb + blockIdx.x], v[k], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);  // stores complete before the arrival
    const int gs = max(kArriveGroup, (nb + kArriveSlots - 1) / kArriveSlots);
    const int ng = (nb + gs - 1) / gs;
    bool last = false;
    if (ng == 1) {
      last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) == (uint32_t)(nb - 1);
    } else {
      const int g = blockIdx.x / gs;
      const uint32_t members = (uint32_t)min(gs, nb - g * gs);
      uint32_t* sub = static_cast<uint32_t*>(scratch) + g * kArriveStride;
      if (__hip_atomic_fetch_add(sub, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
          members - 1) {
        __hip_atomic_store(sub, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == (uint32_t)(ng - 1);
      }
    }
    isLast = last;
  }
  __syncthreads();
  if (!isLast) return false;
  R acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = 0;
  constexpr int kB = 8;  // loads in flight per thread and partial array
  for (int b0 = threadIdx.x; b0 < nb; b0 += kB * NT) {
    R w[K][kB];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int j = 0; j < kB; j++) {
        const int b = b0 + j * NT;
        w[k][j] = b < nb ? __hip_atomic_load(&partials[k * nb + b], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)
                         : (R)0;
      }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int j = 0; j < kB; j++)
        if (b0 + j * NT < nb) acc[k] = combine<OP>(acc[k], w[k][j]);
  }
  blockSum<R, K, NT, OP>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) tot[k] = acc[k];
    *counter = 0u;
  }
  return true;
}

// Convergence test, SolverStats::checkConvergence
// (lg/matrix/solver/SolverStats.cpp:33-56).
__device__ __forceinline__ bool convergedTest(const DevState* s, double rF) {
  bool c = rF < (double)s->absTol || rF < s->r0n * (double)s->relRes;
  if (s->relTol > 0) c = c || rF < s->bNorm * (double)s->relTol;
  return c;
}

// Scalar epilogues of the fused kernels. X is the working precision.
template <typename X>
__device__ void finalize(DevState* s, int op, double sum0, double sum1) {
  switch (op) {
    case kFinStore:
      s->aux = sum0;
      break;
    case kFinNormAux:
      s->aux = (double)sqrt((X)sum0);
      break;
    case kFinBNorm:
      s->bNorm = (double)sqrt((X)sum0);
      break;
    case kFinInit: {
      // pbicgstab/Solver.cpp:66-90: initial residual, rho of the first
      // iteration (r0.r == r.r), loop-entry condition.
      double n0 = (double)sqrt((X)sum0);
      s->r0n = n0;
      s->rF = n0;
      s->iter = 0;
      s->converged = 0;
      s->singular = 0;
      s->done = (s->maxIt <= 0) ? 1 : 0;
      if (!s->done) {
        s->rho = sum0;
        if (fabs(sum0) < kVSmall) {
          s->singular = 1;
          s->done = 1;
        }
      }
      break;
    }
    case kFinAlpha:
      // pbicgstab/Solver.cpp:121-122
      s->alpha = (double)((X)s->rho / (X)sum0);
      break;
    case kFinOmega: {
      // pbicgstab/Solver.cpp:151-159
      X om = (X)sum1 / (X)sum0;
      s->omega = (double)om;
      if (fabs((double)om) < kVSmall) {
        s->singular = 1;
        s->done = 1;
      }
      break;
    }
    case kFinUpdate: {
      // pbicgstab/Solver.cpp:170-180 and the loop condition :86-90; the
      // next iteration's rho (:91-95) is the second sum.
      double rF = (double)sqrt((X)sum0);
      s->rF = rF;
      s->converged = convergedTest(s, rF) ? 1 : 0;
      s->rhoOld = s->rho;
      s->iter = s->iter + 1;
      bool term = (s->converged && s->iter >= s->minIt) || s->iter >= s->maxIt;
      if (!term) {
        s->rho = sum1;
        if (fabs(sum1) < kVSmall) {
          s->singular = 1;
          term = true;
        }
      }
      s->done = term ? 1 : 0;
      break;
    }
    case kFinUpdateRho:
      s->rhoOld = s->rho;
      s->iter = s->iter + 1;
      s->rho = sum1;
      break;
    case kFinUpdateNormV: {
      // the rest of kFinUpdate, with the norm handed over
      s->rF = sum0;
      s->converged = convergedTest(s, sum0) ? 1 : 0;
      bool term = (s->converged && s->iter >= s->minIt) || s->iter >= s->maxIt;
      if (!term && fabs(s->rho) < kVSmall) {
        s->singular = 1;
        term = true;
      }
      s->done = term ? 1 : 0;
      break;
    }
    case kFinInitNormV:
      s->r0n = sum0;
      s->rF = sum0;
      break;
    case kFinBNormV:
      s->bNorm = sum0;
      break;
    case kFinNormAuxV:
      s->aux = sum0;
      break;
    default:
      break;
  }
}

}  // namespace gph
