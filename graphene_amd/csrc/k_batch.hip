// Batched PBiCGStab / MPIR kernels: many independent systems stored as one
// block-diagonal matrix (each system a contiguous range of rows and tiles).
//
// SpMV and the ILU(0) substitution run on the union unchanged (no entry
// couples two systems, the tiles of the block-Jacobi preconditioner never
// span two systems). What is per system: the Krylov scalars and the
// convergence state (one DevState each, updated by the scalar epilogue of
// the single-system solver, pbicgstab/Solver.cpp:86-180), the dot products
// (segmented: every chunk of rows lies in one system), and the vector
// updates (masked by the system's `done` flag). Reductions are
// deterministic: each chunk's partial sum is a fixed-order block sum, and a
// system adds its chunks' partials in chunk order.
//
// Every kernel returns at once when the batch is finished (glob->done), so
// a captured graph of several iterations can be replayed past convergence.
#include "kernels.hpp"

namespace gph {

// Rows of chunk c: [chunkRow[c], chunkRow[c+1]), system chunkSys[c].
struct ChunkRange {
  int32_t sys, r0, r1;
};
__device__ __forceinline__ ChunkRange chunkOf(const BatchArgs& b, int c) {
  return ChunkRange{b.chunkSys[c], b.chunkRow[c], b.chunkRow[c + 1]};
}

// Partial sums of K dot products a_k . b_k over one chunk (one block per
// chunk); chunks of finished systems are skipped (their partials unused).
template <typename X, int K>
__global__ __launch_bounds__(kBatchThreads) void batchDotKernel(BatchArgs b, BatchDot<K> d) {
  __shared__ X lds[K * kBatchThreads / 64];
  if (b.glob->done) return;
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  if (b.st[ch.sys].done) return;
  X acc[K];
#pragma unroll
  for (int k = 0; k < K; k++) acc[k] = (X)0;
  for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads) {
#pragma unroll
    for (int k = 0; k < K; k++)
      acc[k] = acc[k] + static_cast<const X*>(d.a[k])[i] * static_cast<const X*>(d.b[k])[i];
  }
  blockSum<X, K, kBatchThreads>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) static_cast<X*>(b.partials)[k * b.C + blockIdx.x] = acc[k];
  }
}

// Sum of one system's chunk partials: up to kSerialChunks chunks one thread
// adds them in chunk order; more chunks, one wave: lane l adds chunks l,
// l + 64, ... in order and the lanes' sums are combined in a fixed shuffle
// tree. Either way the order depends only on the system's own chunk count.
constexpr int kSerialChunks = 32;

template <typename X>
__device__ __forceinline__ X chunkSumSerial(const X* part, int c0, int c1) {
  X acc = (X)0;
#pragma unroll 8
  for (int c = c0; c < c1; c++) acc = acc + part[c];
  return acc;
}
template <typename X>
__device__ __forceinline__ X chunkSumWave(const X* part, int c0, int c1, int lane) {
  X acc = (X)0;
  for (int c = c0 + lane; c < c1; c += 64) acc = acc + part[c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc = acc + __shfl_down(acc, off, 64);
  return acc;  // lane 0's is the total
}

// The chunk partials of every system summed (chunkSum*) and the scalar
// epilogue `op` applied (the single-system finalize); then the block counts
// the systems still running and sets glob->done when none is (one block;
// the counts are integers, so the result is order-free).
template <typename X>
__global__ __launch_bounds__(1024) void batchFinKernel(BatchArgs b, int op, int K) {
  __shared__ int nActive, nDue;
  if (op != kFinInit && b.glob->done) return;
  // device-side refinement: the restart has nothing to do when no outer step
  // is due
  if (b.deviceOuter && op == kFinInit && b.glob->iter == 0) return;
  if (threadIdx.x == 0) {
    nActive = 0;
    nDue = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (int)(blockDim.x >> 6);
  const X* part = static_cast<const X*>(b.partials);
  int mine = 0;
  // device-side refinement: the initialisation restarts the inner solve of
  // the systems whose outer step is due (phase 1, inner state finished)
  const bool restart = b.deviceOuter && op == kFinInit;
  auto skip = [&](int s, const DevState* st) {
    if (restart) return b.phase[s] != 1;
    return st->done && op != kFinInit;
  };
  auto finish = [&](int s, DevState* st, const double (&sum)[2]) {
    if (op == kFinInit && st->done && !restart) {
      // a system excluded from this solve (its outer iteration has
      // converged) stays finished
    } else {
      finalize<X>(st, op, sum[0], sum[1]);
    }
    if (b.phase) {
      if (restart && !st->done) {
        b.phase[s] = 0;
        atomicAdd(&nDue, -1);
      } else if (!restart && st->done && b.phase[s] == 0) {
        b.phase[s] = b.deviceOuter ? 1 : 2;
        if (b.deviceOuter) atomicAdd(&nDue, 1);
      }
    }
    mine += st->done ? 0 : 1;
  };
  // systems with few chunks: one thread each
  for (int s = threadIdx.x; s < b.S; s += blockDim.x) {
    const int c0 = b.sysChunk[s], c1 = b.sysChunk[s + 1];
    if (c1 - c0 > kSerialChunks) continue;
    DevState* st = b.st + s;
    if (skip(s, st)) continue;
    double sum[2] = {0.0, 0.0};
    for (int k = 0; k < K; k++) sum[k] = (double)chunkSumSerial(part + k * b.C, c0, c1);
    finish(s, st, sum);
  }
  // the others: one wave each
  for (int s = wave; s < b.S; s += nw) {
    const int c0 = b.sysChunk[s], c1 = b.sysChunk[s + 1];
    if (c1 - c0 <= kSerialChunks) continue;
    DevState* st = b.st + s;
    if (skip(s, st)) continue;  // the same for every lane of the wave
    double sum[2] = {0.0, 0.0};
    for (int k = 0; k < K; k++) sum[k] = (double)chunkSumWave(part + k * b.C, c0, c1, lane);
    if (lane == 0) finish(s, st, sum);
  }
  // glob->done: no system running (device-side refinement: set by the outer
  // step instead, and glob->iter counts the systems whose outer step is due)
  if (!b.deviceOuter && (op == kFinInit || op == kFinUpdate || op == kFinOmega)) {
    atomicAdd(&nActive, mine);
    __syncthreads();
    if (threadIdx.x == 0) b.glob->done = nActive == 0 ? 1 : 0;
  }
  if (b.deviceOuter) {
    __syncthreads();
    if (threadIdx.x == 0 && nDue != 0) b.glob->iter = b.glob->iter + nDue;
  }
}

// p = r + beta (p - omega v) per active system (iteration > 0), as the
// single-system p-update (pbicgstab/Solver.cpp:98-101).
template <typename X>
__global__ __launch_bounds__(kBatchThreads) void batchUpdatePKernel(BatchArgs b, const X* r,
                                                                    X* p, const X* v) {
  if (b.glob->done) return;
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  const DevState* st = b.st + ch.sys;
  if (st->done || st->iter == 0) return;
  const X beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
  const X omega = (X)st->omega;
  for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads)
    p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// s = r - alpha v per active system (pbicgstab/Solver.cpp:127-129).
template <typename X>
__global__ __launch_bounds__(kBatchThreads) void batchUpdateSKernel(BatchArgs b, const X* r,
                                                                    X* s, const X* v) {
  if (b.glob->done) return;
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  const DevState* st = b.st + ch.sys;
  if (st->done) return;
  const X alpha = (X)st->alpha;
  for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads) s[i] = r[i] - alpha * v[i];
}

// x += alpha y + omega z ; r = s - omega t per active system, with the
// chunk partials of r.r and r0.r (pbicgstab/Solver.cpp:162-172, :91).
template <typename X>
__global__ __launch_bounds__(kBatchThreads) void batchUpdateXRKernel(
    BatchArgs b, X* x, X* r, const X* y, const X* z, const X* s, const X* t, const X* r0) {
  __shared__ X lds[2 * kBatchThreads / 64];
  if (b.glob->done) return;
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  const DevState* st = b.st + ch.sys;
  if (st->done) return;
  const X alpha = (X)st->alpha, omega = (X)st->omega;
  X acc[2] = {(X)0, (X)0};
  for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads) {
    x[i] = (x[i] + alpha * y[i]) + omega * z[i];
    const X ri = s[i] - omega * t[i];
    r[i] = ri;
    acc[0] = acc[0] + ri * ri;
    acc[1] = acc[1] + r0[i] * ri;
  }
  blockSum<X, 2, kBatchThreads>(acc, lds);
  if (threadIdx.x == 0) {
    static_cast<X*>(b.partials)[blockIdx.x] = acc[0];
    static_cast<X*>(b.partials)[b.C + blockIdx.x] = acc[1];
  }
}

// Start of a solve per active system: x = 0, r = r0 = p = rhs
// (pbicgstab/Solver.cpp:66-83 with a zero initial guess), and the chunk
// partials of r.r.
template <typename X>
__global__ __launch_bounds__(kBatchThreads) void batchStartKernel(BatchArgs b, const X* rhs, X* x,
                                                                  X* r, X* r0, X* p) {
  __shared__ X lds[kBatchThreads / 64];
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  if (b.st[ch.sys].done) return;
  X acc[1] = {(X)0};
  for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads) {
    const X v = rhs[i];
    x[i] = (X)0;
    r[i] = v;
    r0[i] = v;
    p[i] = v;
    acc[0] = acc[0] + v * v;
  }
  blockSum<X, 1, kBatchThreads>(acc, lds);
  if (threadIdx.x == 0) static_cast<X*>(b.partials)[blockIdx.x] = acc[0];
}

// phi += d per system that is still refining (iterative-refinement/Solver.cpp:103).
__global__ __launch_bounds__(kBatchThreads) void batchAccumulateKernel(BatchArgs b,
                                                                       const DevState* outer,
                                                                       double* phi,
                                                                       const float* d) {
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  if (outer[ch.sys].done) return;
  for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads)
    phi[i] = phi[i] + (double)d[i];
}

// One system's outer (MPIR) step from the norm rF of its unrounded FP64
// residual: on the first call the initial norm (r0n), afterwards one
// refinement step is counted and convergence tested against it, with the
// loop condition of iterative-refinement/Solver.cpp:89-112. The inner state
// of a system that has finished refining is marked finished, so the next
// inner solve skips it.
__device__ __forceinline__ void outerStep(BatchArgs& b, DevState* o, int s, double rF,
                                          int first) {
  if (first) {
    o->r0n = rF;
    o->rF = rF;
    o->iter = 0;
    o->converged = 0;
    o->done = o->maxIt <= 0 ? 1 : 0;
  } else if (!o->done) {
    o->iter = o->iter + 1;
    o->rF = rF;
    o->converged = convergedTest(o, rF) ? 1 : 0;
    o->done = (o->converged && o->iter >= o->minIt) || o->iter >= o->maxIt ? 1 : 0;
  }
  b.st[s].done = o->done;
}

// Outer residual norms of every system, the chunk partials summed as in
// batchFinKernel (one thread per system with few chunks, one wave per
// system with more).
__global__ __launch_bounds__(1024) void batchOuterKernel(BatchArgs b, DevState* outer,
                                                         int first) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (int)(blockDim.x >> 6);
  const double* part = static_cast<const double*>(b.partials);
  for (int q = 0; q < 2; q++) {  // pass 0: thread per system; pass 1: wave per system
    for (int s = q == 0 ? (int)threadIdx.x : wave; s < b.S;
         s += q == 0 ? (int)blockDim.x : nw) {
      const int c0 = b.sysChunk[s], c1 = b.sysChunk[s + 1];
      if ((c1 - c0 > kSerialChunks) != (q == 1)) continue;
      const double acc = q == 0 ? chunkSumSerial(part, c0, c1) : chunkSumWave(part, c0, c1, lane);
      if (q == 1 && lane != 0) continue;
      outerStep(b, outer + s, s, sqrt(acc), first);
    }
  }
}

// Inner states of systems that have finished refining are marked finished.
__global__ __launch_bounds__(1024) void batchOuterMaskKernel(BatchArgs b, const DevState* outer) {
  for (int s = threadIdx.x; s < b.S; s += blockDim.x)
    if (outer[s].done) b.st[s].done = 1;
}

// FP64 chunk partials of the squared unrounded residual (outer norm).
__global__ __launch_bounds__(kBatchThreads) void batchNorm64Kernel(BatchArgs b,
                                                                   const DevState* outer,
                                                                   const double* r64) {
  __shared__ double lds[kBatchThreads / 64];
  const ChunkRange ch = chunkOf(b, blockIdx.x);
  double acc[1] = {0.0};
  if (!outer[ch.sys].done)
    for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads) acc[0] = acc[0] + r64[i] * r64[i];
  blockSum<double, 1, kBatchThreads>(acc, lds);
  if (threadIdx.x == 0) static_cast<double*>(b.partials)[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(1024) void batchPhaseSyncKernel(BatchArgs b) {
  for (int s = threadIdx.x; s < b.S; s += blockDim.x) b.phase[s] = b.st[s].done ? 2 : 0;
}

// ---- refinement step on the device, per system in phase 1 -------------------
// phi += d (iterative-refinement/Solver.cpp:103)
// (grid-stride over chunks: the grid is small, so that a step with no outer
// step due costs little)
__global__ __launch_bounds__(kBatchThreads) void batchPhaseAccumulateKernel(BatchArgs b,
                                                                            double* phi,
                                                                            const float* d) {
  if (b.glob->done || b.glob->iter == 0) return;
  for (int c = blockIdx.x; c < b.C; c += gridDim.x) {
    const ChunkRange ch = chunkOf(b, c);
    if (b.phase[ch.sys] != 1) continue;
    for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads)
      phi[i] = phi[i] + (double)d[i];
  }
}

// FP64 chunk partials of the squared unrounded residual (0 for chunks of
// systems in another phase)
__global__ __launch_bounds__(kBatchThreads) void batchPhaseNorm64Kernel(BatchArgs b,
                                                                        const double* r64) {
  __shared__ double lds[kBatchThreads / 64];
  if (b.glob->done || b.glob->iter == 0) return;
  for (int c = blockIdx.x; c < b.C; c += gridDim.x) {
    const ChunkRange ch = chunkOf(b, c);
    if (b.phase[ch.sys] != 1) continue;  // uniform in the block
    double acc[1] = {0.0};
    for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads)
      acc[0] = acc[0] + r64[i] * r64[i];
    blockSum<double, 1, kBatchThreads>(acc, lds);
    if (threadIdx.x == 0) static_cast<double*>(b.partials)[c] = acc[0];
  }
}

// Outer step of the systems in phase 1 (iterative-refinement/Solver.cpp:
// 89-112): the norm of the residual, the inner iterations of the finished
// inner solve counted, one refinement iteration counted and convergence
// tested (first: the initial residual); a finished system goes to phase 2.
// Then glob->done = every system finished.
__global__ __launch_bounds__(1024) void batchPhaseOuterKernel(BatchArgs b, DevState* outer,
                                                              int first) {
  __shared__ int nLeft, nGone;
  if (b.glob->done || b.glob->iter == 0) return;
  if (threadIdx.x == 0) {
    nLeft = 0;
    nGone = 0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (int)(blockDim.x >> 6);
  const double* part = static_cast<const double*>(b.partials);
  for (int q = 0; q < 2; q++) {  // pass 0: thread per system; pass 1: wave per system
    for (int s = q == 0 ? (int)threadIdx.x : wave; s < b.S;
         s += q == 0 ? (int)blockDim.x : nw) {
      const int c0 = b.sysChunk[s], c1 = b.sysChunk[s + 1];
      if ((c1 - c0 > kSerialChunks) != (q == 1)) continue;
      if (b.phase[s] != 1) continue;  // the same for every lane of the wave
      const double acc = q == 0 ? chunkSumSerial(part, c0, c1) : chunkSumWave(part, c0, c1, lane);
      if (q == 1 && lane != 0) continue;
      const double rF = sqrt(acc);
      DevState* o = outer + s;
      if (first) {
        o->r0n = rF;
        o->rF = rF;
        o->iter = 0;
        o->converged = 0;
        o->innerTotal = 0;
        o->done = o->maxIt <= 0 ? 1 : 0;
      } else {
        o->innerTotal = o->innerTotal + b.st[s].iter;
        o->iter = o->iter + 1;
        o->rF = rF;
        o->converged = convergedTest(o, rF) ? 1 : 0;
        o->done = (o->converged && o->iter >= o->minIt) || o->iter >= o->maxIt ? 1 : 0;
      }
      if (o->done) {
        b.phase[s] = 2;
        b.st[s].done = 1;
        atomicAdd(&nGone, 1);
      }
    }
  }
  __syncthreads();
  int mine = 0;
  for (int s = threadIdx.x; s < b.S; s += blockDim.x) mine += b.phase[s] != 2 ? 1 : 0;
  atomicAdd(&nLeft, mine);
  __syncthreads();
  if (threadIdx.x == 0) {
    b.glob->done = nLeft == 0 ? 1 : 0;
    b.glob->iter = b.glob->iter - nGone;
  }
}

// Inner restart of the systems in phase 1: x = 0, r = r0 = p = rhs (the
// rounded residual) and the chunk partials of r.r for the initialisation.
__global__ __launch_bounds__(kBatchThreads) void batchPhaseStartKernel(BatchArgs b,
                                                                       const float* rhs, float* x,
                                                                       float* r, float* r0,
                                                                       float* p) {
  __shared__ float lds[kBatchThreads / 64];
  if (b.glob->done || b.glob->iter == 0) return;
  for (int c = blockIdx.x; c < b.C; c += gridDim.x) {
    const ChunkRange ch = chunkOf(b, c);
    if (b.phase[ch.sys] != 1) continue;  // uniform in the block
    float acc[1] = {0.0f};
    for (int i = ch.r0 + threadIdx.x; i < ch.r1; i += kBatchThreads) {
      const float v = rhs[i];
      x[i] = 0.0f;
      r[i] = v;
      r0[i] = v;
      p[i] = v;
      acc[0] = acc[0] + v * v;
    }
    blockSum<float, 1, kBatchThreads>(acc, lds);
    if (threadIdx.x == 0) static_cast<float*>(b.partials)[c] = acc[0];
  }
}

// ---------------------------------------------------------------------------
void launchBatchDot1(const BatchArgs& b, const void* a0, const void* b0, hipStream_t s) {
  BatchDot<1> d;
  d.a[0] = a0;
  d.b[0] = b0;
  hipLaunchKernelGGL((batchDotKernel<float, 1>), dim3(b.C), dim3(kBatchThreads), 0, s, b, d);
  GPH_HIP(hipGetLastError());
}
void launchBatchDot2(const BatchArgs& b, const void* a0, const void* b0, const void* a1,
                     const void* b1, hipStream_t s) {
  BatchDot<2> d;
  d.a[0] = a0;
  d.b[0] = b0;
  d.a[1] = a1;
  d.b[1] = b1;
  hipLaunchKernelGGL((batchDotKernel<float, 2>), dim3(b.C), dim3(kBatchThreads), 0, s, b, d);
  GPH_HIP(hipGetLastError());
}
void launchBatchFin(const BatchArgs& b, int op, int k, hipStream_t s) {
  hipLaunchKernelGGL(batchFinKernel<float>, dim3(1), dim3(1024), 0, s, b, op, k);
  GPH_HIP(hipGetLastError());
}
void launchBatchUpdateP(const BatchArgs& b, const float* r, float* p, const float* v,
                        hipStream_t s) {
  hipLaunchKernelGGL(batchUpdatePKernel<float>, dim3(b.C), dim3(kBatchThreads), 0, s, b, r, p, v);
  GPH_HIP(hipGetLastError());
}
void launchBatchUpdateS(const BatchArgs& b, const float* r, float* sv, const float* v,
                        hipStream_t s) {
  hipLaunchKernelGGL(batchUpdateSKernel<float>, dim3(b.C), dim3(kBatchThreads), 0, s, b, r, sv, v);
  GPH_HIP(hipGetLastError());
}
void launchBatchUpdateXR(const BatchArgs& b, float* x, float* r, const float* y, const float* z,
                         const float* sv, const float* t, const float* r0, hipStream_t s) {
  hipLaunchKernelGGL(batchUpdateXRKernel<float>, dim3(b.C), dim3(kBatchThreads), 0, s, b, x, r, y,
                     z, sv, t, r0);
  GPH_HIP(hipGetLastError());
}
void launchBatchStart(const BatchArgs& b, const float* rhs, float* x, float* r, float* r0,
                      float* p, hipStream_t s) {
  hipLaunchKernelGGL(batchStartKernel<float>, dim3(b.C), dim3(kBatchThreads), 0, s, b, rhs, x, r,
                     r0, p);
  GPH_HIP(hipGetLastError());
}
void launchBatchAccumulate(const BatchArgs& b, const DevState* outer, double* phi,
                           const float* d, hipStream_t s) {
  hipLaunchKernelGGL(batchAccumulateKernel, dim3(b.C), dim3(kBatchThreads), 0, s, b, outer, phi,
                     d);
  GPH_HIP(hipGetLastError());
}
void launchBatchOuter(const BatchArgs& b, DevState* outer, const double* r64, int first,
                      hipStream_t s) {
  hipLaunchKernelGGL(batchNorm64Kernel, dim3(b.C), dim3(kBatchThreads), 0, s, b, outer, r64);
  hipLaunchKernelGGL(batchOuterKernel, dim3(1), dim3(1024), 0, s, b, outer, first);
  GPH_HIP(hipGetLastError());
}

// grid of the outer-step chunk kernels (grid-stride over the chunks)
static int phaseGrid(const BatchArgs& b) { return b.C < 1024 ? b.C : 1024; }

void launchBatchPhaseSync(const BatchArgs& b, hipStream_t s) {
  if (!b.phase) return;
  hipLaunchKernelGGL(batchPhaseSyncKernel, dim3(1), dim3(1024), 0, s, b);
  GPH_HIP(hipGetLastError());
}
void launchBatchPhaseAccumulate(const BatchArgs& b, double* phi, const float* d,
                                hipStream_t s) {
  hipLaunchKernelGGL(batchPhaseAccumulateKernel, dim3(phaseGrid(b)), dim3(kBatchThreads), 0, s, b,
                     phi, d);
  GPH_HIP(hipGetLastError());
}
void launchBatchPhaseOuter(const BatchArgs& b, DevState* outer, const double* r64, int first,
                           hipStream_t s) {
  hipLaunchKernelGGL(batchPhaseNorm64Kernel, dim3(phaseGrid(b)), dim3(kBatchThreads), 0, s, b, r64);
  hipLaunchKernelGGL(batchPhaseOuterKernel, dim3(1), dim3(1024), 0, s, b, outer, first);
  GPH_HIP(hipGetLastError());
}
void launchBatchPhaseRestart(const BatchArgs& b, const float* rhs, float* x, float* r, float* r0,
                             float* p, hipStream_t s) {
  hipLaunchKernelGGL(batchPhaseStartKernel, dim3(phaseGrid(b)), dim3(kBatchThreads), 0, s, b, rhs, x,
                     r, r0, p);
  hipLaunchKernelGGL(batchFinKernel<float>, dim3(1), dim3(1024), 0, s, b, (int)kFinInit, 1);
  GPH_HIP(hipGetLastError());
}

void launchBatchOuterMask(const BatchArgs& b, const DevState* outer, hipStream_t s) {
  hipLaunchKernelGGL(batchOuterMaskKernel, dim3(1), dim3(1024), 0, s, b, outer);
  GPH_HIP(hipGetLastError());
}

}  // namespace gph
