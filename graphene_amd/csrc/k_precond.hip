// Tile-local preconditioners for gfx950: ILU(0) factorization and
// substitution, Gauss-Seidel sweeps.
//
// The reference runs these per IPU tile on the tile's non-halo block
// (block-Jacobi across tiles): ILU factorization lg/matrix/solver/ilu/Solver.cpp:134-292,
// substitution :367-550, Gauss-Seidel lg/matrix/solver/gauss-seidel/Solver.cpp:45-86
// (arithmetic of codelet lg/codelet/ipu/matrix/solvers/gauss-seidel.cpp:41-61).
// Here one workgroup owns one tile and walks the tile's level sets (rows of
// a level are independent), with a barrier between levels; the tile's
// solution vector sits in LDS when it fits. Per row the arithmetic is the
// reference's sequential row body, so results equal the sequential
// substitution bit for bit.
#include "kernels.hpp"

namespace gph {

size_t tileSharedBytes(const TileArgs& t, DType xt) {
  return (size_t)t.maxTileRows * dtypeSize(xt);
}

static size_t maxDynamicShared() { return 160 * 1024 - 4096; }

// ---------------------------------------------------------------------------
// ILU(0) factorization (Saad Alg. 10.4, IKJ), per tile.
// Row i: for each k = col(ik) < i in-tile: a_ik /= a_kk; a_ij -= a_ik*a_kj
// for the later in-tile entries j of row i (a_kj = 0 if absent); then
// a_ii -= a_ik*a_ki. ilu/Solver.cpp:167-218. Afterwards diag = 1/diag
// (:290-291). DILU (diagonalBased, :191-216): the off-diagonals stay A's
// and a_ii -= (a_ik*a_ki)/a_kk.
//
// A row's entries are its in-tile columns ascending, then its halo columns;
// the scans stop at the first halo column, so only that in-tile prefix
// takes part. Rows of one forward level depend only on rows of lower
// levels, whose factor rows are final: one launch per level over the
// level's rows of every tile (a (row, tile) list), one wave per row. Lane e holds row i's entry e. For a batch of lower entries the wave
// loads the rows k at once (lane q: entry q of row k), finds every a_kj by
// a binary search of row k's in-tile columns in LDS and a shuffle, and a_ki
// by a ballot; then it applies the updates of the batch's entries in order
// (the subtraction order of the sequential loop, a_kj = 0 subtracted too).
// Rows longer than 64 entries, or with a row k that is, run the sequential
// loop on one lane.
// ---------------------------------------------------------------------------
constexpr int kFactorWaves = 4;  // rows (waves) per block
constexpr int kFactorBatch = 4;  // rows k loaded together (128^3: 8 rows took 7.8 ms in FP64, 4 rows 5.8)

template <typename FT>
__device__ __forceinline__ FT readLane(FT v, int l) {
  if constexpr (sizeof(FT) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

__device__ __forceinline__ uint64_t waveBallot(bool p) { return __ballot(p); }
// Orders the wave's LDS writes before its later LDS reads (and reads before
// later writes) across lanes.
__device__ __forceinline__ void waveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ int firstSet(uint64_t m) { return m ? __builtin_ctzll(m) : 64; }

// The sequential row loop (one lane).
template <typename FT>
__device__ void factorRowSerial(const int32_t* __restrict__ rowPtr, const int32_t* __restrict__ col,
                                FT* off, FT* diag, int t0, int t1, int i, int dilu) {
  const int iEnd = rowPtr[i + 1];
  for (int ik = rowPtr[i]; ik < iEnd; ik++) {
    const int k = col[ik];
    if (k >= i || k < t0) break;  // upper part or halo reached
    const FT a_kk = diag[k];
    const int kEnd = rowPtr[k + 1];
    if (dilu) {
      FT a_ki = 0;
      for (int qq = rowPtr[k]; qq < kEnd; qq++) {
        const int c = col[qq];
        if (c < t0 || c >= t1 || c > i) break;
        if (c == i) a_ki = off[qq];
      }
      diag[i] = diag[i] - off[ik] * a_ki / a_kk;
      continue;
    }
    const FT a_ik = off[ik] / a_kk;
    off[ik] = a_ik;
    int qk = rowPtr[k];
    for (int ij = ik + 1; ij < iEnd; ij++) {
      const int j = col[ij];
      if (j < t0 || j >= t1) break;  // stop at halo coefficients
      while (qk < kEnd) {
        const int c = col[qk];
        if (c < t0 || c >= t1 || c >= j) break;
        qk++;
      }
      const FT a_kj = (qk < kEnd && col[qk] == j) ? off[qk] : (FT)0;
      off[ij] = off[ij] - a_ik * a_kj;
    }
    FT a_ki = 0;
    for (int qq = rowPtr[k]; qq < kEnd; qq++) {
      const int c = col[qq];
      if (c < t0 || c >= t1 || c > i) break;
      if (c == i) a_ki = off[qq];
    }
    diag[i] = diag[i] - a_ik * a_ki;
  }
}

template <typename FT>
__global__ __launch_bounds__(64 * kFactorWaves) void iluFactorLevelKernel(FactorArgs f, FT* off,
                                                                          FT* diag, int dilu,
                                                                          int64_t base,
                                                                          int count) {
  __shared__ int kcols[kFactorWaves][kFactorBatch][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int w = (int)blockIdx.x * kFactorWaves + wave;
  if (w >= count) return;  // the wave's exit; the kernel has no block barrier
  const int i = f.list[2 * (base + w)], tile = f.list[2 * (base + w) + 1];
  const int t0 = f.tileStart[tile], t1 = f.tileStart[tile + 1];
  const int32_t* __restrict__ rowPtr = f.rowPtr;
  const int32_t* __restrict__ col = f.col;
  const int rb = rowPtr[i], n = rowPtr[i + 1] - rb;
  // lower in-tile entries lead the row; lane m < mL holds row k = c's start,
  // length and diagonal
  const int c = lane < n ? col[rb + lane] : -1;
  FT v = lane < n ? off[rb + lane] : (FT)0;
  const bool low = lane < n && c >= t0 && c < i;
  const int mL = firstSet(waveBallot(!low));
  const int kb = low ? rowPtr[c] : 0;
  const int nk = low ? rowPtr[c + 1] - kb : 0;
  const FT dk = low ? diag[c] : (FT)1;
  if (n > 64 || waveBallot(nk > 64)) {
    if (lane == 0) factorRowSerial(rowPtr, col, off, diag, t0, t1, i, dilu);
    return;
  }
  const bool inT = lane < n && c >= t0 && c < t1;
  const int P = firstSet(waveBallot(!inT));  // in-tile prefix of row i
  FT di = diag[i];
  int* kc = &kcols[wave][0][0];
  for (int m0 = 0; m0 < mL; m0 += kFactorBatch) {
    const int nb = min(kFactorBatch, mL - m0);
    FT vk[kFactorBatch];
    {
      int ck[kFactorBatch];
#pragma unroll
      for (int b = 0; b < kFactorBatch; b++) {
        ck[b] = -1;
        vk[b] = (FT)0;
        if (b < nb) {
          const int k0 = __builtin_amdgcn_readlane(kb, m0 + b);
          const int len = __builtin_amdgcn_readlane(nk, m0 + b);
          if (lane < len) {
            ck[b] = col[k0 + lane];
            vk[b] = off[k0 + lane];
          }
        }
      }
      // row k's in-tile prefix in LDS, ascending, padded with INT_MAX
#pragma unroll
      for (int b = 0; b < kFactorBatch; b++) {
        const int Pk = firstSet(waveBallot(!(ck[b] >= t0 && ck[b] < t1)));
        kc[b * 64 + lane] = lane < Pk ? ck[b] : INT_MAX;
      }
    }
    waveSync();
    int pos[kFactorBatch];
#pragma unroll
    for (int b = 0; b < kFactorBatch; b++) pos[b] = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
#pragma unroll
      for (int b = 0; b < kFactorBatch; b++)
        if (kc[b * 64 + pos[b] + s - 1] < c) pos[b] += s;
    }
    FT akj[kFactorBatch], aki[kFactorBatch];
#pragma unroll
    for (int b = 0; b < kFactorBatch; b++) {
      const bool hit = kc[b * 64 + pos[b]] == c;
      const FT g = __shfl(vk[b], pos[b], 64);
      akj[b] = hit ? g : (FT)0;
      const int li = firstSet(waveBallot(kc[b * 64 + lane] == i));
      aki[b] = li < 64 ? readLane(vk[b], li) : (FT)0;
    }
    waveSync();  // LDS reads before the next batch's writes
#pragma unroll
    for (int b = 0; b < kFactorBatch; b++) {
      const int m = m0 + b;
      const FT akk = readLane(dk, m0 + min(b, nb - 1));
      if (b < nb && dilu) {
        di = di - readLane(v, m) * aki[b] / akk;
      } else if (b < nb) {
        const FT aik = readLane(v, m) / akk;
        if (lane == m) v = aik;
        if (lane > m && lane < P) v = v - aik * akj[b];
        di = di - aik * aki[b];
      }
    }
  }
  if (!dilu && lane < P) off[rb + lane] = v;
  if (lane == 0) diag[i] = di;
}

template <typename FT>
__global__ void invertDiagKernel(FT* diag, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = (FT)1 / diag[i];
}

template <typename FT>
static void factorLevels(const FactorArgs& f, const std::vector<int64_t>& levelOff, int64_t n,
                         FT* off, FT* diag, bool dilu, hipStream_t s) {
  for (size_t L = 0; L + 1 < levelOff.size(); L++) {
    const int64_t count = levelOff[L + 1] - levelOff[L];
    if (count <= 0) continue;
    const int64_t blocks = (count + kFactorWaves - 1) / kFactorWaves;
    hipLaunchKernelGGL(iluFactorLevelKernel<FT>, dim3((unsigned)blocks), dim3(64 * kFactorWaves),
                       0, s, f, off, diag, dilu ? 1 : 0, levelOff[L], (int)count);
  }
  if (n > 0)
    hipLaunchKernelGGL(invertDiagKernel<FT>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       diag, n);
  GPH_HIP(hipGetLastError());
}

void launchIluFactor(const FactorArgs& f, const std::vector<int64_t>& levelOff, int64_t n,
                     void* off, void* diag, DType ft, bool dilu, hipStream_t s) {
  if (n > (int64_t)INT32_MAX) fail(kUnsupported, "ILU factorization: too many rows");
  if (ft == DType::F32)
    factorLevels(f, levelOff, n, static_cast<float*>(off), static_cast<float*>(diag), dilu, s);
  else if (ft == DType::F64)
    factorLevels(f, levelOff, n, static_cast<double*>(off), static_cast<double*>(diag), dilu, s);
  else
    fail(kUnsupported, "ILU factorization type must be float32 or float64");
}

// ---------------------------------------------------------------------------
// ILU(0) substitution per tile (ilu/Solver.cpp:380-442, :526-548):
// forward  t = b_i; t -= L_a*x_j for in-tile j < i in stored order; x_i = t
// backward t = 0;  t -= U_a*x_j for in-tile j > i, entries descending;
//          x_i = (x_i + t) * invD_i
// Forward levels ascending, then backward levels (levels of the upper
// triangle) ascending. The right-hand side can be formed on the fly:
// kPreP: p = r + beta*(p - omega*v) (iter > 0), kPreS: s = r - alpha*v.
// ---------------------------------------------------------------------------
template <typename FT, typename X, bool USE_LDS>
__global__ __launch_bounds__(kTileThreads) void iluApplyKernel(
    TileArgs t, const FT* __restrict__ lu, const InvDiag<FT>* __restrict__ invD,
    const X* __restrict__ rhs, X* __restrict__ out, PreOp pre, DevState* st,
    int checkDone, int dilu) {
  extern __shared__ unsigned char smemRaw[];
  if (checkDone && st->done) return;
  const int tile = blockIdx.x;
  const int t0 = t.tileStart[tile], t1 = t.tileStart[tile + 1];
  X* xs = USE_LDS ? reinterpret_cast<X*>(smemRaw) : out + t0;
  const int32_t* __restrict__ rowPtr = t.rowPtr;
  const int32_t* __restrict__ col = t.col;

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

  const int* lp = t.lvlPtr + t.lvlPtrOff[tile];
  const int nL = t.lvlPtrOff[tile + 1] - t.lvlPtrOff[tile] - 1;
  for (int L = 0; L < nL; L++) {
    for (int q = lp[L] + threadIdx.x; q < lp[L + 1]; q += kTileThreads) {
      const int li = t.lvlRows[t0 + q];
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
      const int e = rowPtr[i + 1];
      for (int a = rowPtr[i]; a < e; a++) {
        const int j = col[a];
        if (i < j) break;
        if (j < t0 || j >= t1) break;
        s = s - cvt<X>(lu[a]) * xs[j - t0];
      }
      xs[li] = dilu ? s * cvt<X>(invD[i]) : s;
    }
    __syncthreads();
  }
  const int* up = t.ulvlPtr + t.ulvlPtrOff[tile];
  const int nU = t.ulvlPtrOff[tile + 1] - t.ulvlPtrOff[tile] - 1;
  for (int L = 0; L < nU; L++) {
    for (int q = up[L] + threadIdx.x; q < up[L + 1]; q += kTileThreads) {
      const int li = t.ulvlRows[t0 + q];
      const int i = t0 + li;
      X s = 0;
      const int b = rowPtr[i];
      for (int a = rowPtr[i + 1] - 1; a >= b; a--) {
        const int j = col[a];
        if (j < t0 || j >= t1) continue;
        if (j < i) break;
        s = s - cvt<X>(lu[a]) * xs[j - t0];
      }
      xs[li] = dilu ? xs[li] + s * cvt<X>(invD[i]) : (xs[li] + s) * cvt<X>(invD[i]);
    }
    __syncthreads();
  }
  if (USE_LDS)
    for (int q = threadIdx.x; q < t1 - t0; q += kTileThreads)
      out[t0 + q] = xs[q];
}

template <typename FT, typename X>
static void iluApplyTyped(const TileArgs& t, const void* lu, const void* invD,
                          const void* rhs, void* out, const PreOp& pre,
                          DevState* st, int checkDone, int dilu, hipStream_t s) {
  const size_t shm = (size_t)t.maxTileRows * sizeof(X);
  if (shm <= maxDynamicShared()) {
    auto k = iluApplyKernel<FT, X, true>;
    if (shm > 64 * 1024)
      GPH_HIP(hipFuncSetAttribute((const void*)k,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm));
    hipLaunchKernelGGL(k, dim3(t.T), dim3(kTileThreads), shm, s, t,
                       static_cast<const FT*>(lu),
                       static_cast<const InvDiag<FT>*>(invD),
                       static_cast<const X*>(rhs), static_cast<X*>(out), pre,
                       st, checkDone, dilu);
  } else {
    hipLaunchKernelGGL((iluApplyKernel<FT, X, false>), dim3(t.T),
                       dim3(kTileThreads), 0, s, t,
                       static_cast<const FT*>(lu),
                       static_cast<const InvDiag<FT>*>(invD),
                       static_cast<const X*>(rhs), static_cast<X*>(out), pre,
                       st, checkDone, dilu);
  }
  GPH_HIP(hipGetLastError());
}

void launchIluApply(const TileArgs& t, const void* lu, const void* invD,
                    DType ft, const void* rhs, void* out, DType xt,
                    const PreOp& pre, DevState* st, int checkDone, bool dilu,
                    hipStream_t s) {
  if (t.T <= 0) return;
  const int d = dilu ? 1 : 0;
  if (ft == DType::F32 && xt == DType::F32)
    iluApplyTyped<float, float>(t, lu, invD, rhs, out, pre, st, checkDone, d, s);
  else if (ft == DType::F16 && xt == DType::F32)
    iluApplyTyped<__half, float>(t, lu, invD, rhs, out, pre, st, checkDone, d, s);
  else if (ft == DType::F64 && xt == DType::F64)
    iluApplyTyped<double, double>(t, lu, invD, rhs, out, pre, st, checkDone, d, s);
  else if (ft == DType::F32 && xt == DType::F64)
    iluApplyTyped<float, double>(t, lu, invD, rhs, out, pre, st, checkDone, d, s);
  else
    fail(kUnsupported, std::string("unsupported ILU apply types: factors ") +
                           dtypeName(ft) + ", vectors " + dtypeName(xt));
}

// ---------------------------------------------------------------------------
// Gauss-Seidel sweep per tile. Forward: rows in local order,
// x_i = (b_i - sum_{j != i} a_ij x_j) / a_ii with the new value for in-tile
// j < i and the previous value (xin) otherwise (halo included). Backward:
// reverse local order, new values for in-tile j > i.
// ---------------------------------------------------------------------------
template <typename VT, typename X, bool BACKWARD, bool USE_LDS>
__global__ __launch_bounds__(kTileThreads) void gsSweepKernel(
    TileArgs t, const VT* __restrict__ off, const VT* __restrict__ diag,
    const X* __restrict__ b, const X* __restrict__ xin, X* __restrict__ xout,
    DevState* st, int checkDone) {
  extern __shared__ unsigned char smemRaw[];
  if (checkDone && st->done) return;
  const int tile = blockIdx.x;
  const int t0 = t.tileStart[tile], t1 = t.tileStart[tile + 1];
  X* xs = USE_LDS ? reinterpret_cast<X*>(smemRaw) : xout + t0;
  const int32_t* __restrict__ rowPtr = t.rowPtr;
  const int32_t* __restrict__ col = t.col;
  const int* lp = BACKWARD ? t.ulvlPtr + t.ulvlPtrOff[tile]
                           : t.lvlPtr + t.lvlPtrOff[tile];
  const int32_t* rows = BACKWARD ? t.ulvlRows : t.lvlRows;
  const int nL = BACKWARD
                     ? t.ulvlPtrOff[tile + 1] - t.ulvlPtrOff[tile] - 1
                     : t.lvlPtrOff[tile + 1] - t.lvlPtrOff[tile] - 1;
  for (int L = 0; L < nL; L++) {
    for (int q = lp[L] + threadIdx.x; q < lp[L + 1]; q += kTileThreads) {
      const int li = rows[t0 + q];
      const int i = t0 + li;
      X s = b[i];
      const int e = rowPtr[i + 1];
      auto xOf = [&](int j) {
        const bool inTile = j >= t0 && j < t1;
        const bool fresh = inTile && (BACKWARD ? (j > i) : (j < i));
        return fresh ? xs[j - t0] : xin[j];
      };
      int a = rowPtr[i];
      // 8 entries per step: the column and value loads, then the x reads,
      // then the subtractions in stored order
      for (; a + 8 <= e; a += 8) {
        int c[8];
        VT v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          c[k] = col[a + k];
          v[k] = off[a + k];
        }
        X xv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) xv[k] = xOf(c[k]);
#pragma unroll
        for (int k = 0; k < 8; k++) s = s - cvt<X>(v[k]) * xv[k];
      }
      for (; a < e; a++) s = s - cvt<X>(off[a]) * xOf(col[a]);
      xs[li] = s / cvt<X>(diag[i]);
    }
    __syncthreads();
  }
  if (USE_LDS)
    for (int q = threadIdx.x; q < t1 - t0; q += kTileThreads)
      xout[t0 + q] = xs[q];
}

template <typename VT, typename X>
static void gsTyped(const TileArgs& t, const void* off, const void* diag,
                    const void* b, const void* xin, void* xout, bool backward,
                    DevState* st, int checkDone, hipStream_t s) {
  const size_t shm = (size_t)t.maxTileRows * sizeof(X);
  const bool lds = shm <= maxDynamicShared();
  auto launch = [&](auto kern, size_t bytes) {
    if (bytes > 64 * 1024)
      GPH_HIP(hipFuncSetAttribute((const void*)kern,
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes));
    hipLaunchKernelGGL(kern, dim3(t.T), dim3(kTileThreads), bytes, s, t,
                       static_cast<const VT*>(off),
                       static_cast<const VT*>(diag), static_cast<const X*>(b),
                       static_cast<const X*>(xin), static_cast<X*>(xout), st,
                       checkDone);
  };
  if (backward) {
    if (lds)
      launch(gsSweepKernel<VT, X, true, true>, shm);
    else
      launch(gsSweepKernel<VT, X, true, false>, 0);
  } else {
    if (lds)
      launch(gsSweepKernel<VT, X, false, true>, shm);
    else
      launch(gsSweepKernel<VT, X, false, false>, 0);
  }
  GPH_HIP(hipGetLastError());
}

void launchGsSweep(const TileArgs& t, const void* off, const void* diag,
                   DType vt, const void* b, const void* xin, void* xout,
                   DType xt, bool backward, DevState* st, int checkDone,
                   hipStream_t s) {
  if (t.T <= 0) return;
  if (vt == DType::F32 && xt == DType::F32)
    gsTyped<float, float>(t, off, diag, b, xin, xout, backward, st, checkDone, s);
  else if (vt == DType::F32 && xt == DType::F64)
    gsTyped<float, double>(t, off, diag, b, xin, xout, backward, st, checkDone, s);
  else if (vt == DType::F64 && xt == DType::F64)
    gsTyped<double, double>(t, off, diag, b, xin, xout, backward, st, checkDone, s);
  else if (vt == DType::F16 && xt == DType::F32)
    gsTyped<__half, float>(t, off, diag, b, xin, xout, backward, st, checkDone, s);
  else
    fail(kUnsupported, "unsupported Gauss-Seidel types");
}

}  // namespace gph
