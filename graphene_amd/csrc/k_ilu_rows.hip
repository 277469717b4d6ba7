// ILU(0) substitution with factor records in registers (gfx950).
//
// Same arithmetic as ilu/Solver.cpp:380-442 / :526-548 (block-Jacobi over
// tiles, forward then backward substitution per tile). One wave per tile;
// the tile's x (right-hand side with the fused PBiCGStab p- or s-update),
// 1/diag, record words and step headers are staged in LDS. The wave's 64
// lanes are 4 rows of 16; a step is up to 16 columns of rows of one level
// (stream format in ilu_pack.hpp). Each lane holds one chunk of 8 entries of
// a record: a short record (<= 16 entries) uses lane rows 0 and 1 (or 2 and
// 3) of its column, a long record (<= 32 entries) all four rows in the order
// 0, 1, 3, 2. A step's records are stored one after the other, padded to
// whole chunks or, in a dense stream, to even lengths: there a lane loads
// the 8 entries from its chunk's start and drops the products of entries
// past its record's end (they belong to the next record); either way the
// step's loads read one contiguous range.
// Every lane reads the x values of its entries' columns from LDS
// (all at once: they belong to earlier levels) and forms its products; then
// the row's partial sum passes from lane to lane in entry order
// (v_permlane16_swap between rows 0-1 / 2-3, v_permlane32_swap between rows
// 1-3 and 2-0), each lane subtracting its products in the reference's order,
// and the record's last lane finishes the row and writes it (bit-exact with
// the sequential loop). The long-record passes run only in steps that hold
// a long record. Dropped products and a record's padding entry (value +0,
// the +0 slot) are +0; lanes without a chunk load nothing (their buffer
// offsets are out of range, so the loads return +0 values and slot 0):
// subtracting +0 leaves the sum unchanged. A row longer than 32 entries
// continues in the same column of the next steps, the lanes carrying the
// partial sum in a register. A step's LDS write precedes the next step's reads in the wave's
// program order, and LDS executes a wave's operations in order, so no
// barrier separates levels.
//
// Every lane's chunk (8 values and 16-bit column offsets) sits in a ring of
// D register slots: the loads of step k + D are issued when step k is done,
// so memory latency is covered by D steps of work. Steps past a tile's end
// have zero headers (no records), so the loop needs no bounds tests and
// every ring slot is refilled on every path.
#include "kernels.hpp"

namespace gph {

struct RowsLds {
  size_t xs, invd, meta, hdr, total;
};

// x (slot 0 = +0, row i = slot i + 1) | 1/diag (same slots) | record words
// (+ 64 spare) | step headers (+ 2D + 1 zero headers)
__host__ __device__ inline RowsLds rowsLayout(const RowsArgs& p, size_t xsize, int D) {
  RowsLds L;
  size_t o = 0;
  L.xs = o;
  o += ((size_t)(p.maxRows + 1) * xsize + 15) & ~(size_t)15;
  L.invd = o;
  o += ((size_t)(p.maxRows + 1) * xsize + 15) & ~(size_t)15;
  L.meta = o;
  o += (size_t)(p.maxRecs + 64) * 4;
  L.hdr = o;
  o += (size_t)(p.maxSteps + 2 * D + 1) * 4;
  L.total = (o + 15) & ~(size_t)15;
  return L;
}

template <typename FT>
constexpr int rowsDepth() {
  return sizeof(FT) == 8 ? 6 : 8;
}

size_t rowsSharedBytes(const RowsArgs& p, DType ft, DType xt) {
  const int D = ft == DType::F64 ? rowsDepth<double>() : rowsDepth<float>();
  return rowsLayout(p, dtypeSize(xt), D).total;
}

// 16 bytes with 4-byte alignment: the row streams are dword aligned only.
struct Words4 {
  uint32_t w[4];
};

template <typename FT>
__device__ __forceinline__ FT rowVal(const Words4* v, int j);
template <>
__device__ __forceinline__ float rowVal<float>(const Words4* v, int j) {
  return __uint_as_float(v[j >> 2].w[j & 3]);
}
template <>
__device__ __forceinline__ __half rowVal<__half>(const Words4* v, int j) {
  const uint32_t w = v[j >> 3].w[(j >> 1) & 3];
  return __ushort_as_half((unsigned short)((j & 1) ? w >> 16 : w & 0xffffu));
}
template <>
__device__ __forceinline__ double rowVal<double>(const Words4* v, int j) {
  const Words4& q = v[j >> 1];
  const int h = (j & 1) * 2;
  return __hiloint2double((int)q.w[h + 1], (int)q.w[h]);
}

// Buffer resource word 3 of gfx9 raw buffers (32-bit data format).
static constexpr int kBufferFlags = 0x00020000;

__device__ __forceinline__ Words4 bufLoad16(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return Words4{{(uint32_t)v[0], (uint32_t)v[1], (uint32_t)v[2], (uint32_t)v[3]}};
}

// Lanes 0-31 of v moved to lanes 32-63 (lanes 0-31 keep their own), and
// lanes 32-63 moved to lanes 0-31: gfx950 v_permlane32_swap. Lane rows of
// 16 (v_permlane16_swap): rows 0 and 2 moved to rows 1 and 3 (rowUp), rows
// 1 and 3 moved to rows 0 and 2 (rowDown).
__device__ __forceinline__ float lowToHigh(float v) {
  const uint32_t u = __float_as_uint(v);
  return __uint_as_float(__builtin_amdgcn_permlane32_swap(u, u, false, false)[0]);
}
__device__ __forceinline__ float highToLow(float v) {
  const uint32_t u = __float_as_uint(v);
  return __uint_as_float(__builtin_amdgcn_permlane32_swap(u, u, false, false)[1]);
}
__device__ __forceinline__ double lowToHigh(double v) {
  const uint32_t lo = (uint32_t)__double2loint(v), hi = (uint32_t)__double2hiint(v);
  return __hiloint2double((int)__builtin_amdgcn_permlane32_swap(hi, hi, false, false)[0],
                          (int)__builtin_amdgcn_permlane32_swap(lo, lo, false, false)[0]);
}
__device__ __forceinline__ double highToLow(double v) {
  const uint32_t lo = (uint32_t)__double2loint(v), hi = (uint32_t)__double2hiint(v);
  return __hiloint2double((int)__builtin_amdgcn_permlane32_swap(hi, hi, false, false)[1],
                          (int)__builtin_amdgcn_permlane32_swap(lo, lo, false, false)[1]);
}

__device__ __forceinline__ float rowUp(float v) {
  const uint32_t u = __float_as_uint(v);
  return __uint_as_float(__builtin_amdgcn_permlane16_swap(u, u, false, false)[0]);
}
__device__ __forceinline__ float rowDown(float v) {
  const uint32_t u = __float_as_uint(v);
  return __uint_as_float(__builtin_amdgcn_permlane16_swap(u, u, false, false)[1]);
}
__device__ __forceinline__ double rowUp(double v) {
  const uint32_t lo = (uint32_t)__double2loint(v), hi = (uint32_t)__double2hiint(v);
  return __hiloint2double((int)__builtin_amdgcn_permlane16_swap(hi, hi, false, false)[0],
                          (int)__builtin_amdgcn_permlane16_swap(lo, lo, false, false)[0]);
}
__device__ __forceinline__ double rowDown(double v) {
  const uint32_t lo = (uint32_t)__double2loint(v), hi = (uint32_t)__double2hiint(v);
  return __hiloint2double((int)__builtin_amdgcn_permlane16_swap(hi, hi, false, false)[1],
                          (int)__builtin_amdgcn_permlane16_swap(lo, lo, false, false)[1]);
}

// Copies n words to LDS and zero-fills up to nz, kStageW loads in flight.
static constexpr int kStageW = 8;
__device__ __forceinline__ void stageWords(uint32_t* dst, const uint32_t* __restrict__ src,
                                           int n, int nz) {
  const int lane = threadIdx.x;
  for (int q0 = 0; q0 < nz; q0 += kStageW * kRowsThreads) {
    uint32_t a[kStageW];
#pragma unroll
    for (int k = 0; k < kStageW; k++) {
      const int q = q0 + k * kRowsThreads + lane;
      a[k] = n > 0 ? src[q < n ? q : n - 1] : 0u;  // clamped: loads unconditional
    }
#pragma unroll
    for (int k = 0; k < kStageW; k++) {
      const int q = q0 + k * kRowsThreads + lane;
      if (q < nz) dst[q] = q < n ? a[k] : 0u;
    }
  }
}

template <typename DT, typename X, typename F>
__device__ __forceinline__ void stageRows(int t0, int nt, const DT* __restrict__ invD, X* xs,
                                          X* invd, F f, X* wb) {
  const int lane = threadIdx.x;
  for (int q0 = 0; q0 < nt; q0 += kStageW * kRowsThreads) {
    X a[kStageW];
    DT dv[kStageW];
#pragma unroll
    for (int k = 0; k < kStageW; k++) {
      const int q = q0 + k * kRowsThreads + lane;
      const int i = t0 + (q < nt ? q : nt - 1);
      a[k] = f(i);
      dv[k] = invD[i];
    }
#pragma unroll
    for (int k = 0; k < kStageW; k++) {
      const int q = q0 + k * kRowsThreads + lane;
      if (q < nt) {
        if (wb) wb[t0 + q] = a[k];
        xs[q + 1] = a[k];
        invd[q + 1] = cvt<X>(dv[k]);
      }
    }
  }
}

// DENSE: records padded to even lengths, so a lane's 8-entry window may
// extend past its record (those products are dropped); else records are
// whole chunks.
template <typename FT, typename X, bool DENSE>
__global__ __launch_bounds__(kRowsThreads) void iluRowsKernel(
    RowsArgs p, const InvDiag<FT>* __restrict__ invD, const X* __restrict__ rhs,
    X* __restrict__ out, PreOp pre, DevState* st, int checkDone) {
  constexpr int D = rowsDepth<FT>();
  constexpr int NV = (kRowsChunk * (int)sizeof(FT)) / 16;  // 16-byte value words per lane
  extern __shared__ __align__(16) unsigned char smem[];
  if (checkDone && st->done) return;
  int tile = blockIdx.x;
  if (p.nrhs > 1) {
    // (tile, right-hand side) pairs, the pairs of a tile on one XCD (its
    // factor streams are read from that XCD's L2 after the first)
    const int lb = xcdRemap(blockIdx.x, gridDim.x);
    tile = lb / p.nrhs;
    const int j = lb - tile * p.nrhs;
    if (p.rhsState && p.rhsState[j].done) return;
    rhs += j * p.rhsStride;
    out += j * p.rhsStride;
  }
  // batched solves: the tile's system has finished
  if (p.tileSys && p.sysPhase[p.tileSys[tile]] != 0) return;
  const RowsLds L = rowsLayout(p, sizeof(X), D);
  const int lane = threadIdx.x;
  X* xs = reinterpret_cast<X*>(smem + L.xs);
  X* invd = reinterpret_cast<X*>(smem + L.invd);
  uint32_t* meta = reinterpret_cast<uint32_t*>(smem + L.meta);
  uint32_t* hdr = reinterpret_cast<uint32_t*>(smem + L.hdr);

  // step headers and record words (zero-filled past the tile's own),
  // right-hand side (with the fused PBiCGStab update) and 1/diag into LDS
  const int t0 = p.tileStart[tile], nt = p.tileStart[tile + 1] - t0;
  const int h0 = p.hdrOff[tile], nS = p.hdrOff[tile + 1] - h0;
  const int64_t m0 = p.metaOff[tile];
  const int nRec = (int)(p.metaOff[tile + 1] - m0);
  stageWords(meta, p.meta + m0, nRec, nRec + 64);
  stageWords(hdr, p.hdr + h0, nS, p.maxSteps + 2 * D + 1);
  const X* r = static_cast<const X*>(pre.r);
  X* pv = static_cast<X*>(pre.p);
  const X* vv = static_cast<const X*>(pre.v);
  if (pre.kind == kPreP) {
    if (st->iter != 0) {
      const X beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
      const X omega = (X)st->omega;
      stageRows(t0, nt, invD, xs, invd,
                [&](int i) { return r[i] + beta * (pv[i] - omega * vv[i]); }, pv);
    } else {
      stageRows(t0, nt, invD, xs, invd, [&](int i) { return pv[i]; }, (X*)nullptr);
    }
  } else if (pre.kind == kPreS) {
    const X alpha = (X)st->alpha;
    stageRows(t0, nt, invD, xs, invd, [&](int i) { return r[i] - alpha * vv[i]; }, pv);
  } else {
    stageRows(t0, nt, invD, xs, invd, [&](int i) { return rhs[i]; }, (X*)nullptr);
  }
  if (lane == 0) xs[0] = (X)0;  // the zero slot
  __syncthreads();
  if (nS == 0) return;

  // this lane: lane row lr, column c; its record word is 2c (rows 0-1) or
  // 2c + 1 (rows 2-3); its chunk of a short record is lr & 1, of a long
  // record (rows in the order 0, 1, 3, 2) lr ^ (lr >> 1)
  const int lr = lane >> 4;
  const int midx = 2 * (lane & 15) + (lane >> 5);
  const int chShort = lr & 1, chLong = lr ^ (lr >> 1);
  const bool lastShort = (lr & 1) != 0, lastLong = lr == 2;
  // the tile's streams (and the read-ahead past their end) as buffers
  const uint32_t e0 = p.elemOff[tile];
  const uint32_t nElem = p.elemOff[tile + 1] - e0 + kRowsTail;
  const __amdgpu_buffer_rsrc_t vrs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<FT*>(static_cast<const FT*>(p.val) + e0), 0, (int)(nElem * sizeof(FT)),
      kBufferFlags);
  const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint16_t*>(p.col + e0), 0, (int)(nElem * 2u), kBufferFlags);
  const unsigned char* xb = reinterpret_cast<const unsigned char*>(xs);
  const int dilu = p.dilu;
  Words4 vr[D][NV], cr[D];
  uint32_t mr[D], hr[D];
  int nr[DENSE ? D : 1];  // dense: entries of the lane's chunk inside its record
  X carry = (X)0;   // partial sum of a row continued in the next step
  int ldRow = 0;    // first record word of the next step to load
  uint32_t ldElem = 0;  // its first element
  auto header = [&](int k) { return __builtin_amdgcn_readfirstlane(hdr[k]); };
  // loads of a step's chunks into ring slot u: a lane's chunk starts
  // kRowsChunk * ch entries into its record (the record's first element is
  // in its record word); the step's records are contiguous, so a wave-wide
  // load reads one range; lanes without a chunk take an offset outside the
  // buffer: their loads return zeros without a memory access
  auto issue = [&](int u, uint32_t h, uint32_t mw) {
    const int R = (int)((h >> 12) & 63u);
    const uint32_t m = midx < R ? mw : 0u;
    const int ch = (m & kRowsLongBit) ? chLong : chShort;
    const int rem = (int)((m >> kRowsLenShift) & 63u) - kRowsChunk * ch;
    const bool has = rem > 0;
    if constexpr (DENSE) nr[u] = rem < 0 ? 0 : (rem > kRowsChunk ? kRowsChunk : rem);
    const uint32_t e =
        has ? ldElem + (m & 511u) + (uint32_t)(kRowsChunk * ch) : 0x8000000u;
    const int vo = (int)(e * sizeof(FT)), co = (int)(e * 2u);
#pragma unroll
    for (int i = 0; i < NV; i++) vr[u][i] = bufLoad16(vrs, vo, 16 * i);
    cr[u] = bufLoad16(crs, co, 0);
    mr[u] = m;
    hr[u] = h;
    ldRow += R;
    ldElem += h & 0xfffu;
  };
  // first fill, slot by slot in the steady state's order (the compiler's
  // load waits then hold for every pass)
#pragma unroll
  for (int u = 0; u < D; u++) {
    issue(u, header(u), meta[ldRow + midx]);
    __builtin_amdgcn_sched_barrier(0);
  }
  for (int k0 = 0; k0 < nS; k0 += D) {
#pragma unroll
    for (int u = 0; u < D; u++) {
      // descriptors of the step that refills this slot
      const uint32_t hn = header(k0 + u + D);
      const uint32_t mn = meta[ldRow + midx];
      const uint32_t h = hr[u], m = mr[u];
      const bool bwd = (h >> 31) != 0, anyLong = ((h >> 30) & 1u) != 0;
      const int loc = (int)(m >> kRowsSlotShift);  // the row's slot (0: no record)
      const bool lng = (m & kRowsLongBit) != 0, cont = (m & kRowsContBit) != 0;
      const bool partial = (m & kRowsPartialBit) != 0;
      X xv[kRowsChunk];
#pragma unroll
      for (int j = 0; j < kRowsChunk; j++) {
        const uint32_t w = cr[u].w[j >> 1];
        const uint32_t bo = (j & 1) ? w >> 16 : w & 0xffffu;  // 4 * slot
        xv[j] = *reinterpret_cast<const X*>(xb + bo * (sizeof(X) / 4));
      }
      const X pX = xs[loc], pD = invd[loc];
      // all reads issued before the first product waits on one of them
      __builtin_amdgcn_sched_barrier(0);
      X prod[kRowsChunk];
#pragma unroll
      for (int j = 0; j < kRowsChunk; j++) prod[j] = cvt<X>(rowVal<FT>(vr[u], j)) * xv[j];
      // dense records: entries past a record's end belong to other records,
      // +0 instead
      if constexpr (DENSE) {
#pragma unroll
        for (int j = 0; j < kRowsChunk; j++) prod[j] = j < nr[u] ? prod[j] : (X)0;
      }
      // chunk 0 (rows 0, 2), chunk 1 (rows 1, 3), in entry order
      X acc = cont ? carry : bwd ? (X)0 : pX;
#pragma unroll
      for (int j = 0; j < kRowsChunk; j++) acc = acc - prod[j];
      acc = rowUp(acc);
#pragma unroll
      for (int j = 0; j < kRowsChunk; j++) acc = acc - prod[j];
      if (anyLong) {
        // long records: chunk 2 (row 3), chunk 3 (row 2)
        X t = lowToHigh(acc);
#pragma unroll
        for (int j = 0; j < kRowsChunk; j++) t = t - prod[j];
        if (lng && lr == 3) acc = t;
        X v = rowDown(acc);
#pragma unroll
        for (int j = 0; j < kRowsChunk; j++) v = v - prod[j];
        if (lng && lr == 2) acc = v;
        carry = highToLow(acc);  // row 2 to row 0 for a continued row
      }
      X res;
      if (bwd)
        res = dilu ? pX + acc * pD : (pX + acc) * pD;
      else
        res = dilu ? acc * pD : acc;
      if ((lng ? lastLong : lastShort) && loc != 0 && !partial) xs[loc] = res;
      issue(u, hn, mn);
    }
  }
  __syncthreads();
  for (int q = lane; q < nt; q += kRowsThreads) out[t0 + q] = xs[q + 1];
}

template <typename FT, typename X>
static void rowsTyped(const RowsArgs& p, const void* invD, const void* rhs, void* out,
                      const PreOp& pre, DevState* st, int checkDone, hipStream_t s) {
  const size_t shm = rowsLayout(p, sizeof(X), rowsDepth<FT>()).total;
  if (p.dense && sizeof(FT) == 8) fail(kUnsupported, "dense ILU(0) records: float64 factors");
  auto k = p.dense ? iluRowsKernel<FT, X, sizeof(FT) != 8> : iluRowsKernel<FT, X, false>;
  if (shm > 64 * 1024)
    GPH_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shm));
  if (p.nrhs > 1 && pre.kind != kPreNone)
    fail(kUnsupported, "several right-hand sides: no fused vector update");
  hipLaunchKernelGGL(k, dim3(p.T * (p.nrhs > 1 ? p.nrhs : 1)), dim3(kRowsThreads), shm, s, p,
                     static_cast<const InvDiag<FT>*>(invD), static_cast<const X*>(rhs),
                     static_cast<X*>(out), pre, st, checkDone);
  GPH_HIP(hipGetLastError());
}

void launchIluRows(const RowsArgs& p, const void* invD, DType ft, const void* rhs,
                   void* out, DType xt, const PreOp& pre, DevState* st,
                   int checkDone, hipStream_t s) {
  if (p.T <= 0) return;
  if (ft == DType::F32 && xt == DType::F32)
    rowsTyped<float, float>(p, invD, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F16 && xt == DType::F32)
    rowsTyped<__half, float>(p, invD, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F64 && xt == DType::F64)
    rowsTyped<double, double>(p, invD, rhs, out, pre, st, checkDone, s);
  else if (ft == DType::F32 && xt == DType::F64)
    rowsTyped<float, double>(p, invD, rhs, out, pre, st, checkDone, s);
  else
    fail(kUnsupported, std::string("unsupported row-per-lane ILU types: factors ") +
                           dtypeName(ft) + ", vectors " + dtypeName(xt));
}

}  // namespace gph
