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
//     The level's entries are stored column-major (entry j of row r at
//     j*R + r, short rows padded with +0 products), so the row sums read the
//     buffer with a uniform trip count and without bank conflicts.
#include <climits>

#include "kernels.hpp"

namespace gph {

static constexpr int kDepth = 16;  // must equal kPackDepth of ilu_pack.hpp

struct PackedLds {
  size_t xs, invd, prod, tab[2], rows[2], total;
};

// LDS: x (+ zero slot), 1/diag, level products, and per direction the level
// tables (slab end, first row slot, width: 3 x (maxLevels+2) ints, the last
// entry a spare that read-ahead may touch) and the local id of every row
// slot (u16).
__host__ __device__ inline PackedLds packedLayout(const PackedArgs& p,
                                                  size_t xsize) {
  PackedLds L;
  size_t o = 0;
  L.xs = o;
  o += (size_t)(p.maxRows + 1) * xsize;
  L.invd = o;
  o += (size_t)p.maxRows * xsize;
  L.prod = o;
  // a level's products, or the kDepth*64 entries written by passes outside
  // the levels (ring fill, ring overrun)
  const int prodEntries =
      p.maxLevelEntries + 64 > kDepth * 64 ? p.maxLevelEntries + 64 : kDepth * 64;
  o += (size_t)prodEntries * xsize;
  o = (o + 3) & ~(size_t)3;
  for (int d = 0; d < 2; d++) {
    L.tab[d] = o;
    o += (size_t)3 * (p.maxLevels + 2) * 4;
  }
  for (int d = 0; d < 2; d++) {
    L.rows[d] = o;
    o += (size_t)p.maxRows * 2;
  }
  L.total = (o + 15) & ~(size_t)15;
  return L;
}

size_t packedSharedBytes(const PackedArgs& p, DType xt) {
  return packedLayout(p, dtypeSize(xt)).total;
}

// Setup loads are issued kStage per lane before any of them is used, so a
// tile's staging costs a few memory latencies rather than one per row block.
static constexpr int kStage = 8;

template <typename FT, typename X, bool BACKWARD>
__device__ __forceinline__ void packedSweep(const PackedDirArgs& d, int tile,
                                            int nL, X* xs, const X* invd,
                                            X* prod, const uint16_t* rows,
                                            const int* tab, int tabStride,
                                            int dilu) {
  const int lane = threadIdx.x;
  const int* slabEnd = tab;
  const int* rowStart = tab + tabStride;
  const int* width = tab + 2 * tabStride;
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
  // refill sources: slab sb + kDepth + u of pass sb (advanced per pass)
  const uint32_t* __restrict__ cpn = cp;
  const uint32_t* __restrict__ hpn = hp;
  const FT* __restrict__ vpn = vp;

  // The level's tables, and for its first 64 rows the local id, the current
  // x value and 1/diag, are read one level ahead so that the row sums only
  // wait for the products. Row r0 + lane of a level is held by lane `lane`.
  // Products of slab s go to prod[(s - lvlFirst)*64 + lane]; the slabs of
  // the first pass (s < 0, lvlFirst = -kDepth) and those past the last level
  // (lvlFirst = S) land in the first kDepth*64 entries, which no row sum of
  // theirs reads.
  int lvl = 0, lvlFirst = -kDepth, r0 = 0;
  int lvlEnd = slabEnd[0], r1 = rowStart[1], W = width[0];
  const int ntm = rowStart[nL] > 0 ? rowStart[nL] - 1 : 0;  // last row slot (index clamp)
  int pLoc = rows[lane < ntm ? lane : ntm];
  X pX = xs[pLoc], pD = invd[pLoc];
  X xNext = xs[0];  // x at the column of the next slot, read one slot ahead
  for (int sb = -kDepth; sb < Stot; sb += kDepth) {
    int pb = (sb - lvlFirst) * kPackedThreads + lane;  // pb + u*64 >= 0
#pragma unroll
    for (int u = 0; u < kDepth; u++) {
      const int s = sb + u;
      X a;
      if constexpr (kHalfVals)
        a = cvt<X>(__ushort_as_half(
            (unsigned short)((u & 1) ? hw[u >> 1] >> 16 : hw[u >> 1] & 0xffffu)));
      else
        a = cvt<X>(v[u]);
      const X xu = xNext;
      // column of slot u+1 (the next pass's slot 0 sits in cw[0], refilled
      // at slot 1 of this pass); its x is read before this slot's product
      // is stored, and read again if a level ends in between
      const int un = (u + 1) % kDepth;  // constant after unrolling
      const uint32_t cn = (un & 1) ? cw[un >> 1] >> 16 : cw[un >> 1] & 0xffffu;
      xNext = xs[cn];
      prod[pb + u * kPackedThreads] = a * xu;
      if (u & 1) {  // both slots of the word consumed: refill it
        cw[u >> 1] = cpn[(u >> 1) * kPackedThreads];
        if constexpr (kHalfVals) hw[u >> 1] = hpn[(u >> 1) * kPackedThreads];
      }
      if constexpr (!kHalfVals) v[u] = vpn[u * kPackedThreads];
      if (s + 1 == lvlEnd) {
        __syncthreads();
        const int R = r1 - r0;
        // rows r0 .. r0+63 from the prefetched registers
        if (lane < R) {
          const X* pp = prod + lane;
          X acc = BACKWARD ? (X)0 : pX;
          int j = 0;
          // issue 8 reads, then subtract in stored order
          for (; j + 8 <= W; j += 8) {
            X pv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) pv[q] = pp[(j + q) * R];
#pragma unroll
            for (int q = 0; q < 8; q++) acc = acc - pv[q];
          }
          for (; j < W; j++) acc = acc - pp[j * R];
          if (BACKWARD)
            xs[pLoc] = dilu ? pX + acc * pD : (pX + acc) * pD;
          else
            xs[pLoc] = dilu ? acc * pD : acc;
        }
        // further rows of wide levels
        for (int r = lane + kPackedThreads; r < R; r += kPackedThreads) {
          const int loc = rows[r0 + r];
          const X* pp = prod + r;
          const X x0 = xs[loc], d0 = invd[loc];
          X acc = BACKWARD ? (X)0 : x0;
          for (int j = 0; j < W; j++) acc = acc - pp[j * R];
          if (BACKWARD)
            xs[loc] = dilu ? x0 + acc * d0 : (x0 + acc) * d0;
          else
            xs[loc] = dilu ? acc * d0 : acc;
        }
        __syncthreads();
        // next level: tables and the first 64 rows' values
        lvl++;
        r0 = r1;
        lvlFirst = lvlEnd;
        lvlEnd = slabEnd[lvl];
        r1 = rowStart[lvl + 1];
        W = width[lvl];
        pLoc = rows[r0 + lane < ntm ? r0 + lane : ntm];
        pX = xs[pLoc];
        pD = invd[pLoc];
        pb = (sb - lvlFirst) * kPackedThreads + lane;
        xNext = xs[cn];
      }
    }
    cpn += (kDepth / 2) * kPackedThreads;
    hpn += (kDepth / 2) * kPackedThreads;
    vpn += kDepth * kPackedThreads;
  }
  __syncthreads();
}

// Stages a tile's rows into LDS: right-hand side value f(i) (written back
// to `wb` when given: the fused p- or s-update), 1/diag, row slots.
template <typename FT, typename X>
struct PackedStage {
  int t0, nt;
  const InvDiag<FT>* invD;
  const uint16_t* locF;
  const uint16_t* locB;
  X* xs;
  X* invd;
  uint16_t* rowsF;
  uint16_t* rowsB;
  template <typename F>
  __device__ __forceinline__ void run(F f, X* wb) const {
    const int lane = threadIdx.x;
    for (int q0 = 0; q0 < nt; q0 += kStage * kPackedThreads) {
      X a[kStage];
      InvDiag<FT> dv[kStage];
      uint16_t rf[kStage], rb[kStage];
#pragma unroll
      for (int k = 0; k < kStage; k++) {
        const int q = q0 + k * kPackedThreads + lane;
        const int i = t0 + (q < nt ? q : nt - 1);  // clamped: loads unconditional
        a[k] = f(i);
        dv[k] = invD[i];
        rf[k] = locF[i];
        rb[k] = locB[i];
      }
#pragma unroll
      for (int k = 0; k < kStage; k++) {
        const int q = q0 + k * kPackedThreads + lane;
        if (q < nt) {
          if (wb) wb[t0 + q] = a[k];
          xs[q] = a[k];
          invd[q] = cvt<X>(dv[k]);
          rowsF[q] = rf[k];
          rowsB[q] = rb[k];
        }
      }
    }
  }
};

template <typename FT, typename X>
__global__ __launch_bounds__(kPackedThreads) void iluPackedKernel(
    PackedArgs p, const InvDiag<FT>* __restrict__ invD, const X* __restrict__ rhs,
    X* __restrict__ out, PreOp pre, DevState* st, int checkDone) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (checkDone && st->done) return;
  if (p.groupSys) {  // batched: skip groups with no inner solve running
    bool run = false;
    for (int k = p.groupSys[2 * blockIdx.x]; k <= p.groupSys[2 * blockIdx.x + 1]; k++)
      run = run || p.sysPhase[k] == 0;
    if (!run) return;
  }
  const PackedLds L = packedLayout(p, sizeof(X));
  X* xs = reinterpret_cast<X*>(smem + L.xs);
  X* invd = reinterpret_cast<X*>(smem + L.invd);
  X* prod = reinterpret_cast<X*>(smem + L.prod);
  int* tabF = reinterpret_cast<int*>(smem + L.tab[0]);
  int* tabB = reinterpret_cast<int*>(smem + L.tab[1]);
  uint16_t* rowsF = reinterpret_cast<uint16_t*>(smem + L.rows[0]);
  uint16_t* rowsB = reinterpret_cast<uint16_t*>(smem + L.rows[1]);
  const int ts = p.maxLevels + 2;  // table stride (one spare entry)
  const int tile = blockIdx.x, lane = threadIdx.x;
  const int t0 = p.tileStart[tile], nt = p.tileStart[tile + 1] - t0;
  const int loF = p.fwd.lvlOff[tile], nLF = p.fwd.lvlOff[tile + 1] - loF - 1;
  const int loB = p.bwd.lvlOff[tile], nLB = p.bwd.lvlOff[tile + 1] - loB - 1;

  // level tables of both directions
  const int nTab = (nLF > nLB ? nLF : nLB) + 1;
  for (int q0 = 0; q0 < nTab; q0 += kStage * kPackedThreads) {
    int a[kStage][6];
    // loads are unconditional (clamped index) so they issue back to back
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const int q = q0 + k * kPackedThreads + lane;
      const int qf = loF + (q < nLF ? q : nLF), qb = loB + (q < nLB ? q : nLB);
      a[k][0] = p.fwd.slabEnd[qf];
      a[k][1] = p.fwd.rowStart[qf];
      a[k][2] = p.fwd.width[qf];
      a[k][3] = p.bwd.slabEnd[qb];
      a[k][4] = p.bwd.rowStart[qb];
      a[k][5] = p.bwd.width[qb];
    }
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const int q = q0 + k * kPackedThreads + lane;
      if (q <= nLF) {
        tabF[q] = a[k][0];
        tabF[ts + q] = a[k][1];
        tabF[2 * ts + q] = a[k][2];
      }
      if (q <= nLB) {
        tabB[q] = a[k][3];
        tabB[ts + q] = a[k][4];
        tabB[2 * ts + q] = a[k][5];
      }
    }
  }

  // right-hand side (with the fused PBiCGStab update), 1/diag and the row
  // slots of both directions
  const X* r = static_cast<const X*>(pre.r);
  X* pv = static_cast<X*>(pre.p);
  const X* vv = static_cast<const X*>(pre.v);
  const PackedStage<FT, X> stg{t0, nt, invD, p.fwd.rowLoc, p.bwd.rowLoc, xs, invd, rowsF, rowsB};
  if (pre.kind == kPreP) {
    if (st->iter != 0) {
      const X beta = ((X)st->rho / (X)st->rhoOld) * ((X)st->alpha / (X)st->omega);
      const X omega = (X)st->omega;
      stg.run([&](int i) { return r[i] + beta * (pv[i] - omega * vv[i]); }, pv);
    } else {
      stg.run([&](int i) { return pv[i]; }, (X*)nullptr);
    }
  } else if (pre.kind == kPreS) {
    const X alpha = (X)st->alpha;
    stg.run([&](int i) { return r[i] - alpha * vv[i]; }, pv);
  } else {
    stg.run([&](int i) { return rhs[i]; }, (X*)nullptr);
  }
  if (lane == 0) xs[nt] = (X)0;
  __syncthreads();
  packedSweep<FT, X, false>(p.fwd, tile, nLF, xs, invd, prod, rowsF, tabF, ts,
                            p.dilu);
  packedSweep<FT, X, true>(p.bwd, tile, nLB, xs, invd, prod, rowsB, tabB, ts,
                           p.dilu);
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
                     static_cast<const InvDiag<FT>*>(invD), static_cast<const X*>(rhs),
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
