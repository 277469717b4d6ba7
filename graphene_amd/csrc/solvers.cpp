// Solver tree implementation (see solvers.hpp).
#include "solvers.hpp"

#include "gs_pack.hpp"
#include "ilu_pack.hpp"

#include <chrono>
#include <cstdio>
#include <cmath>

namespace graphene::matrix::solver {

using gph::DevBuf;
using gph::DevState;
using gph::DType;
using gph::fail;
using gph::kInvalid;
using gph::kUnsupported;

// ---------------------------------------------------------------------------
void SolverStats::print() const {
  std::fprintf(stderr,
               "[graphene_amd] %s: Solved on %d tiles, %s norm, Initial residual "
               "= %.9g, Final residual = %.9g, Iterations = %lld, Converged = %s, "
               "Singular = %s, Current Execution Time = %.6f s\n",
               solverName.c_str(), numTiles, normToString(norm).c_str(),
               initialResidual, finalResidual, (long long)iterations,
               converged ? "true" : "false", singular ? "true" : "false",
               seconds);
}

void Solver::applyDevice(const Vec&, const Vec&, const gph::PreOp&, DevState*,
                         hipStream_t) {
  fail(kUnsupported, name() + " cannot be used as a preconditioner");
}

void Solver::applyDeviceMulti(const Vec& rhs, const Vec& out, int k, int64_t stride,
                              const DevState* states, DevState* st, hipStream_t s) {
  (void)states;
  const size_t es = gph::dtypeSize(out.type);
  for (int j = 0; j < k; j++) {
    const Vec r{static_cast<char*>(rhs.ptr) + (size_t)j * stride * gph::dtypeSize(rhs.type),
                rhs.type};
    const Vec o{static_cast<char*>(out.ptr) + (size_t)j * stride * es, out.type};
    applyDevice(r, o, gph::PreOp{}, st, s);
  }
}

Vec Workspace::get(int slot, DType t, int64_t n) {
  if ((int)e_.size() <= slot) e_.resize((size_t)slot + 1);
  Entry& e = e_[(size_t)slot];
  if (!e.buf || e.type != t || e.n < n) {
    e.buf.alloc((size_t)std::max<int64_t>(n, 1) * gph::dtypeSize(t));
    e.type = t;
    e.n = n;
  }
  return Vec{e.buf.get(), t};
}
size_t Workspace::bytes() const {
  size_t b = 0;
  for (auto& e : e_) b += e.buf.bytes();
  return b;
}

// gph::NormKind of a configured norm (lg/matrix/Norm.hpp). L1 here is the
// sum of magnitudes and LINF the largest magnitude; the reference's L1 sums
// signed values (Tensor.cpp:430-431) and its LINF reduction is unavailable.
static int normKind(VectorNorm n) {
  switch (n) {
    case VectorNorm::L1:
      return gph::kNormL1;
    case VectorNorm::LINF:
      return gph::kNormLinf;
    default:
      return gph::kNormL2;
  }
}

static double nowSeconds() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Host-pinned copy of a DevState for asynchronous read-back.
class PinnedState {
 public:
  // slots 0 and 1: states copied back after each chunk; 2: initial state
  PinnedState() { GPH_HIP(hipHostMalloc((void**)&p_, 3 * sizeof(DevState), 0)); }
  ~PinnedState() {
    if (p_) (void)hipHostFree(p_);
  }
  DevState* operator[](int i) { return p_ + i; }

 private:
  DevState* p_ = nullptr;
};

// ---------------------------------------------------------------------------
// ILU(0): factorization in the constructor (ilu/Solver.cpp:44-51), apply per
// tile. DILU (diagonalBased, ilu/Solver.cpp:191-216, :405-441) keeps A's
// off-diagonals as factors and only modifies the diagonal.
// ---------------------------------------------------------------------------
class IluSolver : public Solver {
 public:
  IluSolver(const DeviceMatrix& A, std::shared_ptr<ilu::Configuration> c)
      : Solver(A), cfg_(std::move(c)) {
    ft_ = cfg_->factorType.value_or(DType::F32);
    factorize();
  }
  std::string name() const override { return cfg_->solverName(); }
  bool usesInitialGuess() const override { return false; }
  bool supportsDeviceApply() const override { return true; }

  SolverStats solve(const Vec& x, const Vec& b) override {
    double t0 = nowSeconds();
    hipStream_t s = A_.stream();
    applyDevice(b, x, gph::PreOp{}, nullptr, s);
    GPH_HIP(hipStreamSynchronize(s));
    SolverStats st;
    st.solverName = name();
    st.norm = VectorNorm::None;
    st.numTiles = A_.numTiles();
    st.iterations = 1;  // ilu/Solver.cpp:570-573
    st.seconds = nowSeconds() - t0;
    return st;
  }

  void applyDevice(const Vec& rhs, const Vec& out, const gph::PreOp& pre,
                   DevState* st, hipStream_t s) override {
    const void* lu = lu_.get();
    const void* inv = invD_.get();
    DType ft = ft_;
    // float16 factors: float16 off-diagonals, float32 inverse diagonal
    if (ft_ == DType::F16) lu = lu16_.get();
    if (rows_ && gph::rowsSharedBytes(rargs_, ft, out.type) <= kMaxPackedShared) {
      gph::launchIluRows(rargs_, inv, ft, rhs.ptr, out.ptr, out.type, pre, st, st ? 1 : 0, s);
      return;
    }
    if (level_ && gph::levelSharedBytes(largs_, out.type) <= kMaxPackedShared) {
      gph::launchIluLevel(largs_, ft, rhs.ptr, out.ptr, out.type, pre, st, st ? 1 : 0, s);
      return;
    }
    if (packed_ && gph::packedSharedBytes(pargs_, out.type) <= kMaxPackedShared) {
      gph::launchIluPacked(pargs_, inv, ft, rhs.ptr, out.ptr, out.type, pre, st,
                           st ? 1 : 0, s);
      return;
    }
    if (block_) {
      gph::launchIluBlock(bargs_, ft, rhs.ptr, out.ptr, out.type, pre, st, st ? 1 : 0, s);
      return;
    }
    gph::launchIluApply(A_.tileArgs(), lu, inv, ft, rhs.ptr, out.ptr, out.type,
                        pre, st, st ? 1 : 0, cfg_->diagonalBased, s);
  }

  void applyDeviceMulti(const Vec& rhs, const Vec& out, int k, int64_t stride,
                        const DevState* states, DevState* st, hipStream_t s) override {
    DType ft = ft_;
    if (rows_ && gph::rowsSharedBytes(rargs_, ft, out.type) <= kMaxPackedShared) {
      gph::RowsArgs a = rargs_;
      a.nrhs = k;
      a.rhsStride = stride;
      a.rhsState = states;
      gph::launchIluRows(a, invD_.get(), ft, rhs.ptr, out.ptr, out.type, gph::PreOp{}, st,
                         st ? 1 : 0, s);
      return;
    }
    Solver::applyDeviceMulti(rhs, out, k, stride, states, st, s);
  }

  void setSystemSkip(const int32_t* phases,
                     const std::vector<int64_t>& systemRowStart) override {
    if (rows_) {  // the system of every tile (systems are tile ranges)
      const std::vector<int64_t>& ts = A_.hostMatrix().tileStart;
      std::vector<int32_t> sys(ts.size() - 1);
      size_t k = 0;
      for (size_t t = 0; t + 1 < ts.size(); t++) {
        while (k + 2 < systemRowStart.size() && systemRowStart[k + 1] <= ts[t]) k++;
        sys[t] = (int32_t)k;
      }
      tileSys_.upload(sys.data(), sys.size(), A_.stream());
      GPH_HIP(hipStreamSynchronize(A_.stream()));
      rargs_.tileSys = tileSys_.as<int32_t>();
      rargs_.sysPhase = phases;
      return;
    }
    if (!packed_) return;
    // systems of each wave group (groups and systems are row ranges)
    const size_t G = groupRowStart_.size() - 1;
    std::vector<int32_t> gs(2 * G);
    size_t k = 0;
    for (size_t g = 0; g < G; g++) {
      while (k + 1 < systemRowStart.size() - 1 && systemRowStart[k + 1] <= groupRowStart_[g]) k++;
      size_t h = k;
      while (h + 1 < systemRowStart.size() - 1 && systemRowStart[h + 1] < groupRowStart_[g + 1])
        h++;
      gs[2 * g] = (int32_t)k;
      gs[2 * g + 1] = (int32_t)h;
    }
    groupSys_.upload(gs.data(), gs.size(), A_.stream());
    GPH_HIP(hipStreamSynchronize(A_.stream()));
    pargs_.groupSys = groupSys_.as<int32_t>();
    pargs_.sysPhase = phases;
  }

  size_t factorBytes() const {
    return lu_.bytes() + invD_.bytes() + lu16_.bytes() +
           fcol_.bytes() + fval_.bytes() + bcol_.bytes() + bval_.bytes() + lbytes_.bytes() +
           rcol_.bytes() + rval_.bytes() + rmeta_.bytes() + blk_[2].bytes() + blk_[3].bytes() +
           blk_[6].bytes() + blk_[7].bytes();
  }

 private:
  void factorize() {
    hipStream_t s = A_.stream();
    const int64_t nnz = A_.nnzOff(), n = A_.n();
    DType work = ft_ == DType::F64 ? DType::F64 : DType::F32;
    lu_.alloc((size_t)std::max<int64_t>(nnz, 1) * gph::dtypeSize(work));
    invD_.alloc((size_t)n * gph::dtypeSize(work));
    gph::launchCast(A_.crsOff(DType::F32), DType::F32, lu_.get(), work, nnz, s);
    gph::launchCast(A_.crsDiag(DType::F32), DType::F32, invD_.get(), work, n, s);
    std::vector<int32_t> list;
    std::vector<int64_t> levelOff;
    host::forwardLevelRows(A_.hostMatrix(), list, levelOff);
    gph::DevBuf listD;
    listD.upload(list.data(), std::max<size_t>(list.size(), 1), s);
    gph::FactorArgs fa;
    fa.tileStart = A_.tileArgs().tileStart;
    fa.rowPtr = A_.crsRowPtr();
    fa.col = A_.crsCol();
    fa.list = listD.as<int32_t>();
    gph::launchIluFactor(fa, levelOff, n, lu_.get(), invD_.get(), work, cfg_->diagonalBased, s);
    if (ft_ == DType::F16) {
      lu16_.alloc((size_t)std::max<int64_t>(nnz, 1) * 2);
      gph::launchCast(lu_.get(), work, lu16_.get(), DType::F16, nnz, s);
    }
    if (cfg_->substitution == "level") buildLevel(s);
    if (cfg_->substitution == "rows") buildRows(s);
    if (cfg_->substitution == "block") buildBlock(s);
    // also the fallback of the other kernels
    if (!level_ && !rows_ && !block_ && cfg_->substitution != "tile") buildPacked(s);
    // tiles too large for the one-wave kernels: the workgroup-per-tile kernel
    if (cfg_->substitution != "tile" && !block_) {
      const DType xt = ft_ == DType::F64 ? DType::F64 : DType::F32;
      const bool fits =
          (rows_ && gph::rowsSharedBytes(rargs_, ft_, xt) <= kMaxPackedShared) ||
          (level_ && gph::levelSharedBytes(largs_, xt) <= kMaxPackedShared) ||
          (packed_ && gph::packedSharedBytes(pargs_, xt) <= kMaxPackedShared);
      if (!fits) buildBlock(s);
    }
    GPH_HIP(hipStreamSynchronize(s));
  }

  // Level records of the factors for the lane-per-row kernel (built on the
  // host from the device factors).
  // (Not for float16 factors, whose float32 inverse diagonal the records'
  // factor-sized 1/diag field cannot hold: they run in the packed kernel.)
  void buildLevel(hipStream_t s) {
    if (ft_ == DType::F16) return;
    const int64_t nnz = A_.nnzOff(), n = A_.n();
    const size_t es = gph::dtypeSize(ft_);
    std::vector<uint8_t> lu((size_t)std::max<int64_t>(nnz, 1) * es), inv((size_t)n * es);
    const void* luD = ft_ == DType::F16 ? lu16_.get() : lu_.get();
    const void* invDv = invD_.get();
    GPH_HIP(hipMemcpyAsync(lu.data(), luD, (size_t)nnz * es, hipMemcpyDeviceToHost, s));
    GPH_HIP(hipMemcpyAsync(inv.data(), invDv, (size_t)n * es, hipMemcpyDeviceToHost, s));
    GPH_HIP(hipStreamSynchronize(s));
    IluLevelStream L = packIluLevels(A_.hostMatrix(), std::max(1, cfg_->rowsPerWave), lu.data(),
                                     inv.data(), (int)es, cfg_->diagonalBased, gph::kLvlMaxStep,
                                     gph::kLvlRing, gph::kLvlTail, 2);
    if (!L.ok) return;
    lbytes_.upload(L.bytes.data(), L.bytes.size(), s);
    lgoff_.upload(L.pairOff.data(), L.pairOff.size(), s);
    lsteps_.upload(L.steps.data(), L.steps.size(), s);
    lgstart_.upload(L.groupStart.data(), L.groupStart.size(), s);
    largs_.G = (int32_t)L.groupStart.size() - 1;
    largs_.P = (int32_t)L.pairOff.size() - 1;
    largs_.groupStart = lgstart_.as<int32_t>();
    largs_.pairOff = lgoff_.as<int64_t>();
    largs_.steps = lsteps_.as<int32_t>();
    largs_.bytes = lbytes_.as<uint8_t>();
    largs_.maxRows = L.maxRows;
    largs_.dilu = cfg_->diagonalBased ? 1 : 0;
    level_ = true;
  }

  // Row-record stream of the factors (built on the host from the device
  // factors).
  void buildRows(hipStream_t s) {
    static_assert(kRowsMaxLen == gph::kRowsMaxLen && kRowsTail == gph::kRowsTail &&
                      kRowsPad == gph::kRowsChunk && kRowsLong == gph::kRowsLongBit &&
                      kRowsCont == gph::kRowsContBit && kRowsPartial == gph::kRowsPartialBit &&
                      kRowsLenShift == gph::kRowsLenShift &&
                      kRowsSlotShift == gph::kRowsSlotShift,
                  "record format of stream and kernel");
    const int64_t nnz = A_.nnzOff(), n = A_.n();
    const size_t es = gph::dtypeSize(ft_);
    std::vector<uint8_t> lu((size_t)std::max<int64_t>(nnz, 1) * es);
    const void* luD = ft_ == DType::F16 ? lu16_.get() : lu_.get();
    GPH_HIP(hipMemcpyAsync(lu.data(), luD, (size_t)nnz * es, hipMemcpyDeviceToHost, s));
    GPH_HIP(hipStreamSynchronize(s));
    (void)n;
    // dense records when rows average more than one chunk of in-tile
    // entries per direction
    const host::HostMatrix& hm = A_.hostMatrix();
    int64_t inTile = 0;
    for (int32_t t = 0; t < hm.numTiles; t++)
      for (int64_t i = hm.tileStart[(size_t)t]; i < hm.tileStart[(size_t)t + 1]; i++)
        for (int64_t a = hm.rowPtr[(size_t)i]; a < hm.rowPtr[(size_t)i + 1]; a++)
          inTile += hm.col[(size_t)a] >= hm.tileStart[(size_t)t] &&
                    hm.col[(size_t)a] < hm.tileStart[(size_t)t + 1];
    const bool dense = inTile > 2 * (int64_t)kRowsPad * n;
    IluRowStream R = packIluRows(hm, lu.data(), (int)es, dense);
    if (!R.ok) return;
    rcol_.upload(R.col.data(), R.col.size(), s);
    rval_.upload(R.val.data(), R.val.size(), s);
    relem_.upload(R.elemOff.data(), R.elemOff.size(), s);
    rhoff_.upload(R.hdrOff.data(), R.hdrOff.size(), s);
    rhdr_.upload(R.hdr.data(), std::max<size_t>(R.hdr.size(), 1), s);
    rmeta_.upload(R.meta.data(), std::max<size_t>(R.meta.size(), 1), s);
    rmoff_.upload(R.metaOff.data(), R.metaOff.size(), s);
    rargs_.T = A_.numTiles();
    rargs_.tileStart = A_.tileArgs().tileStart;
    rargs_.maxRows = R.maxRows;
    rargs_.maxSteps = R.maxSteps;
    if (R.maxRecs > (1 << 30)) return;
    rargs_.maxRecs = (int32_t)R.maxRecs;
    rargs_.metaOff = rmoff_.as<int64_t>();
    rargs_.dilu = cfg_->diagonalBased ? 1 : 0;
    rargs_.dense = R.dense ? 1 : 0;
    rargs_.elemOff = relem_.as<uint32_t>();
    rargs_.hdrOff = rhoff_.as<int32_t>();
    rargs_.hdr = rhdr_.as<uint32_t>();
    rargs_.meta = rmeta_.as<uint32_t>();
    rargs_.col = rcol_.as<uint16_t>();
    rargs_.val = rval_.get();
    rows_ = true;
  }

  // Level-major stream of the factors for the workgroup-per-tile kernel
  // (built on the host from the device factors).
  void buildBlock(hipStream_t s) {
    const int64_t nnz = A_.nnzOff();
    const size_t es = gph::dtypeSize(ft_);
    std::vector<uint8_t> lu((size_t)std::max<int64_t>(nnz, 1) * es);
    const void* luD = ft_ == DType::F16 ? lu16_.get() : lu_.get();
    GPH_HIP(hipMemcpyAsync(lu.data(), luD, (size_t)nnz * es, hipMemcpyDeviceToHost, s));
    GPH_HIP(hipStreamSynchronize(s));
    const size_t ies = ft_ == DType::F64 ? 8 : 4;  // InvDiag of the factor type
    std::vector<uint8_t> inv((size_t)std::max<int64_t>(A_.n(), 1) * ies);
    GPH_HIP(hipMemcpyAsync(inv.data(), invD_.get(), (size_t)A_.n() * ies, hipMemcpyDeviceToHost,
                           s));
    GPH_HIP(hipStreamSynchronize(s));
    static_assert(kBlockSlots == gph::kBlockThreads, "block stream format");
    IluBlockStream B = packIluBlock(A_.hostMatrix(), lu.data(), (int)es, inv.data(), (int)ies);
    if (!B.ok || gph::blockInfoBytes(B.maxLevels) > kMaxPackedShared) return;
    auto dir = [&](const IluBlockDirection& d, DevBuf* b, gph::BlockDirArgs& a) {
      b[0].upload(d.lvlBase.data(), d.lvlBase.size(), s);
      b[1].upload(d.info.data(), d.info.size(), s);
      b[2].upload(d.rows.data(), d.rows.size(), s);
      b[3].upload(d.ents.data(), d.ents.size(), s);
      a.lvlBase = b[0].as<int32_t>();
      a.info = b[1].as<int32_t>();
      a.rows = b[2].get();
      a.ents = b[3].get();
    };
    dir(B.fwd, blk_, bargs_.fwd);
    dir(B.bwd, blk_ + 4, bargs_.bwd);
    bargs_.maxLevels = B.maxLevels;
    bargs_.T = A_.numTiles();
    bargs_.tileStart = A_.tileArgs().tileStart;
    bargs_.maxRows = B.maxRows;
    bargs_.width = B.width;
    bargs_.dilu = cfg_->diagonalBased ? 1 : 0;
    block_ = true;
  }

  // Level-major packed copy of the factors for the one-wave-per-tile kernel.
  void buildPacked(hipStream_t s) {
    PackedILU P = packILU(A_.hostMatrix(), std::max(1, cfg_->rowsPerWave));
    if (!P.ok) return;
    groupStart_.upload(P.groupStart.data(), P.groupStart.size(), s);
    groupRowStart_.assign(P.groupStart.begin(), P.groupStart.end());
    const void* src = ft_ == DType::F16 ? lu16_.get() : lu_.get();
    const size_t es = gph::dtypeSize(ft_);
    auto dir = [&](const PackedDirection& d, DevBuf& col, DevBuf& val, DevBuf& so,
                   DevBuf& lo, DevBuf& se, DevBuf& rs, DevBuf& ri, DevBuf& wd,
                   gph::PackedDirArgs& a) {
      if ((d.col.size() / kPackSlab) % 2 != 0)
        fail(gph::kRuntime, "packed ILU stream has an odd number of slabs");
      const std::vector<uint16_t> colDev = pairSlabs(d.col);
      col.upload(colDev.data(), colDev.size(), s);
      // 16-bit factor values use the paired slab order as well
      const std::vector<int32_t> srcDev = es == 2 ? pairSlabs(d.src) : d.src;
      DevBuf idx;
      idx.upload(srcDev.data(), srcDev.size(), s);
      val.alloc(d.src.size() * es);
      gph::launchGather(src, ft_, idx.as<int32_t>(), val.get(), (int64_t)d.src.size(), s);
      GPH_HIP(hipStreamSynchronize(s));
      so.upload(d.slabOff.data(), d.slabOff.size(), s);
      lo.upload(d.lvlOff.data(), d.lvlOff.size(), s);
      se.upload(d.slabEnd.data(), d.slabEnd.size(), s);
      rs.upload(d.rowStart.data(), d.rowStart.size(), s);
      ri.upload(d.rowLoc.data(), d.rowLoc.size(), s);
      wd.upload(d.width.data(), d.width.size(), s);
      a.col = col.as<uint16_t>();
      a.val = val.get();
      a.slabOff = so.as<int32_t>();
      a.lvlOff = lo.as<int32_t>();
      a.slabEnd = se.as<int32_t>();
      a.rowStart = rs.as<int32_t>();
      a.rowLoc = ri.as<uint16_t>();
      a.width = wd.as<int32_t>();
    };
    dir(P.fwd, fcol_, fval_, fso_, flo_, fse_, frs_, fri_, fwd_, pargs_.fwd);
    dir(P.bwd, bcol_, bval_, bso_, blo_, bse_, brs_, bri_, bwd_, pargs_.bwd);
    pargs_.T = (int32_t)P.groupStart.size() - 1;
    pargs_.tileStart = groupStart_.as<int32_t>();
    pargs_.maxRows = P.maxRows;
    pargs_.dilu = cfg_->diagonalBased ? 1 : 0;
    pargs_.maxLevels = std::max(P.fwd.maxLevels, P.bwd.maxLevels);
    pargs_.maxLevelEntries = std::max(P.fwd.maxLevelEntries, P.bwd.maxLevelEntries);
    packed_ = true;
  }

  static constexpr size_t kMaxPackedShared = 156 * 1024;
  DevBuf groupStart_, groupSys_;
  std::vector<int64_t> groupRowStart_;
  std::shared_ptr<ilu::Configuration> cfg_;
  DType ft_;
  DevBuf lu_, invD_, lu16_;
  bool packed_ = false, level_ = false, rows_ = false, block_ = false;
  gph::BlockArgs bargs_;
  DevBuf blk_[8];
  gph::RowsArgs rargs_;
  DevBuf rcol_, rval_, relem_, rhoff_, rhdr_, rmeta_, rmoff_, tileSys_;
  gph::PackedArgs pargs_;
  gph::LevelArgs largs_;
  DevBuf lbytes_, lgoff_, lsteps_, lgstart_;
  DevBuf fcol_, fval_, fso_, flo_, fse_, frs_, fri_, fwd_;
  DevBuf bcol_, bval_, bso_, blo_, bse_, brs_, bri_, bwd_;
};

// ---------------------------------------------------------------------------
// Gauss-Seidel (gauss-seidel/Solver.cpp:45-160) with the codelet's use of b.
// ---------------------------------------------------------------------------
class GaussSeidelSolver : public Solver {
 public:
  GaussSeidelSolver(const DeviceMatrix& A,
                    std::shared_ptr<gaussseidel::Configuration> c)
      : Solver(A), cfg_(std::move(c)) {
    buildStream(false, fwd_, fbuf_);
    if (cfg_->symmetric) buildStream(true, bwd_, bbuf_);
  }
  std::string name() const override { return cfg_->solverName(); }
  bool usesInitialGuess() const override { return true; }
  bool supportsDeviceApply() const override { return true; }

  // One "iteration" = forward sweep (+ backward sweep when symmetric).
  void sweeps(const Vec& x, const Vec& b, int count, DevState* st,
              hipStream_t s) {
    Vec tmp = ws_.get(0, x.type, A_.n());
    auto tiles = A_.tileArgs();
    const void* off = A_.crsOff(DType::F32);
    const void* dg = A_.crsDiag(DType::F32);
    const int chk = st ? 1 : 0;
    auto sweep = [&](bool bwd, const void* xin, void* xout) {
      const gph::GsArgs& g = bwd ? bwd_ : fwd_;
      if (g.T > 0 && gph::gsStreamSharedBytes(g, x.type) <= kMaxGsShared)
        gph::launchGsStream(g, dg, b.ptr, xin, xout, x.type, st, chk, s);
      else
        gph::launchGsSweep(tiles, off, dg, DType::F32, b.ptr, xin, xout, x.type, bwd, st, chk, s);
    };
    for (int k = 0; k < count; k++) {
      sweep(false, x.ptr, tmp.ptr);
      if (cfg_->symmetric) {
        sweep(true, tmp.ptr, x.ptr);
      } else {
        gph::launchCast(tmp.ptr, x.type, x.ptr, x.type, A_.n(), s);
      }
    }
  }

  void applyDevice(const Vec& rhs, const Vec& out, const gph::PreOp& pre,
                   DevState* st, hipStream_t s) override {
    Vec b = rhs;
    if (pre.kind != gph::kPreNone) {
      gph::VecArgs v;
      v.n = A_.n();
      v.a = pre.r;
      v.c = pre.v;
      v.o1 = pre.p;
      v.state = st;
      if (pre.kind == gph::kPreP)
        gph::launchUpdateP(v, out.type, s);
      else
        gph::launchUpdateS(v, out.type, s);
      b = Vec{pre.p, out.type};
    }
    gph::launchFill(out.ptr, out.type, A_.n(), 0.0, s);  // usesInitialGuess
    int n = cfg_->numFixedIterations > 0 ? cfg_->numFixedIterations
                                         : std::max(1, cfg_->numSweeps);
    sweeps(out, b, n, st, s);
  }

  SolverStats solve(const Vec& x, const Vec& b) override {
    double t0 = nowSeconds();
    hipStream_t s = A_.stream();
    SolverStats st;
    st.solverName = name();
    st.norm = cfg_->norm;
    st.numTiles = A_.numTiles();
    if (cfg_->numFixedIterations > 0) {
      sweeps(x, b, cfg_->numFixedIterations, nullptr, s);
      GPH_HIP(hipStreamSynchronize(s));
      st.iterations = cfg_->numFixedIterations;
      st.seconds = nowSeconds() - t0;
      return st;
    }
    const int kind = normKind(cfg_->norm);
    double bNorm = 0;
    if (cfg_->relTolerance > 0) bNorm = A_.norm(b, kind, s);
    Vec r = ws_.get(1, x.type, A_.n());
    auto resNorm = [&]() {
      A_.residual(x, b, r, x.type, DType::F32, s);
      return A_.norm(r, kind, s);
    };
    st.initialResidual = resNorm();
    st.finalResidual = st.initialResidual;
    while (!((st.converged && st.iterations >= cfg_->minIterations) ||
             st.iterations >= cfg_->maxIterations)) {
      sweeps(x, b, std::max(1, cfg_->numSweeps), nullptr, s);
      st.iterations += std::max(1, cfg_->numSweeps);
      st.finalResidual = resNorm();
      st.converged = st.finalResidual < cfg_->absTolerance ||
                     st.finalResidual < st.initialResidual * cfg_->relResidual ||
                     (cfg_->relTolerance > 0 &&
                      st.finalResidual < bNorm * cfg_->relTolerance);
      if (cfg_->printPerformanceEachIteration) st.print();
    }
    st.seconds = nowSeconds() - t0;
    if (cfg_->printPerformanceAfterSolve) st.print();
    return st;
  }

 private:
  // Jagged-diagonal level stream of one sweep direction (gs_pack.hpp),
  // values gathered from the device's float32 matrix values; none when a
  // tile's x and a level's products do not fit in LDS (the row-block sweep
  // kernel runs then).
  void buildStream(bool backward, gph::GsArgs& a, DevBuf* buf) {
    if ((size_t)(A_.tileArgs().maxTileRows + 1) * 8 > kMaxGsShared) return;
    static_assert(kGsFresh == gph::kGsFreshBit, "stream format");
    // levels capped so that 8 workgroups (the most of 256 threads a CU
    // holds) fit a CU's LDS with float32 vectors: no tile of the last round
    // of a 16-tiles-per-CU launch waits for an LDS slot
    const int64_t maxRows = A_.tileArgs().maxTileRows;
    const int64_t perBlock = 160 * 1024 / 8 - (gph::kGsMaxWidth + 1) * 4;
    const int64_t cap = perBlock / 4 - maxRows >= 512 ? perBlock / 4 - maxRows : INT64_MAX;
    GsStream g = packGs(A_.hostMatrix(), backward, cap);
    if (!g.ok) return;
    hipStream_t s = A_.stream();
    buf[0].upload(g.lvlBase.data(), g.lvlBase.size(), s);
    buf[1].upload(g.lvl.data(), g.lvl.size(), s);
    buf[2].upload(g.diag.data(), g.diag.size(), s);
    buf[3].upload(g.slot.data(), g.slot.size(), s);
    buf[4].upload(g.code.data(), g.code.size(), s);
    buf[6].upload(g.pos.data(), g.pos.size(), s);
    DevBuf idx;
    idx.upload(g.src.data(), g.src.size(), s);
    buf[5].alloc(g.src.size() * sizeof(float));
    gph::launchGather(A_.crsOff(DType::F32), DType::F32, idx.as<int32_t>(), buf[5].get(),
                      (int64_t)g.src.size(), s);
    GPH_HIP(hipStreamSynchronize(s));
    a.T = A_.numTiles();
    a.maxRows = g.maxRows;
    a.tileStart = A_.tileArgs().tileStart;
    a.lvlBase = buf[0].as<int32_t>();
    a.lvl = buf[1].as<int32_t>();
    a.diagStart = buf[2].as<int32_t>();
    a.slot = buf[3].as<uint32_t>();
    a.code = buf[4].as<uint32_t>();
    a.val = buf[5].as<float>();
    a.pos = buf[6].as<uint16_t>();
    a.maxLevelEntries = g.maxLevelEntries;
    static_assert(gph::kGsMaxWidth == 255, "stream row length field");
    if (gph::gsStreamSharedBytes(a, DType::F32) > kMaxGsShared) a.T = 0;  // row-block kernel
  }

  static constexpr size_t kMaxGsShared = 156 * 1024;
  std::shared_ptr<gaussseidel::Configuration> cfg_;
  Workspace ws_;
  gph::GsArgs fwd_, bwd_;
  DevBuf fbuf_[7], bbuf_[7];
};

// ---------------------------------------------------------------------------
// PBiCGStab (pbicgstab/Solver.cpp:48-194). Per iteration five kernels:
//   K1 p-update fused into y = M^-1 p
//   K2 v = A y with r0.v            -> alpha
//   K3 s-update fused into z = M^-1 s
//   K4 t = A z with t.t and t.s     -> omega
//   K5 x, r update with r.r, r0.r   -> convergence, next rho
// ---------------------------------------------------------------------------
class PBiCGStabSolver : public Solver {
 public:
  PBiCGStabSolver(const DeviceMatrix& A,
                  std::shared_ptr<pbicgstab::Configuration> c)
      : Solver(A), cfg_(std::move(c)), kind_(normKind(cfg_->norm)) {
    if (cfg_->preconditioner) {
      pre_ = Solver::createSolver(A, cfg_->preconditioner);
      if (!pre_->supportsDeviceApply())
        fail(kUnsupported, pre_->name() + " cannot precondition PBiCGStab");
    }
    state_.alloc(sizeof(DevState));
    GPH_HIP(hipMemset(state_.get(), 0, sizeof(DevState)));
    partials_.alloc(A.partialsBytes());
    GPH_HIP(hipMemset(partials_.get(), 0, gph::kArriveBytes));  // arrival counters
  }
  std::string name() const override { return cfg_->solverName(); }
  bool usesInitialGuess() const override { return true; }

  SolverStats solve(const Vec& xIn, const Vec& bIn) override {
    const double t0 = nowSeconds();
    hipStream_t s = A_.stream();
    const int64_t n = A_.n();
    const DType X = cfg_->workingType.value_or(xIn.type);
    if (X != DType::F32 && X != DType::F64)
      fail(kUnsupported, "PBiCGStab working type must be float32 or float64");
    V_ = cfg_->matrixValueType.value_or(DType::F32);

    // bring x and b to the working type
    Vec x = xIn, b = bIn;
    if (x.type != X) {
      x = ws_.get(8, X, n);
      gph::launchCast(xIn.ptr, xIn.type, x.ptr, X, n, s);
    }
    if (b.type != X) {
      b = ws_.get(9, X, n);
      gph::launchCast(bIn.ptr, bIn.type, b.ptr, X, n, s);
    }
    r_ = ws_.get(0, X, n);
    r0_ = ws_.get(1, X, n);
    p_ = ws_.get(2, X, n);
    v_ = ws_.get(3, X, n);
    y_ = ws_.get(4, X, n);
    s_ = ws_.get(5, X, n);
    z_ = ws_.get(6, X, n);
    t_ = ws_.get(7, X, n);
    X_ = X;
    x_ = x;

    DevState init{};
    init.absTol = cfg_->absTolerance;
    init.relTol = cfg_->relTolerance;
    init.relRes = cfg_->relResidual;
    init.minIt = cfg_->minIterations;
    init.maxIt = cfg_->maxIterations;
    DevState* h = pinned_[2];
    *h = init;
    GPH_HIP(hipMemcpyAsync(state_.get(), h, sizeof(DevState),
                           hipMemcpyHostToDevice, s));
    DevState* st = state_.as<DevState>();

    // ||x|| (reference quirk, pbicgstab/Solver.cpp:69-70) or ||b||
    if (cfg_->relTolerance > 0) {
      gph::VecArgs v;
      v.n = (int32_t)n;
      v.a = cfg_->relToleranceFromB ? b.ptr : x.ptr;
      v.state = st;
      v.partials = partials_.get();
      if (kind_ == gph::kNormL2) {
        v.fin = gph::kFinBNorm;
        gph::launchNorm2(v, X, s);
      } else {
        v.fin = gph::kFinBNormV;
        gph::launchNormKind(v, X, kind_, s);
      }
    }
    // r = b - A x ; r0 = p = r ; ||r|| and rho (pbicgstab/Solver.cpp:66-83)
    {
      gph::SpmvArgs a = A_.spmvArgs(V_);
      a.x = x.ptr;
      a.b = b.ptr;
      a.out = r_.ptr;
      a.out2 = r0_.ptr;
      a.out3 = p_.ptr;
      a.state = st;
      a.partials = partials_.get();
      a.fin = gph::kFinInit;
      gph::launchSpmv(a, V_, X, X, X, X, gph::kModeResidual, gph::kRedSelf, s);
      if (kind_ != gph::kNormL2) {  // r0n = rF = ||r|| in the configured norm
        gph::VecArgs v;
        v.n = (int32_t)n;
        v.a = r_.ptr;
        v.state = st;
        v.partials = partials_.get();
        v.fin = gph::kFinInitNormV;
        gph::launchNormKind(v, X, kind_, s);
      }
    }

    int chunk = cfg_->iterationsPerSync > 0 ? cfg_->iterationsPerSync : 4;
    if (cfg_->printPerformanceEachIteration) chunk = 1;
    const bool pipelined = !cfg_->printPerformanceEachIteration;
    // pipelined polling: keep one chunk in flight while checking the previous
    hipEvent_t* ev = ev_.get();
    int cur = 0;
    int64_t launched = 0;
    int size[2] = {chunk, chunk};
    auto enqueueChunk = [&](int slot, int m) {
      for (int k = 0; k < m; k++) enqueueIteration(st, s);
      launched += m;
      size[slot] = m;
      GPH_HIP(hipMemcpyAsync(pinned_[slot], st, sizeof(DevState),
                             hipMemcpyDeviceToHost, s));
      GPH_HIP(hipEventRecord(ev[slot], s));
    };
    // Length of the next chunk from the residual history: iterations
    // enqueued after convergence exit at once but still cost a launch each
    // (about 5 us per kernel), so near the predicted end the chunks shrink
    // to one iteration. The prediction extrapolates the mean log-residual
    // decrease per iteration so far; it only sets chunk lengths, never
    // the iterates.
    int next = chunk;
    auto plan = [&](const DevState& h, int inFlight) {
      next = chunk;
      if (h.iter <= 0 || !(h.rF > 0) || !(h.rF < h.r0n)) return;
      double target = std::max((double)h.absTol, h.r0n * (double)h.relRes);
      if (h.relTol > 0) target = std::max(target, h.bNorm * (double)h.relTol);
      if (!(target > 0)) return;
      const double rate = std::log(h.r0n / h.rF) / h.iter;
      double m = std::log(h.rF / target) / rate;
      m = std::max(m, (double)(h.minIt - h.iter));
      const double left = m - inFlight;
      if (left < chunk) next = std::max(1, (int)std::ceil(left));
    };
    enqueueChunk(cur, chunk);
    SolverStats stats;
    stats.solverName = name();
    stats.norm = cfg_->norm;
    stats.numTiles = A_.numTiles();
    DevState last{};
    for (;;) {
      const int other = cur ^ 1;
      if (pipelined) enqueueChunk(other, next);
      GPH_HIP(hipEventSynchronize(ev[cur]));
      DevState hs = *pinned_[cur];
      if (pipelined) plan(hs, size[other]);
      if (cfg_->printPerformanceEachIteration) {
        stats.initialResidual = hs.r0n;
        stats.finalResidual = hs.rF;
        stats.iterations = hs.iter;
        stats.converged = hs.converged;
        stats.singular = hs.singular;
        stats.seconds = nowSeconds() - t0;
        stats.print();
      }
      if (hs.done) {
        last = hs;
        break;
      }
      if (pipelined)
        cur = other;
      else
        enqueueChunk(cur, chunk);
    }
    // The state read at convergence is final: kernels enqueued after it exit
    // before touching it.
    DevState hs = last;
    if (!deferSync_) {
      GPH_HIP(hipStreamSynchronize(s));
      GPH_HIP(hipMemcpy(pinned_[0], st, sizeof(DevState), hipMemcpyDeviceToHost));
      hs = *pinned_[0];
    }
    if (x.ptr != xIn.ptr) {
      gph::launchCast(x.ptr, X, xIn.ptr, xIn.type, n, s);
      if (!deferSync_) GPH_HIP(hipStreamSynchronize(s));
    }
    stats.initialResidual = hs.r0n;
    stats.finalResidual = hs.rF;
    stats.iterations = hs.iter;
    stats.converged = hs.converged != 0;
    stats.singular = hs.singular != 0;
    stats.seconds = nowSeconds() - t0;
    lastLaunched_ = launched;
    if (cfg_->printPerformanceAfterSolve) stats.print();
    return stats;
  }

  int64_t lastLaunchedIterations() const { return lastLaunched_; }
  void setDeferredSync(bool v) override { deferSync_ = v; }

 private:
  void enqueueIteration(DevState* st, hipStream_t s) {
    const int32_t n = A_.n();
    // K1: p = r + beta (p - omega v) ; y = M^-1 p
    gph::PreOp pp;
    pp.kind = gph::kPreP;
    pp.r = r_.ptr;
    pp.p = p_.ptr;
    pp.v = v_.ptr;
    if (pre_) {
      pre_->applyDevice(p_, y_, pp, st, s);
    } else {
      gph::VecArgs v;
      v.n = n;
      v.a = r_.ptr;
      v.c = v_.ptr;
      v.o1 = p_.ptr;
      v.o2 = y_.ptr;
      v.state = st;
      gph::launchUpdateP(v, X_, s);
    }
    // K2: v = A y ; alpha = rho / (r0 . v)
    {
      gph::SpmvArgs a = A_.spmvArgs(V_);
      a.x = y_.ptr;
      a.out = v_.ptr;
      a.w = r0_.ptr;
      a.state = st;
      a.partials = partials_.get();
      a.fin = gph::kFinAlpha;
      a.checkDone = 1;
      gph::launchSpmv(a, V_, X_, X_, X_, X_, gph::kModeSpmv, gph::kRedW, s);
    }
    // K3: s = r - alpha v ; z = M^-1 s
    gph::PreOp ps;
    ps.kind = gph::kPreS;
    ps.r = r_.ptr;
    ps.p = s_.ptr;
    ps.v = v_.ptr;
    if (pre_) {
      pre_->applyDevice(s_, z_, ps, st, s);
    } else {
      gph::VecArgs v;
      v.n = n;
      v.a = r_.ptr;
      v.c = v_.ptr;
      v.o1 = s_.ptr;
      v.o2 = z_.ptr;
      v.state = st;
      gph::launchUpdateS(v, X_, s);
    }
    // K4: t = A z ; omega = (t . s) / (t . t)
    {
      gph::SpmvArgs a = A_.spmvArgs(V_);
      a.x = z_.ptr;
      a.out = t_.ptr;
      a.w = s_.ptr;
      a.state = st;
      a.partials = partials_.get();
      a.fin = gph::kFinOmega;
      a.checkDone = 1;
      gph::launchSpmv(a, V_, X_, X_, X_, X_, gph::kModeSpmv, gph::kRedSelfW, s);
    }
    // K5: x += alpha y + omega z ; r = s - omega t ; ||r||, next rho
    {
      gph::VecArgs v;
      v.n = n;
      v.a = y_.ptr;
      v.b = z_.ptr;
      v.c = s_.ptr;
      v.d = t_.ptr;
      v.o1 = x_.ptr;
      v.o2 = r_.ptr;
      v.state = st;
      v.partials = partials_.get();
      // L2: the kernel's r.r gives ||r||; other norms: rho only, then a
      // norm kernel finishes the iteration
      if (kind_ != gph::kNormL2) v.fin = gph::kFinUpdateRho;
      gph::launchUpdateXR(v, r0_.ptr, X_, s);
      if (kind_ != gph::kNormL2) {
        gph::VecArgs w;
        w.n = n;
        w.a = r_.ptr;
        w.state = st;
        w.partials = partials_.get();
        w.fin = gph::kFinUpdateNormV;
        w.checkDone = 1;
        gph::launchNormKind(w, X_, kind_, s);
      }
    }
  }

  std::shared_ptr<pbicgstab::Configuration> cfg_;
  int kind_;  // gph::NormKind of cfg_->norm
  std::unique_ptr<Solver> pre_;
  Workspace ws_;
  DevBuf state_, partials_;
  PinnedState pinned_;
  DType X_ = DType::F32, V_ = DType::F32;
  Vec r_, r0_, p_, v_, y_, s_, z_, t_, x_;
  int64_t lastLaunched_ = 0;
  bool deferSync_ = false;
  struct Events {  // chunk completion events, reused across solves
    hipEvent_t e[2] = {nullptr, nullptr};
    Events() {
      GPH_HIP(hipEventCreateWithFlags(&e[0], hipEventDisableTiming));
      GPH_HIP(hipEventCreateWithFlags(&e[1], hipEventDisableTiming));
    }
    ~Events() {
      for (hipEvent_t v : e)
        if (v) (void)hipEventDestroy(v);
    }
    hipEvent_t* get() { return e; }
  } ev_;
};

// ---------------------------------------------------------------------------
// Mixed-precision iterative refinement (iterative-refinement/Solver.cpp:48-123).
// Extended precision float64, or twofloat32: phi in double words, the
// residual and the update in double-word arithmetic (k_dword.hip). The
// residual norm is the FP64 norm of the unrounded residual (the reference
// takes the norm of the residual after rounding to the working type,
// :80-81 and :110).
// ---------------------------------------------------------------------------
class IterativeRefinementSolver : public Solver {
 public:
  IterativeRefinementSolver(const DeviceMatrix& A,
                            std::shared_ptr<iterativerefinement::Configuration> c)
      : Solver(A), cfg_(std::move(c)) {
    inner_ = Solver::createSolver(A, cfg_->innerSolver);
    inner_->setDeferredSync(true);  // every outer step ends in a synchronizing read
    state_.alloc(sizeof(DevState));
    GPH_HIP(hipMemset(state_.get(), 0, sizeof(DevState)));
    partials_.alloc(A.partialsBytes());
    GPH_HIP(hipMemset(partials_.get(), 0, gph::kArriveBytes));  // arrival counters
  }
  std::string name() const override { return cfg_->solverName(); }
  bool usesInitialGuess() const override { return true; }

  SolverStats solve(const Vec& x, const Vec& b) override {
    const double t0 = nowSeconds();
    hipStream_t s = A_.stream();
    const int64_t n = A_.n();
    const DType E = cfg_->extendedPrecisionType.value_or(x.type);
    // x in twofloat32 is the solution itself (phi); the working type is then
    // float32 unless configured
    const DType W = cfg_->workingType.value_or(x.type == DType::TF32 ? DType::F32 : x.type);
    const DType V = cfg_->matrixValueType.value_or(DType::F32);
    const bool dw = E == DType::TF32;  // double-word extended precision
    if (dw && (W != DType::F32 || V != DType::F32 || b.type != DType::F32))
      fail(kUnsupported,
           "twofloat32 refinement: working type, matrix values and b must be float32");
    Vec phi = x;
    if (x.type != E) {
      phi = ws_.get(0, E, n);
      gph::launchCast(x.ptr, x.type, phi.ptr, E, n, s);
    }
    Vec r = ws_.get(1, W, n);
    Vec d = ws_.get(2, W, n);
    DevState* st = state_.as<DevState>();

    SolverStats stats;
    stats.solverName = name();
    stats.norm = cfg_->norm;
    stats.numTiles = A_.numTiles();
    const int kind = normKind(cfg_->norm);
    double bNorm = 0;
    if (cfg_->relTolerance > 0) bNorm = A_.norm(b, kind, s);
    // the unrounded residual (float64), for norms other than L2
    Vec r64 = kind != gph::kNormL2 ? ws_.get(3, DType::F64, n) : Vec{};

    // r = fl_W(b - A phi) accumulated in E, norm of the unrounded residual
    auto residual = [&]() {
      gph::SpmvArgs a = A_.spmvArgs(V);
      a.x = phi.ptr;
      a.b = b.ptr;
      a.out = r.ptr;
      a.state = st;
      a.partials = partials_.get();
      a.fin = gph::kFinNormAux;
      if (r64.ptr) a.out64 = static_cast<double*>(r64.ptr);
      if (dw)
        gph::launchResidualDW(a, s);
      else
        gph::launchSpmv(a, V, E, b.type, W, E, gph::kModeResidual, gph::kRedSelf, s);
      if (r64.ptr) {
        gph::VecArgs v;
        v.n = (int32_t)n;
        v.a = r64.ptr;
        v.state = st;
        v.partials = partials_.get();
        v.fin = gph::kFinNormAuxV;
        gph::launchNormKind(v, DType::F64, kind, s);
      }
      double out = 0;
      GPH_HIP(hipMemcpyAsync(&out, &st->aux, sizeof(double),
                             hipMemcpyDeviceToHost, s));
      GPH_HIP(hipStreamSynchronize(s));
      return out;
    };
    stats.initialResidual = residual();
    stats.finalResidual = stats.initialResidual;
    while (!((stats.converged && stats.iterations >= cfg_->minIterations) ||
             stats.iterations >= cfg_->maxIterations)) {
      stats.iterations++;
      if (inner_->usesInitialGuess()) gph::launchFill(d.ptr, W, n, 0.0, s);
      SolverStats in = inner_->solve(d, r);
      stats.innerIterations += in.iterations;
      if (E == DType::F64) {
        gph::launchAccumulate(static_cast<double*>(phi.ptr), d.ptr, W, n, s);
      } else if (dw) {
        gph::launchAccumulateDW(phi.ptr, d.ptr, W, n, s);
      } else {
        fail(kUnsupported, "IterativeRefinement needs float64 or twofloat32 extended precision");
      }
      stats.finalResidual = residual();
      stats.converged =
          stats.finalResidual < (double)cfg_->absTolerance ||
          stats.finalResidual < stats.initialResidual * (double)cfg_->relResidual ||
          (cfg_->relTolerance > 0 &&
           stats.finalResidual < bNorm * (double)cfg_->relTolerance);
      if (cfg_->printPerformanceEachIteration) {
        stats.seconds = nowSeconds() - t0;
        stats.print();
      }
    }
    if (phi.ptr != x.ptr) gph::launchCast(phi.ptr, E, x.ptr, x.type, n, s);
    GPH_HIP(hipStreamSynchronize(s));
    stats.seconds = nowSeconds() - t0;
    if (cfg_->printPerformanceAfterSolve) stats.print();
    return stats;
  }

 private:
  std::shared_ptr<iterativerefinement::Configuration> cfg_;
  std::unique_ptr<Solver> inner_;
  Workspace ws_;
  DevBuf state_, partials_;
};

// restarter/Solver.cpp:40-56
class RestarterSolver : public Solver {
 public:
  RestarterSolver(const DeviceMatrix& A, std::shared_ptr<restarter::Configuration> c)
      : Solver(A), cfg_(std::move(c)) {
    inner_ = Solver::createSolver(A, cfg_->innerSolver);
  }
  std::string name() const override { return cfg_->solverName(); }
  bool usesInitialGuess() const override { return inner_->usesInitialGuess(); }
  SolverStats solve(const Vec& x, const Vec& b) override {
    SolverStats st = inner_->solve(x, b);
    int restarts = 0;
    int64_t total = st.iterations;
    while (restarts < cfg_->maxRestarts && !st.converged) {
      restarts++;
      st = inner_->solve(x, b);
      total += st.iterations;
    }
    st.innerIterations = total;
    return st;
  }

 private:
  std::shared_ptr<restarter::Configuration> cfg_;
  std::unique_ptr<Solver> inner_;
};

// Solver.cpp:32-56
std::unique_ptr<Solver> Solver::createSolver(const DeviceMatrix& A,
                                             std::shared_ptr<Configuration> c) {
  if (auto g = std::dynamic_pointer_cast<gaussseidel::Configuration>(c))
    return std::make_unique<GaussSeidelSolver>(A, g);
  if (auto ir = std::dynamic_pointer_cast<iterativerefinement::Configuration>(c))
    return std::make_unique<IterativeRefinementSolver>(A, ir);
  if (auto il = std::dynamic_pointer_cast<ilu::Configuration>(c))
    return std::make_unique<IluSolver>(A, il);
  if (auto pb = std::dynamic_pointer_cast<pbicgstab::Configuration>(c))
    return std::make_unique<PBiCGStabSolver>(A, pb);
  if (auto rs = std::dynamic_pointer_cast<restarter::Configuration>(c))
    return std::make_unique<RestarterSolver>(A, rs);
  fail(kInvalid, "Unknown solver: " + (c ? c->solverName() : std::string("null")));
}

}  // namespace graphene::matrix::solver
