// Batched solves (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <chrono>

#include "config.hpp"
#include "ilu_pack.hpp"
#include "kernels.hpp"

namespace graphene::batch {

using gph::DType;
using gph::fail;
using graphene::Json;
using matrix::Vec;
namespace solver = matrix::solver;

namespace {
constexpr int kGraphIterations = 4;  // inner iterations per graph launch
constexpr int kOuterGraphSteps = 8;  // steps per graph launch, outer step on the device

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

gph::DevState stateFor(float absTol, float relRes, int minIt, int maxIt) {
  gph::DevState s{};
  s.absTol = absTol;
  s.relTol = 0.0f;
  s.relRes = relRes;
  s.minIt = minIt;
  s.maxIt = maxIt;
  return s;
}
}  // namespace

// Configuration: MPIR(PBiCGStab(ILU)) or PBiCGStab(ILU); sets the outer and
// inner initial states and returns the ILU configuration.
std::shared_ptr<solver::ilu::Configuration> BatchSolver::parseConfig(
    const std::string& configJson) {
  auto cfg = solver::Configuration::fromJSON(configJson);
  // "outerStep": "device" (default) or "host": where the refinement's outer
  // step runs (a key of the batch only)
  if (const Json* o = Json::parse(configJson).get("outerStep")) {
    if (o->kind != Json::String || (o->str != "device" && o->str != "host"))
      fail(gph::kInvalid, "outerStep: \"device\" or \"host\"");
    outerOnDevice_ = o->str == "device";
  }
  std::shared_ptr<solver::pbicgstab::Configuration> pb;
  if (auto ir = std::dynamic_pointer_cast<solver::iterativerefinement::Configuration>(cfg)) {
    mpir_ = true;
    if (ir->extendedPrecisionType.value_or(DType::F64) != DType::F64 ||
        ir->workingType.value_or(DType::F32) != DType::F32 || ir->relTolerance > 0)
      fail(gph::kUnsupported,
           "batched refinement: float64 extended precision, float32 working type, no "
           "relTolerance");
    outerInit_ = stateFor(ir->absTolerance, ir->relResidual, ir->minIterations,
                          ir->maxIterations);
    pb = std::dynamic_pointer_cast<solver::pbicgstab::Configuration>(ir->innerSolver);
  } else {
    pb = std::dynamic_pointer_cast<solver::pbicgstab::Configuration>(cfg);
  }
  if (!pb) fail(gph::kUnsupported, "batched solves use PBiCGStab (optionally inside refinement)");
  if (pb->workingType.value_or(DType::F32) != DType::F32 || pb->relTolerance > 0 ||
      pb->matrixValueType.value_or(DType::F32) != DType::F32)
    fail(gph::kUnsupported, "batched PBiCGStab: float32 working type and values, no relTolerance");
  auto ilu = std::dynamic_pointer_cast<solver::ilu::Configuration>(pb->preconditioner);
  if (!ilu) fail(gph::kUnsupported, "batched PBiCGStab needs an ILU preconditioner");
  innerInit_ = stateFor(pb->absTolerance, pb->relResidual, pb->minIterations, pb->maxIterations);

  return ilu;
}

// Chunks (consecutive rows of one system, at most kBatchChunk each), states,
// vectors of sysRow_[S] rows and the reduction scratch.
void BatchSolver::setupChunks(hipStream_t st) {
  // chunks: consecutive rows of one system, at most kBatchChunk each
  std::vector<int32_t> chunkSys, chunkRow, sysChunk((size_t)S_ + 1);
  for (int32_t s = 0; s < S_; s++) {
    sysChunk[(size_t)s] = (int32_t)chunkSys.size();
    for (int64_t r = sysRow_[(size_t)s]; r < sysRow_[(size_t)s + 1]; r += gph::kBatchChunk) {
      chunkSys.push_back(s);
      chunkRow.push_back((int32_t)r);
    }
  }
  sysChunk[(size_t)S_] = (int32_t)chunkSys.size();
  chunkRow.push_back((int32_t)sysRow_[(size_t)S_]);
  C_ = (int32_t)chunkSys.size();
  chunkSys_.upload(chunkSys.data(), chunkSys.size(), st);
  chunkRow_.upload(chunkRow.data(), chunkRow.size(), st);
  sysChunk_.upload(sysChunk.data(), sysChunk.size(), st);
  inner_.alloc((size_t)S_ * sizeof(gph::DevState));
  outer_.alloc((size_t)S_ * sizeof(gph::DevState));
  glob_.alloc(sizeof(gph::DevState));
  partials_.alloc((size_t)2 * C_ * sizeof(double));
  const size_t n = (size_t)sysRow_[(size_t)S_];
  for (gph::DevBuf* b : {&r_, &r0_, &p_, &v_, &y_, &s_, &z_, &t_, &d_, &rw_}) b->alloc(n * 4);
  if (mpir_) r64_.alloc(n * 8);
  if (mpir_ && !multi_) {
    phi_.alloc(n * 8);
    b_.alloc(n * 4);
  }
  phase_.alloc((size_t)S_ * sizeof(int32_t));
  args_.S = S_;
  args_.C = C_;
  args_.chunkSys = chunkSys_.as<int32_t>();
  args_.chunkRow = chunkRow_.as<int32_t>();
  args_.sysChunk = sysChunk_.as<int32_t>();
  args_.st = inner_.as<gph::DevState>();
  args_.glob = glob_.as<gph::DevState>();
  args_.partials = partials_.get();
  args_.phase = phase_.as<int32_t>();
  GPH_HIP(hipHostMalloc((void**)&pinned_, 2 * sizeof(gph::DevState), 0));
  gph::DevState zero{};
  GPH_HIP(hipMemcpyAsync(glob_.get(), &zero, sizeof(zero), hipMemcpyHostToDevice, st));
}

BatchSolver::BatchSolver(const matrix::DeviceMatrix& A, const std::vector<int32_t>& sysTile,
                         const std::string& configJson)
    : A_(A) {
  const auto& h = A.hostMatrix();
  S_ = (int32_t)sysTile.size() - 1;
  if (S_ < 1) fail(gph::kInvalid, "a batch needs at least one system");
  if (sysTile.front() != 0 || sysTile.back() != h.numTiles)
    fail(gph::kInvalid, "system tile ranges must cover the matrix's tiles");
  for (int32_t s = 0; s < S_; s++)
    if (sysTile[(size_t)s + 1] <= sysTile[(size_t)s])
      fail(gph::kInvalid, "every system needs at least one tile");
  sysRow_.resize((size_t)S_ + 1);
  for (int32_t s = 0; s <= S_; s++) sysRow_[(size_t)s] = h.tileStart[(size_t)sysTile[(size_t)s]];
  // block-diagonal: every entry of a system's rows lies in the system
  for (int32_t s = 0; s < S_; s++)
    for (int64_t i = sysRow_[(size_t)s]; i < sysRow_[(size_t)s + 1]; i++)
      for (int64_t a = h.rowPtr[(size_t)i]; a < h.rowPtr[(size_t)i + 1]; a++) {
        const int64_t j = h.col[(size_t)a];
        if (j < sysRow_[(size_t)s] || j >= sysRow_[(size_t)s + 1])
          fail(gph::kInvalid, "a matrix entry couples two systems of the batch");
      }

  auto ilu = parseConfig(configJson);
  hipStream_t st = A.stream();
  setupChunks(st);
  const size_t n = (size_t)A.n();

  // ILU substitution kernel: the configured one, or the fastest on this
  // union (three timed applications each) of the row-per-lane kernel and
  // the packed kernel with a few rows-per-wave groupings
  const bool tune = configJson.find("rowsPerWave") == std::string::npos &&
                    configJson.find("substitution") == std::string::npos;
  std::vector<std::pair<std::string, int32_t>> cands;
  if (tune) {
    cands.push_back({"rows", ilu->rowsPerWave});
    // packed groupings that differ (equal groupings are tried once)
    std::vector<std::vector<int32_t>> seen;
    for (int32_t c : {512, 1024, 2048, 4096}) {
      std::vector<int32_t> gr = matrix::groupTiles(h, c);
      if (std::find(seen.begin(), seen.end(), gr) == seen.end()) {
        seen.push_back(std::move(gr));
        cands.push_back({"packed", c});
      }
    }
  } else {
    cands.push_back({ilu->substitution, ilu->rowsPerWave});
  }
  double best = 1e30;
  for (const auto& [sub, c] : cands) {
    auto ic = std::make_shared<solver::ilu::Configuration>(*ilu);
    ic->substitution = sub;
    ic->rowsPerWave = c;
    auto S = solver::Solver::createSolver(A, ic);
    const Vec in{r_.get(), DType::F32}, o{y_.get(), DType::F32};
    gph::launchFill(r_.get(), DType::F32, (int64_t)n, 1.0, st);
    S->applyDevice(in, o, gph::PreOp{}, nullptr, st);  // warm-up
    GPH_HIP(hipStreamSynchronize(st));
    const double t0 = now();
    for (int k = 0; k < 3; k++) S->applyDevice(in, o, gph::PreOp{}, nullptr, st);
    GPH_HIP(hipStreamSynchronize(st));
    const double dt = now() - t0;
    if (dt < best) {
      best = dt;
      rowsPerWave_ = sub == "packed" ? c : 0;
      substitution_ = sub;
      ilu_ = std::move(S);
    }
  }
  // work of finished systems is skipped: ILU wave groups and SpMV slices
  // with no inner solve running return at once
  ilu_->setSystemSkip(args_.phase, sysRow_);
  {
    const gph::SpmvArgs a = A.spmvArgs(DType::F32);
    if (a.numSlices > 0) {
      auto sysOf = [&](int64_t r) {
        return (int32_t)(std::upper_bound(sysRow_.begin(), sysRow_.end(), r) - sysRow_.begin() - 1);
      };
      std::vector<int32_t> ss(2 * (size_t)a.numSlices);
      for (int64_t sl = 0; sl < a.numSlices; sl++) {
        // rows of the slice: 64 consecutive rows, or (rows sorted by length
        // inside 4096-row windows) the slice's window
        int64_t r0 = sl * 64, r1 = std::min<int64_t>(n, r0 + 64);
        if (a.perm) {
          r0 = (r0 / 4096) * 4096;
          r1 = std::min<int64_t>(n, r0 + 4096);
        }
        ss[2 * (size_t)sl] = sysOf(std::min<int64_t>(r0, n - 1));
        ss[2 * (size_t)sl + 1] = sysOf(std::max<int64_t>(r0, r1 - 1));
      }
      sliceSys_.upload(ss.data(), ss.size(), st);
    }
  }
  // warm-up of the SpMV launch path outside any capture
  A.spmv(Vec{y_.get(), DType::F32}, Vec{v_.get(), DType::F32}, DType::F32, DType::F32, st);
  GPH_HIP(hipStreamSynchronize(st));
}

BatchSolver::BatchSolver(const matrix::DeviceMatrix& A, int32_t nrhs,
                         const std::string& configJson,
                         std::shared_ptr<solver::Solver> ilu)
    : A_(A), multi_(true) {
  if (nrhs < 1 || nrhs > 8)
    fail(gph::kInvalid, "1 to 8 right-hand sides, not " + std::to_string(nrhs));
  if (A.sellArgs(DType::F32).numSlices <= 0)
    fail(gph::kUnsupported, "several right-hand sides need the sliced matrix layout");
  S_ = nrhs;
  const int64_t n = A.n();
  sysRow_.resize((size_t)S_ + 1);
  for (int32_t j = 0; j <= S_; j++) sysRow_[(size_t)j] = j * n;
  if ((int64_t)S_ * n >= (int64_t)INT32_MAX)
    fail(gph::kInvalid, "right-hand sides times rows exceed the int32 row range");
  auto iluCfg = parseConfig(configJson);
  GPH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  setupChunks(stream_);
  if (ilu) {
    ilu_ = std::move(ilu);
  } else {
    ilu_ = solver::Solver::createSolver(A, iluCfg);
  }
  substitution_ = iluCfg->substitution;
  // warm-up of the launch paths outside any capture
  const Vec vy{y_.get(), DType::F32}, vv{v_.get(), DType::F32};
  gph::launchFill(y_.get(), DType::F32, (int64_t)S_ * n, 0.0, stream_);
  gph::SpmvArgs a = A.sellArgs(DType::F32);
  a.x = y_.get();
  a.out = v_.get();
  gph::launchSpmvMulti(a, S_, n, nullptr, nullptr, stream_);
  ilu_->applyDeviceMulti(vv, vy, S_, n, nullptr, nullptr, stream_);
  GPH_HIP(hipStreamSynchronize(stream_));
}

size_t BatchSolver::deviceBytes() const {
  size_t b = 0;
  for (const gph::DevBuf* q : {&chunkSys_, &chunkRow_, &sysChunk_, &inner_, &outer_, &glob_,
                               &partials_, &sliceSys_, &r_, &r0_, &p_, &v_, &y_, &s_, &z_, &t_,
                               &d_, &rw_, &r64_, &phase_, &phi_, &b_})
    b += q->bytes();
  return b;
}

BatchSolver::~BatchSolver() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (outerExec_) (void)hipGraphExecDestroy(outerExec_);
  if (outerGraph_) (void)hipGraphDestroy(outerGraph_);
  if (pinned_) (void)hipHostFree(pinned_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// One PBiCGStab iteration for every running system (pbicgstab/Solver.cpp:91-180).
void BatchSolver::enqueueIteration(hipStream_t s) {
  const Vec vp{p_.get(), DType::F32}, vy{y_.get(), DType::F32}, vv{v_.get(), DType::F32};
  const Vec vs{s_.get(), DType::F32}, vz{z_.get(), DType::F32}, vt{t_.get(), DType::F32};
  gph::DevState* glob = glob_.as<gph::DevState>();
  const int64_t n = A_.n();
  auto spmv = [&](const Vec& x, const Vec& y) {
    if (multi_) {  // every right-hand side in one pass over the matrix
      gph::SpmvArgs a = A_.sellArgs(DType::F32);
      a.x = x.ptr;
      a.out = y.ptr;
      gph::launchSpmvMulti(a, S_, n, args_.st, glob, s);
      return;
    }
    gph::SpmvArgs a = A_.spmvArgs(DType::F32);
    a.x = x.ptr;
    a.out = y.ptr;
    a.state = glob;
    a.checkDone = 1;
    if (sliceSys_) {
      a.sysPhase = args_.phase;
      a.sliceSys = sliceSys_.as<int32_t>();
      a.runPhase = 0;
    }
    gph::launchSpmv(a, DType::F32, DType::F32, DType::F32, DType::F32, DType::F32,
                    gph::kModeSpmv, gph::kRedNone, s);
  };
  float* r = r_.as<float>();
  gph::launchBatchUpdateP(args_, r, p_.as<float>(), v_.as<float>(), s);
  auto ilu = [&](const Vec& in, const Vec& out) {
    if (multi_)
      ilu_->applyDeviceMulti(in, out, S_, n, args_.st, glob, s);
    else
      ilu_->applyDevice(in, out, gph::PreOp{}, glob, s);
  };
  ilu(vp, vy);
  spmv(vy, vv);
  gph::launchBatchDot1(args_, r0_.get(), v_.get(), s);
  gph::launchBatchFin(args_, gph::kFinAlpha, 1, s);
  gph::launchBatchUpdateS(args_, r, s_.as<float>(), v_.as<float>(), s);
  ilu(vs, vz);
  spmv(vz, vt);
  gph::launchBatchDot2(args_, t_.get(), t_.get(), t_.get(), s_.get(), s);
  gph::launchBatchFin(args_, gph::kFinOmega, 2, s);
  gph::launchBatchUpdateXR(args_, xInner_, r, y_.as<float>(), z_.as<float>(), s_.as<float>(),
                           t_.as<float>(), r0_.as<float>(), s);
  gph::launchBatchFin(args_, gph::kFinUpdate, 2, s);
}

void BatchSolver::captureGraph(hipStream_t s) {
  GPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kGraphIterations; k++) enqueueIteration(s);
  GPH_HIP(hipStreamEndCapture(s, &graph_));
  GPH_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

// Refinement step of the systems whose inner solve has finished (phase 1),
// on the device: phi += d, the FP64 residual of their SpMV slices, norm and
// convergence, inner restart (first: the initial residual of every system,
// nothing accumulated).
void BatchSolver::enqueueOuterStep(hipStream_t s, int first) {
  gph::DevState* glob = glob_.as<gph::DevState>();
  if (!first) gph::launchBatchPhaseAccumulate(args_, phi_.as<double>(), d_.as<float>(), s);
  gph::SpmvArgs a = A_.spmvArgs(DType::F32);
  a.x = phi_.get();
  a.b = b_.get();
  a.out = rw_.get();
  a.out64 = r64_.as<double>();
  a.state = glob;
  a.checkDone = 1;
  a.sysPhase = args_.phase;
  a.sliceSys = sliceSys_.as<int32_t>();
  a.runPhase = 1;
  gph::launchSpmv(a, DType::F32, DType::F64, DType::F32, DType::F32, DType::F64,
                  gph::kModeResidual, gph::kRedNone, s);
  gph::launchBatchPhaseOuter(args_, outer_.as<gph::DevState>(), r64_.as<double>(), first, s);
  gph::launchBatchPhaseRestart(args_, rw_.as<float>(), d_.as<float>(), r_.as<float>(),
                               r0_.as<float>(), p_.as<float>(), s);
}

// Batched refinement with the outer step on the device: every system moves
// between its inner solve (phase 0), its outer step (phase 1) and finished
// (phase 2) on its own; a captured graph of kOuterGraphSteps steps (one
// PBiCGStab iteration of the running systems, then the outer step of the
// systems that need one) is replayed until every system has finished.
void BatchSolver::solveDeviceOuter(const Vec& x, const Vec& b) {
  hipStream_t s = stream();
  const int64_t n = sysRow_[(size_t)S_];
  args_.deviceOuter = 1;
  std::vector<gph::DevState> init((size_t)S_, innerInit_);
  std::vector<gph::DevState> oinit((size_t)S_, outerInit_);
  std::vector<int32_t> ph((size_t)S_, 1);
  gph::DevState g0{};
  g0.iter = S_;  // every system's first outer step (initial residual) is due
  GPH_HIP(hipMemcpyAsync(inner_.get(), init.data(), init.size() * sizeof(gph::DevState),
                         hipMemcpyHostToDevice, s));
  GPH_HIP(hipMemcpyAsync(outer_.get(), oinit.data(), oinit.size() * sizeof(gph::DevState),
                         hipMemcpyHostToDevice, s));
  GPH_HIP(hipMemcpyAsync(phase_.get(), ph.data(), ph.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, s));
  GPH_HIP(hipMemcpyAsync(glob_.get(), &g0, sizeof(g0), hipMemcpyHostToDevice, s));
  GPH_HIP(hipMemcpyAsync(b_.get(), b.ptr, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  gph::launchFill(phi_.get(), DType::F64, n, 0.0, s);
  xInner_ = d_.as<float>();
  enqueueOuterStep(s, 1);  // initial residuals, first inner starts
  if (!outerExec_) {
    GPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kOuterGraphSteps; k++) {
      enqueueIteration(s);
      enqueueOuterStep(s, 0);
    }
    GPH_HIP(hipStreamEndCapture(s, &outerGraph_));
    GPH_HIP(hipGraphInstantiate(&outerExec_, outerGraph_, nullptr, nullptr, 0));
  }
  // two graph launches in flight; the flag copy of the older one is checked
  hipEvent_t ev[2];
  GPH_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  GPH_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  auto launch = [&](int slot) {
    GPH_HIP(hipGraphLaunch(outerExec_, s));
    launches_++;
    GPH_HIP(hipMemcpyAsync(pinned_ + slot, glob_.get(), sizeof(gph::DevState),
                           hipMemcpyDeviceToHost, s));
    GPH_HIP(hipEventRecord(ev[slot], s));
  };
  int cur = 0;
  launch(cur);
  for (;;) {
    launch(cur ^ 1);
    GPH_HIP(hipEventSynchronize(ev[cur]));
    if (pinned_[cur].done) break;
    cur ^= 1;
  }
  GPH_HIP(hipMemcpyAsync(x.ptr, phi_.get(), (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  GPH_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  args_.deviceOuter = 0;
}

// Batched PBiCGStab from x = 0 on the systems whose inner state is not
// finished; adds each system's iteration count to innerIts.
void BatchSolver::innerSolve(const void* rhs, void* x, std::vector<int64_t>& innerIts) {
  hipStream_t s = stream();
  xInner_ = static_cast<float*>(x);  // always d_ (captured in the graph)
  gph::launchBatchPhaseSync(args_, s);
  gph::launchBatchStart(args_, static_cast<const float*>(rhs), static_cast<float*>(x),
                        r_.as<float>(), r0_.as<float>(), p_.as<float>(), s);
  gph::launchBatchFin(args_, gph::kFinInit, 1, s);
  if (!exec_) captureGraph(s);
  // two graph launches in flight; the flag copy of the older one is checked
  hipEvent_t ev[2];
  GPH_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  GPH_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  auto launch = [&](int slot) {
    GPH_HIP(hipGraphLaunch(exec_, s));
    launches_++;
    GPH_HIP(hipMemcpyAsync(pinned_ + slot, glob_.get(), sizeof(gph::DevState),
                           hipMemcpyDeviceToHost, s));
    GPH_HIP(hipEventRecord(ev[slot], s));
  };
  int cur = 0;
  launch(cur);
  for (;;) {
    launch(cur ^ 1);
    GPH_HIP(hipEventSynchronize(ev[cur]));
    if (pinned_[cur].done) break;
    cur ^= 1;
  }
  GPH_HIP(hipStreamSynchronize(s));
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  std::vector<gph::DevState> st((size_t)S_);
  GPH_HIP(hipMemcpyAsync(st.data(), inner_.get(), st.size() * sizeof(gph::DevState),
                         hipMemcpyDeviceToHost, s));
  GPH_HIP(hipStreamSynchronize(s));
  for (int32_t k = 0; k < S_; k++) innerIts[(size_t)k] += st[(size_t)k].iter;
}

std::vector<SystemStats> BatchSolver::solve(const Vec& x, const Vec& b) {
  const double t0 = now();
  hipStream_t s = stream();
  const int64_t n = sysRow_[(size_t)S_];  // rows of all systems (right-hand sides)
  if (b.type != DType::F32) fail(gph::kUnsupported, "batched solve: b must be float32");
  if (x.type != (mpir_ ? DType::F64 : DType::F32))
    fail(gph::kUnsupported, mpir_ ? "batched refinement: x must be float64"
                                  : "batched PBiCGStab: x must be float32");
  std::vector<gph::DevState> init((size_t)S_, innerInit_);
  GPH_HIP(hipMemcpyAsync(inner_.get(), init.data(), init.size() * sizeof(gph::DevState),
                         hipMemcpyHostToDevice, s));
  std::vector<int64_t> innerIts((size_t)S_, 0);
  std::vector<SystemStats> out((size_t)S_);
  if (!mpir_) {
    innerSolve(b.ptr, d_.get(), innerIts);
    GPH_HIP(hipMemcpyAsync(x.ptr, d_.get(), (size_t)n * 4, hipMemcpyDeviceToDevice, s));
    std::vector<gph::DevState> st((size_t)S_);
    GPH_HIP(hipMemcpyAsync(st.data(), inner_.get(), st.size() * sizeof(gph::DevState),
                           hipMemcpyDeviceToHost, s));
    GPH_HIP(hipStreamSynchronize(s));
    for (int32_t k = 0; k < S_; k++) {
      const auto& q = st[(size_t)k];
      out[(size_t)k] = SystemStats{q.iter, 0, q.r0n, q.rF, q.converged != 0, q.singular != 0};
    }
    seconds_ = now() - t0;
    return out;
  }
  if (!multi_ && sliceSys_ && outerOnDevice_) {
    // refinement, outer step on the device
    solveDeviceOuter(x, b);
    std::vector<gph::DevState> os((size_t)S_);
    GPH_HIP(hipMemcpy(os.data(), outer_.get(), os.size() * sizeof(gph::DevState),
                      hipMemcpyDeviceToHost));
    for (int32_t k = 0; k < S_; k++) {
      const auto& q = os[(size_t)k];
      out[(size_t)k] = SystemStats{q.iter, q.innerTotal, q.r0n, q.rF, q.converged != 0, false};
    }
    seconds_ = now() - t0;
    return out;
  }
  // refinement, outer step on the host (several right-hand sides of one
  // system, or a matrix without the sliced layout): phi = 0 (x is
  // overwritten), r = fl32(b - A phi) in FP64
  std::vector<gph::DevState> oinit((size_t)S_, outerInit_);
  GPH_HIP(hipMemcpyAsync(outer_.get(), oinit.data(), oinit.size() * sizeof(gph::DevState),
                         hipMemcpyHostToDevice, s));
  gph::launchFill(x.ptr, DType::F64, n, 0.0, s);
  gph::DevState* outer = outer_.as<gph::DevState>();
  auto residual = [&](int first) {
    // the matrix's own rows once (a batch), or once per right-hand side
    const int64_t m = A_.n();
    for (int32_t j = 0; j < (multi_ ? S_ : 1); j++) {
      gph::SpmvArgs a = A_.spmvArgs(DType::F32);
      a.x = static_cast<double*>(x.ptr) + j * m;
      a.b = static_cast<const float*>(b.ptr) + j * m;
      a.out = rw_.as<float>() + j * m;
      a.out64 = r64_.as<double>() + j * m;
      gph::launchSpmv(a, DType::F32, DType::F64, DType::F32, DType::F32, DType::F64,
                      gph::kModeResidual, gph::kRedNone, s);
    }
    gph::launchBatchOuter(args_, outer, r64_.as<double>(), first, s);
  };
  residual(1);
  std::vector<gph::DevState> os((size_t)S_);
  for (;;) {
    GPH_HIP(hipMemcpyAsync(os.data(), outer_.get(), os.size() * sizeof(gph::DevState),
                           hipMemcpyDeviceToHost, s));
    GPH_HIP(hipStreamSynchronize(s));
    bool any = false;
    for (const auto& q : os) any = any || !q.done;
    if (!any) break;
    // inner solve on the systems still refining (their inner states were
    // re-opened by the outer step), then phi += d and the new residual
    GPH_HIP(hipMemcpyAsync(inner_.get(), init.data(), init.size() * sizeof(gph::DevState),
                           hipMemcpyHostToDevice, s));
    gph::launchBatchOuterMask(args_, outer, s);
    innerSolve(rw_.get(), d_.get(), innerIts);
    gph::launchBatchAccumulate(args_, outer, static_cast<double*>(x.ptr), d_.as<float>(), s);
    residual(0);
  }
  for (int32_t k = 0; k < S_; k++) {
    const auto& q = os[(size_t)k];
    out[(size_t)k] =
        SystemStats{q.iter, innerIts[(size_t)k], q.r0n, q.rF, q.converged != 0, false};
  }
  seconds_ = now() - t0;
  return out;
}

}  // namespace graphene::batch
