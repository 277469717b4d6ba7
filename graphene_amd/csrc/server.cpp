// Request server (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <thread>

#include "batch.hpp"
#include "config.hpp"

namespace graphene::serve {

using gph::DevBuf;
using gph::DType;
using matrix::DeviceMatrix;
using matrix::Vec;
using Clock = std::chrono::steady_clock;

namespace {
constexpr int32_t kMaxRhs = 8;
}

// One request's workspace: a block solver of k right-hand sides (or a
// single-system solver for the sequential path) and the request buffers
// x_true (f64), A x_true (f64), b and x, k blocks of n rows each.
struct Server::Slot {
  int32_t k = 1;
  size_t bytes = 0;
  uint64_t lastUse = 0;
  std::unique_ptr<batch::BatchSolver> block;
  std::unique_ptr<matrix::solver::Solver> single;
  DevBuf xt, b64, b, x;
  hipStream_t stream() const { return block ? block->stream() : nullptr; }
};

struct Server::System {
  const DeviceMatrix* A = nullptr;
  std::string json;
  std::shared_ptr<matrix::solver::Configuration> cfg;
  bool blockPath = false;                      // ILU-preconditioned, block solves
  DType xType = DType::F64, bType = DType::F32;  // solution / right-hand side
  std::shared_ptr<matrix::solver::Solver> ilu;  // factorized once, shared
  std::vector<std::unique_ptr<Slot>> idle;
  int32_t busy = 0;                           // slots in use
};

Server::Server(int32_t workers, size_t hbmBudgetBytes)
    : workers_(std::max(1, workers)), budget_(hbmBudgetBytes) {}

Server::~Server() = default;

int32_t Server::addSystem(const DeviceMatrix& A, const std::string& configJson) {
  namespace sv = matrix::solver;
  auto cfg = sv::Configuration::fromJSON(configJson);
  auto sys = std::make_unique<System>();
  sys->A = &A;
  sys->json = configJson;
  sys->cfg = cfg;
  // solution and right-hand side types: FP64 solution and FP32 right-hand
  // side for refinement (the extended-precision x of
  // iterative-refinement/Solver.cpp:84), the working type otherwise
  std::shared_ptr<sv::pbicgstab::Configuration> pb;
  if (auto p = std::dynamic_pointer_cast<sv::pbicgstab::Configuration>(cfg)) {
    pb = p;
    sys->xType = p->workingType.value_or(DType::F32);
    sys->bType = sys->xType;
  } else if (auto ir = std::dynamic_pointer_cast<sv::iterativerefinement::Configuration>(cfg)) {
    pb = std::dynamic_pointer_cast<sv::pbicgstab::Configuration>(ir->innerSolver);
    sys->xType = DType::F64;
    sys->bType = DType::F32;
    sys->blockPath = ir->extendedPrecisionType.value_or(DType::F64) == DType::F64 &&
                     ir->workingType.value_or(DType::F32) == DType::F32 &&
                     ir->relTolerance <= 0 && ir->norm == graphene::VectorNorm::L2;
  } else {
    gph::fail(gph::kUnsupported, "server systems solve with PBiCGStab or IterativeRefinement");
  }
  if (pb) {
    auto ilu = std::dynamic_pointer_cast<sv::ilu::Configuration>(pb->preconditioner);
    const bool pbOk = pb->workingType.value_or(DType::F32) == DType::F32 &&
                      pb->relTolerance <= 0 &&
                      pb->matrixValueType.value_or(DType::F32) == DType::F32 &&
                      pb->norm == graphene::VectorNorm::L2;
    const bool direct = std::dynamic_pointer_cast<sv::pbicgstab::Configuration>(cfg) != nullptr;
    sys->blockPath = ilu && pbOk && A.sellArgs(DType::F32).numSlices > 0 &&
                     (direct ? sys->xType == DType::F32 : sys->blockPath);
    if (sys->blockPath) sys->ilu = sv::Solver::createSolver(A, ilu);
  } else {
    sys->blockPath = false;
  }
  systems_.push_back(std::move(sys));
  return (int32_t)systems_.size() - 1;
}

size_t Server::footprint(int32_t system) const { return footprint(system, 1); }

size_t Server::footprint(int32_t system, int32_t k) const {
  if (system < 0 || system >= (int32_t)systems_.size())
    gph::fail(gph::kInvalid, "unknown system id");
  const System& sy = *systems_[(size_t)system];
  const size_t n = (size_t)sy.A->n();
  const size_t xs = gph::dtypeSize(sy.xType), bs = gph::dtypeSize(sy.bType);
  // request buffers: x_true, A x_true (f64), b, x
  const size_t req = (size_t)k * n * (8 + 8 + bs + xs);
  if (sy.blockPath) {
    // ten f32 Krylov vectors + the f64 outer residual, chunk tables, states
    const size_t chunks = ((size_t)k * n + gph::kBatchChunk - 1) / gph::kBatchChunk;
    return req + (size_t)k * n * (40 + 8) + chunks * 32 + (size_t)k * 512 + 65536;
  }
  // the single-system solvers' workspace: about ten vectors of the working type
  return req + n * 10 * xs + 65536;
}

size_t Server::residentBytes() const { return resident_; }

namespace {

struct Req {
  int64_t id;
};

}  // namespace

std::vector<RequestResult> Server::run(const std::vector<int32_t>& system,
                                       const std::vector<int32_t>& nRhs,
                                       const std::vector<double>& arrival,
                                       const std::vector<uint64_t>& seed,
                                       const std::vector<double*>& x0Out) {
  const size_t nr = system.size();
  if (nRhs.size() != nr || arrival.size() != nr || seed.size() != nr ||
      (!x0Out.empty() && x0Out.size() != nr))
    gph::fail(gph::kInvalid, "request arrays differ in length");
  for (size_t i = 0; i < nr; i++) {
    if (system[i] < 0 || system[i] >= (int32_t)systems_.size())
      gph::fail(gph::kInvalid, "request names an unknown system");
    if (nRhs[i] < 1 || nRhs[i] > kMaxRhs)
      gph::fail(gph::kInvalid, "a request needs 1 to 8 right-hand sides");
    if (footprint(system[i], nRhs[i]) > budget_)
      gph::fail(gph::kInvalid, "a request's footprint exceeds the HBM budget");
  }
  std::vector<RequestResult> res(nr);
  std::vector<Clock::time_point> admitted(nr);
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req> queue;  // arrived, not admitted (arrival order)
  size_t done = 0;
  std::exception_ptr err;
  int device = 0;
  GPH_HIP(hipGetDevice(&device));

  std::vector<int64_t> order(nr);
  for (size_t i = 0; i < nr; i++) order[i] = (int64_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return arrival[(size_t)a] < arrival[(size_t)b]; });
  const Clock::time_point t0 = Clock::now();
  auto at = [&](double sec) {
    return t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(sec));
  };

  // Admission (under mu): the first arrived request that can start now. It
  // takes an idle slot of its system with its k, or a new slot if the
  // resident slots plus the new one fit the budget once idle slots (least
  // recently used first) are evicted. Returns the request id, with the
  // slot (or a null slot to create, bytes reserved) and the evicted slots.
  struct Admit {
    int64_t id = -1;
    std::unique_ptr<Slot> slot;
    size_t reserve = 0;
    std::vector<std::unique_ptr<Slot>> evicted;
  };
  auto admit = [&]() {
    Admit a;
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      const size_t i = (size_t)it->id;
      System& sy = *systems_[(size_t)system[i]];
      const int32_t k = nRhs[i];
      if (!sy.blockPath && sy.busy > 0) continue;  // sequential path: one at a time
      // an idle slot with this k
      auto si = std::find_if(sy.idle.begin(), sy.idle.end(),
                             [&](const std::unique_ptr<Slot>& s) { return s->k == k; });
      if (si != sy.idle.end()) {
        a.slot = std::move(*si);
        sy.idle.erase(si);
      } else {
        const size_t need = footprint(system[i], k);
        size_t idleBytes = 0;
        for (auto& s : systems_)
          for (auto& q : s->idle) idleBytes += q->bytes;
        if (resident_ + need > budget_ + idleBytes) continue;  // not even after eviction
        while (resident_ + need > budget_) {  // evict the least recently used idle slot
          System* oldest = nullptr;
          size_t oi = 0;
          for (auto& s : systems_)
            for (size_t q = 0; q < s->idle.size(); q++)
              if (!oldest || s->idle[q]->lastUse < oldest->idle[oi]->lastUse) {
                oldest = s.get();
                oi = q;
              }
          resident_ -= oldest->idle[oi]->bytes;
          a.evicted.push_back(std::move(oldest->idle[oi]));
          oldest->idle.erase(oldest->idle.begin() + (ptrdiff_t)oi);
        }
        resident_ += need;
        a.reserve = need;
      }
      sy.busy++;
      a.id = it->id;
      queue.erase(it);
      return a;
    }
    return a;
  };

  auto makeSlot = [&](System& sy, int32_t k) {
    auto s = std::make_unique<Slot>();
    s->k = k;
    const size_t n = (size_t)sy.A->n();
    if (sy.blockPath) {
      s->block = std::make_unique<batch::BatchSolver>(*sy.A, k, sy.json, sy.ilu);
    } else {
      s->single = matrix::solver::Solver::createSolver(*sy.A, sy.cfg);
    }
    s->xt.alloc((size_t)k * n * 8);
    s->b64.alloc((size_t)k * n * 8);
    s->b.alloc((size_t)k * n * gph::dtypeSize(sy.bType));
    s->x.alloc((size_t)k * n * gph::dtypeSize(sy.xType));
    return s;
  };

  auto serve = [&](System& sy, Slot& sl, size_t i, RequestResult& r) {
    const DeviceMatrix& A = *sy.A;
    const int64_t n = A.n();
    const int32_t k = nRhs[i];
    hipStream_t s = sl.block ? sl.stream() : A.stream();
    // b_j = fl(A x_true_j) for every right-hand side
    for (int32_t j = 0; j < k; j++) {
      const uint64_t sd = seed[i] * 0x100000001B3ull + (uint64_t)j;
      double* xt = sl.xt.as<double>() + (size_t)j * n;
      double* b64 = sl.b64.as<double>() + (size_t)j * n;
      gph::launchUniform(xt, n, sd, s);
      A.spmv(Vec{xt, DType::F64}, Vec{b64, DType::F64}, DType::F64, DType::F32, s);
      gph::launchCast(b64, DType::F64,
                      static_cast<char*>(sl.b.get()) + (size_t)j * n * gph::dtypeSize(sy.bType),
                      sy.bType, n, s);
    }
    r.converged = 1;
    if (sl.block) {
      auto st = sl.block->solve(Vec{sl.x.get(), sy.xType}, Vec{sl.b.get(), DType::F32});
      for (const auto& q : st) {
        r.iterations += q.iterations;
        r.innerIterations += q.innerIterations;
        r.converged &= q.converged ? 1 : 0;
        // x starts at 0: the initial residual is ||b||
        const double rel = q.initialResidual > 0 ? q.finalResidual / q.initialResidual
                                                 : q.finalResidual;
        r.relResidual = std::max(r.relResidual, rel);
      }
    } else {
      for (int32_t j = 0; j < k; j++) {
        void* xj = static_cast<char*>(sl.x.get()) + (size_t)j * n * gph::dtypeSize(sy.xType);
        void* bj = static_cast<char*>(sl.b.get()) + (size_t)j * n * gph::dtypeSize(sy.bType);
        gph::launchFill(xj, sy.xType, n, 0.0, s);
        const auto q = sl.single->solve(Vec{xj, sy.xType}, Vec{bj, sy.bType});
        r.iterations += q.iterations;
        r.innerIterations += q.innerIterations;
        r.converged &= q.converged ? 1 : 0;
        const double rel = q.initialResidual > 0 ? q.finalResidual / q.initialResidual
                                                 : q.finalResidual;
        r.relResidual = std::max(r.relResidual, rel);
      }
    }
    if (!x0Out.empty() && x0Out[i]) {
      if (sy.xType == DType::F64) {
        GPH_HIP(hipMemcpyAsync(x0Out[i], sl.x.get(), (size_t)n * 8, hipMemcpyDeviceToHost, s));
      } else {
        gph::launchCast(sl.x.get(), sy.xType, sl.xt.get(), DType::F64, n, s);
        GPH_HIP(hipMemcpyAsync(x0Out[i], sl.xt.get(), (size_t)n * 8, hipMemcpyDeviceToHost, s));
      }
    }
    GPH_HIP(hipStreamSynchronize(s));
  };

  auto worker = [&]() {
    try {
      GPH_HIP(hipSetDevice(device));
    } catch (...) {
      std::lock_guard<std::mutex> g(mu);
      if (!err) err = std::current_exception();
      cv.notify_all();
      return;
    }
    for (;;) {
      Admit a;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (err || done == nr) return;
          a = admit();
          if (a.id >= 0) break;
          cv.wait(lk);
        }
      }
      a.evicted.clear();  // frees the evicted slots' memory (outside the lock)
      const size_t i = (size_t)a.id;
      System& sy = *systems_[(size_t)system[i]];
      admitted[i] = Clock::now();
      RequestResult r;
      try {
        if (!a.slot) {
          a.slot = makeSlot(sy, nRhs[i]);
          const size_t actual = a.slot->bytes =
              (a.slot->block ? a.slot->block->deviceBytes() : 0) + a.slot->xt.bytes() +
              a.slot->b64.bytes() + a.slot->b.bytes() + a.slot->x.bytes();
          std::lock_guard<std::mutex> g(mu);
          resident_ = resident_ - a.reserve + actual;
        }
        serve(sy, *a.slot, i, r);
      } catch (...) {
        std::lock_guard<std::mutex> g(mu);
        if (!err) err = std::current_exception();
      }
      const Clock::time_point fin = Clock::now();
      {
        std::lock_guard<std::mutex> g(mu);
        const Clock::time_point arr = at(arrival[i]);
        r.latency = std::chrono::duration<double>(fin - arr).count();
        r.wait = std::chrono::duration<double>(admitted[i] - arr).count();
        r.service = std::chrono::duration<double>(fin - admitted[i]).count();
        res[i] = r;
        sy.busy--;
        if (a.slot) {
          if (a.slot->bytes == 0) {  // creation failed: release the reservation
            resident_ -= a.reserve;
          } else {
            a.slot->lastUse = ++clock_;
            sy.idle.push_back(std::move(a.slot));
          }
        } else {
          resident_ -= a.reserve;
        }
        done++;
      }
      cv.notify_all();
    }
  };

  std::vector<std::thread> pool;
  for (int32_t w = 0; w < workers_; w++) pool.emplace_back(worker);
  // release the requests at their arrival times
  for (size_t q = 0; q < nr; q++) {
    const int64_t id = order[q];
    std::this_thread::sleep_until(at(arrival[(size_t)id]));
    std::lock_guard<std::mutex> g(mu);
    if (err) break;
    queue.push_back(Req{id});
    cv.notify_all();
  }
  {
    std::unique_lock<std::mutex> lk(mu);
    if (err) done = nr;  // stop the workers
    cv.notify_all();
  }
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
  return res;
}

}  // namespace graphene::serve
