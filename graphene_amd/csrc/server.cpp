// Request server (see server.hpp).
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <thread>

#include "config.hpp"

namespace graphene::serve {

using gph::DevBuf;
using gph::DType;
using matrix::DeviceMatrix;
using matrix::Vec;
using Clock = std::chrono::steady_clock;

struct Server::System {
  const DeviceMatrix* A = nullptr;
  std::unique_ptr<matrix::solver::Solver> solver;
  DType xType = DType::F64, bType = DType::F32;
  DevBuf xt, b64, b, x;  // x_true (f64), A x_true (f64), b, solution
  size_t footprint = 0;
  bool busy = false;
};

Server::Server(int32_t workers, size_t hbmBudgetBytes)
    : workers_(std::max(1, workers)), budget_(hbmBudgetBytes) {}

Server::~Server() = default;

int32_t Server::addSystem(const DeviceMatrix& A, const std::string& configJson) {
  auto cfg = matrix::solver::Configuration::fromJSON(configJson);
  auto sys = std::make_unique<System>();
  sys->A = &A;
  // solution and right-hand side types: FP64 solution and FP32 right-hand
  // side for iterative refinement (the extended-precision x of
  // iterative-refinement/Solver.cpp:84), the working type otherwise
  if (auto* pb = dynamic_cast<matrix::solver::pbicgstab::Configuration*>(cfg.get())) {
    sys->xType = pb->workingType.value_or(DType::F32);
    sys->bType = sys->xType;
  } else if (dynamic_cast<matrix::solver::iterativerefinement::Configuration*>(cfg.get())) {
    sys->xType = DType::F64;
    sys->bType = DType::F32;
  } else {
    gph::fail(gph::kUnsupported, "server systems solve with PBiCGStab or IterativeRefinement");
  }
  sys->solver = matrix::solver::Solver::createSolver(A, cfg);
  const size_t n = (size_t)A.n();
  sys->xt.alloc(n * 8);
  sys->b64.alloc(n * 8);
  sys->b.alloc(n * gph::dtypeSize(sys->bType));
  sys->x.alloc(n * 8);
  // Krylov workspace (about ten vectors) plus the request vectors
  sys->footprint = n * (10 * gph::dtypeSize(sys->xType) + 16 + gph::dtypeSize(sys->bType) + 8);
  systems_.push_back(std::move(sys));
  return (int32_t)systems_.size() - 1;
}

size_t Server::footprint(int32_t system) const {
  if (system < 0 || system >= (int32_t)systems_.size())
    gph::fail(gph::kInvalid, "unknown system id");
  return systems_[(size_t)system]->footprint;
}

namespace {

struct Req {
  int64_t id;
  Clock::time_point arrival;
};

// One right-hand side: b = fl(A x_true), x = 0, solve; returns stats and
// ||b||.
matrix::solver::SolverStats solveOne(const DeviceMatrix& A, matrix::solver::Solver& S,
                                     DType xType, DType bType, DevBuf& xt, DevBuf& b64,
                                     DevBuf& b, DevBuf& x, uint64_t seed, double& bnorm) {
  hipStream_t s = A.stream();
  const int64_t n = A.n();
  gph::launchUniform(xt.as<double>(), n, seed, s);
  A.spmv(Vec{xt.get(), DType::F64}, Vec{b64.get(), DType::F64}, DType::F64, DType::F32, s);
  gph::launchCast(b64.get(), DType::F64, b.get(), bType, n, s);
  bnorm = A.norm2(Vec{b64.get(), DType::F64}, s);
  gph::launchFill(x.get(), xType, n, 0.0, s);
  return S.solve(Vec{x.get(), xType}, Vec{b.get(), bType});
}

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
    if (nRhs[i] < 1) gph::fail(gph::kInvalid, "a request needs at least one right-hand side");
    if (footprint(system[i]) > budget_)
      gph::fail(gph::kInvalid, "a request's footprint exceeds the HBM budget");
  }
  std::vector<RequestResult> res(nr);
  std::vector<Clock::time_point> admitted(nr);
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Req> queue;  // arrived, not admitted (arrival order)
  size_t released = 0, done = 0, inflight = 0;
  std::exception_ptr err;
  int device = 0;
  GPH_HIP(hipGetDevice(&device));

  // arrival order
  std::vector<int64_t> order(nr);
  for (size_t i = 0; i < nr; i++) order[i] = (int64_t)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return arrival[(size_t)a] < arrival[(size_t)b]; });
  const Clock::time_point t0 = Clock::now();
  auto at = [&](double sec) {
    return t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(sec));
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
      int64_t id = -1;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (err || done == nr) return;
          // first arrived request whose system is idle and fits the budget
          for (auto it = queue.begin(); it != queue.end(); ++it) {
            System& sy = *systems_[(size_t)system[(size_t)it->id]];
            if (!sy.busy && inflight + sy.footprint <= budget_) {
              id = it->id;
              sy.busy = true;
              inflight += sy.footprint;
              queue.erase(it);
              break;
            }
          }
          if (id >= 0) break;
          cv.wait(lk);
        }
      }
      const size_t i = (size_t)id;
      System& sy = *systems_[(size_t)system[i]];
      admitted[i] = Clock::now();
      RequestResult r;
      r.converged = 1;
      try {
        for (int32_t k = 0; k < nRhs[i]; k++) {
          double bnorm = 0;
          const uint64_t sd = seed[i] * 0x100000001B3ull + (uint64_t)k;
          matrix::solver::SolverStats st = solveOne(*sy.A, *sy.solver, sy.xType, sy.bType,
                                                    sy.xt, sy.b64, sy.b, sy.x, sd, bnorm);
          r.iterations += st.iterations;
          r.innerIterations += st.innerIterations;
          r.converged &= st.converged ? 1 : 0;
          const double rel = bnorm > 0 ? st.finalResidual / bnorm : st.finalResidual;
          r.relResidual = std::max(r.relResidual, rel);
          if (k == 0 && !x0Out.empty() && x0Out[i]) {
            hipStream_t s = sy.A->stream();
            const int64_t n = sy.A->n();
            if (sy.xType == DType::F64) {
              GPH_HIP(hipMemcpyAsync(x0Out[i], sy.x.get(), (size_t)n * 8,
                                     hipMemcpyDeviceToHost, s));
            } else {
              gph::launchCast(sy.x.get(), sy.xType, sy.xt.get(), DType::F64, n, s);
              GPH_HIP(hipMemcpyAsync(x0Out[i], sy.xt.get(), (size_t)n * 8,
                                     hipMemcpyDeviceToHost, s));
            }
            GPH_HIP(hipStreamSynchronize(s));
          }
        }
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
        sy.busy = false;
        inflight -= sy.footprint;
        done++;
      }
      cv.notify_all();
    }
  };

  std::vector<std::thread> pool;
  for (int32_t w = 0; w < workers_; w++) pool.emplace_back(worker);
  // release the requests at their arrival times
  for (size_t k = 0; k < nr; k++) {
    const int64_t id = order[k];
    std::this_thread::sleep_until(at(arrival[(size_t)id]));
    std::lock_guard<std::mutex> g(mu);
    if (err) break;
    queue.push_back(Req{id, at(arrival[(size_t)id])});
    released++;
    cv.notify_all();
  }
  {
    std::unique_lock<std::mutex> lk(mu);
    if (err) done = nr;  // stop the workers
    cv.notify_all();
  }
  for (auto& t : pool) t.join();
  if (err) std::rethrow_exception(err);
  (void)released;
  return res;
}

}  // namespace graphene::serve
