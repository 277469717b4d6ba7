// Request server: continuous batching of solve requests on one GPU.
//
// No reference counterpart (the reference solves one system per program).
// Systems (device matrix + solver configuration) are registered once and
// stay resident with their factors. A request names a system, a number of
// right-hand sides k (1-8) and a seed; each right-hand side is
// b = fl32(A x_true), x_true ~ U[-1,1) from the seed, solved from x = 0.
//
// Requests are admitted by `workers` host threads in arrival order as slots
// and memory free up. A slot is the Krylov workspace and request buffers of
// one request: for refinement or PBiCGStab with an ILU(0) preconditioner, a
// block solver of the request's k right-hand sides (each product reads the
// matrix once for all of them, each tile's substitution runs for all of them
// in one launch; batch.hpp), on a stream of its own, so several requests on
// the same system overlap on the device, sharing the system's factors. Slots
// stay resident after a request and are reused by the next request of the
// same system and k; the budget counts every resident slot, and a request
// that does not fit evicts idle slots (least recently used first) before it
// waits. Other configurations (e.g. a Gauss-Seidel preconditioner) run their
// right-hand sides one after the other, one request at a time per system. A
// result depends only on the request (deterministic kernels, per-slot
// state), not on what else runs.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device_matrix.hpp"
#include "solvers.hpp"

namespace graphene::serve {

struct RequestResult {
  double latency = 0;      // finish - arrival (s)
  double wait = 0;         // admission - arrival (s)
  double service = 0;      // finish - admission (s)
  int64_t iterations = 0;  // summed over right-hand sides (outer)
  int64_t innerIterations = 0;
  double relResidual = 0;  // largest final ||r|| / ||b|| over the right-hand sides
  int32_t converged = 0;   // all right-hand sides converged
};

class Server {
 public:
  Server(int32_t workers, size_t hbmBudgetBytes);
  ~Server();

  // Registers a resident system; the matrix must outlive the server.
  // Returns the system id.
  int32_t addSystem(const matrix::DeviceMatrix& A, const std::string& configJson);

  // Device bytes of a slot for one right-hand side of the system.
  size_t footprint(int32_t system) const;
  // ... and for k right-hand sides.
  size_t footprint(int32_t system, int32_t k) const;
  // Device bytes of the resident slots.
  size_t residentBytes() const;

  // Runs the requests; arrival[i] is seconds after the start. x0Out[i]
  // (optional, N doubles) receives the solution of request i's first
  // right-hand side in the matrix's (tile-major) order.
  std::vector<RequestResult> run(const std::vector<int32_t>& system,
                                 const std::vector<int32_t>& nRhs,
                                 const std::vector<double>& arrival,
                                 const std::vector<uint64_t>& seed,
                                 const std::vector<double*>& x0Out);

 private:
  struct System;
  struct Slot;
  std::vector<std::unique_ptr<System>> systems_;
  int32_t workers_;
  size_t budget_;
  size_t resident_ = 0;
  uint64_t clock_ = 0;  // use stamps of idle slots
};

}  // namespace graphene::serve
