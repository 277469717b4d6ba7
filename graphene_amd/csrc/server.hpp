// Request server: continuous batching of solve requests on one GPU.
//
// No reference counterpart (the reference solves one system per program).
// Systems (device matrix + solver configuration) are registered once and stay
// resident with their factors and Krylov workspace. Requests (system,
// number of right-hand sides, seed) arrive on a schedule; worker threads
// admit a request when its system is idle and the HBM footprint of the
// requests in flight stays within the budget, so requests that arrive while
// others run start as soon as a slot and memory free up, and their kernels
// overlap on the device (every system has its own stream). Each right-hand
// side is b = fl32(A x_true), x_true ~ U[-1,1) from the request seed, solved
// from x = 0; a result does not depend on what else runs (deterministic
// kernels, per-system state).
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
  // Returns the system id. Its footprint is the solver's workspace plus the
  // request vectors (the matrix itself is resident either way).
  int32_t addSystem(const matrix::DeviceMatrix& A, const std::string& configJson);

  size_t footprint(int32_t system) const;

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
  std::vector<std::unique_ptr<System>> systems_;
  int32_t workers_;
  size_t budget_;
};

}  // namespace graphene::serve
