// Solver tree on one GPU: PBiCGStab, mixed-precision iterative refinement,
// ILU(0), Gauss-Seidel, restarter.
//
// Interface of lg/matrix/solver/Solver.hpp:41-44 (solve(x, b) -> SolverStats,
// name(), usesInitialGuess()) and factory Solver::createSolver
// (lg/matrix/solver/Solver.cpp:32-56). The reference builds a device-side
// program with cf::While loops; here the host enqueues fused kernels on a
// HIP stream and the loop state (rho, alpha, omega, iteration, convergence)
// stays in HBM, checked by the host every few iterations.
#pragma once

#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "config.hpp"
#include "device_matrix.hpp"

namespace graphene::matrix::solver {

// lg/matrix/solver/SolverStats.hpp:29-69
struct SolverStats {
  std::string solverName;
  VectorNorm norm = VectorNorm::L2;
  int32_t numTiles = 1;
  double initialResidual = std::numeric_limits<double>::infinity();
  double finalResidual = std::numeric_limits<double>::infinity();
  int64_t iterations = 0;
  bool converged = false;
  bool singular = false;
  int64_t innerIterations = 0;
  double seconds = 0;
  void print() const;  // SolverStats.cpp:65-117
};

class Solver {
 public:
  explicit Solver(const DeviceMatrix& A) : A_(A) {}
  virtual ~Solver() = default;
  virtual SolverStats solve(const Vec& x, const Vec& b) = 0;
  virtual std::string name() const = 0;
  virtual bool usesInitialGuess() const = 0;

  // Preconditioner use inside a Krylov iteration: out = M^-1 rhs, where the
  // right-hand side is formed on the fly by `pre` (p- or s-update of
  // PBiCGStab). Kernels skip their work once st->done is set.
  virtual bool supportsDeviceApply() const { return false; }
  // Inner solves of a refinement loop: the solve may return while kernels
  // that exit at once (enqueued past convergence) are still queued on the
  // matrix's stream; the caller orders its own work on that stream and
  // synchronizes before it hands results out.
  virtual void setDeferredSync(bool v) { (void)v; }
  // Batched solves: the matrix is block-diagonal with system k on rows
  // systemRowStart[k] .. systemRowStart[k+1]-1; work of systems with no
  // inner solve running (phases[k] != 0) may be skipped by applyDevice
  // (optional).
  virtual void setSystemSkip(const int32_t* phases,
                             const std::vector<int64_t>& systemRowStart) {
    (void)phases;
    (void)systemRowStart;
  }
  virtual void applyDevice(const Vec& rhs, const Vec& out, const gph::PreOp& pre,
                           gph::DevState* st, hipStream_t s);
  // Several right-hand sides of the same matrix: out_j = M^-1 rhs_j for
  // j < k, vectors at rhs/out + j * stride elements; vector j may be skipped
  // when states[j].done, and every vector when st->done. The solver's own
  // data is only read, so concurrent calls on different streams are safe.
  virtual void applyDeviceMulti(const Vec& rhs, const Vec& out, int k, int64_t stride,
                                const gph::DevState* states, gph::DevState* st,
                                hipStream_t s);

  static std::unique_ptr<Solver> createSolver(
      const DeviceMatrix& A, std::shared_ptr<Configuration> config);

  const DeviceMatrix& matrix() const { return A_; }

 protected:
  const DeviceMatrix& A_;
};

// Workspace of device vectors keyed by slot.
class Workspace {
 public:
  Vec get(int slot, gph::DType t, int64_t n);
  size_t bytes() const;

 private:
  struct Entry {
    gph::DevBuf buf;
    gph::DType type = gph::DType::F32;
    int64_t n = 0;
  };
  std::vector<Entry> e_;
};

}  // namespace graphene::matrix::solver
