// Batched solves: many independent systems held as one block-diagonal
// matrix, solved together.
//
// No reference counterpart (the reference solves one system per program);
// each system follows the single-system algorithm: MPIR with an FP32
// PBiCGStab + ILU(0) inner solve and FP64 residual and update
// (iterative-refinement/Solver.cpp:48-123, pbicgstab/Solver.cpp:48-194), or
// FP32 PBiCGStab + ILU(0) alone, from a zero initial guess.
//
// The systems are consecutive tile ranges of one device matrix, with no
// entry coupling two systems (checked). SpMV and the block-Jacobi ILU(0)
// run once per step on the union; per system are the Krylov scalars, the
// convergence state, the (segmented, deterministic) dot products and the
// masked vector updates. Refinement runs its outer step on the device: each
// system moves on its own between its inner solve (phase 0), its outer step
// (phase 1: phi += d, the FP64 residual of its rows, norm, convergence and
// the inner restart) and finished (phase 2); one PBiCGStab iteration of the
// running systems followed by the outer step of the systems that need one
// is captured in a HIP graph of several such steps and replayed until every
// system has finished ("outerStep": "host": the inner solves of all systems
// end before the host enqueues their outer steps). Work of systems in
// another phase is skipped and a finished batch makes every kernel return
// at once. Setup picks the ILU substitution kernel (and the
// packed kernel's rows per wave) on the union by a few timed applications.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "device_matrix.hpp"
#include "solvers.hpp"

namespace graphene::batch {

struct SystemStats {
  int64_t iterations = 0;       // outer (MPIR) or PBiCGStab iterations
  int64_t innerIterations = 0;  // PBiCGStab iterations over all outer steps
  double initialResidual = 0, finalResidual = 0;
  bool converged = false, singular = false;
};

class BatchSolver {
 public:
  // systemTileStart: S+1 tile offsets; config: IterativeRefinement with a
  // PBiCGStab inner solver, or PBiCGStab, each with an ILU preconditioner.
  BatchSolver(const matrix::DeviceMatrix& A, const std::vector<int32_t>& systemTileStart,
              const std::string& configJson);
  // nrhs (1..8) right-hand sides of the one system A: the vectors are nrhs
  // blocks of A.n() rows (block j = right-hand side j), the matrix streams
  // are read once per product for all of them and every tile's ILU(0)
  // substitution runs for each of them in one launch. Runs on a stream of
  // its own; `ilu` (optional) is a factorized ILU preconditioner of A shared
  // with other solvers (its data is only read).
  BatchSolver(const matrix::DeviceMatrix& A, int32_t nrhs, const std::string& configJson,
              std::shared_ptr<matrix::solver::Solver> ilu);
  ~BatchSolver();

  // device memory held by this solver (vectors, states, scratch)
  size_t deviceBytes() const;
  hipStream_t stream() const { return multi_ ? stream_ : A_.stream(); }

  int32_t numSystems() const { return S_; }
  int64_t systemRowStart(int32_t s) const { return sysRow_[(size_t)s]; }
  bool refinement() const { return mpir_; }
  int32_t tunedRowsPerWave() const { return rowsPerWave_; }
  const std::string& tunedSubstitution() const { return substitution_; }

  // MPIR: x float64, b float32. PBiCGStab: x and b float32. Tile-major
  // order; x is overwritten (the solve starts from zero).
  std::vector<SystemStats> solve(const matrix::Vec& x, const matrix::Vec& b);
  double lastSeconds() const { return seconds_; }
  int64_t lastGraphLaunches() const { return launches_; }

 private:
  std::shared_ptr<matrix::solver::ilu::Configuration> parseConfig(const std::string& json);
  void setupChunks(hipStream_t st);
  void innerSolve(const void* rhs, void* x, std::vector<int64_t>& innerIts);
  void enqueueIteration(hipStream_t s);
  void captureGraph(hipStream_t s);
  void enqueueOuterStep(hipStream_t s, int first);
  void solveDeviceOuter(const matrix::Vec& x, const matrix::Vec& b);

  const matrix::DeviceMatrix& A_;
  int32_t S_ = 0, C_ = 0;
  std::vector<int64_t> sysRow_;
  bool mpir_ = false;
  int32_t rowsPerWave_ = 0;
  std::string substitution_;
  gph::DevState innerInit_{}, outerInit_{};
  std::shared_ptr<matrix::solver::Solver> ilu_;
  bool multi_ = false;  // nrhs right-hand sides of one system
  bool outerOnDevice_ = true;  // refinement: outer step on the device
  hipStream_t stream_ = nullptr;
  gph::DevBuf chunkSys_, chunkRow_, sysChunk_, inner_, outer_, glob_, partials_, sliceSys_;
  gph::DevBuf r_, r0_, p_, v_, y_, s_, z_, t_, d_, rw_, r64_;
  gph::DevBuf phase_;    // per system: 0 inner solve, 1 outer step due, 2 finished
  gph::DevBuf phi_, b_;  // refinement on the device: solution and right-hand side
  hipGraph_t outerGraph_ = nullptr;
  hipGraphExec_t outerExec_ = nullptr;
  gph::BatchArgs args_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  gph::DevState* pinned_ = nullptr;  // 2 host copies of the batch flag
  float* xInner_ = nullptr;          // current inner solution (captured)
  const float* rhsInner_ = nullptr;
  double seconds_ = 0;
  int64_t launches_ = 0;
};

}  // namespace graphene::batch
