// Solver configuration tree parsed from JSON with comments.
//
// Mirrors lg/matrix/solver/Configuration.cpp:103-136 (dispatch on "type")
// and the per-solver Configuration structs (pbicgstab/Configuration.hpp,
// iterative-refinement/Configuration.hpp, ilu/Configuration.hpp,
// gauss-seidel/Configuration.hpp, restarter/Configuration.hpp): same keys,
// same defaults, unknown keys ignored, errors for unknown solver types,
// unknown type names and invalid norms. Keys marked "extension" have no
// reference counterpart.
#pragma once

#include <climits>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "gph_common.hpp"

namespace graphene {

// Minimal JSON value (objects keep insertion order irrelevant; keys unique).
struct Json {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  const Json* get(const std::string& k) const {
    if (kind != Object) return nullptr;
    auto it = obj.find(k);
    return it == obj.end() ? nullptr : &it->second;
  }
  // JSON with // and /* */ comments (nlohmann::json::parse(..., ignore_comments=true),
  // applications/benchmark/benchmark.cpp:117-118).
  static Json parse(const std::string& text);
};

enum class VectorNorm { L1, L2, LINF, None };   // lg/matrix/Norm.hpp
VectorNorm parseVectorNorm(const std::string& s);  // Norm.cpp:24-34
std::string normToString(VectorNorm n);
gph::DType parseType(const std::string& s);        // lg/common/Type.cpp:69-88

namespace matrix::solver {

enum class MultiColorMode { Off, On, Auto };

struct Configuration {
  virtual ~Configuration() = default;
  virtual std::string solverName() const = 0;
  static std::shared_ptr<Configuration> fromPTree(const Json& config);
  static std::shared_ptr<Configuration> fromJSON(const std::string& text);
};

namespace pbicgstab {
struct Configuration : solver::Configuration {
  float absTolerance = 0;
  float relTolerance = 0;
  float relResidual = 0;
  int maxIterations = INT_MAX;
  int minIterations = 0;
  bool printPerformanceAfterSolve = false;
  bool printPerformanceEachIteration = false;
  bool verbose = false;
  VectorNorm norm = VectorNorm::L2;
  std::optional<gph::DType> workingType;
  std::shared_ptr<solver::Configuration> preconditioner;
  // extensions
  std::optional<gph::DType> matrixValueType;  // storage type of A's values
  bool relToleranceFromB = false;  // ||b|| instead of ||x|| for relTolerance
  int iterationsPerSync = 0;       // iterations enqueued per host check
  explicit Configuration(const Json& c);
  std::string solverName() const override {
    return preconditioner ? "PBiCGStab(" + preconditioner->solverName() + ")"
                          : "BiCGStab";
  }
};
}  // namespace pbicgstab

namespace iterativerefinement {
struct Configuration : solver::Configuration {
  float absTolerance = 0;
  float relTolerance = 0;
  float relResidual = 0;
  int maxIterations = INT_MAX;
  int minIterations = 0;
  bool printPerformanceAfterSolve = false;
  bool printPerformanceEachIteration = false;
  VectorNorm norm = VectorNorm::L2;
  std::shared_ptr<solver::Configuration> innerSolver;
  std::optional<gph::DType> extendedPrecisionType;
  // extensions
  std::optional<gph::DType> workingType;       // inner precision (default x's)
  std::optional<gph::DType> matrixValueType;   // values for the residual
  explicit Configuration(const Json& c);
  std::string solverName() const override { return "IterativeRefinement"; }
};
}  // namespace iterativerefinement

namespace ilu {
struct Configuration : solver::Configuration {
  MultiColorMode solveMulticolor = MultiColorMode::Auto;
  MultiColorMode factorizeMulticolor = MultiColorMode::Auto;
  bool diagonalBased = false;
  // extension: storage type of the factors (default: matrix type float32)
  std::optional<gph::DType> factorType;
  // extension: rows per wave group of the substitution kernels
  int rowsPerWave = 1024;
  // extension: substitution kernel: "rows" (lanes per row, factor rows in
  // registers; the packed kernel where its formats do not fit), "packed"
  // (product slabs), "level" (lane per row over level records), "block"
  // (a workgroup per tile, any tile size; also where the one-wave kernels do
  // not fit), "tile" (a workgroup per tile, unpacked factors)
  std::string substitution = "rows";
  explicit Configuration(const Json& c);
  std::string solverName() const override {
    return diagonalBased ? "DILU" : "ILU(0)";
  }
};
}  // namespace ilu

namespace gaussseidel {
struct Configuration : solver::Configuration {
  float absTolerance = 0;
  float relTolerance = 0;
  float relResidual = 0;
  int maxIterations = INT_MAX;
  int minIterations = 0;
  int numSweeps = 1;
  int numFixedIterations = 0;
  MultiColorMode solveMulticolor = MultiColorMode::Auto;
  bool printPerformanceAfterSolve = false;
  bool printPerformanceEachIteration = false;
  VectorNorm norm = VectorNorm::L2;
  std::optional<gph::DType> workingType;
  // extension: forward sweep followed by a backward sweep
  bool symmetric = false;
  explicit Configuration(const Json& c);
  std::string solverName() const override {
    return symmetric ? "SymmetricGaussSeidel" : "GaussSeidel";
  }
};
}  // namespace gaussseidel

namespace restarter {
struct Configuration : solver::Configuration {
  int maxRestarts = 1;
  std::shared_ptr<solver::Configuration> innerSolver;
  explicit Configuration(const Json& c);
  std::string solverName() const override {
    return "Restarter of " + innerSolver->solverName();
  }
};
}  // namespace restarter

}  // namespace matrix::solver
}  // namespace graphene
