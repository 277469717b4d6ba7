// `benchmark`: the drop-in counterpart of applications/benchmark/benchmark.cpp.
//
// Same command line (benchmark.cpp:85-113):
//   benchmark <config.json> -t|--tiles N [-e|--emulate B]
//             (-m|--matrix file.mtx | -p|--poisson nx,ny,nz) [-d|--profile dir]
// and the same config keys (benchmark.cpp:163-183): "x" / "b" (number, or a
// MatrixMarket file in global row order), "x0" (number, default 0),
// "benchmark" ("solve" or "spmv"), "solver" (the solver configuration tree).
// Differences: the matrix is split into N tiles by the library's graph
// partitioner (recursive BFS bisection; METIS is not available), or by
// --partition contiguous|FILE (FILE: one tile id per row, global order);
// --emulate is accepted and ignored (the run is on the GPU); --profile writes
// profile.json (timings and solver statistics) instead of Poplar profiling
// data.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "device_matrix.hpp"
#include "host_matrix.hpp"
#include "solvers.hpp"

namespace {

using graphene::Json;
using graphene::matrix::DeviceMatrix;
using graphene::matrix::MatrixFormat;
using graphene::matrix::Vec;
namespace host = graphene::matrix::host;
namespace solver = graphene::matrix::solver;

struct Cli {
  std::string config, matrix, poisson, profile, partition = "graph";
  long tiles = -1;
  bool emulate = false;
};

[[noreturn]] void usage(const std::string& msg) {
  std::fprintf(stderr,
               "%s\nUsage: benchmark <config> -t,--tiles N [-e,--emulate B] "
               "(-m,--matrix FILE | -p,--poisson nx,ny,nz) [-d,--profile DIR] "
               "[--partition graph|contiguous|FILE]\n",
               msg.c_str());
  std::exit(2);
}

Cli parseArgs(int argc, char** argv) {
  Cli c;
  auto value = [&](int& i, const std::string& a, const char* longName) -> std::string {
    const std::string eq = std::string(longName) + "=";
    if (a.rfind(eq, 0) == 0) return a.substr(eq.size());
    if (i + 1 >= argc) usage("missing value for " + a);
    return argv[++i];
  };
  auto is = [](const std::string& a, const char* s, const char* l) {
    return a == s || a == l || a.rfind(std::string(l) + "=", 0) == 0;
  };
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (is(a, "-t", "--tiles")) {
      c.tiles = std::strtol(value(i, a, "--tiles").c_str(), nullptr, 10);
    } else if (is(a, "-e", "--emulate")) {
      const std::string v = value(i, a, "--emulate");
      c.emulate = v == "1" || v == "true" || v == "on";
    } else if (is(a, "-m", "--matrix")) {
      c.matrix = value(i, a, "--matrix");
    } else if (is(a, "-p", "--poisson")) {
      c.poisson = value(i, a, "--poisson");
    } else if (is(a, "-d", "--profile")) {
      c.profile = value(i, a, "--profile");
    } else if (is(a, "--partition", "--partition")) {
      c.partition = value(i, a, "--partition");
    } else if (a == "-h" || a == "--help") {
      usage("Graphene Benchmark Tool");
    } else if (!a.empty() && a[0] == '-') {
      usage("unknown option " + a);
    } else if (c.config.empty()) {
      c.config = a;
    } else {
      usage("unexpected argument " + a);
    }
  }
  if (c.config.empty()) usage("config is required");
  if (!std::filesystem::is_regular_file(c.config)) usage("config: file does not exist: " + c.config);
  if (c.tiles < 1) usage("--tiles is required (a positive number)");
  if (c.matrix.empty() == c.poisson.empty())
    usage("exactly one of --matrix and --poisson is required");
  if (!c.matrix.empty() && !std::filesystem::is_regular_file(c.matrix))
    usage("--matrix: file does not exist: " + c.matrix);
  if (!c.profile.empty() && !std::filesystem::is_directory(c.profile))
    usage("--profile: directory does not exist: " + c.profile);
  if (c.partition != "graph" && c.partition != "contiguous" &&
      !std::filesystem::is_regular_file(c.partition))
    usage("--partition: graph, contiguous or an existing file: " + c.partition);
  return c;
}

// "nx,ny,nz" (benchmark.cpp parsePoissonConfig)
void parsePoisson(const std::string& s, int64_t& nx, int64_t& ny, int64_t& nz) {
  if (std::sscanf(s.c_str(), "%ld,%ld,%ld", &nx, &ny, &nz) != 3 || nx < 1 || ny < 1 || nz < 1)
    throw std::invalid_argument("Invalid poisson config: " + s + " (expected nx,ny,nz)");
}

// MatrixMarket vector in global order: a vector object (array or coordinate)
// or the first column of a matrix object
// (DistributedTileLayout.cpp:185-206, MatrixMarket.cpp:67-107).
std::vector<float> loadVector(const std::string& path, int64_t n) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("Failed to open file " + path);
  std::string line;
  std::getline(f, line);
  std::istringstream hs(line);
  std::string banner, object, format, field;
  hs >> banner >> object >> format >> field;
  for (auto* s : {&object, &format, &field})
    for (auto& ch : *s) ch = (char)std::tolower((unsigned char)ch);
  if (banner != "%%MatrixMarket") throw std::runtime_error(path + ": not a MatrixMarket file");
  if (object == "matrix" && format == "coordinate") {
    f.close();
    host::TripletMatrix m = host::loadTripletMatrixFromFile(path);
    if ((int64_t)m.nrows != n) throw std::runtime_error(path + ": vector length differs from the matrix");
    std::vector<float> v((size_t)n, 0.0f);
    for (size_t k = 0; k < m.size(); k++)
      if (m.cols[k] == 0) v[(size_t)m.rows[k]] = m.vals[k];
    return v;
  }
  while (std::getline(f, line))
    if (!line.empty() && line[0] != '%') break;
  std::istringstream ss(line);
  std::vector<float> v((size_t)n, 0.0f);
  if (format == "array") {
    int64_t rows = 0, cols = 1;
    ss >> rows;
    if (object == "matrix") ss >> cols;
    if (rows != n) throw std::runtime_error(path + ": vector length differs from the matrix");
    for (int64_t i = 0; i < n; i++)  // column-major: the first column comes first
      if (!(f >> v[(size_t)i])) throw std::runtime_error(path + ": truncated");
  } else if (object == "vector" && format == "coordinate") {
    int64_t rows = 0, nnz = 0;
    ss >> rows >> nnz;
    if (rows != n) throw std::runtime_error(path + ": vector length differs from the matrix");
    for (int64_t k = 0; k < nnz; k++) {
      int64_t i = 0;
      double val = 1.0;
      if (!(f >> i)) throw std::runtime_error(path + ": truncated");
      if (field != "pattern") f >> val;
      if (i < 1 || i > n) throw std::runtime_error(path + ": index out of range");
      v[(size_t)(i - 1)] = (float)val;
    }
  } else {
    throw std::runtime_error(path + ": unsupported MatrixMarket layout " + object + " " + format);
  }
  return v;
}

// Number (constant vector) or file name; global order -> tile-major.
std::vector<float> vectorField(const Json& j, const std::string& name, const host::HostMatrix& h) {
  const int64_t n = h.nrows;
  if (j.kind == Json::Number) return std::vector<float>((size_t)n, (float)j.num);
  if (j.kind == Json::String) {
    const std::vector<float> g = loadVector(j.str, n);
    std::vector<float> l((size_t)n);
    for (int64_t p = 0; p < n; p++) l[(size_t)p] = g[(size_t)h.newToOld[(size_t)p]];
    return l;
  }
  throw std::runtime_error("Invalid config value for vector " + name);
}

double seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct DevVec {
  void* p = nullptr;
  explicit DevVec(size_t bytes) { GPH_HIP(hipMalloc(&p, bytes)); }
  ~DevVec() { (void)hipFree(p); }
};

void summary(const char* name, const std::vector<float>& v) {
  double s2 = 0;
  for (float a : v) s2 += (double)a * a;
  std::printf("%s: n=%zu, ||.||_2 = %.9g, [", name, v.size(), std::sqrt(s2));
  for (size_t i = 0; i < v.size() && i < 4; i++) std::printf("%s%.9g", i ? ", " : "", v[i]);
  std::printf("%s]\n", v.size() > 4 ? ", ..." : "");
}

int run(const Cli& cli) {
  std::ifstream cf(cli.config);
  std::stringstream text;
  text << cf.rdbuf();
  const Json config = Json::parse(text.str());
  if (cli.emulate) std::printf("--emulate: ignored, running on the GPU\n");

  const double t0 = seconds();
  host::TripletMatrix coo;
  if (!cli.matrix.empty()) {
    coo = host::loadTripletMatrixFromFile(cli.matrix);
  } else {
    int64_t nx, ny, nz;
    parsePoisson(cli.poisson, nx, ny, nz);
    coo = host::generate3DPoissonTripletMatrix(nx, ny, nz);
  }
  // partition (CRSHostMatrix.cpp:139-271 calls METIS here)
  std::vector<int32_t> r2t;
  const int32_t T = (int32_t)cli.tiles;
  if (T > 1 && cli.partition == "graph") {
    r2t = host::partitionGraph(coo.nrows, coo.rows, coo.cols, T);
  } else if (T > 1 && cli.partition != "contiguous") {
    std::ifstream pf(cli.partition);
    r2t.resize((size_t)coo.nrows);
    for (auto& t : r2t)
      if (!(pf >> t) || t < 0 || t >= T)
        throw std::runtime_error("--partition file: needs one tile id in [0, tiles) per row");
  }
  std::printf("Using %d tiles (%s partition)\n", T,
              T == 1 ? "single" : cli.partition == "graph" || cli.partition == "contiguous"
                                      ? cli.partition.c_str() : "file");
  std::shared_ptr<host::HostMatrix> h =
      host::buildHostMatrix(std::move(coo), T, r2t.empty() ? nullptr : r2t.data());
  DeviceMatrix A(h, MatrixFormat::CRS);
  const int64_t n = A.n();
  const double tSetup = seconds() - t0;
  std::printf("Matrix: %ld rows, %ld off-diagonal entries, %d tiles (%.2f s)\n", (long)n,
              (long)A.nnzOff(), A.numTiles(), tSetup);

  hipStream_t s = A.stream();
  DevVec xd((size_t)n * 4), bd((size_t)n * 4);
  auto upload = [&](const std::vector<float>& v, void* d) {
    GPH_HIP(hipMemcpyAsync(d, v.data(), v.size() * 4, hipMemcpyHostToDevice, s));
    GPH_HIP(hipStreamSynchronize(s));
  };
  auto download = [&](void* d) {
    std::vector<float> v((size_t)n);
    GPH_HIP(hipMemcpyAsync(v.data(), d, v.size() * 4, hipMemcpyDeviceToHost, s));
    GPH_HIP(hipStreamSynchronize(s));
    return v;
  };
  const Vec x{xd.p, gph::DType::F32}, b{bd.p, gph::DType::F32};
  // b from the config, or b = A x (benchmark.cpp:163-171)
  if (const Json* jb = config.get("b")) {
    upload(vectorField(*jb, "b", *h), bd.p);
  } else if (const Json* jx = config.get("x")) {
    upload(vectorField(*jx, "x", *h), xd.p);
    A.spmv(x, b, gph::DType::F32, gph::DType::F32, s);
  } else {
    throw std::runtime_error("Either b or x must be provided in the config");
  }
  // x = x0 (benchmark.cpp:173-174)
  float x0 = 0.0f;
  if (const Json* j = config.get("x0")) {
    if (j->kind != Json::Number) throw std::runtime_error("Invalid config value for x0");
    x0 = (float)j->num;
  }
  upload(std::vector<float>((size_t)n, x0), xd.p);

  const Json* jbench = config.get("benchmark");
  const std::string bench = jbench && jbench->kind == Json::String ? jbench->str : "";
  std::ostringstream prof;
  prof << "{\"rows\": " << n << ", \"nnz_off\": " << A.nnzOff() << ", \"tiles\": " << A.numTiles()
       << ", \"setup_s\": " << tSetup << ", \"benchmark\": \"" << bench << "\"";
  if (bench == "solve") {
    summary("b", download(bd.p));
    const Json* js = config.get("solver");
    if (!js) throw std::runtime_error("solve benchmark needs a \"solver\" configuration");
    auto cfg = solver::Configuration::fromPTree(*js);
    const double tc = seconds();
    auto S = solver::Solver::createSolver(A, cfg);
    const double tCreate = seconds() - tc;
    const solver::SolverStats st = S->solve(x, b);
    st.print();
    const std::vector<float> xf = download(xd.p);
    summary("x final", xf);
    prof << ", \"solver\": \"" << st.solverName << "\", \"solver_create_s\": " << tCreate
         << ", \"solve_s\": " << st.seconds << ", \"iterations\": " << st.iterations
         << ", \"inner_iterations\": " << st.innerIterations
         << ", \"initial_residual\": " << st.initialResidual
         << ", \"final_residual\": " << st.finalResidual
         << ", \"converged\": " << (st.converged ? "true" : "false")
         << ", \"singular\": " << (st.singular ? "true" : "false");
  } else if (bench == "spmv") {
    DevVec yd((size_t)n * 4);
    const Vec y{yd.p, gph::DType::F32};
    A.spmv(x, y, gph::DType::F32, gph::DType::F32, s);  // warm-up
    hipEvent_t e0, e1;
    GPH_HIP(hipEventCreate(&e0));
    GPH_HIP(hipEventCreate(&e1));
    const int reps = 20;
    GPH_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++) A.spmv(x, y, gph::DType::F32, gph::DType::F32, s);
    GPH_HIP(hipEventRecord(e1, s));
    GPH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GPH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double us = 1e3 * ms / reps;
    const double bytes = 8.0 * (double)A.nnzOff() + 16.0 * (double)n + 4.0;
    std::printf("SpMV: %.2f us per call, %.1f GB/s (8 nnz_off + 16 N + 4 bytes)\n", us,
                bytes / us * 1e-3);
    summary("y", download(yd.p));
    prof << ", \"spmv_us\": " << us << ", \"spmv_GBs\": " << bytes / us * 1e-3;
  } else {
    throw std::runtime_error("Unknown benchmark: " + bench);
  }
  prof << "}\n";
  if (!cli.profile.empty()) {
    const std::filesystem::path out = std::filesystem::path(cli.profile) / "profile.json";
    std::ofstream(out) << prof.str();
    std::printf("Profile written to %s\n", out.string().c_str());
  }
  std::printf("Done!\n");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const Cli cli = parseArgs(argc, argv);
  try {
    return run(cli);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
