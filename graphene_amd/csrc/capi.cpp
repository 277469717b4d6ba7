// extern "C" boundary of the library (declarations in include/graphene_amd.h).
#include <cstring>
#include <memory>
#include <string>

#include "../../include/graphene_amd.h"
#include "batch.hpp"
#include "config.hpp"
#include "device_matrix.hpp"
#include "host_matrix.hpp"
#include "server.hpp"
#include "solvers.hpp"

using namespace graphene;
using namespace graphene::matrix;
namespace hostns = graphene::matrix::host;

struct gph_host_matrix {
  std::shared_ptr<hostns::HostMatrix> m;
};
struct gph_matrix {
  std::unique_ptr<DeviceMatrix> A;
  bool async = false;  // gph_spmv / gph_residual return without synchronizing
};
struct gph_server {
  std::unique_ptr<graphene::serve::Server> S;
};
struct gph_batch {
  std::unique_ptr<graphene::batch::BatchSolver> B;
};
struct gph_solver {
  const gph_matrix* A;
  std::shared_ptr<solver::Configuration> cfg;
  std::unique_ptr<solver::Solver> S;
};

namespace {
thread_local std::string g_lastError;

template <typename F>
int guard(F&& f) {
  try {
    f();
    g_lastError.clear();
    return GPH_OK;
  } catch (const gph::Error& e) {
    g_lastError = e.what();
    return e.code;
  } catch (const std::bad_alloc& e) {
    g_lastError = std::string("out of host memory: ") + e.what();
    return GPH_ERR_RUNTIME;
  } catch (const std::exception& e) {
    g_lastError = e.what();
    return GPH_ERR_RUNTIME;
  } catch (...) {
    g_lastError = "unknown error";
    return GPH_ERR_RUNTIME;
  }
}

gph::DType dt(gph_type t) {
  switch (t) {
    case GPH_FLOAT16:
      return gph::DType::F16;
    case GPH_FLOAT32:
      return gph::DType::F32;
    case GPH_FLOAT64:
      return gph::DType::F64;
    case GPH_TWOFLOAT32:
      return gph::DType::TF32;
  }
  gph::fail(gph::kInvalid, "invalid element type " + std::to_string((int)t));
}

void need(const void* p, const char* what) {
  if (!p) gph::fail(gph::kInvalid, std::string(what) + " must not be NULL");
}

int emitTriplets(hostns::TripletMatrix&& m, int64_t* count, int64_t* rows,
                 int64_t* cols, float* vals) {
  need(count, "count");
  if (rows) {
    if (*count < (int64_t)m.size())
      gph::fail(gph::kInvalid, "output buffers too small");
    need(cols, "cols");
    need(vals, "vals");
    std::memcpy(rows, m.rows.data(), m.size() * sizeof(int64_t));
    std::memcpy(cols, m.cols.data(), m.size() * sizeof(int64_t));
    std::memcpy(vals, m.vals.data(), m.size() * sizeof(float));
  }
  *count = (int64_t)m.size();
  return 0;
}

gph_host_matrix* makeHost(hostns::TripletMatrix&& m, int32_t T,
                          const int32_t* r2t) {
  auto* h = new gph_host_matrix;
  try {
    h->m = hostns::buildHostMatrix(std::move(m), T, r2t);
  } catch (...) {
    delete h;
    throw;
  }
  return h;
}

template <typename T>
void permute(const hostns::HostMatrix& m, const void* in, void* out,
             bool toLocal) {
  const T* a = static_cast<const T*>(in);
  T* b = static_cast<T*>(out);
  for (int64_t p = 0; p < m.nrows; p++) {
    int64_t g = m.newToOld[(size_t)p];
    if (toLocal)
      b[p] = a[g];
    else
      b[g] = a[p];
  }
}
}  // namespace

extern "C" {

const char* gph_last_error(void) { return g_lastError.c_str(); }
int32_t gph_version(void) { return 100; }

int gph_generate_poisson3d(int64_t nx, int64_t ny, int64_t nz, int64_t* count,
                           int64_t* rows, int64_t* cols, float* vals) {
  return guard([&] {
    emitTriplets(hostns::generate3DPoissonTripletMatrix(nx, ny, nz), count,
                 rows, cols, vals);
  });
}
int gph_generate_convdiff3d27(int64_t nx, int64_t ny, int64_t nz,
                              double peclet, uint64_t seed, int64_t* count,
                              int64_t* rows, int64_t* cols, float* vals) {
  return guard([&] {
    emitTriplets(hostns::generate3DConvectionDiffusion27(nx, ny, nz, peclet, seed),
                 count, rows, cols, vals);
  });
}
int gph_generate_convdiff2d9(int64_t nx, int64_t ny, double diffusion,
                             double peclet, uint64_t seed, int64_t* count,
                             int64_t* rows, int64_t* cols, float* vals) {
  return guard([&] {
    emitTriplets(
        hostns::generate2DConvectionDiffusion9(nx, ny, diffusion, peclet, seed),
        count, rows, cols, vals);
  });
}
int gph_generate_powerlaw(int64_t n, int32_t min_len, int32_t max_len,
                          double exponent, int32_t band, double local_fraction,
                          uint64_t seed, int64_t* count, int64_t* rows,
                          int64_t* cols, float* vals) {
  return guard([&] {
    emitTriplets(hostns::generatePowerLaw(n, min_len, max_len, exponent, band,
                                          local_fraction, seed),
                 count, rows, cols, vals);
  });
}
int gph_uniform_pm1(uint64_t seed, int64_t n, float* out) {
  return guard([&] {
    need(out, "out");
    hostns::uniformPm1(seed, n, out);
  });
}

int gph_partition_grid3d(int64_t nx, int64_t ny, int64_t nz, int64_t bx,
                         int64_t by, int64_t bz, int32_t* row_to_tile,
                         int32_t* num_tiles) {
  return guard([&] {
    need(row_to_tile, "row_to_tile");
    auto v = hostns::partitionGrid3D(nx, ny, nz, bx, by, bz, num_tiles);
    std::memcpy(row_to_tile, v.data(), v.size() * sizeof(int32_t));
  });
}

int gph_partition_graph(int64_t nrows, int64_t nnz, const int64_t* rows,
                        const int64_t* cols, int32_t num_tiles,
                        int32_t* row_to_tile) {
  return guard([&] {
    need(row_to_tile, "row_to_tile");
    if (nnz > 0) {
      need(rows, "rows");
      need(cols, "cols");
    }
    std::vector<int64_t> r(rows, rows + nnz), c(cols, cols + nnz);
    auto v = hostns::partitionGraph(nrows, r, c, num_tiles);
    std::memcpy(row_to_tile, v.data(), v.size() * sizeof(int32_t));
  });
}

int gph_host_matrix_from_coo(int64_t nrows, int64_t nnz, const int64_t* rows,
                             const int64_t* cols, const float* vals,
                             int32_t num_tiles, const int32_t* row_to_tile,
                             gph_host_matrix** out) {
  return guard([&] {
    need(out, "out");
    if (nnz > 0) {
      need(rows, "rows");
      need(cols, "cols");
      need(vals, "vals");
    }
    hostns::TripletMatrix m;
    m.nrows = m.ncols = nrows;
    m.rows.assign(rows, rows + nnz);
    m.cols.assign(cols, cols + nnz);
    m.vals.assign(vals, vals + nnz);
    *out = makeHost(std::move(m), num_tiles, row_to_tile);
  });
}
int gph_host_matrix_poisson3d(int64_t nx, int64_t ny, int64_t nz,
                              int32_t num_tiles, const int32_t* row_to_tile,
                              gph_host_matrix** out) {
  return guard([&] {
    need(out, "out");
    *out = makeHost(hostns::generate3DPoissonTripletMatrix(nx, ny, nz),
                    num_tiles, row_to_tile);
  });
}
int gph_host_matrix_convdiff3d27(int64_t nx, int64_t ny, int64_t nz,
                                 double peclet, uint64_t seed,
                                 int32_t num_tiles, const int32_t* row_to_tile,
                                 gph_host_matrix** out) {
  return guard([&] {
    need(out, "out");
    *out = makeHost(
        hostns::generate3DConvectionDiffusion27(nx, ny, nz, peclet, seed),
        num_tiles, row_to_tile);
  });
}
int gph_host_matrix_powerlaw(int64_t n, int32_t min_len, int32_t max_len,
                             double exponent, int32_t band,
                             double local_fraction, uint64_t seed,
                             int32_t num_tiles, const int32_t* row_to_tile,
                             gph_host_matrix** out) {
  return guard([&] {
    need(out, "out");
    *out = makeHost(hostns::generatePowerLaw(n, min_len, max_len, exponent,
                                             band, local_fraction, seed),
                    num_tiles, row_to_tile);
  });
}
int gph_host_matrix_load_mtx(const char* path, int32_t num_tiles,
                             const int32_t* row_to_tile,
                             gph_host_matrix** out) {
  return guard([&] {
    need(path, "path");
    need(out, "out");
    *out = makeHost(hostns::loadTripletMatrixFromFile(path), num_tiles,
                    row_to_tile);
  });
}
void gph_host_matrix_destroy(gph_host_matrix* h) { delete h; }

int gph_host_matrix_get_info(const gph_host_matrix* h,
                             gph_host_matrix_info* info) {
  return guard([&] {
    need(h, "matrix");
    need(info, "info");
    const auto& m = *h->m;
    info->nrows = m.nrows;
    info->nnz_off = m.nnzOff();
    info->num_halo = m.numHalo();
    info->num_tiles = m.numTiles;
    int32_t mc = 0, mr = 0;
    for (int32_t t = 0; t < m.numTiles; t++) {
      mc = std::max(mc, m.numColors[(size_t)t]);
      mr = std::max<int32_t>(mr, (int32_t)m.tileRows(t));
    }
    info->max_colors = mc;
    info->max_tile_rows = mr;
    info->num_color_entries = (int64_t)m.colorStart.size();
    info->num_upper_color_entries = (int64_t)m.upperColorStart.size();
    info->multicolor_recommended = m.multicolorRecommended ? 1 : 0;
  });
}

int gph_host_matrix_export(const gph_host_matrix* h, int64_t* tile_start,
                           int64_t* new_to_old, int32_t* num_interior,
                           int32_t* num_separator_regions, int64_t* row_ptr,
                           int32_t* col, int32_t* col_local, float* off,
                           float* diag, int64_t* halo_start,
                           int64_t* halo_global) {
  return guard([&] {
    need(h, "matrix");
    const auto& m = *h->m;
    auto cp = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(tile_start, m.tileStart);
    cp(new_to_old, m.newToOld);
    cp(num_interior, m.numInterior);
    cp(num_separator_regions, m.numSeparatorRegions);
    cp(row_ptr, m.rowPtr);
    cp(col, m.col);
    cp(col_local, m.colLocal);
    cp(off, m.offDiag);
    cp(diag, m.diag);
    cp(halo_start, m.haloStart);
    cp(halo_global, m.haloGlobal);
  });
}

int gph_host_matrix_export_levels(const gph_host_matrix* h, int32_t* level,
                                  int32_t* color_sort_addr,
                                  int64_t* color_start, int32_t* num_colors) {
  return guard([&] {
    need(h, "matrix");
    const auto& m = *h->m;
    auto cp = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(level, m.level);
    cp(color_sort_addr, m.colorSortAddr);
    cp(color_start, m.colorStart);
    cp(num_colors, m.numColors);
  });
}

int gph_host_matrix_export_ldu(const gph_host_matrix* h, int64_t* num_faces,
                               int32_t* lower_addr, int32_t* upper_addr,
                               int32_t* owner_start, int32_t* neighbour_start,
                               int32_t* neighbour_col_ind, float* upper,
                               float* lower) {
  return guard([&] {
    need(h, "matrix");
    need(num_faces, "num_faces");
    hostns::LDUHost L = hostns::buildLDU(*h->m);
    *num_faces = L.numFaces;
    auto cp = [](auto* dst, const auto& v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(lower_addr, L.lowerAddr);
    cp(upper_addr, L.upperAddr);
    cp(owner_start, L.ownerStart);
    cp(neighbour_start, L.neighbourStart);
    cp(neighbour_col_ind, L.neighbourColInd);
    cp(upper, L.upper);
    cp(lower, L.lower);
  });
}

int gph_host_matrix_to_local(const gph_host_matrix* h, gph_type t,
                             const void* global_in, void* local_out) {
  return guard([&] {
    need(h, "matrix");
    need(global_in, "input");
    need(local_out, "output");
    switch (dt(t)) {
      case gph::DType::F32: permute<float>(*h->m, global_in, local_out, true); break;
      case gph::DType::F64: permute<double>(*h->m, global_in, local_out, true); break;
      case gph::DType::F16: permute<uint16_t>(*h->m, global_in, local_out, true); break;
    }
  });
}
int gph_host_matrix_to_global(const gph_host_matrix* h, gph_type t,
                              const void* local_in, void* global_out) {
  return guard([&] {
    need(h, "matrix");
    need(local_in, "input");
    need(global_out, "output");
    switch (dt(t)) {
      case gph::DType::F32: permute<float>(*h->m, local_in, global_out, false); break;
      case gph::DType::F64: permute<double>(*h->m, local_in, global_out, false); break;
      case gph::DType::F16: permute<uint16_t>(*h->m, local_in, global_out, false); break;
    }
  });
}

int gph_device_count(int32_t* count) {
  return guard([&] {
    need(count, "count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
  });
}
int gph_malloc(size_t bytes, void** ptr) {
  return guard([&] {
    need(ptr, "ptr");
    GPH_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  });
}
int gph_free(void* ptr) {
  return guard([&] {
    if (ptr) GPH_HIP(hipFree(ptr));
  });
}
int gph_memcpy_htod(void* dst, const void* src, size_t bytes) {
  return guard([&] { GPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int gph_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  return guard([&] { GPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int gph_memset(void* dst, int value, size_t bytes) {
  return guard([&] { GPH_HIP(hipMemset(dst, value, bytes)); });
}
int gph_synchronize(void) {
  return guard([&] { GPH_HIP(hipDeviceSynchronize()); });
}
int gph_vector_cast(gph_type in_type, const void* in, gph_type out_type,
                    void* out, int64_t n) {
  return guard([&] {
    if (n < 0) gph::fail(gph::kInvalid, "negative length");
    if (n == 0) return;
    need(in, "in");
    need(out, "out");
    gph::launchCast(in, dt(in_type), out, dt(out_type), n, nullptr);
    GPH_HIP(hipStreamSynchronize(nullptr));
  });
}

int gph_matrix_create(const gph_host_matrix* h, gph_format fmt,
                      int32_t rows_per_block, gph_matrix** out) {
  return guard([&] {
    need(h, "host matrix");
    need(out, "out");
    if (fmt != GPH_CRS && fmt != GPH_LDU)
      gph::fail(gph::kInvalid, "invalid matrix format");
    auto* A = new gph_matrix;
    try {
      A->A = std::make_unique<DeviceMatrix>(
          h->m, fmt == GPH_CRS ? MatrixFormat::CRS : MatrixFormat::LDU,
          rows_per_block);
    } catch (...) {
      delete A;
      throw;
    }
    *out = A;
  });
}
void gph_matrix_destroy(gph_matrix* A) { delete A; }
int gph_matrix_set_stream(gph_matrix* A, void* stream, int32_t synchronous) {
  return guard([&] {
    need(A, "matrix");
    GPH_HIP(hipStreamSynchronize(A->A->stream()));  // work already enqueued ends first
    A->A->setStream(static_cast<hipStream_t>(stream));
    A->async = stream != nullptr && !synchronous;
  });
}
int gph_stream_create(void** stream) {
  return guard([&] {
    need(stream, "stream");
    hipStream_t s = nullptr;
    GPH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
  });
}
int gph_stream_destroy(void* stream) {
  return guard([&] {
    need(stream, "stream");
    GPH_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  });
}
int gph_stream_synchronize(void* stream) {
  return guard([&] { GPH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream))); });
}
int gph_matrix_set_sliced(gph_matrix* A, int32_t sliced) {
  return guard([&] {
    need(A, "matrix");
    A->A->setUseSell(sliced != 0);
  });
}
int gph_matrix_device_bytes(const gph_matrix* A, size_t* bytes) {
  return guard([&] {
    need(A, "matrix");
    need(bytes, "bytes");
    *bytes = A->A->deviceBytes();
  });
}

int gph_spmv(const gph_matrix* A, gph_type value_type, gph_type x_type,
             const void* x, gph_type y_type, void* y, gph_type acc_type) {
  return guard([&] {
    need(A, "matrix");
    need(x, "x");
    need(y, "y");
    hipStream_t s = A->A->stream();
    A->A->spmv(Vec{const_cast<void*>(x), dt(x_type)}, Vec{y, dt(y_type)},
               dt(acc_type), dt(value_type), s);
    if (!A->async) GPH_HIP(hipStreamSynchronize(s));
  });
}
int gph_residual(const gph_matrix* A, gph_type value_type, gph_type x_type,
                 const void* x, gph_type b_type, const void* b,
                 gph_type r_type, void* r, gph_type acc_type, double* r64) {
  return guard([&] {
    need(A, "matrix");
    need(x, "x");
    need(b, "b");
    need(r, "r");
    hipStream_t s = A->A->stream();
    A->A->residual(Vec{const_cast<void*>(x), dt(x_type)},
                   Vec{const_cast<void*>(b), dt(b_type)}, Vec{r, dt(r_type)},
                   dt(acc_type), dt(value_type), s, r64);
    if (!A->async) GPH_HIP(hipStreamSynchronize(s));
  });
}
int gph_vector_norm(const gph_matrix* A, gph_type t, const void* x,
                    double* out) {
  return guard([&] {
    need(A, "matrix");
    need(x, "x");
    need(out, "out");
    *out = A->A->norm2(Vec{const_cast<void*>(x), dt(t)}, A->A->stream());
  });
}
int gph_vector_norm_kind(const gph_matrix* A, gph_norm norm, gph_type t,
                         const void* x, double* out) {
  return guard([&] {
    need(A, "matrix");
    need(x, "x");
    need(out, "out");
    if (norm != GPH_NORM_L2 && norm != GPH_NORM_L1 && norm != GPH_NORM_LINF)
      gph::fail(gph::kInvalid, "invalid norm " + std::to_string((int)norm));
    static_assert((int)GPH_NORM_L1 == gph::kNormL1 && (int)GPH_NORM_LINF == gph::kNormLinf &&
                      (int)GPH_NORM_L2 == gph::kNormL2,
                  "norm codes of the C-ABI and the kernels");
    *out = A->A->norm(Vec{const_cast<void*>(x), dt(t)}, (int)norm, A->A->stream());
  });
}
int gph_vector_norm_noalloc(const gph_matrix* A, gph_type t, const void* x,
                            double* out) {
  return guard([&] {
    need(A, "matrix");
    need(x, "x");
    need(out, "out");
    *out = gph::norm2NoAlloc(x, dt(t), A->A->n(), A->A->stream());
  });
}
int gph_bench_spmv(const gph_matrix* A, gph_type value_type, gph_type x_type,
                   const void* x, void* y, int32_t iters, double* avg_ms) {
  return guard([&] {
    need(A, "matrix");
    need(avg_ms, "avg_ms");
    if (iters < 1) gph::fail(gph::kInvalid, "iters must be >= 1");
    hipStream_t s = A->A->stream();
    gph::DType v = dt(value_type), xt = dt(x_type);
    Vec xv{const_cast<void*>(x), xt}, yv{y, xt};
    A->A->spmv(xv, yv, xt, v, s);  // warm-up
    hipEvent_t e0, e1;
    GPH_HIP(hipEventCreate(&e0));
    GPH_HIP(hipEventCreate(&e1));
    GPH_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < iters; i++) A->A->spmv(xv, yv, xt, v, s);
    GPH_HIP(hipEventRecord(e1, s));
    GPH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GPH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
  });
}

int gph_solver_create(const gph_matrix* A, const char* json, gph_solver** out) {
  return guard([&] {
    need(A, "matrix");
    need(json, "json");
    need(out, "out");
    auto* S = new gph_solver;
    try {
      S->A = A;
      S->cfg = solver::Configuration::fromJSON(json);
      S->S = solver::Solver::createSolver(*A->A, S->cfg);
    } catch (...) {
      delete S;
      throw;
    }
    *out = S;
  });
}
void gph_solver_destroy(gph_solver* S) { delete S; }

int gph_solve(gph_solver* S, gph_type x_type, void* x, gph_type b_type,
              const void* b, gph_solver_stats* stats) {
  return guard([&] {
    need(S, "solver");
    need(x, "x");
    need(b, "b");
    solver::SolverStats st = S->S->solve(Vec{x, dt(x_type)},
                                         Vec{const_cast<void*>(b), dt(b_type)});
    if (stats) {
      stats->initial_residual = st.initialResidual;
      stats->final_residual = st.finalResidual;
      stats->iterations = st.iterations;
      stats->inner_iterations = st.innerIterations;
      stats->converged = st.converged ? 1 : 0;
      stats->singular = st.singular ? 1 : 0;
      stats->seconds = st.seconds;
      std::memset(stats->name, 0, sizeof(stats->name));
      std::strncpy(stats->name, st.solverName.c_str(), sizeof(stats->name) - 1);
    }
  });
}
int gph_bench_apply(gph_solver* S, gph_type t, const void* rhs, void* out,
                    int32_t iters, double* avg_ms) {
  return guard([&] {
    need(S, "solver");
    need(rhs, "rhs");
    need(out, "out");
    need(avg_ms, "avg_ms");
    if (iters < 1) gph::fail(gph::kInvalid, "iters must be >= 1");
    if (!S->S->supportsDeviceApply())
      gph::fail(gph::kUnsupported, S->S->name() + " has no preconditioner apply");
    hipStream_t s = S->S->matrix().stream();
    Vec in{const_cast<void*>(rhs), dt(t)}, o{out, dt(t)};
    S->S->applyDevice(in, o, gph::PreOp{}, nullptr, s);  // warm-up
    hipEvent_t e0, e1;
    GPH_HIP(hipEventCreate(&e0));
    GPH_HIP(hipEventCreate(&e1));
    GPH_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < iters; i++) S->S->applyDevice(in, o, gph::PreOp{}, nullptr, s);
    GPH_HIP(hipEventRecord(e1, s));
    GPH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GPH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
  });
}
int gph_solver_name(const gph_solver* S, char* buf, size_t len) {
  return guard([&] {
    need(S, "solver");
    need(buf, "buf");
    std::string n = S->S->name();
    if (len == 0) gph::fail(gph::kInvalid, "buffer length 0");
    std::strncpy(buf, n.c_str(), len - 1);
    buf[len - 1] = 0;
  });
}
int gph_solver_uses_initial_guess(const gph_solver* S, int32_t* out) {
  return guard([&] {
    need(S, "solver");
    need(out, "out");
    *out = S->S->usesInitialGuess() ? 1 : 0;
  });
}
int gph_config_check(const char* json, char* name_buf, size_t len) {
  return guard([&] {
    need(json, "json");
    auto c = solver::Configuration::fromJSON(json);
    if (name_buf && len > 0) {
      std::string n = c->solverName();
      std::strncpy(name_buf, n.c_str(), len - 1);
      name_buf[len - 1] = 0;
    }
  });
}

int gph_batch_create(const gph_matrix* A, int32_t n_systems,
                     const int32_t* system_tile_start, const char* json,
                     gph_batch** out) {
  return guard([&] {
    need(A, "matrix");
    need(system_tile_start, "system_tile_start");
    need(json, "json");
    need(out, "out");
    if (n_systems < 1) gph::fail(gph::kInvalid, "n_systems must be >= 1");
    std::vector<int32_t> st(system_tile_start, system_tile_start + n_systems + 1);
    auto* b = new gph_batch;
    try {
      b->B = std::make_unique<graphene::batch::BatchSolver>(*A->A, st, json);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}
int gph_batch_create_rhs(const gph_matrix* A, int32_t n_rhs, const char* json,
                         gph_batch** out) {
  return guard([&] {
    need(A, "matrix");
    need(json, "json");
    need(out, "out");
    auto* b = new gph_batch;
    try {
      b->B = std::make_unique<graphene::batch::BatchSolver>(*A->A, n_rhs, json, nullptr);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}
void gph_batch_destroy(gph_batch* b) { delete b; }
int gph_batch_info(const gph_batch* b, int32_t* n_systems, int32_t* refinement,
                   int32_t* rows_per_wave) {
  return guard([&] {
    need(b, "batch");
    if (n_systems) *n_systems = b->B->numSystems();
    if (refinement) *refinement = b->B->refinement() ? 1 : 0;
    if (rows_per_wave) *rows_per_wave = b->B->tunedRowsPerWave();
  });
}
int gph_batch_solve(gph_batch* b, gph_type x_type, void* x, const void* b_f32,
                    gph_solver_stats* stats, double* seconds) {
  return guard([&] {
    need(b, "batch");
    need(x, "x");
    need(b_f32, "b");
    auto r = b->B->solve(Vec{x, dt(x_type)}, Vec{const_cast<void*>(b_f32), gph::DType::F32});
    if (stats) {
      for (size_t k = 0; k < r.size(); k++) {
        gph_solver_stats& o = stats[k];
        std::memset(&o, 0, sizeof(o));
        o.initial_residual = r[k].initialResidual;
        o.final_residual = r[k].finalResidual;
        o.iterations = r[k].iterations;
        o.inner_iterations = r[k].innerIterations;
        o.converged = r[k].converged ? 1 : 0;
        o.singular = r[k].singular ? 1 : 0;
        o.seconds = b->B->lastSeconds();
        std::strncpy(o.name, b->B->refinement() ? "batched IterativeRefinement"
                                                : "batched PBiCGStab",
                     sizeof(o.name) - 1);
      }
    }
    if (seconds) *seconds = b->B->lastSeconds();
  });
}

}  // extern "C"

int gph_server_create(int32_t workers, size_t hbm_budget_bytes, gph_server** out) {
  return guard([&] {
    need(out, "out");
    if (workers < 1) gph::fail(gph::kInvalid, "workers must be >= 1");
    auto s = std::make_unique<gph_server>();
    s->S = std::make_unique<graphene::serve::Server>(workers, hbm_budget_bytes);
    *out = s.release();
  });
}
void gph_server_destroy(gph_server* s) { delete s; }
int gph_server_add_system(gph_server* s, const gph_matrix* A, const char* json,
                          int32_t* system_id) {
  return guard([&] {
    need(s, "server");
    need(A, "matrix");
    need(json, "json");
    need(system_id, "system_id");
    *system_id = s->S->addSystem(*A->A, json);
  });
}
int gph_server_footprint(const gph_server* s, int32_t system_id, size_t* bytes) {
  return guard([&] {
    need(s, "server");
    need(bytes, "bytes");
    *bytes = s->S->footprint(system_id);
  });
}
int gph_server_resident_bytes(const gph_server* s, size_t* bytes) {
  return guard([&] {
    need(s, "server");
    need(bytes, "bytes");
    *bytes = s->S->residentBytes();
  });
}
int gph_server_run(gph_server* s, int64_t n, const int32_t* system,
                   const int32_t* n_rhs, const double* arrival_s,
                   const uint64_t* seed, double* const* x0_out,
                   gph_request_result* results) {
  return guard([&] {
    need(s, "server");
    if (n < 0) gph::fail(gph::kInvalid, "negative request count");
    if (n == 0) return;
    need(system, "system");
    need(n_rhs, "n_rhs");
    need(arrival_s, "arrival_s");
    need(seed, "seed");
    need(results, "results");
    std::vector<double*> xo;
    if (x0_out) xo.assign(x0_out, x0_out + n);
    auto r = s->S->run(std::vector<int32_t>(system, system + n),
                       std::vector<int32_t>(n_rhs, n_rhs + n),
                       std::vector<double>(arrival_s, arrival_s + n),
                       std::vector<uint64_t>(seed, seed + n), xo);
    for (int64_t i = 0; i < n; i++) {
      gph_request_result& o = results[i];
      o.latency = r[(size_t)i].latency;
      o.wait = r[(size_t)i].wait;
      o.service = r[(size_t)i].service;
      o.iterations = r[(size_t)i].iterations;
      o.inner_iterations = r[(size_t)i].innerIterations;
      o.rel_residual = r[(size_t)i].relResidual;
      o.converged = r[(size_t)i].converged;
      o.pad = 0;
    }
  });
}
