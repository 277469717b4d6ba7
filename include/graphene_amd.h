/*
 * graphene_amd: C-ABI of the MI355X sparse-solver hot path (CRS/LDU SpMV,
 * PBiCGStab, ILU(0) and Gauss-Seidel preconditioning, mixed-precision
 * iterative refinement).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the esa-tu-darmstadt/graphene checkout; lg/ = libgraphene/libgraphene/).
 * The reference is a C++ API with no FFI of its own; these functions are what
 * an FFI for that API binds: opaque handles, plain pointers and sizes, int
 * status (0 = OK), gph_last_error() for the message. No exception crosses the
 * boundary.
 *
 * Vectors are device pointers in the matrix's tile-major row order (the
 * reference's distributed vector layout, DistributedTileLayout.cpp:132-154):
 * use gph_host_matrix_to_local / gph_host_matrix_to_global to convert from
 * and to the natural row order. Vectors "with halo" and "without halo" are
 * the same vector here: halo cells are read from their owner's slot.
 */
#ifndef GRAPHENE_AMD_H
#define GRAPHENE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GPH_OK = 0,
  GPH_ERR_INVALID = 1,
  GPH_ERR_RUNTIME = 2,
  GPH_ERR_HIP = 3,
  GPH_ERR_UNSUPPORTED = 4
} gph_status;

/* Element types, names of lg/common/Type.cpp:69-88. GPH_TWOFLOAT32 is the
 * double-word type two<float>: per element a float pair {hi, lo} (8 bytes)
 * whose value is hi + lo. It is the extended precision of refinement
 * ("extendedPrecisionType": "twofloat32") and the x and accumulation type of
 * gph_residual; casts convert float32/float64 to and from it. */
typedef enum {
  GPH_FLOAT16 = 1,
  GPH_FLOAT32 = 2,
  GPH_FLOAT64 = 3,
  GPH_TWOFLOAT32 = 4
} gph_type;

/* lg/matrix/MatrixFormat.hpp */
typedef enum { GPH_CRS = 0, GPH_LDU = 1 } gph_format;

typedef struct gph_host_matrix gph_host_matrix;
typedef struct gph_matrix gph_matrix;
typedef struct gph_solver gph_solver;

/* lg/matrix/solver/SolverStats.hpp:29-69 (+ total inner iterations and the
 * wall time of the solve). */
typedef struct {
  double initial_residual;
  double final_residual;
  int64_t iterations;
  int64_t inner_iterations;
  int32_t converged;
  int32_t singular;
  double seconds;
  char name[96];
} gph_solver_stats;

typedef struct {
  int64_t nrows;
  int64_t nnz_off;
  int64_t num_halo;
  int32_t num_tiles;
  int32_t max_colors;
  int64_t num_color_entries;
  int64_t num_upper_color_entries;
  int32_t multicolor_recommended;
  int32_t max_tile_rows;
} gph_host_matrix_info;

/* ---- errors --------------------------------------------------------------- */
/* Message of the last failing call on this thread ("" if none). */
const char* gph_last_error(void);
int32_t gph_version(void);

/* ---- triplet generators (COO) --------------------------------------------- *
 * Query the count with rows == NULL, then call again with buffers of that
 * size. Values are float32 (the reference generates Type::FLOAT32).         */
/* generate3DPoissonTripletMatrix, lg/matrix/host/details/Poisson.cpp:26-69 */
int gph_generate_poisson3d(int64_t nx, int64_t ny, int64_t nz, int64_t* count,
                           int64_t* rows, int64_t* cols, float* vals);
/* No reference counterpart: 27-point convection-diffusion (upwind, cell
 * Peclet `peclet`, seeded direction). */
int gph_generate_convdiff3d27(int64_t nx, int64_t ny, int64_t nz,
                              double peclet, uint64_t seed, int64_t* count,
                              int64_t* rows, int64_t* cols, float* vals);
/* No reference counterpart: 2D 9-point convection-diffusion. */
int gph_generate_convdiff2d9(int64_t nx, int64_t ny, double diffusion,
                             double peclet, uint64_t seed, int64_t* count,
                             int64_t* rows, int64_t* cols, float* vals);
/* No reference counterpart: unstructured power-law matrix. */
int gph_generate_powerlaw(int64_t n, int32_t min_len, int32_t max_len,
                          double exponent, int32_t band, double local_fraction,
                          uint64_t seed, int64_t* count, int64_t* rows,
                          int64_t* cols, float* vals);
/* Portable RNG: U[-1,1) floats from a seed. */
int gph_uniform_pm1(uint64_t seed, int64_t n, float* out);

/* ---- partitions (stand-ins for the METIS call, CRSHostMatrix.cpp:139-271) */
int gph_partition_grid3d(int64_t nx, int64_t ny, int64_t nz, int64_t bx,
                         int64_t by, int64_t bz, int32_t* row_to_tile,
                         int32_t* num_tiles);
/* Stand-in for the METIS partition of CRSHostMatrix.cpp:139-271 (METIS is
 * not available): recursive breadth-first bisection of the symmetrized
 * pattern of a COO matrix into num_tiles parts weighted by row length + 1;
 * deterministic. row_to_tile receives nrows entries. */
int gph_partition_graph(int64_t nrows, int64_t nnz, const int64_t* rows,
                        const int64_t* cols, int32_t num_tiles,
                        int32_t* row_to_tile);

/* ---- host matrices -------------------------------------------------------- */
/* HostMatrix(CRS, triplets, numTiles), lg/matrix/host/HostMatrix.cpp:31-44 and
 * CRSHostMatrix::CRSHostMatrix, CRSHostMatrix.cpp:63-85. row_to_tile may be
 * NULL (one tile, or contiguous row ranges when num_tiles > 1). */
int gph_host_matrix_from_coo(int64_t nrows, int64_t nnz, const int64_t* rows,
                             const int64_t* cols, const float* vals,
                             int32_t num_tiles, const int32_t* row_to_tile,
                             gph_host_matrix** out);
/* host::generate3DPoissonMatrix, lg/matrix/host/HostMatrix.hpp:92-96 */
int gph_host_matrix_poisson3d(int64_t nx, int64_t ny, int64_t nz,
                              int32_t num_tiles, const int32_t* row_to_tile,
                              gph_host_matrix** out);
/* No reference counterpart: 27-point convection-diffusion host matrix. */
int gph_host_matrix_convdiff3d27(int64_t nx, int64_t ny, int64_t nz,
                                 double peclet, uint64_t seed,
                                 int32_t num_tiles, const int32_t* row_to_tile,
                                 gph_host_matrix** out);
/* No reference counterpart: power-law host matrix. */
int gph_host_matrix_powerlaw(int64_t n, int32_t min_len, int32_t max_len,
                             double exponent, int32_t band,
                             double local_fraction, uint64_t seed,
                             int32_t num_tiles, const int32_t* row_to_tile,
                             gph_host_matrix** out);
/* host::loadMatrixFromFile, lg/matrix/host/HostMatrix.hpp:88-91 */
int gph_host_matrix_load_mtx(const char* path, int32_t num_tiles,
                             const int32_t* row_to_tile,
                             gph_host_matrix** out);
void gph_host_matrix_destroy(gph_host_matrix* h);
int gph_host_matrix_get_info(const gph_host_matrix* h,
                             gph_host_matrix_info* info);
/* Structure export (any pointer may be NULL). Sizes: tile_start T+1,
 * new_to_old n, num_interior T, num_separator_regions T, row_ptr n+1,
 * col/col_local/off nnz_off, diag n, halo_start T+1, halo_global num_halo. */
int gph_host_matrix_export(const gph_host_matrix* h, int64_t* tile_start,
                           int64_t* new_to_old, int32_t* num_interior,
                           int32_t* num_separator_regions, int64_t* row_ptr,
                           int32_t* col, int32_t* col_local, float* off,
                           float* diag, int64_t* halo_start,
                           int64_t* halo_global);
/* Level sets (calculateRowColors / calculateColorAddressings,
 * CRSHostMatrix.cpp:527-675): level n, color_sort_addr n, color_start
 * num_color_entries, num_colors T. */
int gph_host_matrix_export_levels(const gph_host_matrix* h, int32_t* level,
                                  int32_t* color_sort_addr,
                                  int64_t* color_start, int32_t* num_colors);
/* LDU addressing (LDUAddressing.hpp:30-63), sizes: faces = nnz_off / 2.
 * lower_addr/upper_addr/upper faces, owner_start n+1, neighbour_start n+1,
 * neighbour_col_ind/lower faces. */
int gph_host_matrix_export_ldu(const gph_host_matrix* h, int64_t* num_faces,
                               int32_t* lower_addr, int32_t* upper_addr,
                               int32_t* owner_start, int32_t* neighbour_start,
                               int32_t* neighbour_col_ind, float* upper,
                               float* lower);
/* Permute a vector between natural row order and tile-major order
 * (DistributedTileLayout::decomposeVector, DistributedTileLayout.cpp:157-183). */
int gph_host_matrix_to_local(const gph_host_matrix* h, gph_type t,
                             const void* global_in, void* local_out);
int gph_host_matrix_to_global(const gph_host_matrix* h, gph_type t,
                              const void* local_in, void* global_out);

/* ---- device memory -------------------------------------------------------- */
int gph_device_count(int32_t* count);
int gph_malloc(size_t bytes, void** ptr);
int gph_free(void* ptr);
int gph_memcpy_htod(void* dst, const void* src, size_t bytes);
int gph_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int gph_memset(void* dst, int value, size_t bytes);
int gph_synchronize(void);
/* No reference counterpart: HIP streams for callers without HIP bindings
 * (a void* is a hipStream_t). */
int gph_stream_create(void** stream);
int gph_stream_destroy(void* stream);
int gph_stream_synchronize(void* stream);
/* out = (out_type) in, n elements, one streaming kernel on the default stream
 * (Expression::cast, libgraphene/dsl/tensor/Expression.hpp:152). With
 * in_type == out_type it is a plain copy kernel. */
int gph_vector_cast(gph_type in_type, const void* in, gph_type out_type,
                    void* out, int64_t n);

/* ---- device matrices ------------------------------------------------------ */
/* HostMatrix::copyToTile, lg/matrix/host/HostMatrix.hpp:53 (CRS) and the
 * LDUMatrix constructor, lg/matrix/details/ldu/LDUMatrix.hpp:49-57 (LDU).
 * rows_per_block: SpMV rows per workgroup (0 = automatic). */
int gph_matrix_create(const gph_host_matrix* h, gph_format fmt,
                      int32_t rows_per_block, gph_matrix** out);
void gph_matrix_destroy(gph_matrix* A);
/* No reference counterpart: kernel form of the row gather, 1 = sliced ELL,
 * 0 = row blocks staged in LDS. Both give identical results. Without a call
 * the form is chosen per matrix: sliced ELL, except for rows of strongly
 * varying length whose columns mostly span more than 2^16 (row blocks; a
 * CRS matrix of that kind also keeps its entries in column order per row
 * block, 10 more bytes per off-diagonal). */
int gph_matrix_set_sliced(gph_matrix* A, int32_t sliced);
/* No reference counterpart (the reference runs on a Poplar engine): the
 * stream on which A's operations and those of solvers created on A are
 * enqueued, a caller's hipStream_t, or NULL for the matrix's own stream;
 * A's work is then ordered after the caller's earlier work on that stream.
 * synchronous = 0: gph_spmv and gph_residual return once their kernels are
 * enqueued (the caller synchronizes the stream before reading y or r);
 * every other call returns after its work is done. The stream must outlive
 * its use by A. */
int gph_matrix_set_stream(gph_matrix* A, void* stream, int32_t synchronous);
int gph_matrix_device_bytes(const gph_matrix* A, size_t* bytes);

/* y = A*x (Matrix::operator*, lg/matrix/Matrix.hpp:95-97; MatrixBase::spmv,
 * MatrixBase.hpp:66-74). Matrix values used in `value_type`. */
int gph_spmv(const gph_matrix* A, gph_type value_type, gph_type x_type,
             const void* x, gph_type y_type, void* y, gph_type acc_type);
/* r = b - A*x (Matrix::residual, lg/matrix/Matrix.hpp:85-93). r64 (optional)
 * receives the unrounded float64 residual when acc_type is float64, the
 * double-word residual rounded to float64 when acc_type is twofloat32 (then
 * x is twofloat32, b, r and value_type float32). */
int gph_residual(const gph_matrix* A, gph_type value_type, gph_type x_type,
                 const void* x, gph_type b_type, const void* b,
                 gph_type r_type, void* r, gph_type acc_type, double* r64);
/* ||x||_2 (Matrix::vectorNorm, lg/matrix/Matrix.hpp:115-117). */
int gph_vector_norm(const gph_matrix* A, gph_type t, const void* x,
                    double* out);
/* Norms of lg/matrix/Norm.hpp:24-34 (names "L1", "L2", "LINF" in solver
 * configurations). L1 is the sum of magnitudes and LINF the largest
 * magnitude (the reference's L1 sums signed values, Tensor.cpp:430-431, and
 * its LINF reduction is unavailable). */
typedef enum { GPH_NORM_L2 = 0, GPH_NORM_L1 = 1, GPH_NORM_LINF = 2 } gph_norm;
/* ||x|| in the given norm (Matrix::vectorNorm(norm, x),
 * lg/matrix/Matrix.hpp:115-117). */
int gph_vector_norm_kind(const gph_matrix* A, gph_norm norm, gph_type t,
                         const void* x, double* out);
/* No reference counterpart: ||x||_2 as gph_vector_norm (same kernel, same
 * result bits), for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops). It calls no hipMalloc/hipFree: the
 * reduction scratch is a static device array of 33,024 bytes per GPU, held
 * until the process ends. n is the matrix's row count, at most 2^31 - 1.
 * Calls that overlap on one GPU share that array and race: issue them one
 * at a time per GPU. */
int gph_vector_norm_noalloc(const gph_matrix* A, gph_type t, const void* x,
                            double* out);
/* Mean device time (ms) of `iters` back-to-back SpMV launches, HIP events on
 * the launch stream. */
int gph_bench_spmv(const gph_matrix* A, gph_type value_type, gph_type x_type,
                   const void* x, void* y, int32_t iters, double* avg_ms);

/* ---- solvers -------------------------------------------------------------- */
/* solver::Configuration::fromJSON (lg/matrix/solver/Configuration.cpp:126-136)
 * + Solver::createSolver (lg/matrix/solver/Solver.cpp:32-56). ILU factorizes
 * here, as the reference does in its constructor. `json` is JSON with
 * comments. */
int gph_solver_create(const gph_matrix* A, const char* json, gph_solver** out);
void gph_solver_destroy(gph_solver* S);
/* Solver::solve(x, b) (lg/matrix/solver/Solver.hpp:41); Matrix::solve
 * (lg/matrix/Matrix.cpp:26-30) is create + solve. */
int gph_solve(gph_solver* S, gph_type x_type, void* x, gph_type b_type,
              const void* b, gph_solver_stats* stats);
/* Mean device time (ms) of `iters` back-to-back preconditioner applications
 * out = M^-1 rhs of an ILU or Gauss-Seidel solver (the apply inside
 * pbicgstab/Solver.cpp's iteration), HIP events on the launch stream. No
 * reference counterpart: a timing hook for the substitution kernels. */
int gph_bench_apply(gph_solver* S, gph_type t, const void* rhs, void* out,
                    int32_t iters, double* avg_ms);
/* Solver::name(), Solver::usesInitialGuess() */
int gph_solver_name(const gph_solver* S, char* buf, size_t len);
int gph_solver_uses_initial_guess(const gph_solver* S, int32_t* out);
/* Validates a configuration without a matrix (CPU-only). */
int gph_config_check(const char* json, char* name_buf, size_t len);

/* ---- request server (continuous batching) ---------------------------------
 * No reference counterpart: the serving extension of the solve path. Systems
 * (a device matrix + a PBiCGStab or IterativeRefinement configuration) are
 * registered once and stay resident with their factors and workspace;
 * requests (system, right-hand sides, seed) arrive on a schedule and are
 * admitted in arrival order by `workers` host threads. A request runs on a
 * slot: the Krylov workspace and buffers of its 1-8 right-hand sides, solved
 * as one block on a stream of the slot's own (ILU-preconditioned
 * configurations; others run their right-hand sides in turn, one request
 * per system at a time), so requests on one system overlap and share its
 * factors. Slots stay resident for later requests of the same system and
 * right-hand-side count; every resident slot counts against
 * `hbm_budget_bytes`, and a request that does not fit evicts idle slots,
 * least recently used first, before it waits. Each right-hand side is
 * b = fl(A x_true), x_true ~ U[-1,1) from the seed, solved from x = 0. */
typedef struct gph_server gph_server;
typedef struct {
  double latency;      /* finish - arrival, seconds */
  double wait;         /* admission - arrival */
  double service;      /* finish - admission */
  int64_t iterations;  /* summed over the right-hand sides */
  int64_t inner_iterations;
  double rel_residual; /* largest final ||r|| / ||b|| */
  int32_t converged;   /* all right-hand sides converged */
  int32_t pad;
} gph_request_result;
int gph_server_create(int32_t workers, size_t hbm_budget_bytes, gph_server** out);
void gph_server_destroy(gph_server* s);
/* The matrix must outlive the server. */
int gph_server_add_system(gph_server* s, const gph_matrix* A, const char* json,
                          int32_t* system_id);
/* Device bytes of a slot for one right-hand side of the system. */
int gph_server_footprint(const gph_server* s, int32_t system_id, size_t* bytes);
/* Device bytes of the slots resident now (between runs). */
int gph_server_resident_bytes(const gph_server* s, size_t* bytes);
/* Runs n requests; arrival_s[i] = seconds after the start; x0_out (nullable,
 * or an array of n nullable pointers to n_rows doubles) receives the solution
 * of request i's first right-hand side in the matrix's tile-major order. */
int gph_server_run(gph_server* s, int64_t n, const int32_t* system,
                   const int32_t* n_rhs, const double* arrival_s,
                   const uint64_t* seed, double* const* x0_out,
                   gph_request_result* results);

/* ---- batched solves ---------------------------------------------------------
 * No reference counterpart: many independent systems solved together. The
 * systems are consecutive tile ranges of one device matrix (tiles
 * system_tile_start[k] .. system_tile_start[k+1]-1 form system k) with no
 * entry coupling two systems. json: IterativeRefinement (float64 extended,
 * float32 working type) around PBiCGStab, or PBiCGStab alone, each with an
 * ILU preconditioner; every system follows the single-system algorithm from
 * x = 0 with its own scalars and convergence. The iterations run as a
 * captured HIP graph replayed until every system has finished. With
 * refinement the outer step (phi += d, FP64 residual, norm, convergence,
 * inner restart) runs on the device too, per system as soon as its inner
 * solve has finished; "outerStep": "host" in the top-level configuration
 * instead ends every system's inner solve before the outer steps, which the
 * host drives. The ILU substitution kernel (and the packed kernel's rows per
 * wave) is picked by timing at creation unless the ILU configuration sets
 * "substitution" or "rowsPerWave". */
typedef struct gph_batch gph_batch;
int gph_batch_create(const gph_matrix* A, int32_t n_systems,
                     const int32_t* system_tile_start, const char* json,
                     gph_batch** out);
/* No reference counterpart: n_rhs (1..8) right-hand sides of one system A,
 * solved together as a batch whose systems are the right-hand sides: x and b
 * hold n_rhs blocks of A's rows (block j = right-hand side j, tile-major
 * order). Each product reads the matrix streams once for all right-hand
 * sides and each tile's ILU(0) substitution runs for all of them in one
 * launch; every right-hand side keeps its own scalars and convergence. The
 * matrix needs its sliced layout (every CRS and LDU device matrix has one).
 * Solve, info and destroy as for gph_batch_create. */
int gph_batch_create_rhs(const gph_matrix* A, int32_t n_rhs, const char* json,
                         gph_batch** out);
void gph_batch_destroy(gph_batch* b);
/* refinement: 1 for IterativeRefinement (x float64), 0 for PBiCGStab (x
 * float32); rows_per_wave: the packed ILU kernel's rows per wave chosen at
 * creation, 0 when the row-per-lane kernel was chosen. */
int gph_batch_info(const gph_batch* b, int32_t* n_systems, int32_t* refinement,
                   int32_t* rows_per_wave);
/* x (overwritten) and b (float32) in the matrix's tile-major order; stats
 * (nullable) receives n_systems entries (iterations = outer iterations for
 * refinement, inner_iterations = all PBiCGStab iterations); seconds
 * (nullable) the wall time of the solve. */
int gph_batch_solve(gph_batch* b, gph_type x_type, void* x, const void* b_f32,
                    gph_solver_stats* stats, double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* GRAPHENE_AMD_H */
