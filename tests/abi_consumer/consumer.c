/* A C99 consumer of the drop-in boundary: it includes only the public header
 * include/graphene_amd.h and links libgraphene_amd.so, as a binding on the
 * reference side would (the steps of applications/benchmark/benchmark.cpp:
 * 144-183: generate a Poisson matrix, copy it to the device, b = A x, solve).
 *
 *   consumer host   host-side calls only (no GPU needed)
 *   consumer gpu    also the device calls: SpMV, a PBiCGStab + ILU(0) solve,
 *                   and an SpMV on a stream of the caller's
 *
 * Prints "consumer ok" and exits 0 on success. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "graphene_amd.h"

#define CHECK(call)                                                     \
  do {                                                                  \
    int rc_ = (call);                                                   \
    if (rc_ != GPH_OK) {                                                \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, gph_last_error()); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

static int fail(const char* what) {
  fprintf(stderr, "check failed: %s\n", what);
  return 1;
}

int main(int argc, char** argv) {
  const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
  const int64_t nx = 8, ny = 7, nz = 6, n = nx * ny * nz;
  if (gph_version() <= 0) return fail("version");

  /* triplet generator: count query, then the triplets */
  int64_t count = 0;
  CHECK(gph_generate_poisson3d(nx, ny, nz, &count, NULL, NULL, NULL));
  if (count != n + 2 * ((nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)))
    return fail("Poisson triplet count");

  /* host matrix (generate3DPoissonMatrix) and its structure */
  gph_host_matrix* h = NULL;
  CHECK(gph_host_matrix_poisson3d(nx, ny, nz, 1, NULL, &h));
  gph_host_matrix_info info;
  CHECK(gph_host_matrix_get_info(h, &info));
  if (info.nrows != n || info.nnz_off != count - n || info.num_tiles != 1)
    return fail("host matrix info");

  /* configuration parsing and error reporting */
  char name[96];
  const char* cfg =
      "{\"type\": \"PBiCGStab\", \"relResidual\": 1e-6, \"maxIterations\": 200,"
      " \"preconditioner\": {\"type\": \"ILU\"}}  // comments are allowed";
  CHECK(gph_config_check(cfg, name, sizeof name));
  if (strstr(name, "PBiCGStab") == NULL) return fail("solver name");
  if (gph_config_check("{\"type\": \"Nope\"}", name, sizeof name) != GPH_ERR_INVALID ||
      strstr(gph_last_error(), "Nope") == NULL)
    return fail("error path");

  if (gpu) {
    int32_t devices = 0;
    CHECK(gph_device_count(&devices));
    if (devices < 1) return fail("no device");
    gph_matrix* A = NULL;
    CHECK(gph_matrix_create(h, GPH_CRS, 0, &A));
    float* hx = malloc((size_t)n * sizeof(float));
    float* hy = malloc((size_t)n * sizeof(float));
    for (int64_t i = 0; i < n; i++) hx[i] = 1.0f;
    void *x = NULL, *b = NULL;
    CHECK(gph_malloc((size_t)n * 4, &x));
    CHECK(gph_malloc((size_t)n * 4, &b));
    CHECK(gph_memcpy_htod(x, hx, (size_t)n * 4));
    CHECK(gph_spmv(A, GPH_FLOAT32, GPH_FLOAT32, x, GPH_FLOAT32, b, GPH_FLOAT32));
    CHECK(gph_memcpy_dtoh(hy, b, (size_t)n * 4));
    /* row sum of the 7-point Poisson row: 6 minus its in-domain neighbours
     * (one tile: the tile-major order is the natural order) */
    for (int64_t k = 0; k < nz; k++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
          const int nb = (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
          if (hy[(k * ny + j) * nx + i] != (float)(6 - nb)) return fail("SpMV row sum");
        }
    /* solve A x = b from x = 0: x = 1 */
    CHECK(gph_memset(x, 0, (size_t)n * 4));
    gph_solver* S = NULL;
    CHECK(gph_solver_create(A, cfg, &S));
    gph_solver_stats st;
    CHECK(gph_solve(S, GPH_FLOAT32, x, GPH_FLOAT32, b, &st));
    CHECK(gph_memcpy_dtoh(hx, x, (size_t)n * 4));
    double err = 0;
    for (int64_t i = 0; i < n; i++) err = fmax(err, fabs(hx[i] - 1.0));
    if (!st.converged || st.singular || st.iterations < 1 || err > 1e-4)
      return fail("solve");
    /* the same SpMV on a stream of the caller's, returning once enqueued */
    void* stream = NULL;
    CHECK(gph_stream_create(&stream));
    CHECK(gph_matrix_set_stream(A, stream, 0));
    CHECK(gph_memset(b, 0, (size_t)n * 4));
    CHECK(gph_spmv(A, GPH_FLOAT32, GPH_FLOAT32, x, GPH_FLOAT32, b, GPH_FLOAT32));
    CHECK(gph_stream_synchronize(stream));
    CHECK(gph_memcpy_dtoh(hy, b, (size_t)n * 4));
    CHECK(gph_matrix_set_stream(A, NULL, 1));
    CHECK(gph_stream_destroy(stream));
    /* x is the solution (1 to 1e-4): A x gives the row sums again */
    for (int64_t k = 0; k < nz; k++)
      for (int64_t j = 0; j < ny; j++)
        for (int64_t i = 0; i < nx; i++) {
          const int nb = (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1) + (k > 0) + (k < nz - 1);
          if (fabs(hy[(k * ny + j) * nx + i] - (6 - nb)) > 1e-2)
            return fail("SpMV on the caller's stream");
        }
    gph_solver_destroy(S);
    CHECK(gph_free(x));
    CHECK(gph_free(b));
    gph_matrix_destroy(A);
    free(hx);
    free(hy);
  }
  gph_host_matrix_destroy(h);
  printf("consumer ok\n");
  return 0;
}
