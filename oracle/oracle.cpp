// CPU restatement of Graphene's sparse-solver hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the
// smoke() entry and bench.py's cpu_baseline leg load it. The product
// (libgraphene_amd.so) never links, loads or calls it.
//
// Every function follows the reference file cited beside it (paths relative to
// the reference checkout, "lg/" = libgraphene/libgraphene/). Floating point is
// evaluated strictly in source order with FP contraction disabled
// (-ffp-contract=off), so per-row sums are reproducible bit for bit.
//
// Data layout shared by every function below ("tile-major" layout):
//   rows are numbered tile by tile; inside a tile in the reference's local
//   order (interior rows, then separator regions). Column indices are
//   tile-major row positions, stored per row in the reference's *local*
//   column order (in-tile columns ascending, then halo columns in halo order).
//   The halo exchange of the reference (lg/matrix/details/MatrixBase.cpp:36-109)
//   copies the owner's value into the halo slot, which is the same as reading
//   the owner's position directly.
//
// Parity pins (tests/test_oracle.py): Poisson closed form and Kronecker
// construction, scipy COO->CSR, scipy FP64 SpMV with a Higham bound, LDU == CRS,
// ILU0 (L*U)_ij == A_ij on the pattern plus dense brute force, triangular
// solves vs scipy, GS fixed point and dense textbook sweep, level-set
// validity and closed-form level counts, reduction known-answer tests of
// lg/dsl/tensor/tests/TensorTests.cpp, BiCGStab vs scipy.sparse.linalg.bicgstab
// (no preconditioner), MPIR final FP64 residual.

#include <omp.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#define ORC_API extern "C"

// ---------------------------------------------------------------------------
// Portable RNG (no reference counterpart; fixed by this project so that host
// generators, CPU tests and GPU tests draw identical inputs).
// splitmix64 seeds xoshiro256**; U[-1,1) float = ((u >> 40) * 2^-24) * 2 - 1.
// ---------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Xoshiro {
  uint64_t s[4];
};
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static void xo_seed(Xoshiro* r, uint64_t seed) {
  uint64_t sm = seed;
  for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&sm);
}
static uint64_t xo_next(Xoshiro* r) {
  uint64_t* s = r->s;
  uint64_t result = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}
static float xo_uniform_pm1(Xoshiro* r) {
  float u = (float)(xo_next(r) >> 40) * (1.0f / 16777216.0f);
  return u * 2.0f - 1.0f;
}

ORC_API void orc_rng_uniform_pm1(uint64_t seed, int64_t n, float* out) {
  Xoshiro r;
  xo_seed(&r, seed);
  for (int64_t i = 0; i < n; i++) out[i] = xo_uniform_pm1(&r);
}

ORC_API void orc_rng_u64(uint64_t seed, int64_t n, uint64_t* out) {
  Xoshiro r;
  xo_seed(&r, seed);
  for (int64_t i = 0; i < n; i++) out[i] = xo_next(&r);
}

// ---------------------------------------------------------------------------
// 7-point 3D Poisson triplets. lg/matrix/host/details/Poisson.cpp:26-69
// Push order per row: diag, i-1, i+1, j-1, j+1, k-1, k+1 (Poisson.cpp:42-60).
// rows == nullptr -> count only.
// ---------------------------------------------------------------------------
ORC_API int64_t orc_poisson7(int64_t nx, int64_t ny, int64_t nz, int64_t* rows,
                     int64_t* cols, float* vals) {
  int64_t c = 0;
  auto push = [&](int64_t r, int64_t col, float v) {
    if (rows) {
      rows[c] = r;
      cols[c] = col;
      vals[c] = v;
    }
    c++;
  };
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t row = k * nx * ny + j * nx + i;
        push(row, row, 6.0f);
        if (i > 0) push(row, row - 1, -1.0f);
        if (i < nx - 1) push(row, row + 1, -1.0f);
        if (j > 0) push(row, row - nx, -1.0f);
        if (j < ny - 1) push(row, row + nx, -1.0f);
        if (k > 0) push(row, row - nx * ny, -1.0f);
        if (k < nz - 1) push(row, row + nx * ny, -1.0f);
      }
  return c;
}

// ---------------------------------------------------------------------------
// COO -> CRS with split diagonal.
// sortTripletMatrx: lg/matrix/host/details/CoordinateFormat.cpp:29-40 (sort by
// row, then column). convertToCRS: lg/matrix/host/details/crs/CRSHostMatrix.cpp:87-137
// (diagonal into a dense array, off-diagonals into rowPtr/colInd/offDiag).
// Duplicates are summed (Matrix Market convention; the reference de-duplicates
// via std::set later, CRSHostMatrix.cpp:411-419 - deliberate deviation).
// Returns nnz_off.
// ---------------------------------------------------------------------------
ORC_API int64_t orc_coo_to_crs(int64_t n, int64_t nnz, const int64_t* rows,
                       const int64_t* cols, const float* vals, int64_t* rowPtr,
                       int32_t* colInd, float* offDiag, float* diag) {
  std::vector<int64_t> idx(nnz);
  for (int64_t i = 0; i < nnz; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
    if (rows[a] != rows[b]) return rows[a] < rows[b];
    return cols[a] < cols[b];
  });
  for (int64_t i = 0; i < n; i++) diag[i] = 0.0f;
  int64_t c = 0;
  rowPtr[0] = 0;
  int64_t lastRow = 0;
  for (int64_t t = 0; t < nnz; t++) {
    int64_t e = idx[t];
    int64_t r = rows[e], col = cols[e];
    if (r == col) {
      diag[r] = diag[r] + vals[e];
      continue;
    }
    while (lastRow != r) {
      rowPtr[++lastRow] = c;
    }
    if (c > 0 && (int64_t)colInd[c - 1] == col && lastRow == r &&
        rowPtr[r] < c) {
      offDiag[c - 1] = offDiag[c - 1] + vals[e];
      continue;
    }
    colInd[c] = (int32_t)col;
    offDiag[c] = vals[e];
    c++;
  }
  while (lastRow < n) rowPtr[++lastRow] = c;
  return c;
}

// ---------------------------------------------------------------------------
// Tile layout for a given row->tile map.
// calculateTileLayouts: CRSHostMatrix.cpp:273-354; region bookkeeping and local
// order: lg/matrix/host/DistributedTileLayout.cpp:63-126 (interior rows
// ascending, then separator regions in first-encounter order with cells
// ascending, then halo regions).
// Halo regions of a tile are ordered by (source tile, source region index):
// the reference's order depends on TBB scheduling (CRSHostMatrix.cpp:310-346),
// this fixes one deterministic order.
// calculateLocalAddressings: CRSHostMatrix.cpp:379-432 (columns as sorted
// local ids; halo ids are the largest). decomposeOffDiagCoefficients:
// CRSHostMatrix.cpp:465-525 (values permuted with their columns).
//
// Outputs (caller allocated):
//   tileStart[T+1], newToOld[n], numInterior[T], numSeparatorRegions[T],
//   haloStart[T+1], haloGlobal[cap=nnz_off], lrowPtr[n+1], lcolLocal[nnz_off],
//   lcol[nnz_off] (tile-major position of the column), loff[nnz_off], ldiag[n].
// Returns total halo count, or -1 on error.
// ---------------------------------------------------------------------------
ORC_API int64_t orc_tile_layout(int64_t n, const int64_t* rowPtr, const int32_t* colInd,
                        const float* offDiag, const float* diag, int32_t T,
                        const int32_t* rowToTile, int64_t* tileStart,
                        int64_t* newToOld, int32_t* numInterior,
                        int32_t* numSeparatorRegions, int64_t* haloStart,
                        int64_t* haloGlobal, int64_t* lrowPtr,
                        int32_t* lcolLocal, int64_t* lcol, float* loff,
                        float* ldiag) {
  struct Region {
    std::vector<int32_t> dst;
    std::vector<int64_t> cells;
  };
  std::vector<std::vector<int64_t>> interior(T);
  std::vector<std::vector<Region>> seps(T);
  for (int64_t r = 0; r < n; r++)
    if (rowToTile[r] < 0 || rowToTile[r] >= T) return -1;
  // Destination tiles of a row: tiles owning its columns
  // (CRSHostMatrix.cpp:289-299) united with tiles whose rows reference it.
  // The two coincide for structurally symmetric patterns; the union keeps
  // nonsymmetric patterns well defined (the reference throws there).
  std::vector<std::set<int32_t>> refBy(n);
  for (int64_t r = 0; r < n; r++)
    for (int64_t a = rowPtr[r]; a < rowPtr[r + 1]; a++)
      if (rowToTile[colInd[a]] != rowToTile[r])
        refBy[colInd[a]].insert(rowToTile[r]);
  for (int64_t r = 0; r < n; r++) {
    int32_t t = rowToTile[r];
    std::set<int32_t> nb = refBy[r];
    for (int64_t a = rowPtr[r]; a < rowPtr[r + 1]; a++) {
      int32_t ct = rowToTile[colInd[a]];
      if (ct != t) nb.insert(ct);
    }
    if (nb.empty()) {
      interior[t].push_back(r);
    } else {
      std::vector<int32_t> key(nb.begin(), nb.end());
      Region* reg = nullptr;
      for (auto& R : seps[t])
        if (R.dst == key) {
          reg = &R;
          break;
        }
      if (!reg) {
        seps[t].push_back(Region{key, {}});
        reg = &seps[t].back();
      }
      reg->cells.push_back(r);
    }
  }
  // halo lists per destination tile: sources ascending, regions ascending
  std::vector<std::vector<int64_t>> halo(T);
  for (int32_t s = 0; s < T; s++)
    for (auto& R : seps[s])
      for (int32_t d : R.dst)
        halo[d].insert(halo[d].end(), R.cells.begin(), R.cells.end());

  std::vector<int64_t> oldToNew(n, -1);
  int64_t pos = 0;
  for (int32_t t = 0; t < T; t++) {
    tileStart[t] = pos;
    numInterior[t] = (int32_t)interior[t].size();
    numSeparatorRegions[t] = (int32_t)seps[t].size();
    for (int64_t r : interior[t]) {
      newToOld[pos] = r;
      oldToNew[r] = pos++;
    }
    for (auto& R : seps[t])
      for (int64_t r : R.cells) {
        newToOld[pos] = r;
        oldToNew[r] = pos++;
      }
  }
  tileStart[T] = pos;
  if (pos != n) return -1;
  int64_t h = 0;
  for (int32_t t = 0; t < T; t++) {
    haloStart[t] = h;
    for (int64_t g : halo[t]) haloGlobal[h++] = g;
  }
  haloStart[T] = h;

  // local addressing
  int64_t c = 0;
  lrowPtr[0] = 0;
  for (int32_t t = 0; t < T; t++) {
    int64_t nt = tileStart[t + 1] - tileStart[t];
    std::map<int64_t, int64_t> haloLocal;
    for (int64_t q = haloStart[t]; q < haloStart[t + 1]; q++)
      haloLocal[haloGlobal[q]] = nt + (q - haloStart[t]);
    for (int64_t lr = 0; lr < nt; lr++) {
      int64_t p = tileStart[t] + lr;
      int64_t g = newToOld[p];
      ldiag[p] = diag[g];
      struct E {
        int64_t local, pos;
        float v;
      };
      std::vector<E> es;
      for (int64_t a = rowPtr[g]; a < rowPtr[g + 1]; a++) {
        int64_t gc = colInd[a];
        int64_t local;
        if (rowToTile[gc] == t) {
          local = oldToNew[gc] - tileStart[t];
        } else {
          auto it = haloLocal.find(gc);
          if (it == haloLocal.end()) return -1;
          local = it->second;
        }
        es.push_back(E{local, oldToNew[gc], offDiag[a]});
      }
      std::sort(es.begin(), es.end(),
                [](const E& x, const E& y) { return x.local < y.local; });
      for (auto& e : es) {
        lcolLocal[c] = (int32_t)e.local;
        lcol[c] = e.pos;
        loff[c] = e.v;
        c++;
      }
      lrowPtr[p + 1] = c;
    }
  }
  return h;
}

// ---------------------------------------------------------------------------
// Level sets ("colors"). calculateRowColors: CRSHostMatrix.cpp:527-614 and
// calculateColorAddressings: CRSHostMatrix.cpp:616-675.
// level(i) = 0 if row i has no in-tile lower neighbour, else
// 1 + max level(j) over in-tile j < i (the layered Kahn sweep of the
// reference assigns exactly the longest-path depth).
// colorSortAddr: local row ids grouped by level, ascending inside a level.
// colorStart has numColors[t]+1 entries per tile, concatenated (cap n+T).
// Returns the total number of colour entries written to colorStart.
// ---------------------------------------------------------------------------
ORC_API int64_t orc_levels(int64_t n, int32_t T, const int64_t* tileStart,
                   const int64_t* rowPtr, const int64_t* col, int32_t* level,
                   int32_t* colorSortAddr, int64_t* colorStart,
                   int32_t* numColors) {
  int64_t cs = 0;
  for (int32_t t = 0; t < T; t++) {
    int64_t t0 = tileStart[t], t1 = tileStart[t + 1];
    int32_t maxL = -1;
    for (int64_t p = t0; p < t1; p++) {
      int32_t L = 0;
      for (int64_t a = rowPtr[p]; a < rowPtr[p + 1]; a++) {
        int64_t j = col[a];
        if (j < t0 || j >= t1) break;  // halo ids sort last
        if (j >= p) break;             // reached the upper triangle
        L = std::max(L, level[j] + 1);
      }
      level[p] = L;
      maxL = std::max(maxL, L);
    }
    numColors[t] = maxL + 1;
    int64_t w = t0;
    for (int32_t L = 0; L <= maxL; L++) {
      colorStart[cs++] = w - t0;
      for (int64_t p = t0; p < t1; p++)
        if (level[p] == L) colorSortAddr[w++] = (int32_t)(p - t0);
    }
    colorStart[cs++] = w - t0;
  }
  return cs;
}

// ---------------------------------------------------------------------------
// SpMV and residual. lg/matrix/details/crs/CRSMatrix.cpp:62-80 (spmv) and
// :107-125 (residual): sum = diag_i*x_i, then += off_a*x_col(a) in stored
// (local column) order; residual = b_i - sum; cast to the destination type.
// ---------------------------------------------------------------------------
template <typename V, typename X, typename A>
static A row_sum(int64_t p, const int64_t* rowPtr, const int64_t* col,
                 const V* off, const V* diag, const X* x) {
  A s = (A)diag[p] * (A)x[p];
  for (int64_t a = rowPtr[p]; a < rowPtr[p + 1]; a++)
    s = s + (A)off[a] * (A)x[col[a]];
  return s;
}

ORC_API void orc_spmv_f32(int64_t n, const int64_t* rowPtr, const int64_t* col,
                  const float* off, const float* diag, const float* x,
                  float* y) {
  for (int64_t p = 0; p < n; p++)
    y[p] = row_sum<float, float, float>(p, rowPtr, col, off, diag, x);
}
ORC_API void orc_spmv_f64(int64_t n, const int64_t* rowPtr, const int64_t* col,
                  const float* off, const float* diag, const double* x,
                  double* y) {
  for (int64_t p = 0; p < n; p++)
    y[p] = row_sum<float, double, double>(p, rowPtr, col, off, diag, x);
}
ORC_API void orc_residual_f32(int64_t n, const int64_t* rowPtr, const int64_t* col,
                      const float* off, const float* diag, const float* x,
                      const float* b, float* r) {
  for (int64_t p = 0; p < n; p++)
    r[p] = b[p] - row_sum<float, float, float>(p, rowPtr, col, off, diag, x);
}
// MPIR residual (iterative-refinement/Solver.cpp:79,107 with
// CRSMatrix.cpp:107-125): accumulate in FP64, b promoted, stored both rounded
// to FP32 (r32) and unrounded (r64, may be null).
ORC_API void orc_residual_mixed(int64_t n, const int64_t* rowPtr, const int64_t* col,
                        const float* off, const float* diag, const double* x,
                        const float* b, float* r32, double* r64) {
  for (int64_t p = 0; p < n; p++) {
    double r = (double)b[p] -
               row_sum<float, double, double>(p, rowPtr, col, off, diag, x);
    if (r32) r32[p] = (float)r;
    if (r64) r64[p] = r;
  }
}

// ---------------------------------------------------------------------------
// Double-word ("twofloat32", two<float>) arithmetic: a value is hi + lo, two
// floats. The reference's extended precision on hardware without FP64
// (iterative-refinement/Solver.cpp:64-110 with the residuals of
// CRSMatrix.cpp:107-125 and LDUMatrix.cpp:130-159 instantiated for
// two<float>); its arithmetic comes from libtwofloat 0.2.0 (conanfile.py:64),
// which the reference checkout does not contain, so this restates the
// published algorithms instead (Joldes, Muller, Popescu, ACM TOMS 44(2),
// 2017; u = 2^-24):
//   sum:     2Sum(xh, yh) -> (sh, sl); 2Sum(xl, yl) -> (th, tl); c = sl + th;
//            Fast2Sum(sh, c) -> (vh, vl); w = tl + vl; Fast2Sum(vh, w)
//            (relative error <= 3u^2 + 13u^3)
//   product: ch = xh*yh, cl1 = fma(xh, yh, -ch); t0 = xl*yl;
//            t1 = fma(xh, yl, t0); cl2 = fma(xl, yh, t1); cl3 = cl1 + cl2;
//            Fast2Sum(ch, cl3)  (relative error <= 5u^2)
//   difference: sum with the negated operand; float -> two: (f, 0);
//   two -> float: hi + lo.
// Which variants libtwofloat's default operators use is not known here
// (parity unpinned against libtwofloat; the tests pin these against exact
// rational arithmetic).
// ---------------------------------------------------------------------------
struct Two {
  float h = 0.0f, l = 0.0f;
  Two() = default;
  Two(float v) : h(v), l(0.0f) {}
  Two(float a, float b) : h(a), l(b) {}
  float eval() const { return h + l; }
  double dbl() const { return (double)h + (double)l; }
};
static_assert(sizeof(Two) == 8, "two<float> is a float pair");
static void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  const float a1 = s - b;
  const float b1 = s - a1;
  e = (a - a1) + (b - b1);
}
static void fast_two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  e = b - (s - a);
}
static Two operator+(const Two& x, const Two& y) {
  float sh, sl, th, tl, vh, vl, zh, zl;
  two_sum(x.h, y.h, sh, sl);
  two_sum(x.l, y.l, th, tl);
  const float c = sl + th;
  fast_two_sum(sh, c, vh, vl);
  const float w = tl + vl;
  fast_two_sum(vh, w, zh, zl);
  return Two(zh, zl);
}
static Two operator-(const Two& x, const Two& y) { return x + Two(-y.h, -y.l); }
static Two operator*(const Two& x, const Two& y) {
  const float ch = x.h * y.h;
  const float cl1 = std::fma(x.h, y.h, -ch);
  const float t0 = x.l * y.l;
  const float t1 = std::fma(x.h, y.l, t0);
  const float cl2 = std::fma(x.l, y.h, t1);
  const float cl3 = cl1 + cl2;
  float zh, zl;
  fast_two_sum(ch, cl3, zh, zl);
  return Two(zh, zl);
}
// primitives exposed for the exact-arithmetic pins
ORC_API void orc_dw_two_sum(float a, float b, float* out) { two_sum(a, b, out[0], out[1]); }
ORC_API void orc_dw_two_prod(float a, float b, float* out) {
  out[0] = a * b;
  out[1] = std::fma(a, b, -out[0]);
}
ORC_API void orc_dw_add(int64_t n, const float* x, const float* y, float* z) {
  for (int64_t i = 0; i < n; i++) {
    const Two r = Two(x[2 * i], x[2 * i + 1]) + Two(y[2 * i], y[2 * i + 1]);
    z[2 * i] = r.h;
    z[2 * i + 1] = r.l;
  }
}
ORC_API void orc_dw_mul(int64_t n, const float* x, const float* y, float* z) {
  for (int64_t i = 0; i < n; i++) {
    const Two r = Two(x[2 * i], x[2 * i + 1]) * Two(y[2 * i], y[2 * i + 1]);
    z[2 * i] = r.h;
    z[2 * i + 1] = r.l;
  }
}
// r = fl32(two(b) - (two(diag) * x_i + sum two(a_ij) * x_j)), x as (hi, lo)
// pairs; r64 (optional) = the double-word residual rounded to float64.
ORC_API void orc_residual_dw(int64_t n, const int64_t* rowPtr, const int64_t* col,
                             const float* off, const float* diag, const float* xhl,
                             const float* b, float* r32, double* r64) {
  const Two* x = reinterpret_cast<const Two*>(xhl);
  for (int64_t p = 0; p < n; p++) {
    const Two r = Two(b[p]) - row_sum<float, Two, Two>(p, rowPtr, col, off, diag, x);
    if (r32) r32[p] = r.eval();
    if (r64) r64[p] = r.dbl();
  }
}

// ---------------------------------------------------------------------------
// LDU addressing built from a structurally symmetric CRS (no reference
// builder exists; follows lg/matrix/details/ldu/LDUAddressing.hpp:30-63):
// faces enumerate (l, u) with l < u in row-major order of l, then u ascending;
// lowerAddr[f] = l (owner), upperAddr[f] = u (neighbour), upper[f] = A[l][u];
// ownerStart[c] = first face owned by c; losort lists faces by neighbour
// (stable in face order): neighbourStart[c], neighbourColInd[s] = face;
// lower[s] = A[upperAddr[f]][lowerAddr[f]] stored in losort order
// (LDUMatrix.cpp:89 indexes lowerCoeffs by s).
// Returns number of faces, or -1 if the pattern is not symmetric.
// ---------------------------------------------------------------------------
ORC_API int64_t orc_ldu_build(int64_t n, const int64_t* rowPtr, const int64_t* col,
                      const float* off, int64_t* lowerAddr, int64_t* upperAddr,
                      int64_t* ownerStart, int64_t* neighbourStart,
                      int64_t* neighbourColInd, float* upper, float* lower) {
  auto find = [&](int64_t r, int64_t c) -> int64_t {
    for (int64_t a = rowPtr[r]; a < rowPtr[r + 1]; a++)
      if (col[a] == c) return a;
    return -1;
  };
  int64_t F = 0;
  for (int64_t l = 0; l < n; l++) {
    ownerStart[l] = F;
    std::vector<int64_t> us;
    for (int64_t a = rowPtr[l]; a < rowPtr[l + 1]; a++)
      if (col[a] > l) us.push_back(a);
    std::sort(us.begin(), us.end(),
              [&](int64_t x, int64_t y) { return col[x] < col[y]; });
    for (int64_t a : us) {
      lowerAddr[F] = l;
      upperAddr[F] = col[a];
      upper[F] = off[a];
      F++;
    }
  }
  ownerStart[n] = F;
  std::vector<int64_t> cnt(n + 1, 0);
  for (int64_t f = 0; f < F; f++) cnt[upperAddr[f] + 1]++;
  for (int64_t c = 0; c < n; c++) cnt[c + 1] += cnt[c];
  for (int64_t c = 0; c <= n; c++) neighbourStart[c] = cnt[c];
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t f = 0; f < F; f++) {
    int64_t s = fill[upperAddr[f]]++;
    neighbourColInd[s] = f;
    int64_t a = find(upperAddr[f], lowerAddr[f]);
    if (a < 0) return -1;
    lower[s] = off[a];
  }
  // every lower-triangle entry must be covered by a face
  for (int64_t r = 0; r < n; r++)
    for (int64_t a = rowPtr[r]; a < rowPtr[r + 1]; a++)
      if (col[a] < r && find(col[a], r) < 0) return -1;
  return F;
}

// LDU SpMV / residual. lg/matrix/details/ldu/LDUMatrix.cpp:62-99 / :126-165:
// sum = diag_c*x_c; + upper[f]*x[upperAddr[f]] over owned faces ascending;
// + lower[s]*x[lowerAddr[neighbourColInd[s]]] over losort entries ascending.
template <typename X, typename A>
static A ldu_row(int64_t c, const int64_t* lowerAddr, const int64_t* upperAddr,
                 const int64_t* ownerStart, const int64_t* neighbourStart,
                 const int64_t* neighbourColInd, const float* upper,
                 const float* lower, const float* diag, const X* x) {
  A s = (A)diag[c] * (A)x[c];
  for (int64_t f = ownerStart[c]; f < ownerStart[c + 1]; f++)
    s = s + (A)upper[f] * (A)x[upperAddr[f]];
  for (int64_t q = neighbourStart[c]; q < neighbourStart[c + 1]; q++)
    s = s + (A)lower[q] * (A)x[lowerAddr[neighbourColInd[q]]];
  return s;
}
ORC_API void orc_ldu_spmv_f32(int64_t n, const int64_t* lowerAddr,
                      const int64_t* upperAddr, const int64_t* ownerStart,
                      const int64_t* neighbourStart,
                      const int64_t* neighbourColInd, const float* upper,
                      const float* lower, const float* diag, const float* x,
                      float* y) {
  for (int64_t c = 0; c < n; c++)
    y[c] = ldu_row<float, float>(c, lowerAddr, upperAddr, ownerStart,
                                 neighbourStart, neighbourColInd, upper, lower,
                                 diag, x);
}
ORC_API void orc_ldu_residual_mixed(int64_t n, const int64_t* lowerAddr,
                            const int64_t* upperAddr, const int64_t* ownerStart,
                            const int64_t* neighbourStart,
                            const int64_t* neighbourColInd, const float* upper,
                            const float* lower, const float* diag,
                            const double* x, const float* b, float* r32,
                            double* r64) {
  for (int64_t c = 0; c < n; c++) {
    double r = (double)b[c] -
               ldu_row<double, double>(c, lowerAddr, upperAddr, ownerStart,
                                       neighbourStart, neighbourColInd, upper,
                                       lower, diag, x);
    if (r32) r32[c] = (float)r;
    if (r64) r64[c] = r;
  }
}

ORC_API void orc_ldu_residual_dw(int64_t n, const int64_t* lowerAddr,
                                 const int64_t* upperAddr, const int64_t* ownerStart,
                                 const int64_t* neighbourStart,
                                 const int64_t* neighbourColInd, const float* upper,
                                 const float* lower, const float* diag,
                                 const float* xhl, const float* b, float* r32,
                                 double* r64) {
  const Two* x = reinterpret_cast<const Two*>(xhl);
  for (int64_t c = 0; c < n; c++) {
    const Two r = Two(b[c]) - ldu_row<Two, Two>(c, lowerAddr, upperAddr, ownerStart,
                                                neighbourStart, neighbourColInd, upper,
                                                lower, diag, x);
    if (r32) r32[c] = r.eval();
    if (r64) r64[c] = r.dbl();
  }
}

// ---------------------------------------------------------------------------
// ILU(0) factorization per tile, in place. lg/matrix/solver/ilu/Solver.cpp:134-292
// (row body :167-218, sequential driver :271-287, diagonal inversion :290-291;
// intended semantics also in codelet lg/codelet/ipu/matrix/solvers/ilu/ilu-factorize.cpp:241-315).
// Rows 1..n_t-1 of each tile in local order:
//   for ik in row i with k = col(ik) < i:
//     a_ik /= a_kk
//     for ij > ik with j in-tile: a_ij -= a_ik * a(k, j)     (a(k,j) = 0 if absent)
//     a_ii -= a_ik * a(k, i)
// then diag := 1 / diag.
// DILU (diagonalBased, :191-216): off-diagonals stay as in A, only
//     a_ii -= (a_ik * a(k, i)) / a_kk
// with a_kk the already updated diagonal of row k.
// ---------------------------------------------------------------------------
template <typename F, bool DILU = false>
static void ilu0_factor(int64_t n, int32_t T, const int64_t* tileStart,
                        const int64_t* rowPtr, const int64_t* col, F* off,
                        F* diag) {
  auto value = [&](int64_t k, int64_t j) -> F {
    F v = 0;
    for (int64_t a = rowPtr[k]; a < rowPtr[k + 1]; a++)
      if (col[a] == j) v = off[a];
    return v;
  };
  // tiles are independent blocks: one thread per tile (same results)
#pragma omp parallel for schedule(dynamic, 4)
  for (int32_t t = 0; t < T; t++) {
    int64_t t0 = tileStart[t], t1 = tileStart[t + 1];
    for (int64_t i = t0 + 1; i < t1; i++) {
      int64_t iEnd = rowPtr[i + 1];
      for (int64_t ik = rowPtr[i]; ik < iEnd; ik++) {
        int64_t k = col[ik];
        if (k >= i || k < t0) break;  // k >= i; halo columns sort last
        F a_kk = diag[k];
        if (DILU) {
          F a_ki = value(k, i);
          diag[i] = diag[i] - off[ik] * a_ki / a_kk;
          continue;
        }
        off[ik] = off[ik] / a_kk;
        F a_ik = off[ik];
        for (int64_t ij = ik + 1; ij < iEnd; ij++) {
          int64_t j = col[ij];
          if (j < t0 || j >= t1) break;  // stop at halo coefficients
          F a_kj = value(k, j);
          off[ij] = off[ij] - a_ik * a_kj;
        }
        F a_ki = value(k, i);
        diag[i] = diag[i] - a_ik * a_ki;
      }
    }
  }
  for (int64_t p = 0; p < n; p++) diag[p] = (F)1 / diag[p];
}
ORC_API void orc_ilu0_factor_f32(int64_t n, int32_t T, const int64_t* tileStart,
                         const int64_t* rowPtr, const int64_t* col, float* off,
                         float* diag) {
  ilu0_factor<float>(n, T, tileStart, rowPtr, col, off, diag);
}
ORC_API void orc_ilu0_factor_f64(int64_t n, int32_t T, const int64_t* tileStart,
                         const int64_t* rowPtr, const int64_t* col, double* off,
                         double* diag) {
  ilu0_factor<double>(n, T, tileStart, rowPtr, col, off, diag);
}
// DILU factorization: only diag (in place, inverted at the end) changes.
ORC_API void orc_dilu_factor_f32(int64_t n, int32_t T, const int64_t* tileStart,
                                 const int64_t* rowPtr, const int64_t* col,
                                 const float* off, float* diag) {
  ilu0_factor<float, true>(n, T, tileStart, rowPtr, col, const_cast<float*>(off),
                           diag);
}
ORC_API void orc_dilu_factor_f64(int64_t n, int32_t T, const int64_t* tileStart,
                                 const int64_t* rowPtr, const int64_t* col,
                                 const double* off, double* diag) {
  ilu0_factor<double, true>(n, T, tileStart, rowPtr, col, const_cast<double*>(off),
                            diag);
}

// ILU(0) substitution per tile. lg/matrix/solver/ilu/Solver.cpp:380-442 (row
// bodies) and :526-548 (sequential driver); codelet ilu-solve.cpp:30-104.
// forward, i ascending:  t = b_i; for a ascending: j = col; break if j > i or
//                        halo; t -= L_a * x_j;  x_i = t
// backward, i descending: t = 0; for a descending: break if j < i; skip halo;
//                        t -= U_a * x_j;  x_i = (x_i + t) * invD_i
// Arithmetic in the vector type X (factors promoted).
// DILU (diagonalBased, :405-441): lu = A's off-diagonals, forward
// x_i = t * invD_i, backward x_i = x_i + t * invD_i.
template <typename F, typename X, bool DILU = false>
static void ilu0_apply(int32_t T, const int64_t* tileStart,
                       const int64_t* rowPtr, const int64_t* col, const F* lu,
                       const F* invD, const X* b, X* x) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int32_t t = 0; t < T; t++) {
    int64_t t0 = tileStart[t], t1 = tileStart[t + 1];
    for (int64_t i = t0; i < t1; i++) {
      X s = b[i - 0];
      for (int64_t a = rowPtr[i]; a < rowPtr[i + 1]; a++) {
        int64_t j = col[a];
        if (i < j) break;
        if (j < t0 || j >= t1) break;
        s = s - (X)lu[a] * x[j];
      }
      x[i] = DILU ? s * (X)invD[i] : s;
    }
    for (int64_t i = t1 - 1; i >= t0; i--) {
      X s = 0;
      for (int64_t a = rowPtr[i + 1] - 1; a >= rowPtr[i]; a--) {
        int64_t j = col[a];
        // halo ids sort last in local order and are skipped; the first
        // in-tile column below i ends the upper part
        if (j < t0 || j >= t1) continue;
        if (j < i) break;
        s = s - (X)lu[a] * x[j];
      }
      x[i] = DILU ? x[i] + s * (X)invD[i] : (x[i] + s) * (X)invD[i];
    }
  }
}
ORC_API void orc_ilu0_apply_f32(int32_t T, const int64_t* tileStart,
                        const int64_t* rowPtr, const int64_t* col,
                        const float* lu, const float* invD, const float* b,
                        float* x) {
  ilu0_apply<float, float>(T, tileStart, rowPtr, col, lu, invD, b, x);
}
ORC_API void orc_ilu0_apply_f64(int32_t T, const int64_t* tileStart,
                        const int64_t* rowPtr, const int64_t* col,
                        const double* lu, const double* invD, const double* b,
                        double* x) {
  ilu0_apply<double, double>(T, tileStart, rowPtr, col, lu, invD, b, x);
}
ORC_API void orc_dilu_apply_f32(int32_t T, const int64_t* tileStart,
                                const int64_t* rowPtr, const int64_t* col,
                                const float* off, const float* invD,
                                const float* b, float* x) {
  ilu0_apply<float, float, true>(T, tileStart, rowPtr, col, off, invD, b, x);
}
ORC_API void orc_dilu_apply_f64(int32_t T, const int64_t* tileStart,
                                const int64_t* rowPtr, const int64_t* col,
                                const double* off, const double* invD,
                                const double* b, double* x) {
  ilu0_apply<double, double, true>(T, tileStart, rowPtr, col, off, invD, b, x);
}

// ---------------------------------------------------------------------------
// Gauss-Seidel sweep per tile, sequential in local order.
// Structure: lg/matrix/solver/gauss-seidel/Solver.cpp:45-86 (halo exchanged
// before the sweep, one sequential pass per tile over all stored columns,
// halo included). Arithmetic: codelet lg/codelet/ipu/matrix/solvers/gauss-seidel.cpp:41-61
// (t = b_i; t -= off*x_col; x_i = t / diag_i). The reference's CodeDSL kernel
// seeds t with diag_i and ignores b; this restates the codelet.
// backward != 0 runs the rows in reverse local order (symmetric GS; no
// reference counterpart). Halo values are the values before the sweep.
// ---------------------------------------------------------------------------
template <typename X>
static void gs_sweep(int64_t n, int32_t T, const int64_t* tileStart,
                     const int64_t* rowPtr, const int64_t* col,
                     const float* off, const float* diag, const X* b, X* x,
                     int backward) {
  std::vector<X> old(x, x + n);
  for (int32_t t = 0; t < T; t++) {
    int64_t t0 = tileStart[t], t1 = tileStart[t + 1];
    for (int64_t q = 0; q < t1 - t0; q++) {
      int64_t i = backward ? t1 - 1 - q : t0 + q;
      X s = b[i];
      for (int64_t a = rowPtr[i]; a < rowPtr[i + 1]; a++) {
        int64_t j = col[a];
        X xj = (j >= t0 && j < t1) ? x[j] : old[j];
        s = s - (X)off[a] * xj;
      }
      x[i] = s / (X)diag[i];
    }
  }
}
ORC_API void orc_gs_sweep_f32(int64_t n, int32_t T, const int64_t* tileStart,
                      const int64_t* rowPtr, const int64_t* col,
                      const float* off, const float* diag, const float* b,
                      float* x, int backward) {
  gs_sweep<float>(n, T, tileStart, rowPtr, col, off, diag, b, x, backward);
}
ORC_API void orc_gs_sweep_f64(int64_t n, int32_t T, const int64_t* tileStart,
                      const int64_t* rowPtr, const int64_t* col,
                      const float* off, const float* diag, const double* b,
                      double* x, int backward) {
  gs_sweep<double>(n, T, tileStart, rowPtr, col, off, diag, b, x, backward);
}

// ---------------------------------------------------------------------------
// Reductions. lg/dsl/tensor/Expression.cpp:729-769 (4-stage driver),
// :655-691 (per worker), :597-653 (per tile), :695-727 (grouped), chunking
// :472-485 (6 workers, chunk = floor(n/6/4)*4, worker 5 takes the rest),
// grouping lg/common/Shape.hpp:336-351. Accumulation in the operand type.
// ---------------------------------------------------------------------------
template <typename R>
static R reduce_seq(const R* v, int64_t m, bool workers) {
  if (!workers) {
    R s = 0;
    for (int64_t i = 0; i < m; i++) s = s + v[i];
    return s;
  }
  int64_t chunk = ((m / 6) / 4) * 4;
  R part[6];
  for (int w = 0; w < 6; w++) {
    int64_t s0 = chunk * w, s1 = (w == 5) ? m : s0 + chunk;
    R s = 0;
    for (int64_t i = s0; i < s1; i++) s = s + v[i];
    part[w] = s;
  }
  R s = 0;
  for (int w = 0; w < 6; w++) s = s + part[w];
  return s;
}

template <typename R>
static R reduce_add(const R* v, int32_t T, const int64_t* tileStart,
                    int square) {
  int64_t N = tileStart[T] - tileStart[0];
  std::vector<R> sq;
  const R* src = v;
  if (square) {
    sq.resize(N);
    for (int64_t i = 0; i < N; i++) sq[i] = v[i] * v[i];
    src = sq.data();
  }
  // stages 1+2: per tile
  std::vector<R> tileVal(T);
  for (int32_t t = 0; t < T; t++) {
    int64_t m = tileStart[t + 1] - tileStart[t];
    const R* tv = src + (tileStart[t] - tileStart[0]);
    if (N > 6) {
      tileVal[t] = reduce_seq(tv, m, true);
    } else {
      tileVal[t] = reduce_seq(tv, m, false);
    }
  }
  // grouped stages over tile ids; partials live on group leader tiles
  std::vector<int64_t> owner(T);
  for (int32_t t = 0; t < T; t++) owner[t] = t;
  std::vector<R> vals = tileVal;
  const int64_t groups[4] = {368, 736, 1472, 999999999};
  for (int g = 0; g < 4; g++) {
    if (vals.size() == 1) break;
    int64_t G = groups[g];
    std::vector<R> nv;
    std::vector<int64_t> no;
    size_t i = 0;
    while (i < vals.size()) {
      int64_t leader = (owner[i] / G) * G;
      std::vector<R> grp;
      while (i < vals.size() && (owner[i] / G) * G == leader) {
        grp.push_back(vals[i]);
        i++;
      }
      nv.push_back(reduce_seq(grp.data(), (int64_t)grp.size(),
                              grp.size() > 6));
      no.push_back(leader);
    }
    vals.swap(nv);
    owner.swap(no);
  }
  return vals[0];
}
ORC_API float orc_reduce_add_f32(const float* v, int32_t T, const int64_t* tileStart,
                         int square) {
  return reduce_add<float>(v, T, tileStart, square);
}
ORC_API double orc_reduce_add_f64(const double* v, int32_t T,
                          const int64_t* tileStart, int square) {
  return reduce_add<double>(v, T, tileStart, square);
}

// ---------------------------------------------------------------------------
// Solvers on the tile-major layout.
// ---------------------------------------------------------------------------
struct OrcSystem {
  int64_t n;
  int32_t T;
  std::vector<int64_t> tileStart, rowPtr, col;
  std::vector<float> off, diag;
  // lazily built ILU factors
  bool hasF32 = false, hasF64 = false;
  std::vector<float> lu32, invD32;
  std::vector<double> lu64, invD64;
  // lazily built DILU diagonals (off-diagonals are A's)
  std::vector<float> diluInvD32;
  std::vector<double> dilu64, diluInvD64;
};

ORC_API void* orc_system_create(int64_t n, int32_t T, const int64_t* tileStart,
                        const int64_t* rowPtr, const int64_t* col,
                        const float* off, const float* diag) {
  OrcSystem* s = new OrcSystem;
  s->n = n;
  s->T = T;
  s->tileStart.assign(tileStart, tileStart + T + 1);
  s->rowPtr.assign(rowPtr, rowPtr + n + 1);
  int64_t nnz = rowPtr[n];
  s->col.assign(col, col + nnz);
  s->off.assign(off, off + nnz);
  s->diag.assign(diag, diag + n);
  return s;
}
ORC_API void orc_system_destroy(void* h) { delete (OrcSystem*)h; }

// Solver configuration: mirrors the fields of
// lg/matrix/solver/pbicgstab/Configuration.hpp and gauss-seidel/Configuration.hpp.
struct OrcPBiCGStabCfg {
  float absTolerance, relTolerance, relResidual;
  int32_t maxIterations, minIterations;
  int32_t working64;      // 1: working type float64
  int32_t precond;        // 0 none, 1 ILU0, 2 Gauss-Seidel, 3 DILU
  int32_t factor64;       // ILU factors computed/stored in float64
  int32_t gsSweeps;       // numFixedIterations of the GS preconditioner
  int32_t gsSymmetric;    // forward + backward per sweep
  int32_t bNormFromX;     // reference quirk (pbicgstab/Solver.cpp:69-70)
  int32_t norm;           // 0 L2, 1 L1 (sum of magnitudes), 2 LINF (largest magnitude)
};

// Vector norm of kind `norm` in the element type R (Tensor::norm,
// lg/dsl/tensor/Tensor.cpp:425-439): L2 squares in the reference's
// reduction order and takes the root; L1 sums magnitudes in the same order
// (the reference sums signed values, Tensor.cpp:430-431; the textbook norm
// is the build's documented choice); LINF is the largest magnitude (exact,
// order-free).
template <typename R>
static R vec_norm(const R* u, int64_t n, int32_t T, const int64_t* tsp, int32_t norm) {
  if (norm == 2) {
    R m = 0;
    for (int64_t i = 0; i < n; i++) m = std::max(m, (R)std::fabs(u[i]));
    return m;
  }
  if (norm == 1) {
    std::vector<R> a(n);
    for (int64_t i = 0; i < n; i++) a[i] = (R)std::fabs(u[i]);
    return reduce_add<R>(a.data(), T, tsp, 0);
  }
  return (R)std::sqrt(reduce_add<R>(u, T, tsp, 1));
}
struct OrcStats {
  double initialResidual, finalResidual;
  int32_t iterations, converged, singular, pad;
};

static void ensure_factors(OrcSystem* s, bool f64) {
  int64_t nnz = s->rowPtr[s->n];
  if (f64 && !s->hasF64) {
    s->lu64.assign(s->off.begin(), s->off.end());
    s->invD64.assign(s->diag.begin(), s->diag.end());
    ilu0_factor<double>(s->n, s->T, s->tileStart.data(), s->rowPtr.data(),
                        s->col.data(), s->lu64.data(), s->invD64.data());
    s->hasF64 = true;
  }
  if (!f64 && !s->hasF32) {
    s->lu32.assign(s->off.begin(), s->off.end());
    s->invD32.assign(s->diag.begin(), s->diag.end());
    ilu0_factor<float>(s->n, s->T, s->tileStart.data(), s->rowPtr.data(),
                       s->col.data(), s->lu32.data(), s->invD32.data());
    s->hasF32 = true;
  }
  (void)nnz;
}

static void ensure_dilu(OrcSystem* s, bool f64) {
  if (f64 && s->diluInvD64.empty()) {
    s->dilu64.assign(s->off.begin(), s->off.end());
    s->diluInvD64.assign(s->diag.begin(), s->diag.end());
    ilu0_factor<double, true>(s->n, s->T, s->tileStart.data(), s->rowPtr.data(),
                              s->col.data(), s->dilu64.data(), s->diluInvD64.data());
  }
  if (!f64 && s->diluInvD32.empty()) {
    s->diluInvD32.assign(s->diag.begin(), s->diag.end());
    ilu0_factor<float, true>(s->n, s->T, s->tileStart.data(), s->rowPtr.data(),
                             s->col.data(), s->off.data(), s->diluInvD32.data());
  }
}

template <typename X>
static void precond_apply(OrcSystem* s, const OrcPBiCGStabCfg* c, const X* b,
                          X* y) {
  int64_t n = s->n;
  if (c->precond == 0) {
    for (int64_t i = 0; i < n; i++) y[i] = b[i];
  } else if (c->precond == 1) {
    if (c->factor64) {
      ensure_factors(s, true);
      ilu0_apply<double, X>(s->T, s->tileStart.data(), s->rowPtr.data(),
                            s->col.data(), s->lu64.data(), s->invD64.data(), b,
                            y);
    } else {
      ensure_factors(s, false);
      ilu0_apply<float, X>(s->T, s->tileStart.data(), s->rowPtr.data(),
                           s->col.data(), s->lu32.data(), s->invD32.data(), b,
                           y);
    }
  } else if (c->precond == 3) {
    if (c->factor64) {
      ensure_dilu(s, true);
      ilu0_apply<double, X, true>(s->T, s->tileStart.data(), s->rowPtr.data(),
                                  s->col.data(), s->dilu64.data(),
                                  s->diluInvD64.data(), b, y);
    } else {
      ensure_dilu(s, false);
      ilu0_apply<float, X, true>(s->T, s->tileStart.data(), s->rowPtr.data(),
                                 s->col.data(), s->off.data(),
                                 s->diluInvD32.data(), b, y);
    }
  } else {
    // Gauss-Seidel uses the initial guess: y = 0 (pbicgstab/Solver.cpp:107-109)
    for (int64_t i = 0; i < n; i++) y[i] = 0;
    for (int k = 0; k < c->gsSweeps; k++) {
      gs_sweep<X>(n, s->T, s->tileStart.data(), s->rowPtr.data(),
                  s->col.data(), s->off.data(), s->diag.data(), b, y, 0);
      if (c->gsSymmetric)
        gs_sweep<X>(n, s->T, s->tileStart.data(), s->rowPtr.data(),
                    s->col.data(), s->off.data(), s->diag.data(), b, y, 1);
    }
  }
}

// SolverStats::checkConvergence, lg/matrix/solver/SolverStats.cpp:33-56.
static bool converged_test(double rF, double r0, double bNorm,
                           const float absT, float relT, float relR) {
  bool c = rF < (double)absT || rF < r0 * (double)relR;
  if (relT > 0) c = c || rF < bNorm * (double)relT;
  return c;
}

// PBiCGStab. lg/matrix/solver/pbicgstab/Solver.cpp:48-194, exact operation
// sequence; singularity threshold Traits<float>::vsmall() = 1e-37
// (lg/common/Traits.hpp:38-46, SolverStats.cpp:58-63).
template <typename X>
static void pbicgstab(OrcSystem* s, const OrcPBiCGStabCfg* c, X* x,
                      const X* b, OrcStats* st) {
  int64_t n = s->n;
  const int64_t* tsp = s->tileStart.data();
  const int32_t T = s->T;
  // element loops run on all threads (each element computed as in a serial
  // loop); the reductions stay serial in the reference's order
  auto spmv = [&](const X* in, X* out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++)
      out[p] = row_sum<float, X, X>(p, s->rowPtr.data(), s->col.data(),
                                    s->off.data(), s->diag.data(), in);
  };
  auto dot = [&](const X* u, const X* v) {
    std::vector<X> t(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) t[i] = u[i] * v[i];
    return reduce_add<X>(t.data(), T, tsp, 0);
  };
  auto norm = [&](const X* u) { return vec_norm<X>(u, n, T, tsp, c->norm); };
  std::vector<X> r(n), r0(n), p(n), v(n), y(n), z(n), sv(n), t(n);
  for (int64_t i = 0; i < n; i++)
    r[i] = b[i] - row_sum<float, X, X>(i, s->rowPtr.data(), s->col.data(),
                                       s->off.data(), s->diag.data(), x);
  double bNorm = 0;
  if (c->relTolerance > 0) bNorm = c->bNormFromX ? norm(x) : norm(b);
  double r0n = norm(r.data());
  st->initialResidual = r0n;
  st->finalResidual = r0n;
  st->iterations = 0;
  st->converged = 0;
  st->singular = 0;
  r0 = r;
  p = r;
  X rhoOld = 0, alpha = 0, omega = 0;
  const double vsmall = 1e-37;
  while (!((st->converged && st->iterations >= c->minIterations) ||
           st->iterations >= c->maxIterations || st->singular)) {
    X rho = dot(r0.data(), r.data());
    st->singular = st->singular || std::fabs((double)rho) < vsmall;
    if (st->singular) break;
    if (st->iterations != 0) {
      X beta = (rho / rhoOld) * (alpha / omega);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++)
        p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
    precond_apply<X>(s, c, p.data(), y.data());
    spmv(y.data(), v.data());
    alpha = rho / dot(r0.data(), v.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) sv[i] = r[i] - alpha * v[i];
    precond_apply<X>(s, c, sv.data(), z.data());
    spmv(z.data(), t.data());
    X tt = dot(t.data(), t.data());
    omega = dot(t.data(), sv.data()) / tt;
    st->singular = st->singular || std::fabs((double)omega) < vsmall;
    if (st->singular) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      x[i] = (x[i] + alpha * y[i]) + omega * z[i];
      r[i] = sv[i] - omega * t[i];
    }
    st->finalResidual = norm(r.data());
    st->converged = converged_test(st->finalResidual, st->initialResidual,
                                   bNorm, c->absTolerance, c->relTolerance,
                                   c->relResidual);
    rhoOld = rho;
    st->iterations++;
  }
}

ORC_API float orc_vec_norm_f32(const float* u, int64_t n, int32_t T, const int64_t* tsp,
                               int32_t norm) {
  return vec_norm<float>(u, n, T, tsp, norm);
}
ORC_API double orc_vec_norm_f64(const double* u, int64_t n, int32_t T, const int64_t* tsp,
                                int32_t norm) {
  return vec_norm<double>(u, n, T, tsp, norm);
}

// Threads of the parallel loops (the results do not depend on it).
ORC_API void orc_set_threads(int32_t n) { omp_set_num_threads(n > 0 ? n : 1); }
ORC_API int32_t orc_max_threads(void) { return omp_get_max_threads(); }

ORC_API int32_t orc_pbicgstab_f32(void* h, const OrcPBiCGStabCfg* c, float* x,
                          const float* b, OrcStats* st) {
  pbicgstab<float>((OrcSystem*)h, c, x, b, st);
  return 0;
}
ORC_API int32_t orc_pbicgstab_f64(void* h, const OrcPBiCGStabCfg* c, double* x,
                          const double* b, OrcStats* st) {
  pbicgstab<double>((OrcSystem*)h, c, x, b, st);
  return 0;
}

// Mixed-precision iterative refinement. lg/matrix/solver/iterative-refinement/Solver.cpp:48-123.
// phi (FP64) is the solution; working precision of the inner solver is
// float32 (inner64 = 0) or float64. normFromRounded = 1 takes the norm of the
// FP32-stored residual as the reference does (:80-81, :110); 0 takes the FP64
// norm of the unrounded residual.
struct OrcMPIRCfg {
  float absTolerance, relTolerance, relResidual;
  int32_t maxIterations, minIterations;
  int32_t normFromRounded;
  OrcPBiCGStabCfg inner;
  int32_t norm;  // as OrcPBiCGStabCfg::norm
};
ORC_API int32_t orc_mpir(void* h, const OrcMPIRCfg* c, double* phi, const float* b,
                 OrcStats* st, int32_t* innerIterations) {
  OrcSystem* s = (OrcSystem*)h;
  int64_t n = s->n;
  const int32_t T = s->T;
  const int64_t* tsp = s->tileStart.data();
  std::vector<float> r32(n);
  std::vector<double> r64(n);
  auto resid = [&]() {
    orc_residual_mixed(n, s->rowPtr.data(), s->col.data(), s->off.data(),
                       s->diag.data(), phi, b, r32.data(), r64.data());
    if (c->normFromRounded) return (double)vec_norm<float>(r32.data(), n, T, tsp, c->norm);
    return vec_norm<double>(r64.data(), n, T, tsp, c->norm);
  };
  double bNorm = 0;
  if (c->relTolerance > 0) {
    std::vector<double> bd(b, b + n);
    bNorm = vec_norm<double>(bd.data(), n, T, tsp, c->norm);
  }
  st->initialResidual = resid();
  st->finalResidual = st->initialResidual;
  st->iterations = 0;
  st->converged = 0;
  st->singular = 0;
  int32_t inner = 0;
  while (!((st->converged && st->iterations >= c->minIterations) ||
           st->iterations >= c->maxIterations)) {
    st->iterations++;
    OrcStats ist;
    if (c->inner.working64) {
      std::vector<double> d(n, 0.0);
      pbicgstab<double>(s, &c->inner, d.data(), r64.data(), &ist);
      for (int64_t i = 0; i < n; i++) phi[i] = phi[i] + d[i];
    } else {
      std::vector<float> d(n, 0.0f);
      pbicgstab<float>(s, &c->inner, d.data(), r32.data(), &ist);
      for (int64_t i = 0; i < n; i++) phi[i] = phi[i] + (double)d[i];
    }
    inner += ist.iterations;
    st->finalResidual = resid();
    st->converged = converged_test(st->finalResidual, st->initialResidual,
                                   bNorm, c->absTolerance, c->relTolerance,
                                   c->relResidual);
  }
  if (innerIterations) *innerIterations = inner;
  return 0;
}

// MPIR with extendedPrecisionType "twofloat32" (iterative-refinement/
// Solver.cpp:48-123): phi as double words ((hi, lo) float pairs), the
// residual in double-word arithmetic (orc_residual_dw), the update
// phi = phi + two(d). Norm of the double-word residual rounded to float64,
// as the FP64 variant above uses the unrounded residual.
ORC_API int32_t orc_mpir_dw(void* h, const OrcMPIRCfg* c, float* phihl, const float* b,
                            OrcStats* st, int32_t* innerIterations) {
  OrcSystem* s = (OrcSystem*)h;
  int64_t n = s->n;
  const int32_t T = s->T;
  const int64_t* tsp = s->tileStart.data();
  Two* phi = reinterpret_cast<Two*>(phihl);
  std::vector<float> r32(n);
  std::vector<double> r64(n);
  auto resid = [&]() {
    orc_residual_dw(n, s->rowPtr.data(), s->col.data(), s->off.data(), s->diag.data(),
                    phihl, b, r32.data(), r64.data());
    if (c->normFromRounded) return (double)vec_norm<float>(r32.data(), n, T, tsp, c->norm);
    return vec_norm<double>(r64.data(), n, T, tsp, c->norm);
  };
  double bNorm = 0;
  if (c->relTolerance > 0) {
    std::vector<double> bd(b, b + n);
    bNorm = vec_norm<double>(bd.data(), n, T, tsp, c->norm);
  }
  st->initialResidual = resid();
  st->finalResidual = st->initialResidual;
  st->iterations = 0;
  st->converged = 0;
  st->singular = 0;
  int32_t inner = 0;
  while (!((st->converged && st->iterations >= c->minIterations) ||
           st->iterations >= c->maxIterations)) {
    st->iterations++;
    OrcStats ist;
    std::vector<float> d(n, 0.0f);
    pbicgstab<float>(s, &c->inner, d.data(), r32.data(), &ist);
    for (int64_t i = 0; i < n; i++) phi[i] = phi[i] + Two(d[i]);
    inner += ist.iterations;
    st->finalResidual = resid();
    st->converged = converged_test(st->finalResidual, st->initialResidual,
                                   bNorm, c->absTolerance, c->relTolerance,
                                   c->relResidual);
  }
  if (innerIterations) *innerIterations = inner;
  return 0;
}

