// Host-side matrix preparation (see host_matrix.hpp).
#include "host_matrix.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>

#include "gph_common.hpp"
#include "parallel.hpp"

namespace graphene::matrix::host {

using gph::fail;
using gph::kInvalid;
using gph::kRuntime;

// ---------------------------------------------------------------------------
// RNG: splitmix64 seeding of xoshiro256** (public-domain reference algorithms
// by Vigna). Fixed here so that every component draws identical inputs.
// ---------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}
Rng::Rng(uint64_t seed) {
  uint64_t sm = seed;
  for (auto& w : s) w = splitmix64(sm);
}
uint64_t Rng::next() {
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
float Rng::uniformPm1() {
  float u = (float)(next() >> 40) * (1.0f / 16777216.0f);
  return u * 2.0f - 1.0f;
}
double Rng::uniform01d() { return (double)(next() >> 11) * 0x1.0p-53; }

void uniformPm1(uint64_t seed, int64_t n, float* out) {
  Rng r(seed);
  for (int64_t i = 0; i < n; i++) out[i] = r.uniformPm1();
}

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------

// lg/matrix/host/details/Poisson.cpp:26-69. Row k*nx*ny + j*nx + i; diagonal
// 6, -1 to every in-domain face neighbour; push order diag, i-1, i+1, j-1,
// j+1, k-1, k+1.
TripletMatrix generate3DPoissonTripletMatrix(int64_t nx, int64_t ny,
                                             int64_t nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) fail(kInvalid, "Poisson: empty grid");
  TripletMatrix m;
  m.nrows = m.ncols = nx * ny * nz;
  size_t cap = (size_t)(7 * m.nrows);
  m.rows.reserve(cap);
  m.cols.reserve(cap);
  m.vals.reserve(cap);
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t row = k * nx * ny + j * nx + i;
        m.push(row, row, 6.0f);
        if (i > 0) m.push(row, row - 1, -1.0f);
        if (i < nx - 1) m.push(row, row + 1, -1.0f);
        if (j > 0) m.push(row, row - nx, -1.0f);
        if (j < ny - 1) m.push(row, row + nx, -1.0f);
        if (k > 0) m.push(row, row - nx * ny, -1.0f);
        if (k < nz - 1) m.push(row, row + nx * ny, -1.0f);
      }
  return m;
}

// Unit direction from a seed: v ~ U[-1,1)^dim (doubles), redrawn until
// 0.1 <= |v| <= 1, then normalised.
static void unitDirection(Rng& r, int dim, double* u) {
  for (;;) {
    double v[3] = {0, 0, 0}, nrm = 0;
    for (int d = 0; d < dim; d++) {
      v[d] = r.uniform01d() * 2.0 - 1.0;
      nrm += v[d] * v[d];
    }
    nrm = std::sqrt(nrm);
    if (nrm >= 0.1 && nrm <= 1.0) {
      for (int d = 0; d < dim; d++) u[d] = v[d] / nrm;
      return;
    }
  }
}

// 3D convection-diffusion, 27-point stencil, no reference counterpart.
// Diffusion: 27-point Laplacian (diagonal 26, -1 for each of the 26
// neighbours). Convection: first-order upwind along the unit direction u
// drawn from `seed`, cell Peclet number `peclet`: for every axis d the
// upwind face neighbour (-e_d if u_d > 0, +e_d if u_d < 0) gets
// -peclet*|u_d| and the diagonal +peclet*|u_d|. Dirichlet boundaries are
// eliminated: out-of-domain neighbours are dropped, the diagonal keeps its
// full value. Coefficients are formed in FP64 and rounded to FP32. Entries of
// a row are pushed diagonal first, then neighbours in ascending column order.
TripletMatrix generate3DConvectionDiffusion27(int64_t nx, int64_t ny,
                                              int64_t nz, double peclet,
                                              uint64_t seed) {
  if (nx <= 0 || ny <= 0 || nz <= 0) fail(kInvalid, "convdiff27: empty grid");
  Rng rng(seed);
  double u[3];
  unitDirection(rng, 3, u);
  double coef[27];
  double dg = 26.0;
  for (int q = 0; q < 27; q++) coef[q] = -1.0;
  auto idx = [](int di, int dj, int dk) {
    return (dk + 1) * 9 + (dj + 1) * 3 + (di + 1);
  };
  for (int d = 0; d < 3; d++) {
    double a = peclet * std::fabs(u[d]);
    int s = u[d] > 0 ? -1 : 1;
    int o[3] = {0, 0, 0};
    o[d] = s;
    if (u[d] != 0.0) coef[idx(o[0], o[1], o[2])] -= a;
    dg += a;
  }
  const float diagF = (float)dg;
  float cf[27];
  for (int q = 0; q < 27; q++) cf[q] = (float)coef[q];

  TripletMatrix m;
  m.nrows = m.ncols = nx * ny * nz;
  size_t cap = (size_t)(27 * m.nrows);
  m.rows.reserve(cap);
  m.cols.reserve(cap);
  m.vals.reserve(cap);
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t row = k * nx * ny + j * nx + i;
        m.push(row, row, diagF);
        for (int dk = -1; dk <= 1; dk++) {
          int64_t kk = k + dk;
          if (kk < 0 || kk >= nz) continue;
          for (int dj = -1; dj <= 1; dj++) {
            int64_t jj = j + dj;
            if (jj < 0 || jj >= ny) continue;
            for (int di = -1; di <= 1; di++) {
              int64_t ii = i + di;
              if (ii < 0 || ii >= nx) continue;
              if (di == 0 && dj == 0 && dk == 0) continue;
              m.push(row, kk * nx * ny + jj * nx + ii, cf[idx(di, dj, dk)]);
            }
          }
        }
      }
  return m;
}

// 2D convection-diffusion, 9-point stencil, no reference counterpart.
// diffusion*(9-point Laplacian: diagonal 8, -1 for the 8 neighbours) plus
// first-order upwind convection with cell Peclet `peclet` along a seeded unit
// direction (coefficient peclet*diffusion*|u_d| on the upwind face
// neighbour and the diagonal). Dirichlet boundaries eliminated.
TripletMatrix generate2DConvectionDiffusion9(int64_t nx, int64_t ny,
                                             double diffusion, double peclet,
                                             uint64_t seed) {
  if (nx <= 0 || ny <= 0) fail(kInvalid, "convdiff9: empty grid");
  Rng rng(seed);
  double u[3];
  unitDirection(rng, 2, u);
  double coef[9];
  double dg = 8.0 * diffusion;
  for (int q = 0; q < 9; q++) coef[q] = -diffusion;
  auto idx = [](int di, int dj) { return (dj + 1) * 3 + (di + 1); };
  for (int d = 0; d < 2; d++) {
    double a = peclet * diffusion * std::fabs(u[d]);
    int s = u[d] > 0 ? -1 : 1;
    int o[2] = {0, 0};
    o[d] = s;
    if (u[d] != 0.0) coef[idx(o[0], o[1])] -= a;
    dg += a;
  }
  const float diagF = (float)dg;
  float cf[9];
  for (int q = 0; q < 9; q++) cf[q] = (float)coef[q];
  TripletMatrix m;
  m.nrows = m.ncols = nx * ny;
  for (int64_t j = 0; j < ny; ++j)
    for (int64_t i = 0; i < nx; ++i) {
      int64_t row = j * nx + i;
      m.push(row, row, diagF);
      for (int dj = -1; dj <= 1; dj++) {
        int64_t jj = j + dj;
        if (jj < 0 || jj >= ny) continue;
        for (int di = -1; di <= 1; di++) {
          int64_t ii = i + di;
          if (ii < 0 || ii >= nx) continue;
          if (di == 0 && dj == 0) continue;
          m.push(row, jj * nx + ii, cf[idx(di, dj)]);
        }
      }
    }
  return m;
}

// Unstructured power-law matrix, no reference counterpart.
// Row length L (stored entries including the diagonal) drawn from the
// discrete truncated power law P(L) ~ L^-exponent on [minLen, maxLen]; the
// L-1 off-diagonal columns are distinct and != row: with probability
// localFraction uniform in [row-band, row+band] (redrawn when outside the
// matrix), otherwise uniform in [0, n). Off-diagonal values -U[0,1),
// diagonal 1.1 * sum|off| (FP64, rounded to FP32). Rows are drawn in order
// from one RNG stream.
TripletMatrix generatePowerLaw(int64_t n, int32_t minLen, int32_t maxLen,
                               double exponent, int32_t band,
                               double localFraction, uint64_t seed) {
  if (n <= 1 || minLen < 1 || maxLen < minLen || band < 1)
    fail(kInvalid, "power-law generator: invalid parameters");
  Rng rng(seed);
  std::vector<double> cdf;
  double tot = 0;
  for (int32_t L = minLen; L <= maxLen; L++) {
    tot += std::pow((double)L, -exponent);
    cdf.push_back(tot);
  }
  TripletMatrix m;
  m.nrows = m.ncols = n;
  std::vector<int64_t> cols;
  for (int64_t r = 0; r < n; r++) {
    double uu = rng.uniform01d() * tot;
    int32_t L = maxLen;
    for (size_t q = 0; q < cdf.size(); q++)
      if (uu < cdf[q]) {
        L = minLen + (int32_t)q;
        break;
      }
    int32_t want = std::min<int64_t>(L - 1, n - 1);
    cols.clear();
    double asum = 0;
    std::vector<float> vals;
    while ((int32_t)cols.size() < want) {
      int64_t c;
      if (rng.uniform01d() < localFraction) {
        int64_t off = (int64_t)(rng.next() % (uint64_t)(2 * band + 1)) - band;
        c = r + off;
        if (c < 0 || c >= n) continue;
      } else {
        c = (int64_t)(rng.next() % (uint64_t)n);
      }
      if (c == r) continue;
      if (std::find(cols.begin(), cols.end(), c) != cols.end()) continue;
      float v = -(float)rng.uniform01d();
      cols.push_back(c);
      vals.push_back(v);
      asum += std::fabs((double)v);
    }
    m.push(r, r, (float)(1.1 * asum));
    for (size_t q = 0; q < cols.size(); q++) m.push(r, cols[q], vals[q]);
  }
  return m;
}

// Matrix Market coordinate reader (real / integer / pattern; general,
// symmetric, skew-symmetric), 1-based indices, symmetric storage mirrored.
TripletMatrix loadTripletMatrixFromFile(const std::string& path) {
  std::ifstream in(path);
  if (!in.is_open()) fail(kInvalid, "Failed to open file " + path);
  std::string line;
  if (!std::getline(in, line)) fail(kInvalid, "empty Matrix Market file");
  std::string lower = line;
  for (auto& ch : lower) ch = (char)std::tolower((unsigned char)ch);
  if (lower.rfind("%%matrixmarket", 0) != 0)
    fail(kInvalid, "missing %%MatrixMarket header");
  if (lower.find("coordinate") == std::string::npos)
    fail(gph::kUnsupported, "only coordinate Matrix Market files");
  bool pattern = lower.find("pattern") != std::string::npos;
  bool symmetric = lower.find(" symmetric") != std::string::npos ||
                   lower.find("hermitian") != std::string::npos;
  bool skew = lower.find("skew-symmetric") != std::string::npos;
  if (lower.find("complex") != std::string::npos)
    fail(kInvalid, "complex Matrix Market files are not supported");
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] != '%') break;
  }
  std::istringstream hs(line);
  int64_t nr, nc, ne;
  if (!(hs >> nr >> nc >> ne)) fail(kInvalid, "bad Matrix Market size line");
  TripletMatrix m;
  m.nrows = nr;
  m.ncols = nc;
  m.rows.reserve(ne);
  m.cols.reserve(ne);
  m.vals.reserve(ne);
  for (int64_t e = 0; e < ne; e++) {
    int64_t r, c;
    double v = 1.0;
    if (!(in >> r >> c)) fail(kInvalid, "truncated Matrix Market file");
    if (!pattern && !(in >> v)) fail(kInvalid, "truncated Matrix Market file");
    r -= 1;
    c -= 1;
    if (r < 0 || r >= nr || c < 0 || c >= nc)
      fail(kInvalid, "Matrix Market index out of range");
    m.push(r, c, (float)v);
    if ((symmetric || skew) && r != c) m.push(c, r, (float)(skew ? -v : v));
  }
  return m;
}

// ---------------------------------------------------------------------------
// Partitioners
// ---------------------------------------------------------------------------
std::vector<int32_t> partitionGrid3D(int64_t nx, int64_t ny, int64_t nz,
                                     int64_t bx, int64_t by, int64_t bz,
                                     int32_t* numTiles) {
  if (bx <= 0 || by <= 0 || bz <= 0) fail(kInvalid, "block size must be > 0");
  int64_t tx = (nx + bx - 1) / bx, ty = (ny + by - 1) / by,
          tz = (nz + bz - 1) / bz;
  std::vector<int32_t> r2t((size_t)(nx * ny * nz));
  for (int64_t k = 0; k < nz; k++)
    for (int64_t j = 0; j < ny; j++)
      for (int64_t i = 0; i < nx; i++)
        r2t[(size_t)(k * nx * ny + j * nx + i)] =
            (int32_t)(((k / bz) * ty + (j / by)) * tx + (i / bx));
  if (numTiles) *numTiles = (int32_t)(tx * ty * tz);
  return r2t;
}

// Contiguous row ranges with balanced weight (row length + 1, the vertex
// weight the reference hands to METIS, CRSHostMatrix.cpp:197).
std::vector<int32_t> partitionContiguous(const std::vector<int64_t>& rowPtr,
                                         int32_t numTiles) {
  int64_t n = (int64_t)rowPtr.size() - 1;
  if (numTiles < 1) fail(kInvalid, "numTiles must be >= 1");
  std::vector<int32_t> r2t((size_t)n);
  double total = (double)(rowPtr[n] + n);
  double acc = 0;
  for (int64_t r = 0; r < n; r++) {
    int32_t t = (int32_t)std::min<double>(numTiles - 1,
                                          std::floor(acc / total * numTiles));
    r2t[(size_t)r] = t;
    acc += (double)(rowPtr[r + 1] - rowPtr[r] + 1);
  }
  return r2t;
}

std::vector<int32_t> partitionGraph(int64_t n, const std::vector<int64_t>& rows,
                                    const std::vector<int64_t>& cols, int32_t numTiles) {
  if (numTiles < 1) fail(kInvalid, "numTiles must be >= 1");
  if (numTiles > n && n > 0) fail(kInvalid, "more tiles than rows");
  if (rows.size() != cols.size()) fail(kInvalid, "rows and cols differ in length");
  // symmetrized adjacency without self loops (CRS, ascending, unique)
  std::vector<int64_t> deg((size_t)n + 1, 0), weight((size_t)n, 1);
  for (size_t k = 0; k < rows.size(); k++) {
    const int64_t r = rows[k], c = cols[k];
    if (r < 0 || r >= n || c < 0 || c >= n) fail(kInvalid, "index out of range");
    weight[(size_t)r]++;
    if (r != c) {
      deg[(size_t)r + 1]++;
      deg[(size_t)c + 1]++;
    }
  }
  for (int64_t i = 0; i < n; i++) deg[(size_t)i + 1] += deg[(size_t)i];
  std::vector<int64_t> adj((size_t)deg[(size_t)n]);
  {
    std::vector<int64_t> pos(deg.begin(), deg.end() - 1);
    for (size_t k = 0; k < rows.size(); k++) {
      const int64_t r = rows[k], c = cols[k];
      if (r == c) continue;
      adj[(size_t)pos[(size_t)r]++] = c;
      adj[(size_t)pos[(size_t)c]++] = r;
    }
  }
  std::vector<int64_t> ptr((size_t)n + 1, 0);
  {
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++) {
      const int64_t a = deg[(size_t)i], b = deg[(size_t)i + 1];
      std::sort(adj.begin() + a, adj.begin() + b);
      ptr[(size_t)i] = w;
      for (int64_t k = a; k < b; k++)
        if (k == a || adj[(size_t)k] != adj[(size_t)k - 1]) adj[(size_t)w++] = adj[(size_t)k];
    }
    ptr[(size_t)n] = w;
    adj.resize((size_t)w);
  }
  // recursive bisection over vertex sets; part id = set label
  std::vector<int32_t> label((size_t)n, 0), r2t((size_t)n, 0);
  std::vector<int64_t> order, queue, dist((size_t)n, -1);
  struct Job {
    std::vector<int64_t> verts;
    int32_t first, parts;
  };
  std::vector<Job> stack;
  {
    Job all;
    all.verts.resize((size_t)n);
    for (int64_t i = 0; i < n; i++) all.verts[(size_t)i] = i;
    all.first = 0;
    all.parts = numTiles;
    stack.push_back(std::move(all));
  }
  int32_t nextLabel = 1;
  // BFS inside the set `lab` from `src`; returns the visit order
  auto bfs = [&](int64_t src, int32_t lab, std::vector<int64_t>& out) {
    out.clear();
    queue.clear();
    queue.push_back(src);
    dist[(size_t)src] = 0;
    for (size_t h = 0; h < queue.size(); h++) {
      const int64_t v = queue[h];
      out.push_back(v);
      for (int64_t k = ptr[(size_t)v]; k < ptr[(size_t)v + 1]; k++) {
        const int64_t u = adj[(size_t)k];
        if (label[(size_t)u] == lab && dist[(size_t)u] < 0) {
          dist[(size_t)u] = dist[(size_t)v] + 1;
          queue.push_back(u);
        }
      }
    }
    for (int64_t v : out) dist[(size_t)v] = -1;
  };
  while (!stack.empty()) {
    Job job = std::move(stack.back());
    stack.pop_back();
    if (job.parts == 1) {
      for (int64_t v : job.verts) r2t[(size_t)v] = job.first;
      continue;
    }
    const int32_t lab = label[(size_t)job.verts.front()];
    int64_t total = 0;
    for (int64_t v : job.verts) total += weight[(size_t)v];
    const int32_t p1 = job.parts / 2;
    const double target = (double)total * p1 / job.parts;
    // pseudo-peripheral start: the last vertex reached from the first
    // vertex, twice
    int64_t src = job.verts.front();
    for (int rep = 0; rep < 2; rep++) {
      bfs(src, lab, order);
      src = order.back();
    }
    // first part: BFS order from src, continued over the set's other
    // components in index order, until the target weight
    const int32_t labA = nextLabel++, labB = nextLabel++;
    int64_t acc = 0;
    Job a, b;
    a.first = job.first;
    a.parts = p1;
    b.first = job.first + p1;
    b.parts = job.parts - p1;
    auto takeA = [&](int64_t v) {
      label[(size_t)v] = labA;
      a.verts.push_back(v);
      acc += weight[(size_t)v];
    };
    const int64_t need = (int64_t)a.parts;  // every part needs a vertex
    bfs(src, lab, order);
    size_t oi = 0, vi = 0;
    auto remainingVerts = [&]() { return (int64_t)job.verts.size() - (int64_t)a.verts.size(); };
    while ((double)acc < target || (int64_t)a.verts.size() < need) {
      if (remainingVerts() <= (int64_t)b.parts) break;  // leave one per part of B
      if (oi < order.size()) {
        const int64_t v = order[oi++];
        if (label[(size_t)v] == lab) takeA(v);
        continue;
      }
      // next component: the smallest unvisited vertex of the set
      while (vi < job.verts.size() && label[(size_t)job.verts[vi]] != lab) vi++;
      if (vi >= job.verts.size()) break;
      bfs(job.verts[vi], lab, order);
      oi = 0;
    }
    for (int64_t v : job.verts)
      if (label[(size_t)v] == lab) {
        label[(size_t)v] = labB;
        b.verts.push_back(v);
      }
    std::sort(a.verts.begin(), a.verts.end());
    stack.push_back(std::move(b));
    stack.push_back(std::move(a));
  }
  return r2t;
}

// ---------------------------------------------------------------------------
// COO -> CRS (CoordinateFormat.cpp:29-40 + CRSHostMatrix.cpp:87-137):
// sorted by (row, col), diagonal split out, duplicates summed.
// ---------------------------------------------------------------------------
struct GlobalCRS {
  std::vector<int64_t> rowPtr;
  std::vector<int32_t> col;
  std::vector<float> off;
  std::vector<float> diag;
};

static GlobalCRS convertToCRS(TripletMatrix&& coo) {
  const int64_t n = coo.nrows;
  if (coo.ncols != n) fail(kInvalid, "matrix must be square");
  if (n <= 0) fail(kInvalid, "matrix must have at least one row");
  if (n >= (int64_t)INT32_MAX) fail(kInvalid, "too many rows for int32 ids");
  const int64_t ne = (int64_t)coo.size();
  const int64_t* rows = coo.rows.data();
  const int64_t* cols = coo.cols.data();
  // index checks, and whether the triplets are already in row order (the
  // generators and unions of them emit them so)
  constexpr int64_t kGrain = 1 << 20;
  const int64_t nch = (ne + kGrain - 1) / kGrain;
  std::vector<uint8_t> unsortedChunk((size_t)std::max<int64_t>(nch, 1), 0);
  gph::parallelFor(nch, 1, [&](int64_t k) {
    const int64_t b = k * kGrain, e = std::min(ne, b + kGrain);
    for (int64_t i = b; i < e; i++) {
      if (rows[i] < 0 || rows[i] >= n || cols[i] < 0 || cols[i] >= n)
        fail(kInvalid, "triplet index out of range");
      if (i > 0 && rows[i] < rows[i - 1]) unsortedChunk[(size_t)k] = 1;
    }
  });
  const bool sorted =
      std::find(unsortedChunk.begin(), unsortedChunk.end(), 1) == unsortedChunk.end();
  // cnt[r]: first triplet of row r in `ord` (the identity when sorted)
  std::vector<int64_t> cnt((size_t)n + 1, 0);
  std::vector<int64_t> ord;
  if (sorted) {
    gph::parallelFor(nch, 1, [&](int64_t k) {
      const int64_t b = k * kGrain, e = std::min(ne, b + kGrain);
      for (int64_t i = b; i < e; i++) {
        const int64_t prev = i == 0 ? -1 : rows[i - 1];
        for (int64_t r = prev + 1; r <= rows[i]; r++) cnt[(size_t)r] = i;
      }
    });
    for (int64_t r = ne > 0 ? rows[ne - 1] + 1 : 0; r <= n; r++) cnt[(size_t)r] = ne;
  } else {
    for (int64_t e = 0; e < ne; e++) cnt[(size_t)rows[e] + 1]++;
    for (int64_t r = 0; r < n; r++) cnt[(size_t)r + 1] += cnt[(size_t)r];
    ord.resize((size_t)ne);
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < ne; e++) ord[(size_t)fill[(size_t)rows[e]]++] = e;
  }
  // per row: entries in column order (stable), the diagonal summed apart,
  // duplicates summed; row ranges in parallel, then concatenated
  GlobalCRS g;
  g.rowPtr.assign((size_t)n + 1, 0);
  g.diag.assign((size_t)n, 0.0f);
  constexpr int64_t kRowGrain = 1 << 14;
  const int64_t nrc = (n + kRowGrain - 1) / kRowGrain;
  struct Piece {
    std::vector<int32_t> col;
    std::vector<float> off;
  };
  std::vector<Piece> pieces((size_t)nrc);
  gph::parallelFor(nrc, 1, [&](int64_t k) {
    const int64_t r0 = k * kRowGrain, r1 = std::min(n, r0 + kRowGrain);
    Piece& pc = pieces[(size_t)k];
    pc.col.reserve((size_t)(cnt[(size_t)r1] - cnt[(size_t)r0]));
    pc.off.reserve((size_t)(cnt[(size_t)r1] - cnt[(size_t)r0]));
    std::vector<std::pair<int64_t, float>> rowE;
    for (int64_t r = r0; r < r1; r++) {
      rowE.clear();
      for (int64_t q = cnt[(size_t)r]; q < cnt[(size_t)r + 1]; q++) {
        const int64_t e = sorted ? q : ord[(size_t)q];
        rowE.emplace_back(cols[e], coo.vals[(size_t)e]);
      }
      // stable: insertion for short rows (no allocation), merge sort beyond
      if (rowE.size() <= 64) {
        for (size_t q = 1; q < rowE.size(); q++) {
          const auto x = rowE[q];
          size_t z = q;
          for (; z > 0 && rowE[z - 1].first > x.first; z--) rowE[z] = rowE[z - 1];
          rowE[z] = x;
        }
      } else {
        std::stable_sort(rowE.begin(), rowE.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
      }
      const size_t rowStart = pc.col.size();
      for (size_t q = 0; q < rowE.size(); q++) {
        const int64_t c = rowE[q].first;
        const float v = rowE[q].second;
        if (c == r) {
          g.diag[(size_t)r] = g.diag[(size_t)r] + v;
          continue;
        }
        if (pc.col.size() > rowStart && pc.col.back() == c) {
          pc.off.back() = pc.off.back() + v;
          continue;
        }
        pc.col.push_back((int32_t)c);
        pc.off.push_back(v);
      }
      g.rowPtr[(size_t)r + 1] = (int64_t)(pc.col.size() - rowStart);
    }
  });
  std::vector<int64_t> base((size_t)nrc + 1, 0);
  for (int64_t k = 0; k < nrc; k++)
    base[(size_t)k + 1] = base[(size_t)k] + (int64_t)pieces[(size_t)k].col.size();
  if (base[(size_t)nrc] >= (int64_t)INT32_MAX)
    fail(kInvalid, "too many non-zeros for int32 offsets");
  g.col.resize((size_t)base[(size_t)nrc]);
  g.off.resize((size_t)base[(size_t)nrc]);
  gph::parallelFor(nrc, 1, [&](int64_t k) {
    const int64_t r0 = k * kRowGrain, r1 = std::min(n, r0 + kRowGrain);
    Piece& pc = pieces[(size_t)k];
    std::copy(pc.col.begin(), pc.col.end(), g.col.begin() + base[(size_t)k]);
    std::copy(pc.off.begin(), pc.off.end(), g.off.begin() + base[(size_t)k]);
    // row lengths to offsets (rowPtr[r0] is the piece's base)
    int64_t acc = base[(size_t)k];
    for (int64_t r = r0; r < r1; r++) {
      acc += g.rowPtr[(size_t)r + 1];
      g.rowPtr[(size_t)r + 1] = acc;
    }
    Piece().col.swap(pc.col);
    Piece().off.swap(pc.off);
  });
  coo = TripletMatrix();
  return g;
}

// ---------------------------------------------------------------------------
// Tile layouts (CRSHostMatrix.cpp:273-354, DistributedTileLayout.cpp:63-126),
// local addressing (CRSHostMatrix.cpp:379-432) and value decomposition
// (CRSHostMatrix.cpp:433-525). Rows are bucketed by tile instead of scanning
// every row per tile (same per-tile ascending order, O(n) instead of O(T n)).
// Halo regions of a tile are ordered by (source tile, source region index).
// ---------------------------------------------------------------------------
static void computeLayout(HostMatrix& m, const GlobalCRS& g) {
  const int64_t n = m.nrows;
  const int32_t T = m.numTiles;
  m.tileStart.assign((size_t)T + 1, 0);
  m.newToOld.assign((size_t)n, 0);
  m.oldToNew.assign((size_t)n, 0);
  m.numInterior.assign((size_t)T, 0);
  m.numSeparatorRegions.assign((size_t)T, 0);
  m.haloStart.assign((size_t)T + 1, 0);
  m.haloGlobal.clear();
  const int32_t* rowTile = m.rowToTile.data();

  if (T == 1) {
    // CRSHostMatrix.cpp:151-158: a single tile holds every row, all interior.
    std::iota(m.newToOld.begin(), m.newToOld.end(), 0);
    std::iota(m.oldToNew.begin(), m.oldToNew.end(), 0);
    m.tileStart[1] = n;
    m.numInterior[0] = (int32_t)n;
  } else {
    struct Region {
      std::vector<int32_t> dst;
      std::vector<int64_t> cells;
    };
    std::vector<std::vector<int64_t>> interior((size_t)T);
    std::vector<std::vector<Region>> seps((size_t)T);
    // Destination tiles of a row: tiles owning its columns
    // (CRSHostMatrix.cpp:289-299) united with tiles whose rows reference it;
    // identical for structurally symmetric patterns, and well defined for
    // nonsymmetric ones (where the reference's local addressing throws).
    // The referencing tiles of each row (a transpose; their order within a
    // row does not matter, the destination set is sorted below).
    std::vector<int64_t> refPtr((size_t)n + 1, 0);
    constexpr int64_t kRows = 1 << 14;
    const int64_t nrc = (n + kRows - 1) / kRows;
    gph::parallelFor(nrc, 1, [&](int64_t k) {
      for (int64_t r = k * kRows; r < std::min(n, (k + 1) * kRows); r++)
        for (int64_t a = g.rowPtr[(size_t)r]; a < g.rowPtr[(size_t)r + 1]; a++) {
          const int32_t c = g.col[(size_t)a];
          if (rowTile[c] != rowTile[r])
            __atomic_fetch_add(&refPtr[(size_t)c + 1], 1, __ATOMIC_RELAXED);
        }
    });
    for (int64_t r = 0; r < n; r++) refPtr[(size_t)r + 1] += refPtr[(size_t)r];
    std::vector<int32_t> refTile((size_t)refPtr[(size_t)n]);
    {
      std::vector<int64_t> fill(refPtr.begin(), refPtr.end() - 1);
      gph::parallelFor(nrc, 1, [&](int64_t k) {
        for (int64_t r = k * kRows; r < std::min(n, (k + 1) * kRows); r++)
          for (int64_t a = g.rowPtr[(size_t)r]; a < g.rowPtr[(size_t)r + 1]; a++) {
            const int32_t c = g.col[(size_t)a];
            if (rowTile[c] != rowTile[r])
              refTile[(size_t)__atomic_fetch_add(&fill[(size_t)c], 1, __ATOMIC_RELAXED)] =
                  rowTile[r];
          }
      });
    }
    // rows of each tile in ascending order
    std::vector<int64_t> tileRowPtr((size_t)T + 1, 0), tileRows((size_t)n);
    for (int64_t r = 0; r < n; r++) tileRowPtr[(size_t)rowTile[r] + 1]++;
    for (int32_t t = 0; t < T; t++) tileRowPtr[(size_t)t + 1] += tileRowPtr[(size_t)t];
    {
      std::vector<int64_t> fill(tileRowPtr.begin(), tileRowPtr.end() - 1);
      for (int64_t r = 0; r < n; r++) tileRows[(size_t)fill[(size_t)rowTile[r]]++] = r;
    }
    // per tile: interior rows, then separator regions (rows with the same
    // destination set) in order of first appearance
    gph::parallelFor(T, 1, [&](int64_t tt) {
      const int32_t t = (int32_t)tt;
      std::map<std::vector<int32_t>, int32_t> lookup;
      std::vector<int32_t> key;
      for (int64_t q = tileRowPtr[(size_t)t]; q < tileRowPtr[(size_t)t + 1]; q++) {
        const int64_t r = tileRows[(size_t)q];
        key.clear();
        for (int64_t a = g.rowPtr[(size_t)r]; a < g.rowPtr[(size_t)r + 1]; a++) {
          const int32_t ct = rowTile[g.col[(size_t)a]];
          if (ct != t) key.push_back(ct);
        }
        for (int64_t z = refPtr[(size_t)r]; z < refPtr[(size_t)r + 1]; z++)
          key.push_back(refTile[(size_t)z]);
        if (key.empty()) {
          interior[(size_t)t].push_back(r);
          continue;
        }
        std::sort(key.begin(), key.end());
        key.erase(std::unique(key.begin(), key.end()), key.end());
        auto it = lookup.find(key);
        int32_t reg;
        if (it == lookup.end()) {
          reg = (int32_t)seps[(size_t)t].size();
          seps[(size_t)t].push_back(Region{key, {}});
          lookup.emplace(key, reg);
        } else {
          reg = it->second;
        }
        seps[(size_t)t][(size_t)reg].cells.push_back(r);
      }
    });
    // halo of tile d: the regions that list d, by source tile, then region
    std::vector<std::vector<std::pair<int32_t, int32_t>>> into((size_t)T);
    for (int32_t s = 0; s < T; s++)
      for (int32_t reg = 0; reg < (int32_t)seps[(size_t)s].size(); reg++)
        for (int32_t d : seps[(size_t)s][(size_t)reg].dst) into[(size_t)d].push_back({s, reg});
    std::vector<int64_t> haloSize((size_t)T, 0);
    for (int32_t d = 0; d < T; d++)
      for (const auto& sr : into[(size_t)d])
        haloSize[(size_t)d] += (int64_t)seps[(size_t)sr.first][(size_t)sr.second].cells.size();
    int64_t pos = 0, h = 0;
    for (int32_t t = 0; t < T; t++) {
      m.tileStart[(size_t)t] = pos;
      m.numInterior[(size_t)t] = (int32_t)interior[(size_t)t].size();
      m.numSeparatorRegions[(size_t)t] = (int32_t)seps[(size_t)t].size();
      pos += tileRowPtr[(size_t)t + 1] - tileRowPtr[(size_t)t];
      if (m.tileStart[(size_t)t] == pos)
        fail(kInvalid, "tile " + std::to_string(t) + " holds no rows");
      m.haloStart[(size_t)t] = h;
      h += haloSize[(size_t)t];
    }
    m.tileStart[(size_t)T] = pos;
    m.haloStart[(size_t)T] = h;
    m.haloGlobal.resize((size_t)h);
    gph::parallelFor(T, 1, [&](int64_t tt) {
      const size_t t = (size_t)tt;
      int64_t p = m.tileStart[t];
      for (int64_t r : interior[t]) {
        m.newToOld[(size_t)p] = r;
        m.oldToNew[(size_t)r] = p++;
      }
      for (auto& R : seps[t])
        for (int64_t r : R.cells) {
          m.newToOld[(size_t)p] = r;
          m.oldToNew[(size_t)r] = p++;
        }
      int64_t q = m.haloStart[t];
      for (const auto& sr : into[t])
        for (int64_t r : seps[(size_t)sr.first][(size_t)sr.second].cells)
          m.haloGlobal[(size_t)q++] = r;
    });
  }

  // local addressing + values in local column order, tile by tile: row
  // offsets first (row lengths in the new order), then each tile's rows
  // with its halo ids looked up in a sorted copy of its halo list
  const int64_t nnz = g.rowPtr[(size_t)n];
  m.rowPtr.assign((size_t)n + 1, 0);
  m.col.resize((size_t)nnz);
  m.colLocal.resize((size_t)nnz);
  m.offDiag.resize((size_t)nnz);
  m.diag.resize((size_t)n);
  for (int64_t p = 0; p < n; p++) {
    const int64_t gr = m.newToOld[(size_t)p];
    m.rowPtr[(size_t)p + 1] = m.rowPtr[(size_t)p] + (g.rowPtr[(size_t)gr + 1] - g.rowPtr[(size_t)gr]);
  }
  struct E {
    int64_t local;
    int32_t pos;
    float v;
  };
  gph::parallelFor(T, 1, [&](int64_t tt) {
    const int32_t t = (int32_t)tt;
    const int64_t t0 = m.tileStart[(size_t)t];
    const int64_t nt = m.tileStart[(size_t)t + 1] - t0;
    std::vector<std::pair<int64_t, int64_t>> halo;  // (global row, local id)
    if (T > 1) {
      for (int64_t q = m.haloStart[(size_t)t]; q < m.haloStart[(size_t)t + 1]; q++)
        halo.push_back({m.haloGlobal[(size_t)q], nt + (q - m.haloStart[(size_t)t])});
      std::sort(halo.begin(), halo.end());
    }
    std::vector<E> es;
    for (int64_t lr = 0; lr < nt; lr++) {
      const int64_t p = t0 + lr;
      const int64_t gr = m.newToOld[(size_t)p];
      m.diag[(size_t)p] = g.diag[(size_t)gr];
      es.clear();
      for (int64_t a = g.rowPtr[(size_t)gr]; a < g.rowPtr[(size_t)gr + 1]; a++) {
        const int64_t gc = g.col[(size_t)a];
        int64_t local = -1;
        if (T == 1 || rowTile[gc] == t) {
          local = m.oldToNew[(size_t)gc] - t0;
        } else {
          auto it = std::lower_bound(halo.begin(), halo.end(), std::make_pair(gc, (int64_t)-1));
          if (it != halo.end() && it->first == gc) local = it->second;
        }
        if (local < 0) fail(kRuntime, "internal: column missing from the halo");
        es.push_back(E{local, (int32_t)m.oldToNew[(size_t)gc], g.off[(size_t)a]});
      }
      if (T > 1)
        std::sort(es.begin(), es.end(),
                  [](const E& x, const E& y) { return x.local < y.local; });
      int64_t c = m.rowPtr[(size_t)p];
      for (const E& e : es) {
        m.col[(size_t)c] = e.pos;
        m.colLocal[(size_t)c] = (int32_t)e.local;
        m.offDiag[(size_t)c] = e.v;
        c++;
      }
    }
  });
}

// Level sets (CRSHostMatrix.cpp:527-675): forward level of a row = 1 + the
// largest level of its in-tile lower neighbours (0 if none), which is the
// layer the reference's Kahn sweep assigns. Rows are grouped by level,
// ascending local id inside a level. Backward levels do the same over the
// in-tile upper neighbours (rows descending) for the backward sweeps.
static void computeLevels(HostMatrix& m) {
  const int64_t n = m.nrows;
  const int32_t T = m.numTiles;
  m.level.assign((size_t)n, 0);
  m.colorSortAddr.assign((size_t)n, 0);
  m.numColors.assign((size_t)T, 0);
  m.colorStart.clear();
  m.colorStartOffset.assign((size_t)T + 1, 0);
  m.upperLevelRows.assign((size_t)n, 0);
  m.upperColorStart.clear();
  m.upperColorStartOffset.assign((size_t)T + 1, 0);
  m.numUpperColors.assign((size_t)T, 0);
  std::vector<int32_t> ulev((size_t)n, 0);
  // per tile (independent): levels, rows grouped by level, level starts
  std::vector<std::vector<int64_t>> cStart((size_t)T), uStart((size_t)T);
  std::vector<int64_t> parTile((size_t)T, 0);
  gph::parallelFor(T, 1, [&](int64_t tt) {
    const int32_t t = (int32_t)tt;
    const int64_t t0 = m.tileStart[(size_t)t], t1 = m.tileStart[(size_t)t + 1];
    int32_t maxL = 0;
    for (int64_t p = t0; p < t1; p++) {
      int32_t L = 0;
      for (int64_t a = m.rowPtr[(size_t)p]; a < m.rowPtr[(size_t)p + 1]; a++) {
        int64_t j = m.col[(size_t)a];
        if (j < t0 || j >= t1 || j >= p) break;  // halo ids sort last
        L = std::max(L, m.level[(size_t)j] + 1);
      }
      m.level[(size_t)p] = L;
      maxL = std::max(maxL, L);
    }
    int32_t maxU = 0;
    for (int64_t p = t1 - 1; p >= t0; p--) {
      int32_t L = 0;
      for (int64_t a = m.rowPtr[(size_t)p]; a < m.rowPtr[(size_t)p + 1]; a++) {
        int64_t j = m.col[(size_t)a];
        if (j < t0 || j >= t1) break;
        if (j > p) L = std::max(L, ulev[(size_t)j] + 1);
      }
      ulev[(size_t)p] = L;
      maxU = std::max(maxU, L);
    }
    std::vector<int64_t> cnt;
    auto group = [&](const std::vector<int32_t>& lv, int32_t maxLev,
                     std::vector<int32_t>& rows, std::vector<int64_t>& starts,
                     bool countPar) {
      cnt.assign((size_t)maxLev + 2, 0);
      for (int64_t p = t0; p < t1; p++) cnt[(size_t)lv[(size_t)p] + 1]++;
      for (int32_t L = 0; L <= maxLev; L++) {
        if (countPar && cnt[(size_t)L + 1] > 12) parTile[(size_t)t] += cnt[(size_t)L + 1];
        cnt[(size_t)L + 1] += cnt[(size_t)L];
      }
      for (int32_t L = 0; L <= maxLev + 1; L++) starts.push_back(cnt[(size_t)L]);
      std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
      for (int64_t p = t0; p < t1; p++)
        rows[(size_t)(t0 + fill[(size_t)lv[(size_t)p]]++)] = (int32_t)(p - t0);
    };
    m.numColors[(size_t)t] = maxL + 1;
    group(m.level, maxL, m.colorSortAddr, cStart[(size_t)t], true);
    m.numUpperColors[(size_t)t] = maxU + 1;
    group(ulev, maxU, m.upperLevelRows, uStart[(size_t)t], false);
  });
  int64_t parRows = 0;
  for (int32_t t = 0; t < T; t++) {
    m.colorStartOffset[(size_t)t] = (int64_t)m.colorStart.size();
    m.colorStart.insert(m.colorStart.end(), cStart[(size_t)t].begin(), cStart[(size_t)t].end());
    m.upperColorStartOffset[(size_t)t] = (int64_t)m.upperColorStart.size();
    m.upperColorStart.insert(m.upperColorStart.end(), uStart[(size_t)t].begin(),
                             uStart[(size_t)t].end());
    parRows += parTile[(size_t)t];
  }
  m.colorStartOffset[(size_t)T] = (int64_t)m.colorStart.size();
  m.upperColorStartOffset[(size_t)T] = (int64_t)m.upperColorStart.size();
  // "multicolor recommended" if > 90% of rows sit in levels with > 12 rows
  m.multicolorRecommended = (double)parRows > 0.9 * (double)n;
}

std::shared_ptr<HostMatrix> buildHostMatrix(TripletMatrix&& coo,
                                            int32_t numTiles,
                                            const int32_t* rowToTile) {
  if (numTiles < 1) fail(kInvalid, "numTiles must be >= 1");
  auto m = std::make_shared<HostMatrix>();
  m->nrows = coo.nrows;
  GlobalCRS g = convertToCRS(std::move(coo));
  const int64_t n = m->nrows;
  if (numTiles > n) fail(kInvalid, "more tiles than rows");
  if (numTiles == 1) {
    m->rowToTile.assign((size_t)n, 0);
  } else if (rowToTile) {
    m->rowToTile.assign(rowToTile, rowToTile + n);
    // compact away empty tiles (CRSHostMatrix.cpp:238-264)
    std::vector<int64_t> perTile((size_t)numTiles, 0);
    for (int32_t t : m->rowToTile) {
      if (t < 0 || t >= numTiles) fail(kInvalid, "rowToTile entry out of range");
      perTile[(size_t)t]++;
    }
    std::vector<int32_t> shift((size_t)numTiles, -1);
    int32_t k = 0;
    for (int32_t t = 0; t < numTiles; t++)
      if (perTile[(size_t)t] > 0) shift[(size_t)t] = k++;
    for (auto& t : m->rowToTile) t = shift[(size_t)t];
    numTiles = k;
  } else {
    m->rowToTile = partitionContiguous(g.rowPtr, numTiles);
  }
  m->numTiles = numTiles;
  computeLayout(*m, g);
  computeLevels(*m);
  return m;
}

void forwardLevelRows(const HostMatrix& m, std::vector<int32_t>& list,
                      std::vector<int64_t>& levelOff) {
  const int32_t T = (int32_t)m.numColors.size();
  int32_t nl = 0;
  for (int32_t t = 0; t < T; t++) nl = std::max(nl, m.numColors[(size_t)t]);
  levelOff.assign((size_t)nl + 1, 0);
  for (int32_t t = 0; t < T; t++) {
    const int64_t* cs = m.colorStart.data() + m.colorStartOffset[(size_t)t];
    for (int32_t L = 0; L < m.numColors[(size_t)t]; L++) levelOff[(size_t)L + 1] += cs[L + 1] - cs[L];
  }
  for (int32_t L = 0; L < nl; L++) levelOff[(size_t)L + 1] += levelOff[(size_t)L];
  list.assign(2 * (size_t)levelOff[(size_t)nl], 0);
  std::vector<int64_t> at(levelOff.begin(), levelOff.end() - 1);
  for (int32_t t = 0; t < T; t++) {
    const int64_t t0 = m.tileStart[(size_t)t];
    const int64_t* cs = m.colorStart.data() + m.colorStartOffset[(size_t)t];
    for (int32_t L = 0; L < m.numColors[(size_t)t]; L++)
      for (int64_t q = cs[L]; q < cs[L + 1]; q++) {
        const size_t r = (size_t)at[(size_t)L]++;
        list[2 * r] = (int32_t)(t0 + m.colorSortAddr[(size_t)(t0 + q)]);
        list[2 * r + 1] = t;
      }
  }
}

LDUHost buildLDU(const HostMatrix& m) {
  const int64_t n = m.nrows;
  LDUHost L;
  L.ownerStart.assign((size_t)n + 1, 0);
  L.neighbourStart.assign((size_t)n + 1, 0);
  constexpr int64_t kRows = 4096;
  // faces: owner ascending, neighbour ascending (rows in parallel: counts,
  // offsets, then each row's upper entries sorted into place)
  std::vector<int64_t> up((size_t)n + 1, 0), low((size_t)n + 1, 0);
  gph::parallelFor(n, kRows, [&](int64_t l) {
    int64_t cu = 0, cl = 0;
    for (int64_t a = m.rowPtr[(size_t)l]; a < m.rowPtr[(size_t)l + 1]; a++) {
      if (m.col[(size_t)a] > l) cu++;
      if (m.col[(size_t)a] < l) cl++;
    }
    up[(size_t)l + 1] = cu;
    low[(size_t)l + 1] = cl;
  });
  for (int64_t l = 0; l < n; l++) {
    up[(size_t)l + 1] += up[(size_t)l];
    low[(size_t)l + 1] += low[(size_t)l];
  }
  L.numFaces = up[(size_t)n];
  // every face (l, u) must appear as the lower entry (u, l): counts first
  if (low[(size_t)n] != L.numFaces)
    fail(kInvalid, "LDU format needs a structurally symmetric matrix");
  if (L.numFaces >= (int64_t)INT32_MAX) fail(kInvalid, "too many faces for int32 ids");
  for (int64_t l = 0; l <= n; l++) {
    L.ownerStart[(size_t)l] = (int32_t)up[(size_t)l];
    L.neighbourStart[(size_t)l] = (int32_t)low[(size_t)l];
  }
  L.lowerAddr.resize((size_t)L.numFaces);
  L.upperAddr.resize((size_t)L.numFaces);
  L.upper.resize((size_t)L.numFaces);
  gph::parallelFor(n, kRows, [&](int64_t l) {
    std::vector<std::pair<int32_t, float>> ups;
    for (int64_t a = m.rowPtr[(size_t)l]; a < m.rowPtr[(size_t)l + 1]; a++)
      if (m.col[(size_t)a] > l) ups.emplace_back(m.col[(size_t)a], m.offDiag[(size_t)a]);
    std::sort(ups.begin(), ups.end(),
              [](const auto& x, const auto& y) { return x.first < y.first; });
    int64_t f = up[(size_t)l];
    for (auto& u : ups) {
      L.lowerAddr[(size_t)f] = (int32_t)l;
      L.upperAddr[(size_t)f] = u.first;
      L.upper[(size_t)f] = u.second;
      f++;
    }
  });
  // losort: per neighbour cell u, its faces (l, u) by face index, i.e. by
  // owner l ascending; with them the lower coefficients A[u][l]
  L.neighbourColInd.assign((size_t)L.numFaces, 0);
  L.lower.assign((size_t)L.numFaces, 0.0f);
  gph::parallelFor(n, kRows, [&](int64_t u) {
    std::vector<std::pair<int32_t, float>> lows;
    for (int64_t a = m.rowPtr[(size_t)u]; a < m.rowPtr[(size_t)u + 1]; a++)
      if (m.col[(size_t)a] < u) lows.emplace_back(m.col[(size_t)a], m.offDiag[(size_t)a]);
    std::sort(lows.begin(), lows.end(),
              [](const auto& x, const auto& y) { return x.first < y.first; });
    int64_t s = low[(size_t)u];
    for (auto& lw : lows) {
      const int32_t l = lw.first;
      // face (l, u): search l's owned faces
      const int32_t f0 = L.ownerStart[(size_t)l], f1 = L.ownerStart[(size_t)l + 1];
      auto it = std::lower_bound(L.upperAddr.begin() + f0, L.upperAddr.begin() + f1,
                                 (int32_t)u);
      if (it == L.upperAddr.begin() + f1 || *it != (int32_t)u)
        fail(kInvalid, "LDU format needs a structurally symmetric matrix");
      L.neighbourColInd[(size_t)s] = (int32_t)(it - L.upperAddr.begin());
      L.lower[(size_t)s] = lw.second;
      s++;
    }
  });
  return L;
}

}  // namespace graphene::matrix::host
