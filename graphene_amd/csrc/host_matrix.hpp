// Host-side matrix preparation: triplet generators and loaders, COO->CRS,
// row partitions, the per-tile layout (interior / separator / halo regions),
// local addressing, level sets and the LDU addressing.
//
// Mirrors lg/matrix/host/** of the reference: generate3DPoissonMatrix and
// loadMatrixFromFile (HostMatrix.hpp:88-96), CRSHostMatrix (CRSHostMatrix.cpp:63-85).
// Instead of one address space per IPU tile, the result is a single
// "tile-major" numbering: rows are stored tile after tile, each tile in the
// reference's local order, so the tile vectors of the reference are
// contiguous slices of one HBM vector and the halo exchange becomes a direct
// read of the owner's slot.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace graphene::matrix::host {

// COO triplets in structure-of-arrays form (CoordinateFormat.hpp:29-37).
struct TripletMatrix {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> rows, cols;
  std::vector<float> vals;
  void push(int64_t r, int64_t c, float v) {
    rows.push_back(r);
    cols.push_back(c);
    vals.push_back(v);
  }
  size_t size() const { return vals.size(); }
};

// Generators. generate3DPoissonTripletMatrix follows Poisson.cpp:26-69; the
// others have no reference counterpart (definitions in the .cpp file).
TripletMatrix generate3DPoissonTripletMatrix(int64_t nx, int64_t ny,
                                             int64_t nz);
TripletMatrix generate3DConvectionDiffusion27(int64_t nx, int64_t ny,
                                              int64_t nz, double peclet,
                                              uint64_t seed);
TripletMatrix generate2DConvectionDiffusion9(int64_t nx, int64_t ny,
                                             double diffusion, double peclet,
                                             uint64_t seed);
TripletMatrix generatePowerLaw(int64_t n, int32_t minLen, int32_t maxLen,
                               double exponent, int32_t band,
                               double localFraction, uint64_t seed);
// Matrix Market coordinate files (MatrixMarket.cpp:32-66; symmetric files are
// generalised like fast_matrix_market's default).
TripletMatrix loadTripletMatrixFromFile(const std::string& path);

// Portable RNG shared with the tests: splitmix64-seeded xoshiro256**.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed);
  uint64_t next();
  float uniformPm1();      // ((u >> 40) * 2^-24) * 2 - 1
  double uniform01d();     // (u >> 11) * 2^-53
};
void uniformPm1(uint64_t seed, int64_t n, float* out);

// Row partitions (stand-ins for the METIS call of CRSHostMatrix.cpp:139-271;
// METIS is not available here, these are this library's own deterministic
// partitioners).
std::vector<int32_t> partitionGrid3D(int64_t nx, int64_t ny, int64_t nz,
                                     int64_t bx, int64_t by, int64_t bz,
                                     int32_t* numTiles);
std::vector<int32_t> partitionContiguous(const std::vector<int64_t>& rowPtr,
                                         int32_t numTiles);
// Graph partition of the symmetrized pattern of a COO matrix (rows/cols
// 0-based, diagonal and duplicates allowed) into numTiles parts by
// recursive breadth-first bisection: each split grows the first part from a
// pseudo-peripheral vertex in BFS order until it holds its share of the
// weight (row length + 1, as METIS gets in CRSHostMatrix.cpp:197).
// Deterministic (neighbours visited in ascending order).
std::vector<int32_t> partitionGraph(int64_t nrows, const std::vector<int64_t>& rows,
                                    const std::vector<int64_t>& cols, int32_t numTiles);

// Prepared host matrix (CRSHostMatrix of the reference).
class HostMatrix {
 public:
  int64_t nrows = 0;
  int32_t numTiles = 1;
  std::string name = "A";

  // partition and layout
  std::vector<int32_t> rowToTile;
  std::vector<int64_t> tileStart;   // T+1, tile-major row offsets
  std::vector<int64_t> newToOld;    // tile-major position -> global row
  std::vector<int64_t> oldToNew;    // global row -> tile-major position
  std::vector<int32_t> numInterior; // per tile
  std::vector<int32_t> numSeparatorRegions;
  std::vector<int64_t> haloStart;   // T+1
  std::vector<int64_t> haloGlobal;  // halo cells (global ids) per tile

  // tile-major CRS, columns in the reference's local order
  std::vector<int64_t> rowPtr;      // n+1
  std::vector<int32_t> col;         // tile-major positions
  std::vector<int32_t> colLocal;    // local ids (>= n_t for halo)
  std::vector<float> offDiag;
  std::vector<float> diag;

  // level sets (calculateRowColors / calculateColorAddressings)
  std::vector<int32_t> level;            // forward level per row
  std::vector<int32_t> colorSortAddr;    // local ids grouped by level
  std::vector<int64_t> colorStart;       // concatenated, numColors[t]+1 each
  std::vector<int64_t> colorStartOffset; // T+1 offsets into colorStart
  std::vector<int32_t> numColors;
  // backward levels (upper triangle) used by the GPU backward sweeps
  std::vector<int32_t> upperLevelRows;     // local ids grouped by level
  std::vector<int64_t> upperColorStart;
  std::vector<int64_t> upperColorStartOffset;
  std::vector<int32_t> numUpperColors;
  bool multicolorRecommended = false;

  int64_t nnzOff() const { return rowPtr.empty() ? 0 : rowPtr.back(); }
  int64_t numHalo() const { return haloStart.empty() ? 0 : haloStart.back(); }
  int64_t tileRows(int32_t t) const { return tileStart[t + 1] - tileStart[t]; }
};

// Builds the prepared matrix. rowToTile may be null (numTiles == 1 or the
// contiguous partitioner). Follows CRSHostMatrix::CRSHostMatrix
// (CRSHostMatrix.cpp:63-85).
std::shared_ptr<HostMatrix> buildHostMatrix(TripletMatrix&& coo,
                                            int32_t numTiles,
                                            const int32_t* rowToTile);

// Rows of the forward levels of all tiles, level by level (level 0 of every
// tile, then level 1, ...): list[2r] = row (tile-major), list[2r+1] = its
// tile; levelOff[L]..levelOff[L+1]-1 are level L's entries.
void forwardLevelRows(const HostMatrix& m, std::vector<int32_t>& list,
                      std::vector<int64_t>& levelOff);

// LDU addressing over the tile-major numbering (LDUAddressing.hpp:30-63).
struct LDUHost {
  int64_t numFaces = 0;
  std::vector<int32_t> lowerAddr, upperAddr;      // per face
  std::vector<int32_t> ownerStart;                // n+1
  std::vector<int32_t> neighbourStart;            // n+1 (losortStart)
  std::vector<int32_t> neighbourColInd;           // losort -> face
  std::vector<float> upper;                       // per face
  std::vector<float> lower;                       // per losort position
};
// Throws if the pattern is not structurally symmetric.
LDUHost buildLDU(const HostMatrix& m);

}  // namespace graphene::matrix::host
