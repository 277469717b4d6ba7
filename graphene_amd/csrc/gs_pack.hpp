// Jagged-diagonal level stream of the matrix for the Gauss-Seidel sweep
// kernel (k_gs.hip).
//
// A sweep of tile t visits the tile's rows level by level (forward: levels
// of the lower in-tile triangle; backward: levels of the upper one); the
// rows of a level are independent. Per direction, each level's rows are
// sorted by length (longest first, stable) into slots; level {first slot,
// rows R, first diagonal, width W (its longest row)}. Entry j of every row
// longer than j forms diagonal j: diagonal j holds entry j of slots 0..c_j-1
// (c_j = rows longer than j) at diag[first diagonal + j] + slot, diagonal
// j+1 follows it (diag[.. + W] ends the level). Consecutive slots' entries
// are adjacent, so a wave's loads of one entry index are coalesced, and
// nothing is padded. Per slot: tile-local row id << 8 | row length. An
// entry's code says where its x comes from: bit 31 set, the tile-local id
// of a row this sweep has already updated (in-tile j < i forward, j > i
// backward: read from the sweep's own copy); clear, the tile-major column
// (the previous x). The entries of a row keep their stored order, which is
// the order the sweep subtracts them in. Within a level the stream holds
// the entries in gather order (previous-x columns ascending, then the
// values this sweep wrote), each with its jagged-diagonal position in the
// level (pos), where its product is staged: the gathers of neighbouring
// lanes share cache lines, the subtraction order is unchanged.
#pragma once

#include <climits>
#include <cstdint>
#include <vector>

#include "host_matrix.hpp"

namespace graphene::matrix {

constexpr uint32_t kGsFresh = 0x80000000u;

struct GsStream {
  std::vector<int32_t> lvlBase;  // T+1: first level of each tile
  std::vector<int32_t> lvl;      // per level: first slot, rows, first diagonal, width
  std::vector<int32_t> diag;     // per level: width + 1 diagonal starts
  std::vector<uint32_t> slot;    // per slot: local row << 8 | length
  std::vector<uint32_t> code;    // per entry: x source
  std::vector<int32_t> src;      // per entry: CRS position of its value
  std::vector<uint16_t> pos;     // per entry: jagged-diagonal position in its level
  int32_t maxRows = 0;           // largest tile
  int32_t maxLevelEntries = 0;   // entries of the largest level
  bool ok = false;  // false: rows longer than 255, tiles over 2^24 rows, levels over 2^16
                    // entries, or over 32-bit offsets
};

// Levels of more than maxLevelEntries entries are split into consecutive
// levels of at most that many (or of one row).
GsStream packGs(const host::HostMatrix& m, bool backward,
                int64_t maxLevelEntries = INT64_MAX);

}  // namespace graphene::matrix
