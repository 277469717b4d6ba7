// Level-major packed layout of the ILU(0) factors for the gfx950
// substitution kernel. Consecutive tiles are grouped into one wave's unit
// of work (the tiles are independent blocks, so a group is a block-diagonal
// "super tile" whose levels are the unions of its tiles' levels).
//
// For every tile and direction (forward: strictly lower in-tile entries in
// ascending column order; backward: strictly upper in-tile entries in
// descending column order, as ilu/Solver.cpp:380-442 visits them), the
// entries are grouped by level and stored level after level; every level
// starts on a 64-entry "slab" boundary so one wave loads a slab with one
// coalesced instruction. Within a level of R rows and width W (the longest
// row), entry j of row r sits at j*R + r; shorter rows are padded with zero
// entries that read a zero slot. Subtracting their +0 products leaves every
// partial sum unchanged, so the substitution keeps the reference's
// arithmetic exactly. Per row slot: the row's local id.
#pragma once

#include <cstdint>
#include <vector>

#include "host_matrix.hpp"

namespace graphene::matrix {

constexpr int kPackSlab = 64;   // entries per slab (one per lane)
constexpr int kPackDepth = 16;  // slabs in flight per wave (prefetch ring)

struct PackedDirection {
  std::vector<uint16_t> col;      // local column id per slot (zero slot = n_t)
  std::vector<int32_t> src;       // CRS position of the entry, -1 = zero
  std::vector<int32_t> slabOff;   // T+1: first slab of each tile's stream
  std::vector<int32_t> lvlOff;    // T+1: offset of each tile's level table
  std::vector<int32_t> slabEnd;   // per tile: numLevels (cumulative) + sentinel
  std::vector<int32_t> rowStart;  // per tile: numLevels+1 entries
  std::vector<int32_t> width;     // per tile: row width of each level (+ 0)
  std::vector<uint16_t> rowLoc;   // per row slot (level order): local row id
  int32_t maxLevels = 0;
  int32_t maxLevelEntries = 0;
};

struct PackedILU {
  PackedDirection fwd, bwd;
  std::vector<int32_t> groupStart;  // G+1 row offsets of the wave groups
  int32_t maxRows = 0;
  bool ok = false;  // false: a group exceeds the 16-bit local ids
};

// Groups consecutive tiles into units of about targetRows rows (one wave
// each) and packs both substitution directions. Every group's stream has an
// even number of slabs.
PackedILU packILU(const host::HostMatrix& m, int64_t targetRows);

// Device order of a 16-bit slot array: slabs 2p and 2p+1 are interleaved so
// that lane l loads both of its slots with one 32-bit load,
// out[p*128 + 2l + (s&1)] = in[s*64 + l].
template <typename T>
std::vector<T> pairSlabs(const std::vector<T>& in) {
  std::vector<T> out(in.size());
  const size_t nSlabs = in.size() / kPackSlab;
  for (size_t s = 0; s < nSlabs; s++)
    for (size_t l = 0; l < (size_t)kPackSlab; l++)
      out[(s >> 1) * 2 * kPackSlab + 2 * l + (s & 1)] = in[s * kPackSlab + l];
  return out;
}

}  // namespace graphene::matrix
