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

// Consecutive tiles grouped into units of about targetRows rows: G+1 tile
// boundaries.
std::vector<int32_t> groupTiles(const host::HostMatrix& m, int64_t targetRows);

// Level-record stream of the ILU(0) factors for the lane-per-row kernel.
//
// Groups are taken in pairs (A = group 2p, B = group 2p+1); a wave runs both
// with lanes 0-31 on A and 32-63 on B. Per pair, the forward then the
// backward substitution is a sequence of steps; a step holds up to 32 rows
// of the current level of each half (rows of a level are independent). A
// step record is 16-byte aligned:
//   header (16 B): u16 Rm (row slots per half), u16 W (entries per row
//                  slot, even), u32 bytes (record size), u32 flags (bit 0:
//                  per-row 1/diag present), u32 0
//   u16 local row id [2][Rm]                     (padded to 16 B)
//   factor 1/diag [2][Rm]        (flag bit 0)    (padded to 16 B)
//   u16 local column [2][Rm][W]  (row-major)     (padded to 16 B)
//   factor value [2][Rm][W]      (row-major)     (padded to 16 B)
// Slot q = h*Rm + r is row r of half h. Entries of a row are in the order
// the reference's substitution visits them (forward: strictly lower in-tile
// entries ascending; backward: strictly upper in-tile entries descending,
// ilu/Solver.cpp:380-442). Unused entries and empty slots refer to the zero
// slot (local id maxRows, where x stays +0) with value 0: their +0 products
// leave every partial sum unchanged. A record is at most maxStepBytes and
// does not cross a multiple of `ring` bytes from its pair's start (the step
// takes fewer rows, or a filler record with Rm = 0 pads to the boundary).
struct IluLevelStream {
  std::vector<uint8_t> bytes;        // all pairs, then a zero tail
  std::vector<int64_t> pairOff;      // P+1 byte offsets (multiples of 16)
  std::vector<int32_t> steps;        // 2P: forward, backward step counts
  std::vector<int32_t> groupStart;   // G+1 row offsets
  int32_t maxRows = 0;               // largest group (the zero slot's id)
  bool ok = false;  // false: 16-bit ids or the record size limit exceeded
};

// lu: factor off-diagonals in CRS order, invD: 1/diag (or the DILU
// diagonal), both of element size es (2, 4 or 8 bytes). diluFwd: forward
// records carry 1/diag too. tail: zero bytes appended (read-ahead).
IluLevelStream packIluLevels(const host::HostMatrix& m, int64_t targetRows,
                             const void* lu, const void* invD, int es, bool diluFwd,
                             int64_t maxStepBytes, int64_t ring, int64_t tail,
                             int64_t wAlign);

// Row stream of the ILU(0) factors for the one-wave-per-tile substitution
// kernel (k_ilu_rows.hip), whose 64 lanes form 4 rows of 16.
//
// Per tile, the forward then the backward substitution is a sequence of
// steps, each holding rows of one level (the rows of a level are
// independent). A step has up to kRowsColumns columns; a column holds either
// two short records (at most kRowsShortLen entries each: lane rows 0-1 and
// 2-3) or one long record (at most kRowsMaxLen entries, all four lane rows).
// A record is a row's entries (forward: strictly lower in-tile entries
// ascending; backward: strictly upper in-tile entries descending, the order
// ilu/Solver.cpp:380-442 visits them), stored contiguously and padded with
// zero entries to a multiple of kRowsPad, or (a dense stream; not for
// float64 factors) to an even number of entries (one zero entry for an odd
// row: 16-bit columns then start on 4-byte boundaries). A step stores its records one after the
// other in record-word order. Lane l (lane row l >> 4, column l & 15) takes
// record word 2 (l & 15) + (l >> 5) and the record's chunk of kRowsPad
// entries (l >> 4) & 1 of a short record, chunk 0, 1, 3, 2 for lane rows
// 0-3 of a long record: it loads the kRowsPad entries from the chunk's
// start and uses the ones inside the record (the rest belong to the next
// record, step or tile, or to the zero tail; the kernel does not subtract
// their products). An entry is its column as a 16-bit byte offset into the
// tile's x slots (4 * slot; slot 0 holds +0, row i of the tile is slot
// i + 1) and the factor value, at the same element index of their streams.
// A zero entry (value +0, column slot 0) adds a +0 product, which leaves
// every partial sum unchanged. A row with more than kRowsMaxLen entries
// takes the same column of consecutive steps of its level, kRowsMaxLen
// entries per step; the lanes carry the partial sum from one step to the
// next (flags: continue from the carried sum; partial, the row is not
// finished). Rows are placed longest first.
//   step header (u32): elements of the step (bits 0-11), record words R =
//     2 * columns (12-17), a long record present (30), backward (31)
//   record words (u32, two per column, the same word twice for a long
//     record, 0 for an empty half): the record's first element relative to
//     the step's (bits 0-8), its stored length (9-14), long (15),
//     continue (16), partial (17), the row's slot (18-31)
// A tile's steps use elements [elemOff[t], elemOff[t+1]), headers
// [hdrOff[t], hdrOff[t+1]) and record words [metaOff[t], metaOff[t+1]).
// The element streams end with kRowsTail elements of zero entries.
constexpr int kRowsShortLen = 16;
constexpr int kRowsMaxLen = 32;
constexpr int kRowsPad = 8;
constexpr int kRowsColumns = 16;
constexpr int kRowsTail = 64;
constexpr uint32_t kRowsLong = 1u << 15;
constexpr uint32_t kRowsCont = 1u << 16;
constexpr uint32_t kRowsPartial = 1u << 17;
constexpr int kRowsLenShift = 9;
constexpr int kRowsSlotShift = 18;
struct IluRowStream {
  std::vector<uint16_t> col;       // byte offsets (4 * local id)
  std::vector<uint8_t> val;        // factor values (es bytes each)
  std::vector<uint32_t> elemOff;   // T+1
  std::vector<int32_t> hdrOff;     // T+1
  std::vector<int64_t> metaOff;    // T+1
  std::vector<uint32_t> hdr;
  std::vector<uint32_t> meta;
  int32_t maxRows = 0;             // largest tile (the zero slot's id)
  int32_t maxSteps = 0;            // most steps of one tile (both directions)
  int64_t maxRecs = 0;             // most records of one tile
  bool dense = false;              // records padded to even lengths only
  bool ok = false;  // false: ids or the stream too large for the formats
};

// dense: records padded to an even length (element size 2 or 4) instead of
// whole chunks. Fewer bytes for rows longer than a chunk, at the price of a
// per-entry test in the kernel and more registers: worth it when the rows
// average more than kRowsPad in-tile entries per direction (3D 27-point
// tiles: 22.5k instead of 28.2k elements per 16x8x8 tile, apply 95 -> 89 us;
// 2D 9-point tiles of 4-entry rows: 10 % slower).
IluRowStream packIluRows(const host::HostMatrix& m, const void* lu, int es, bool dense);

// Level-major stream of the ILU(0) factors for the workgroup-per-tile
// kernel (tiles of any size; x in LDS when it fits).
//
// Per direction (forward: strictly lower in-tile entries in stored,
// ascending order; backward: strictly upper in-tile entries in descending
// order, as ilu/Solver.cpp:380-442 visits them) and tile, the tile's levels
// one after the other, a level of more than kBlockSlots rows split into
// pieces of at most that many (consecutive "levels" of the stream). A level
// of R rows: info {first slot, R, first entry, W (its longest row)}; slot s
// is a record (local row id, the row's 1/d); entry j < width of slot s is
// record first entry + j*R + s (local column, factor value). Every level has
// `width` entries per slot (the largest W of the stream rounded up to 4, 8
// or 16); entries past a row's end refer to the zero slot (local id = the
// tile's row count, where the kernel keeps x = +0) with value +0, and
// subtracting their +0 products leaves every partial sum unchanged.
// Records are 8 bytes (32-bit id and a float32 value or float16 value bits)
// or 16 bytes (id, 0, float64 value) for float64 factors.
constexpr int kBlockMaxWidth = 16;
constexpr int kBlockSlots = 1024;  // rows of a level piece (one per thread)

struct IluBlockDirection {
  std::vector<int32_t> lvlBase;  // T+1: each tile's first level
  std::vector<int32_t> info;     // 4 per level
  std::vector<uint8_t> rows;     // per slot: (row, 1/d) record
  std::vector<uint8_t> ents;     // per entry: (column, value) record
};

struct IluBlockStream {
  IluBlockDirection fwd, bwd;
  int32_t width = 0;      // entries per slot (4, 8 or 16)
  int32_t maxRows = 0;    // largest tile
  int32_t maxLevels = 0;  // most levels of one tile, both directions together
  bool ok = false;        // false: a level wider than kBlockMaxWidth, or too large
};

// lu: factor values (es bytes each, CRS order); invD: 1/d per row (ies
// bytes: 8 for float64 factors, else 4).
IluBlockStream packIluBlock(const host::HostMatrix& m, const void* lu, int es, const void* invD,
                            int ies);

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
