// Level-major packed ILU(0) layout (see ilu_pack.hpp).
#include "ilu_pack.hpp"

#include <algorithm>
#include <climits>

namespace graphene::matrix {

// Consecutive tiles are grouped until a group holds about `targetRows`
// rows; a group is one wave's unit of work. Tiles are independent blocks,
// so level L of a group is the union of its tiles' level-L rows and the
// group's level count is the largest tile level count.
static std::vector<int32_t> groupTiles(const host::HostMatrix& m,
                                       int64_t targetRows) {
  std::vector<int32_t> g;
  g.push_back(0);
  int64_t rows = 0;
  for (int32_t t = 0; t < m.numTiles; t++) {
    const int64_t nt = m.tileRows(t);
    if (rows > 0 && rows + nt > targetRows) {
      g.push_back(t);
      rows = 0;
    }
    rows += nt;
  }
  g.push_back(m.numTiles);
  return g;
}

static void packDirection(const host::HostMatrix& m,
                          const std::vector<int32_t>& groups, bool backward,
                          PackedDirection& d, bool& ok) {
  const int32_t G = (int32_t)groups.size() - 1;
  d.slabOff.assign((size_t)G + 1, 0);
  d.lvlOff.assign((size_t)G + 1, 0);
  d.rowLoc.assign((size_t)m.nrows, 0);
  const std::vector<int32_t>& rows = backward ? m.upperLevelRows : m.colorSortAddr;
  const std::vector<int64_t>& starts = backward ? m.upperColorStart : m.colorStart;
  const std::vector<int64_t>& startOff =
      backward ? m.upperColorStartOffset : m.colorStartOffset;
  int64_t slab = 0;
  std::vector<int64_t> lvRows;   // rows of the level (matrix ids)
  std::vector<int32_t> lvEnts;   // their entries (CRS positions), row after row
  std::vector<int64_t> lvOff;    // R+1 offsets into lvEnts
  for (int32_t gi = 0; gi < G; gi++) {
    const int32_t ta = groups[(size_t)gi], tb = groups[(size_t)gi + 1];
    const int64_t g0 = m.tileStart[(size_t)ta];
    const int64_t ng = m.tileStart[(size_t)tb] - g0;
    if (ng + 1 > 65535) {
      ok = false;
      return;
    }
    d.slabOff[(size_t)gi] = (int32_t)slab;
    d.lvlOff[(size_t)gi] = (int32_t)d.slabEnd.size();
    int32_t nL = 0;
    for (int32_t t = ta; t < tb; t++)
      nL = std::max<int32_t>(
          nL, (int32_t)(startOff[(size_t)t + 1] - startOff[(size_t)t] - 1));
    d.maxLevels = std::max(d.maxLevels, nL);
    int64_t groupSlab = 0;
    int64_t slot = 0;  // row slot within the group, level order
    for (int32_t L = 0; L < nL; L++) {
      lvRows.clear();
      lvEnts.clear();
      lvOff.assign(1, 0);
      for (int32_t t = ta; t < tb; t++) {
        const int64_t t0 = m.tileStart[(size_t)t], t1 = m.tileStart[(size_t)t + 1];
        const int64_t* ls = starts.data() + startOff[(size_t)t];
        const int32_t tL = (int32_t)(startOff[(size_t)t + 1] - startOff[(size_t)t] - 1);
        if (L >= tL) continue;
        for (int64_t q = ls[L]; q < ls[L + 1]; q++) {
          const int64_t i = t0 + rows[(size_t)(t0 + q)];
          const int64_t a0 = m.rowPtr[(size_t)i], a1 = m.rowPtr[(size_t)i + 1];
          if (!backward) {
            for (int64_t a = a0; a < a1; a++) {
              const int64_t j = m.col[(size_t)a];
              if (i < j || j < t0 || j >= t1) break;
              lvEnts.push_back((int32_t)a);
            }
          } else {
            for (int64_t a = a1 - 1; a >= a0; a--) {
              const int64_t j = m.col[(size_t)a];
              if (j < t0 || j >= t1) continue;
              if (j < i) break;
              lvEnts.push_back((int32_t)a);
            }
          }
          lvRows.push_back(i);
          lvOff.push_back((int64_t)lvEnts.size());
        }
      }
      const int64_t R = (int64_t)lvRows.size();
      int64_t W = 0;
      for (int64_t r = 0; r < R; r++)
        W = std::max<int64_t>(W, lvOff[(size_t)r + 1] - lvOff[(size_t)r]);
      d.rowStart.push_back((int32_t)slot);
      d.width.push_back((int32_t)W);
      for (int64_t r = 0; r < R; r++)
        d.rowLoc[(size_t)(g0 + slot + r)] = (uint16_t)(lvRows[(size_t)r] - g0);
      slot += R;
      // entry j of row r at position j*R + r; short rows padded with zero
      // entries (value 0 at the zero slot ng), which subtract +0
      for (int64_t j = 0; j < W; j++)
        for (int64_t r = 0; r < R; r++) {
          if (j < lvOff[(size_t)r + 1] - lvOff[(size_t)r]) {
            const int32_t a = lvEnts[(size_t)(lvOff[(size_t)r] + j)];
            d.col.push_back((uint16_t)(m.col[(size_t)a] - g0));
            d.src.push_back(a);
          } else {
            d.col.push_back((uint16_t)ng);
            d.src.push_back(-1);
          }
        }
      const int64_t e = W * R;
      d.maxLevelEntries = std::max<int32_t>(d.maxLevelEntries, (int32_t)e);
      const int64_t ns = (e + kPackSlab - 1) / kPackSlab;
      for (int64_t p = e; p < ns * kPackSlab; p++) {  // pad the level
        d.col.push_back((uint16_t)ng);
        d.src.push_back(-1);
      }
      groupSlab += ns;
      d.slabEnd.push_back((int32_t)groupSlab);
    }
    if (slot != ng) {
      ok = false;
      return;
    }
    d.rowStart.push_back((int32_t)ng);
    d.width.push_back(0);
    d.slabEnd.push_back(INT32_MAX);  // sentinel: the tables hold nL+1 entries
    if (groupSlab % 2) {  // streams start on an even slab (pairSlabs)
      for (int64_t p = 0; p < kPackSlab; p++) {
        d.col.push_back((uint16_t)ng);
        d.src.push_back(-1);
      }
      groupSlab++;
    }
    // The prefetch ring reads up to two depths past a group's last slab;
    // those reads land in the next group's stream (values unused) or in the
    // padding after the last group, so no bounds test is needed in the kernel.
    slab += groupSlab;
    if ((slab + 2 * kPackDepth) * kPackSlab > INT32_MAX) {
      ok = false;
      return;
    }
  }
  for (int64_t p = 0; p < 2 * kPackDepth * kPackSlab; p++) {
    d.col.push_back(0);
    d.src.push_back(-1);
  }
  d.slabOff[(size_t)G] = (int32_t)slab;
  d.lvlOff[(size_t)G] = (int32_t)d.slabEnd.size();
}

PackedILU packILU(const host::HostMatrix& m, int64_t targetRows) {
  PackedILU p;
  p.ok = true;
  std::vector<int32_t> groups = groupTiles(m, targetRows);
  p.groupStart.resize(groups.size());
  for (size_t q = 0; q < groups.size(); q++) {
    p.groupStart[q] = (int32_t)m.tileStart[(size_t)groups[q]];
    if (q > 0)
      p.maxRows = std::max(p.maxRows, p.groupStart[q] - p.groupStart[q - 1]);
  }
  packDirection(m, groups, false, p.fwd, p.ok);
  if (p.ok) packDirection(m, groups, true, p.bwd, p.ok);
  return p;
}

}  // namespace graphene::matrix
