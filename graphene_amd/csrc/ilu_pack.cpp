// Level-major packed ILU(0) layout (see ilu_pack.hpp).
#include "ilu_pack.hpp"

#include <algorithm>
#include <climits>

namespace graphene::matrix {

static void packDirection(const host::HostMatrix& m, bool backward,
                          PackedDirection& d, bool& ok) {
  const int32_t T = m.numTiles;
  d.slabOff.assign((size_t)T + 1, 0);
  d.lvlOff.assign((size_t)T + 1, 0);
  d.rowInfo.assign((size_t)m.nrows * 2, 0);
  const std::vector<int32_t>& rows = backward ? m.upperLevelRows : m.colorSortAddr;
  const std::vector<int64_t>& starts = backward ? m.upperColorStart : m.colorStart;
  const std::vector<int64_t>& startOff =
      backward ? m.upperColorStartOffset : m.colorStartOffset;
  int64_t slab = 0;
  std::vector<int32_t> ents;
  for (int32_t t = 0; t < T; t++) {
    const int64_t t0 = m.tileStart[(size_t)t], t1 = m.tileStart[(size_t)t + 1];
    const int64_t nt = t1 - t0;
    if (nt + 1 > 65535) {
      ok = false;
      return;
    }
    d.slabOff[(size_t)t] = (int32_t)slab;
    d.lvlOff[(size_t)t] = (int32_t)d.slabEnd.size();
    const int64_t* ls = starts.data() + startOff[(size_t)t];
    const int32_t nL = (int32_t)(startOff[(size_t)t + 1] - startOff[(size_t)t] - 1);
    d.maxLevels = std::max(d.maxLevels, nL);
    int64_t tileSlab = 0;
    for (int32_t L = 0; L < nL; L++) {
      d.rowStart.push_back((int32_t)ls[L]);
      int64_t e = 0;  // entries of this level
      for (int64_t q = ls[L]; q < ls[L + 1]; q++) {
        const int64_t li = rows[(size_t)(t0 + q)];
        const int64_t i = t0 + li;
        ents.clear();
        const int64_t a0 = m.rowPtr[(size_t)i], a1 = m.rowPtr[(size_t)i + 1];
        if (!backward) {
          for (int64_t a = a0; a < a1; a++) {
            const int64_t j = m.col[(size_t)a];
            if (i < j || j < t0 || j >= t1) break;
            ents.push_back((int32_t)a);
          }
        } else {
          for (int64_t a = a1 - 1; a >= a0; a--) {
            const int64_t j = m.col[(size_t)a];
            if (j < t0 || j >= t1) continue;
            if (j < i) break;
            ents.push_back((int32_t)a);
          }
        }
        const int64_t len = ents.empty() ? 1 : (int64_t)ents.size();
        // row slot q (level order) holds this row's descriptor
        d.rowInfo[(size_t)(t0 + q) * 2] = (uint32_t)li | ((uint32_t)len << 16);
        d.rowInfo[(size_t)(t0 + q) * 2 + 1] = (uint32_t)e;
        if (ents.empty()) {
          d.col.push_back((uint16_t)nt);
          d.src.push_back(-1);
        } else {
          for (int32_t a : ents) {
            d.col.push_back((uint16_t)(m.col[(size_t)a] - t0));
            d.src.push_back(a);
          }
        }
        e += len;
      }
      d.maxLevelEntries = std::max<int32_t>(d.maxLevelEntries, (int32_t)e);
      const int64_t ns = (e + kPackSlab - 1) / kPackSlab;
      for (int64_t p = e; p < ns * kPackSlab; p++) {  // pad the level
        d.col.push_back((uint16_t)nt);
        d.src.push_back(-1);
      }
      tileSlab += ns;
      d.slabEnd.push_back((int32_t)tileSlab);
    }
    d.rowStart.push_back((int32_t)nt);
    d.slabEnd.push_back(INT32_MAX);  // sentinel: both tables hold nL+1 entries
    // The prefetch ring reads up to two depths past a tile's last slab; those
    // reads land in the next tile's stream (values unused) or in the padding
    // after the last tile, so no bounds test is needed in the kernel.
    slab += tileSlab;
    if ((slab + 2 * kPackDepth) * kPackSlab > INT32_MAX) {
      ok = false;
      return;
    }
  }
  for (int64_t p = 0; p < 2 * kPackDepth * kPackSlab; p++) {
    d.col.push_back(0);
    d.src.push_back(-1);
  }
  d.slabOff[(size_t)T] = (int32_t)slab;
  d.lvlOff[(size_t)T] = (int32_t)d.slabEnd.size();
}

PackedILU packILU(const host::HostMatrix& m) {
  PackedILU p;
  p.ok = true;
  for (int32_t t = 0; t < m.numTiles; t++)
    p.maxRows = std::max<int32_t>(p.maxRows, (int32_t)m.tileRows(t));
  packDirection(m, false, p.fwd, p.ok);
  if (p.ok) packDirection(m, true, p.bwd, p.ok);
  return p;
}

}  // namespace graphene::matrix
