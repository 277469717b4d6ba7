// Level-major packed ILU(0) layout (see ilu_pack.hpp).
#include "ilu_pack.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <cstdio>

#include "parallel.hpp"

namespace graphene::matrix {

// Consecutive tiles are grouped until a group holds about `targetRows`
// rows; a group is one wave's unit of work. Tiles are independent blocks,
// so level L of a group is the union of its tiles' level-L rows and the
// group's level count is the largest tile level count.
std::vector<int32_t> groupTiles(const host::HostMatrix& m, int64_t targetRows) {
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
  // groups in parts of consecutive groups, packed in parallel (a group's
  // tables are relative to its own stream), then concatenated
  struct Part {
    std::vector<uint16_t> col;
    std::vector<int32_t> src, slabEnd, rowStart, width;
    std::vector<int64_t> slabs, levels;  // per group of the part
    int32_t maxLevels = 0, maxLevelEntries = 0;
    bool ok = true;
  };
  constexpr int32_t kGroupsPerPart = 8;
  const int32_t nParts = (G + kGroupsPerPart - 1) / kGroupsPerPart;
  std::vector<Part> parts((size_t)nParts);
  gph::parallelFor(nParts, 1, [&](int64_t pi) {
    Part& P = parts[(size_t)pi];
    std::vector<int64_t> lvRows;   // rows of the level (matrix ids)
    std::vector<int32_t> lvEnts;   // their entries (CRS positions), row after row
    std::vector<int64_t> lvOff;    // R+1 offsets into lvEnts
    for (int32_t gi = (int32_t)pi * kGroupsPerPart;
         gi < std::min(G, (int32_t)(pi + 1) * kGroupsPerPart) && P.ok; gi++) {
      const int32_t ta = groups[(size_t)gi], tb = groups[(size_t)gi + 1];
      const int64_t g0 = m.tileStart[(size_t)ta];
      const int64_t ng = m.tileStart[(size_t)tb] - g0;
      if (ng + 1 > 65535) {
        P.ok = false;
        break;
      }
      const size_t lv0 = P.slabEnd.size();
      int32_t nL = 0;
      for (int32_t t = ta; t < tb; t++)
        nL = std::max<int32_t>(
            nL, (int32_t)(startOff[(size_t)t + 1] - startOff[(size_t)t] - 1));
      P.maxLevels = std::max(P.maxLevels, nL);
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
        P.rowStart.push_back((int32_t)slot);
        P.width.push_back((int32_t)W);
        for (int64_t r = 0; r < R; r++)
          d.rowLoc[(size_t)(g0 + slot + r)] = (uint16_t)(lvRows[(size_t)r] - g0);
        slot += R;
        // entry j of row r at position j*R + r; short rows padded with zero
        // entries (value 0 at the zero slot ng), which subtract +0
        for (int64_t j = 0; j < W; j++)
          for (int64_t r = 0; r < R; r++) {
            if (j < lvOff[(size_t)r + 1] - lvOff[(size_t)r]) {
              const int32_t a = lvEnts[(size_t)(lvOff[(size_t)r] + j)];
              P.col.push_back((uint16_t)(m.col[(size_t)a] - g0));
              P.src.push_back(a);
            } else {
              P.col.push_back((uint16_t)ng);
              P.src.push_back(-1);
            }
          }
        const int64_t e = W * R;
        P.maxLevelEntries = std::max<int32_t>(P.maxLevelEntries, (int32_t)e);
        const int64_t ns = (e + kPackSlab - 1) / kPackSlab;
        for (int64_t p = e; p < ns * kPackSlab; p++) {  // pad the level
          P.col.push_back((uint16_t)ng);
          P.src.push_back(-1);
        }
        groupSlab += ns;
        P.slabEnd.push_back((int32_t)groupSlab);
      }
      if (slot != ng) {
        P.ok = false;
        break;
      }
      P.rowStart.push_back((int32_t)ng);
      P.width.push_back(0);
      P.slabEnd.push_back(INT32_MAX);  // sentinel: the tables hold nL+1 entries
      if (groupSlab % 2) {  // streams start on an even slab (pairSlabs)
        for (int64_t p = 0; p < kPackSlab; p++) {
          P.col.push_back((uint16_t)ng);
          P.src.push_back(-1);
        }
        groupSlab++;
      }
      P.slabs.push_back(groupSlab);
      P.levels.push_back((int64_t)(P.slabEnd.size() - lv0));
    }
  });
  int64_t slab = 0, lvl = 0;
  for (int32_t pi = 0; pi < nParts; pi++) {
    const Part& P = parts[(size_t)pi];
    if (!P.ok) {
      ok = false;
      return;
    }
    d.maxLevels = std::max(d.maxLevels, P.maxLevels);
    d.maxLevelEntries = std::max(d.maxLevelEntries, P.maxLevelEntries);
    for (size_t q = 0; q < P.slabs.size(); q++) {
      const size_t gi = (size_t)pi * kGroupsPerPart + q;
      d.slabOff[gi] = (int32_t)slab;
      d.lvlOff[gi] = (int32_t)lvl;
      slab += P.slabs[q];
      lvl += P.levels[q];
      // The prefetch ring reads up to two depths past a group's last slab;
      // those reads land in the next group's stream (values unused) or in the
      // padding after the last group, so no bounds test is needed in the kernel.
      if ((slab + 2 * kPackDepth) * kPackSlab > INT32_MAX) {
        ok = false;
        return;
      }
    }
  }
  for (const Part& P : parts) {
    d.col.insert(d.col.end(), P.col.begin(), P.col.end());
    d.src.insert(d.src.end(), P.src.begin(), P.src.end());
    d.slabEnd.insert(d.slabEnd.end(), P.slabEnd.begin(), P.slabEnd.end());
    d.rowStart.insert(d.rowStart.end(), P.rowStart.begin(), P.rowStart.end());
    d.width.insert(d.width.end(), P.width.begin(), P.width.end());
  }
  for (int64_t p = 0; p < 2 * kPackDepth * kPackSlab; p++) {
    d.col.push_back(0);
    d.src.push_back(-1);
  }
  d.slabOff[(size_t)G] = (int32_t)slab;
  d.lvlOff[(size_t)G] = (int32_t)d.slabEnd.size();
}

// Rows of level L of a group (matrix ids, level order) and their entries
// (CRS positions, in substitution order), row after row.
static void groupLevel(const host::HostMatrix& m, int32_t ta, int32_t tb, int32_t L,
                       bool backward, std::vector<int64_t>& lvRows,
                       std::vector<int32_t>& lvEnts, std::vector<int64_t>& lvOff) {
  const std::vector<int32_t>& rows = backward ? m.upperLevelRows : m.colorSortAddr;
  const std::vector<int64_t>& starts = backward ? m.upperColorStart : m.colorStart;
  const std::vector<int64_t>& startOff =
      backward ? m.upperColorStartOffset : m.colorStartOffset;
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
}

static int32_t groupLevels(const host::HostMatrix& m, int32_t ta, int32_t tb, bool backward) {
  const std::vector<int64_t>& startOff =
      backward ? m.upperColorStartOffset : m.colorStartOffset;
  int32_t nL = 0;
  for (int32_t t = ta; t < tb; t++)
    nL = std::max<int32_t>(nL, (int32_t)(startOff[(size_t)t + 1] - startOff[(size_t)t] - 1));
  return nL;
}

IluLevelStream packIluLevels(const host::HostMatrix& m, int64_t targetRows,
                             const void* lu, const void* invD, int es, bool diluFwd,
                             int64_t maxStepBytes, int64_t ring, int64_t tail, int64_t wAlign) {
  IluLevelStream o;
  o.ok = true;
  const std::vector<int32_t> groups = groupTiles(m, targetRows);
  const int32_t G = (int32_t)groups.size() - 1;
  const int32_t P = (G + 1) / 2;
  o.groupStart.resize((size_t)G + 1);
  for (int32_t q = 0; q <= G; q++) {
    o.groupStart[(size_t)q] = (int32_t)m.tileStart[(size_t)groups[(size_t)q]];
    if (q > 0)
      o.maxRows = std::max(o.maxRows, o.groupStart[(size_t)q] - o.groupStart[(size_t)q - 1]);
  }
  if (o.maxRows + 1 > 65535) {
    o.ok = false;
    return o;
  }
  const uint16_t zero = (uint16_t)o.maxRows;  // zero slot of every half
  o.pairOff.assign((size_t)P + 1, 0);
  o.steps.assign((size_t)2 * P, 0);
  const uint8_t* luB = static_cast<const uint8_t*>(lu);
  const uint8_t* invB = static_cast<const uint8_t*>(invD);
  auto pad16 = [](int64_t b) { return (b + 15) & ~(int64_t)15; };
  auto put = [&](size_t at, const void* src, size_t n) {
    std::memcpy(o.bytes.data() + at, src, n);
  };
  // one half's current level: rows (matrix ids) and entries, and the next
  // row of it to place
  struct Half {
    int32_t ta = 0, tb = 0;
    int64_t g0 = 0;
    int32_t nL = 0, L = -1;
    int64_t r0 = 0;
    std::vector<int64_t> rows, off;
    std::vector<int32_t> ents;
    int64_t R() const { return (int64_t)rows.size(); }
    int64_t len(int64_t r) const { return off[(size_t)r + 1] - off[(size_t)r]; }
  };
  for (int32_t pi = 0; pi < P; pi++) {
    o.pairOff[(size_t)pi] = (int64_t)o.bytes.size();
    for (int dir = 0; dir < 2; dir++) {
      const bool backward = dir == 1;
      const bool withInv = backward || diluFwd;
      Half H[2];
      for (int h = 0; h < 2; h++) {
        const int32_t gi = 2 * pi + h;
        if (gi >= G) continue;
        H[h].ta = groups[(size_t)gi];
        H[h].tb = groups[(size_t)gi + 1];
        H[h].g0 = m.tileStart[(size_t)H[h].ta];
        H[h].nL = groupLevels(m, H[h].ta, H[h].tb, backward);
      }
      auto advance = [&](Half& x) {  // next level with rows left, if any
        while (x.r0 >= x.R() && x.L + 1 < x.nL) {
          x.L++;
          x.r0 = 0;
          groupLevel(m, x.ta, x.tb, x.L, backward, x.rows, x.ents, x.off);
        }
        return x.r0 < x.R();
      };
      int64_t placed[2] = {0, 0};
      for (;;) {
        const bool has0 = advance(H[0]), has1 = advance(H[1]);
        if (!has0 && !has1) break;
        // rows of this step per half (up to 32 each), the same count slot
        // range for both halves; W = longest row of both, rounded up to even
        int64_t Rs = 32;
        int64_t cnt[2], W = 0, bytes = 0;
        auto sizeOf = [&](int64_t rs) {
          W = 0;
          for (int h = 0; h < 2; h++) {
            cnt[h] = (h == 0 ? has0 : has1) ? std::min<int64_t>(rs, H[h].R() - H[h].r0) : 0;
            for (int64_t r = 0; r < cnt[h]; r++) W = std::max(W, H[h].len(H[h].r0 + r));
          }
          W = (W + wAlign - 1) / wAlign * wAlign;
          const int64_t rm = std::max(cnt[0], cnt[1]);
          return 16 + pad16(2 * 2 * rm) + (withInv ? pad16(es * 2 * rm) : 0) +
                 pad16(2 * 2 * rm * W) + pad16(es * 2 * rm * W);
        };
        bytes = sizeOf(Rs);
        while (bytes > maxStepBytes && Rs > 1) bytes = sizeOf(Rs = (Rs + 1) / 2);
        if (bytes > maxStepBytes || bytes > ring || W > 65534) {
          o.ok = false;
          return o;
        }
        const int64_t pos = ((int64_t)o.bytes.size() - o.pairOff[(size_t)pi]) % ring;
        if (pos + bytes > ring) {
          while (Rs > 1 && pos + bytes > ring) bytes = sizeOf(--Rs);
          if (pos + bytes > ring) {  // not even one row fits: filler record
            const size_t at = o.bytes.size();
            o.bytes.resize(at + (size_t)(ring - pos), 0);
            const uint32_t fb = (uint32_t)(ring - pos);
            put(at + 4, &fb, 4);
            o.steps[(size_t)(2 * pi + dir)]++;
            continue;
          }
        }
        const int64_t rm = std::max(cnt[0], cnt[1]);
        const size_t at = o.bytes.size();
        o.bytes.resize(at + (size_t)bytes, 0);
        const uint16_t hR = (uint16_t)rm, hW = (uint16_t)W;
        const uint32_t hB = (uint32_t)bytes, hF = withInv ? 1u : 0u;
        put(at, &hR, 2);
        put(at + 2, &hW, 2);
        put(at + 4, &hB, 4);
        put(at + 8, &hF, 4);
        const size_t qLoc = at + 16;
        const size_t qInv = qLoc + (size_t)pad16(4 * rm);
        const size_t qc = qInv + (withInv ? (size_t)pad16(es * 2 * rm) : 0);
        const size_t qv = qc + (size_t)pad16(4 * rm * W);
        for (int h = 0; h < 2; h++)
          for (int64_t r = 0; r < rm; r++) {
            const size_t slot = (size_t)(h * rm + r);
            uint16_t loc = zero;  // empty slot: zero row, zero entries
            if (r < cnt[h]) {
              const Half& x = H[h];
              const int64_t row = x.r0 + r;
              loc = (uint16_t)(x.rows[(size_t)row] - x.g0);
              if (withInv)
                put(qInv + (size_t)es * slot, invB + (size_t)es * (size_t)x.rows[(size_t)row],
                    (size_t)es);
              for (int64_t j = 0; j < x.len(row); j++) {
                const int32_t a = x.ents[(size_t)(x.off[(size_t)row] + j)];
                const uint16_t c = (uint16_t)(m.col[(size_t)a] - x.g0);
                put(qc + 2 * (slot * (size_t)W + (size_t)j), &c, 2);
                put(qv + (size_t)es * (slot * (size_t)W + (size_t)j), luB + (size_t)es * (size_t)a,
                    (size_t)es);
              }
              for (int64_t j = x.len(row); j < W; j++)  // zero products
                put(qc + 2 * (slot * (size_t)W + (size_t)j), &zero, 2);
            } else {
              for (int64_t j = 0; j < W; j++) put(qc + 2 * (slot * (size_t)W + (size_t)j), &zero, 2);
            }
            put(qLoc + 2 * slot, &loc, 2);
          }
        for (int h = 0; h < 2; h++) {
          H[h].r0 += cnt[h];
          placed[h] += cnt[h];
        }
        o.steps[(size_t)(2 * pi + dir)]++;
      }
      for (int h = 0; h < 2; h++) {
        const int32_t gi = 2 * pi + h;
        const int64_t ng = gi < G ? o.groupStart[(size_t)gi + 1] - o.groupStart[(size_t)gi] : 0;
        if (placed[h] != ng) {
          o.ok = false;
          return o;
        }
      }
    }
  }
  o.pairOff[(size_t)P] = (int64_t)o.bytes.size();
  o.bytes.resize(o.bytes.size() + (size_t)pad16(tail), 0);
  return o;
}

IluRowStream packIluRows(const host::HostMatrix& m, const void* lu, int es, bool dense) {
  IluRowStream o;
  o.dense = dense && es != 8;
  const int32_t T = m.numTiles;
  for (int32_t t = 0; t < T; t++) o.maxRows = std::max<int32_t>(o.maxRows, (int32_t)m.tileRows(t));
  if (4 * (int64_t)(o.maxRows + 1) > 65535) return o;
  const uint16_t zero = 0;  // slot 0 holds +0; row i of a tile is slot i + 1
  const uint8_t* luB = static_cast<const uint8_t*>(lu);
  const int64_t gran = o.dense ? 2 : kRowsPad;  // record length granule
  o.elemOff.assign((size_t)T + 1, 0);
  o.hdrOff.assign((size_t)T + 1, 0);
  o.metaOff.assign((size_t)T + 1, 0);
  // one record: row r's entries [e0, e0 + l) (r < 0: none); a column holds
  // one long record (rec[0], `lng`) or up to two short ones
  struct Rec {
    int32_t r = -1;
    int64_t e0 = 0, l = 0;
    uint32_t flags = 0;  // kRowsCont | kRowsPartial
  };
  struct Column {
    bool lng = false;
    Rec rec[2];
  };
  // tiles in parts of consecutive tiles, packed in parallel (a part's words
  // are relative to its step and record starts), then concatenated
  struct Part {
    std::vector<uint16_t> col;
    std::vector<uint8_t> val;
    std::vector<uint32_t> hdr, meta;
    std::vector<int64_t> elem0, hdr0, meta0;  // per tile of the part
  };
  constexpr int32_t kTilesPerPart = 16;
  const int32_t nParts = (T + kTilesPerPart - 1) / kTilesPerPart;
  std::vector<Part> parts((size_t)nParts);
  gph::parallelFor(nParts, 1, [&](int64_t pi) {
    Part& P = parts[(size_t)pi];
    std::vector<int64_t> rows, off;
    std::vector<int32_t> ents, order;
    std::vector<Column> step;
    for (int32_t t = (int32_t)pi * kTilesPerPart; t < std::min(T, (int32_t)(pi + 1) * kTilesPerPart);
         t++) {
      P.elem0.push_back((int64_t)P.col.size());
      P.hdr0.push_back((int64_t)P.hdr.size());
      P.meta0.push_back((int64_t)P.meta.size());
      const int64_t t0 = m.tileStart[(size_t)t];
      for (int dir = 0; dir < 2; dir++) {
        const bool backward = dir == 1;
        const int32_t nL = groupLevels(m, t, t + 1, backward);
        for (int32_t L = 0; L < nL; L++) {
          groupLevel(m, t, t + 1, L, backward, rows, ents, off);
          const int32_t R = (int32_t)rows.size();
          order.resize((size_t)R);
          for (int32_t r = 0; r < R; r++) order[(size_t)r] = r;
          auto len = [&](int32_t r) { return off[(size_t)r + 1] - off[(size_t)r]; };
          std::stable_sort(order.begin(), order.end(),
                           [&](int32_t a, int32_t b) { return len(a) > len(b); });
          int32_t nLong = 0;
          while (nLong < R && len(order[(size_t)nLong]) > kRowsShortLen) nLong++;
          int32_t next = nLong;  // next short row
          auto shortRec = [&]() {
            Rec c;
            if (next < R) {
              c.r = order[(size_t)next++];
              c.l = len(c.r);
            }
            return c;
          };
          // fills the free columns with short rows and emits the step
          auto emit = [&](std::vector<Column>& cols) {
            for (Column& c : cols)
              if (!c.lng && c.rec[0].r < 0) {
                c.rec[0] = shortRec();
                c.rec[1] = shortRec();
              }
            while ((int32_t)cols.size() < kRowsColumns && next < R) {
              Column c;
              c.rec[0] = shortRec();
              c.rec[1] = shortRec();
              cols.push_back(c);
            }
            while (!cols.empty() && !cols.back().lng && cols.back().rec[0].r < 0) cols.pop_back();
            // records in record-word order, each contiguous, padded to whole
            // chunks or (dense) to an even length; a long record's two words
            // share its entries
            bool anyLong = false;
            int64_t elems = 0;
            for (const Column& c : cols) {
              uint32_t w[2] = {0u, 0u};
              for (int hh = 0; hh < (c.lng ? 1 : 2); hh++) {
                const Rec& rc = c.rec[hh];
                if (rc.r < 0) continue;
                const int64_t stored = (rc.l + gran - 1) / gran * gran;
                const int64_t slot = rows[(size_t)rc.r] - t0 + 1;
                w[hh] = (uint32_t)elems | (uint32_t)stored << kRowsLenShift |
                        (c.lng ? kRowsLong : 0u) | rc.flags | (uint32_t)slot << kRowsSlotShift;
                for (int64_t e = 0; e < stored; e++) {
                  if (e < rc.l) {
                    const int32_t a = ents[(size_t)(off[(size_t)rc.r] + rc.e0 + e)];
                    P.col.push_back((uint16_t)(4 * (m.col[(size_t)a] - t0 + 1)));
                    P.val.insert(P.val.end(), luB + (size_t)a * es, luB + (size_t)(a + 1) * es);
                  } else {
                    P.col.push_back(zero);
                    P.val.insert(P.val.end(), (size_t)es, 0);
                  }
                }
                elems += stored;
              }
              if (c.lng) w[1] = w[0];
              P.meta.push_back(w[0]);
              P.meta.push_back(w[1]);
              anyLong = anyLong || c.lng;
            }
            P.hdr.push_back((uint32_t)elems | (uint32_t)(2 * cols.size()) << 12 |
                            (anyLong ? 1u << 30 : 0u) | (backward ? 1u << 31 : 0u));
            cols.clear();
          };
          // long rows, one column each, kRowsColumns at a time; a row longer
          // than kRowsMaxLen continues in the same column of the next steps
          for (int32_t c0 = 0; c0 < nLong; c0 += kRowsColumns) {
            const int32_t n = std::min<int32_t>(kRowsColumns, nLong - c0);
            const int64_t chunks = (len(order[(size_t)c0]) + kRowsMaxLen - 1) / kRowsMaxLen;
            for (int64_t k = 0; k < chunks; k++) {
              for (int32_t q = 0; q < n; q++) {
                const int32_t r = order[(size_t)(c0 + q)];
                const int64_t nk = (len(r) + kRowsMaxLen - 1) / kRowsMaxLen;
                Column c;
                if (k < nk) {
                  c.lng = true;
                  c.rec[0].r = r;
                  c.rec[0].e0 = k * kRowsMaxLen;
                  c.rec[0].l = std::min<int64_t>(kRowsMaxLen, len(r) - c.rec[0].e0);
                  c.rec[0].flags = (k > 0 ? kRowsCont : 0u) | (k + 1 < nk ? kRowsPartial : 0u);
                }
                step.push_back(c);
              }
              emit(step);
            }
          }
          while (next < R) emit(step);
        }
      }
    }
  });
  int64_t nCol = 0, nHdr = 0, nMeta = 0;
  for (const Part& P : parts) {
    for (size_t q = 0; q < P.elem0.size(); q++) {
      const size_t t = (size_t)(&P - parts.data()) * kTilesPerPart + q;
      o.elemOff[t] = (uint32_t)(nCol + P.elem0[q]);
      o.hdrOff[t] = (int32_t)(nHdr + P.hdr0[q]);
      o.metaOff[t] = nMeta + P.meta0[q];
      const int64_t h1 = q + 1 < P.hdr0.size() ? P.hdr0[q + 1] : (int64_t)P.hdr.size();
      const int64_t m1 = q + 1 < P.meta0.size() ? P.meta0[q + 1] : (int64_t)P.meta.size();
      o.maxSteps = std::max<int32_t>(o.maxSteps, (int32_t)(h1 - P.hdr0[q]));
      o.maxRecs = std::max<int64_t>(o.maxRecs, m1 - P.meta0[q]);
    }
    nCol += (int64_t)P.col.size();
    nHdr += (int64_t)P.hdr.size();
    nMeta += (int64_t)P.meta.size();
  }
  if (nCol > (int64_t)UINT32_MAX - 2 * kRowsTail || nHdr > INT32_MAX) return o;
  o.elemOff[(size_t)T] = (uint32_t)nCol;
  o.hdrOff[(size_t)T] = (int32_t)nHdr;
  o.metaOff[(size_t)T] = nMeta;
  o.col.resize((size_t)nCol + kRowsTail, zero);
  o.val.resize(((size_t)nCol + kRowsTail) * (size_t)es, 0);
  o.hdr.resize((size_t)nHdr);
  o.meta.resize((size_t)nMeta);
  gph::parallelFor(nParts, 1, [&](int64_t pi) {
    Part& P = parts[(size_t)pi];
    const size_t t0 = (size_t)pi * kTilesPerPart;
    std::copy(P.col.begin(), P.col.end(), o.col.begin() + o.elemOff[t0]);
    std::copy(P.val.begin(), P.val.end(), o.val.begin() + (size_t)o.elemOff[t0] * (size_t)es);
    std::copy(P.hdr.begin(), P.hdr.end(), o.hdr.begin() + o.hdrOff[t0]);
    std::copy(P.meta.begin(), P.meta.end(), o.meta.begin() + o.metaOff[t0]);
    Part().col.swap(P.col);
    Part().val.swap(P.val);
  });
  o.ok = true;
  return o;
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

// The workgroup-per-tile stream: the entries of every level piece in the
// substitution's order (pass 1), then the records at the common width.
namespace {
struct BlockPiece {
  int32_t tile = 0;
  int64_t q0 = 0, R = 0;  // rows [q0, q0 + R) of the tile's level order
  int32_t W = 0;
  std::vector<int64_t> entOff;  // R + 1
  std::vector<int64_t> ents;    // CRS positions
};
}  // namespace

static std::vector<BlockPiece> blockPieces(const host::HostMatrix& m, bool bwd, int32_t& maxW,
                                           std::vector<int32_t>& lvlBase) {
  std::vector<BlockPiece> out;
  const int32_t T = m.numTiles;
  lvlBase.assign((size_t)T + 1, 0);
  for (int32_t t = 0; t < T; t++) {
    lvlBase[(size_t)t] = (int32_t)out.size();
    const int64_t t0 = m.tileStart[(size_t)t], nt = m.tileRows(t);
    const int64_t* cs = bwd ? m.upperColorStart.data() + m.upperColorStartOffset[(size_t)t]
                            : m.colorStart.data() + m.colorStartOffset[(size_t)t];
    const int32_t nl = bwd ? m.numUpperColors[(size_t)t] : m.numColors[(size_t)t];
    const int32_t* rows = bwd ? m.upperLevelRows.data() : m.colorSortAddr.data();
    for (int32_t L = 0; L < nl; L++)
      for (int64_t q0 = cs[L]; q0 < cs[L + 1]; q0 += kBlockSlots) {
        BlockPiece p;
        p.tile = t;
        p.q0 = q0;
        p.R = std::min<int64_t>(kBlockSlots, cs[L + 1] - q0);
        p.entOff.push_back(0);
        for (int64_t q = 0; q < p.R; q++) {
          const int32_t li = rows[(size_t)(t0 + q0 + q)];
          const int64_t i = t0 + li;
          if (!bwd) {
            for (int64_t a = m.rowPtr[(size_t)i]; a < m.rowPtr[(size_t)i + 1]; a++) {
              const int32_t c = m.colLocal[(size_t)a];
              if (c >= nt || c > li) break;  // halo or upper part reached
              p.ents.push_back(a);
            }
          } else {
            for (int64_t a = m.rowPtr[(size_t)i + 1] - 1; a >= m.rowPtr[(size_t)i]; a--) {
              const int32_t c = m.colLocal[(size_t)a];
              if (c >= nt) continue;  // halo
              if (c < li) break;
              p.ents.push_back(a);
            }
          }
          p.entOff.push_back((int64_t)p.ents.size());
          p.W = std::max<int32_t>(p.W, (int32_t)(p.entOff[(size_t)q + 1] - p.entOff[(size_t)q]));
        }
        maxW = std::max(maxW, p.W);
        out.push_back(std::move(p));
      }
  }
  lvlBase[(size_t)T] = (int32_t)out.size();
  return out;
}

static bool emitBlockDirection(const host::HostMatrix& m, const std::vector<BlockPiece>& pieces,
                               const uint8_t* lu, int es, const uint8_t* inv, int ies, int width,
                               bool bwd, IluBlockDirection& d) {
  const size_t rs = ies == 8 ? 16 : 8;  // row record
  const size_t er = es == 8 ? 16 : 8;   // entry record
  const int32_t* rows = bwd ? m.upperLevelRows.data() : m.colorSortAddr.data();
  int64_t slots = 0, ents = 0;
  for (const BlockPiece& p : pieces) {
    if (slots + p.R > INT32_MAX || ents + (int64_t)width * p.R > INT32_MAX) return false;
    d.info.insert(d.info.end(), {(int32_t)slots, (int32_t)p.R, (int32_t)ents, p.W});
    slots += p.R;
    ents += (int64_t)width * p.R;
  }
  d.rows.assign((size_t)std::max<int64_t>(slots, 1) * rs, 0);
  d.ents.assign((size_t)std::max<int64_t>(ents, 1) * er, 0);
  for (size_t k = 0; k < pieces.size(); k++) {
    const BlockPiece& p = pieces[k];
    const int64_t t0 = m.tileStart[(size_t)p.tile];
    const uint32_t zero = (uint32_t)m.tileRows(p.tile);
    const int64_t slot0 = d.info[4 * k], ent0 = d.info[4 * k + 2];
    for (int64_t q = 0; q < p.R; q++) {
      const int32_t li = rows[(size_t)(t0 + p.q0 + q)];
      uint8_t* rr = &d.rows[(size_t)(slot0 + q) * rs];
      std::memcpy(rr, &li, 4);
      std::memcpy(rr + rs - (size_t)ies, inv + (size_t)(t0 + li) * (size_t)ies, (size_t)ies);
      const int64_t e0 = p.entOff[(size_t)q], n = p.entOff[(size_t)q + 1] - e0;
      for (int64_t j = 0; j < width; j++) {
        uint8_t* er_ = &d.ents[(size_t)(ent0 + j * p.R + q) * er];
        if (j < n) {
          const int64_t a = p.ents[(size_t)(e0 + j)];
          const uint32_t c = (uint32_t)m.colLocal[(size_t)a];
          std::memcpy(er_, &c, 4);
          std::memcpy(er_ + er - (size_t)(es == 8 ? 8 : 4), lu + (size_t)a * (size_t)es, (size_t)es);
        } else {
          std::memcpy(er_, &zero, 4);  // the zero slot, value +0
        }
      }
    }
  }
  if (d.info.empty()) d.info.assign(4, 0);
  return true;
}

IluBlockStream packIluBlock(const host::HostMatrix& m, const void* lu, int es, const void* invD,
                            int ies) {
  IluBlockStream o;
  for (int32_t t = 0; t < m.numTiles; t++)
    o.maxRows = std::max<int32_t>(o.maxRows, (int32_t)m.tileRows(t));
  int32_t maxW = 0;
  const std::vector<BlockPiece> f = blockPieces(m, false, maxW, o.fwd.lvlBase);
  const std::vector<BlockPiece> b = blockPieces(m, true, maxW, o.bwd.lvlBase);
  if (maxW > kBlockMaxWidth) return o;
  o.width = maxW <= 4 ? 4 : maxW <= 8 ? 8 : 16;
  const uint8_t* luB = static_cast<const uint8_t*>(lu);
  const uint8_t* invB = static_cast<const uint8_t*>(invD);
  o.ok = emitBlockDirection(m, f, luB, es, invB, ies, o.width, false, o.fwd) &&
         emitBlockDirection(m, b, luB, es, invB, ies, o.width, true, o.bwd);
  for (int32_t t = 0; o.ok && t < m.numTiles; t++)
    o.maxLevels = std::max<int32_t>(
        o.maxLevels, o.fwd.lvlBase[(size_t)t + 1] - o.fwd.lvlBase[(size_t)t] +
                         o.bwd.lvlBase[(size_t)t + 1] - o.bwd.lvlBase[(size_t)t]);
  return o;
}

}  // namespace graphene::matrix
