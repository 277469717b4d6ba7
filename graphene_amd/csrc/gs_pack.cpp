// Jagged-diagonal level stream for the Gauss-Seidel sweep (gs_pack.hpp).
#include "gs_pack.hpp"

#include <algorithm>
#include <climits>

#include "parallel.hpp"

namespace graphene::matrix {

GsStream packGs(const host::HostMatrix& m, bool backward, int64_t maxLevelEntries) {
  GsStream o;
  const int32_t T = m.numTiles;
  o.lvlBase.assign((size_t)T + 1, 0);
  for (int32_t t = 0; t < T; t++) o.maxRows = std::max<int32_t>(o.maxRows, (int32_t)m.tileRows(t));
  if (o.maxRows >= (1 << 24)) return o;
  // tiles in parts of consecutive tiles, packed in parallel with part-local
  // indices (first slot, first diagonal, diagonal starts), then concatenated
  // with the part's offsets added
  struct Part {
    std::vector<int32_t> lvl, diag, lvlCount;
    std::vector<uint32_t> slot, code;
    std::vector<int32_t> src;
    std::vector<uint16_t> pos;
    int32_t maxLevelEntries = 0;
    bool ok = true;
  };
  constexpr int32_t kTilesPerPart = 16;
  const int32_t nParts = (T + kTilesPerPart - 1) / kTilesPerPart;
  std::vector<Part> parts((size_t)nParts);
  gph::parallelFor(nParts, 1, [&](int64_t pi) {
    Part& P = parts[(size_t)pi];
    std::vector<int32_t> all, order, perm, ts;
    std::vector<uint32_t> tc;
    for (int32_t t = (int32_t)pi * kTilesPerPart; t < std::min(T, (int32_t)(pi + 1) * kTilesPerPart);
         t++) {
      const size_t lv0 = P.lvl.size();
      [&] {
        const int64_t t0 = m.tileStart[(size_t)t], nt = m.tileRows(t);
        const int64_t* cs = backward ? m.upperColorStart.data() + m.upperColorStartOffset[(size_t)t]
                                     : m.colorStart.data() + m.colorStartOffset[(size_t)t];
        const int32_t nl = backward ? m.numUpperColors[(size_t)t] : m.numColors[(size_t)t];
        const int32_t* rows = backward ? m.upperLevelRows.data() : m.colorSortAddr.data();
        auto len = [&](int32_t li) {
          return m.rowPtr[(size_t)(t0 + li) + 1] - m.rowPtr[(size_t)(t0 + li)];
        };
        for (int32_t L = 0; L < nl; L++) {
          all.assign(rows + t0 + cs[L], rows + t0 + cs[L + 1]);
          std::stable_sort(all.begin(), all.end(),
                           [&](int32_t a, int32_t b) { return len(a) > len(b); });
          // a level of more than maxLevelEntries entries is streamed as several
          // consecutive levels (its rows are independent of each other)
          for (size_t qa = 0, qb = 0; qa < all.size(); qa = qb) {
            int64_t acc = 0;
            for (qb = qa; qb < all.size() && (qb == qa || acc + len(all[qb]) <= maxLevelEntries); qb++)
              acc += len(all[qb]);
            order.assign(all.begin() + (ptrdiff_t)qa, all.begin() + (ptrdiff_t)qb);
            const int64_t R = (int64_t)order.size();
            const int64_t W = R ? len(order[0]) : 0;
            if (W > 255) {
              P.ok = false;
              return;
            }
            if ((int64_t)P.slot.size() + R > INT32_MAX ||
                (int64_t)P.code.size() + W * R > INT32_MAX) {
              P.ok = false;
              return;
            }
            P.lvl.insert(P.lvl.end(), {(int32_t)P.slot.size(), (int32_t)R, (int32_t)P.diag.size(),
                                       (int32_t)W});
            for (int64_t q = 0; q < R; q++)
              P.slot.push_back(((uint32_t)order[(size_t)q] << 8) | (uint32_t)len(order[(size_t)q]));
            for (int64_t j = 0; j < W; j++) {
              P.diag.push_back((int32_t)P.code.size());
              for (int64_t q = 0; q < R; q++) {
                const int32_t li = order[(size_t)q];
                if (len(li) <= j) break;  // sorted: the rows after are shorter too
                const int64_t a = m.rowPtr[(size_t)(t0 + li)] + j;
                const int32_t cl = m.colLocal[(size_t)a];
                const bool fresh = cl < nt && (backward ? cl > li : cl < li);
                P.code.push_back(fresh ? (kGsFresh | (uint32_t)cl) : (uint32_t)m.col[(size_t)a]);
                P.src.push_back((int32_t)a);
              }
            }
            P.diag.push_back((int32_t)P.code.size());
            const int64_t e0 = P.diag[(size_t)P.lvl[P.lvl.size() - 2]], ne = (int64_t)P.code.size() - e0;
            if (ne > 65536) {
              P.ok = false;
              return;
            }
            P.maxLevelEntries = std::max<int32_t>(P.maxLevelEntries, (int32_t)ne);
            // the level's entries in gather order: previous-x columns ascending,
            // then the values this sweep wrote, by row; pos = jagged-diagonal slot
            perm.resize((size_t)ne);
            for (int64_t k = 0; k < ne; k++) perm[(size_t)k] = (int32_t)k;
            std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
              const uint32_t ca = P.code[(size_t)(e0 + a)], cb = P.code[(size_t)(e0 + b)];
              const bool fa = (ca & kGsFresh) != 0, fb = (cb & kGsFresh) != 0;
              return fa != fb ? fb : ca < cb;
            });
            tc.resize((size_t)ne);
            ts.resize((size_t)ne);
            for (int64_t k = 0; k < ne; k++) {
              tc[(size_t)k] = P.code[(size_t)(e0 + perm[(size_t)k])];
              ts[(size_t)k] = P.src[(size_t)(e0 + perm[(size_t)k])];
              P.pos.push_back((uint16_t)perm[(size_t)k]);
            }
            std::copy(tc.begin(), tc.end(), P.code.begin() + e0);
            std::copy(ts.begin(), ts.end(), P.src.begin() + e0);
          }
        }
      }();
      if (!P.ok) return;
      P.lvlCount.push_back((int32_t)((P.lvl.size() - lv0) / 4));
    }
  });
  int64_t nLvl = 0, nSlot = 0, nDiag = 0, nCode = 0;
  for (int32_t pi = 0; pi < nParts; pi++) {
    const Part& P = parts[(size_t)pi];
    if (!P.ok) return o;
    for (size_t q = 0; q < P.lvlCount.size(); q++) {
      o.lvlBase[(size_t)pi * kTilesPerPart + q] = (int32_t)nLvl;
      nLvl += P.lvlCount[q];
    }
    o.maxLevelEntries = std::max(o.maxLevelEntries, P.maxLevelEntries);
    nSlot += (int64_t)P.slot.size();
    nDiag += (int64_t)P.diag.size();
    nCode += (int64_t)P.code.size();
  }
  if (nSlot > INT32_MAX || nDiag > INT32_MAX || nCode > INT32_MAX || 4 * nLvl > INT32_MAX)
    return o;
  o.lvlBase[(size_t)T] = (int32_t)nLvl;
  o.lvl.resize((size_t)(4 * nLvl));
  o.slot.resize((size_t)nSlot);
  o.diag.resize((size_t)nDiag);
  o.code.resize((size_t)nCode);
  o.src.resize((size_t)nCode);
  o.pos.resize((size_t)nCode);
  std::vector<int64_t> bLvl((size_t)nParts + 1, 0), bSlot((size_t)nParts + 1, 0),
      bDiag((size_t)nParts + 1, 0), bCode((size_t)nParts + 1, 0);
  for (int32_t pi = 0; pi < nParts; pi++) {
    const Part& P = parts[(size_t)pi];
    bLvl[(size_t)pi + 1] = bLvl[(size_t)pi] + (int64_t)P.lvl.size();
    bSlot[(size_t)pi + 1] = bSlot[(size_t)pi] + (int64_t)P.slot.size();
    bDiag[(size_t)pi + 1] = bDiag[(size_t)pi] + (int64_t)P.diag.size();
    bCode[(size_t)pi + 1] = bCode[(size_t)pi] + (int64_t)P.code.size();
  }
  gph::parallelFor(nParts, 1, [&](int64_t pi) {
    Part& P = parts[(size_t)pi];
    const int32_t sOff = (int32_t)bSlot[(size_t)pi], dOff = (int32_t)bDiag[(size_t)pi],
                  cOff = (int32_t)bCode[(size_t)pi];
    for (size_t q = 0; q < P.lvl.size(); q += 4) {
      int32_t* L = &o.lvl[(size_t)bLvl[(size_t)pi] + q];
      L[0] = P.lvl[q] + sOff;
      L[1] = P.lvl[q + 1];
      L[2] = P.lvl[q + 2] + dOff;
      L[3] = P.lvl[q + 3];
    }
    for (size_t q = 0; q < P.diag.size(); q++) o.diag[(size_t)dOff + q] = P.diag[q] + cOff;
    std::copy(P.slot.begin(), P.slot.end(), o.slot.begin() + sOff);
    std::copy(P.code.begin(), P.code.end(), o.code.begin() + cOff);
    std::copy(P.src.begin(), P.src.end(), o.src.begin() + cOff);
    std::copy(P.pos.begin(), P.pos.end(), o.pos.begin() + cOff);
  });
  if (o.lvl.empty()) o.lvl.assign(4, 0);
  if (o.diag.empty()) o.diag.assign(1, 0);
  if (o.slot.empty()) o.slot.assign(1, 0);
  if (o.code.empty()) {
    o.code.assign(1, kGsFresh);
    o.src.assign(1, -1);
    o.pos.assign(1, 0);
  }
  o.ok = true;
  return o;
}

}  // namespace graphene::matrix
