// Jagged-diagonal level stream for the Gauss-Seidel sweep (gs_pack.hpp).
#include "gs_pack.hpp"

#include <algorithm>
#include <climits>

namespace graphene::matrix {

GsStream packGs(const host::HostMatrix& m, bool backward, int64_t maxLevelEntries) {
  GsStream o;
  const int32_t T = m.numTiles;
  o.lvlBase.assign((size_t)T + 1, 0);
  for (int32_t t = 0; t < T; t++) o.maxRows = std::max<int32_t>(o.maxRows, (int32_t)m.tileRows(t));
  if (o.maxRows >= (1 << 24)) return o;
  o.code.reserve((size_t)m.nnzOff());
  o.src.reserve((size_t)m.nnzOff());
  std::vector<int32_t> all, order, perm, ts;
  std::vector<uint32_t> tc;
  o.pos.reserve((size_t)m.nnzOff());
  for (int32_t t = 0; t < T; t++) {
    o.lvlBase[(size_t)t] = (int32_t)(o.lvl.size() / 4);
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
        if (W > 255) return o;
        if ((int64_t)o.slot.size() + R > INT32_MAX ||
            (int64_t)o.code.size() + W * R > INT32_MAX)
          return o;
        o.lvl.insert(o.lvl.end(), {(int32_t)o.slot.size(), (int32_t)R, (int32_t)o.diag.size(),
                                   (int32_t)W});
        for (int64_t q = 0; q < R; q++)
          o.slot.push_back(((uint32_t)order[(size_t)q] << 8) | (uint32_t)len(order[(size_t)q]));
        for (int64_t j = 0; j < W; j++) {
          o.diag.push_back((int32_t)o.code.size());
          for (int64_t q = 0; q < R; q++) {
            const int32_t li = order[(size_t)q];
            if (len(li) <= j) break;  // sorted: the rows after are shorter too
            const int64_t a = m.rowPtr[(size_t)(t0 + li)] + j;
            const int32_t cl = m.colLocal[(size_t)a];
            const bool fresh = cl < nt && (backward ? cl > li : cl < li);
            o.code.push_back(fresh ? (kGsFresh | (uint32_t)cl) : (uint32_t)m.col[(size_t)a]);
            o.src.push_back((int32_t)a);
          }
        }
        o.diag.push_back((int32_t)o.code.size());
        const int64_t e0 = o.diag[(size_t)o.lvl[o.lvl.size() - 2]], ne = (int64_t)o.code.size() - e0;
        if (ne > 65536) return o;
        o.maxLevelEntries = std::max<int32_t>(o.maxLevelEntries, (int32_t)ne);
        // the level's entries in gather order: previous-x columns ascending,
        // then the values this sweep wrote, by row; pos = jagged-diagonal slot
        perm.resize((size_t)ne);
        for (int64_t k = 0; k < ne; k++) perm[(size_t)k] = (int32_t)k;
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
          const uint32_t ca = o.code[(size_t)(e0 + a)], cb = o.code[(size_t)(e0 + b)];
          const bool fa = (ca & kGsFresh) != 0, fb = (cb & kGsFresh) != 0;
          return fa != fb ? fb : ca < cb;
        });
        tc.resize((size_t)ne);
        ts.resize((size_t)ne);
        for (int64_t k = 0; k < ne; k++) {
          tc[(size_t)k] = o.code[(size_t)(e0 + perm[(size_t)k])];
          ts[(size_t)k] = o.src[(size_t)(e0 + perm[(size_t)k])];
          o.pos.push_back((uint16_t)perm[(size_t)k]);
        }
        std::copy(tc.begin(), tc.end(), o.code.begin() + e0);
        std::copy(ts.begin(), ts.end(), o.src.begin() + e0);
      }
    }
  }
  o.lvlBase[(size_t)T] = (int32_t)(o.lvl.size() / 4);
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
