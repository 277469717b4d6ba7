// Device-resident matrix (see device_matrix.hpp).
#include "device_matrix.hpp"

#include <algorithm>
#include <cmath>

#include "parallel.hpp"

namespace graphene::matrix {

using gph::DevBuf;
using gph::DType;
using gph::fail;

static std::vector<int32_t> toI32(const std::vector<int64_t>& v) {
  std::vector<int32_t> o(v.size());
  for (size_t i = 0; i < v.size(); i++) {
    if (v[i] > INT32_MAX || v[i] < INT32_MIN)
      fail(gph::kInvalid, "index does not fit int32");
    o[i] = (int32_t)v[i];
  }
  return o;
}

template <typename T>
static void up(DevBuf& b, const std::vector<T>& v, hipStream_t s) {
  b.upload(v.data(), v.size(), s);
}

DeviceMatrix::DeviceMatrix(std::shared_ptr<const host::HostMatrix> host,
                           MatrixFormat fmt, int32_t rowsPerBlock)
    : host_(std::move(host)), fmt_(fmt) {
  const host::HostMatrix& h = *host_;
  n_ = (int32_t)h.nrows;
  T_ = h.numTiles;
  nnz_ = h.nnzOff();
  GPH_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  hipStream_t s = stream_;

  up(rowPtr_, toI32(h.rowPtr), s);
  up(col_, h.col, s);
  up(off32_, h.offDiag, s);
  up(diag32_, h.diag, s);

  std::vector<int64_t> rowLen((size_t)n_);
  for (int32_t r = 0; r < n_; r++) rowLen[(size_t)r] = h.rowPtr[(size_t)r + 1] - h.rowPtr[(size_t)r];
  buildSell(h.rowPtr, h.col, h.offDiag, sellCrs_);
  const std::vector<int32_t> br =
      buildRowBlocks(rowLen, rowsPerBlock, blockRowCrs_, numBlocksCrs_);
  // matrices that run in the row-block kernel by default (irregular rows,
  // far columns) gather x in column order (4M power-law rows, blocks of up
  // to 256 rows: 173 us; in stored order 237 us with 64-row blocks, the
  // best size there, 284 us with 256)
  if (fmt_ == MatrixFormat::CRS && !(sellCrs_.numSlices > 0 && sellCrs_.preferred))
    buildColumnOrder(br);

  if (fmt_ == MatrixFormat::LDU) {
    host::LDUHost L = host::buildLDU(h);
    numFaces_ = L.numFaces;
    std::vector<int32_t> lowerCol((size_t)L.numFaces);
    for (int64_t q = 0; q < L.numFaces; q++)
      lowerCol[(size_t)q] = L.lowerAddr[(size_t)L.neighbourColInd[(size_t)q]];
    up(ownerStart_, L.ownerStart, s);
    up(upperCol_, L.upperAddr, s);
    up(upper32_, L.upper, s);
    up(nbrStart_, L.neighbourStart, s);
    up(lowerCol_, lowerCol, s);
    up(lower32_, L.lower, s);
    std::vector<int64_t> len((size_t)n_);
    for (int32_t r = 0; r < n_; r++)
      len[(size_t)r] = (int64_t)(L.ownerStart[(size_t)r + 1] - L.ownerStart[(size_t)r]) +
                       (int64_t)(L.neighbourStart[(size_t)r + 1] - L.neighbourStart[(size_t)r]);
    buildRowBlocks(len, rowsPerBlock, blockRowLdu_, numBlocksLdu_);
    // per row: owned faces (upper, ascending), then losort faces (lower)
    std::vector<int64_t> rp((size_t)n_ + 1, 0);
    std::vector<int32_t> cc;
    std::vector<float> vv;
    cc.reserve((size_t)2 * L.numFaces);
    vv.reserve((size_t)2 * L.numFaces);
    for (int32_t r = 0; r < n_; r++) {
      for (int32_t f = L.ownerStart[(size_t)r]; f < L.ownerStart[(size_t)r + 1]; f++) {
        cc.push_back(L.upperAddr[(size_t)f]);
        vv.push_back(L.upper[(size_t)f]);
      }
      for (int32_t q = L.neighbourStart[(size_t)r]; q < L.neighbourStart[(size_t)r + 1]; q++) {
        cc.push_back(lowerCol[(size_t)q]);
        vv.push_back(L.lower[(size_t)q]);
      }
      rp[(size_t)r + 1] = (int64_t)cc.size();
    }
    buildSell(rp, cc, vv, sellLdu_);
  }

  // tile structure and level sets
  up(tileStart_, toI32(h.tileStart), s);
  up(lvlRows_, h.colorSortAddr, s);
  up(lvlPtr_, toI32(h.colorStart), s);
  up(lvlPtrOff_, toI32(h.colorStartOffset), s);
  up(ulvlRows_, h.upperLevelRows, s);
  up(ulvlPtr_, toI32(h.upperColorStart), s);
  up(ulvlPtrOff_, toI32(h.upperColorStartOffset), s);
  maxTileRows_ = 0;
  for (int32_t t = 0; t < T_; t++)
    maxTileRows_ = std::max<int32_t>(maxTileRows_, (int32_t)h.tileRows(t));

  scratchState_.alloc(sizeof(gph::DevState));
  GPH_HIP(hipMemsetAsync(scratchState_.get(), 0, sizeof(gph::DevState), s));
  int32_t maxBlocks = std::max({numBlocksCrs_, numBlocksLdu_, 4096, gph::kSellMaxGrid});
  partialsBytes_ = gph::kArriveBytes + (size_t)maxBlocks * 2 * sizeof(double);
  scratchPartials_.alloc(partialsBytes_);
  GPH_HIP(hipMemsetAsync(scratchPartials_.get(), 0, gph::kArriveBytes, s));
  GPH_HIP(hipStreamSynchronize(s));
}

DeviceMatrix::~DeviceMatrix() {
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

// Contiguous row blocks of at most 256 rows whose entries fit one staging
// chunk (kSpmvCap) where possible; a row longer than a chunk forms its own
// block and is streamed in several chunks.
std::vector<int32_t> DeviceMatrix::buildRowBlocks(const std::vector<int64_t>& rowLen,
                                                  int32_t rowsPerBlock, DevBuf& out,
                                                  int32_t& nb) {
  const int32_t maxRows =
      (rowsPerBlock > 0) ? std::min(rowsPerBlock, gph::kSpmvThreads)
                         : gph::kSpmvThreads;
  std::vector<int32_t> br;
  br.push_back(0);
  int64_t acc = 0;
  int32_t rows = 0;
  for (int32_t r = 0; r < n_; r++) {
    int64_t l = rowLen[(size_t)r];
    if (rows > 0 && (rows >= maxRows || acc + l > gph::kSpmvCap)) {
      br.push_back(r);
      acc = 0;
      rows = 0;
    }
    acc += l;
    rows++;
  }
  br.push_back(n_);
  nb = (int32_t)br.size() - 1;
  out.upload(br.data(), br.size(), stream_);
  return br;
}

// The CRS stream of every row-block chunk (the kSpmvCap-entry pieces the
// row-block kernel stages at a time) reordered by column, ties in stored
// order, with each entry's position in its chunk: the kernel gathers x in
// column order, so neighbouring lanes read neighbouring x, and writes each
// product to its stored position; the row sums are unchanged.
void DeviceMatrix::buildColumnOrder(const std::vector<int32_t>& br) {
  static_assert(gph::kSpmvCap <= 65536, "chunk positions are 16-bit");
  const host::HostMatrix& h = *host_;
  std::vector<int32_t> gc((size_t)nnz_);
  std::vector<uint16_t> gp((size_t)nnz_);
  std::vector<float> gv((size_t)nnz_);
  gph::parallelFor((int64_t)br.size() - 1, 64, [&](int64_t b) {
    std::vector<int32_t> idx;
    const int64_t e0 = h.rowPtr[(size_t)br[(size_t)b]], e1 = h.rowPtr[(size_t)br[(size_t)b + 1]];
    for (int64_t c0 = e0; c0 < e1; c0 += gph::kSpmvCap) {
      const int32_t cn = (int32_t)std::min<int64_t>(gph::kSpmvCap, e1 - c0);
      idx.resize((size_t)cn);
      for (int32_t k = 0; k < cn; k++) idx[(size_t)k] = k;
      std::stable_sort(idx.begin(), idx.end(), [&](int32_t p, int32_t q) {
        return h.col[(size_t)(c0 + p)] < h.col[(size_t)(c0 + q)];
      });
      for (int32_t k = 0; k < cn; k++) {
        const size_t src = (size_t)(c0 + idx[(size_t)k]);
        gc[(size_t)c0 + k] = h.col[src];
        gp[(size_t)c0 + k] = (uint16_t)idx[(size_t)k];
        gv[(size_t)c0 + k] = h.offDiag[src];
      }
    }
  });
  up(gcol_, gc, stream_);
  up(gpos_, gp, stream_);
  up(gval32_, gv, stream_);
}

// Sliced ELL: 64 rows per slice, slot (k, lane) = sliceOff[s] + 64k + lane.
// Rows keep their stored entry order; padding slots hold value 0 and a valid
// column. Column indices are compressed per slice: when a slice's columns
// span less than 2^16, they are stored as 16-bit offsets from sliceBase[s]
// (entry (k, lane) at 16-bit word sliceColW[s] + 64k + lane); otherwise as
// 32-bit columns (sliceBase[s] = -1, entry at word sliceColW[s] + 2(64k +
// lane)). Structured-grid matrices in tile order have most slices narrow. When row lengths vary strongly (max >
// 2x mean), rows are sorted by length inside windows of 4096 rows so that a
// slice holds rows of similar length (perm gives each lane's row). Such a
// matrix with more than 10 % wide slices runs in the row-block kernel unless
// the caller asks for the sliced form.
void DeviceMatrix::buildSell(const std::vector<int64_t>& rowPtr,
                             const std::vector<int32_t>& cols,
                             const std::vector<float>& vals, Sell& out) {
  const int64_t n = n_;
  int64_t maxLen = 0;
  for (int64_t r = 0; r < n; r++)
    maxLen = std::max(maxLen, rowPtr[(size_t)r + 1] - rowPtr[(size_t)r]);
  if (maxLen > 65535) return;  // row-block kernel only
  const double mean = n > 0 ? (double)rowPtr[(size_t)n] / (double)n : 0.0;
  const bool sorted = maxLen > 2.0 * mean + 1.0;
  const int64_t numSlices = (n + 63) / 64;
  std::vector<int32_t> perm((size_t)numSlices * 64, (int32_t)n);
  for (int64_t r = 0; r < n; r++) perm[(size_t)r] = (int32_t)r;
  if (sorted) {
    const int64_t W = 4096;
    gph::parallelFor((n + W - 1) / W, 16, [&](int64_t wi) {
      const int64_t w0 = wi * W, w1 = std::min(n, w0 + W);
      std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int32_t a, int32_t b) {
        return rowPtr[(size_t)a + 1] - rowPtr[(size_t)a] > rowPtr[(size_t)b + 1] - rowPtr[(size_t)b];
      });
    });
  }
  std::vector<int32_t> sliceOff((size_t)numSlices + 1, 0);
  std::vector<uint16_t> len((size_t)numSlices * 64, 0);
  // per slice (in parallel): row lengths, width, column range
  std::vector<int64_t> width((size_t)numSlices, 0);
  std::vector<int32_t> loC((size_t)numSlices, 0), hiC((size_t)numSlices, 0);
  gph::parallelFor(numSlices, 1024, [&](int64_t sl) {
    int64_t K = 0;
    int32_t lo = INT32_MAX, hi = -1;
    for (int l = 0; l < 64; l++) {
      int32_t r = perm[(size_t)(sl * 64 + l)];
      if (r >= n) continue;
      int64_t ln = rowPtr[(size_t)r + 1] - rowPtr[(size_t)r];
      len[(size_t)(sl * 64 + l)] = (uint16_t)ln;
      K = std::max(K, ln);
      for (int64_t e = rowPtr[(size_t)r]; e < rowPtr[(size_t)r + 1]; e++) {
        lo = std::min(lo, cols[(size_t)e]);
        hi = std::max(hi, cols[(size_t)e]);
      }
    }
    if (hi < 0) lo = hi = 0;  // empty slice
    width[(size_t)sl] = K;
    loC[(size_t)sl] = lo;
    hiC[(size_t)sl] = hi;
  });
  int64_t slots = 0;
  for (int64_t sl = 0; sl < numSlices; sl++) {
    sliceOff[(size_t)sl] = (int32_t)slots;
    slots += 64 * width[(size_t)sl];
    if (slots > INT32_MAX) return;
  }
  sliceOff[(size_t)numSlices] = (int32_t)slots;
  // column range per slice -> 16-bit or 32-bit column words
  std::vector<int32_t> base((size_t)numSlices, -1);
  std::vector<uint32_t> colW((size_t)numSlices + 1, 0);
  uint64_t words = 0;
  int64_t wide = 0;
  for (int64_t sl = 0; sl < numSlices; sl++) {
    const int32_t lo = loC[(size_t)sl], hi = hiC[(size_t)sl];
    const int64_t sw = sliceOff[(size_t)sl + 1] - sliceOff[(size_t)sl];
    colW[(size_t)sl] = (uint32_t)words;
    if ((int64_t)hi - lo < 65536) {
      base[(size_t)sl] = lo;
      words += (uint64_t)sw;
    } else {
      words += 2 * (uint64_t)sw;
      wide++;
    }
    if (words > UINT32_MAX) return;
  }
  colW[(size_t)numSlices] = (uint32_t)words;
  std::vector<uint16_t> c((size_t)std::max<uint64_t>(words, 2), 0);
  std::vector<float> v((size_t)std::max<int64_t>(slots, 1), 0.0f);
  gph::parallelFor(numSlices, 1024, [&](int64_t sl) {
    const int32_t b = base[(size_t)sl];
    for (int l = 0; l < 64; l++) {
      int32_t r = perm[(size_t)(sl * 64 + l)];
      const int64_t ln = r < n ? rowPtr[(size_t)r + 1] - rowPtr[(size_t)r] : 0;
      const int64_t K = (sliceOff[(size_t)sl + 1] - sliceOff[(size_t)sl]) / 64;
      for (int64_t k = 0; k < K; k++) {
        const size_t slot = (size_t)(sliceOff[(size_t)sl] + 64 * k + l);
        // padding: the slice base (narrow) or column 0 (wide)
        int32_t cc = b >= 0 ? b : 0;
        if (k < ln) {
          cc = cols[(size_t)(rowPtr[(size_t)r] + k)];
          v[slot] = vals[(size_t)(rowPtr[(size_t)r] + k)];
        }
        const size_t w = (size_t)colW[(size_t)sl] + (size_t)(64 * k + l) * (b >= 0 ? 1 : 2);
        if (b >= 0) {
          c[w] = (uint16_t)(cc - b);
        } else {
          c[w] = (uint16_t)((uint32_t)cc & 0xffffu);
          c[w + 1] = (uint16_t)((uint32_t)cc >> 16);
        }
      }
    }
  });
  hipStream_t s = stream_;
  out.sliceOff.upload(sliceOff.data(), sliceOff.size(), s);
  out.rowLen.upload(len.data(), len.size(), s);
  if (sorted) out.perm.upload(perm.data(), perm.size(), s);
  out.sliceBase.upload(base.data(), base.size(), s);
  out.sliceColW.upload(colW.data(), colW.size(), s);
  out.col.upload(c.data(), c.size(), s);
  out.val32.upload(v.data(), v.size(), s);
  out.slots = slots;
  out.preferred = !(sorted && 10 * wide > numSlices);
  out.numSlices = (int32_t)numSlices;
}

const void* DeviceMatrix::sellVal(const Sell& s, DType vt) const {
  switch (vt) {
    case DType::F32:
      return s.val32.get();
    case DType::F64:
      return ensureCopy(s.val32, s.val64, vt, s.slots);
    case DType::F16:
      return ensureCopy(s.val32, s.val16, vt, s.slots);
  }
  return nullptr;
}

const void* DeviceMatrix::ensureCopy(const DevBuf& src32, DevBuf& dst,
                                     DType vt, int64_t count) const {
  std::lock_guard<std::mutex> lk(copyMutex_);
  if (!dst) {
    dst.alloc((size_t)std::max<int64_t>(count, 1) * gph::dtypeSize(vt));
    gph::launchCast(src32.get(), DType::F32, dst.get(), vt, count, stream_);
    GPH_HIP(hipStreamSynchronize(stream_));
  }
  return dst.get();
}

const void* DeviceMatrix::crsOff(DType vt) const {
  switch (vt) {
    case DType::F32:
      return off32_.get();
    case DType::F64:
      return ensureCopy(off32_, off64_, vt, nnz_);
    case DType::F16:
      return ensureCopy(off32_, off16_, vt, nnz_);
  }
  return nullptr;
}
const void* DeviceMatrix::crsDiag(DType vt) const {
  switch (vt) {
    case DType::F32:
      return diag32_.get();
    case DType::F64:
      return ensureCopy(diag32_, diag64_, vt, n_);
    case DType::F16:
      return ensureCopy(diag32_, diag16_, vt, n_);
  }
  return nullptr;
}

gph::SpmvArgs DeviceMatrix::sellArgs(DType vt) const {
  gph::SpmvArgs a;
  a.n = n_;
  a.diag = crsDiag(vt);
  const Sell& sl = fmt_ == MatrixFormat::CRS ? sellCrs_ : sellLdu_;
  if (sl.numSlices > 0) {
    a.numSlices = sl.numSlices;
    a.sliceOff = sl.sliceOff.as<int32_t>();
    a.rowLen = sl.rowLen.as<uint16_t>();
    a.perm = sl.perm ? sl.perm.as<int32_t>() : nullptr;
    a.sliceBase = sl.sliceBase.as<int32_t>();
    a.sliceColW = sl.sliceColW.as<uint32_t>();
    a.scol = sl.col.as<uint16_t>();
    a.sval = sellVal(sl, vt);
  }
  return a;
}

gph::SpmvArgs DeviceMatrix::spmvArgs(DType vt) const {
  gph::SpmvArgs a;
  a.n = n_;
  a.diag = crsDiag(vt);
  {
    const Sell& sl = fmt_ == MatrixFormat::CRS ? sellCrs_ : sellLdu_;
    if (sellActive(sl)) {
      a.numSlices = sl.numSlices;
      a.sliceOff = sl.sliceOff.as<int32_t>();
      a.rowLen = sl.rowLen.as<uint16_t>();
      a.perm = sl.perm ? sl.perm.as<int32_t>() : nullptr;
      a.sliceBase = sl.sliceBase.as<int32_t>();
      a.sliceColW = sl.sliceColW.as<uint32_t>();
      a.scol = sl.col.as<uint16_t>();
      a.sval = sellVal(sl, vt);
      return a;
    }
  }
  if (fmt_ == MatrixFormat::CRS) {
    a.numBlocks = numBlocksCrs_;
    a.blockRow = blockRowCrs_.as<int32_t>();
    a.nstreams = 1;
    a.ptr[0] = rowPtr_.as<int32_t>();
    a.col[0] = col_.as<int32_t>();
    a.val[0] = crsOff(vt);
    if (gcol_) {
      a.gcol = gcol_.as<int32_t>();
      a.gpos = gpos_.as<uint16_t>();
      a.gval = vt == DType::F32   ? gval32_.get()
               : vt == DType::F64 ? ensureCopy(gval32_, gval64_, vt, nnz_)
                                  : ensureCopy(gval32_, gval16_, vt, nnz_);
    }
  } else {
    a.numBlocks = numBlocksLdu_;
    a.blockRow = blockRowLdu_.as<int32_t>();
    a.nstreams = 2;
    a.ptr[0] = ownerStart_.as<int32_t>();
    a.col[0] = upperCol_.as<int32_t>();
    a.ptr[1] = nbrStart_.as<int32_t>();
    a.col[1] = lowerCol_.as<int32_t>();
    switch (vt) {
      case DType::F32:
        a.val[0] = upper32_.get();
        a.val[1] = lower32_.get();
        break;
      case DType::F64:
        a.val[0] = ensureCopy(upper32_, upper64_, vt, numFaces_);
        a.val[1] = ensureCopy(lower32_, lower64_, vt, numFaces_);
        break;
      case DType::F16:
        a.val[0] = ensureCopy(upper32_, upper16_, vt, numFaces_);
        a.val[1] = ensureCopy(lower32_, lower16_, vt, numFaces_);
        break;
    }
  }
  return a;
}

gph::TileArgs DeviceMatrix::tileArgs() const {
  gph::TileArgs t;
  t.T = T_;
  t.tileStart = tileStart_.as<int32_t>();
  t.lvlRows = lvlRows_.as<int32_t>();
  t.lvlPtr = lvlPtr_.as<int32_t>();
  t.lvlPtrOff = lvlPtrOff_.as<int32_t>();
  t.ulvlRows = ulvlRows_.as<int32_t>();
  t.ulvlPtr = ulvlPtr_.as<int32_t>();
  t.ulvlPtrOff = ulvlPtrOff_.as<int32_t>();
  t.rowPtr = rowPtr_.as<int32_t>();
  t.col = col_.as<int32_t>();
  t.maxTileRows = maxTileRows_;
  return t;
}

void DeviceMatrix::spmv(const Vec& x, const Vec& y, DType acc, DType vt,
                        hipStream_t s) const {
  gph::SpmvArgs a = spmvArgs(vt);
  a.x = x.ptr;
  a.out = y.ptr;
  gph::launchSpmv(a, vt, x.type, y.type, y.type, acc, gph::kModeSpmv,
                  gph::kRedNone, s);
}

void DeviceMatrix::residual(const Vec& x, const Vec& b, const Vec& r,
                            DType acc, DType vt, hipStream_t s,
                            double* r64) const {
  gph::SpmvArgs a = spmvArgs(vt);
  a.x = x.ptr;
  a.b = b.ptr;
  a.out = r.ptr;
  a.out64 = r64;
  if (acc == DType::TF32 || x.type == DType::TF32) {
    // double-word residual: x in twofloat32, b, r and the values float32
    if (acc != DType::TF32 || x.type != DType::TF32 || b.type != DType::F32 ||
        r.type != DType::F32 || vt != DType::F32)
      gph::fail(gph::kUnsupported,
                "twofloat32 residual: x and accumulation twofloat32, b, r and values float32");
    gph::launchResidualDW(a, s);
    return;
  }
  gph::launchSpmv(a, vt, x.type, b.type, r.type, acc, gph::kModeResidual,
                  gph::kRedNone, s);
}

double DeviceMatrix::norm(const Vec& x, int kind, hipStream_t s) const {
  if (kind == gph::kNormL2) return norm2(x, s);
  gph::VecArgs v;
  v.n = n_;
  v.a = x.ptr;
  v.state = scratchState();
  v.partials = scratchPartials();
  v.fin = gph::kFinNormAuxV;
  gph::launchNormKind(v, x.type, kind, s);
  double out = 0;
  GPH_HIP(hipMemcpyAsync(&out, &scratchState()->aux, sizeof(double),
                         hipMemcpyDeviceToHost, s));
  GPH_HIP(hipStreamSynchronize(s));
  return out;
}

double DeviceMatrix::norm2(const Vec& x, hipStream_t s) const {
  gph::VecArgs v;
  v.n = n_;
  v.a = x.ptr;
  v.state = scratchState();
  v.partials = scratchPartials();
  v.fin = gph::kFinNormAux;
  gph::launchNorm2(v, x.type, s);
  double out = 0;
  GPH_HIP(hipMemcpyAsync(&out, &scratchState()->aux, sizeof(double),
                         hipMemcpyDeviceToHost, s));
  GPH_HIP(hipStreamSynchronize(s));
  return out;
}

size_t DeviceMatrix::deviceBytes() const {
  size_t b = 0;
  for (const DevBuf* p : std::initializer_list<const DevBuf*>
       {&rowPtr_, &col_, &off32_, &diag32_, &off64_, &diag64_, &off16_,
        &diag16_, &blockRowCrs_, &ownerStart_, &upperCol_, &upper32_,
        &nbrStart_, &lowerCol_, &lower32_, &upper64_, &lower64_, &upper16_,
        &lower16_, &blockRowLdu_, &tileStart_, &lvlRows_, &lvlPtr_,
        &lvlPtrOff_, &ulvlRows_, &ulvlPtr_, &ulvlPtrOff_, &scratchState_,
        &scratchPartials_, &gcol_, &gpos_, &gval32_, &gval64_, &gval16_})
    b += p->bytes();
  for (const Sell* sl : {&sellCrs_, &sellLdu_})
    for (const DevBuf* p : std::initializer_list<const DevBuf*>{
             &sl->sliceOff, &sl->rowLen, &sl->perm, &sl->sliceBase, &sl->sliceColW,
             &sl->col, &sl->val32,
             &sl->val64, &sl->val16})
      b += p->bytes();
  return b;
}

}  // namespace graphene::matrix
