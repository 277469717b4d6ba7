// Device-resident matrix: the counterpart of HostMatrix::copyToTile and of
// the reference's CRSMatrix / LDUMatrix (lg/matrix/details/crs/CRSMatrix.hpp:35-66,
// lg/matrix/details/ldu/LDUMatrix.hpp:49-57) plus the operations of
// MatrixBase (spmv, residual, vectorNorm; lg/matrix/details/MatrixBase.hpp:66-82).
#pragma once

#include <memory>
#include <mutex>

#include "gph_common.hpp"
#include "host_matrix.hpp"
#include "kernels.hpp"

namespace graphene::matrix {

enum class MatrixFormat : int32_t { CRS = 0, LDU = 1 };  // MatrixFormat.hpp

// A device vector: pointer + element type (the reference's Tensor in the
// tile-major numbering; vectors "with halo" and "without halo" coincide).
struct Vec {
  void* ptr = nullptr;
  gph::DType type = gph::DType::F32;
};

class DeviceMatrix {
 public:
  DeviceMatrix(std::shared_ptr<const host::HostMatrix> host, MatrixFormat fmt,
               int32_t rowsPerBlock = 0);
  ~DeviceMatrix();
  DeviceMatrix(const DeviceMatrix&) = delete;
  DeviceMatrix& operator=(const DeviceMatrix&) = delete;

  const host::HostMatrix& hostMatrix() const { return *host_; }
  std::shared_ptr<const host::HostMatrix> hostPtr() const { return host_; }
  MatrixFormat format() const { return fmt_; }
  int32_t n() const { return n_; }
  int64_t nnzOff() const { return nnz_; }
  int32_t numTiles() const { return T_; }
  // the stream of the matrix's work and of the solvers built on it: the
  // caller's (setStream) or the matrix's own
  hipStream_t stream() const { return user_ ? user_ : stream_; }
  void setStream(hipStream_t s) { user_ = s; }

  // Row-gather arguments for values stored in `vt` (copies in float16 /
  // float64 are made on first use).
  gph::SpmvArgs spmvArgs(gph::DType vt) const;
  // The sliced-ELL form of the matrix's row streams, whichever form
  // spmvArgs prefers (numSlices = 0 if the matrix has none).
  gph::SpmvArgs sellArgs(gph::DType vt) const;
  gph::TileArgs tileArgs() const;
  const void* crsOff(gph::DType vt) const;   // CRS off-diagonal values
  const void* crsDiag(gph::DType vt) const;
  const int32_t* crsRowPtr() const { return rowPtr_.as<int32_t>(); }
  const int32_t* crsCol() const { return col_.as<int32_t>(); }

  // y = A*x, accumulated in `acc` (Matrix::operator*, MatrixBase::spmv).
  void spmv(const Vec& x, const Vec& y, gph::DType acc, gph::DType vt,
            hipStream_t s) const;
  // r = b - A*x accumulated in `acc` (MatrixBase::residual).
  void residual(const Vec& x, const Vec& b, const Vec& r, gph::DType acc,
                gph::DType vt, hipStream_t s, double* r64 = nullptr) const;
  // ||x||_2 (MatrixBase::vectorNorm), synchronous.
  double norm2(const Vec& x, hipStream_t s) const;
  // ||x|| in the norm `kind` (gph::NormKind), synchronous.
  double norm(const Vec& x, int kind, hipStream_t s) const;

  // scratch for standalone reductions
  gph::DevState* scratchState() const { return scratchState_.as<gph::DevState>(); }
  void* scratchPartials() const { return scratchPartials_.get(); }
  size_t partialsBytes() const { return partialsBytes_; }

  size_t deviceBytes() const;

  // Kernel form of the row gather: sliced ELL (default) or row blocks.
  // Without a call, the sliced form is used where it is the faster one: rows
  // of similar length whose columns mostly fit the 16-bit slice offsets.
  // Rows of strongly varying length with far columns (power-law graphs)
  // run faster in the row-block kernel (4M rows, mean 14 entries, 10 %
  // uniformly random columns: 173 us against 460 us on MI355X).
  void setUseSell(bool v) { sellMode_ = v ? 1 : 0; }
  bool useSell() const { return sellActive(fmt_ == MatrixFormat::CRS ? sellCrs_ : sellLdu_); }

 private:
  std::vector<int32_t> buildRowBlocks(const std::vector<int64_t>& rowLen,
                                      int32_t rowsPerBlock, gph::DevBuf& out,
                                      int32_t& nb);
  void buildColumnOrder(const std::vector<int32_t>& blockRow);
  const void* ensureCopy(const gph::DevBuf& src32, gph::DevBuf& dst, gph::DType vt,
                         int64_t count) const;

  std::shared_ptr<const host::HostMatrix> host_;
  MatrixFormat fmt_;
  int32_t n_ = 0, T_ = 0;
  int64_t nnz_ = 0;
  hipStream_t stream_ = nullptr;
  hipStream_t user_ = nullptr;  // caller's stream (not owned)

  gph::DevBuf rowPtr_, col_, off32_, diag32_;
  mutable gph::DevBuf off64_, diag64_, off16_, diag16_;
  gph::DevBuf blockRowCrs_;
  int32_t numBlocksCrs_ = 0;
  // row-block chunks of the CRS stream in column order (irregular matrices):
  // column, position in the chunk, value
  gph::DevBuf gcol_, gpos_, gval32_;
  mutable gph::DevBuf gval64_, gval16_;

  // LDU: owner-face stream and resolved losort stream
  int64_t numFaces_ = 0;
  gph::DevBuf ownerStart_, upperCol_, upper32_, nbrStart_, lowerCol_, lower32_;
  mutable gph::DevBuf upper64_, lower64_, upper16_, lower16_;
  gph::DevBuf blockRowLdu_;
  int32_t numBlocksLdu_ = 0;

  gph::DevBuf tileStart_, lvlRows_, lvlPtr_, lvlPtrOff_, ulvlRows_, ulvlPtr_,
      ulvlPtrOff_;
  int32_t maxTileRows_ = 0;

  // sliced-ELL copies of the row streams (CRS order / LDU order)
  struct Sell {
    int32_t numSlices = 0;
    // col: column words, 16-bit offsets from sliceBase for slices whose
    // columns span less than 2^16, full 32-bit columns otherwise
    gph::DevBuf sliceOff, rowLen, perm, sliceBase, sliceColW, col, val32;
    mutable gph::DevBuf val64, val16;
    int64_t slots = 0;
    bool preferred = true;  // the default kernel form for this matrix
  };
  void buildSell(const std::vector<int64_t>& rowPtr, const std::vector<int32_t>& cols,
                 const std::vector<float>& vals, Sell& out);
  const void* sellVal(const Sell& s, gph::DType vt) const;
  Sell sellCrs_, sellLdu_;
  int sellMode_ = -1;  // -1: automatic, 0: row blocks, 1: sliced
  bool sellActive(const Sell& sl) const {
    return sl.numSlices > 0 && (sellMode_ == 1 || (sellMode_ < 0 && sl.preferred));
  }

  gph::DevBuf scratchState_, scratchPartials_;
  size_t partialsBytes_ = 0;
  mutable std::mutex copyMutex_;
};

}  // namespace graphene::matrix
