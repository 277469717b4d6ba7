// Host-side launch interface of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gph_common.hpp"
#include "k_common.hpp"

namespace gph {

constexpr int kSpmvThreads = 256;   // one row per thread, <= 256 rows/block
constexpr int kSellMaxGrid = 4096;  // bound of the sliced-ELL grid (resident blocks)
constexpr int kSpmvCap = 2048;      // staged products per chunk
constexpr int kVecThreads = 256;
constexpr int kTileThreads = 256;

// Row-gather operator in "streams" form: CRS has one stream
// (rowPtr, col, offDiag); LDU has two (owner faces: ownerStart, upperAddr,
// upper; losort faces: neighbourStart, resolved lower column, lower). Rows
// sum diag*x first, then each stream in stored order.
struct SpmvArgs {
  int32_t n = 0;
  int32_t numBlocks = 0;
  const int32_t* blockRow = nullptr;  // numBlocks+1 row boundaries
  int32_t nstreams = 1;
  const int32_t* ptr[2] = {nullptr, nullptr};
  const int32_t* col[2] = {nullptr, nullptr};
  const void* val[2] = {nullptr, nullptr};
  const void* diag = nullptr;
  // Row-block form, one stream (CRS): each staging chunk's entries ordered
  // by column, with each entry's position in its chunk, so the x gathers
  // of neighbouring lanes share cache lines; used instead of col[0] /
  // val[0] for the products when set (the row sums keep the stored order)
  const int32_t* gcol = nullptr;
  const uint16_t* gpos = nullptr;
  const void* gval = nullptr;
  const void* x = nullptr;
  const void* b = nullptr;      // residual right-hand side
  void* out = nullptr;
  void* out2 = nullptr;         // optional copies of out
  void* out3 = nullptr;
  double* out64 = nullptr;      // optional unrounded FP64 result
  const void* w = nullptr;      // dot partner
  DevState* state = nullptr;
  void* partials = nullptr;     // grid reduction scratch: arrival counters, then sums
  int32_t fin = kFinNone;
  int32_t checkDone = 0;
  int32_t remap = 1;
  // Sliced-ELL form (64 rows per slice, entries column-major inside a
  // slice): used when numSlices > 0.
  int32_t numSlices = 0;
  const int32_t* sliceOff = nullptr;   // numSlices+1, in slots
  const uint16_t* rowLen = nullptr;    // numSlices*64 entries per lane
  const int32_t* perm = nullptr;       // lane -> row (null: slice*64+lane)
  const int32_t* sliceBase = nullptr;  // column base (16-bit slices) or -1
  const uint32_t* sliceColW = nullptr; // first column word of each slice
  const uint16_t* scol = nullptr;      // column words
  const void* sval = nullptr;
  // optional (batched solves): slice s is processed only when one of the
  // systems sliceSys[2s] .. sliceSys[2s+1] is in phase runPhase
  // (sysPhase[k]: 0 inner solve running, 1 outer step due, 2 finished)
  const int32_t* sysPhase = nullptr;
  const int32_t* sliceSys = nullptr;
  int32_t runPhase = 0;
};

enum SpmvMode : int32_t { kModeSpmv = 0, kModeResidual = 1 };
enum RedKind : int32_t { kRedNone = 0, kRedSelf = 1, kRedW = 2, kRedSelfW = 3 };

// Launches the row-gather kernel. vt: value type; xt: x type; bt: b type
// (residual); ot: output type; at: accumulation type.
void launchSpmv(const SpmvArgs& a, DType vt, DType xt, DType bt, DType ot,
                DType at, int mode, int red, hipStream_t s);
// y_j = A x_j for k (1..8) float32 vectors of the sliced layout at
// x/out + j*stride, the matrix streams read once (k_multi.hip); vector j is
// skipped when states[j].done, every vector when glob->done (both optional).
void launchSpmvMulti(const SpmvArgs& a, int k, int64_t stride, const DevState* states,
                     const DevState* glob, hipStream_t s);

// --- vector kernels -------------------------------------------------------
struct VecArgs {
  int32_t n = 0;
  const void* a = nullptr;
  const void* b = nullptr;
  const void* c = nullptr;
  const void* d = nullptr;
  void* o1 = nullptr;
  void* o2 = nullptr;
  DevState* state = nullptr;
  void* partials = nullptr;     // grid reduction scratch: arrival counters, then sums
  int32_t fin = kFinNone;
  int32_t checkDone = 0;
};
int vecGrid(int64_t n);
// x += alpha*y + omega*z ; r = s - omega*t ; sums r.r and r0.r  (kFinUpdate)
//   a=y, b=z, c=s, d=t, o1=x, o2=r, w=r0 passed in `extra`
void launchUpdateXR(const VecArgs& v, const void* r0, DType x, hipStream_t s);
// p = r + beta*(p - omega*v) when iter > 0; o2 (optional) = p
//   a=r, c=v, o1=p
void launchUpdateP(const VecArgs& v, DType x, hipStream_t s);
// s = r - alpha*v ; o2 (optional) = s   (a=r, c=v, o1=s)
void launchUpdateS(const VecArgs& v, DType x, hipStream_t s);
// sum of squares of a (type t), finalised with v.fin into the state
void launchNorm2(const VecArgs& v, DType t, hipStream_t s);
// sum |a| (kNormL1) or max |a| (kNormLinf), finalised with v.fin (one of the
// ...V ops, which take the finished norm)
void launchNormKind(const VecArgs& v, DType t, int kind, hipStream_t s);
void launchFill(void* x, DType t, int64_t n, double value, hipStream_t s);
// --- double-word ("twofloat32") kernels (k_dword.hip) ----------------------
// r = fl32(b - A x) in double-word arithmetic (x: TF32, b, r and the matrix
// values float32); out64 (optional) = the residual rounded to float64;
// with state and fin: sum of its squares finalised with fin
void launchResidualDW(const SpmvArgs& a, hipStream_t s);
// phi (TF32) += d (float32), in double words
void launchAccumulateDW(void* phi, const void* d, DType dt, int64_t n, hipStream_t s);
// casts float32 / float64 <-> TF32; false when neither type is TF32
bool launchCastDW(const void* in, DType it, void* out, DType ot, int64_t n, hipStream_t s);
// ||x||_2 with the reduction scratch in a static device array instead of an
// allocation (k_noalloc.hip); synchronous on stream s.
double norm2NoAlloc(const void* x, DType t, int64_t n, hipStream_t s);
size_t noAllocScratchBytes();
// x_i = U[-1,1) from a counter-based hash of (seed, i) (device-side RHS)
void launchUniform(double* x, int64_t n, uint64_t seed, hipStream_t s);
void launchCast(const void* in, DType it, void* out, DType ot, int64_t n,
                hipStream_t s);
// phi(f64) += d (dt)
void launchAccumulate(double* phi, const void* d, DType dt, int64_t n,
                      hipStream_t s);

// --- tile-local preconditioners -------------------------------------------
struct TileArgs {
  int32_t T = 0;
  const int32_t* tileStart = nullptr;   // T+1
  const int32_t* lvlRows = nullptr;     // local ids grouped by forward level
  const int32_t* lvlPtr = nullptr;      // per tile numColors+1, relative
  const int32_t* lvlPtrOff = nullptr;   // T+1
  const int32_t* ulvlRows = nullptr;    // backward levels
  const int32_t* ulvlPtr = nullptr;
  const int32_t* ulvlPtrOff = nullptr;
  const int32_t* rowPtr = nullptr;
  const int32_t* col = nullptr;
  int32_t maxTileRows = 0;
};

// Right-hand side pre-operation fused into the preconditioner sweep.
enum PreOpKind : int32_t { kPreNone = 0, kPreP = 1, kPreS = 2 };
struct PreOp {
  int32_t kind = kPreNone;
  const void* r = nullptr;
  void* p = nullptr;   // p (kPreP, in/out) or s (kPreS, out)
  const void* v = nullptr;
};

// Rows of the forward levels of all tiles, level by level: list[2r] = row,
// list[2r+1] = its tile.
struct FactorArgs {
  const int32_t* tileStart = nullptr;
  const int32_t* rowPtr = nullptr;
  const int32_t* col = nullptr;
  const int32_t* list = nullptr;
};
// In-place ILU(0) factorization of copies of off/diag (ft = f32 or f64),
// diagonal inverted at the end. dilu: diagonal-based ILU, only diag changes.
// levelOff[L]..levelOff[L+1]-1: the list entries of forward level L; n:
// rows of the matrix.
void launchIluFactor(const FactorArgs& f, const std::vector<int64_t>& levelOff, int64_t n,
                     void* off, void* diag, DType ft, bool dilu, hipStream_t s);
// out = (LU)^-1 rhs per tile; ft factor type, xt vector type. dilu: the
// factors are A's off-diagonals and the DILU diagonal.
void launchIluApply(const TileArgs& t, const void* lu, const void* invD,
                    DType ft, const void* rhs, void* out, DType xt,
                    const PreOp& pre, DevState* st, int checkDone, bool dilu,
                    hipStream_t s);
// One Gauss-Seidel sweep per tile, xout != xin.
void launchGsSweep(const TileArgs& t, const void* off, const void* diag,
                   DType vt, const void* b, const void* xin, void* xout,
                   DType xt, bool backward, DevState* st, int checkDone,
                   hipStream_t s);

size_t tileSharedBytes(const TileArgs& t, DType xt);

// Level-major packed ILU(0) factors (one direction).
struct PackedDirArgs {
  const uint16_t* col = nullptr;     // local column per slot
  const void* val = nullptr;         // factor value per slot
  const int32_t* slabOff = nullptr;  // T+1
  const int32_t* lvlOff = nullptr;   // T+1
  const int32_t* slabEnd = nullptr;  // per tile nL+1 (last = sentinel)
  const int32_t* rowStart = nullptr; // per tile nL+1
  const int32_t* width = nullptr;    // per tile nL+1: entries per row slot
  const uint16_t* rowLoc = nullptr;  // per row slot: local row id
};
struct PackedArgs {
  int32_t T = 0;
  const int32_t* tileStart = nullptr;
  int32_t maxRows = 0, maxLevels = 0, maxLevelEntries = 0;
  int32_t dilu = 0;  // diagonal-based ILU: x_i = t*invD_i / x_i + t*invD_i
  PackedDirArgs fwd, bwd;
  // optional (batched solves): group g is skipped when the systems
  // groupSys[2g] .. groupSys[2g+1] have no inner solve running
  // (sysPhase[k] != 0)
  const int32_t* sysPhase = nullptr;
  const int32_t* groupSys = nullptr;
};
constexpr int kPackedThreads = 64;  // one wave per tile
size_t packedSharedBytes(const PackedArgs& p, DType xt);
// out = (LU)^-1 rhs with the packed factors; one wave per tile.
void launchIluPacked(const PackedArgs& p, const void* invD, DType ft,
                     const void* rhs, void* out, DType xt, const PreOp& pre,
                     DevState* st, int checkDone, hipStream_t s);
// Level-record ILU(0) substitution (lane per row, two groups per wave,
// record stream through an LDS ring filled by LDS-DMA; record format in
// ilu_pack.hpp).
struct LevelArgs {
  int32_t G = 0;                        // groups
  int32_t P = 0;                        // group pairs (one wave each)
  const int32_t* groupStart = nullptr;  // G+1 row offsets
  const int64_t* pairOff = nullptr;     // P+1 byte offsets of the records
  const int32_t* steps = nullptr;       // 2P: forward, backward step counts
  const uint8_t* bytes = nullptr;       // record stream (+ zero tail)
  int32_t maxRows = 0;                  // also the zero slot's local id
  int32_t dilu = 0;
};
constexpr int kLvlChunk = 1024;        // bytes per LDS-DMA wave load (64 x 16 B)
constexpr int kLvlRing = 16384;        // LDS ring bytes
constexpr int kLvlLag = 10;            // chunks in flight beyond the current step
// Largest step record (the ring holds a step, its successor's header and a
// partial chunk on each side) and the zero tail after the stream.
constexpr int64_t kLvlMaxStep = kLvlRing - 2 * kLvlChunk;
constexpr int64_t kLvlTail = kLvlRing + kLvlChunk;
size_t levelSharedBytes(const LevelArgs& p, DType xt);
void launchIluLevel(const LevelArgs& p, DType ft, const void* rhs, void* out,
                    DType xt, const PreOp& pre, DevState* st, int checkDone,
                    hipStream_t s);
// Row-record ILU(0) substitution (k_ilu_rows.hip): one wave per tile, up to
// 8 factor entries of a record per lane, the lanes of a record handing the
// row's partial sum on in entry order; each lane's entries are loaded into
// a register ring several steps ahead (stream format in ilu_pack.hpp).
struct RowsArgs {
  int32_t T = 0, maxRows = 0, maxSteps = 0, maxRecs = 0;
  int32_t dilu = 0;
  int32_t dense = 0;  // records padded to even lengths (else whole chunks)
  const int32_t* tileStart = nullptr;  // T+1
  const uint32_t* elemOff = nullptr;   // T+1 element offsets of the tiles' streams
  const int32_t* hdrOff = nullptr;     // T+1 offsets of the tiles' step headers
  const uint32_t* hdr = nullptr;       // step headers
  const int64_t* metaOff = nullptr;    // T+1 offsets of the tiles' record words
  const uint32_t* meta = nullptr;      // record words
  const uint16_t* col = nullptr;       // column byte offsets (4 * local id)
  const void* val = nullptr;           // factor values
  // optional (batched solves): tile t is skipped when system tileSys[t]
  // has no inner solve running (sysPhase[tileSys[t]] != 0)
  const int32_t* sysPhase = nullptr;
  const int32_t* tileSys = nullptr;
  // several right-hand sides of the same matrix in one launch: vector j is
  // rhs/out + j * rhsStride; it is skipped when rhsState[j].done (optional)
  int32_t nrhs = 1;
  int64_t rhsStride = 0;
  const DevState* rhsState = nullptr;
};
constexpr int kRowsThreads = 64;
constexpr int kRowsMaxLen = 32;  // entries of a record in one step (long record)
constexpr int kRowsChunk = 8;    // entries per lane
constexpr int kRowsTail = 64;    // padding elements after the streams
// record word flags
constexpr uint32_t kRowsLongBit = 1u << 15;
constexpr uint32_t kRowsContBit = 1u << 16;
constexpr uint32_t kRowsPartialBit = 1u << 17;
constexpr int kRowsLenShift = 9;     // record word: stored length (6 bits)
constexpr int kRowsSlotShift = 18;   // record word: the row's slot (14 bits)
size_t rowsSharedBytes(const RowsArgs& p, DType ft, DType xt);
void launchIluRows(const RowsArgs& p, const void* invD, DType ft, const void* rhs,
                   void* out, DType xt, const PreOp& pre, DevState* st,
                   int checkDone, hipStream_t s);
// Workgroup-per-tile ILU(0) substitution (k_ilu_block.hip) for tiles of any
// size: 1024 threads walk the tile's levels with a barrier per level, x in
// LDS when it fits, the level-major factor stream (ilu_pack.hpp,
// packIluBlock) prefetched several levels ahead into registers.
constexpr int kBlockThreads = 1024;
struct BlockDirArgs {
  const int32_t* lvlBase = nullptr;  // T+1 first level of each tile
  const int32_t* info = nullptr;     // per level {first slot, rows, first entry, width}
  const void* rows = nullptr;        // per slot: (local row id, 1/d) record
  const void* ents = nullptr;        // per entry: (local column, factor value) record
};
struct BlockArgs {
  int32_t T = 0, maxRows = 0, width = 0, dilu = 0;
  int32_t maxLevels = 0;  // most levels of one tile (both directions)
  const int32_t* tileStart = nullptr;
  BlockDirArgs fwd, bwd;
};
// LDS of the level tables of one tile
__host__ __device__ inline size_t blockInfoBytes(int32_t maxLevels) {
  return (size_t)maxLevels * 16;
}
size_t blockSharedBytes(const BlockArgs& a, DType xt);
void launchIluBlock(const BlockArgs& a, DType ft, const void* rhs, void* out, DType xt,
                    const PreOp& pre, DevState* st, int checkDone, hipStream_t s);
// Gauss-Seidel sweep over the jagged-diagonal level stream (k_gs.hip,
// stream format in gs_pack.hpp): a workgroup per tile, the sweep's x in
// LDS, per level the products entry-parallel, then the row sums.
constexpr int kGsThreads = 256;
constexpr int kGsMaxWidth = 255;  // longest row of the stream
constexpr uint32_t kGsFreshBit = 0x80000000u;
struct GsArgs {
  int32_t T = 0, maxRows = 0;
  int32_t maxLevelEntries = 0;         // entries of the largest level
  const int32_t* tileStart = nullptr;
  const int32_t* lvlBase = nullptr;    // T+1
  const int32_t* lvl = nullptr;        // per level: first slot, rows, first diagonal, width
  const int32_t* diagStart = nullptr;  // per level: width + 1 diagonal starts
  const uint32_t* slot = nullptr;      // per slot: tile-local row << 8 | length
  const uint32_t* code = nullptr;      // per entry: x source
  const float* val = nullptr;          // per entry: matrix value
  const uint16_t* pos = nullptr;       // per entry: position of its product in the level
};
size_t gsStreamSharedBytes(const GsArgs& a, DType xt);
void launchGsStream(const GsArgs& a, const void* diag, const void* b, const void* xin,
                    void* xout, DType xt, DevState* st, int checkDone, hipStream_t s);
// Batched solves over a block-diagonal union of systems (k_batch.hip):
// chunks of rows, each inside one system; per-system solver states.
constexpr int kBatchThreads = 256;
constexpr int kBatchChunk = 2048;  // rows per chunk (at most)
struct BatchArgs {
  int32_t S = 0;                      // systems
  int32_t C = 0;                      // chunks
  const int32_t* chunkSys = nullptr;  // C: system of each chunk
  const int32_t* chunkRow = nullptr;  // C+1: first row of each chunk
  const int32_t* sysChunk = nullptr;  // S+1: first chunk of each system
  DevState* st = nullptr;             // S inner-solver states
  DevState* glob = nullptr;           // done = no system running
  void* partials = nullptr;           // 2*C sums
  // per system: 0 inner solve running, 1 outer (refinement) step due,
  // 2 finished (optional: the skip masks of SpMV slices and ILU tiles)
  int32_t* phase = nullptr;
  // refinement with the outer step on the device: a finished inner solve
  // moves its system to phase 1 and glob->done is set by the outer step
  int32_t deviceOuter = 0;
};
template <int K>
struct BatchDot {
  const void* a[K];
  const void* b[K];
};
void launchBatchDot1(const BatchArgs& b, const void* a0, const void* b0, hipStream_t s);
void launchBatchDot2(const BatchArgs& b, const void* a0, const void* b0, const void* a1,
                     const void* b1, hipStream_t s);
void launchBatchFin(const BatchArgs& b, int op, int k, hipStream_t s);
void launchBatchUpdateP(const BatchArgs& b, const float* r, float* p, const float* v,
                        hipStream_t s);
void launchBatchUpdateS(const BatchArgs& b, const float* r, float* sv, const float* v,
                        hipStream_t s);
void launchBatchUpdateXR(const BatchArgs& b, float* x, float* r, const float* y, const float* z,
                         const float* sv, const float* t, const float* r0, hipStream_t s);
void launchBatchStart(const BatchArgs& b, const float* rhs, float* x, float* r, float* r0,
                      float* p, hipStream_t s);
void launchBatchAccumulate(const BatchArgs& b, const DevState* outer, double* phi,
                           const float* d, hipStream_t s);
void launchBatchOuter(const BatchArgs& b, DevState* outer, const double* r64, int first,
                      hipStream_t s);
// Device-side refinement step of the systems in phase 1 (their inner solve
// has finished): phi += d (launchBatchPhaseAccumulate); after the FP64
// residual of their rows: per-system norm, refinement convergence, phase 2
// for the finished ones and glob->done when every system is (outer step,
// launchBatchPhaseOuter; first: the initial residual of every system); the
// inner restart of those still refining: d = 0, r = r0 = p = fl32(residual)
// and the inner initialisation, phase 0 (launchBatchPhaseRestart).
// phase[s] = 2 if the inner state of system s is finished, else 0 (start of
// an inner solve driven by the host)
void launchBatchPhaseSync(const BatchArgs& b, hipStream_t s);
void launchBatchPhaseAccumulate(const BatchArgs& b, double* phi, const float* d, hipStream_t s);
void launchBatchPhaseOuter(const BatchArgs& b, DevState* outer, const double* r64, int first,
                           hipStream_t s);
void launchBatchPhaseRestart(const BatchArgs& b, const float* rhs, float* x, float* r, float* r0,
                             float* p, hipStream_t s);
void launchBatchOuterMask(const BatchArgs& b, const DevState* outer, hipStream_t s);
// dst[k] = idx[k] < 0 ? 0 : src[idx[k]]  (same element type t)
void launchGather(const void* src, DType t, const int32_t* idx, void* dst,
                  int64_t n, hipStream_t s);

}  // namespace gph
