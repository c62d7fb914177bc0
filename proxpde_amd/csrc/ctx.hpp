// Context of the assembly / operator-apply library: device buffers of one FE
// space on one mesh, its CSR pattern and matrices, and the HIP stream.
#pragma once

#include "../../include/proxpde_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace pxg
{

struct Error: std::runtime_error
{
  int code;
  Error(int c, std::string const & m): std::runtime_error(m), code(c) {}
};

#define PXG_HIP(call)                                                                              \
  do                                                                                               \
  {                                                                                                \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::pxg::Error(e_ == hipErrorOutOfMemory ? PXG_E_OOM : PXG_E_HIP,                        \
                         std::string(#call) + ": " + hipGetErrorString(e_));                       \
  } while (0)

#define PXG_REQUIRE(cond, code, msg)                                                               \
  do                                                                                               \
  {                                                                                                \
    if (!(cond))                                                                                   \
      throw ::pxg::Error(code, msg);                                                               \
  } while (0)

// owning device buffer (hipMalloc / hipFree), resized without preserving content
template <typename T>
struct DevBuf
{
  T * p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(DevBuf const &) = delete;
  DevBuf & operator=(DevBuf const &) = delete;
  DevBuf(DevBuf && o) noexcept: p(o.p), n(o.n)
  {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf & operator=(DevBuf && o) noexcept
  {
    if (this != &o)
    {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release()
  {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void resize(size_t m)
  {
    if (m == n && p)
      return;
    release();
    if (m)
      PXG_HIP(hipMalloc(&p, m * sizeof(T)));
    n = m;
  }
  void upload(T const * h, size_t m, hipStream_t s)
  {
    resize(m);
    if (m)
      PXG_HIP(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T * h, size_t m, hipStream_t s) const
  {
    if (m)
      PXG_HIP(hipMemcpyAsync(h, p, m * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// FE tables uploaded once per (FE type, quadrature)
struct FeTables
{
  int feType = 0, geo = 0, dim = 0, n = 0, npe = 0, nq = 0;
  std::vector<double> phi;     // nq x n
  std::vector<double> grad;    // nq x n x dim   (reference gradients)
  std::vector<double> geoGrad; // nq x npe x dim (vertex mapping gradients)
  std::vector<double> xi;      // nq x dim
  std::vector<double> w;       // nq
  DevBuf<double> d_phi, d_grad, d_geoGrad, d_xi, d_w;
  DevBuf<float> d_phi32, d_grad32;
};

} // namespace pxg

struct pxg_ctx
{
  int device = 0;
  hipStream_t stream = nullptr;
  bool ownStream = false;
  std::string err;

  pxg::FeTables fe;
  // mesh
  int64_t npts = 0, nelem = 0;
  int npe = 0;
  pxg::DevBuf<double> xyz;
  pxg::DevBuf<uint32_t> conn;
  // dofs
  int64_t ndofs = 0;
  pxg::DevBuf<uint32_t> elemMap;
  // Dirichlet data
  bool hasBC = false;
  pxg::DevBuf<uint8_t> bcFlag;
  pxg::DevBuf<double> bcVal;
  // coefficient field
  bool hasKappa = false;
  pxg::DevBuf<double> kappa;
  uint64_t kappaStamp = 0; // counts the times kappa was set (state derived from it checks it)
  // structural pattern and maps
  int64_t nnzS = 0;
  pxg::DevBuf<int64_t> rowptrS;
  pxg::DevBuf<int32_t> colS;
  pxg::DevBuf<uint32_t> segPtr;  // nnzS + 1 offsets into contrib
  pxg::DevBuf<uint32_t> contrib; // E*n*n contribution ids sorted by (row, col, element)
  pxg::DevBuf<uint32_t> emapNnz; // E*n*n -> structural nnz id
  pxg::DevBuf<uint32_t> diagPos; // structural position of each diagonal entry
  pxg::DevBuf<uint32_t> diagC;   // compact position of each diagonal in the pruned layout (~0u: pruned)
  pxg::DevBuf<uint32_t> keep, pos; // close(): keep flags and their exclusive scan
  pxg::DevBuf<unsigned char> scanTmp;
  pxg::DevBuf<uint32_t> dofSeg;  // ndofs + 1 offsets into dofContrib
  pxg::DevBuf<uint32_t> dofContrib; // E*n (element, local) ids sorted by (dof, element)
  // accumulated values (structural) and rhs
  int valPrecision = PXG_FP64;
  bool valsDirty = false;
  bool valsZero = false;  // structural values are logically zero (not yet written)
  bool keepFresh = false; // keep flags of prune(1e-16) match the structural values
  bool val32Valid = false; // FP32 copy of the closed values is current
  bool val64Valid = false; // FP64 copy of the closed values is current
  // pruned layout (keep, pos, rowptr, col) of the last close with the default
  // threshold: a rebuild after clear writes kept entries straight into it
  bool layoutValid = false;
  int layoutPrecision = PXG_FP64;
  // structural values stored split: kept entries in val / val32 at pos[], the
  // pruned ones in valS / valS32
  bool valsSplit = false;
  pxg::DevBuf<int> changed; // direct scatter: some prune flag differs from the layout
  // stage blocks of the rebuild into the pruned layout: runs of rows with at
  // most a fixed number of element-matrix rows, staged in LDS; per structural
  // entry its contribution offset in the block and keep flag, per
  // contribution its LDS slot
  bool stageValid = false;
  int64_t nStage = 0;
  int stagePairs = 0; // element-matrix rows staged per block
  bool stagedFused = false; // the last build's staged rebuild wrote dinv and the lifting
  pxg::DevBuf<int32_t> stageR0;   // nStage + 1 first rows
  pxg::DevBuf<uint32_t> stageBlk; // 8 words per block
  pxg::DevBuf<uint16_t> stageInfo, stageSid;
  // position of every element-matrix row (e, i) in dofContrib order; a build
  // that rebuilds through the stage blocks has its element kernel write row
  // (e, i) at that position with stride n (kePairOrder), so a stage block
  // reads its rows as one contiguous run
  pxg::DevBuf<uint32_t> pairPos;
  bool kePairOrder = false;
  uint64_t stamp = 0;      // counts closes: changes whenever the closed matrix may have changed
  pxg::DevBuf<double> valS;
  pxg::DevBuf<float> valS32;
  pxg::DevBuf<double> b;
  // closed (pruned) matrix
  int64_t nnz = 0;
  bool closed = false;
  pxg::DevBuf<int64_t> rowptr;
  pxg::DevBuf<int32_t> col;
  pxg::DevBuf<double> val;
  pxg::DevBuf<float> val32;
  pxg::DevBuf<double> dinv;
  pxg::DevBuf<float> dinv32;
  uint64_t patStamp = 0; // counts changes of the closed pattern (rowptr, col)
  pxg::DevBuf<int32_t> blockRow; // row blocks of the stream SpMV kernel
  int64_t nStreamBlocks = 0;
  uint64_t blockStamp = 0; // pattern stamp the row blocks were built for (0: none)
  // block-packed copy of the closed matrix for the SpMV / Jacobi kernels
  bool packWanted = false;
  bool packOk = false;      // every row block spans fewer than 2^18 columns
  uint64_t packStamp = 0;   // pattern stamp of the index part
  uint64_t packValStamp = 0, packVal32Stamp = 0; // matrix stamps of the value parts
  pxg::DevBuf<int32_t> packBase;
  pxg::DevBuf<uint16_t> packLo;
  pxg::DevBuf<uint16_t> packHi;
  pxg::DevBuf<int> packFlag;
  pxg::DevBuf<double> packVal;
  pxg::DevBuf<float> packVal32;
  // work buffers
  pxg::DevBuf<double> geoBuf; // E*nq*8
  pxg::DevBuf<double> keBuf;  // E*n*n (fp64)
  pxg::DevBuf<float> keBuf32; // E*n*n (fp32)
  pxg::DevBuf<double> feBuf;  // E*n
  pxg::DevBuf<double> qwork;  // E*nq scratch
  pxg::DevBuf<double> mfBuf;  // matrix-free element results [E][n][nb]
  // batches of the sum-factorised apply in 2x2x2 element cubes (matfree_sf.hip)
  int cubeState = 0; // 0: not built, 1: built, -1: the mesh keeps the element buffer
  int64_t nCube = 0, nPartDof = 0, nPart = 0;
  pxg::DevBuf<uint32_t> cubeElem; // [G][8] element of each corner, ~0u empty
  pxg::DevBuf<uint32_t> cubeOut;  // [G][n][8] where each corner's node goes
  pxg::DevBuf<uint32_t> partDof, partSeg, partIds; // dofs left partial, their slots
  pxg::DevBuf<double> partBuf;    // [nPart][nb]
  // the same cubes for the one-vector FP32 operator: slots in chunks of 64
  int64_t nChunk = 0, nPartDofF = 0;
  pxg::DevBuf<uint32_t> cubeIdC;    // [chunk][n][64] dof of each slot's node, bit 31: written by the slot
  pxg::DevBuf<uint32_t> cubeCmC;    // [chunk][64] complete nodes of the slot, bit 31: slot has an element
  pxg::DevBuf<uint32_t> cubeSlotOf; // slot of each element
  pxg::DevBuf<uint32_t> partDofF, partSegF, partIdsF; // partial values at cube * (2P-1)^3 + node
  pxg::DevBuf<int> status;    // device error flag
  bool statusPending = false; // status of the last build not read yet (read by close)
  pxg::DevBuf<int> workCounter; // element counter of the persistent element kernels
  pxg::DevBuf<double> jacTmp;   // pxg_jacobi: the other vector of the sweeps
  pxg::DevBuf<float> jacTmp32;
  pxg::DevBuf<unsigned char> sortTmp;
  // kernel timing with HIP events on the context stream (off by default)
  bool prof = false;
  struct ProfRec
  {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> profPending;
  std::vector<hipEvent_t> eventPool;
  std::map<std::string, std::pair<double, int64_t>> profAcc;
};

namespace pxg
{
// records the launches issued in its lifetime under `name` when profiling is on
struct ProfScope
{
  pxg_ctx * c;
  std::string name;
  hipEvent_t a = nullptr;
  ProfScope(pxg_ctx * ctx, char const * n): c(ctx), name(n)
  {
    if (c->prof)
    {
      a = take();
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~ProfScope()
  {
    if (c->prof && a)
    {
      hipEvent_t b = take();
      (void)hipEventRecord(b, c->stream);
      c->profPending.push_back({name, a, b});
    }
  }
  hipEvent_t take()
  {
    if (!c->eventPool.empty())
    {
      hipEvent_t e = c->eventPool.back();
      c->eventPool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
};

// resolve pending event pairs into accumulated milliseconds
inline void profFlush(pxg_ctx * c)
{
  for (auto & r: c->profPending)
  {
    float ms = 0.f;
    (void)hipEventSynchronize(r.b);
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto & acc = c->profAcc[r.name];
    acc.first += ms;
    acc.second += 1;
    c->eventPool.push_back(r.a);
    c->eventPool.push_back(r.b);
  }
  c->profPending.clear();
}
} // namespace pxg
