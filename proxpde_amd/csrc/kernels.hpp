// Host-side launchers of the device kernels (declared here, defined in the .hip files).
#pragma once

#include "ctx.hpp"
#include "fe.hpp"

namespace pxg
{

// assembly.hip
void buildDofAdjacency(pxg_ctx * c);
void buildPattern(pxg_ctx * c);
void assembleLhs(pxg_ctx * c, int nterms, pxg_term const * terms, int precision);
void assembleRhs(pxg_ctx * c, int kind, double const * params, double const * qvaluesHost);
void qpoints(pxg_ctx * c, double * outHost);
void geometryOut(pxg_ctx * c, double * detJ, double * JxW);
void closeMatrix(pxg_ctx * c, double eps);
// structural values back from the split storage of a direct rebuild
void unsplitValues(pxg_ctx * c);
void clearValues(pxg_ctx * c);
// FP32 copy of closed FP64 values / FP64 copy of closed FP32 values, made on first use
void ensureVal32(pxg_ctx * c);
void ensureVal64(pxg_ctx * c);
// element_mfma.hip: fused geometry + MFMA element kernel; false if not applicable
bool elementMfma(pxg_ctx * c, int flags, double cS, double cSF, double cM, double cMF, int precision);
// element_th_mfma.hip: Taylor-Hood Q2-Q1 hexahedron on the matrix cores,
// coefficients of stiffness, tensor stiffness, mass, grad, div; false if the
// context holds another element
bool elementThMfma(pxg_ctx * c, double kS, double kT, double kM, double kG, double kD, int precision);
// element_exact.hip: linear elements (P1 line, Q1 quad, Q1 hex, P1 tet)
// computed in FP64 in the reference's evaluation order, geometry fused, into
// keBuf (PXG_FP64) or rounded into keBuf32 (PXG_FP32, exact-arithmetic zeros
// flushed); false if not applicable
bool elementExact(pxg_ctx * c, int nterms, pxg_term const * terms, int precision);

// matfree_sf.hip: sum-factorised matrix-free kappa-stiffness operator of the
// Q1 / Q2 hexahedra with their tensor Gauss rules (element buffer [E][n][nb],
// reduced per dof in element order by the gather)
bool sfAvailable(pxg_ctx const * c);
void sfPrepare(pxg_ctx * c);
template <typename T>
void sfElementApply(pxg_ctx * c, double const * kappa, int nb, T const * x, T * buf, bool useBC, hipStream_t s);
// batches of two or more vectors straight into y (element cubes + the sum of
// the partial values per dof); false if the mesh has no cube structure
bool sfApplyCubes(pxg_ctx * c, double const * kappa, int nb, double const * x, double * y, bool useBC,
                  hipStream_t s);
// element cubes of the context's mesh (built on first use); false if the
// mesh has no cube structure
bool sfCubesReady(pxg_ctx * c);
// [E][6][n] FP32 metrics to the slot-major layout of the cubes ([chunk][6][n][64])
void sfMetricSlots(pxg_ctx * c, float const * met, float * metC, hipStream_t s);
// one vector through the FP32 operator of slot-major metrics, straight into
// y with the epilogue of sfGather (modes 0 to 2); part: nCube (2P-1)^3 floats
void sfApplyPaCubes(pxg_ctx * c, int mode, float const * metC, float const * x, float const * b, float const * dinv,
                    double omega, float * y, float * part, bool useBC, hipStream_t s);
void sfElementGrad(pxg_ctx * c, int nb, double const * x, double const * ybar, double * buf, bool useBC,
                   hipStream_t s);
template <typename T>
void sfElementDiag(pxg_ctx * c, double const * kappa, T * buf, hipStream_t s);
// per-point metrics (kappa, geometry, weights) of every element, [E][6][n],
// and the apply of one vector that reads them instead of recomputing them
template <typename T>
void sfMetric(pxg_ctx * c, double const * kappa, T * met, hipStream_t s);
template <typename T>
void sfApplyPA(pxg_ctx * c, T const * met, T const * x, T * buf, bool useBC, hipStream_t s);
// per-dof gather of an [E][n] element buffer; mode 0: y = Ax, 1: y = b - Ax,
// 2: y = x + omega dinv (b - Ax), 3: y = 1 / Ax
template <typename T>
void sfGather(pxg_ctx * c, int mode, T const * buf, T const * x, T const * b, T const * dinv, double omega, T * y,
              hipStream_t s);

// Row stride of the element-matrix buffers (keBuf, keBuf32): rows of 16 or
// more entries are padded to a multiple of 16 entries (128 B in FP64, 64 B in
// FP32), so an element-matrix row fills whole cache lines and no line holds
// parts of two rows.  Two rows of one element belong to two global rows,
// which the scatter gathers at different times; a shared line would be
// fetched twice.  Smaller element matrices are kept dense.
constexpr int keRowStride(int n) { return n >= 16 ? (n + 15) / 16 * 16 : n; }

// Work counters of the persistent element kernels (element_mfma.hip,
// matfree.hip): blocks b with the same b % P share one counter over a
// contiguous range of the elements, one counter per 256-byte line.
constexpr int kWorkCounters = 8;
constexpr int kWorkCounterStride = 64;

// sparse.hip: CSR kernels on raw device arrays
struct CsrView
{
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int64_t const * rowptr = nullptr;
  int32_t const * col = nullptr;
  double const * val = nullptr;
  float const * val32 = nullptr;
  double const * dinv = nullptr;
  float const * dinv32 = nullptr;
  // row blocks of the stream kernel (block q: rows blockRow[q] .. blockRow[q+1]-1,
  // at most kStreamSlots entries); null: the row-group kernel is used
  int32_t const * blockRow = nullptr;
  int64_t nblocks = 0;
  // block-packed copy for the stream kernel (null: not built or not current):
  // column of slot k of block q = pbase[q] + (plo[k] | hi2(k) << 16), values
  // in kStreamSlots slots per block
  int32_t const * pbase = nullptr;
  uint16_t const * plo = nullptr;
  uint16_t const * phi = nullptr;
  double const * pval = nullptr;
  float const * pval32 = nullptr;
};

// entries staged in LDS by one block of the stream kernel
constexpr int kStreamSlots = 2048;
// builds the row-block table of a CSR matrix (device rowptr); nblocks = 0
// when some row is too long for the stream kernel.  Synchronises the stream.
void buildRowBlocks(int64_t nrows, int64_t nnz, int64_t const * rowptr, DevBuf<int32_t> & blockRow,
                    int64_t & nblocks, hipStream_t s);
// row blocks of the closed matrix of a context, built on first use after a
// change of the closed pattern
void ensureRowBlocks(pxg_ctx * c);
// block-packed copy of the closed matrix in `precision` when the context asks
// for one (pxg_csr_pack) and every block spans fewer than 2^18 columns
void ensurePack(pxg_ctx * c, int precision);
// values in `precision`, row blocks and (if asked for) the packed copy: what
// the SpMV / Jacobi kernels read
void ensureOperator(pxg_ctx * c, int precision);

CsrView viewOf(pxg_ctx const * c);

// y = A x
template <typename T>
void spmv(CsrView const & A, T const * x, T * y, hipStream_t s);
// y = b - A x
template <typename T>
void residual(CsrView const & A, T const * b, T const * x, T * y, hipStream_t s);
// y = x + omega D^-1 (b - A x)    (y != x)
template <typename T>
void jacobiStep(CsrView const & A, T const * b, T const * x, T * y, double omega, hipStream_t s);
// y += A x
template <typename T>
void spmvAdd(CsrView const & A, T const * x, T * y, hipStream_t s);

// deterministic FP64 dot product into a device scalar
template <typename T>
void dot(int64_t n, T const * a, T const * b, double * outDev, double * partial, hipStream_t s);

// vector helpers
template <typename T>
void fill(int64_t n, T * x, T v, hipStream_t s);
template <typename Tout, typename Tin>
void convert(int64_t n, Tin const * x, Tout * y, hipStream_t s);
template <typename T>
void jacobiDiag(int64_t n, T const * dinv, T const * b, T * x, double omega, hipStream_t s);
// x += alpha p ; r -= alpha q with alpha = sc[ia] / sc[ib]  (scalars on device)
template <typename T>
void cgUpdate(int64_t n, T * x, T const * p, T * r, T const * q, double const * sc, int ia, int ib, hipStream_t s);
// x += alpha p (FP64 iterate) ; r -= alpha q (FP32 residual), alpha = sc[ia] / sc[ib]
void cgUpdateMixed(int64_t n, double * x, float const * p, float * r, float const * q, double const * sc, int ia,
                   int ib, hipStream_t s);
// p = z + beta p with beta = sc[ia] / sc[ib]
template <typename T>
void cgDirection(int64_t n, T * p, T const * z, double const * sc, int ia, int ib, hipStream_t s);
// x += y (mixed precision)
template <typename T, typename U>
void addTo(int64_t n, T * x, U const * y, hipStream_t s);
// y = a*x + b*y
void axpby(int64_t n, double a, double const * x, double b, double * y, hipStream_t s);

constexpr int kDotBlocks = 1024;

} // namespace pxg
