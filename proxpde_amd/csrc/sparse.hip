// CSR operator kernels on gfx950: SpMV, residual, weighted Jacobi, transfer
// operators (P x, P^T r as CSR products) and the vector kernels of CG.
//
// SpMV follows SparseMatrix<RowMajor> * Vec of the reference's multigrid test
// (test/multigrid.cpp:78,81,97).  Two kernels, both with a summation order
// fixed by the matrix and the launch shape (bitwise repeatable):
//  * stream kernel (the CSR-stream scheme of CSR-Adaptive, Greathouse and
//    Daga, SC'14): a block takes the rows whose first entry falls in one
//    window of `cap` entries (at most kStreamSlots entries in all), streams
//    their column/value arrays fully coalesced, multiplies by the gathered x
//    into LDS, then reduces each row from LDS with a group of 8 lanes;
//  * packed stream kernel: the same blocks and reduction over a block-packed
//    copy of the matrix (one base column per block, 16 + 2 bits of column
//    offset per entry, values in fixed slots of kStreamSlots per block), so
//    an entry streams 6.25 (FP32) or 10.25 (FP64) bytes instead of 8 or 12;
//    the products, and so the results, are bitwise those of the stream kernel;
//  * row-group kernel (no row-block table, or rows too long): one group of
//    LPR lanes per row, lanes stride the row, butterfly reduction.
#include "kernels.hpp"

#include <algorithm>
#include <cstdlib>

namespace pxg
{

namespace
{
constexpr int kBlock = 256;

inline unsigned blocksFor(int64_t work, int block = kBlock)
{
  int64_t g = (work + block - 1) / block;
  return (unsigned)std::max<int64_t>(g, 1);
}

enum : int
{
  kModeApply = 0,    // y = A x
  kModeResidual = 1, // y = b - A x
  kModeJacobi = 2,   // y = x + omega dinv (b - A x)
  kModeAdd = 3,      // y += A x
};

template <typename T, int LPR, int MODE>
__global__ void __launch_bounds__(kBlock)
    k_csr(int64_t nrows, int64_t const * __restrict__ rowptr, int32_t const * __restrict__ col,
          T const * __restrict__ val, T const * __restrict__ x, T const * __restrict__ b,
          T const * __restrict__ dinv, T omega, T * __restrict__ y)
{
  int64_t const gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t const row = gid / LPR;
  int const lane = (int)(gid % LPR);
  if (row >= nrows)
    return; // whole LPR groups leave together
  int64_t const k0 = rowptr[row], k1 = rowptr[row + 1];
  // up to kUnroll entries per lane are loaded before any is used, so the
  // column/value stream and the x gathers of a lane are all in flight at once;
  // the matrix streams bypass the caches (non-temporal), keeping x resident
  constexpr int kUnroll = 8;
  int const len = (int)(k1 - k0);
  T s = T(0);
  if (len > 0)
  {
    // indices clamped into the row: the loads carry no condition, so they
    // are all issued before the first use
    int32_t cc[kUnroll];
    T vv[kUnroll];
#pragma unroll
    for (int m = 0; m < kUnroll; ++m)
    {
      int64_t const k = k0 + min(lane + m * LPR, len - 1);
      cc[m] = __builtin_nontemporal_load(col + k);
      vv[m] = __builtin_nontemporal_load(val + k);
    }
    T xv[kUnroll];
#pragma unroll
    for (int m = 0; m < kUnroll; ++m)
      xv[m] = x[cc[m]];
#pragma unroll
    for (int m = 0; m < kUnroll; ++m)
      s += lane + m * LPR < len ? vv[m] * xv[m] : T(0);
  }
  for (int64_t k = k0 + lane + (int64_t)kUnroll * LPR; k < k1; k += LPR)
    s += val[k] * x[col[k]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1)
    s += __shfl_xor(s, off, LPR);
  if (lane == 0)
  {
    if constexpr (MODE == kModeApply)
      y[row] = s;
    else if constexpr (MODE == kModeResidual)
      y[row] = b[row] - s;
    else if constexpr (MODE == kModeJacobi)
      y[row] = x[row] + omega * dinv[row] * (b[row] - s);
    else
      y[row] += s;
  }
}

template <typename T, int MODE>
__device__ inline void csrOut(int64_t row, T s, T const * __restrict__ x, T const * __restrict__ b,
                              T const * __restrict__ dinv, T omega, T * __restrict__ y)
{
  if constexpr (MODE == kModeApply)
    y[row] = s;
  else if constexpr (MODE == kModeResidual)
    y[row] = b[row] - s;
  else if constexpr (MODE == kModeJacobi)
    y[row] = x[row] + omega * dinv[row] * (b[row] - s);
  else
    y[row] += s;
}

// Row block of a workgroup.  Workgroups are dispatched to the 8 XCDs in
// turn (workgroup b to XCD b % 8); each XCD has its own L2.  Mapping the
// workgroups of XCD x to the x-th contiguous run of row blocks keeps the x
// entries an XCD gathers within a narrow column band, so they stay in its
// L2 instead of every XCD streaming all of x.  A bijection for any count.
__device__ inline unsigned xcdRowBlock()
{
  constexpr unsigned kXcd = 8;
  unsigned const nb = gridDim.x, b = blockIdx.x;
  unsigned const x = b % kXcd, k = b / kXcd, base = nb / kXcd, extra = nb % kXcd;
  return x * base + min(x, extra) + k;
}

// row sums of the products staged in LDS (entry k of the block in prod[k]):
// G lanes per row (chosen from the mean row length); the lanes of one group
// run the same rows, so the group stays converged
template <typename T, int MODE, int G>
__device__ inline void streamReduce(T const * prod, int64_t r0, int64_t r1, int64_t base,
                                    int64_t const * __restrict__ rowptr, T const * __restrict__ x,
                                    T const * __restrict__ b, T const * __restrict__ dinv, T omega,
                                    T * __restrict__ y)
{
  int const t = threadIdx.x;
  int const lane = t % G;
  for (int64_t row = r0 + t / G; row < r1; row += kBlock / G)
  {
    int const k0 = (int)(rowptr[row] - base), k1 = (int)(rowptr[row + 1] - base);
    // rows up to kRed * G entries in one unrolled pass of independent reads
    constexpr int kRed = G >= 8 ? 16 : 8;
    T s = T(0);
#pragma unroll
    for (int j = 0; j < kRed; ++j)
    {
      int const k = k0 + lane + j * G;
      T const v = prod[min(k, kStreamSlots - 1)];
      s += k < k1 ? v : T(0);
    }
    for (int k = k0 + lane + kRed * G; k < k1; k += G)
      s += prod[k];
    if constexpr (G > 1)
    {
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1)
        s += __shfl_xor(s, off, G);
    }
    if (lane == 0)
      csrOut<T, MODE>(row, s, x, b, dinv, omega, y);
  }
}

template <typename T, int MODE, int G>
__global__ void __launch_bounds__(kBlock)
    k_csr_stream(int32_t const * __restrict__ blockRow, int64_t const * __restrict__ rowptr,
                 int32_t const * __restrict__ col, T const * __restrict__ val, T const * __restrict__ x,
                 T const * __restrict__ b, T const * __restrict__ dinv, T omega, T * __restrict__ y)
{
  constexpr int kSlots = kStreamSlots / kBlock; // entries per thread
  __shared__ T prod[kStreamSlots];
  unsigned const qb = xcdRowBlock();
  int64_t const r0 = blockRow[qb], r1 = blockRow[qb + 1];
  if (r0 == r1)
    return; // a window covered by a row that started earlier
  int64_t const base = rowptr[r0], end = rowptr[r1];
  int const t = threadIdx.x;
  if (end > base) // block-uniform
  {
    // every load is issued unconditionally (indices clamped into the block,
    // out-of-range slots masked to zero) so all of them are in flight at once
    int32_t cc[kSlots];
    T vv[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m)
    {
      int64_t const k = min(base + t + m * kBlock, end - 1);
      cc[m] = __builtin_nontemporal_load(col + k);
      vv[m] = __builtin_nontemporal_load(val + k);
    }
    T xv[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m)
      xv[m] = x[cc[m]];
#pragma unroll
    for (int m = 0; m < kSlots; ++m)
      prod[t + m * kBlock] = base + t + m * kBlock < end ? vv[m] * xv[m] : T(0);
  }
  __syncthreads();
  streamReduce<T, MODE, G>(prod, r0, r1, base, rowptr, x, b, dinv, omega, y);
}

// packed stream kernel: thread t takes the slots t + m kBlock of its block
// (m < 8) like the stream kernel, so value, offset and x loads of a wave are
// consecutive; the 2 high offset bits of those eight slots share one 16-bit
// word.  Slot indices are clamped into the block's entries.
template <typename T, int MODE, int G>
__global__ void __launch_bounds__(kBlock)
    k_csr_packed(int32_t const * __restrict__ blockRow, int64_t const * __restrict__ rowptr,
                 int32_t const * __restrict__ pbase, uint16_t const * __restrict__ plo,
                 uint16_t const * __restrict__ phi, T const * __restrict__ pval, T const * __restrict__ x,
                 T const * __restrict__ b, T const * __restrict__ dinv, T omega, T * __restrict__ y)
{
  constexpr int kSlots = kStreamSlots / kBlock;
  static_assert(kSlots == 8, "eight 2-bit fields per 16-bit word");
  __shared__ T prod[kStreamSlots];
  unsigned const qb = xcdRowBlock();
  int64_t const r0 = blockRow[qb], r1 = blockRow[qb + 1];
  if (r0 == r1)
    return;
  int64_t const base = rowptr[r0];
  int const len = (int)(rowptr[r1] - base);
  int const t = threadIdx.x;
  if (len > 0) // block-uniform
  {
    size_t const blk = (size_t)qb * kStreamSlots;
    uint32_t const hi = phi[(size_t)qb * kBlock + t];
    int32_t const cb = pbase[qb];
    T vv[kSlots];
    int32_t cc[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m)
    {
      int const k = min(t + m * kBlock, len - 1);
      vv[m] = __builtin_nontemporal_load(pval + blk + k);
      uint32_t const lo = __builtin_nontemporal_load(plo + blk + k);
      // a slot past the entries reads the low bits of the last entry and its
      // own high-bit field, which is zero: a column between cb and that
      // entry's column, so in range; its product is masked below
      cc[m] = cb + (int32_t)(lo | (((hi >> (2 * m)) & 3u) << 16));
    }
    T xv[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m)
      xv[m] = x[cc[m]];
#pragma unroll
    for (int m = 0; m < kSlots; ++m)
      prod[t + m * kBlock] = t + m * kBlock < len ? vv[m] * xv[m] : T(0);
  }
  __syncthreads();
  streamReduce<T, MODE, G>(prod, r0, r1, base, rowptr, x, b, dinv, omega, y);
}

// index part of the block-packed copy: per block the smallest column, per
// slot the low 16 bits of (col - base), and for thread t of the kernel the 2
// high bits of its slots t + m kBlock in bits 2m of one 16-bit word;
// bad = 1 when some block spans 2^18 columns or more
__global__ void __launch_bounds__(kBlock)
    k_pack_index(int32_t const * __restrict__ blockRow, int64_t const * __restrict__ rowptr,
                 int32_t const * __restrict__ col, int32_t * __restrict__ pbase, uint16_t * __restrict__ plo,
                 uint16_t * __restrict__ phi, int * __restrict__ bad)
{
  __shared__ int smin[kBlock], smax[kBlock];
  int64_t const r0 = blockRow[blockIdx.x], r1 = blockRow[blockIdx.x + 1];
  int64_t const base = rowptr[r0];
  int const len = (int)(rowptr[r1] - base);
  int const t = threadIdx.x;
  int mn = 0x7fffffff, mx = -1;
  for (int k = t; k < len; k += kBlock)
  {
    int const c = col[base + k];
    mn = min(mn, c);
    mx = max(mx, c);
  }
  smin[t] = mn;
  smax[t] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1)
  {
    if (t < off)
    {
      smin[t] = min(smin[t], smin[t + off]);
      smax[t] = max(smax[t], smax[t + off]);
    }
    __syncthreads();
  }
  int const cb = len > 0 ? smin[0] : 0;
  if (t == 0)
  {
    pbase[blockIdx.x] = cb;
    if (len > 0 && (int64_t)smax[0] - cb >= (1 << 18))
      atomicOr(bad, 1);
  }
  uint32_t h = 0;
  size_t const blk = (size_t)blockIdx.x * kStreamSlots;
#pragma unroll
  for (int m = 0; m < kStreamSlots / kBlock; ++m)
  {
    int const k = t + m * kBlock;
    uint32_t const off = k < len ? (uint32_t)(col[base + k] - cb) : 0u;
    plo[blk + k] = (uint16_t)(off & 0xffffu);
    h |= ((off >> 16) & 3u) << (2 * m);
  }
  phi[(size_t)blockIdx.x * kBlock + t] = (uint16_t)h;
}

// value part of the block-packed copy (slots past a block's entries: 0)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_pack_values(int32_t const * __restrict__ blockRow, int64_t const * __restrict__ rowptr,
                  T const * __restrict__ val, T * __restrict__ pval)
{
  int64_t const r0 = blockRow[blockIdx.x], r1 = blockRow[blockIdx.x + 1];
  int64_t const base = rowptr[r0];
  int const len = (int)(rowptr[r1] - base);
  size_t const slot = (size_t)blockIdx.x * kStreamSlots;
  for (int k = threadIdx.x; k < kStreamSlots; k += kBlock)
    pval[slot + k] = k < len ? val[base + k] : T(0);
}

__global__ void k_max_row_len(int64_t nrows, int64_t const * __restrict__ rowptr, int * __restrict__ out)
{
  __shared__ int sh[kBlock];
  int m = 0;
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kBlock)
    m = max(m, (int)min<int64_t>(rowptr[r + 1] - rowptr[r], 1 << 30));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1)
  {
    if ((int)threadIdx.x < off)
      sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[blockIdx.x] = sh[0];
}

// blockRow[q] = first row whose first entry is >= q * cap (q = 0 .. nb);
// row r takes the q with rowptr[r-1] < q * cap <= rowptr[r] (rowptr[-1] = -inf,
// rowptr[nrows] = +inf), so every q is written exactly once
__global__ void k_block_rows(int64_t nrows, int64_t const * __restrict__ rowptr, int64_t cap, int64_t nb,
                             int32_t * __restrict__ blockRow)
{
  int64_t const r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r > nrows)
    return;
  int64_t const lo = r == 0 ? 0 : rowptr[r - 1] / cap + 1;
  int64_t const hi = r == nrows ? nb : min(rowptr[r] / cap, nb);
  for (int64_t q = lo; q <= hi; ++q)
    blockRow[q] = (int32_t)r;
}

// lanes per row from the mean row length; PXG_SPMV_LPR (4..64) overrides the
// choice for launch-shape tuning
int lanesPerRow(double mean)
{
  static int const forced = []
  {
    char const * e = std::getenv("PXG_SPMV_LPR");
    int v = e ? std::atoi(e) : 0;
    return (v == 4 || v == 8 || v == 16 || v == 32 || v == 64) ? v : 0;
  }();
  if (forced)
    return forced;
  if (mean > 192)
    return 64;
  if (mean > 96)
    return 32;
  if (mean > 40)
    return 16;
  if (mean > 12)
    return 8;
  return 4;
}

// PXG_SPMV_KERNEL=rowgroup selects the row-group kernel (launch-shape tuning)
bool rowGroupForced()
{
  char const * e = std::getenv("PXG_SPMV_KERNEL");
  return e && e[0] == 'r';
}

template <typename T>
T const * packValOf(CsrView const & A)
{
  if constexpr (sizeof(T) == 8)
    return A.pval;
  else
    return A.pval32;
}

template <typename T, int MODE>
void launchCsr(CsrView const & A, T const * val, T const * x, T const * b, T const * dinv, double omega, T * y,
               hipStream_t s)
{
  if (A.nrows == 0)
    return;
  T const om = (T)omega;
  double const mean = A.nnz > 0 ? (double)A.nnz / (double)A.nrows : 1.0;
  T const * pval = packValOf<T>(A);
  if (A.blockRow && A.nblocks > 0 && pval && A.plo && !rowGroupForced())
  {
    unsigned const g = (unsigned)A.nblocks;
    if (mean > 48)
      k_csr_packed<T, MODE, 8><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.pbase, A.plo, A.phi, pval, x, b, dinv,
                                                    om, y);
    else if (mean > 20)
      k_csr_packed<T, MODE, 4><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.pbase, A.plo, A.phi, pval, x, b, dinv,
                                                    om, y);
    else if (mean > 8)
      k_csr_packed<T, MODE, 2><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.pbase, A.plo, A.phi, pval, x, b, dinv,
                                                    om, y);
    else
      k_csr_packed<T, MODE, 1><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.pbase, A.plo, A.phi, pval, x, b, dinv,
                                                    om, y);
    PXG_HIP(hipGetLastError());
    return;
  }
  if (A.blockRow && A.nblocks > 0 && !rowGroupForced())
  {
    unsigned const g = (unsigned)A.nblocks;
    // lanes per row in the reduction: about 4-8 entries per lane
    if (mean > 48)
      k_csr_stream<T, MODE, 8><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.col, val, x, b, dinv, om, y);
    else if (mean > 20)
      k_csr_stream<T, MODE, 4><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.col, val, x, b, dinv, om, y);
    else if (mean > 8)
      k_csr_stream<T, MODE, 2><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.col, val, x, b, dinv, om, y);
    else
      k_csr_stream<T, MODE, 1><<<g, kBlock, 0, s>>>(A.blockRow, A.rowptr, A.col, val, x, b, dinv, om, y);
    PXG_HIP(hipGetLastError());
    return;
  }
  int lpr = lanesPerRow(mean);
  unsigned const g = blocksFor(A.nrows * lpr);
  switch (lpr)
  {
  case 4: k_csr<T, 4, MODE><<<g, kBlock, 0, s>>>(A.nrows, A.rowptr, A.col, val, x, b, dinv, om, y); break;
  case 8: k_csr<T, 8, MODE><<<g, kBlock, 0, s>>>(A.nrows, A.rowptr, A.col, val, x, b, dinv, om, y); break;
  case 16: k_csr<T, 16, MODE><<<g, kBlock, 0, s>>>(A.nrows, A.rowptr, A.col, val, x, b, dinv, om, y); break;
  case 32: k_csr<T, 32, MODE><<<g, kBlock, 0, s>>>(A.nrows, A.rowptr, A.col, val, x, b, dinv, om, y); break;
  default: k_csr<T, 64, MODE><<<g, kBlock, 0, s>>>(A.nrows, A.rowptr, A.col, val, x, b, dinv, om, y); break;
  }
  PXG_HIP(hipGetLastError());
}

template <typename T>
T const * valOf(CsrView const & A);
template <>
double const * valOf<double>(CsrView const & A)
{
  return A.val;
}
template <>
float const * valOf<float>(CsrView const & A)
{
  return A.val32;
}
template <typename T>
T const * dinvOf(CsrView const & A);
template <>
double const * dinvOf<double>(CsrView const & A)
{
  return A.dinv;
}
template <>
float const * dinvOf<float>(CsrView const & A)
{
  return A.dinv32;
}

// deterministic dot product: fixed grid, per-block tree in FP64, then a
// single-block pass over the partials
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_dot_partial(int64_t n, T const * __restrict__ a, T const * __restrict__ b, double * __restrict__ partial)
{
  __shared__ double sh[kBlock];
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    s += (double)a[i] * (double)b[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1)
  {
    if ((int)threadIdx.x < off)
      sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(kBlock) k_dot_final(int nparts, double const * __restrict__ partial, double * out)
{
  __shared__ double sh[kBlock];
  double s = 0.;
  for (int i = threadIdx.x; i < nparts; i += kBlock)
    s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1)
  {
    if ((int)threadIdx.x < off)
      sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    *out = sh[0];
}

template <typename T>
__global__ void k_fill(int64_t n, T * x, T v)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    x[i] = v;
}

template <typename Tout, typename Tin>
__global__ void k_convert(int64_t n, Tin const * __restrict__ x, Tout * __restrict__ y)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    y[i] = (Tout)x[i];
}

template <typename T>
__global__ void k_jacobi_diag(int64_t n, T const * __restrict__ dinv, T const * __restrict__ b, T * __restrict__ x,
                              T omega)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    x[i] = omega * dinv[i] * b[i];
}

template <typename T>
__global__ void k_cg_update(int64_t n, T * __restrict__ x, T const * __restrict__ p, T * __restrict__ r,
                            T const * __restrict__ q, double const * __restrict__ sc, int ia, int ib)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  // a zero denominator means the search direction vanished (exact convergence)
  T const alpha = sc[ib] != 0. ? (T)(sc[ia] / sc[ib]) : T(0);
  x[i] += alpha * p[i];
  r[i] -= alpha * q[i];
}

// FP64 iterate, FP32 residual: x += alpha p (in FP64) ; r -= alpha q
__global__ void k_cg_update_mixed(int64_t n, double * __restrict__ x, float const * __restrict__ p,
                                  float * __restrict__ r, float const * __restrict__ q, double const * __restrict__ sc,
                                  int ia, int ib)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  double const alpha = sc[ib] != 0. ? sc[ia] / sc[ib] : 0.;
  x[i] += alpha * (double)p[i];
  r[i] -= (float)alpha * q[i];
}

template <typename T>
__global__ void k_cg_direction(int64_t n, T * __restrict__ p, T const * __restrict__ z, double const * __restrict__ sc,
                               int ia, int ib)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n)
    return;
  T const beta = sc[ib] != 0. ? (T)(sc[ia] / sc[ib]) : T(0);
  p[i] = z[i] + beta * p[i];
}

template <typename T, typename U>
__global__ void k_add_to(int64_t n, T * __restrict__ x, U const * __restrict__ y)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    x[i] += (T)y[i];
}

__global__ void k_axpby(int64_t n, double a, double const * __restrict__ x, double b, double * __restrict__ y)
{
  int64_t const i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    y[i] = a * x[i] + b * y[i];
}

} // namespace

void buildRowBlocks(int64_t nrows, int64_t nnz, int64_t const * rowptr, DevBuf<int32_t> & blockRow,
                    int64_t & nblocks, hipStream_t s)
{
  nblocks = 0;
  if (nrows <= 0 || nnz <= 0 || nrows >= (int64_t)1 << 31)
    return;
  constexpr int kParts = 256;
  DevBuf<int> part;
  part.resize(kParts);
  k_max_row_len<<<kParts, kBlock, 0, s>>>(nrows, rowptr, part.p);
  PXG_HIP(hipGetLastError());
  int h[kParts];
  PXG_HIP(hipMemcpyAsync(h, part.p, sizeof(h), hipMemcpyDeviceToHost, s));
  PXG_HIP(hipStreamSynchronize(s));
  int maxLen = 0;
  for (int v: h)
    maxLen = std::max(maxLen, v);
  if (maxLen > kStreamSlots / 2)
    return; // rows too long for one block's staging: row-group kernel
  // a window of cap entries plus the tail of its last row fits the staging
  int64_t const cap = kStreamSlots - std::max(maxLen, 1);
  int64_t const nb = nnz / cap + 1;
  blockRow.resize(nb + 1);
  k_block_rows<<<blocksFor(nrows + 1), kBlock, 0, s>>>(nrows, rowptr, cap, nb, blockRow.p);
  PXG_HIP(hipGetLastError());
  PXG_HIP(hipStreamSynchronize(s));
  nblocks = nb;
}

void ensureRowBlocks(pxg_ctx * c)
{
  if (!c->closed || c->blockStamp == c->patStamp)
    return;
  buildRowBlocks(c->ndofs, c->nnz, c->rowptr.p, c->blockRow, c->nStreamBlocks, c->stream);
  c->blockStamp = c->patStamp;
}

void ensurePack(pxg_ctx * c, int precision)
{
  if (!c->packWanted || !c->closed)
    return;
  ensureRowBlocks(c);
  if (c->nStreamBlocks <= 0)
    return;
  hipStream_t s = c->stream;
  unsigned const g = (unsigned)c->nStreamBlocks;
  size_t const slots = (size_t)c->nStreamBlocks * kStreamSlots;
  if (c->packStamp != c->patStamp)
  {
    c->packBase.resize(c->nStreamBlocks);
    c->packLo.resize(slots);
    c->packHi.resize((size_t)c->nStreamBlocks * kBlock);
    c->packFlag.resize(1);
    PXG_HIP(hipMemsetAsync(c->packFlag.p, 0, sizeof(int), s));
    k_pack_index<<<g, kBlock, 0, s>>>(c->blockRow.p, c->rowptr.p, c->col.p, c->packBase.p, c->packLo.p,
                                      c->packHi.p, c->packFlag.p);
    PXG_HIP(hipGetLastError());
    int bad = 1;
    PXG_HIP(hipMemcpyAsync(&bad, c->packFlag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    PXG_HIP(hipStreamSynchronize(s));
    c->packOk = !bad;
    c->packStamp = c->patStamp;
    c->packValStamp = c->packVal32Stamp = 0;
  }
  if (!c->packOk)
    return;
  if (precision == PXG_FP32 && c->packVal32Stamp != c->stamp)
  {
    ensureVal32(c);
    c->packVal32.resize(slots);
    k_pack_values<float><<<g, kBlock, 0, s>>>(c->blockRow.p, c->rowptr.p, c->val32.p, c->packVal32.p);
    PXG_HIP(hipGetLastError());
    c->packVal32Stamp = c->stamp;
  }
  if (precision == PXG_FP64 && c->packValStamp != c->stamp)
  {
    ensureVal64(c);
    c->packVal.resize(slots);
    k_pack_values<double><<<g, kBlock, 0, s>>>(c->blockRow.p, c->rowptr.p, c->val.p, c->packVal.p);
    PXG_HIP(hipGetLastError());
    c->packValStamp = c->stamp;
  }
}

void ensureOperator(pxg_ctx * c, int precision)
{
  if (precision == PXG_FP32)
    ensureVal32(c);
  else
    ensureVal64(c);
  ensureRowBlocks(c);
  ensurePack(c, precision);
}

CsrView viewOf(pxg_ctx const * c)
{
  CsrView v;
  v.nrows = c->ndofs;
  v.ncols = c->ndofs;
  v.nnz = c->nnz;
  v.rowptr = c->rowptr.p;
  v.col = c->col.p;
  v.val = c->val.p;
  v.val32 = c->val32.p;
  v.dinv = c->dinv.p;
  v.dinv32 = c->dinv32.p;
  bool const blocks = c->nStreamBlocks > 0 && c->blockStamp == c->patStamp;
  v.blockRow = blocks ? c->blockRow.p : nullptr;
  v.nblocks = blocks ? c->nStreamBlocks : 0;
  bool const pack = blocks && c->packWanted && c->packOk && c->packStamp == c->patStamp;
  if (pack)
  {
    v.pbase = c->packBase.p;
    v.plo = c->packLo.p;
    v.phi = c->packHi.p;
    v.pval = c->packValStamp == c->stamp ? c->packVal.p : nullptr;
    v.pval32 = c->packVal32Stamp == c->stamp ? c->packVal32.p : nullptr;
  }
  return v;
}

template <typename T>
void spmv(CsrView const & A, T const * x, T * y, hipStream_t s)
{
  launchCsr<T, kModeApply>(A, valOf<T>(A), x, nullptr, nullptr, 0., y, s);
}
template <typename T>
void residual(CsrView const & A, T const * b, T const * x, T * y, hipStream_t s)
{
  launchCsr<T, kModeResidual>(A, valOf<T>(A), x, b, nullptr, 0., y, s);
}
template <typename T>
void jacobiStep(CsrView const & A, T const * b, T const * x, T * y, double omega, hipStream_t s)
{
  launchCsr<T, kModeJacobi>(A, valOf<T>(A), x, b, dinvOf<T>(A), omega, y, s);
}
template <typename T>
void spmvAdd(CsrView const & A, T const * x, T * y, hipStream_t s)
{
  launchCsr<T, kModeAdd>(A, valOf<T>(A), x, nullptr, nullptr, 0., y, s);
}

template <typename T>
void dot(int64_t n, T const * a, T const * b, double * outDev, double * partial, hipStream_t s)
{
  k_dot_partial<T><<<kDotBlocks, kBlock, 0, s>>>(n, a, b, partial);
  k_dot_final<<<1, kBlock, 0, s>>>(kDotBlocks, partial, outDev);
  PXG_HIP(hipGetLastError());
}

template <typename T>
void fill(int64_t n, T * x, T v, hipStream_t s)
{
  k_fill<T><<<blocksFor(n), kBlock, 0, s>>>(n, x, v);
  PXG_HIP(hipGetLastError());
}

template <typename Tout, typename Tin>
void convert(int64_t n, Tin const * x, Tout * y, hipStream_t s)
{
  k_convert<Tout, Tin><<<blocksFor(n), kBlock, 0, s>>>(n, x, y);
  PXG_HIP(hipGetLastError());
}

template <typename T>
void jacobiDiag(int64_t n, T const * dinv, T const * b, T * x, double omega, hipStream_t s)
{
  k_jacobi_diag<T><<<blocksFor(n), kBlock, 0, s>>>(n, dinv, b, x, (T)omega);
  PXG_HIP(hipGetLastError());
}

template <typename T>
void cgUpdate(int64_t n, T * x, T const * p, T * r, T const * q, double const * sc, int ia, int ib, hipStream_t s)
{
  k_cg_update<T><<<blocksFor(n), kBlock, 0, s>>>(n, x, p, r, q, sc, ia, ib);
  PXG_HIP(hipGetLastError());
}

void cgUpdateMixed(int64_t n, double * x, float const * p, float * r, float const * q, double const * sc, int ia,
                   int ib, hipStream_t s)
{
  k_cg_update_mixed<<<blocksFor(n), kBlock, 0, s>>>(n, x, p, r, q, sc, ia, ib);
  PXG_HIP(hipGetLastError());
}

template <typename T>
void cgDirection(int64_t n, T * p, T const * z, double const * sc, int ia, int ib, hipStream_t s)
{
  k_cg_direction<T><<<blocksFor(n), kBlock, 0, s>>>(n, p, z, sc, ia, ib);
  PXG_HIP(hipGetLastError());
}

template <typename T, typename U>
void addTo(int64_t n, T * x, U const * y, hipStream_t s)
{
  k_add_to<T, U><<<blocksFor(n), kBlock, 0, s>>>(n, x, y);
  PXG_HIP(hipGetLastError());
}

void axpby(int64_t n, double a, double const * x, double b, double * y, hipStream_t s)
{
  k_axpby<<<blocksFor(n), kBlock, 0, s>>>(n, a, x, b, y);
  PXG_HIP(hipGetLastError());
}

#define PXG_INST(T)                                                                                \
  template void spmv<T>(CsrView const &, T const *, T *, hipStream_t);                             \
  template void residual<T>(CsrView const &, T const *, T const *, T *, hipStream_t);              \
  template void jacobiStep<T>(CsrView const &, T const *, T const *, T *, double, hipStream_t);    \
  template void spmvAdd<T>(CsrView const &, T const *, T *, hipStream_t);                          \
  template void dot<T>(int64_t, T const *, T const *, double *, double *, hipStream_t);            \
  template void fill<T>(int64_t, T *, T, hipStream_t);                                             \
  template void jacobiDiag<T>(int64_t, T const *, T const *, T *, double, hipStream_t);            \
  template void cgUpdate<T>(int64_t, T *, T const *, T *, T const *, double const *, int, int,     \
                            hipStream_t);                                                          \
  template void cgDirection<T>(int64_t, T *, T const *, double const *, int, int, hipStream_t);
PXG_INST(double)
PXG_INST(float)
template void convert<float, double>(int64_t, double const *, float *, hipStream_t);
template void convert<double, float>(int64_t, float const *, double *, hipStream_t);
template void addTo<double, float>(int64_t, double *, float const *, hipStream_t);
template void addTo<double, double>(int64_t, double *, double const *, hipStream_t);
template void addTo<float, float>(int64_t, float *, float const *, hipStream_t);

} // namespace pxg
