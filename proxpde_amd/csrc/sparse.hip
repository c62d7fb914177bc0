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

template <typename T, int MODE, int G>
__global__ void __launch_bounds__(kBlock)
    k_csr_stream(int32_t const * __restrict__ blockRow, int64_t const * __restrict__ rowptr,
                 int32_t const * __restrict__ col, T const * __restrict__ val, T const * __restrict__ x,
                 T const * __restrict__ b, T const * __restrict__ dinv, T omega, T * __restrict__ y)
{
  constexpr int kSlots = kStreamSlots / kBlock; // entries per thread
  // G lanes per row in the reduction (chosen from the mean row length)
  __shared__ T prod[kStreamSlots];
  int64_t const r0 = blockRow[blockIdx.x], r1 = blockRow[blockIdx.x + 1];
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
  int const lane = t % G;
  // the lanes of one group run the same rows, so the group stays converged
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

template <typename T, int MODE>
void launchCsr(CsrView const & A, T const * val, T const * x, T const * b, T const * dinv, double omega, T * y,
               hipStream_t s)
{
  if (A.nrows == 0)
    return;
  T const om = (T)omega;
  double const mean = A.nnz > 0 ? (double)A.nnz / (double)A.nrows : 1.0;
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
  if (!c->closed || c->blockStamp == c->stamp)
    return;
  buildRowBlocks(c->ndofs, c->nnz, c->rowptr.p, c->blockRow, c->nStreamBlocks, c->stream);
  c->blockStamp = c->stamp;
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
  bool const blocks = c->nStreamBlocks > 0 && c->blockStamp == c->stamp;
  v.blockRow = blocks ? c->blockRow.p : nullptr;
  v.nblocks = blocks ? c->nStreamBlocks : 0;
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
