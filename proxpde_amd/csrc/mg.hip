// Geometric multigrid-preconditioned CG on the assembled level matrices.
//
// The reference keeps only the transfer operators in the library
// (include/multigrid.hpp:12-182: Prolongator / Restrictor from embedding
// matrices) and its V-cycle in a test (test/multigrid.cpp:18-112, GMRES
// smoother + LU coarse solve).  Here the cycle is the symmetric one a CG
// preconditioner needs:
//   pre-smooth  nu1 weighted-Jacobi sweeps from x = 0
//   residual    r = b - A x
//   restrict    b_c = P^T r          (R = P^T, not the row-normalised R)
//   recurse     e_c = V(b_c), coarsest level: direct solve in FP64 (dense
//               inverse of its free-dof block, made on the GPU by Gauss-Jordan
//               elimination whenever the coarse matrix changes and applied as
//               one matrix-vector product, like the reference test's LU coarse
//               solve); above kDenseMax free dofs Jacobi-CG to coarse_rtol
//   prolongate  x += P e_c
//   post-smooth nu2 sweeps
// P(f, c) is the coarse basis function c evaluated at fine dof point f in the
// fine element's parent (the values of the reference's embedding matrices),
// with rows of constrained fine dofs and columns of constrained coarse dofs
// removed so Dirichlet rows never receive a coarse correction.
// FP64 mode runs everything in double.  FP32 mode runs the preconditioned CG
// and the cycle in float (FP64 dot products, FP64 coarsest solve) on an FP64
// iterate, refined by replacing the CG residual with the FP64 residual
// b - A x (see solveIR).
#include "kernels.hpp"

#include <cstdlib>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

namespace pxg
{
void mfApply(pxg_ctx * c, double const * kappa, int nb, double const * x, double * y, bool useBC);
}

using namespace pxg;

struct pxg_mg
{
  std::string err;
  int nlev = 0;
  std::vector<pxg_ctx *> ctx;
  std::vector<std::array<int, 3>> dims;
  double omega = 2. / 3., coarseRtol = 1e-10;
  int nu1 = 2, nu2 = 2;
  hipStream_t s = nullptr;
  struct Level
  {
    int64_t n = 0;
    // prolongation from level+1 and its transpose
    std::vector<int64_t> hPrp;
    std::vector<int32_t> hPcol;
    std::vector<double> hPval;
    DevBuf<int64_t> Prp, Rrp;
    DevBuf<int32_t> Pcol, Rcol;
    DevBuf<double> Pval, Rval;
    DevBuf<float> Pval32, Rval32;
    DevBuf<int32_t> Pblk, Rblk; // row blocks of the stream SpMV kernel
    CsrView P, R;
    DevBuf<double> b, x, t, r;
    DevBuf<float> b32, x32, t32, r32;
    // applied matrix-free (mode PXG_FP32_MF): FP32 per-point metrics of the
    // level's current kappa field, element buffer, inverse diagonal
    uint64_t mfKappaStamp = 0;
    DevBuf<float> met32, mfBuf32, dinvMf32;
    // element cubes of the level's mesh: slot-major metrics and the partial
    // values of the one-vector operator (met32 and mfBuf32 then released)
    bool cubes = false;
    DevBuf<float> metC, partF;
  };
  std::vector<Level> lv;
  DevBuf<double> sc, partial;
  DevBuf<double> r, z, p, q;
  DevBuf<float> r32, z32, p32, q32;
  // coarsest-level CG vectors
  DevBuf<double> cr, cz, cp, cq;
  // coarsest-level direct solve: inverse of the free-dof block
  bool dense = false;
  uint64_t denseStamp = 0;
  int64_t nfree = 0;
  DevBuf<int32_t> freeIdx;  // coarse dof -> free index, -1 for constrained dofs
  DevBuf<int32_t> freeDofs; // free index -> coarse dof
  DevBuf<double> inv, invTmp;
  DevBuf<double> gjP, gjR; // pivot-block inverse and scaled pivot rows of the blocked inversion
  DevBuf<int> flag;
  // levels [0, mfLevels) applied matrix-free in this solve (mode
  // PXG_FP32_MF: all but the coarsest); element buffer of the FP64
  // matrix-free residual on level 0
  int mfLevels = 0;
  DevBuf<double> mfBuf64;
  // orders the cycle stream after work issued on the level contexts' own
  // streams (value casts, packed copies made on first use)
  hipEvent_t join = nullptr;
  // residual norms read one cycle late: copied to pinned host memory behind
  // an event, read while the next preconditioner cycle runs
  double * hNorm = nullptr;
  hipEvent_t normEv = nullptr;
  ~pxg_mg()
  {
    if (join)
      (void)hipEventDestroy(join);
    if (normEv)
      (void)hipEventDestroy(normEv);
    if (hNorm)
      (void)hipHostFree(hNorm);
  }
};

namespace
{
template <typename T>
struct Buf;
template <>
struct Buf<double>
{
  static double * b(pxg_mg::Level & L) { return L.b.p; }
  static double * x(pxg_mg::Level & L) { return L.x.p; }
  static double * t(pxg_mg::Level & L) { return L.t.p; }
  static double * r(pxg_mg::Level & L) { return L.r.p; }
};
template <>
struct Buf<float>
{
  static float * b(pxg_mg::Level & L) { return L.b32.p; }
  static float * x(pxg_mg::Level & L) { return L.x32.p; }
  static float * t(pxg_mg::Level & L) { return L.t32.p; }
  static float * r(pxg_mg::Level & L) { return L.r32.p; }
};

// work of the levels issued on their own streams (ensureOperator,
// ensureVal64) completes before the cycle stream reads it
void joinLevels(pxg_mg * m)
{
  for (auto * c: m->ctx)
    if (c->stream != m->s)
    {
      if (!m->join)
        PXG_HIP(hipEventCreateWithFlags(&m->join, hipEventDisableTiming));
      PXG_HIP(hipEventRecord(m->join, c->stream));
      PXG_HIP(hipStreamWaitEvent(m->s, m->join, 0));
    }
}

// largest free-dof block inverted densely on the coarsest level (8 n^2 bytes,
// n^3 work in n / 32 blocked Gauss-Jordan steps)
constexpr int64_t kDenseMax = 4096;

__global__ void k_dense_scatter(int64_t n, int32_t const * __restrict__ freeIdx, int64_t const * __restrict__ rowptr,
                                int32_t const * __restrict__ col, double const * __restrict__ val, int64_t nf,
                                double * __restrict__ D)
{
  int64_t const r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n)
    return;
  int64_t const i = freeIdx[r];
  if (i < 0)
    return;
  for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
  {
    int64_t const j = freeIdx[col[k]];
    if (j >= 0)
      D[i * nf + j] = val[k];
  }
}

// one Gauss-Jordan step on pivot k, A -> B (out of place, so every thread
// reads the pivot row and column as they were before the step).  After steps
// 0..n-1 an SPD matrix has become its inverse; the pivots of an SPD matrix
// stay positive, so no pivoting is needed.
// Blocked Gauss-Jordan inversion in place (the sweep operator; no pivoting,
// the free block of an SPD coarse matrix), pivots in blocks K of kGjBlock:
//   P = A_KK^-1,  R = P A_K,  then out of place
//   B_KK = P,  B_Kj = R_j,  B_iK = -A_iK P,  B_ij = A_ij - A_iK R_j  (i, j not in K)
// Three launches per block of pivots instead of one per pivot.
constexpr int kGjBlock = 32;

// P = A_KK^-1 of the pivot block [k0, k0 + nb): one block, unblocked
// Gauss-Jordan in LDS
__global__ void __launch_bounds__(1024) k_gj_pivot(int64_t nf, int64_t k0, int nb, double const * __restrict__ A,
                                                  double * __restrict__ P)
{
  __shared__ double d[kGjBlock][kGjBlock + 1];
  __shared__ double prow[kGjBlock], pcol[kGjBlock];
  int const t = threadIdx.x;
  for (int e = t; e < nb * nb; e += blockDim.x)
    d[e / nb][e % nb] = A[(k0 + e / nb) * nf + k0 + e % nb];
  __syncthreads();
  for (int k = 0; k < nb; ++k)
  {
    double const rp = 1. / d[k][k];
    if (t < nb)
    {
      prow[t] = d[k][t];
      pcol[t] = d[t][k];
    }
    __syncthreads();
    for (int e = t; e < nb * nb; e += blockDim.x)
    {
      int const i = e / nb, j = e % nb;
      double v;
      if (i == k)
        v = (j == k) ? rp : prow[j] * rp;
      else if (j == k)
        v = -pcol[i] * rp;
      else
        v = d[i][j] - pcol[i] * (prow[j] * rp);
      d[i][j] = v;
    }
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += blockDim.x)
    P[e] = d[e / nb][e % nb];
}

// R = P A_K (nb x nf; columns in K unused)
__global__ void k_gj_rows(int64_t nf, int64_t k0, int nb, double const * __restrict__ A,
                          double const * __restrict__ P, double * __restrict__ R)
{
  int64_t const t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nb * nf)
    return;
  int64_t const m = t / nf, j = t - m * nf;
  double v = 0.;
  for (int l = 0; l < nb; ++l)
    v += P[m * nb + l] * A[(k0 + l) * nf + j];
  R[t] = v;
}

// B from A, P and R (see above)
__global__ void k_gj_update(int64_t nf, int64_t k0, int nb, double const * __restrict__ A,
                            double const * __restrict__ P, double const * __restrict__ R, double * __restrict__ B)
{
  int64_t const t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nf * nf)
    return;
  int64_t const i = t / nf, j = t - i * nf;
  bool const ik = i >= k0 && i < k0 + nb, jk = j >= k0 && j < k0 + nb;
  double v;
  if (ik)
    v = jk ? P[(i - k0) * nb + (j - k0)] : R[(i - k0) * nf + j];
  else if (jk)
  {
    v = 0.;
    for (int l = 0; l < nb; ++l)
      v -= A[i * nf + k0 + l] * P[l * nb + (j - k0)];
  }
  else
  {
    v = A[t];
    for (int l = 0; l < nb; ++l)
      v -= A[i * nf + k0 + l] * R[l * nf + j];
  }
  B[t] = v;
}

// one block: flag = 1 if a diagonal entry of the inverse is not positive and
// finite (the matrix was not SPD)
__global__ void k_inverse_check(int64_t nf, double const * __restrict__ inv, int * __restrict__ flag)
{
  int bad = 0;
  for (int64_t i = threadIdx.x; i < nf; i += blockDim.x)
  {
    double const d = inv[i * nf + i];
    bad |= !(d > 0. && isfinite(d));
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0)
    *flag = bad;
}

// x = A^-1 b on the coarsest level: one wave per row; constrained rows are
// diagonal (x = b / a_rr), free rows take a dot with their inverse row
__global__ void k_dense_apply(int64_t n, int64_t nf, int32_t const * __restrict__ freeIdx,
                              int32_t const * __restrict__ freeDofs, double const * __restrict__ inv,
                              double const * __restrict__ dinv, double const * __restrict__ b, double * __restrict__ x)
{
  int const lane = threadIdx.x & 63;
  int64_t const r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (r >= n)
    return;
  int64_t const i = freeIdx[r];
  if (i < 0)
  {
    if (lane == 0)
      x[r] = dinv[r] * b[r];
    return;
  }
  double const * row = inv + i * nf;
  double acc = 0.;
  for (int64_t j = lane; j < nf; j += 64)
    acc += row[j] * b[freeDofs[j]];
  for (int o = 32; o > 0; o >>= 1)
    acc += __shfl_xor(acc, o);
  if (lane == 0)
    x[r] = acc;
}

double readScalar(pxg_mg * m, int i)
{
  double h = 0.;
  PXG_HIP(hipMemcpyAsync(&h, m->sc.p + i, sizeof(double), hipMemcpyDeviceToHost, m->s));
  PXG_HIP(hipStreamSynchronize(m->s));
  return h;
}

// an asynchronous read of scalar i (takeScalar waits for it): the host
// reads the norm of one iteration while the stream runs the cycle after it,
// so the stream is not left empty while the host decides
void postScalar(pxg_mg * m, int i)
{
  if (!m->hNorm)
    PXG_HIP(hipHostMalloc((void **)&m->hNorm, sizeof(double), hipHostMallocDefault));
  if (!m->normEv)
    PXG_HIP(hipEventCreateWithFlags(&m->normEv, hipEventDisableTiming));
  PXG_HIP(hipMemcpyAsync(m->hNorm, m->sc.p + i, sizeof(double), hipMemcpyDeviceToHost, m->s));
  PXG_HIP(hipEventRecord(m->normEv, m->s));
}

double takeScalar(pxg_mg * m)
{
  PXG_HIP(hipEventSynchronize(m->normEv));
  return *m->hNorm;
}

// the residual norm expected of the next iteration from the last two (the
// cycle's contraction); a test that this may pass is made before the cycle
double nextNormGuess(double r1, double r2) { return r2 > 0. ? r1 * (r1 / r2) : r1; }
constexpr double kLagSafety = 3.;

void copyScalar(pxg_mg * m, int dst, int src)
{
  PXG_HIP(hipMemcpyAsync(m->sc.p + dst, m->sc.p + src, sizeof(double), hipMemcpyDeviceToDevice, m->s));
}

// host construction of P for one level pair (structured hex meshes)
void buildProlongation(pxg_mg * m, int l)
{
  pxg_ctx * F = m->ctx[l];
  pxg_ctx * Cc = m->ctx[l + 1];
  auto const & df = m->dims[l];
  auto const & dc = m->dims[l + 1];
  PXG_REQUIRE(F->fe.geo == kHex && Cc->fe.geo == kHex && F->fe.feType == Cc->fe.feType, PXG_E_INVALID,
              "multigrid levels must be hex meshes of one FE type");
  PXG_REQUIRE(df[0] == 2 * dc[0] && df[1] == 2 * dc[1] && df[2] == 2 * dc[2], PXG_E_INVALID,
              "element counts must halve per level");
  PXG_REQUIRE((int64_t)df[0] * df[1] * df[2] == F->nelem && (int64_t)dc[0] * dc[1] * dc[2] == Cc->nelem,
              PXG_E_INVALID, "level dims do not match the meshes");
  hipStream_t s = m->s;
  FeInfo const & fe = feInfo(F->fe.feType);
  int const n = fe.n;
  // host copies of the setup data
  std::vector<double> xf(3 * F->npts), xc(3 * Cc->npts);
  std::vector<uint32_t> cf(F->nelem * 8), cc(Cc->nelem * 8), ef(F->nelem * n), ec(Cc->nelem * n);
  std::vector<uint8_t> bf(F->ndofs), bcn(Cc->ndofs);
  F->xyz.download(xf.data(), xf.size(), F->stream);
  F->conn.download(cf.data(), cf.size(), F->stream);
  F->elemMap.download(ef.data(), ef.size(), F->stream);
  if (F->hasBC)
    F->bcFlag.download(bf.data(), bf.size(), F->stream);
  Cc->xyz.download(xc.data(), xc.size(), Cc->stream);
  Cc->conn.download(cc.data(), cc.size(), Cc->stream);
  Cc->elemMap.download(ec.data(), ec.size(), Cc->stream);
  if (Cc->hasBC)
    Cc->bcFlag.download(bcn.data(), bcn.size(), Cc->stream);
  PXG_HIP(hipStreamSynchronize(F->stream));
  PXG_HIP(hipStreamSynchronize(Cc->stream));
  int64_t const nf = F->ndofs;
  std::vector<int64_t> rowStart(nf + 1, 0);
  std::vector<std::vector<std::pair<int32_t, double>>> rows(nf);
  std::vector<uint8_t> seen(nf, 0);
  double vf[24], vc[24], pf[27 * 3];
  double const zero[3] = {0., 0., 0.};
  for (int64_t e = 0; e < F->nelem; ++e)
  {
    int64_t const i = e % df[0], j = (e / df[0]) % df[1], k = e / ((int64_t)df[0] * df[1]);
    int64_t const pe = (i / 2) + (int64_t)dc[0] * ((j / 2) + (int64_t)dc[1] * (k / 2));
    for (int v = 0; v < 8; ++v)
      for (int d = 0; d < 3; ++d)
      {
        vf[3 * v + d] = xf[3ull * cf[e * 8 + v] + d];
        vc[3 * v + d] = xc[3ull * cc[pe * 8 + v] + d];
      }
    fe.dofPoints(vf, pf);
    // affine map of the parent parallelepiped: x = mid + J xi
    double mid[3] = {0., 0., 0.}, J[3][3] = {{0.}};
    for (int v = 0; v < 8; ++v)
    {
      double g[3];
      geoGrad(kHex, v, zero, g);
      for (int r = 0; r < 3; ++r)
      {
        mid[r] += 0.125 * vc[3 * v + r];
        for (int c = 0; c < 3; ++c)
          J[r][c] += vc[3 * v + r] * g[c];
      }
    }
    double const det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int a = 0; a < n; ++a)
    {
      uint32_t const dofF = ef[e * n + a];
      if (seen[dofF])
        continue;
      seen[dofF] = 1;
      if (bf[dofF])
        continue; // constrained fine row: no coarse correction
      double xi[3];
      for (int r = 0; r < 3; ++r)
      {
        double t = 0.;
        for (int c = 0; c < 3; ++c)
          t += Ji[r][c] * (pf[3 * a + c] - mid[c]);
        double const h = std::round(2. * t);
        xi[r] = std::fabs(2. * t - h) < 1e-8 ? 0.5 * h : t;
      }
      auto & row = rows[dofF];
      for (int b = 0; b < n; ++b)
      {
        uint32_t const dofC = ec[pe * n + b];
        if (bcn[dofC])
          continue;
        double const val = fe.phi(b, xi);
        if (std::fabs(val) > 1e-12)
          row.push_back({(int32_t)dofC, val});
      }
      std::sort(row.begin(), row.end());
    }
  }
  pxg_mg::Level & L = m->lv[l];
  L.hPrp.assign(nf + 1, 0);
  for (int64_t r = 0; r < nf; ++r)
    L.hPrp[r + 1] = L.hPrp[r] + (int64_t)rows[r].size();
  int64_t const nnz = L.hPrp[nf];
  L.hPcol.resize(nnz);
  L.hPval.resize(nnz);
  for (int64_t r = 0; r < nf; ++r)
    for (size_t k = 0; k < rows[r].size(); ++k)
    {
      L.hPcol[L.hPrp[r] + k] = rows[r][k].first;
      L.hPval[L.hPrp[r] + k] = rows[r][k].second;
    }
  rows.clear();
  // R = P^T (counting sort by column keeps rows of R in increasing fine index)
  int64_t const ncr = Cc->ndofs;
  std::vector<int64_t> rrp(ncr + 1, 0);
  for (int64_t k = 0; k < nnz; ++k)
    rrp[L.hPcol[k] + 1]++;
  for (int64_t r = 0; r < ncr; ++r)
    rrp[r + 1] += rrp[r];
  std::vector<int32_t> rcol(nnz);
  std::vector<double> rval(nnz);
  {
    std::vector<int64_t> pos(rrp.begin(), rrp.end() - 1);
    for (int64_t r = 0; r < nf; ++r)
      for (int64_t k = L.hPrp[r]; k < L.hPrp[r + 1]; ++k)
      {
        int64_t const o = pos[L.hPcol[k]]++;
        rcol[o] = (int32_t)r;
        rval[o] = L.hPval[k];
      }
  }
  std::vector<float> p32(L.hPval.begin(), L.hPval.end()), r32(rval.begin(), rval.end());
  L.Prp.upload(L.hPrp.data(), L.hPrp.size(), s);
  L.Pcol.upload(L.hPcol.data(), nnz, s);
  L.Pval.upload(L.hPval.data(), nnz, s);
  L.Pval32.upload(p32.data(), nnz, s);
  L.Rrp.upload(rrp.data(), rrp.size(), s);
  L.Rcol.upload(rcol.data(), nnz, s);
  L.Rval.upload(rval.data(), nnz, s);
  L.Rval32.upload(r32.data(), nnz, s);
  PXG_HIP(hipStreamSynchronize(s));
  L.P = CsrView{nf, ncr, nnz, L.Prp.p, L.Pcol.p, L.Pval.p, L.Pval32.p, nullptr, nullptr};
  L.R = CsrView{ncr, nf, nnz, L.Rrp.p, L.Rcol.p, L.Rval.p, L.Rval32.p, nullptr, nullptr};
  buildRowBlocks(nf, nnz, L.Prp.p, L.Pblk, L.P.nblocks, s);
  buildRowBlocks(ncr, nnz, L.Rrp.p, L.Rblk, L.R.nblocks, s);
  L.P.blockRow = L.P.nblocks > 0 ? L.Pblk.p : nullptr;
  L.R.blockRow = L.R.nblocks > 0 ? L.Rblk.p : nullptr;
}

// operator of level l in the cycle: the assembled CSR matrix, or in mode
// PXG_FP32_MF (every level but the coarsest) the partially assembled
// matrix-free operator (FP32)
template <typename T>
struct LevelOp
{
  pxg_mg * m;
  int l;
  CsrView A;
  bool mf;
  LevelOp(pxg_mg * mg, int lev): m(mg), l(lev), A(viewOf(mg->ctx[lev])), mf(lev < mg->mfLevels)
  {
    if (mf && sizeof(T) != 4)
      throw Error(PXG_E_INVALID, "the matrix-free levels are FP32");
  }
  int64_t n() const { return A.nrows; }
  T const * dinv() const
  {
    if (mf)
      return (T const *)m->lv[l].dinvMf32.p;
    return sizeof(T) == 8 ? (T const *)A.dinv : (T const *)A.dinv32;
  }
  // element kernel into the element buffer, then the gather with its epilogue
  void mfRun(int mode, T const * x, T const * b, double omega, T * y, hipStream_t s) const
  {
    pxg_ctx * c = m->ctx[l];
    auto & L = m->lv[l];
    if (L.cubes)
    {
      sfApplyPaCubes(c, mode, L.metC.p, (float const *)x, (float const *)b, L.dinvMf32.p, omega, (float *)y,
                     L.partF.p, true, s);
      return;
    }
    sfApplyPA<float>(c, L.met32.p, (float const *)x, L.mfBuf32.p, true, s);
    sfGather<float>(c, mode, L.mfBuf32.p, (float const *)x, (float const *)b, L.dinvMf32.p, omega, (float *)y, s);
  }
  void apply(T const * x, T * y, hipStream_t s) const
  {
    if (mf)
      mfRun(0, x, nullptr, 0., y, s);
    else
      spmv<T>(A, x, y, s);
  }
  void residual(T const * b, T const * x, T * r, hipStream_t s) const
  {
    if (mf)
      mfRun(1, x, b, 0., r, s);
    else
      ::pxg::residual<T>(A, b, x, r, s);
  }
  void jacobi(T const * b, T const * x, T * y, double omega, hipStream_t s) const
  {
    if (mf)
      mfRun(2, x, b, omega, y, s);
    else
      jacobiStep<T>(A, b, x, y, omega, s);
  }
};

// smoothing sweeps, ping-ponging between x and t, result left in x
template <typename T>
void smooth(LevelOp<T> const & A, T const * b, T * x, T * t, int sweeps, double omega, bool fromZero, hipStream_t s)
{
  int64_t const n = A.n();
  if (sweeps <= 0)
  {
    if (fromZero)
      fill<T>(n, x, T(0), s);
    return;
  }
  T * cur = x;
  int steps = sweeps;
  if (fromZero)
  {
    // the first sweep from zero is x = omega D^-1 b; place it so that the
    // remaining sweeps end in x
    steps = sweeps - 1;
    cur = (steps % 2 == 0) ? x : t;
    jacobiDiag<T>(n, A.dinv(), b, cur, omega, s);
  }
  for (int k = 0; k < steps; ++k)
  {
    T * oth = (cur == x) ? t : x;
    A.jacobi(b, cur, oth, omega, s);
    cur = oth;
  }
  if (cur != x)
    PXG_HIP(hipMemcpyAsync(x, cur, n * sizeof(T), hipMemcpyDeviceToDevice, s));
}

// smallest level applied through the element cubes (two wavefronts of 64
// elements per SIMD); PXG_MF_CUBE_MIN_ELEMS sets another bound (tests run
// small levels through the cubes with it)
constexpr int64_t kMfCubeMinElems = 131072;
int64_t mfCubeMinElems()
{
  char const * e = std::getenv("PXG_MF_CUBE_MIN_ELEMS");
  return e && *e ? std::atoll(e) : kMfCubeMinElems;
}

// metrics and inverse diagonal of matrix-free level l for the current kappa
// field of its context (recomputed when kappa was set again)
void prepareMf(pxg_mg * m, int l)
{
  pxg_ctx * c = m->ctx[l];
  auto & L = m->lv[l];
  PXG_REQUIRE(c->hasKappa, PXG_E_STATE, "the matrix-free levels need their kappa field (pxg_set_coefficient)");
  PXG_REQUIRE(sfAvailable(c), PXG_E_INVALID,
              "the matrix-free levels need Q1 / Q2 hexahedra with their Gauss rule");
  sfPrepare(c);
  if (L.mfKappaStamp == c->kappaStamp && (L.cubes ? L.metC.p : L.met32.p))
    return;
  hipStream_t s = m->s;
  int64_t const E = c->nelem, n = c->fe.n;
  L.met32.resize((size_t)E * 6 * n);
  L.mfBuf32.resize((size_t)E * n);
  L.dinvMf32.resize(c->ndofs);
  sfMetric<float>(c, c->kappa.p, L.met32.p, s);
  sfElementDiag<float>(c, c->kappa.p, L.mfBuf32.p, s);
  sfGather<float>(c, 3, L.mfBuf32.p, nullptr, nullptr, nullptr, 0., L.dinvMf32.p, s);
  // (the cube kernel runs one lane per element: on small levels too few
  // wavefronts to cover its latency, where the pencil kernel's nine lanes per
  // element do better)
  L.cubes = E >= mfCubeMinElems() && sfCubesReady(c);
  if (L.cubes)
  {
    int const ld = c->fe.feType == PXG_FE_HEX_Q2 ? 5 : 3; // cube lattice side
    L.metC.resize((size_t)c->nChunk * 6 * n * 64);
    L.partF.resize(std::max<size_t>(1, (size_t)c->nCube * ld * ld * ld));
    sfMetricSlots(c, L.met32.p, L.metC.p, s);
    PXG_HIP(hipStreamSynchronize(s));
    L.met32.release();
    L.mfBuf32.release();
  }
  L.mfKappaStamp = c->kappaStamp;
}

// (re)builds the dense inverse of the coarsest free-dof block when the coarse
// matrix has changed since the last factorisation
void factorCoarse(pxg_mg * m)
{
  pxg_ctx * c = m->ctx.back();
  hipStream_t s = m->s;
  ensureVal64(c);
  joinLevels(m);
  if (m->denseStamp == c->stamp && c->stamp != 0)
    return;
  int64_t const n = c->ndofs;
  std::vector<uint8_t> bc(n, 0);
  if (c->hasBC)
  {
    c->bcFlag.download(bc.data(), n, s);
    PXG_HIP(hipStreamSynchronize(s));
  }
  std::vector<int32_t> idx(n, -1), dofs;
  for (int64_t r = 0; r < n; ++r)
    if (!bc[r])
    {
      idx[r] = (int32_t)dofs.size();
      dofs.push_back((int32_t)r);
    }
  int64_t const nf = (int64_t)dofs.size();
  m->dense = false;
  m->denseStamp = c->stamp;
  if (nf > kDenseMax)
    return; // iterative coarse solve
  m->nfree = nf;
  m->freeIdx.upload(idx.data(), n, s);
  if (nf == 0)
  {
    m->dense = true;
    m->denseStamp = c->stamp;
    return;
  }
  m->freeDofs.upload(dofs.data(), nf, s);
  m->inv.resize(nf * nf);
  m->invTmp.resize(nf * nf);
  m->flag.resize(1);
  CsrView const A = viewOf(c);
  int const bs = 256;
  unsigned const g2 = (unsigned)((nf * nf + bs - 1) / bs);
  fill<double>(nf * nf, m->inv.p, 0., s);
  k_dense_scatter<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(n, m->freeIdx.p, A.rowptr, A.col, A.val, nf, m->inv.p);
  double * a = m->inv.p, * b = m->invTmp.p;
  m->gjP.resize(kGjBlock * kGjBlock);
  m->gjR.resize(kGjBlock * nf);
  for (int64_t k0 = 0; k0 < nf; k0 += kGjBlock)
  {
    int const nb = (int)std::min<int64_t>(kGjBlock, nf - k0);
    k_gj_pivot<<<1, 1024, 0, s>>>(nf, k0, nb, a, m->gjP.p);
    k_gj_rows<<<(unsigned)((nb * nf + bs - 1) / bs), bs, 0, s>>>(nf, k0, nb, a, m->gjP.p, m->gjR.p);
    k_gj_update<<<g2, bs, 0, s>>>(nf, k0, nb, a, m->gjP.p, m->gjR.p, b);
    std::swap(a, b);
  }
  if (a != m->inv.p)
    PXG_HIP(hipMemcpyAsync(m->inv.p, a, nf * nf * sizeof(double), hipMemcpyDeviceToDevice, s));
  k_inverse_check<<<1, bs, 0, s>>>(nf, m->inv.p, m->flag.p);
  PXG_HIP(hipGetLastError());
  int bad = 0;
  PXG_HIP(hipMemcpyAsync(&bad, m->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  PXG_HIP(hipStreamSynchronize(s));
  // a singular or indefinite coarse block (no Dirichlet dofs, pure Neumann)
  // keeps the iterative coarse solve
  m->dense = !bad;
  m->denseStamp = c->stamp;
}

// coarsest level: dense direct solve, or Jacobi-preconditioned CG in FP64
// from zero to coarseRtol when the free block is too large to invert
void coarseSolve(pxg_mg * m, double const * b, double * x)
{
  pxg_ctx * c = m->ctx.back();
  CsrView const A = viewOf(c);
  int64_t const n = A.nrows;
  hipStream_t s = m->s;
  if (m->dense)
  {
    k_dense_apply<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, m->nfree, m->freeIdx.p, m->freeDofs.p, m->inv.p,
                                                        A.dinv, b, x);
    PXG_HIP(hipGetLastError());
    return;
  }
  double * r = m->cr.p, * z = m->cz.p, * p = m->cp.p, * q = m->cq.p;
  double * sc = m->sc.p;
  fill<double>(n, x, 0., s);
  PXG_HIP(hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  dot<double>(n, b, b, sc + 8, m->partial.p, s);
  double const bn = std::sqrt(readScalar(m, 8));
  if (bn == 0.)
    return;
  jacobiDiag<double>(n, A.dinv, r, z, 1.0, s);
  PXG_HIP(hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  dot<double>(n, r, z, sc + 9, m->partial.p, s);
  int const maxit = 4000;
  for (int it = 1; it <= maxit; ++it)
  {
    spmv<double>(A, p, q, s);
    dot<double>(n, p, q, sc + 10, m->partial.p, s);
    cgUpdate<double>(n, x, p, r, q, sc, 9, 10, s);
    if (it % 8 == 0)
    {
      dot<double>(n, r, r, sc + 11, m->partial.p, s);
      if (std::sqrt(readScalar(m, 11)) <= m->coarseRtol * bn)
        return;
    }
    jacobiDiag<double>(n, A.dinv, r, z, 1.0, s);
    dot<double>(n, r, z, sc + 12, m->partial.p, s);
    cgDirection<double>(n, p, z, sc, 12, 9, s);
    copyScalar(m, 9, 12);
  }
}

template <typename T>
void vcycle(pxg_mg * m, int l, T const * b, T * x)
{
  hipStream_t s = m->s;
  if (l == m->nlev - 1)
  {
    pxg_mg::Level & L = m->lv[l];
    if constexpr (sizeof(T) == 8)
      coarseSolve(m, (double const *)b, (double *)x);
    else
    {
      convert<double, float>(L.n, (float const *)b, L.b.p, s);
      coarseSolve(m, L.b.p, L.x.p);
      convert<float, double>(L.n, L.x.p, (float *)x, s);
    }
    return;
  }
  pxg_mg::Level & L = m->lv[l];
  pxg_mg::Level & Lc = m->lv[l + 1];
  LevelOp<T> const A(m, l);
  T * t = Buf<T>::t(L);
  T * r = Buf<T>::r(L);
  smooth<T>(A, b, x, t, m->nu1, m->omega, true, s);
  A.residual(b, x, r, s);
  T * bc = Buf<T>::b(Lc);
  T * xc = Buf<T>::x(Lc);
  spmv<T>(L.R, r, bc, s);
  vcycle<T>(m, l + 1, bc, xc);
  spmvAdd<T>(L.P, xc, x, s);
  smooth<T>(A, b, x, t, m->nu2, m->omega, false, s);
}

// preconditioned CG in precision T on level 0; b, x in T; returns iterations
template <typename T>
int pcg(pxg_mg * m, T const * b, T * x, double rtol, int maxit, double & relres, bool zeroStart)
{
  hipStream_t s = m->s;
  LevelOp<T> const A(m, 0);
  int64_t const n = A.n();
  T * r = sizeof(T) == 8 ? (T *)m->r.p : (T *)m->r32.p;
  T * z = sizeof(T) == 8 ? (T *)m->z.p : (T *)m->z32.p;
  T * p = sizeof(T) == 8 ? (T *)m->p.p : (T *)m->p32.p;
  T * q = sizeof(T) == 8 ? (T *)m->q.p : (T *)m->q32.p;
  double * sc = m->sc.p;
  if (zeroStart)
  {
    fill<T>(n, x, T(0), s);
    PXG_HIP(hipMemcpyAsync(r, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
  }
  else
    A.residual(b, x, r, s);
  dot<T>(n, b, b, sc + 0, m->partial.p, s);
  dot<T>(n, r, r, sc + 1, m->partial.p, s);
  double const bn = std::sqrt(readScalar(m, 0));
  double rn = std::sqrt(readScalar(m, 1));
  relres = bn > 0. ? rn / bn : 0.;
  if (relres <= rtol || bn == 0.)
    return 0;
  vcycle<T>(m, 0, r, z);
  PXG_HIP(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, s));
  dot<T>(n, r, z, sc + 2, m->partial.p, s);
  // the norm of an iteration is read after its next cycle is queued (the
  // cycle changes z, p and scalars only, so x is still that iterate when the
  // test passes); it is read before the cycle when the contraction so far
  // says the test may pass, which saves the cycle
  double r1 = rn, r2 = -1.;
  int it = 0;
  while (it < maxit)
  {
    A.apply(p, q, s);
    dot<T>(n, p, q, sc + 3, m->partial.p, s);
    cgUpdate<T>(n, x, p, r, q, sc, 2, 3, s);
    ++it;
    dot<T>(n, r, r, sc + 1, m->partial.p, s);
    bool const lag = nextNormGuess(r1, r2) > kLagSafety * rtol * bn;
    if (!lag)
    {
      rn = std::sqrt(readScalar(m, 1));
      PXG_REQUIRE(std::isfinite(rn), PXG_E_NUMERIC, "CG breakdown (non-finite residual)");
      relres = rn / bn;
      if (relres <= rtol)
        break;
    }
    else
      postScalar(m, 1);
    vcycle<T>(m, 0, r, z);
    dot<T>(n, r, z, sc + 4, m->partial.p, s);
    cgDirection<T>(n, p, z, sc, 4, 2, s);
    copyScalar(m, 2, 4);
    if (lag)
    {
      rn = std::sqrt(takeScalar(m));
      PXG_REQUIRE(std::isfinite(rn), PXG_E_NUMERIC, "CG breakdown (non-finite residual)");
      relres = rn / bn;
      if (relres <= rtol)
        break;
    }
    r2 = r1;
    r1 = rn;
  }
  return it;
}

// FP64 residual r = b - A x on level 0: assembled FP64 matrix, or the
// matrix-free kappa-stiffness operator when level 0 holds FP32 values
void residual64(pxg_mg * m, double const * b, double const * x, double * r)
{
  pxg_ctx * c = m->ctx[0];
  if (m->mfLevels > 0)
  {
    // the FP64 sum-factorised operator (geometry recomputed), same Dirichlet form
    m->mfBuf64.resize((size_t)c->nelem * c->fe.n);
    sfElementApply<double>(c, c->kappa.p, 1, x, m->mfBuf64.p, true, m->s);
    sfGather<double>(c, 1, m->mfBuf64.p, x, b, nullptr, 0., r, m->s);
  }
  else if (c->valPrecision == PXG_FP64 && c->closed)
    residual<double>(viewOf(c), b, x, r, m->s);
  else
  {
    PXG_REQUIRE(c->hasKappa, PXG_E_STATE, "FP32 level 0 needs the kappa field for the FP64 residual");
    PXG_HIP(hipStreamSynchronize(m->s));
    mfApply(c, c->kappa.p, 1, x, r, true);
    axpby(c->ndofs, 1.0, b, -1.0, r, m->s);
  }
}

// FP32 + FP64 refinement without restarts: preconditioned CG on FP32
// vectors and matrices (FP64 inner products) whose iterate is kept in FP64,
// with the recursive FP32 residual replaced by the FP64 residual b - A x
// whenever it has fallen by kReplace since the last replacement or reaches
// the tolerance (reliable updates, van der Vorst and Ye, SISC 22 (2000)).
// Every refinement is thereby one more CG step instead of a restarted inner
// solve; convergence is decided on the FP64 residual only.  The FP32
// recursive residual drifts from the FP64 one by the rounding of the FP32
// matrix (and, with the matrix-free FP64 residual, by the difference of the
// two operators), about 1e-7 of the residual at the last replacement: a
// hundredfold drop keeps that gap near 1e-5 of the current residual.  (Drops
// of 1e-4 between replacements let the gap reach the residual and cost CG
// 4-8 times the iterations.)
int solveIR(pxg_mg * m, double const * b, double * x, double rtol, int maxit, double & relres)
{
  constexpr double kReplace = 1e-2;
  hipStream_t s = m->s;
  int64_t const n = m->lv[0].n;
  LevelOp<float> const A(m, 0);
  double * r64 = m->r.p;
  float * r = m->r32.p, * z = m->z32.p, * p = m->p32.p, * q = m->q32.p;
  double * sc = m->sc.p;
  dot<double>(n, b, b, sc + 5, m->partial.p, s);
  double const bn = std::sqrt(readScalar(m, 5));
  if (bn == 0.)
  {
    fill<double>(n, x, 0., s);
    relres = 0.;
    return 0;
  }
  // true residual of x: its norm (the convergence measure) and its FP32
  // rounding (the CG residual from here on)
  auto trueNorm = [&]() -> double
  {
    residual64(m, b, x, r64);
    dot<double>(n, r64, r64, sc + 6, m->partial.p, s);
    double const rn = std::sqrt(readScalar(m, 6));
    PXG_REQUIRE(std::isfinite(rn), PXG_E_NUMERIC, "non-finite residual in refinement");
    return rn;
  };
  auto trueResidual = [&]() -> double
  {
    double const rn = trueNorm();
    convert<float, double>(n, r64, r, s);
    return rn;
  };
  double last = trueResidual();
  relres = last / bn;
  if (relres <= rtol)
    return 0;
  vcycle<float>(m, 0, r, z);
  PXG_HIP(hipMemcpyAsync(p, z, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  dot<float>(n, r, z, sc + 2, m->partial.p, s);
  // The recursive norm of an iteration is read after its next cycle is
  // queued (see pcg), unless the contraction so far says that it may call
  // for the true residual; a replacement found late waits for the next
  // iteration (it must come before a cycle), a convergence test found late
  // is made at once on the true norm (x is still that iterate)
  auto wanted = [&](double rr) { return rr <= rtol * bn || rr <= kReplace * last; };
  double r1 = last, r2 = -1.;
  bool replaceNext = false;
  int it = 0;
  while (it < maxit)
  {
    A.apply(p, q, s);
    dot<float>(n, p, q, sc + 3, m->partial.p, s);
    cgUpdateMixed(n, x, p, r, q, sc, 2, 3, s);
    ++it;
    dot<float>(n, r, r, sc + 1, m->partial.p, s);
    bool const lag = !replaceNext && !wanted(nextNormGuess(r1, r2) / kLagSafety);
    double rr = 0.;
    if (!lag)
    {
      rr = std::sqrt(readScalar(m, 1));
      PXG_REQUIRE(std::isfinite(rr), PXG_E_NUMERIC, "CG breakdown (non-finite residual)");
      if (replaceNext || wanted(rr))
      {
        replaceNext = false;
        last = trueResidual();
        relres = last / bn;
        if (relres <= rtol)
          break;
      }
    }
    else
      postScalar(m, 1);
    vcycle<float>(m, 0, r, z);
    dot<float>(n, r, z, sc + 4, m->partial.p, s);
    cgDirection<float>(n, p, z, sc, 4, 2, s);
    copyScalar(m, 2, 4);
    if (lag)
    {
      rr = std::sqrt(takeScalar(m));
      PXG_REQUIRE(std::isfinite(rr), PXG_E_NUMERIC, "CG breakdown (non-finite residual)");
      if (rr <= rtol * bn)
      {
        double const tn = trueNorm();
        if (tn <= rtol * bn)
        {
          last = tn;
          relres = tn / bn;
          break;
        }
        replaceNext = true;
      }
      else if (rr <= kReplace * last)
        replaceNext = true;
    }
    r2 = r1;
    r1 = rr;
  }
  return it;
}
} // namespace

extern "C" {

int pxg_mg_create(pxg_mg ** out, int nlev, pxg_ctx ** levels, const int * dims, double omega, int nu1, int nu2,
                  double coarseRtol)
{
  if (!out || nlev < 1 || !levels || !dims)
    return PXG_E_INVALID;
  *out = nullptr;
  auto * m = new (std::nothrow) pxg_mg;
  if (!m)
    return PXG_E_OOM;
  try
  {
    m->nlev = nlev;
    m->omega = omega;
    m->nu1 = nu1;
    m->nu2 = nu2;
    m->coarseRtol = coarseRtol;
    for (int l = 0; l < nlev; ++l)
    {
      PXG_REQUIRE(levels[l] && levels[l]->closed, PXG_E_STATE, "levels must be closed contexts");
      // every level matrix is applied several times per cycle: the levels
      // keep block-packed copies (pxg_csr_pack) for the smoother and residual
      levels[l]->packWanted = true;
      m->ctx.push_back(levels[l]);
      m->dims.push_back({dims[3 * l], dims[3 * l + 1], dims[3 * l + 2]});
    }
    PXG_HIP(hipSetDevice(levels[0]->device));
    m->s = levels[0]->stream;
    m->lv.resize(nlev);
    for (int l = 0; l < nlev; ++l)
    {
      auto & L = m->lv[l];
      L.n = levels[l]->ndofs;
      L.b.resize(L.n);
      L.x.resize(L.n);
      L.t.resize(L.n);
      L.r.resize(L.n);
      L.b32.resize(L.n);
      L.x32.resize(L.n);
      L.t32.resize(L.n);
      L.r32.resize(L.n);
    }
    for (int l = 0; l + 1 < nlev; ++l)
      buildProlongation(m, l);
    factorCoarse(m);
    int64_t const n0 = m->lv[0].n, nc = m->lv.back().n;
    m->sc.resize(16);
    m->partial.resize(kDotBlocks);
    m->r.resize(n0);
    m->z.resize(n0);
    m->p.resize(n0);
    m->q.resize(n0);
    m->r32.resize(n0);
    m->z32.resize(n0);
    m->p32.resize(n0);
    m->q32.resize(n0);
    m->cr.resize(nc);
    m->cz.resize(nc);
    m->cp.resize(nc);
    m->cq.resize(nc);
  }
  catch (Error const & e)
  {
    levels[0]->err = e.what();
    delete m;
    return e.code;
  }
  catch (std::exception const & e)
  {
    levels[0]->err = e.what();
    delete m;
    return PXG_E_INVALID;
  }
  *out = m;
  return PXG_OK;
}

int pxg_mg_destroy(pxg_mg * m)
{
  delete m;
  return PXG_OK;
}

const char * pxg_mg_last_error(const pxg_mg * m) { return m ? m->err.c_str() : "null multigrid"; }

int pxg_mg_update(pxg_mg * m)
{
  if (!m)
    return PXG_E_INVALID;
  try
  {
    PXG_HIP(hipSetDevice(m->ctx[0]->device));
    m->s = m->ctx[0]->stream;
    // the coarsest matrix is factored here; the others are checked by the
    // solve (only the coarsest need be assembled for PXG_FP32_MF)
    PXG_REQUIRE(m->ctx.back()->closed, PXG_E_STATE, "the coarsest level must be closed");
    factorCoarse(m);
    m->err.clear();
    return PXG_OK;
  }
  catch (Error const & e)
  {
    m->err = e.what();
    return e.code;
  }
}

int pxg_mg_solve(pxg_mg * m, int mode, const double * b, double * x, double rtol, int maxit, int * iters,
                 double * relres)
{
  if (!m || !b || !x)
    return PXG_E_INVALID;
  try
  {
    PXG_HIP(hipSetDevice(m->ctx[0]->device));
    m->s = m->ctx[0]->stream; // follow the finest context (its stream may have been replaced)
    PXG_REQUIRE(mode == PXG_FP64 || mode == PXG_FP32 || mode == PXG_FP32_MF, PXG_E_INVALID, "solve mode");
    bool const mf = mode == PXG_FP32_MF;
    PXG_REQUIRE(!mf || m->nlev >= 2, PXG_E_INVALID, "the matrix-free levels need an assembled coarsest level");
    int const nmf = mf ? m->nlev - 1 : 0;
    for (int l = nmf; l < m->nlev; ++l)
      PXG_REQUIRE(m->ctx[l]->closed, PXG_E_STATE, "levels must be closed");
    factorCoarse(m);
    double rr = 0.;
    int it = 0;
    if (mode != PXG_FP64)
    {
      for (int l = nmf; l < m->nlev; ++l)
        ensureOperator(m->ctx[l], PXG_FP32);
      joinLevels(m);
      struct Mf // the mode holds for this solve only
      {
        pxg_mg * m;
        ~Mf() { m->mfLevels = 0; }
      } guard{m};
      m->mfLevels = nmf;
      for (int l = 0; l < nmf; ++l)
        prepareMf(m, l);
      it = solveIR(m, b, x, rtol, maxit, rr);
    }
    else
    {
      for (auto * c: m->ctx)
        ensureOperator(c, PXG_FP64);
      joinLevels(m);
      it = pcg<double>(m, b, x, rtol, maxit, rr, false);
    }
    PXG_HIP(hipStreamSynchronize(m->s));
    if (iters)
      *iters = it;
    if (relres)
      *relres = rr;
    m->err.clear();
    if (rr > rtol)
    {
      m->err = "iteration limit reached before the tolerance";
      return PXG_E_NOCONV;
    }
    return PXG_OK;
  }
  catch (Error const & e)
  {
    m->err = e.what();
    return e.code;
  }
}

int pxg_mg_vcycle(pxg_mg * m, const double * b, double * x)
{
  if (!m || !b || !x)
    return PXG_E_INVALID;
  try
  {
    PXG_HIP(hipSetDevice(m->ctx[0]->device));
    m->s = m->ctx[0]->stream; // follow the finest context (its stream may have been replaced)
    for (auto * c: m->ctx)
    {
      PXG_REQUIRE(c->closed, PXG_E_STATE, "levels must be closed");
      ensureOperator(c, PXG_FP64);
    }
    factorCoarse(m);
    joinLevels(m);
    vcycle<double>(m, 0, b, x);
    PXG_HIP(hipStreamSynchronize(m->s));
    return PXG_OK;
  }
  catch (Error const & e)
  {
    m->err = e.what();
    return e.code;
  }
}

int pxg_mg_prolongation(pxg_mg * m, int l, int64_t * rowptr, int32_t * col, double * val, int64_t * nnz)
{
  if (!m || l < 0 || l + 1 >= m->nlev)
    return PXG_E_INVALID;
  auto const & L = m->lv[l];
  if (nnz)
    *nnz = (int64_t)L.hPcol.size();
  if (rowptr)
    std::memcpy(rowptr, L.hPrp.data(), L.hPrp.size() * sizeof(int64_t));
  if (col)
    std::memcpy(col, L.hPcol.data(), L.hPcol.size() * sizeof(int32_t));
  if (val)
    std::memcpy(val, L.hPval.data(), L.hPval.size() * sizeof(double));
  return PXG_OK;
}

// Restrictor of the reference (include/multigrid.hpp:108-175): the entries of
// P^T, every row scaled to sum one (sum in increasing fine index, then each
// value divided by it).  The cycle itself uses P^T; this is the reference's
// operator as a host CSR, like pxg_mg_prolongation (rows and columns of
// constrained dofs are the ones P leaves out).
int pxg_mg_restriction(pxg_mg * m, int l, int64_t * rowptr, int32_t * col, double * val, int64_t * nnz)
{
  if (!m || l < 0 || l + 1 >= m->nlev)
    return PXG_E_INVALID;
  auto const & L = m->lv[l];
  int64_t const nf = (int64_t)L.hPrp.size() - 1, np = (int64_t)L.hPcol.size(), nc = m->lv[l + 1].n;
  if (nnz)
    *nnz = np;
  if (!rowptr && !col && !val)
    return PXG_OK;
  std::vector<int64_t> rp(nc + 1, 0);
  for (int64_t k = 0; k < np; ++k)
    rp[L.hPcol[k] + 1]++;
  for (int64_t r = 0; r < nc; ++r)
    rp[r + 1] += rp[r];
  std::vector<int32_t> rc(np);
  std::vector<double> rv(np);
  std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
  for (int64_t r = 0; r < nf; ++r)
    for (int64_t k = L.hPrp[r]; k < L.hPrp[r + 1]; ++k)
    {
      int64_t const o = pos[L.hPcol[k]]++;
      rc[o] = (int32_t)r;
      rv[o] = L.hPval[k];
    }
  for (int64_t r = 0; r < nc; ++r)
  {
    double sum = 0.;
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
      sum += rv[k];
    for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
      rv[k] /= sum;
  }
  if (rowptr)
    std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int64_t));
  if (col)
    std::memcpy(col, rc.data(), rc.size() * sizeof(int32_t));
  if (val)
    std::memcpy(val, rv.data(), rv.size() * sizeof(double));
  return PXG_OK;
}

} // extern "C"
