// Matrix-free operator y = K(kappa) x for a batch of vectors and its
// kappa-gradient, with the element geometry recomputed on every call (nothing
// stored between forward and backward).
//
// For element e and quadrature point q (reference gradients g_j(q), vertex
// map J, A = J^T J):
//   v_bq = sum_j g_j(q) x_{b,j}            (reference gradient of x_b)
//   y_{b,i} += sum_q g_i(q) . (kappa_q JxW_q A^{-1} v_bq)
//   d/dkappa_n (ybar . K x) = sum_e sum_q phi_n(q) JxW_q sum_b v_bq(x)^T A^{-1} v_bq(ybar)
// which is the AssemblyStiffnessFE bilinear form (include/assembly_lhs.hpp:93-109)
// applied without forming Ke.  Element results go to an [E][n][nb] buffer and
// are gathered per dof in element order (no float atomics).
#include "kernels.hpp"

#include <algorithm>

namespace pxg
{

namespace
{
constexpr int kMfBlock = 128;
constexpr int kMaxNb = 16;

template <int DIM, int NPE>
__device__ inline double geomAt(int64_t e, int q, double const * __restrict__ xyz,
                                uint32_t const * __restrict__ conn, double const * __restrict__ geoGrad,
                                double (&I)[3][3])
{
  double J[3][DIM];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < DIM; ++c)
      J[r][c] = 0.;
#pragma unroll
  for (int v = 0; v < NPE; ++v)
  {
    uint32_t const pid = conn[e * NPE + v];
    double const x0 = xyz[3ull * pid], x1 = xyz[3ull * pid + 1], x2 = xyz[3ull * pid + 2];
#pragma unroll
    for (int c = 0; c < DIM; ++c)
    {
      double const g = geoGrad[(q * NPE + v) * DIM + c];
      J[0][c] += x0 * g;
      J[1][c] += x1 * g;
      J[2][c] += x2 * g;
    }
  }
  double A[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
#pragma unroll
  for (int a = 0; a < DIM; ++a)
#pragma unroll
    for (int b = 0; b < DIM; ++b)
      A[a][b] = J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b];
  double det;
  if constexpr (DIM == 1)
  {
    det = A[0][0];
    I[0][0] = 1. / det;
  }
  else if constexpr (DIM == 2)
  {
    det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    I[0][0] = A[1][1] / det;
    I[0][1] = I[1][0] = -A[0][1] / det;
    I[1][1] = A[0][0] / det;
  }
  else
  {
    double const c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double const c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double const c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    I[0][0] = c00 / det;
    I[0][1] = I[1][0] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
    I[0][2] = I[2][0] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
    I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
    I[1][2] = I[2][1] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
    I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
  }
  return det;
}

// phase 1: per (element, q) the flux w_bq = kappa_q JxW_q A^{-1} v_bq into LDS;
// phase 2: per (element, i) y_{b,i} = sum_q g_i(q) . w_bq into the element buffer
template <int DIM, int NPE>
__global__ void __launch_bounds__(kMfBlock)
    k_mf_apply(int64_t E, int n, int nq, int nb, int epb, double const * __restrict__ xyz,
               uint32_t const * __restrict__ conn, double const * __restrict__ geoGrad,
               double const * __restrict__ w, double const * __restrict__ phi, double const * __restrict__ grad,
               uint32_t const * __restrict__ emap, double const * __restrict__ kappa,
               uint8_t const * __restrict__ cons, double const * __restrict__ x, double * __restrict__ ybuf,
               int * __restrict__ status)
{
  extern __shared__ double sm[]; // epb * nq * nb * DIM
  int64_t const e0 = (int64_t)blockIdx.x * epb;
  int const tid = threadIdx.x;
  if (tid < epb * nq)
  {
    int const le = tid / nq, q = tid - le * nq;
    int64_t const e = e0 + le;
    if (e < E)
    {
      double I[3][3];
      double const det = geomAt<DIM, NPE>(e, q, xyz, conn, geoGrad, I);
      if (!(det > 0.))
        atomicOr(status, 1);
      double kq = 0.;
      for (int i = 0; i < n; ++i)
        kq += phi[q * n + i] * kappa[emap[e * n + i]];
      double const s = kq * (sqrt(det) * w[q]);
      for (int b = 0; b < nb; ++b)
      {
        double v[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          v[d] = 0.;
        for (int j = 0; j < n; ++j)
        {
          uint32_t const id = emap[e * n + j];
          double const xv = (cons && cons[id]) ? 0. : x[(size_t)id * nb + b];
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            v[d] += grad[(q * n + j) * DIM + d] * xv;
        }
        double * o = sm + ((size_t)(le * nq + q) * nb + b) * DIM;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
        {
          double t = 0.;
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            t += I[a][d] * v[d];
          o[a] = s * t;
        }
      }
    }
  }
  __syncthreads();
  if (tid < epb * n)
  {
    int const le = tid / n, i = tid - le * n;
    int64_t const e = e0 + le;
    if (e < E)
    {
      uint32_t const id = emap[e * n + i];
      bool const ci = cons && cons[id];
      for (int b = 0; b < nb; ++b)
      {
        double y = 0.;
        if (ci)
          y = x[(size_t)id * nb + b]; // unit diagonal per element on constrained rows
        else
          for (int q = 0; q < nq; ++q)
          {
            double const * f = sm + ((size_t)(le * nq + q) * nb + b) * DIM;
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              y += grad[(q * n + i) * DIM + d] * f[d];
          }
        ybuf[((size_t)e * n + i) * nb + b] = y;
      }
    }
  }
}

template <int DIM, int NPE>
__global__ void __launch_bounds__(kMfBlock)
    k_mf_gradk(int64_t E, int n, int nq, int nb, int epb, double const * __restrict__ xyz,
               uint32_t const * __restrict__ conn, double const * __restrict__ geoGrad,
               double const * __restrict__ w, double const * __restrict__ phi, double const * __restrict__ grad,
               uint32_t const * __restrict__ emap, uint8_t const * __restrict__ cons, double const * __restrict__ x,
               double const * __restrict__ ybar, double * __restrict__ gbuf)
{
  extern __shared__ double sm[]; // epb * nq
  int64_t const e0 = (int64_t)blockIdx.x * epb;
  int const tid = threadIdx.x;
  if (tid < epb * nq)
  {
    int const le = tid / nq, q = tid - le * nq;
    int64_t const e = e0 + le;
    double c = 0.;
    if (e < E)
    {
      double I[3][3];
      double const det = geomAt<DIM, NPE>(e, q, xyz, conn, geoGrad, I);
      for (int b = 0; b < nb; ++b)
      {
        double vx[DIM], vy[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          vx[d] = vy[d] = 0.;
        for (int j = 0; j < n; ++j)
        {
          uint32_t const id = emap[e * n + j];
          if (cons && cons[id])
            continue;
          double const xv = x[(size_t)id * nb + b], yv = ybar[(size_t)id * nb + b];
#pragma unroll
          for (int d = 0; d < DIM; ++d)
          {
            double const g = grad[(q * n + j) * DIM + d];
            vx[d] += g * xv;
            vy[d] += g * yv;
          }
        }
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            c += vx[a] * I[a][d] * vy[d];
      }
      c *= sqrt(det) * w[q];
    }
    sm[le * nq + q] = c;
  }
  __syncthreads();
  if (tid < epb * n)
  {
    int const le = tid / n, i = tid - le * n;
    int64_t const e = e0 + le;
    if (e < E)
    {
      double g = 0.;
      for (int q = 0; q < nq; ++q)
        g += phi[q * n + i] * sm[le * nq + q];
      gbuf[(size_t)e * n + i] = g;
    }
  }
}

// y[d][b] = sum over (element, local) entries of dof d, in element order
__global__ void k_gather(int64_t ndofs, int nb, uint32_t const * __restrict__ seg,
                         uint32_t const * __restrict__ ids, double const * __restrict__ buf, double * __restrict__ y)
{
  int64_t const t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndofs * nb)
    return;
  int64_t const d = t / nb;
  int const b = (int)(t - d * nb);
  double acc = 0.;
  for (uint32_t k = seg[d]; k < seg[d + 1]; ++k)
    acc += buf[(size_t)ids[k] * nb + b];
  y[t] = acc;
}

template <typename F>
void dispatchGeoMf(int geo, F && f)
{
  switch (geo)
  {
  case kLine: f(std::integral_constant<int, 1>{}, std::integral_constant<int, 2>{}); break;
  case kQuad: f(std::integral_constant<int, 2>{}, std::integral_constant<int, 4>{}); break;
  case kHex: f(std::integral_constant<int, 3>{}, std::integral_constant<int, 8>{}); break;
  case kTet: f(std::integral_constant<int, 3>{}, std::integral_constant<int, 4>{}); break;
  default: throw Error(PXG_E_INVALID, "unknown geometry");
  }
}

void prepare(pxg_ctx * c, int nb)
{
  PXG_REQUIRE(c->fe.feType && c->nelem && c->ndofs, PXG_E_STATE, "matrix-free needs set_fe, set_mesh, set_dofmap");
  PXG_REQUIRE(nb >= 1 && nb <= kMaxNb, PXG_E_INVALID, "batch size must be 1..16");
  PXG_REQUIRE(c->fe.nq <= kMfBlock && c->fe.n <= kMfBlock, PXG_E_INVALID, "element too large");
  if (!c->dofSeg.p)
    buildDofAdjacency(c);
}
} // namespace

void mfApply(pxg_ctx * c, double const * kappa, int nb, double const * x, double * y, bool useBC)
{
  prepare(c, nb);
  FeTables & T = c->fe;
  int const epb = std::max(1, kMfBlock / std::max(T.nq, T.n));
  int64_t const E = c->nelem;
  unsigned const grid = (unsigned)((E + epb - 1) / epb);
  size_t const sh = (size_t)epb * T.nq * nb * T.dim * sizeof(double);
  DevBuf<double> ybuf;
  ybuf.resize((size_t)E * T.n * nb);
  hipStream_t s = c->stream;
  PXG_HIP(hipMemsetAsync(c->status.p, 0, sizeof(int), s));
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  ProfScope ps(c, "mf_apply");
  dispatchGeoMf(T.geo,
                [&](auto dimC, auto npeC)
                {
                  constexpr int D = decltype(dimC)::value, NPE = decltype(npeC)::value;
                  k_mf_apply<D, NPE><<<grid, kMfBlock, sh, s>>>(E, T.n, T.nq, nb, epb, c->xyz.p, c->conn.p,
                                                               T.d_geoGrad.p, T.d_w.p, T.d_phi.p, T.d_grad.p,
                                                               c->elemMap.p, kappa, cons, x, ybuf.p, c->status.p);
                });
  PXG_HIP(hipGetLastError());
  int64_t const m = c->ndofs * nb;
  k_gather<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(c->ndofs, nb, c->dofSeg.p, c->dofContrib.p, ybuf.p, y);
  PXG_HIP(hipGetLastError());
  int h = 0;
  PXG_HIP(hipMemcpyAsync(&h, c->status.p, sizeof(int), hipMemcpyDeviceToHost, s));
  PXG_HIP(hipStreamSynchronize(s));
  PXG_REQUIRE(h == 0, PXG_E_NUMERIC, "non-positive Jacobian determinant in an element");
}

void mfGradKappa(pxg_ctx * c, int nb, double const * x, double const * ybar, double * gk, bool useBC)
{
  prepare(c, nb);
  FeTables & T = c->fe;
  int const epb = std::max(1, kMfBlock / std::max(T.nq, T.n));
  int64_t const E = c->nelem;
  unsigned const grid = (unsigned)((E + epb - 1) / epb);
  size_t const sh = (size_t)epb * T.nq * sizeof(double);
  DevBuf<double> gbuf;
  gbuf.resize((size_t)E * T.n);
  hipStream_t s = c->stream;
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  ProfScope ps(c, "mf_gradk");
  dispatchGeoMf(T.geo,
                [&](auto dimC, auto npeC)
                {
                  constexpr int D = decltype(dimC)::value, NPE = decltype(npeC)::value;
                  k_mf_gradk<D, NPE><<<grid, kMfBlock, sh, s>>>(E, T.n, T.nq, nb, epb, c->xyz.p, c->conn.p,
                                                               T.d_geoGrad.p, T.d_w.p, T.d_phi.p, T.d_grad.p,
                                                               c->elemMap.p, cons, x, ybar, gbuf.p);
                });
  PXG_HIP(hipGetLastError());
  k_gather<<<(unsigned)((c->ndofs + 255) / 256), 256, 0, s>>>(c->ndofs, 1, c->dofSeg.p, c->dofContrib.p, gbuf.p,
                                                              gk);
  PXG_HIP(hipGetLastError());
  PXG_HIP(hipStreamSynchronize(s));
}

} // namespace pxg

using namespace pxg;

extern "C" {

int pxg_mf_apply(pxg_ctx * c, const double * kappa, int nb, const double * x, double * y, int useBC)
{
  if (!c)
    return PXG_E_INVALID;
  try
  {
    PXG_REQUIRE(kappa && x && y && x != y, PXG_E_INVALID, "mf_apply pointers");
    PXG_HIP(hipSetDevice(c->device));
    mfApply(c, kappa, nb, x, y, useBC != 0);
    return PXG_OK;
  }
  catch (Error const & e)
  {
    c->err = e.what();
    return e.code;
  }
}

int pxg_mf_grad_kappa(pxg_ctx * c, int nb, const double * x, const double * ybar, double * gk, int useBC)
{
  if (!c)
    return PXG_E_INVALID;
  try
  {
    PXG_REQUIRE(x && ybar && gk, PXG_E_INVALID, "mf_grad_kappa pointers");
    PXG_HIP(hipSetDevice(c->device));
    mfGradKappa(c, nb, x, ybar, gk, useBC != 0);
    return PXG_OK;
  }
  catch (Error const & e)
  {
    c->err = e.what();
    return e.code;
  }
}

} // extern "C"
