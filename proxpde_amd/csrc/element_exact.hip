// FP64 element matrices of the linear elements (P1 line, Q1 quad, Q1 hex,
// P1 tet) in the evaluation order of the reference's CurFE::reinit and
// AssemblyStiffness / AssemblyMass, geometry fused in, no multiply-add
// contraction.
//
// On uniform meshes with a constant coefficient many entries of the global
// matrix are zero in exact arithmetic and come out of the element sums as
// rounding residues near 1e-16 -- exactly where Builder::closeMatrix prunes
// (prune(1, 1e-16), reference include/builder.hpp:366-371).  Which residues
// survive depends on every rounding on the way, so the kept pattern is only
// reproducible by doing the same operations in the same order:
//   CurFE::reinit (include/curfe.hpp:50-103), per quadrature point q:
//     J[r][c]  = sum_g x_g[r] dpsi_g(q)[c]          (g ascending, from 0)
//     A        = J^T J, det A and A^-1 by cofactors, detJ = sqrt(det A)
//     JxW      = detJ w_q
//     J+[a][r] = sum_b A^-1[a][b] J[r][b]
//     dphi_i[r]= sum_a dphiRef_i(q)[a] J+[a][r]
//   per term in call order, per q ascending (include/assembly_lhs.hpp:52-65,
//   93-109, 207-221, 268-280):
//     Ke[i][j] += s_q (dphi_i . dphi_j)   or   s_q (phi_i phi_j)
//     s_q = c JxW | kappa_q JxW | c JxW | (c kappa_q) JxW,
//     kappa_q = sum_i phi_i(q) kappa_i                 (FEVar::evaluate)
// then the per-element Dirichlet elimination of Builder::buildLhs
// (include/builder.hpp:83-124).  The global sums (scatter in element order)
// then give the same bits, so the pruned pattern is the same entry for entry.
//
// One block of 128 threads takes EPB elements: phase 0 stages vertex
// coordinates, coefficient values and Dirichlet data in LDS; phase 1 has one
// thread per (element, point) building the geometry and all physical
// gradients of that point into LDS; phase 2 has one thread per (element,
// row) contracting its row.  The geometry never goes to HBM.
#include "kernels.hpp"

#include <algorithm>

// every expression below is evaluated as written: no fused multiply-add
#pragma clang fp contract(off)

namespace pxg
{

namespace
{
constexpr int kExactThreads = 128;
constexpr int kExactMaxTerms = 8;

struct ExactTerms
{
  int n;
  int kind[kExactMaxTerms];
  double coef[kExactMaxTerms];
};

template <int N, int DIM>
__global__ void __launch_bounds__(kExactThreads)
    k_element_exact(int64_t E, int nq, int epb, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
                    double const * __restrict__ grad, double const * __restrict__ phi, double const * __restrict__ w,
                    uint32_t const * __restrict__ emap, double const * __restrict__ kappa, ExactTerms terms,
                    uint8_t const * __restrict__ bcFlag, double const * __restrict__ bcVal, int hasBC,
                    double * __restrict__ ke, double * __restrict__ fe, int * __restrict__ status)
{
  // LDS: vertex coordinates, coefficient and Dirichlet data per element,
  // term weights per (element, point, term), physical gradients per
  // (element, point, basis)
  extern __shared__ double lds[];
  double * sX = lds;                       // epb x N x 3
  double * sK = sX + epb * N * 3;          // epb x N
  double * sH = sK + epb * N;              // epb x N Dirichlet values
  double * sS = sH + epb * N;              // epb x nq x terms.n
  double * sD = sS + epb * nq * terms.n;   // epb x nq x N x 3
  int * sC = reinterpret_cast<int *>(sD + epb * nq * N * 3); // epb x N constrained flags
  int64_t const e0 = (int64_t)blockIdx.x * epb;
  int const ne = (int)std::min<int64_t>(epb, E - e0);
  bool needK = false;
  for (int t = 0; t < terms.n; ++t)
    needK = needK || terms.kind[t] == PXG_TERM_STIFFNESS_FE || terms.kind[t] == PXG_TERM_MASS_FE;

  // phase 0: element data
  for (int k = threadIdx.x; k < ne * N; k += kExactThreads)
  {
    int64_t const ge = e0 * N + k;
    uint32_t const pid = conn[ge];
    sX[3 * k] = xyz[3ull * pid];
    sX[3 * k + 1] = xyz[3ull * pid + 1];
    sX[3 * k + 2] = xyz[3ull * pid + 2];
    uint32_t const id = emap[ge];
    sK[k] = needK ? kappa[id] : 0.;
    bool const con = hasBC && bcFlag[id];
    sC[k] = con;
    sH[k] = con ? bcVal[id] : 0.;
  }
  __syncthreads();

  // phase 1: geometry of one (element, point), gradients of every basis
  for (int k = threadIdx.x; k < ne * nq; k += kExactThreads)
  {
    int const el = k / nq, q = k - el * nq;
    double const * x = sX + el * N * 3;
    double J[3][DIM];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < DIM; ++c)
        J[r][c] = 0.;
#pragma unroll
    for (int g = 0; g < N; ++g)
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          J[r][c] += x[3 * g + r] * grad[(q * N + g) * DIM + c];
    double A[DIM][DIM];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int b = 0; b < DIM; ++b)
      {
        double s = 0.;
#pragma unroll
        for (int r = 0; r < 3; ++r)
          s += J[r][a] * J[r][b];
        A[a][b] = s;
      }
    double det, I[DIM][DIM];
    if constexpr (DIM == 1)
    {
      det = A[0][0];
      I[0][0] = 1. / A[0][0];
    }
    else if constexpr (DIM == 2)
    {
      det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
      I[0][0] = A[1][1] / det;
      I[0][1] = -A[0][1] / det;
      I[1][0] = -A[1][0] / det;
      I[1][1] = A[0][0] / det;
    }
    else
    {
      double const c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
      double const c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
      double const c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
      det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
      I[0][0] = c00 / det;
      I[1][0] = c01 / det;
      I[2][0] = c02 / det;
      I[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
      I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
      I[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
      I[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
      I[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
      I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    }
    if (!(det > 0.))
      atomicOr(status, 1);
    double const JxW = sqrt(det) * w[q];
    double Jp[DIM][3];
#pragma unroll
    for (int a = 0; a < DIM; ++a)
#pragma unroll
      for (int r = 0; r < 3; ++r)
      {
        double s = 0.;
#pragma unroll
        for (int b = 0; b < DIM; ++b)
          s += I[a][b] * J[r][b];
        Jp[a][r] = s;
      }
    double kq = 0.;
    if (needK)
    {
      double const * kl = sK + el * N;
#pragma unroll
      for (int i = 0; i < N; ++i)
        kq += phi[q * N + i] * kl[i];
    }
    double * so = sS + (el * nq + q) * terms.n;
    for (int t = 0; t < terms.n; ++t)
    {
      double const c = terms.coef[t];
      double s;
      switch (terms.kind[t])
      {
      case PXG_TERM_STIFFNESS_FE: s = (kq * JxW) * c; break;
      case PXG_TERM_MASS_FE: s = (c * kq) * JxW; break;
      default: s = c * JxW; break;
      }
      so[t] = s;
    }
    double * d = sD + (el * nq + q) * N * 3;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int r = 0; r < 3; ++r)
      {
        double s = 0.;
#pragma unroll
        for (int a = 0; a < DIM; ++a)
          s += grad[(q * N + i) * DIM + a] * Jp[a][r];
        d[i * 3 + r] = s;
      }
  }
  __syncthreads();

  // phase 2: row i of one element
  for (int k = threadIdx.x; k < ne * N; k += kExactThreads)
  {
    int const el = k / N, i = k - el * N;
    double acc[N];
#pragma unroll
    for (int j = 0; j < N; ++j)
      acc[j] = 0.;
    for (int t = 0; t < terms.n; ++t)
    {
      bool const stiff = terms.kind[t] == PXG_TERM_STIFFNESS || terms.kind[t] == PXG_TERM_STIFFNESS_FE;
      for (int q = 0; q < nq; ++q)
      {
        double const s = sS[(el * nq + q) * terms.n + t];
        if (stiff)
        {
          double const * d = sD + (el * nq + q) * N * 3;
          double const g0 = d[i * 3], g1 = d[i * 3 + 1], g2 = d[i * 3 + 2];
#pragma unroll
          for (int j = 0; j < N; ++j)
          {
            double const v = g0 * d[j * 3] + g1 * d[j * 3 + 1] + g2 * d[j * 3 + 2];
            acc[j] += s * v;
          }
        }
        else
        {
          double const pi = phi[q * N + i];
#pragma unroll
          for (int j = 0; j < N; ++j)
          {
            double const v = pi * phi[q * N + j];
            acc[j] += s * v;
          }
        }
      }
    }
    // Dirichlet elimination: Fe_i = -sum_j Ke_ij h_j on free rows (h_j = 0
    // on free columns adds exact zeros, skipped), unit row and Fe_i = h_i on
    // constrained ones, constrained columns cleared
    double f = 0.;
    if (hasBC)
    {
      double const * h = sH + el * N;
      int const * con = sC + el * N;
      if (con[i])
      {
        f = h[i];
#pragma unroll
        for (int j = 0; j < N; ++j)
          acc[j] = j == i ? 1. : 0.;
      }
      else
      {
        double s = 0.;
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (con[j])
          {
            s += acc[j] * h[j];
            acc[j] = 0.;
          }
        f = 0. - s;
      }
    }
    int64_t const row = e0 * N + k;
    double * out = ke + row * keRowStride(N);
#pragma unroll
    for (int j = 0; j < N; ++j)
      out[j] = acc[j];
    fe[row] = f;
  }
}

template <int N, int DIM>
void launchExact(pxg_ctx * c, ExactTerms const & terms)
{
  FeTables const & T = c->fe;
  int const nq = T.nq;
  int const epb = std::max(1, kExactThreads / std::max(N, nq));
  size_t const lds =
      sizeof(double) * (size_t)epb * (N * 3 + 2 * N + nq * terms.n + nq * N * 3) + sizeof(int) * (size_t)epb * N;
  PXG_REQUIRE(lds <= 65536, PXG_E_INVALID, "quadrature rule too large for the linear element kernel");
  int64_t const E = c->nelem;
  unsigned const grid = (unsigned)((E + epb - 1) / epb);
  c->keBuf.resize((size_t)E * N * keRowStride(N));
  k_element_exact<N, DIM><<<grid, kExactThreads, lds, c->stream>>>(
      E, nq, epb, c->xyz.p, c->conn.p, T.d_grad.p, T.d_phi.p, T.d_w.p, c->elemMap.p, c->kappa.p, terms, c->bcFlag.p,
      c->bcVal.p, c->hasBC, c->keBuf.p, c->feBuf.p, c->status.p);
  PXG_HIP(hipGetLastError());
}
} // namespace

bool elementExact(pxg_ctx * c, int nterms, pxg_term const * terms)
{
  FeTables const & T = c->fe;
  // the linear elements map through their own basis: mapping points are the
  // vertices and the mapping gradients are the basis gradients
  if (T.feType != PXG_FE_LINE_P1 && T.feType != PXG_FE_QUAD_Q1 && T.feType != PXG_FE_HEX_Q1 &&
      T.feType != PXG_FE_TET_P1)
    return false;
  if (nterms > kExactMaxTerms || T.npe != T.n)
    return false;
  ExactTerms et{};
  et.n = nterms;
  for (int k = 0; k < nterms; ++k)
  {
    et.kind[k] = terms[k].kind;
    et.coef[k] = terms[k].coef;
  }
  switch (T.feType)
  {
  case PXG_FE_LINE_P1: launchExact<2, 1>(c, et); break;
  case PXG_FE_QUAD_Q1: launchExact<4, 2>(c, et); break;
  case PXG_FE_HEX_Q1: launchExact<8, 3>(c, et); break;
  case PXG_FE_TET_P1: launchExact<4, 3>(c, et); break;
  }
  return true;
}

} // namespace pxg
