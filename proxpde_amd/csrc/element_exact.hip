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
// One lane per element with the element in registers (coordinates, the
// upper triangle of Ke, the gradients of the current point): no LDS, no
// barriers, the per-point tables read with scalar loads since every lane is
// at the same point.  The geometry never goes to HBM.
#include "kernels.hpp"

#include <algorithm>

// every expression below is evaluated as written: no fused multiply-add
#pragma clang fp contract(off)

namespace pxg
{

namespace
{
constexpr int kExactThreads = 256;
constexpr int kExactMaxTerms = 8;

struct ExactTerms
{
  int n;
  int kind[kExactMaxTerms];
  double coef[kExactMaxTerms];
};

// geometry of one quadrature point: J+ and JxW from the vertex coordinates
// (CurFE::reinit, include/curfe.hpp:50-103, in its order)
template <int N, int DIM>
__device__ inline void exactGeometry(double const (&x)[N][3], double const * __restrict__ gq, double wq,
                                     double (&Jp)[DIM][3], double & JxW, bool & bad)
{
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
        J[r][c] += x[g][r] * gq[g * DIM + c];
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
  bad = bad || !(det > 0.);
  JxW = sqrt(det) * wq;
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
}

// One lane per element, everything in registers: the loops over points run
// in the same order for every lane, so the tables are read with scalar loads.
// Entry (i, j) and (j, i) are the same bits (each product commutes and the
// sums run in the same order), so the upper triangle is accumulated and
// mirrored.
template <typename TO, int N, int DIM>
__global__ void __launch_bounds__(kExactThreads) __attribute__((amdgpu_waves_per_eu(2)))
    k_element_exact(int64_t E, int nq, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
                    double const * __restrict__ grad, double const * __restrict__ phi, double const * __restrict__ w,
                    uint32_t const * __restrict__ emap, double const * __restrict__ kappa, ExactTerms terms,
                    uint8_t const * __restrict__ bcFlag, double const * __restrict__ bcVal, int hasBC,
                    TO * __restrict__ ke, double * __restrict__ fe, int * __restrict__ status,
                    uint32_t const * __restrict__ rowPos)
{
  constexpr int NP = N * (N + 1) / 2; // upper-triangle entries, row-major
  int64_t const e = (int64_t)blockIdx.x * kExactThreads + threadIdx.x;
  if (e >= E)
    return;
  bool needK = false;
  for (int t = 0; t < terms.n; ++t)
    needK = needK || terms.kind[t] == PXG_TERM_STIFFNESS_FE || terms.kind[t] == PXG_TERM_MASS_FE;

  double x[N][3];
  uint32_t id[N];
#pragma unroll
  for (int g = 0; g < N; ++g)
  {
    uint32_t const pid = conn[e * N + g];
    x[g][0] = xyz[3ull * pid];
    x[g][1] = xyz[3ull * pid + 1];
    x[g][2] = xyz[3ull * pid + 2];
    id[g] = emap[e * N + g];
  }
  double kl[N];
#pragma unroll
  for (int i = 0; i < N; ++i)
    kl[i] = needK ? kappa[id[i]] : 0.;

  double acc[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k)
    acc[k] = 0.;
  bool bad = false;
  for (int t = 0; t < terms.n; ++t)
  {
    int const kind = terms.kind[t];
    double const c = terms.coef[t];
    bool const stiff = kind == PXG_TERM_STIFFNESS || kind == PXG_TERM_STIFFNESS_FE;
    for (int q = 0; q < nq; ++q)
    {
      double Jp[DIM][3], JxW;
      exactGeometry<N, DIM>(x, grad + q * N * DIM, w[q], Jp, JxW, bad);
      double const * pq = phi + q * N;
      double kq = 0.;
      if (needK)
      {
#pragma unroll
        for (int i = 0; i < N; ++i)
          kq += pq[i] * kl[i];
      }
      double s;
      switch (kind)
      {
      case PXG_TERM_STIFFNESS_FE: s = (kq * JxW) * c; break;
      case PXG_TERM_MASS_FE: s = (c * kq) * JxW; break;
      default: s = c * JxW; break;
      }
      if (stiff)
      {
        double d[N][3];
        double const * gq = grad + q * N * DIM;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
          for (int r = 0; r < 3; ++r)
          {
            double v = 0.;
#pragma unroll
            for (int a = 0; a < DIM; ++a)
              v += gq[i * DIM + a] * Jp[a][r];
            d[i][r] = v;
          }
        int k = 0;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
          for (int j = i; j < N; ++j, ++k)
          {
            double const v = d[i][0] * d[j][0] + d[i][1] * d[j][1] + d[i][2] * d[j][2];
            acc[k] += s * v;
          }
      }
      else
      {
        int k = 0;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
          for (int j = i; j < N; ++j, ++k)
          {
            double const v = pq[i] * pq[j];
            acc[k] += s * v;
          }
      }
    }
  }
  if (bad)
    atomicOr(status, 1);
  if constexpr (sizeof(TO) == 4)
  {
    // FP32 output: the entries that are zero in exact arithmetic come out as
    // FP64 rounding residues; those below 2^-19 sqrt(Ke_ii Ke_jj) are
    // flushed so the FP32 matrix has the exact-arithmetic sparsity
    float dg[N];
    int k = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j, ++k)
        if (j == i)
          dg[i] = fabsf((float)acc[k]);
    k = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j, ++k)
        if (j != i && fabsf((float)acc[k]) <= 0x1p-19f * sqrtf(dg[i] * dg[j]))
          acc[k] = 0.;
  }

  // Dirichlet elimination per row, then the row store.  Fe_i = -sum_j Ke_ij
  // h_j on free rows (h_j = 0 on free columns adds exact zeros, skipped),
  // unit row and Fe_i = h_i on constrained ones, constrained columns cleared
  // (include/builder.hpp:83-124)
  bool con[N];
  double h[N];
#pragma unroll
  for (int i = 0; i < N; ++i)
  {
    con[i] = hasBC && bcFlag[id[i]];
    h[i] = con[i] ? bcVal[id[i]] : 0.;
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
  {
    double row[N]; // in the output precision
#pragma unroll
    for (int j = 0; j < N; ++j)
    {
      int const a = i < j ? i : j, b = i < j ? j : i;
      row[j] = (double)(TO)acc[a * N - a * (a - 1) / 2 + (b - a)];
    }
    double f = 0.;
    if (hasBC)
    {
      if (con[i])
      {
        f = h[i];
#pragma unroll
        for (int j = 0; j < N; ++j)
          row[j] = j == i ? 1. : 0.;
      }
      else
      {
        double sum = 0.;
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (con[j])
          {
            sum += row[j] * h[j];
            row[j] = 0.;
          }
        f = 0. - sum;
      }
    }
    // row (e, i) in element order, or at its pair position (rows of N values
    // either way: the linear elements have no row padding)
    int64_t const r = rowPos ? (int64_t)rowPos[e * N + i] : e * N + i;
    TO * out = ke + r * N;
    if constexpr (sizeof(TO) == 8 && N % 2 == 0)
    {
#pragma unroll
      for (int j = 0; j < N; j += 2)
        *reinterpret_cast<double2 *>(out + j) = double2{row[j], row[j + 1]};
    }
    else if constexpr (sizeof(TO) == 4 && N % 4 == 0)
    {
#pragma unroll
      for (int j = 0; j < N; j += 4)
        *reinterpret_cast<float4 *>(out + j) = float4{(float)row[j], (float)row[j + 1], (float)row[j + 2],
                                                      (float)row[j + 3]};
    }
    else
    {
#pragma unroll
      for (int j = 0; j < N; ++j)
        out[j] = row[j];
    }
    fe[r] = f;
  }
}

template <int N, int DIM>
void launchExact(pxg_ctx * c, ExactTerms const & terms, int precision)
{
  FeTables const & T = c->fe;
  int64_t const E = c->nelem;
  unsigned const grid = (unsigned)((E + kExactThreads - 1) / kExactThreads);
  size_t const sz = (size_t)E * N * keRowStride(N);
  if (precision == PXG_FP64)
    c->keBuf.resize(sz);
  else
    c->keBuf32.resize(sz);
  if (E == 0)
    return;
  static_assert(keRowStride(N) == N);
  uint32_t const * rowPos = c->kePairOrder ? c->pairPos.p : nullptr;
  if (precision == PXG_FP64)
    k_element_exact<double, N, DIM><<<grid, kExactThreads, 0, c->stream>>>(
        E, T.nq, c->xyz.p, c->conn.p, T.d_grad.p, T.d_phi.p, T.d_w.p, c->elemMap.p, c->kappa.p, terms, c->bcFlag.p,
        c->bcVal.p, c->hasBC, c->keBuf.p, c->feBuf.p, c->status.p, rowPos);
  else
    k_element_exact<float, N, DIM><<<grid, kExactThreads, 0, c->stream>>>(
        E, T.nq, c->xyz.p, c->conn.p, T.d_grad.p, T.d_phi.p, T.d_w.p, c->elemMap.p, c->kappa.p, terms, c->bcFlag.p,
        c->bcVal.p, c->hasBC, c->keBuf32.p, c->feBuf.p, c->status.p, rowPos);
  PXG_HIP(hipGetLastError());
}
} // namespace

bool elementExact(pxg_ctx * c, int nterms, pxg_term const * terms, int precision)
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
  case PXG_FE_LINE_P1: launchExact<2, 1>(c, et, precision); break;
  case PXG_FE_QUAD_Q1: launchExact<4, 2>(c, et, precision); break;
  case PXG_FE_HEX_Q1: launchExact<8, 3>(c, et, precision); break;
  case PXG_FE_TET_P1: launchExact<4, 3>(c, et, precision); break;
  }
  return true;
}

} // namespace pxg
