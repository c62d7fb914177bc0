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
// Each wavefront takes its own EPW elements (64 / max(n, nq)): phase 0
// stages vertex coordinates, coefficient values and Dirichlet data in LDS;
// phase 1 has one lane per (element, point) building the geometry and all
// physical gradients of that point into LDS; phase 2 computes every entry of
// the upper triangle once (lane per (element, row), cyclic pairs) into the
// element matrix in LDS; phase 3 has one lane per (element, row) for the
// Dirichlet elimination and the row store.  The geometry never goes to HBM.
#include "kernels.hpp"

#include <algorithm>

// every expression below is evaluated as written: no fused multiply-add
#pragma clang fp contract(off)

namespace pxg
{

namespace
{
constexpr int kExactWaves = 2; // wavefronts per block; each works on its own elements
constexpr int kExactMaxTerms = 8;

struct ExactTerms
{
  int n;
  int kind[kExactMaxTerms];
  double coef[kExactMaxTerms];
};

__device__ inline void exactSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS doubles per element of the kernel below
__host__ __device__ constexpr int exactLdsPerElem(int N, int nq, int nterms)
{
  return N * 3 + 2 * N + nq * nterms + (nq * N * 3 > N * N ? nq * N * 3 : N * N) + N; // + flags
}

template <int N, int DIM>
__global__ void __launch_bounds__(64 * kExactWaves)
    k_element_exact(int64_t E, int nq, int epw, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
                    double const * __restrict__ grad, double const * __restrict__ phi, double const * __restrict__ w,
                    uint32_t const * __restrict__ emap, double const * __restrict__ kappa, ExactTerms terms,
                    uint8_t const * __restrict__ bcFlag, double const * __restrict__ bcVal, int hasBC,
                    double * __restrict__ ke, double * __restrict__ fe, int * __restrict__ status)
{
  // LDS of a wave: vertex coordinates, coefficient and Dirichlet data per
  // element, term weights per (element, point, term), physical gradients
  // per (element, point, basis) -- reused for the element matrices once
  // the contractions are done -- and the constrained flags
  extern __shared__ double lds[];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int const per = exactLdsPerElem(N, nq, terms.n);
  double * sX = lds + (size_t)wave * epw * per; // epw x N x 3
  double * sK = sX + epw * N * 3;               // epw x N
  double * sH = sK + epw * N;                   // epw x N Dirichlet values
  double * sS = sH + epw * N;                   // epw x nq x terms.n
  double * sD = sS + epw * nq * terms.n;        // epw x nq x N x 3, then epw x N x N
  double * sC = sD + epw * (nq * N * 3 > N * N ? nq * N * 3 : N * N); // epw x N flags (0 / 1)
  double * sKe = sD;
  int64_t const e0 = ((int64_t)blockIdx.x * kExactWaves + wave) * epw;
  int const ne = (int)(e0 < E ? std::min<int64_t>(epw, E - e0) : 0);
  bool needK = false;
  for (int t = 0; t < terms.n; ++t)
    needK = needK || terms.kind[t] == PXG_TERM_STIFFNESS_FE || terms.kind[t] == PXG_TERM_MASS_FE;

  // phase 0: element data
  for (int k = lane; k < ne * N; k += 64)
  {
    int64_t const ge = e0 * N + k;
    uint32_t const pid = conn[ge];
    sX[3 * k] = xyz[3ull * pid];
    sX[3 * k + 1] = xyz[3ull * pid + 1];
    sX[3 * k + 2] = xyz[3ull * pid + 2];
    uint32_t const id = emap[ge];
    sK[k] = needK ? kappa[id] : 0.;
    bool const con = hasBC && bcFlag[id];
    sC[k] = con ? 1. : 0.;
    sH[k] = con ? bcVal[id] : 0.;
  }
  exactSync();

  // phase 1: geometry of one (element, point), gradients of every basis
  for (int k = lane; k < ne * nq; k += 64)
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
  exactSync();

  // phase 2: every entry of the upper triangle once.  Lane (element, i)
  // takes the pairs (i, (i + d) mod N) for d = 0 .. N/2 (d = N/2 only for
  // i < N/2 when N is even): each unordered pair exactly once.  The value of
  // (i, j) and of (j, i) are the same bits (each product is commutative and
  // the sums run in the same order), so both triangles take it.
  constexpr int ND = N / 2 + 1;
  double acc[ND];
  int const el2 = lane / N, i2 = lane - el2 * N;
  bool const live2 = lane < 64 - 64 % N && el2 < ne;
#pragma unroll
  for (int dd = 0; dd < ND; ++dd)
    acc[dd] = 0.;
  if (live2)
  {
    for (int t = 0; t < terms.n; ++t)
    {
      bool const stiff = terms.kind[t] == PXG_TERM_STIFFNESS || terms.kind[t] == PXG_TERM_STIFFNESS_FE;
      for (int q = 0; q < nq; ++q)
      {
        double const s = sS[(el2 * nq + q) * terms.n + t];
        if (stiff)
        {
          double const * d = sD + (el2 * nq + q) * N * 3;
          double const g0 = d[i2 * 3], g1 = d[i2 * 3 + 1], g2 = d[i2 * 3 + 2];
#pragma unroll
          for (int dd = 0; dd < ND; ++dd)
          {
            int const j = (i2 + dd) % N;
            double const v = g0 * d[j * 3] + g1 * d[j * 3 + 1] + g2 * d[j * 3 + 2];
            acc[dd] += s * v;
          }
        }
        else
        {
          double const pi = phi[q * N + i2];
#pragma unroll
          for (int dd = 0; dd < ND; ++dd)
          {
            int const j = (i2 + dd) % N;
            double const v = pi * phi[q * N + j];
            acc[dd] += s * v;
          }
        }
      }
    }
  }
  exactSync(); // the gradients are read: their LDS takes the element matrices
  if (live2)
  {
    double * K = sKe + el2 * N * N;
#pragma unroll
    for (int dd = 0; dd < ND; ++dd)
    {
      if (N % 2 == 0 && dd == N / 2 && i2 >= N / 2)
        continue;
      int const j = (i2 + dd) % N;
      K[i2 * N + j] = acc[dd];
      K[j * N + i2] = acc[dd];
    }
  }
  exactSync();

  // phase 3: row i of one element, Dirichlet elimination, store.  Fe_i =
  // -sum_j Ke_ij h_j on free rows (h_j = 0 on free columns adds exact
  // zeros, skipped), unit row and Fe_i = h_i on constrained ones,
  // constrained columns cleared (include/builder.hpp:83-124)
  for (int k = lane; k < ne * N; k += 64)
  {
    int const el = k / N, i = k - el * N;
    double row[N];
    double const * K = sKe + el * N * N + i * N;
#pragma unroll
    for (int j = 0; j < N; ++j)
      row[j] = K[j];
    double f = 0.;
    if (hasBC)
    {
      double const * h = sH + el * N;
      double const * con = sC + el * N;
      if (con[i] != 0.)
      {
        f = h[i];
#pragma unroll
        for (int j = 0; j < N; ++j)
          row[j] = j == i ? 1. : 0.;
      }
      else
      {
        double s = 0.;
#pragma unroll
        for (int j = 0; j < N; ++j)
          if (con[j] != 0.)
          {
            s += row[j] * h[j];
            row[j] = 0.;
          }
        f = 0. - s;
      }
    }
    int64_t const r = e0 * N + k;
    double * out = ke + r * keRowStride(N);
#pragma unroll
    for (int j = 0; j < N; ++j)
      out[j] = row[j];
    fe[r] = f;
  }
}

template <int N, int DIM>
void launchExact(pxg_ctx * c, ExactTerms const & terms)
{
  FeTables const & T = c->fe;
  int const nq = T.nq;
  // elements per wave: every lane takes a point in the geometry phase and
  // a row in the contraction
  int const epw = std::max(1, 64 / std::max(N, nq));
  size_t const lds = sizeof(double) * (size_t)kExactWaves * epw * exactLdsPerElem(N, nq, terms.n);
  PXG_REQUIRE(lds <= 65536, PXG_E_INVALID, "quadrature rule too large for the linear element kernel");
  int64_t const E = c->nelem;
  int64_t const per = (int64_t)kExactWaves * epw;
  unsigned const grid = (unsigned)((E + per - 1) / per);
  c->keBuf.resize((size_t)E * N * keRowStride(N));
  k_element_exact<N, DIM><<<grid, 64 * kExactWaves, lds, c->stream>>>(
      E, nq, epw, c->xyz.p, c->conn.p, T.d_grad.p, T.d_phi.p, T.d_w.p, c->elemMap.p, c->kappa.p, terms, c->bcFlag.p,
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
