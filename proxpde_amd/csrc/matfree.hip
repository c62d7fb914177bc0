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
constexpr int kMaxNb = 16;
constexpr int kMfWaves = 8; // wavefronts per block (they share the tables)
constexpr int kMaxQ = 64;

__device__ inline void waveSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of one element being processed by a group of SUB lanes
// (NBC vectors of the batch are staged at a time)
template <int DIM, int NPE, int N, int SUB, int NBC, bool GRAD>
struct ElemLds
{
  double xyz[NPE * 3];
  double kap[N];
  int flag[N];
  double X[N][NBC];
  double Y[GRAD ? N : 1][NBC];
  double Wf[GRAD ? 1 : SUB][NBC][DIM];
  double C[GRAD ? SUB : 1];
};

// One group of SUB lanes per element (64 / SUB elements per wavefront), the
// constant tables of the block in LDS.  Per element:
//   gather   lane l < N: dof id, kappa, constraint flag; lane l < 3 NPE: one
//            vertex coordinate; then per chunk of the batch the x (and ybar)
//            values, zero on constrained dofs
//   point    lane q < nq: J, A = J^T J, A^{-1}, JxW, kappa_q; reference
//            gradients v_b = sum_j g_j(q) x_{b,j} read from LDS;
//            apply: w_bq = kappa_q JxW_q A^{-1} v_bq -> LDS
//            grad : c_q += sum_b v_bq(x)^T A^{-1} v_bq(ybar)
//   row      lane i < N: apply: y_{b,i} = sum_q g_i(q) . w_bq (constrained rows:
//            x_{b,i}, the unit diagonal of each element) -> element buffer;
//            grad: g_i = sum_q phi_i(q) JxW_q c_q -> element buffer
// All sums run in a fixed order; the element buffer is reduced per dof in
// element order by k_gather, so the result is bitwise repeatable.
template <int DIM, int NPE, int N, int SUB, int NBC, bool GRAD>
__global__ void __launch_bounds__(64 * kMfWaves)
    k_mf_elem(int64_t E, int nq, int nb, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
              double const * __restrict__ geoGradG, double const * __restrict__ wG, double const * __restrict__ phiG,
              double const * __restrict__ gradG, uint32_t const * __restrict__ emap,
              double const * __restrict__ kappa, uint8_t const * __restrict__ cons, double const * __restrict__ x,
              double const * __restrict__ ybar, double * __restrict__ out, int * __restrict__ status,
              int * __restrict__ work)
{
  constexpr int EPW = 64 / SUB;
  using L = ElemLds<DIM, NPE, N, SUB, NBC, GRAD>;
  // tables for nq <= SUB quadrature points
  __shared__ double sGrad[SUB * N * DIM];
  __shared__ double sPhi[SUB * N];
  __shared__ double sGeo[SUB * NPE * DIM];
  __shared__ double sW[SUB];
  __shared__ L lds[kMfWaves * EPW];
  int const tid = threadIdx.x;
  for (int k = tid; k < nq * N * DIM; k += blockDim.x)
    sGrad[k] = gradG[k];
  for (int k = tid; k < nq * N; k += blockDim.x)
    sPhi[k] = phiG[k];
  for (int k = tid; k < nq * NPE * DIM; k += blockDim.x)
    sGeo[k] = geoGradG[k];
  for (int k = tid; k < nq; k += blockDim.x)
    sW[k] = wG[k];
  __syncthreads();

  int const wave = tid >> 6, lane = tid & 63;
  int const sub = lane / SUB, l = lane - sub * SUB;
  L & M = lds[wave * EPW + sub];
  // gathered data of the next element, loaded one element ahead so the
  // global-memory latency hides behind the current element's work:
  // ids first (conn, emap), values (coordinates, kappa, flag, the first
  // chunk of x / ybar) after the geometry phase
  constexpr int CPL = (NPE * 3 + SUB - 1) / SUB; // coordinates per lane
  uint32_t pPid[CPL], pId = 0;
  double pX[CPL], pK = 0., pXv[NBC], pYv[NBC];
  int pF = 0;
  auto gatherIds = [&](int64_t el)
  {
    if (el < E)
    {
#pragma unroll
      for (int c = 0; c < CPL; ++c)
        if (l + c * SUB < NPE * 3)
          pPid[c] = conn[el * NPE + (l + c * SUB) / 3];
      if (l < N)
        pId = emap[el * N + l];
    }
  };
  auto gatherVals = [&](int64_t el)
  {
    if (el < E)
    {
#pragma unroll
      for (int c = 0; c < CPL; ++c)
        if (l + c * SUB < NPE * 3)
          pX[c] = xyz[3ull * pPid[c] + (l + c * SUB) % 3];
      if (l < N)
      {
        if constexpr (!GRAD)
          pK = kappa[pId];
        pF = cons ? (int)cons[pId] : 0;
#pragma unroll
        for (int bb = 0; bb < NBC; ++bb)
        {
          pXv[bb] = bb < nb ? x[(size_t)pId * nb + bb] : 0.;
          if constexpr (GRAD)
            pYv[bb] = bb < nb ? ybar[(size_t)pId * nb + bb] : 0.;
        }
      }
    }
  };
  // waves take batches of EPW consecutive elements from counters: blocks b
  // with the same b % P (P = min(grid, kWorkCounters)) share a counter over
  // one contiguous range of batches, so waves of unequal issue priority end
  // together; the batch after next is requested one batch ahead
  int const parts = min((int)gridDim.x, kWorkCounters);
  int const part = (int)(blockIdx.x % parts);
  int64_t const nbat = (E + EPW - 1) / EPW;
  int64_t const q0 = nbat * part / parts, q1 = nbat * (part + 1) / parts;
  int * const ctr = work + part * kWorkCounterStride;
  auto grab = [&]() -> int
  {
    int v = 0;
    if (lane == 0)
      v = atomicAdd(ctr, 1);
    return v;
  };
  auto batch = [&](int v) -> int64_t // first element of the batch (wave-uniform), E when done
  {
    int64_t const q = q0 + __builtin_amdgcn_readfirstlane(v);
    return q < q1 ? q * EPW : E;
  };
  int64_t base = batch(grab());
  int64_t nbase = batch(grab());
  gatherIds(base + sub);
  gatherVals(base + sub);
  // the wave loops uniformly; each group checks its own element
  while (base < E)
  {
    int const req = grab(); // batch after next (lane 0)
    int64_t const e = base + sub;
    bool const live = e < E;
    uint32_t id = 0;
    if (live)
    {
#pragma unroll
      for (int c = 0; c < CPL; ++c)
        if (l + c * SUB < NPE * 3)
          M.xyz[l + c * SUB] = pX[c];
      if (l < N)
      {
        id = pId;
        M.kap[l] = pK;
        M.flag[l] = pF;
#pragma unroll
        for (int bb = 0; bb < NBC; ++bb)
        {
          M.X[l][bb] = pF ? 0. : pXv[bb];
          if constexpr (GRAD)
            M.Y[l][bb] = pF ? 0. : pYv[bb];
        }
      }
    }
    waveSync();
    gatherIds(nbase + sub);
    // geometry of point q = l, kept in registers across the batch
    double I00 = 0., I01 = 0., I02 = 0., I11 = 0., I12 = 0., I22 = 0., sq = 0.;
    if (live && l < nq)
    {
      int const q = l;
      double J[3][DIM];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          J[r][c] = 0.;
#pragma unroll
      for (int v = 0; v < NPE; ++v)
#pragma unroll
        for (int c = 0; c < DIM; ++c)
        {
          double const g = sGeo[(q * NPE + v) * DIM + c];
          J[0][c] += M.xyz[3 * v] * g;
          J[1][c] += M.xyz[3 * v + 1] * g;
          J[2][c] += M.xyz[3 * v + 2] * g;
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
        I00 = 1. / det;
      }
      else if constexpr (DIM == 2)
      {
        det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        I00 = A[1][1] / det;
        I01 = -A[0][1] / det;
        I11 = A[0][0] / det;
      }
      else
      {
        double const c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        double const c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
        double const c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
        det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
        I00 = c00 / det;
        I01 = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
        I02 = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
        I11 = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
        I12 = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
        I22 = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
      }
      if (!(det > 0.))
        atomicOr(status, 1);
      double const jxw = sqrt(det) * sW[q];
      if constexpr (GRAD)
        sq = jxw;
      else
      {
        double kq = 0.;
        for (int i = 0; i < N; ++i)
          kq += sPhi[q * N + i] * M.kap[i];
        sq = kq * jxw;
      }
    }
    gatherVals(nbase + sub);
    double cq = 0.; // gradient: sum over the batch at point q
    for (int b0 = 0; b0 < nb; b0 += NBC)
    {
      if (b0 > 0 && live && l < N) // the first chunk came with the prefetch
      {
        bool const fl = M.flag[l] != 0;
#pragma unroll
        for (int bb = 0; bb < NBC; ++bb)
        {
          int const b = b0 + bb;
          M.X[l][bb] = (b < nb && !fl) ? x[(size_t)id * nb + b] : 0.;
          if constexpr (GRAD)
            M.Y[l][bb] = (b < nb && !fl) ? ybar[(size_t)id * nb + b] : 0.;
        }
      }
      waveSync();
      if (live && l < nq)
      {
        int const q = l;
#pragma unroll 1
        for (int bb = 0; bb < NBC; ++bb)
        {
          double v[3] = {0., 0., 0.}, u[3] = {0., 0., 0.};
#pragma unroll 3
          for (int j = 0; j < N; ++j)
          {
            double const xv = M.X[j][bb];
            double yv = 0.;
            if constexpr (GRAD)
              yv = M.Y[j][bb];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
            {
              double const g = sGrad[(q * N + j) * DIM + d];
              v[d] += g * xv;
              if constexpr (GRAD)
                u[d] += g * yv;
            }
          }
          // I v with I symmetric (00 01 02 11 12 22)
          double t[3];
          t[0] = I00 * v[0] + (DIM > 1 ? I01 * v[1] : 0.) + (DIM > 2 ? I02 * v[2] : 0.);
          t[1] = DIM > 1 ? I01 * v[0] + I11 * v[1] + (DIM > 2 ? I12 * v[2] : 0.) : 0.;
          t[2] = DIM > 2 ? I02 * v[0] + I12 * v[1] + I22 * v[2] : 0.;
          if constexpr (GRAD)
          {
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              cq += t[d] * u[d];
          }
          else
          {
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              M.Wf[q][bb][d] = sq * t[d];
          }
        }
      }
      waveSync();
      if constexpr (!GRAD)
      {
        if (live && l < N)
        {
          int const i = l;
          bool const fl = M.flag[i] != 0;
#pragma unroll
          for (int bb = 0; bb < NBC; ++bb)
          {
            int const b = b0 + bb;
            if (b >= nb)
              break;
            double y = 0.;
            if (fl)
              y = x[(size_t)id * nb + b]; // unit diagonal per element on constrained rows
            else
#pragma unroll 3
              for (int q = 0; q < nq; ++q)
              {
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  y += sGrad[(q * N + i) * DIM + d] * M.Wf[q][bb][d];
              }
            out[((size_t)e * N + i) * nb + b] = y;
          }
        }
        waveSync(); // X and Wf are rewritten by the next chunk
      }
    }
    if constexpr (GRAD)
    {
      if (live && l < nq)
        M.C[l] = cq * sq;
      waveSync();
      if (live && l < N)
      {
        double g = 0.;
        for (int q = 0; q < nq; ++q)
          g += sPhi[q * N + l] * M.C[q];
        out[(size_t)e * N + l] = g;
      }
      waveSync();
    }
    base = nbase;
    nbase = batch(req);
  }
}

__global__ void k_reset_work(int * c)
{
  if (threadIdx.x < kWorkCounters)
    c[threadIdx.x * kWorkCounterStride] = 0;
}

// y[d][b] = sum over the (element, local) entries of dof d, in element order;
// the first kMaxC entries are loaded ahead (indices clamped into the segment)
__global__ void k_gather(int64_t ndofs, int nb, uint32_t const * __restrict__ seg,
                         uint32_t const * __restrict__ ids, double const * __restrict__ buf, double * __restrict__ y)
{
  int64_t const t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ndofs * nb)
    return;
  int64_t const d = t / nb;
  int const b = (int)(t - d * nb);
  uint32_t const k0 = seg[d], k1 = seg[d + 1];
  double acc = 0.;
  if (k1 > k0)
  {
    constexpr int kMaxC = 8;
    uint32_t id[kMaxC];
#pragma unroll
    for (int j = 0; j < kMaxC; ++j)
      id[j] = ids[min(k0 + j, k1 - 1)];
    double v[kMaxC];
#pragma unroll
    for (int j = 0; j < kMaxC; ++j)
      v[j] = buf[(size_t)id[j] * nb + b];
#pragma unroll
    for (int j = 0; j < kMaxC; ++j)
      if (k0 + j < k1)
        acc += v[j];
    for (uint32_t k = k0 + kMaxC; k < k1; ++k)
      acc += buf[(size_t)ids[k] * nb + b];
  }
  y[t] = acc;
}

// compile-time element shape per FE type: (DIM, NPE, N)
template <typename F>
void dispatchFeMf(int feType, F && f)
{
  using std::integral_constant;
  switch (feType)
  {
  case PXG_FE_LINE_P1: f(integral_constant<int, 1>{}, integral_constant<int, 2>{}, integral_constant<int, 2>{}); break;
  case PXG_FE_QUAD_Q1: f(integral_constant<int, 2>{}, integral_constant<int, 4>{}, integral_constant<int, 4>{}); break;
  case PXG_FE_QUAD_Q2: f(integral_constant<int, 2>{}, integral_constant<int, 4>{}, integral_constant<int, 9>{}); break;
  case PXG_FE_HEX_Q1: f(integral_constant<int, 3>{}, integral_constant<int, 8>{}, integral_constant<int, 8>{}); break;
  case PXG_FE_HEX_Q2: f(integral_constant<int, 3>{}, integral_constant<int, 8>{}, integral_constant<int, 27>{}); break;
  case PXG_FE_TET_P1: f(integral_constant<int, 3>{}, integral_constant<int, 4>{}, integral_constant<int, 4>{}); break;
  case PXG_FE_TET_P2: f(integral_constant<int, 3>{}, integral_constant<int, 4>{}, integral_constant<int, 10>{}); break;
  default: throw Error(PXG_E_INVALID, "no matrix-free operator for this FE type");
  }
}

void prepare(pxg_ctx * c, int nb)
{
  PXG_REQUIRE(c->fe.feType && c->nelem && c->ndofs, PXG_E_STATE, "matrix-free needs set_fe, set_mesh, set_dofmap");
  PXG_REQUIRE(nb >= 1 && nb <= kMaxNb, PXG_E_INVALID, "batch size must be 1..16");
  PXG_REQUIRE(c->fe.nq <= kMaxQ, PXG_E_INVALID, "at most 64 quadrature points");
  if (!c->dofSeg.p)
    buildDofAdjacency(c);
}

// one launch of the element kernel: lanes per element = the smallest power of
// two (at least 8) holding n and nq; one resident wave of blocks taking batches
// of elements from the work counters
template <bool GRAD>
void launchElem(pxg_ctx * c, double const * kappa, uint8_t const * cons, int nb, double const * x,
                double const * ybar, double * outBuf)
{
  FeTables & T = c->fe;
  int const need = std::max(T.n, T.nq);
  int dev = 0, ncu = 0;
  PXG_HIP(hipGetDevice(&dev));
  PXG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  dispatchFeMf(T.feType,
               [&](auto dimC, auto npeC, auto nC)
               {
                 constexpr int D = decltype(dimC)::value, NPE = decltype(npeC)::value, NN = decltype(nC)::value;
                 auto go = [&](auto subC, auto nbcC)
                 {
                   constexpr int SUB = decltype(subC)::value, NBC = decltype(nbcC)::value;
                   if constexpr (SUB >= NN) // (the dispatch below never picks a smaller group)
                   {
                   auto kern = k_mf_elem<D, NPE, NN, SUB, NBC, GRAD>;
                   int perCu = 0;
                   PXG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kern, 64 * kMfWaves, 0));
                   int64_t const per = (int64_t)kMfWaves * (64 / SUB);
                   int64_t const need = (c->nelem + per - 1) / per;
                   int const grid = (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ncu * std::max(perCu, 1)));
                   PXG_REQUIRE(need < ((int64_t)1 << 31) - 2 * per * grid, PXG_E_INVALID,
                               "element count exceeds the 32-bit work counters");
                   c->workCounter.resize(kWorkCounters * kWorkCounterStride);
                   k_reset_work<<<1, 64, 0, c->stream>>>(c->workCounter.p);
                   kern<<<grid, 64 * kMfWaves, 0, c->stream>>>(c->nelem, T.nq, nb, c->xyz.p, c->conn.p, T.d_geoGrad.p,
                                                               T.d_w.p, T.d_phi.p, T.d_grad.p, c->elemMap.p, kappa,
                                                               cons, x, ybar, outBuf, c->status.p, c->workCounter.p);
                   }
                 };
                 auto pick = [&](auto subC)
                 {
                   if (nb == 1)
                     go(subC, std::integral_constant<int, 1>{});
                   else
                     go(subC, std::integral_constant<int, 4>{});
                 };
                 if (need <= 8)
                   pick(std::integral_constant<int, 8>{});
                 else if (need <= 16)
                   pick(std::integral_constant<int, 16>{});
                 else if (need <= 32)
                   pick(std::integral_constant<int, 32>{});
                 else
                   pick(std::integral_constant<int, 64>{});
               });
  PXG_HIP(hipGetLastError());
}
} // namespace

void mfApply(pxg_ctx * c, double const * kappa, int nb, double const * x, double * y, bool useBC)
{
  prepare(c, nb);
  FeTables & T = c->fe;
  int64_t const E = c->nelem;
  hipStream_t s = c->stream;
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  ProfScope ps(c, "mf_apply");
  PXG_HIP(hipMemsetAsync(c->status.p, 0, sizeof(int), s));
  if (!(sfAvailable(c) && sfApplyCubes(c, kappa, nb, x, y, useBC, s)))
  {
    c->mfBuf.resize(std::max<size_t>(c->mfBuf.n, (size_t)E * T.n * nb));
    if (sfAvailable(c)) // Q1 / Q2 hexahedra: sum-factorised element kernel
      sfElementApply<double>(c, kappa, nb, x, c->mfBuf.p, useBC, s);
    else
      launchElem<false>(c, kappa, cons, nb, x, nullptr, c->mfBuf.p);
    int64_t const m = c->ndofs * nb;
    k_gather<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(c->ndofs, nb, c->dofSeg.p, c->dofContrib.p, c->mfBuf.p, y);
    PXG_HIP(hipGetLastError());
  }
  int h = 0;
  PXG_HIP(hipMemcpyAsync(&h, c->status.p, sizeof(int), hipMemcpyDeviceToHost, s));
  PXG_HIP(hipStreamSynchronize(s));
  PXG_REQUIRE(h == 0, PXG_E_NUMERIC, "non-positive Jacobian determinant in an element");
}

void mfGradKappa(pxg_ctx * c, int nb, double const * x, double const * ybar, double * gk, bool useBC)
{
  prepare(c, nb);
  FeTables & T = c->fe;
  int64_t const E = c->nelem;
  c->mfBuf.resize(std::max<size_t>(c->mfBuf.n, (size_t)E * T.n));
  hipStream_t s = c->stream;
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  ProfScope ps(c, "mf_gradk");
  PXG_HIP(hipMemsetAsync(c->status.p, 0, sizeof(int), s));
  if (sfAvailable(c))
    sfElementGrad(c, nb, x, ybar, c->mfBuf.p, useBC, s);
  else
    launchElem<true>(c, nullptr, cons, nb, x, ybar, c->mfBuf.p);
  k_gather<<<(unsigned)((c->ndofs + 255) / 256), 256, 0, s>>>(c->ndofs, 1, c->dofSeg.p, c->dofContrib.p, c->mfBuf.p,
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
