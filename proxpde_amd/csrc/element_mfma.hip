// Element matrices on the matrix cores (gfx950 MFMA), geometry fused in.
//
// For an element with local basis gradients g_j(q) (reference coordinates),
// per-point metric M_q = s_q (J_q^T J_q)^{-1} and mass weight m_q:
//     Ke = A * C,   A[i][k] = g_i(q,a) | phi_i(q)                     (N x K)
//                   C[k][j] = sum_b M_q^{ab} g_j(q,b) | m_q phi_j(q)  (K x N, per element)
// with k running over the (q, a) pairs (and the mass rows).  Every operand
// value is formed in registers right before its matrix instruction from the
// per-point metric (LDS) and the gradient table (LDS), so no K x N staging
// buffer is needed and more waves fit per CU.  Single-tile elements (n <= 16:
// P2 tet, Q2 quad) step through the points one at a time (rows a = the lane's
// row group); larger ones (Q2 hex) through blocks of four points and one
// component a per step, where a lane's gradient values are both its A and
// the input of its C operands.
// The N x N product runs as v_mfma_f64_16x16x4f64 (FP64) or
// v_mfma_f32_16x16x4f32 (FP32) tiles; only the upper tiles are multiplied
// (Ke is symmetric): the diagonal tiles are stored whole, the off-diagonal
// tile also as its transpose, so the scatter reads whole element rows.  The epilogue applies the per-element
// Dirichlet elimination of Builder::buildLhs (reference
// include/builder.hpp:83-124) on the accumulators and writes Ke plus the
// lifting Fe.
//
// One wavefront per element; kWaves wavefronts per block share the constant
// tables in LDS; waves take elements from a shared counter.
#include "kernels.hpp"
#include "wave.hpp"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace pxg
{

namespace
{

template <typename T>
struct Mfma;
template <>
struct Mfma<double>
{
  using Acc = __attribute__((__vector_size__(4 * sizeof(double)))) double;
  static __device__ inline Acc mma(double a, double b, Acc c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of output element r held by a lane of group g = lane/16 (column lane%16);
  // measured on gfx950: the f64 variant interleaves the four lane groups
  static __device__ inline int row(int r, int g) { return 4 * r + g; }
};
template <>
struct Mfma<float>
{
  using Acc = __attribute__((__vector_size__(4 * sizeof(float)))) float;
  static __device__ inline Acc mma(float a, float b, Acc c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // the f32 variant gives each lane group four consecutive rows
  static __device__ inline int row(int r, int g) { return 4 * g + r; }
};

constexpr int kWaves = 8;
constexpr int kCounters = kWorkCounters;
constexpr int kCounterStride = kWorkCounterStride;

// ordering point for the lanes of one wavefront: LDS accesses of a wave
// execute in order, so a compiler fence plus a wave barrier suffices
__device__ inline void waveSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int DIM, int NQ, bool MASS>
struct Shape
{
  static constexpr int NT = (N + 15) / 16;                // output tiles per side
  static constexpr int KS = NQ * DIM;                     // stiffness rows of C
  static constexpr int K = KS + (MASS ? NQ : 0);          // all rows of C
  static constexpr int STEPS = (K + 3) / 4;               // MFMA k steps of 4
  // single-tile elements (n <= 16): q-major rows, one quadrature point per
  // step with rows a = lane/16 (see k_element_mfma)
  static constexpr bool QM = NT == 1;
  static constexpr int MROWS = QM ? 4 : DIM + (MASS ? 1 : 0); // rows of the C factor per point
  // multi-tile elements: gradient / basis tables and per-point factors padded
  // with zeros to whole tiles of columns and whole blocks of four points, so
  // the operand loop needs no clamps or selects
  static constexpr int NP = QM ? N : 16 * NT;
  static constexpr int NQP = QM ? NQ : 4 * ((NQ + 3) / 4);
};

// per-wave LDS (kept small: two blocks of kWaves waves must fit one CU's LDS)
template <typename T, int N, int NPE, int NQ, int MROWS>
struct WaveLds
{
  double x[NPE * 3];
  double kap[N];
  double bcv[N];
  // rows of the per-point C operand factor: a < DIM the metric row
  // s_q (J^T J)^{-1}[a][:], a = DIM the mass row (m_q, 0, 0) when mass
  // terms are built (always, and zero rows after it, for q-major elements)
  double M[NQ][MROWS][3]; // (NQ: points padded to blocks of four, the padding zero)
  double f[N];  // lifting accumulator of the Dirichlet elimination
  double dg[sizeof(T) == 4 ? N : 1]; // FP32: diagonal of Ke, the flush threshold of noise entries
  int flag[N];
  uint32_t rp[N]; // output row of each element row (element or pair order)
};

template <typename T, int N, int DIM, int NPE, int NQ, bool MASS, bool PO>
__global__ void __launch_bounds__(64 * kWaves, 4)
    k_element_mfma(int64_t E, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
                   uint32_t const * __restrict__ emap, double const * __restrict__ kappa,
                   uint8_t const * __restrict__ bcFlag, double const * __restrict__ bcVal, int hasBC,
                   double const * __restrict__ geoGradG, double const * __restrict__ wG,
                   double const * __restrict__ phiG, double const * __restrict__ gradG,
                   double cS, double cSF, double cM, double cMF, int flags,
                   T * __restrict__ ke, double * __restrict__ fe, int * __restrict__ status,
                   int * __restrict__ work, uint32_t const * __restrict__ rowPos)
{
  using S = Shape<N, DIM, NQ, MASS>;
  using L = WaveLds<T, N, NPE, Shape<N, DIM, NQ, MASS>::NQP, Shape<N, DIM, NQ, MASS>::MROWS>;
  using Acc = typename Mfma<T>::Acc;
  constexpr int NT = S::NT;
  constexpr bool QM = S::QM;
  constexpr int NP = S::NP, NQP = S::NQP;
  bool const needK = (flags & 10) != 0;

  // mapping gradients as [vertex][component][q] and basis values as [node][q]
  // (sPhiT, for the coefficient at the points): the geometry phase has lane q
  // read them, consecutive words across lanes and so no bank conflicts
  __shared__ double sGeo[NQ * NPE * DIM];
  __shared__ double sPhi[NQP * NP]; // phi_j(q) at q NP + j (zero padding)
  __shared__ double sPhiT[N * NQ];
  // reference gradients: g_j(q, b) at (q NP + j) DIM + b (zero padding), or
  // at (q DIM + b) N + j for q-major elements
  __shared__ double sGrad[NQP * NP * DIM];
  __shared__ double sW[NQ];
  __shared__ L lds[kWaves];

  int const tid = threadIdx.x;
  for (int k = tid; k < NQ * NPE * DIM; k += blockDim.x)
  {
    int const q = k / (NPE * DIM), vc = k - q * (NPE * DIM); // source: [q][v][c]
    sGeo[vc * NQ + q] = geoGradG[k];
  }
  for (int k = tid; k < NQ * N; k += blockDim.x)
  {
    int const q = k / N, i = k - q * N;
    sPhiT[i * NQ + q] = phiG[k];
  }
  for (int k = tid; k < NQP * NP; k += blockDim.x)
  {
    int const q = k / NP, j = k - q * NP;
    sPhi[k] = q < NQ && j < N ? phiG[q * N + j] : 0.;
  }
  if constexpr (QM)
  {
    for (int k = tid; k < NQ * N * DIM; k += blockDim.x)
    {
      int const q = k / (N * DIM), j = (k / DIM) % N, b = k % DIM;
      sGrad[(q * DIM + b) * N + j] = gradG[k];
    }
  }
  else
  {
    for (int k = tid; k < NQP * NP * DIM; k += blockDim.x)
    {
      int const q = k / (NP * DIM), j = (k / DIM) % NP, b = k % DIM;
      sGrad[k] = q < NQ && j < N ? gradG[(q * N + j) * DIM + b] : 0.;
    }
  }
  for (int k = tid; k < NQ; k += blockDim.x)
    sW[k] = wG[k];

  int const wave = tid >> 6, lane = tid & 63;
  L & W = lds[wave];
  // the padded points' factors stay zero (the geometry writes q < NQ)
  for (int k = NQ * S::MROWS * 3 + lane; k < NQP * S::MROWS * 3; k += 64)
    (&W.M[0][0][0])[k] = 0.;

  __syncthreads();

  int const kr = lane >> 4, jc = lane & 15;
  // waves own disjoint LDS regions and take elements from counters (waves of
  // one SIMD are issued with different priorities, so a fixed share per wave
  // would leave the slowest waves running alone at the end).  Blocks b with
  // the same b % P (P = min(grid, kCounters)) share a counter over one
  // contiguous range of elements (blocks are placed on the XCDs round-robin,
  // so a range stays in one XCD's L2); every range has blocks, wherever they
  // run.  After the
  // tables are loaded waves synchronise only with themselves.  The element
  // after next is requested one element ahead, so the atomic's latency hides
  // behind an element's work.
  int const parts = min((int)gridDim.x, kCounters);
  int const part = (int)(blockIdx.x % parts);
  int64_t const r0 = E * part / parts, r1 = E * (part + 1) / parts;
  int * const ctr = work + part * kCounterStride;
  auto grab = [&]() -> int
  {
    int v = 0;
    // the counter's address goes through a register the compiler takes as
    // per-lane: an atomic on a uniform address is rewritten into one whose
    // reply is waited for at once, and this one's reply is needed only
    // before the element's stores
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    if (lane == 0)
      v = atomicAdd(ctr + z, 1);
    return v;
  };
  // wave-uniform (scalar registers): lane 0's counter value
  auto elem = [&](int v) -> int64_t
  {
    int64_t const k = r0 + __builtin_amdgcn_readfirstlane(v);
    return k < r1 ? k : E;
  };
  int64_t e = elem(grab());
  int64_t en = elem(grab());
  // gathered data of the element being processed, loaded one element ahead
  double rx = 0., rk = 0., rb = 0.;
  int rf = 0;
  uint32_t rr = 0; // output row of element row `lane` (PO: its pair position)
  auto gatherIds = [&](int64_t el, uint32_t & pid, uint32_t & id)
  {
    if (el < E)
    {
      if (lane < NPE * 3)
        pid = conn[el * NPE + lane / 3];
      if (lane < N)
      {
        id = emap[el * N + lane];
        if constexpr (PO)
          rr = rowPos[el * N + lane];
      }
    }
  };
  auto gatherVals = [&](int64_t el, uint32_t pid, uint32_t id)
  {
    if (el < E)
    {
      if (lane < NPE * 3)
        rx = xyz[3ull * pid + lane % 3];
      if (lane < N)
      {
        rk = needK ? kappa[id] : 0.;
        rf = hasBC ? (int)bcFlag[id] : 0;
        rb = hasBC ? bcVal[id] : 0.; // not dependent on the flag (no serialised load)
      }
    }
  };
  {
    uint32_t pid = 0, id = 0;
    gatherIds(e, pid, id);
    gatherVals(e, pid, id);
  }
  // (no load pending at the loop head)
  asm volatile("" ::"v"(rx), "v"(rk), "v"(rb), "v"(rf), "v"(rr) : "memory");
  while (e < E)
  {
    // ---- 1. element data -> LDS
    if (lane < NPE * 3)
      W.x[lane] = rx;
    if (lane < N)
    {
      W.kap[lane] = rk;
      W.flag[lane] = rf;
      W.bcv[lane] = rf ? rb : 0.;
      W.f[lane] = 0.;
      if constexpr (PO)
        W.rp[lane] = rr;
    }
    waveSync();
    uint32_t npid = 0, nid = 0;
    gatherIds(en, npid, nid);
    // ---- 2. geometry, one quadrature point per lane
    if (lane < NQ)
    {
      int const q = lane;
      // vertex by vertex, so only one vertex's coordinates and mapping
      // gradients are live at a time (register pressure sets the occupancy)
      double J[3][DIM];
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          J[r][c] = 0.;
#pragma unroll 1
      for (int v = 0; v < NPE; ++v)
      {
        double gv[DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c)
          gv[c] = sGeo[(v * DIM + c) * NQ + q];
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
          double const xr = W.x[3 * v + r];
#pragma unroll
          for (int c = 0; c < DIM; ++c)
            J[r][c] += xr * gv[c];
        }
      }
      double A[DIM][DIM];
#pragma unroll
      for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j)
          A[i][j] = J[0][i] * J[0][j] + J[1][i] * J[1][j] + J[2][i] * J[2][j];
      double det, I00 = 0., I01 = 0., I02 = 0., I11 = 0., I12 = 0., I22 = 0.;
      if constexpr (DIM == 2)
      {
        det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        double const r = 1. / det;
        I00 = A[1][1] * r;
        I01 = -A[0][1] * r;
        I11 = A[0][0] * r;
      }
      else
      {
        double const c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
        double const c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
        double const c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
        det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
        double const r = 1. / det;
        I00 = c00 * r;
        I01 = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * r;
        I02 = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * r;
        I11 = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * r;
        I12 = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * r;
        I22 = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * r;
      }
      if (!(det > 0.))
        atomicOr(status, 1);
      double const JxW = sqrt(det) * sW[q];
      double kq = 0.;
      if (needK)
#pragma unroll 3
        for (int i = 0; i < N; ++i)
          kq += sPhiT[i * NQ + q] * W.kap[i];
      double s = 0., m = 0.;
      if (flags & 1)
        s += cS * JxW;
      if (flags & 2)
        s += (kq * JxW) * cSF;
      if (flags & 4)
        m += cM * JxW;
      if (flags & 8)
        m += (cMF * kq) * JxW;
      W.M[q][0][0] = s * I00;
      W.M[q][0][1] = s * I01;
      W.M[q][0][2] = s * I02;
      W.M[q][1][0] = s * I01;
      W.M[q][1][1] = s * I11;
      W.M[q][1][2] = s * I12;
      if constexpr (DIM == 3)
      {
        W.M[q][2][0] = s * I02;
        W.M[q][2][1] = s * I12;
        W.M[q][2][2] = s * I22;
      }
      if constexpr (MASS || QM)
      {
        W.M[q][DIM][0] = m;
        W.M[q][DIM][1] = 0.;
        W.M[q][DIM][2] = 0.;
      }
      if constexpr (QM)
      {
#pragma unroll
        for (int a = DIM + 1; a < 4; ++a)
        {
          W.M[q][a][0] = 0.;
          W.M[q][a][1] = 0.;
          W.M[q][a][2] = 0.;
        }
      }
    }
    waveSync();
    // next element's values and the element after next from the work
    // counter: their latency hides behind the matrix products
    gatherVals(en, npid, nid);
    int const ennReq = grab();
    // ---- 3. Ke tiles on the matrix cores (upper tiles only); the C operand
    //         C[4s + kr][16 tj + jc] is formed in registers per step
    Acc acc[NT][NT];
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
      for (int tj = 0; tj < NT; ++tj)
        acc[ti][tj] = Acc{0, 0, 0, 0};
    if constexpr (QM)
    {
      // q-major (one tile): step q, lane row a = kr.  The lane's dof j = jc
      // gives A[j][4q + a] = g_j(q, a) (phi_j(q) on the mass row) and
      // C[4q + a][j] = row a of the C factor . g_j(q); rows past the mass row
      // are zero in the factor, their A value is any finite table value
      int const jj = jc < N ? jc : N - 1;
      bool const massRow = MASS && kr == DIM;
      auto operands = [&](int q, T & aV, T & bV)
      {
        double const m0 = W.M[q][kr][0], m1 = W.M[q][kr][1], m2 = W.M[q][kr][2];
        double g0 = sGrad[(q * DIM + 0) * N + jj];
        double const g1 = sGrad[(q * DIM + 1) * N + jj];
        double const g2 = DIM == 3 ? sGrad[(q * DIM + (DIM - 1)) * N + jj] : 0.;
        if constexpr (MASS)
          g0 = massRow ? sPhi[q * N + jj] : g0;
        bV = (T)(m0 * g0 + m1 * g1 + m2 * g2);
        aV = (T)(kr == 1 ? g1 : (kr == 2 && DIM == 3) ? g2 : g0);
      };
      T aC, bC;
      operands(0, aC, bC);
#pragma unroll 1
      for (int q = 0; q < NQ; ++q)
      {
        T aN, bN;
        operands(q + 1 < NQ ? q + 1 : q, aN, bN);
        acc[0][0] = Mfma<T>::mma(aC, bC, acc[0][0]);
        aC = aN;
        bC = bN;
      }
    }
    else
    {
      // K ordered by (block of 4 points, component): step (qb, a) takes the
      // rows k = (q, a) for q = 4 qb + kr.  The lane's reference gradients
      // g_{16 t + jc}(q, 0..2) of its two column tiles are at once its A
      // operands (component a) and the input of its C operands
      //   C[(q, a)][16 t + jc] = sum_b M_q[a][b] g_{16 t + jc}(q, b)
      // so per block of points a lane reads 6 gradient values and the point's
      // metric (a broadcast within its 16-lane row group) for 3 (mass: 4)
      // steps of NT (NT + 1) / 2 matrix instructions; padded points (q >=
      // NQ) and columns (j >= N) give zero C values.
      constexpr int QB = (NQ + 3) / 4;
      auto load = [&](int qb, T (&g)[NT][DIM], T (&cv)[NT][DIM], T (&ph)[NT], T (&pm)[NT])
      {
        // padded points (q >= NQ) have zero factors, padded columns (j >=
        // N) zero tables: their C values come out zero
        int const q = 4 * qb + kr;
        double const * m = &W.M[q][0][0];
        double Mq[DIM][DIM];
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
          for (int b = 0; b < DIM; ++b)
            Mq[a][b] = m[3 * a + b];
        double const mm = MASS ? W.M[q][DIM][0] : 0.;
#pragma unroll
        for (int t = 0; t < NT; ++t)
        {
          int const j = 16 * t + jc;
          double const * gp = sGrad + (q * NP + j) * DIM;
          double gd[DIM];
#pragma unroll
          for (int b = 0; b < DIM; ++b)
            gd[b] = gp[b];
#pragma unroll
          for (int a = 0; a < DIM; ++a)
          {
            double c = 0.;
#pragma unroll
            for (int b = 0; b < DIM; ++b)
              c += Mq[a][b] * gd[b];
            g[t][a] = (T)gd[a];
            cv[t][a] = (T)c;
          }
          if constexpr (MASS)
          {
            double const p = sPhi[q * NP + j];
            ph[t] = (T)p;
            pm[t] = (T)(mm * p);
          }
        }
      };
#pragma unroll 1
      for (int qb = 0; qb < QB; ++qb)
      {
        // (other waves of the SIMD cover the LDS latency of the operands)
        T gC[NT][DIM], cC[NT][DIM], pC[NT], qC[NT];
        load(qb, gC, cC, pC, qC);
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
          for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int ti = 0; ti <= tj; ++ti)
              acc[ti][tj] = Mfma<T>::mma(gC[ti][a], cC[tj][a], acc[ti][tj]);
        if constexpr (MASS)
        {
#pragma unroll
          for (int tj = 0; tj < NT; ++tj)
#pragma unroll
            for (int ti = 0; ti <= tj; ++ti)
              acc[ti][tj] = Mfma<T>::mma(pC[ti], qC[tj], acc[ti][tj]);
        }
      }
    }
    // the next element's gathered values and the work counter's reply are
    // taken here, before this element's stores: vector memory operations
    // complete in order, so a wait for them after the stores (at the top of
    // the next element) would also wait for all of the stores
    int64_t const enn = elem(ennReq);
    asm volatile("" ::"v"(rx), "v"(rk), "v"(rb), "v"(rf), "v"(rr) : "memory");
    // ---- 4. the upper tiles go from the accumulators straight to the
    //         row-major store (the off-diagonal one mirrored): tile (ti, tj),
    //         value r of a lane is D[i][j], i = 16 ti + row(r, lane/16),
    //         j = 16 tj + lane%16 (16 consecutive j per row are contiguous).
    //         Dirichlet elimination: rows/columns of constrained dofs become
    //         unit rows; the lifting f = -Ke bcv uses the unmodified Ke, its
    //         lower part through symmetry.
    bool const anyBC = __any(lane < N && W.flag[lane] != 0);
    // row-major rows of NS values in element order, or (PO) of N values at
    // the rows' pair positions
    constexpr size_t ns = PO ? (size_t)N : (size_t)keRowStride(N);
    auto rowOf = [&](int i) -> size_t
    {
      if constexpr (PO)
        return (size_t)W.rp[i];
      else
        return (size_t)(e * N + i);
    };
    if constexpr (sizeof(T) == 4)
    {
      // FP32: entries below 2^-19 sqrt(Ke_ii Ke_jj) are rounding noise of
      // entries that are zero in exact arithmetic (FP64 gives exact zeros or
      // |v| <= 1e-16 there, which close prunes); they are flushed to zero so
      // the FP32 matrix keeps the FP64 sparsity
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          int const i = 16 * t + Mfma<T>::row(r, kr), j = 16 * t + jc;
          if (i == j && i < N)
            W.dg[i] = fabs((double)acc[t][t][r]);
        }
      waveSync();
#pragma unroll
      for (int tj = 0; tj < NT; ++tj)
      {
        int const j = 16 * tj + jc;
        double const dj = W.dg[j < N ? j : N - 1];
#pragma unroll
        for (int ti = 0; ti <= tj; ++ti)
#pragma unroll
          for (int r = 0; r < 4; ++r)
          {
            int const i = 16 * ti + Mfma<T>::row(r, kr);
            double const di = W.dg[i < N ? i : N - 1];
            if (fabs((double)acc[ti][tj][r]) <= 0x1p-19 * sqrt(di * dj))
              acc[ti][tj][r] = T(0);
          }
      }
    }
    double colPart[NT];
#pragma unroll
    for (int tj = 0; tj < NT; ++tj)
    {
      colPart[tj] = 0.;
      int const j = 16 * tj + jc;
      int const jj = j < N ? j : N - 1;
      int const fj = W.flag[jj];
      double const bj = W.bcv[jj];
#pragma unroll
      for (int ti = 0; ti <= tj; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          int const i = 16 * ti + Mfma<T>::row(r, kr);
          bool const valid = i < N && j < N && j >= i;
          T v = acc[ti][tj][r];
          if (anyBC)
          {
            int const ii = i < N ? i : N - 1;
            double const dv = valid ? (double)v : 0.;
            // row i collects Ke_ij bcv_j (j >= i), row j collects Ke_ij bcv_i (j > i)
            double const rp = rowSum16(dv * bj);
            if (jc == 0 && i < N)
              W.f[i] += rp;
            if (j > i)
              colPart[tj] += dv * W.bcv[ii];
            if (W.flag[ii] | fj)
              v = (i == j) ? T(1) : T(0);
          }
          // a diagonal tile holds both triangles (each product computed on
          // its own); the off-diagonal tile is stored again as its transpose
          if (i < N && j < N)
          {
            ke[rowOf(i) * ns + j] = v;
            if (ti != tj)
              ke[rowOf(j) * ns + i] = v;
          }
        }
    }
    if (anyBC)
    {
#pragma unroll
      for (int tj = 0; tj < NT; ++tj)
      {
        double cp = colPart[tj];
        cp += __shfl_xor(cp, 16);
        cp += __shfl_xor(cp, 32);
        int const j = 16 * tj + jc;
        if (kr == 0 && j < N)
          W.f[j] += cp;
      }
      waveSync();
    }
    if (lane < N)
      fe[rowOf(lane)] = anyBC ? (W.flag[lane] ? W.bcv[lane] : 0. - W.f[lane]) : 0.;
    waveSync(); // LDS of this element is reused by the next one
    e = en;
    en = enn;
  }
}

__global__ void k_reset_counters(int * c)
{
  if (threadIdx.x < kCounters)
    c[threadIdx.x * kCounterStride] = 0;
}

template <typename T, int N, int DIM, int NPE, int NQ, bool MASS>
void launchMfma(pxg_ctx * c, int flags, double cS, double cSF, double cM, double cMF, T * ke)
{
  FeTables & tab = c->fe;
  int dev = 0, ncu = 0;
  PXG_HIP(hipGetDevice(&dev));
  PXG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // one resident wave of blocks, each striding over the elements; the
  // pair-order variant exists for the Q2 hexahedron (the rows whose element
  // matrices the rebuild writes in pair order)
  auto go = [&](auto po)
  {
    constexpr bool PO = decltype(po)::value;
    auto * kern = k_element_mfma<T, N, DIM, NPE, NQ, MASS, PO>;
    int perCu = 0;
    PXG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kern, 64 * kWaves, 0));
    int64_t const blocksNeeded = (c->nelem + kWaves - 1) / kWaves;
    int const grid =
        (int)std::max<int64_t>(1, std::min<int64_t>(blocksNeeded, (int64_t)ncu * std::max(perCu, 1)));
    PXG_REQUIRE(c->nelem < ((int64_t)1 << 31) - (int64_t)grid * kWaves * 2, PXG_E_INVALID,
                "element count exceeds the 32-bit work counters");
    PXG_REQUIRE(c->nelem * N < ((int64_t)1 << 32), PXG_E_INVALID, "element rows exceed 32-bit row positions");
    c->workCounter.resize(kCounters * kCounterStride);
    k_reset_counters<<<1, 64, 0, c->stream>>>(c->workCounter.p);
    kern<<<grid, 64 * kWaves, 0, c->stream>>>(c->nelem, c->xyz.p, c->conn.p, c->elemMap.p, c->kappa.p, c->bcFlag.p,
                                              c->bcVal.p, c->hasBC ? 1 : 0, tab.d_geoGrad.p, tab.d_w.p, tab.d_phi.p,
                                              tab.d_grad.p, cS, cSF, cM, cMF, flags, ke, c->feBuf.p, c->status.p,
                                              c->workCounter.p, PO ? c->pairPos.p : nullptr);
  };
  if constexpr (N == 27)
  {
    if (c->kePairOrder)
      go(std::true_type{});
    else
      go(std::false_type{});
  }
  else
  {
    PXG_REQUIRE(!c->kePairOrder, PXG_E_STATE, "pair-order element output is for the hexahedra");
    go(std::false_type{});
  }
  PXG_HIP(hipGetLastError());
}

template <typename T, int N, int DIM, int NPE, int NQ>
void launchMfmaMass(pxg_ctx * c, int flags, double cS, double cSF, double cM, double cMF, T * ke)
{
  if (flags & 12)
    launchMfma<T, N, DIM, NPE, NQ, true>(c, flags, cS, cSF, cM, cMF, ke);
  else
    launchMfma<T, N, DIM, NPE, NQ, false>(c, flags, cS, cSF, cM, cMF, ke);
}

} // namespace

// true when the fused MFMA kernel handles this FE with its recommended
// quadrature size; the caller falls back to the geometry + row kernels otherwise
bool elementMfma(pxg_ctx * c, int flags, double cS, double cSF, double cM, double cMF, int precision)
{
  FeTables const & T = c->fe;
  auto run = [&](auto tag)
  {
    using TT = decltype(tag);
    TT * ke = nullptr;
    size_t const np = (size_t)T.n * keRowStride(T.n); // full row-major element matrix, padded rows
    if constexpr (sizeof(TT) == 8)
    {
      c->keBuf.resize((size_t)c->nelem * np);
      ke = c->keBuf.p;
    }
    else
    {
      c->keBuf32.resize((size_t)c->nelem * np);
      ke = c->keBuf32.p;
    }
    if (T.feType == PXG_FE_HEX_Q2 && T.nq == 27)
      launchMfmaMass<TT, 27, 3, 8, 27>(c, flags, cS, cSF, cM, cMF, ke);
    else if (T.feType == PXG_FE_TET_P2 && T.nq == 11)
      launchMfmaMass<TT, 10, 3, 4, 11>(c, flags, cS, cSF, cM, cMF, ke);
    else if (T.feType == PXG_FE_QUAD_Q2 && T.nq == 9)
      launchMfmaMass<TT, 9, 2, 4, 9>(c, flags, cS, cSF, cM, cMF, ke);
    else
      return false;
    return true;
  };
  if (precision == PXG_FP32)
    return run(float{});
  return run(double{});
}

} // namespace pxg
