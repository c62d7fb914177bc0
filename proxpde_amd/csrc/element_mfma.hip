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
  // multi-tile elements keep a point's factor as its distinct values: the
  // symmetric metric's upper triangle (00 01 02 11 12 22; 2-D: 00 01 11),
  // then the mass factor when mass terms are built
  static constexpr int MV = (DIM == 3 ? 6 : 3) + (MASS ? 1 : 0);
};

// per-wave LDS (kept small: two blocks of kWaves waves must fit one CU's LDS)
template <typename T, int N, int NPE, int NQ, int MROWS, int MV, int NE>
struct WaveLds
{
  // NE elements per wavefront: element h's values at h NPE 3, h N, h NQ MV
  double x[NE * NPE * 3];
  double kap[NE * N];
  double bcv[NE * N];
  // per-point C operand factor (NQ: points padded to blocks of four, the
  // padding zero).  q-major elements: [NQ][MROWS][3], rows a < DIM the metric
  // row s_q (J^T J)^{-1}[a][:], a = DIM the mass row (m_q, 0, 0), zero rows
  // after it.  Multi-tile elements: [NE][NQ][MV], the distinct values.
  double M[NQ * MROWS * 3 > NE * NQ * MV ? NQ * MROWS * 3 : NE * NQ * MV];
  double f[N];  // lifting accumulator of the Dirichlet elimination
  double dg[sizeof(T) == 4 ? N : 1]; // FP32: diagonal of Ke, the flush threshold of noise entries
  int flag[NE * N];
  uint32_t rp[NE * N]; // output row of each element row (element or pair order)
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
  // two elements per wavefront where each one's nodes, coordinates and points
  // fit in half of it (the geometry of both in one pass), in the FP64 kernels
  // that store rows in pair order: the others are at 128 registers already
  // and would spill
  constexpr int NE = !S::QM && NQ <= 32 && N <= 32 && NPE * 3 <= 32 && PO && sizeof(T) == 8 ? 2 : 1;
  using L = WaveLds<T, N, NPE, S::NQP, S::MROWS, S::MV, NE>;
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
  constexpr int MW = NE == 1 ? S::MROWS * 3 : S::MV; // factor values per point
  for (int k = lane; k < NE * NQP * MW; k += 64)
    if ((k / MW) % NQP >= NQ)
      W.M[k] = 0.;

  __syncthreads();

  int const kr = lane >> 4, jc = lane & 15;
  // gather and geometry: lane 32 h + l takes node, coordinate or point l of
  // the wavefront's element h
  int const half = NE == 2 ? lane >> 5 : 0, hl = NE == 2 ? lane & 31 : lane;
  // waves own disjoint LDS regions and take elements from counters (waves of
  // one SIMD are issued with different priorities, so a fixed share per wave
  // would leave the slowest waves running alone at the end).  Blocks b with
  // the same b % P (P = min(grid, kCounters)) share a counter over one
  // contiguous range of elements (blocks are placed on the XCDs round-robin,
  // so a range stays in one XCD's L2); every range has blocks, wherever they
  // run.  After the
  // tables are loaded waves synchronise only with themselves.  A counter step
  // is NE consecutive elements (the rebuild kernels: two, whose geometry runs
  // in one pass, then their matrices one after the other); the step after
  // next is requested one step ahead, so the atomic's latency hides behind a
  // step's work.
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
  // wave-uniform (scalar registers): the first element of lane 0's counter
  // value (NE elements per counter step)
  auto elem = [&](int v) -> int64_t
  {
    int64_t const k = r0 + (int64_t)NE * __builtin_amdgcn_readfirstlane(v);
    return k < r1 ? k : E;
  };
  int64_t e = elem(grab());
  int64_t en = elem(grab());
  // gathered data of the element being processed, loaded one element ahead
  double rx = 0., rk = 0., rb = 0.;
  int rf = 0;
  uint32_t rr = 0; // output row of element row `hl` (PO: its pair position)
  // element el + half of the wavefront's elements el.. (el + 1 only inside
  // the range)
  auto gatherIds = [&](int64_t el, uint32_t & pid, uint32_t & id)
  {
    int64_t const g = el + half;
    if (el < E && (half == 0 || g < r1))
    {
      if (hl < NPE * 3)
        pid = conn[g * NPE + hl / 3];
      if (hl < N)
      {
        id = emap[g * N + hl];
        if constexpr (PO)
          rr = rowPos[g * N + hl];
      }
    }
  };
  auto gatherVals = [&](int64_t el, uint32_t pid, uint32_t id)
  {
    if (el < E && (half == 0 || el + half < r1))
    {
      if (hl < NPE * 3)
        rx = xyz[3ull * pid + hl % 3];
      if (hl < N)
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
    bool const two = NE == 2 && e + 1 < r1; // (the second element's data are valid)
    if (hl < NPE * 3)
      W.x[half * NPE * 3 + hl] = rx;
    if (hl < N)
    {
      int const k = half * N + hl;
      W.kap[k] = rk;
      W.flag[k] = rf;
      W.bcv[k] = rf ? rb : 0.;
      if constexpr (PO)
        W.rp[k] = rr;
    }
    if (lane < N)
      W.f[lane] = 0.;
    waveSync();
    uint32_t npid = 0, nid = 0;
    gatherIds(en, npid, nid);
    // ---- 2. geometry, one quadrature point per lane (per half wavefront)
    if (hl < NQ)
    {
      int const q = hl;
      double const * const X = W.x + half * NPE * 3;
      double const * const kap = W.kap + half * N;
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
          double const xr = X[3 * v + r];
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
      if (!(det > 0.) && (half == 0 || two))
        atomicOr(status, 1);
      double const JxW = sqrt(det) * sW[q];
      double kq = 0.;
      if (needK)
#pragma unroll 3
        for (int i = 0; i < N; ++i)
          kq += sPhiT[i * NQ + q] * kap[i];
      double s = 0., m = 0.;
      if (flags & 1)
        s += cS * JxW;
      if (flags & 2)
        s += (kq * JxW) * cSF;
      if (flags & 4)
        m += cM * JxW;
      if (flags & 8)
        m += (cMF * kq) * JxW;
      if constexpr (QM)
      {
        double * const mq = W.M + q * 12;
        mq[0] = s * I00;
        mq[1] = s * I01;
        mq[2] = s * I02;
        mq[3] = s * I01;
        mq[4] = s * I11;
        mq[5] = s * I12;
        if constexpr (DIM == 3)
        {
          mq[6] = s * I02;
          mq[7] = s * I12;
          mq[8] = s * I22;
        }
        mq[3 * DIM] = m;
        mq[3 * DIM + 1] = 0.;
        mq[3 * DIM + 2] = 0.;
#pragma unroll
        for (int k = 3 * DIM + 3; k < 12; ++k)
          mq[k] = 0.;
      }
      else if constexpr (NE == 1)
      {
        double * const mq = W.M + q * (S::MROWS * 3);
        mq[0] = s * I00;
        mq[1] = s * I01;
        mq[2] = s * I02;
        mq[3] = s * I01;
        mq[4] = s * I11;
        mq[5] = s * I12;
        if constexpr (DIM == 3)
        {
          mq[6] = s * I02;
          mq[7] = s * I12;
          mq[8] = s * I22;
        }
        if constexpr (MASS)
        {
          mq[3 * DIM] = m;
          mq[3 * DIM + 1] = 0.;
          mq[3 * DIM + 2] = 0.;
        }
      }
      else
      {
        double * const mq = W.M + (half * NQP + q) * S::MV;
        mq[0] = s * I00;
        mq[1] = s * I01;
        if constexpr (DIM == 3)
        {
          mq[2] = s * I02;
          mq[3] = s * I11;
          mq[4] = s * I12;
          mq[5] = s * I22;
        }
        else
          mq[2] = s * I11;
        if constexpr (MASS)
          mq[S::MV - 1] = m;
      }
    }
    waveSync();
    // next element's values and the element after next from the work
    // counter: their latency hides behind the matrix products
    gatherVals(en, npid, nid);
    int const ennReq = grab();
    // ---- 3.-4. per element of the wavefront: element h's tables and rows
    int64_t enn = E;
#pragma unroll 1
    for (int h = 0; h < (two ? 2 : 1); ++h)
    {
      int const fb = h * N; // element h's flags, values and output rows
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
          double const * const mr = W.M + (q * 4 + kr) * 3;
          double const m0 = mr[0], m1 = mr[1], m2 = mr[2];
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
          double Mq[DIM][DIM], mm = 0.;
          if constexpr (NE == 1)
          {
            double const * const m = W.M + q * (S::MROWS * 3);
#pragma unroll
            for (int a = 0; a < DIM; ++a)
#pragma unroll
              for (int b = 0; b < DIM; ++b)
                Mq[a][b] = m[3 * a + b];
            if constexpr (MASS)
              mm = m[3 * DIM];
          }
          else
          {
            double const * const m = W.M + (h * NQP + q) * S::MV;
            Mq[0][0] = m[0];
            Mq[0][1] = Mq[1][0] = m[1];
            if constexpr (DIM == 3)
            {
              Mq[0][2] = Mq[2][0] = m[2];
              Mq[1][1] = m[3];
              Mq[1][2] = Mq[2][1] = m[4];
              Mq[2][2] = m[5];
            }
            else
              Mq[1][1] = m[2];
            if constexpr (MASS)
              mm = m[S::MV - 1];
          }
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
      enn = elem(ennReq);
      asm volatile("" ::"v"(rx), "v"(rk), "v"(rb), "v"(rf), "v"(rr) : "memory");
      // ---- 4. the upper tiles go from the accumulators straight to the
      //         row-major store (the off-diagonal one mirrored): tile (ti, tj),
      //         value r of a lane is D[i][j], i = 16 ti + row(r, lane/16),
      //         j = 16 tj + lane%16 (16 consecutive j per row are contiguous).
      //         Dirichlet elimination: rows/columns of constrained dofs become
      //         unit rows; the lifting f = -Ke bcv uses the unmodified Ke, its
      //         lower part through symmetry.
      bool const anyBC = __any(lane < N && W.flag[fb + lane] != 0);
      // row-major rows of NS values in element order, or (PO) of N values at
      // the rows' pair positions
      constexpr size_t ns = PO ? (size_t)N : (size_t)keRowStride(N);
      auto rowOf = [&](int i) -> size_t
      {
        if constexpr (PO)
          return (size_t)W.rp[fb + i];
        else
          return (size_t)((NE == 2 ? e + h : e) * N + i);
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
        int const fj = W.flag[fb + jj];
        double const bj = W.bcv[fb + jj];
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
                colPart[tj] += dv * W.bcv[fb + ii];
              if (W.flag[fb + ii] | fj)
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
      {
        fe[rowOf(lane)] = anyBC ? (W.flag[fb + lane] ? W.bcv[fb + lane] : 0. - W.f[lane]) : 0.;
        if constexpr (NE == 2)
          W.f[lane] = 0.; // the second element's accumulator
      }
      waveSync(); // LDS of this element is reused by the next one
    }
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
