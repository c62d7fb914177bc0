// Taylor-Hood Q2-Q1 hexahedron (89 local dofs: three velocity components of
// 27 Q2 functions, then 8 Q1 pressure functions) on the matrix cores.
//
// With the physical gradients G_q[a][c] = sum_b gref_q[a][b] P_q[b][c]
// (P = (J^T J)^-1 J^T, include/curfe.hpp) and s_q = JxW_q, the element
// matrix of Builder::buildLhs / buildCoupling (include/builder.hpp:31-277)
// with AssemblyStiffness, AssemblyTensorStiffness, AssemblyScalarMass,
// AssemblyGrad and AssemblyDiv (include/assembly_lhs.hpp:36-229,
// include/assembly_coupling.hpp:41-133) is
//   velocity (d, a) x velocity (e, j):
//       [d == e] sum_q s_q ((k_s + k_t) G_q[a] . G_q[j] + k_m phi_a phi_j)
//       + k_t sum_q s_q G_q[a][d] G_q[j][e]
//   velocity (d, a) x pressure k:   k_g sum_q s_q G_q[a][d] psi_k
//   pressure k x velocity (e, j):   k_d sum_q s_q psi_k G_q[j][e]
// Three products on v_mfma_f64_16x16x4f64 (FP32: f32) tiles, one wavefront
// per element:
//   S (27 x 27 as 2 x 2 tiles, K = 3 x 28 points + 28 mass rows): the
//     per-component block, staged in LDS;
//   T (K = 28 points): the velocity rows and columns as 6 tiles of 16, tile t
//     holding component t / 2 and functions 16 (t % 2) + 0..15 (the odd
//     tiles 11 of them), so one tile's operands G_q[a][t / 2] use one column
//     of P.  The coupled block in upper tiles only (it is symmetric), and in
//     the same pass the tiles against the pressure tile (the grad block,
//     without its coefficient); the div block is the grad block mirrored
//     with k_d.  Tiles of one component get the S tile added at the store.
// Operand values are formed in registers from the per-point P (LDS, per
// wavefront) and the reference tables (LDS, per block); points padded to 28
// have P = 0 and JxW = 0.  The epilogue applies the per-element Dirichlet
// elimination of the reference Builder (constrained rows become unit rows,
// constrained columns are cleared into the lifting Fe = -Ke h); the lifting
// sums come from the point operands (sum_j Ke_ij h_j = sum_q A_iq (B h)_q for
// the product blocks, rows of S for the per-component block), in a fixed
// order.
#include "kernels.hpp"
#include "wave.hpp"

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <type_traits>
#include <utility>

namespace pxg
{

namespace
{

constexpr int kThWaves = 4;
constexpr int kNV = 27, kNP = 8, kN = 3 * kNV + kNP, kNQ = 27, kNQP = 28, kNPE = 8;
// Tiled velocity index: tile t (0..5) holds component t / 2, functions
// 16 (t % 2) + 0..15 (odd tiles: 11 functions, then padding); tile 6 holds
// the pressure functions 0..7.  Every tile has one component, so a tile's
// operands share one column of P and the per-component block is added in
// whole tiles.
constexpr int kVT = 6, kPT = 6, kHi = kNV - 16;

// local index (0..88) of row 0 of tile t
__host__ __device__ constexpr int thBase(int t) { return t < kVT ? kNV * (t >> 1) + 16 * (t & 1) : 3 * kNV; }

template <typename T>
struct ThMfma;
template <>
struct ThMfma<double>
{
  using Acc = __attribute__((__vector_size__(4 * sizeof(double)))) double;
  static __device__ inline Acc mma(double a, double b, Acc c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int r, int g) { return 4 * r + g; }
};
template <>
struct ThMfma<float>
{
  using Acc = __attribute__((__vector_size__(4 * sizeof(float)))) float;
  static __device__ inline Acc mma(float a, float b, Acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int r, int g) { return 4 * g + r; }
};

__device__ inline void thSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct ThWaveLds
{
  double geo[kNQP][10]; // P by columns (P[b][c] at c * 3 + b), JxW
  double S[kNV][kNV];   // per-component block (reads of the padded rows and
                        // columns 27..31 stay inside this struct and are
                        // never stored)
  double x[kNPE * 3];
  double bcv[kN];
  double f[kN]; // lifting sums
  int flag[kN];
};
static_assert(offsetof(ThWaveLds, S) + sizeof(double) * (31 * kNV + 31) < sizeof(ThWaveLds), "padded S reads");

template <typename T>
__global__ void __launch_bounds__(64 * kThWaves, 2)
    k_element_th_mfma(int64_t E, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
                      uint32_t const * __restrict__ emap, uint8_t const * __restrict__ bcFlag,
                      double const * __restrict__ bcVal, int hasBC, double const * __restrict__ geoGradG,
                      double const * __restrict__ wG, double const * __restrict__ phiG,
                      double const * __restrict__ gradG, double kS, double kT, double kM, double kG, double kD,
                      T * __restrict__ ke, double * __restrict__ fe, int * __restrict__ status)
{
  using M = ThMfma<T>;
  using Acc = typename M::Acc;
  __shared__ double sGeo[kNQ * kNPE * 3];        // mapping gradients [vertex][component][q]
  __shared__ double sGrad[kNQP * (kNV + 1) * 3]; // Q2 reference gradients [q][a][b], a = 27 the zero function
  __shared__ double sPhi[kNQP * (kNV + 1)];      // Q2 values [q][a], a = 27 zero
  __shared__ double sPsi[kNQP * 16];             // Q1 pressure values [q][k] (zero past 8)
  __shared__ double sW[kNQ];
  __shared__ ThWaveLds lds[kThWaves];
  int const tid = threadIdx.x;
  for (int k = tid; k < kNQ * kNPE * 3; k += blockDim.x)
  {
    int const q = k / (kNPE * 3), vc = k - q * (kNPE * 3);
    sGeo[vc * kNQ + q] = geoGradG[k];
  }
  for (int k = tid; k < kNQP * (kNV + 1) * 3; k += blockDim.x)
  {
    int const q = k / ((kNV + 1) * 3), ab = k - q * ((kNV + 1) * 3); // (a, b) of the first component's functions
    sGrad[k] = q < kNQ && ab / 3 < kNV ? gradG[(q * kN + ab / 3) * 3 + ab % 3] : 0.;
  }
  for (int k = tid; k < kNQP * (kNV + 1); k += blockDim.x)
  {
    int const q = k / (kNV + 1), a = k - q * (kNV + 1);
    sPhi[k] = q < kNQ && a < kNV ? phiG[q * kN + a] : 0.;
  }
  for (int k = tid; k < kNQP * 16; k += blockDim.x)
  {
    int const q = k / 16, j = k - q * 16;
    sPsi[k] = q < kNQ && j < kNP ? phiG[q * kN + 3 * kNV + j] : 0.;
  }
  for (int k = tid; k < kNQ; k += blockDim.x)
    sW[k] = wG[k];
  int const wave = tid >> 6, lane = tid & 63;
  ThWaveLds & W = lds[wave];
  if (lane < 10)
    W.geo[kNQP - 1][lane] = 0.; // the padded point
  __syncthreads();
  int const kr = lane >> 4, jc = lane & 15;
  int const aHi = jc < kHi ? 16 + jc : kNV; // the lane's function in odd tiles (the zero function past 26)
  int const lrow = M::row(0, kr);           // the lane's row offset within a tile
  // the element data (vertex coordinates; dof ids, flags and values of the
  // two local dofs lane and lane + 64) is gathered ahead: the ids two
  // elements ahead, the values they point to one element ahead, both issued
  // at the top of an element and waited for before its first store.  Vector
  // memory operations complete in order, so a wait for a load issued after
  // this element's stores (the top of the next element) would also wait for
  // all of them.
  int64_t const stride = (int64_t)gridDim.x * kThWaves;
  struct Pre
  {
    uint32_t pid = 0, id0 = 0, id1 = 0;
    double x = 0., b0 = 0., b1 = 0.;
    int f0 = 0, f1 = 0;
  };
  auto gatherIds = [&](int64_t el, Pre & p)
  {
    if (el < E)
    {
      p.pid = conn[el * kNPE + (lane < kNPE * 3 ? lane / 3 : 0)];
      p.id0 = emap[el * kN + lane];
      p.id1 = emap[el * kN + (lane + 64 < kN ? lane + 64 : lane)];
    }
  };
  auto gatherVals = [&](int64_t el, Pre & p)
  {
    if (el < E)
    {
      p.x = xyz[3ull * p.pid + lane % 3];
      if (hasBC)
      {
        p.f0 = bcFlag[p.id0];
        p.f1 = bcFlag[p.id1];
        p.b0 = bcVal[p.id0];
        p.b1 = bcVal[p.id1];
      }
    }
  };
  // physical gradients of the lane's functions at point q for the first NT
  // tiles: gd[t] = G_q[a][t / 2] with a = 16 (t % 2) + jc
  auto formG = [&](int q, double (&gd)[kVT], auto ntC)
  {
    constexpr int NT = decltype(ntC)::value;
    double const * gl = sGrad + (q * (kNV + 1) + jc) * 3;
    double const * gh = sGrad + (q * (kNV + 1) + aHi) * 3;
    double const * pc = W.geo[q];
#pragma unroll
    for (int t = 0; t < NT; ++t)
    {
      double const * gr = (t & 1) ? gh : gl;
      int const c = t >> 1;
      gd[t] = gr[0] * pc[3 * c] + gr[1] * pc[3 * c + 1] + gr[2] * pc[3 * c + 2];
    }
  };
  int64_t e = (int64_t)blockIdx.x * kThWaves + wave;
  Pre cur, nxt, ids2;
  gatherIds(e, cur);
  gatherVals(e, cur);
  gatherIds(e + stride, nxt);
  // (so that no load is pending at the loop head either)
  asm volatile("" ::"v"(cur.x), "v"(cur.b0), "v"(cur.b1), "v"(cur.f0), "v"(cur.f1), "v"(nxt.pid), "v"(nxt.id0),
               "v"(nxt.id1)
               : "memory");
  for (; e < E; e += stride)
  {
    // ---- 1. element data
    if (lane < kNPE * 3)
      W.x[lane] = cur.x;
    W.flag[lane] = cur.f0;
    W.bcv[lane] = cur.f0 ? cur.b0 : 0.;
    if (lane + 64 < kN)
    {
      W.flag[lane + 64] = cur.f1;
      W.bcv[lane + 64] = cur.f1 ? cur.b1 : 0.;
    }
    bool const anyBC = __any(cur.f0 != 0) || __any(lane + 64 < kN && cur.f1 != 0);
    gatherVals(e + stride, nxt);
    gatherIds(e + 2 * stride, ids2);
    thSync();
    // ---- 2. geometry, one point per lane: P = (J^T J)^-1 J^T, JxW
    if (lane < kNQ)
    {
      int const q = lane;
      double J[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
#pragma unroll
      for (int v = 0; v < kNPE; ++v)
      {
        double gv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
          gv[c] = sGeo[(v * 3 + c) * kNQ + q];
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
          double const xr = W.x[3 * v + r];
#pragma unroll
          for (int c = 0; c < 3; ++c)
            J[r][c] += xr * gv[c];
        }
      }
      double A[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
          A[a][b] = J[0][a] * J[0][b] + J[1][a] * J[1][b] + J[2][a] * J[2][b];
      double const c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
      double const c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
      double const c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
      double const det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
      double I[3][3];
      I[0][0] = c00 / det;
      I[0][1] = I[1][0] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
      I[0][2] = I[2][0] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
      I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
      I[1][2] = I[2][1] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
      I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
      if (!(det > 0.))
        atomicOr(status, 1);
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          W.geo[q][c * 3 + a] = I[a][0] * J[c][0] + I[a][1] * J[c][1] + I[a][2] * J[c][2];
      W.geo[q][9] = sqrt(det) * sW[q];
    }
    thSync();
    // ---- 3. S, the per-component block (functions a x j of one component),
    //         K over (point, component) and the mass rows; in the same loop,
    //         on elements with constrained dofs, the lifting sums of the T,
    //         grad and div blocks: with h the constrained values and
    //         U_q = sum_(e,j) G_q[j][e] h_(e,j), P_q = sum_k JxW_q psi_k h_k,
    //         row (d, a) gets sum_q G_q[a][d] (k_t JxW_q U_q + k_g P_q) and
    //         pressure row k gets k_d sum_q JxW_q psi_k U_q
    double hv[kVT], hp = 0., lift[kVT], liftP = 0.;
#pragma unroll
    for (int t = 0; t < kVT; ++t)
    {
      hv[t] = (!(t & 1) || jc < kHi) ? W.bcv[thBase(t) + jc] : 0.;
      lift[t] = 0.;
    }
    if (jc < kNP)
      hp = W.bcv[3 * kNV + jc];
    {
      Acc s00 = Acc{0, 0, 0, 0}, s01 = Acc{0, 0, 0, 0}, s11 = Acc{0, 0, 0, 0};
#pragma unroll 1
      for (int qb = 0; qb < kNQP / 4; ++qb)
      {
        int const q = 4 * qb + kr;
        double gd[kVT];
        formG(q, gd, std::integral_constant<int, kVT>{});
        double const jw = W.geo[q][9];
        double const cs = (kS + kT) * jw;
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
          T const g0 = (T)gd[2 * c], g1 = (T)gd[2 * c + 1];
          T const h0 = (T)(cs * gd[2 * c]), h1 = (T)(cs * gd[2 * c + 1]);
          s00 = M::mma(g0, h0, s00);
          s01 = M::mma(g0, h1, s01);
          s11 = M::mma(g1, h1, s11);
        }
        if (kM != 0.)
        {
          double const cm = kM * jw;
          double const p0 = sPhi[q * (kNV + 1) + jc], p1 = sPhi[q * (kNV + 1) + aHi];
          s00 = M::mma((T)p0, (T)(cm * p0), s00);
          s01 = M::mma((T)p0, (T)(cm * p1), s01);
          s11 = M::mma((T)p1, (T)(cm * p1), s11);
        }
        if (anyBC)
        {
          double const ps = jw * sPsi[q * 16 + jc];
          double u = 0.;
#pragma unroll
          for (int t = 0; t < kVT; ++t)
            u += gd[t] * hv[t];
          u = rowSum16(u);
          double const pp = rowSum16(ps * hp);
          double const w = kT * jw * u + kG * pp;
#pragma unroll
          for (int t = 0; t < kVT; ++t)
            lift[t] += gd[t] * w;
          liftP += ps * u;
        }
      }
      // stage S (27 x 27, both triangles)
      auto stage = [&](Acc const & s, int ti, int tj)
      {
#pragma unroll
        for (int r = 0; r < 4; ++r)
        {
          int const a = 16 * ti + M::row(r, kr), j = 16 * tj + jc;
          if (a < kNV && j < kNV)
          {
            W.S[a][j] = (double)s[r];
            if (ti != tj)
              W.S[j][a] = (double)s[r];
          }
        }
      };
      stage(s00, 0, 0);
      stage(s01, 0, 1);
      stage(s11, 1, 1);
    }
    if (anyBC)
    {
      // the four point groups' sums, then rows (t, jc) and pressure jc
#pragma unroll
      for (int t = 0; t < kVT; ++t)
      {
        lift[t] += __shfl_xor(lift[t], 16);
        lift[t] += __shfl_xor(lift[t], 32);
      }
      liftP += __shfl_xor(liftP, 16);
      liftP += __shfl_xor(liftP, 32);
      if (kr == 0)
      {
#pragma unroll
        for (int t = 0; t < kVT; ++t)
          if (!(t & 1) || jc < kHi)
            W.f[thBase(t) + jc] = lift[t];
        if (jc < kNP)
          W.f[3 * kNV + jc] = kD * liftP;
      }
    }
    thSync();
    if (anyBC)
    {
      // the per-component block's share: row (d, a) += sum_j S[a][j] h_(d, j)
      for (int i = lane; i < 3 * kNV; i += 64)
      {
        int const d = (i >= kNV) + (i >= 2 * kNV), a = i - d * kNV;
        double sum = 0.;
#pragma unroll 9
        for (int j = 0; j < kNV; ++j)
          sum += W.S[a][j] * W.bcv[d * kNV + j];
        W.f[i] += sum;
      }
    }
    // ---- 4. T: coupled velocity block (upper tiles) and the grad block, in
    //         three passes over column tiles (5 and the pressure tile, 3-4,
    //         0-2) so at most 12 accumulator tiles are live; each pass forms
    //         its operands again, runs its tiles and stores them (5.)
    constexpr size_t ns = (size_t)keRowStride(kN);
    T * const out = ke + (size_t)e * kN * ns;
    int const sLane = lrow * kNV + jc; // the lane's entry of S tiles, less the tile offsets
    auto pass = [&]<int... Cs>(std::integer_sequence<int, Cs...>)
    {
      constexpr int cols[] = {Cs...}; // the pass's column tiles
      constexpr int NC = sizeof...(Cs);
      constexpr int RT = std::max({(Cs < kPT ? Cs + 1 : kVT)...}); // row tiles 0..RT-1
      Acc acc[RT][NC];
#pragma unroll
      for (int ti = 0; ti < RT; ++ti)
#pragma unroll
        for (int u = 0; u < NC; ++u)
          acc[ti][u] = Acc{0, 0, 0, 0};
      // no coefficient tests in here: a zero k_t gives zero operands h, zero
      // k_g and k_d scale the pressure tile at the store
#pragma unroll 1
      for (int qb = 0; qb < kNQP / 4; ++qb)
      {
        int const q = 4 * qb + kr;
        double gd[kVT];
        formG(q, gd, std::integral_constant<int, RT>{});
        double const jw = W.geo[q][9];
        double const ct = kT * jw;
        T g[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t)
          g[t] = (T)gd[t];
        T const ps = (T)(jw * sPsi[q * 16 + jc]); // pressure column jc (zero past 8)
#pragma unroll
        for (int u = 0; u < NC; ++u)
        {
          int const tj = cols[u];
          if (tj < kPT)
          {
            T const h = (T)(ct * gd[tj < RT ? tj : 0]);
#pragma unroll
            for (int ti = 0; ti < RT; ++ti)
              if (ti <= tj)
                acc[ti][u] = M::mma(g[ti], h, acc[ti][u]);
          }
          else
          {
#pragma unroll
            for (int ti = 0; ti < RT; ++ti)
              acc[ti][u] = M::mma(g[ti], ps, acc[ti][u]);
          }
        }
      }
      // ---- 5. values, Dirichlet elimination, stores.  Tile (ti, tj): value
      //         r of the lane at local (thBase(ti) + row(r, kr), thBase(tj) +
      //         jc); tiles of one component get the S tile added; off-diagonal
      //         tiles are stored again mirrored (the pressure column tile
      //         mirrored is the div block).  Addresses are the lane's base
      //         plus constants.
#pragma unroll
      for (int u = 0; u < NC; ++u)
      {
        int const tj = cols[u];
        int const j = thBase(tj) + jc; // local column
        bool const jv = tj < kPT ? (!(tj & 1) || jc < kHi) : jc < kNP;
        int const fj = anyBC ? W.flag[tj < kPT || jv ? j : 0] : 0;
        T * const colBase = out + (size_t)lrow * ns + j;
        T * const mirBase = out + (size_t)j * ns + lrow;
#pragma unroll
        for (int ti = 0; ti < RT; ++ti)
        {
          if (ti > tj)
            continue;
          bool const sameComp = tj < kPT && (ti >> 1) == (tj >> 1);
#pragma unroll
          for (int r = 0; r < 4; ++r)
          {
            int const i0 = thBase(ti) + M::row(r, 0); // the entry's row less the lane offset
            int const rr = M::row(r, kr);
            bool const valid = (!(ti & 1) || rr < kHi) && jv;
            double v = (double)acc[ti][u][r];
            double vm;
            if (tj < kPT)
            {
              if (sameComp)
                v += (&W.S[0][0])[sLane + (16 * (ti & 1) + M::row(r, 0)) * kNV + 16 * (tj & 1)];
              vm = v;
            }
            else
            {
              vm = kD * v; // div block
              v = kG * v;  // grad block
            }
            if (anyBC)
            {
              int const fi = W.flag[i0 + lrow];
              if (fi | fj)
              {
                v = (fi && ti == tj && rr == jc) ? 1. : 0.;
                vm = 0.;
              }
            }
            if (valid)
            {
              colBase[(size_t)i0 * ns] = (T)v;
              if (ti != tj)
                mirBase[i0] = (T)vm;
            }
          }
        }
      }
    };
    // the gathered data is taken here, before this element's first store
    asm volatile("" ::"v"(nxt.x), "v"(nxt.b0), "v"(nxt.b1), "v"(nxt.f0), "v"(nxt.f1), "v"(ids2.pid), "v"(ids2.id0),
                 "v"(ids2.id1)
                 : "memory");
    pass(std::integer_sequence<int, 5, 6>{});
    pass(std::integer_sequence<int, 3, 4>{});
    pass(std::integer_sequence<int, 0, 1, 2>{});
    // the pressure x pressure block: zero, unit diagonal on constrained rows
    for (int k = lane; k < kNP * kNP; k += 64)
    {
      int const i = 3 * kNV + k / kNP, j = 3 * kNV + k % kNP;
      out[(size_t)i * ns + j] = (T)((i == j && W.flag[i]) ? 1. : 0.);
    }
    for (int i = lane; i < kN; i += 64)
      fe[e * kN + i] = W.flag[i] ? W.bcv[i] : (anyBC ? 0. - W.f[i] : 0.);
    thSync(); // the wave's LDS is rewritten by its next element
    cur = nxt;
    nxt.pid = ids2.pid;
    nxt.id0 = ids2.id0;
    nxt.id1 = ids2.id1;
  }
}

} // namespace

bool elementThMfma(pxg_ctx * c, double kS, double kT, double kM, double kG, double kD, int precision)
{
  FeTables const & tab = c->fe;
  if (tab.feType != PXG_FE_HEX_Q2Q1 || tab.nq != kNQ || tab.n != kN)
    return false;
  int dev = 0, ncu = 0;
  PXG_HIP(hipGetDevice(&dev));
  PXG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  size_t const np = (size_t)kN * keRowStride(kN);
  auto run = [&](auto tag)
  {
    using TT = decltype(tag);
    TT * ke;
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
    auto kern = k_element_th_mfma<TT>;
    int perCu = 0;
    PXG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kern, 64 * kThWaves, 0));
    int64_t const need = (c->nelem + kThWaves - 1) / kThWaves;
    // PXG_TH_MFMA_BLOCKS caps the grid (a wavefront then takes many
    // elements: the gathers ahead of the current element are exercised)
    char const * cap = std::getenv("PXG_TH_MFMA_BLOCKS");
    int64_t const most = cap && *cap ? std::max<int64_t>(1, std::atoll(cap)) : (int64_t)ncu * std::max(perCu, 1) * 4;
    unsigned const grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, most));
    kern<<<grid, 64 * kThWaves, 0, c->stream>>>(c->nelem, c->xyz.p, c->conn.p, c->elemMap.p, c->bcFlag.p, c->bcVal.p,
                                                c->hasBC ? 1 : 0, tab.d_geoGrad.p, tab.d_w.p, tab.d_phi.p, tab.d_grad.p,
                                                kS, kT, kM, kG, kD, ke, c->feBuf.p, c->status.p);
    PXG_HIP(hipGetLastError());
  };
  if (precision == PXG_FP32)
    run(float{});
  else
    run(double{});
  return true;
}

} // namespace pxg
