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
//   S (27 x 27, K = 3 x 28 points + 28 mass rows): the per-component block,
//     staged in LDS and added on the three diagonal blocks at the store;
//   T (velocity rows and columns laid out as 6 tiles of 16, K = 28 points):
//     the coupled block, upper tiles only (it is symmetric), and in the same
//     pass the 6 tiles against the pressure tile (the grad block, without
//     its coefficient); the div block is the grad block mirrored with k_d.
// Operand values are formed in registers from the per-point P (LDS, per
// wavefront) and the reference tables (LDS, per block); points padded to 28
// have P = 0 and JxW = 0.  The epilogue applies the per-element Dirichlet
// elimination of the reference Builder (constrained rows become unit rows,
// constrained columns are cleared into the lifting Fe = -Ke h), with the
// lifting sums in a fixed order.
#include "kernels.hpp"

#include <algorithm>
#include <type_traits>
#include <utility>

namespace pxg
{

namespace
{

constexpr int kThWaves = 4;
constexpr int kNV = 27, kNP = 8, kN = 3 * kNV + kNP, kNQ = 27, kNQP = 28, kNPE = 8;
// tiled index of the T product: velocity 0..80, pressure 96..103
constexpr int kPT = 6; // the pressure tile

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

// local index (0..88) of a tiled index, -1 for padding
__device__ inline int thLocal(int r)
{
  return r < 3 * kNV ? r : (r >= 16 * kPT && r < 16 * kPT + kNP ? 3 * kNV + (r - 16 * kPT) : -1);
}

struct ThWaveLds
{
  double geo[kNQP][10]; // P by columns (P[b][c] at c * 3 + b), JxW
  double S[kNV][kNV];   // per-component block
  double x[kNPE * 3];
  double bcv[kN];
  double f[kN]; // lifting sums
  int flag[kN];
};

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
  __shared__ double sGeo[kNQ * kNPE * 3];     // mapping gradients [vertex][component][q]
  __shared__ double sGrad[kNQP * (kNV + 1) * 3]; // reference gradients of the Q2 functions [q][a][b], a = 27 zero
  __shared__ double sPhi[kNQP * kNV];         // Q2 values [q][a]
  __shared__ double sPsi[kNQP * 16];          // Q1 pressure values [q][k] (k < 8)
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
  for (int k = tid; k < kNQP * kNV; k += blockDim.x)
  {
    int const q = k / kNV, a = k - q * kNV;
    sPhi[k] = q < kNQ ? phiG[q * kN + a] : 0.;
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
  bool const doS = kS + kT != 0. || kM != 0.;
  // the element data (vertex coordinates; dof ids, flags and values of the
  // two local dofs lane and lane + 64) is gathered one element ahead: its
  // dependent loads are in flight while the previous element computes
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
  int64_t e = (int64_t)blockIdx.x * kThWaves + wave;
  Pre cur, nxt;
  gatherIds(e, cur);
  gatherVals(e, cur);
  for (; e < E; e += stride)
  {
    // ---- 1. element data
    if (lane < kNPE * 3)
      W.x[lane] = cur.x;
    W.flag[lane] = cur.f0;
    W.bcv[lane] = cur.f0 ? cur.b0 : 0.;
    W.f[lane] = 0.;
    if (lane + 64 < kN)
    {
      W.flag[lane + 64] = cur.f1;
      W.bcv[lane + 64] = cur.f1 ? cur.b1 : 0.;
      W.f[lane + 64] = 0.;
    }
    thSync();
    gatherIds(e + stride, nxt);
    // ---- 2. geometry, one point per lane: P = (J^T J)^-1 J^T, JxW
    if (lane < kNQ)
    {
      int const q = lane;
      double J[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
#pragma unroll 1
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
    gatherVals(e + stride, nxt);
    // physical gradient G_q[a][c] of Q2 function a at the point of the lane's
    // row group
    auto grad = [&](int q, int a, double const (&P)[3][3], int c) -> double
    {
      double const * g = sGrad + (q * (kNV + 1) + a) * 3;
      return g[0] * P[0][c] + g[1] * P[1][c] + g[2] * P[2][c];
    };
    auto loadP = [&](int q, double (&P)[3][3]) -> double
    {
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          P[a][c] = W.geo[q][c * 3 + a];
      return W.geo[q][9];
    };
    // ---- 3. S: the per-component block (function a x function j), K over
    //         (point, component) and the mass rows
    if (doS)
    {
      Acc s[2][2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
          s[ti][tj] = Acc{0, 0, 0, 0};
#pragma unroll 1
      for (int qb = 0; qb < kNQP / 4; ++qb)
      {
        int const q = 4 * qb + kr;
        double P[3][3];
        double const jw = loadP(q, P);
        double const cs = (kS + kT) * jw, cm = kM * jw;
        T g[2][3], h[2][3], p[2], pm[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
        {
          int const a = 16 * t + jc, aa = a < kNV ? a : kNV - 1;
          bool const v = a < kNV;
#pragma unroll
          for (int c = 0; c < 3; ++c)
          {
            double const gv = v ? grad(q, aa, P, c) : 0.;
            g[t][c] = (T)gv;
            h[t][c] = (T)(cs * gv);
          }
          double const ph = v ? sPhi[q * kNV + aa] : 0.;
          p[t] = (T)ph;
          pm[t] = (T)(cm * ph);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int ti = 0; ti <= tj; ++ti)
              s[ti][tj] = M::mma(g[ti][c], h[tj][c], s[ti][tj]);
        if (kM != 0.)
        {
#pragma unroll
          for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int ti = 0; ti <= tj; ++ti)
              s[ti][tj] = M::mma(p[ti], pm[tj], s[ti][tj]);
        }
      }
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti <= tj; ++ti)
#pragma unroll
          for (int r = 0; r < 4; ++r)
          {
            int const a = 16 * ti + M::row(r, kr), j = 16 * tj + jc;
            if (a < kNV && j < kNV)
            {
              W.S[a][j] = (double)s[ti][tj][r];
              if (ti != tj)
                W.S[j][a] = (double)s[ti][tj][r];
            }
          }
    }
    else
    {
      for (int k = lane; k < kNV * kNV; k += 64)
        (&W.S[0][0])[k] = 0.;
    }
    thSync();
    // ---- 4. T: coupled velocity block (upper tiles) and the grad block, in
    //         three passes over column tiles (5 and the pressure tile, 3-4,
    //         0-2) so at most 12 accumulator tiles are live; each pass forms
    //         its operands again, runs its tiles and stores them (5.)
    bool const anyBC = __any(lane < kN && W.flag[lane] != 0) || __any(lane + 64 < kN && W.flag[lane + 64] != 0);
    constexpr size_t ns = (size_t)keRowStride(kN);
    T * const out = ke + (size_t)e * kN * ns;
    // per-lane operand rows of the six velocity tiles: function a and
    // component d of tiled row 16 t + jc (padding: the zero function)
    int gOff[6], pOff[6];
#pragma unroll
    for (int t = 0; t < 6; ++t)
    {
      int const r = 16 * t + jc;
      bool const v = r < 3 * kNV;
      int const d = v ? r / kNV : 0;
      gOff[t] = (v ? r - d * kNV : kNV) * 3;
      pOff[t] = d * 3;
    }
    int const lrow = M::row(0, kr); // a lane's row offset within a tile
    auto pass = [&]<int... Cs>(std::integer_sequence<int, Cs...>)
    {
      constexpr int cols[] = {Cs...}; // the pass's column tiles
      constexpr int NC = sizeof...(Cs);
      constexpr int RT = std::max({(Cs < kPT ? Cs + 1 : kPT)...}); // row tiles 0..RT-1
      Acc acc[RT][NC];
#pragma unroll
      for (int ti = 0; ti < RT; ++ti)
#pragma unroll
        for (int u = 0; u < NC; ++u)
          acc[ti][u] = Acc{0, 0, 0, 0};
      {
#pragma unroll 1
        for (int qb = 0; qb < kNQP / 4; ++qb)
        {
          int const q = 4 * qb + kr;
          double const jw = W.geo[q][9];
          double const ct = kT * jw;
          double const * gq = sGrad + q * ((kNV + 1) * 3);
          T g[RT], h[RT];
#pragma unroll
          for (int t = 0; t < RT; ++t)
          {
            double const * gr = gq + gOff[t];
            double const * pc = &W.geo[q][pOff[t]];
            double const gv = gr[0] * pc[0] + gr[1] * pc[1] + gr[2] * pc[2];
            g[t] = (T)gv;
            h[t] = (T)(ct * gv);
          }
          T const ps = (T)(jw * sPsi[q * 16 + jc]); // pressure column jc (zero past 8)
          // no coefficient tests in here: a branch around the products makes
          // the compiler move the accumulators out of and back into the
          // matrix registers on every point block (a zero k_t gives zero
          // operands h, zero k_g and k_d scale the pressure tile at the store)
#pragma unroll
          for (int u = 0; u < NC; ++u)
          {
            int const tj = cols[u];
            if (tj < kPT)
            {
#pragma unroll
              for (int ti = 0; ti < RT; ++ti)
                if (ti <= tj)
                  acc[ti][u] = M::mma(g[ti], h[tj < RT ? tj : 0], acc[ti][u]);
            }
            else
            {
#pragma unroll
              for (int ti = 0; ti < RT; ++ti)
                acc[ti][u] = M::mma(g[ti], ps, acc[ti][u]);
            }
          }
        }
      }
      // ---- 5. values, Dirichlet elimination, lifting, stores.  Tile (ti, tj):
      //         value r of the lane at tiled (16 ti + row(r, kr), 16 tj + jc);
      //         off-diagonal tiles are stored again mirrored (the pressure
      //         column tile mirrored is the div block).  Rows of the velocity
      //         tiles are local rows; the index arithmetic per entry is a
      //         constant plus the lane's offsets.
#pragma unroll
      for (int u = 0; u < NC; ++u)
      {
        int const tj = cols[u];
        int const j = tj < kPT ? 16 * tj + jc : 3 * kNV + jc; // local column
        bool const jv = tj < kPT ? j < 3 * kNV : jc < kNP;
        int const jj = jv ? j : 0;
        int const dj = jj / kNV, aj = jj - dj * kNV; // (velocity columns)
        int const fj = jv ? W.flag[jj] : 0;
        double const bj = jv ? W.bcv[jj] : 0.;
        T * const colBase = out + (size_t)lrow * ns + jj;    // + row(r, 0) + 16 ti rows
        T * const mirBase = out + (size_t)jj * ns + lrow;    // + row(r, 0) + 16 ti columns
        double colPart = 0.;
#pragma unroll
        for (int ti = 0; ti < RT; ++ti)
        {
          if (ti > tj)
            continue;
          // does the tile meet a diagonal (component) block of S?
          bool const nearDiag = tj < kPT && (16 * tj) / kNV <= (16 * ti + 15) / kNV && (16 * ti) / kNV <= (16 * tj + 15) / kNV;
#pragma unroll
          for (int r = 0; r < 4; ++r)
          {
            int const i0 = 16 * ti + M::row(r, 0); // the entry's row less the lane offset
            int const i = i0 + lrow;
            bool const iv = i0 + M::row(0, 3) < 3 * kNV || i < 3 * kNV;
            bool const valid = iv && jv;
            double v = (double)acc[ti][u][r];
            double vm = v; // the mirrored entry (j, i)
            if (tj < kPT)
            {
              if (nearDiag)
              {
                int const di = (i >= kNV) + (i >= 2 * kNV);
                if (valid && di == dj)
                  v += W.S[i - di * kNV][aj];
                vm = v;
              }
            }
            else
            {
              vm = kD * v; // div block
              v = kG * v;  // grad block
            }
            if (anyBC)
            {
              int const ii = iv ? i : 0;
              v = valid ? v : 0.;
              vm = valid ? vm : 0.;
              // row i collects Ke_ij bcv_j, row j the mirrored Ke_ji bcv_i
              double rp = v * bj;
#pragma unroll
              for (int off = 8; off > 0; off >>= 1)
                rp += __shfl_xor(rp, off);
              if (jc == 0 && iv && !W.flag[ii])
                W.f[ii] += rp;
              if (ti != tj)
                colPart += vm * W.bcv[ii];
              int const fi = W.flag[ii];
              if (fi | fj)
              {
                v = (fi && i == j) ? 1. : 0.;
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
        if (anyBC)
        {
          double cp = colPart;
          cp += __shfl_xor(cp, 16);
          cp += __shfl_xor(cp, 32);
          if (kr == 0 && jv && !fj)
            W.f[jj] += cp;
        }
      }
    };
    pass(std::integer_sequence<int, 5, 6>{});
    pass(std::integer_sequence<int, 3, 4>{});
    pass(std::integer_sequence<int, 0, 1, 2>{});
    // the pressure x pressure block: zero, unit diagonal on constrained rows
    for (int k = lane; k < kNP * kNP; k += 64)
    {
      int const i = 3 * kNV + k / kNP, j = 3 * kNV + k % kNP;
      out[(size_t)i * ns + j] = (T)((i == j && W.flag[i]) ? 1. : 0.);
    }
    thSync();
    for (int i = lane; i < kN; i += 64)
      fe[e * kN + i] = W.flag[i] ? W.bcv[i] : (anyBC ? 0. - W.f[i] : 0.);
    thSync(); // the wave's LDS is rewritten by its next element
    cur = nxt;
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
    unsigned const grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ncu * std::max(perCu, 1) * 4));
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
