// Sum-factorised matrix-free kappa-stiffness operator for the tensor-product
// hexahedra: Q1 with the 2x2x2 and Q2 with the 3x3x3 Gauss rule (the
// reference's RecommendedQR, include/fe.hpp).  Same bilinear form as
// matfree.hip (AssemblyStiffnessFE, include/assembly_lhs.hpp:93-109, with the
// Dirichlet form C K C + D_c of Builder::buildLhs), same element buffer and
// per-dof gather, but the contractions run one direction at a time:
//
//   basis phi_(i,j,k) = l_i(xi) l_j(eta) l_k(zeta), points q = (qx, qy, qz),
//   B[q][t] = l_t(g_q), D[q][t] = l_t'(g_q) (P x P, P = 2 or 3)
//   reference gradient of u at the points:
//     u_xi = (D x B x B) u, u_eta = (B x D x B) u, u_zeta = (B x B x D) u
//   v_q = M_q grad u_q,  M_q = kappa_q w_q |det J_q| J_q^-1 J_q^-T
//   y = (D^T x B^T x B^T) v_xi + (B^T x D^T x B^T) v_eta + (B^T x B^T x D^T) v_zeta
//
// 6 P^4 multiply-adds per element and vector against 2 P^6 x 3 for the
// dense contraction (Q2: 1,458 against 4,374, plus the metric).
//
// Lanes: P^2 lanes per element hold one pencil each (P values along one
// direction); a contraction along the pencil's direction runs in registers,
// the others after a transpose through the element's LDS region (the lanes
// of an element are in one wavefront: a wave barrier orders the LDS accesses).
// Pencil roles: stage x (j, k), stage y (qx, k), stage z (qx, qy).  Q2 packs
// 7 elements (63 lanes) per wavefront, Q1 16.
//
// Geometry: the trilinear vertex map, which is the reference's isoparametric
// map (the Q2 mapping points are vertex averages, include/reffe.hpp:949-980);
// computed per point by the lane that owns the point's z pencil and kept in
// registers across the batch.
#include "kernels.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <unordered_map>
#include <vector>

namespace pxg
{

namespace
{
constexpr int kSfWaves = 4;

__device__ inline void sfSync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1D tables of the tensor-product element (lexicographic nodes t = 0..P-1
// from xi = -1 to +1), Gauss points and weights, and the reference local
// index of every lexicographic node (i + P (j + P k))
template <int P>
struct SfTab
{
  double B[P][P];
  double D[P][P];
  float Bf[P][P]; // the same in FP32 for the FP32 contractions
  float Df[P][P];
  double g[P];
  double w[P];
  int local[P * P * P];
};

// table entries in the precision of the contraction
template <typename T, int P>
__device__ inline T tB(SfTab<P> const & t, int q, int i)
{
  if constexpr (sizeof(T) == 4)
    return t.Bf[q][i];
  else
    return t.B[q][i];
}
template <typename T, int P>
__device__ inline T tD(SfTab<P> const & t, int q, int i)
{
  if constexpr (sizeof(T) == 4)
    return t.Df[q][i];
  else
    return t.D[q][i];
}

// LDS of one element: two regions of 3 arrays of P^3 values (stage buffers;
// before the batch loop region 1 holds the vertex coordinates)
template <int P>
struct SfLds
{
  double r[2][3][P * P * P];
};

enum : int
{
  kSfApply = 0, // y_e = K_e x (element buffer [E][n][nb])
  kSfGrad = 1,  // g_e = d/dkappa (ybar . K x) per node (element buffer [E][n])
  kSfDiag = 2,  // diagonal of K_e (element buffer [E][n])
  kSfMetric = 3, // per-point metrics M_q of every element ([E][6][n], see k_sf_apply_pa)
};

// metric M_q = kappa_q w_q |det J_q| J_q^-1 J_q^-T of point (qx, qy, qz)
// from the vertex coordinates X (trilinear map); false for a degenerate J
__device__ inline bool sfPointMetric(double const * X, double xi, double eta, double zeta, double wq, double kq,
                                     double (&M)[6])
{
  constexpr int vid[2][2][2] = {{{0, 1}, {3, 2}}, {{4, 5}, {7, 6}}}; // [vz][vy][vx]
  double const lx[2] = {0.5 * (1. - xi), 0.5 * (1. + xi)};
  double const ly[2] = {0.5 * (1. - eta), 0.5 * (1. + eta)};
  double const lz[2] = {0.5 * (1. - zeta), 0.5 * (1. + zeta)};
  double J[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    double d0 = 0., d1 = 0., d2 = 0.;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
      {
        d0 += ly[a] * lz[b] * (X[3 * vid[b][a][1] + r] - X[3 * vid[b][a][0] + r]);
        d1 += lx[a] * lz[b] * (X[3 * vid[b][1][a] + r] - X[3 * vid[b][0][a] + r]);
        d2 += lx[a] * ly[b] * (X[3 * vid[1][b][a] + r] - X[3 * vid[0][b][a] + r]);
      }
    J[r][0] = 0.5 * d0;
    J[r][1] = 0.5 * d1;
    J[r][2] = 0.5 * d2;
  }
  double const c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double const c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double const c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  double const det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  double const id_ = 1. / det;
  double Ji[3][3]; // J^-1[a][r] = cofactor(r, a) / det
  Ji[0][0] = c00 * id_;
  Ji[1][0] = c01 * id_;
  Ji[2][0] = c02 * id_;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id_;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id_;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id_;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id_;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id_;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id_;
  double const s = kq * wq * fabs(det);
  int m = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = a; b < 3; ++b)
      M[m++] = s * (Ji[a][0] * Ji[b][0] + Ji[a][1] * Ji[b][1] + Ji[a][2] * Ji[b][2]);
  return det != 0. && std::isfinite(det);
}

// Data of one element gathered by a lane of its x pencil (j, k): dof ids,
// constraint flags and coefficient values of the pencil's P nodes, and the
// vertex coordinates this lane stages (lanes take the 24 values in turn)
template <int P>
struct SfGather
{
  static constexpr int CPL = (24 + P * P - 1) / (P * P); // coordinates per lane
  uint32_t id[P];
  uint32_t pid[CPL];
  bool fl[P];
  double kn[P];
  double xc[CPL];
};

// Persistent element kernel: wave w takes the batches w, w + W, ... of EPW
// consecutive elements (W = waves in the grid) and loads the dof ids of the
// next batch, then its coordinates, coefficients and flags (nb = 1: also its
// vector values), while the current batch computes.  NBC vectors of the
// batch go through the stages one after the other, with their values loaded
// together and their results written together per node ([E][n][nb], nb
// contiguous values).
template <typename T, int P, int MODE, int NBC>
__global__ void __launch_bounds__(64 * kSfWaves)
    k_sf_elem(int64_t E, int nb, SfTab<P> tab, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
              uint32_t const * __restrict__ emap, double const * __restrict__ kappa,
              uint8_t const * __restrict__ cons, T const * __restrict__ x, T const * __restrict__ ybar,
              T * __restrict__ out, int * __restrict__ status)
{
  constexpr int P2 = P * P, N = P * P * P, EPW = 64 / P2;
  using G = SfGather<P>;
  constexpr int CPL = G::CPL;
  // a lane past the last whole element of the wave (Q2: lane 63) works on a
  // region of its own and stores nothing
  constexpr int SLOTS = (64 + P2 - 1) / P2;
  __shared__ SfLds<P> lds[kSfWaves * SLOTS];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int const slot = lane / P2, pl = lane - slot * P2;
  int const pa = pl % P, pc = pl / P; // the two pencil indices
  SfLds<P> & L = lds[wave * SLOTS + slot];
  auto at = [](int i, int j, int k) { return i + P * (j + P * k); };
  int loc[P]; // reference local index of the lane's x-pencil nodes
#pragma unroll
  for (int i = 0; i < P; ++i)
    loc[i] = tab.local[at(i, pa, pc)];
  int64_t const nbatch = (E + EPW - 1) / EPW;
  int64_t const stride = (int64_t)gridDim.x * kSfWaves;
  int64_t bt = (int64_t)blockIdx.x * kSfWaves + wave;

  auto elemOf = [&](int64_t b) -> int64_t { return b * EPW + slot; };
  auto liveOf = [&](int64_t b) -> bool { return b < nbatch && slot < EPW && elemOf(b) < E; };
  // all loads are unconditional, from indices clamped into the arrays (a
  // lane without an element reads the last one and stores nothing), so the
  // loads of a batch are in flight together with no branch around them
  auto loadIds = [&](G & g, int64_t b)
  {
    int64_t const e = std::min(elemOf(b), E - 1);
#pragma unroll
    for (int i = 0; i < P; ++i)
      g.id[i] = emap[e * N + loc[i]];
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      g.pid[c] = conn[e * 8 + std::min(pl + c * P2, 23) / 3];
  };
  auto loadVals = [&](G & g, int64_t b)
  {
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      g.xc[c] = xyz[3ull * g.pid[c] + std::min(pl + c * P2, 23) % 3];
#pragma unroll
    for (int i = 0; i < P; ++i)
    {
      g.fl[i] = cons ? cons[g.id[i]] != 0 : false;
      if constexpr (MODE != kSfGrad)
        g.kn[i] = kappa[g.id[i]];
    }
  };
  // vector values of a chunk of NBC vectors starting at b0 (zero past nb)
  auto loadVec = [&](T const * v, G const & g, int b0, T (&out)[P][NBC])
  {
#pragma unroll
    for (int i = 0; i < P; ++i)
#pragma unroll
      for (int c = 0; c < NBC; ++c)
      {
        T const t = v[(size_t)g.id[i] * nb + std::min(b0 + c, nb - 1)];
        out[i][c] = b0 + c < nb ? t : T(0);
      }
  };

  G cur, nxt;
  loadIds(cur, bt);
  loadVals(cur, bt);
  T xcur[P][NBC]; // nb = 1: the vector values travel with the prefetch
  if constexpr (NBC == 1 && MODE == kSfApply)
    loadVec(x, cur, 0, xcur);

  while (bt < nbatch)
  {
    bool const live = liveOf(bt);
    int64_t const e = elemOf(bt);
    int64_t const bn = bt + stride;
    loadIds(nxt, bn);
    // vertex coordinates to region 1
#pragma unroll
    for (int c = 0; c < CPL; ++c)
    {
      int const v = pl + c * P2;
      if (v < 24)
        L.r[1][0][v] = cur.xc[c];
    }

    // ---- kappa at the points (apply, diag): interpolation B x B x B ----
    double kq[P];
#pragma unroll
    for (int q = 0; q < P; ++q)
      kq[q] = 1.;
    if constexpr (MODE != kSfGrad)
    {
#pragma unroll
      for (int q = 0; q < P; ++q) // x: lane (j, k)
      {
        double s = 0.;
#pragma unroll
        for (int i = 0; i < P; ++i)
          s += tab.B[q][i] * cur.kn[i];
        L.r[0][0][at(q, pa, pc)] = s;
      }
      sfSync();
#pragma unroll
      for (int q = 0; q < P; ++q) // y: lane (qx, k)
      {
        double s = 0.;
#pragma unroll
        for (int j = 0; j < P; ++j)
          s += tab.B[q][j] * L.r[0][0][at(pa, j, pc)];
        L.r[0][1][at(pa, q, pc)] = s;
      }
      sfSync();
#pragma unroll
      for (int q = 0; q < P; ++q) // z: lane (qx, qy)
      {
        double s = 0.;
#pragma unroll
        for (int k = 0; k < P; ++k)
          s += tab.B[q][k] * L.r[0][1][at(pa, pc, k)];
        kq[q] = s;
      }
    }
    sfSync();

    // ---- metric of the lane's z pencil: points (qx, qy, qz) = (pa, pc, qz) ----
    // M = s (J^T J)^-1 = s J^-1 J^-T, s = kappa_q w_q |det J| (grad: without kappa)
    double M[P][6];
#pragma unroll
    for (int q = 0; q < P; ++q)
      if (!sfPointMetric(L.r[1][0], tab.g[pa], tab.g[pc], tab.g[q], tab.w[pa] * tab.w[pc] * tab.w[q], kq[q], M[q]) &&
          live)
        atomicOr(status, 1);
    sfSync(); // region 1 (coordinates) is a stage buffer from here on
    // the next batch's coordinates, coefficients and flags are in flight
    // during the contractions (nb = 1: its vector values too)
    loadVals(nxt, bn);
    T xnxt[P][NBC];
    if constexpr (NBC == 1 && MODE == kSfApply)
      loadVec(x, nxt, 0, xnxt);

    T * A0 = reinterpret_cast<T *>(L.r[0][0]); // stage buffers: 3 arrays of N values each
    T * A1 = A0 + N;
    T * A2 = A1 + N;
    T * B0 = reinterpret_cast<T *>(L.r[1][0]);
    T * B1 = B0 + N;
    T * B2 = B1 + N;

    // forward stages of one vector (node values u of the lane's x pencil)
    // to the reference gradient at the lane's points (g[q][0..2], q = qz)
    auto forward = [&](T const (&u)[P], T (&g)[P][3])
    {
#pragma unroll
      for (int q = 0; q < P; ++q) // x: lane (j, k)
      {
        T sb = T(0), sd = T(0);
#pragma unroll
        for (int i = 0; i < P; ++i)
        {
          sb += tB<T>(tab, q, i) * u[i];
          sd += tD<T>(tab, q, i) * u[i];
        }
        A0[at(q, pa, pc)] = sb;
        A1[at(q, pa, pc)] = sd;
      }
      sfSync();
#pragma unroll
      for (int q = 0; q < P; ++q) // y: lane (qx, k)
      {
        T bb = T(0), bd = T(0), db = T(0);
#pragma unroll
        for (int j = 0; j < P; ++j)
        {
          T const ub = A0[at(pa, j, pc)], ud = A1[at(pa, j, pc)];
          bb += tB<T>(tab, q, j) * ub;
          bd += tD<T>(tab, q, j) * ub;
          db += tB<T>(tab, q, j) * ud;
        }
        B0[at(pa, q, pc)] = bb;
        B1[at(pa, q, pc)] = bd;
        B2[at(pa, q, pc)] = db;
      }
      sfSync();
#pragma unroll
      for (int q = 0; q < P; ++q) // z: lane (qx, qy)
      {
        T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
        for (int k = 0; k < P; ++k)
        {
          gx += tB<T>(tab, q, k) * B2[at(pa, pc, k)];
          gy += tB<T>(tab, q, k) * B1[at(pa, pc, k)];
          gz += tD<T>(tab, q, k) * B0[at(pa, pc, k)];
        }
        g[q][0] = gx;
        g[q][1] = gy;
        g[q][2] = gz;
      }
    };

    if constexpr (MODE == kSfApply)
    {
      T Mt[P][6];
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int m = 0; m < 6; ++m)
          Mt[q][m] = T(M[q][m]);
      for (int b0 = 0; b0 < nb; b0 += NBC)
      {
        T xv[P][NBC], yv[P][NBC];
        if constexpr (NBC == 1)
        {
#pragma unroll
          for (int i = 0; i < P; ++i)
            xv[i][0] = xcur[i][0];
        }
        else
          loadVec(x, cur, b0, xv);
#pragma unroll
        for (int c = 0; c < NBC; ++c)
        {
          T u[P];
#pragma unroll
          for (int i = 0; i < P; ++i)
            u[i] = cur.fl[i] ? T(0) : xv[i][c];
          T g[P][3];
          forward(u, g); // ends after reading B0..B2; the A buffers are free
          T v[P][3];
#pragma unroll
          for (int q = 0; q < P; ++q)
          {
            v[q][0] = Mt[q][0] * g[q][0] + Mt[q][1] * g[q][1] + Mt[q][2] * g[q][2];
            v[q][1] = Mt[q][1] * g[q][0] + Mt[q][3] * g[q][1] + Mt[q][4] * g[q][2];
            v[q][2] = Mt[q][2] * g[q][0] + Mt[q][4] * g[q][1] + Mt[q][5] * g[q][2];
          }
#pragma unroll
          for (int k = 0; k < P; ++k) // z^T: lane (qx, qy)
          {
            T s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll
            for (int q = 0; q < P; ++q)
            {
              s1 += tB<T>(tab, q, k) * v[q][0];
              s2 += tB<T>(tab, q, k) * v[q][1];
              s3 += tD<T>(tab, q, k) * v[q][2];
            }
            A0[at(pa, pc, k)] = s1;
            A1[at(pa, pc, k)] = s2;
            A2[at(pa, pc, k)] = s3;
          }
          sfSync();
#pragma unroll
          for (int j = 0; j < P; ++j) // y^T: lane (qx, k)
          {
            T sd = T(0), sb = T(0);
#pragma unroll
            for (int q = 0; q < P; ++q)
            {
              sd += tB<T>(tab, q, j) * A0[at(pa, q, pc)];
              sb += tD<T>(tab, q, j) * A1[at(pa, q, pc)] + tB<T>(tab, q, j) * A2[at(pa, q, pc)];
            }
            B0[at(pa, j, pc)] = sd;
            B1[at(pa, j, pc)] = sb;
          }
          sfSync();
#pragma unroll
          for (int i = 0; i < P; ++i) // x^T: lane (j, k)
          {
            T y = T(0);
#pragma unroll
            for (int q = 0; q < P; ++q)
              y += tD<T>(tab, q, i) * B0[at(q, pa, pc)] + tB<T>(tab, q, i) * B1[at(q, pa, pc)];
            // unit diagonal per element on constrained rows
            yv[i][c] = cur.fl[i] ? xv[i][c] : y;
          }
          sfSync(); // the B buffers are rewritten by the next vector
        }
        if (live)
        {
#pragma unroll
          for (int i = 0; i < P; ++i)
          {
            T * o = out + ((size_t)e * N + loc[i]) * nb + b0;
            if (nb == NBC)
            {
#pragma unroll
              for (int c = 0; c < NBC; ++c)
                o[c] = yv[i][c];
            }
            else
            {
#pragma unroll
              for (int c = 0; c < NBC; ++c)
                if (b0 + c < nb)
                  o[c] = yv[i][c];
            }
          }
        }
      }
    }
    else if constexpr (MODE == kSfGrad)
    {
      // c_q = sum_b grad x_b . M grad ybar_b (M without kappa), then the
      // transposed interpolation B^T x B^T x B^T of c
      T cq[P];
#pragma unroll
      for (int q = 0; q < P; ++q)
        cq[q] = T(0);
      for (int b0 = 0; b0 < nb; b0 += NBC)
      {
        T xv[P][NBC], wv[P][NBC];
        loadVec(x, cur, b0, xv);
        loadVec(ybar, cur, b0, wv);
#pragma unroll
        for (int c = 0; c < NBC; ++c)
        {
          T u[P], w[P];
#pragma unroll
          for (int i = 0; i < P; ++i)
          {
            u[i] = cur.fl[i] ? T(0) : xv[i][c];
            w[i] = cur.fl[i] ? T(0) : wv[i][c];
          }
          T gu[P][3], gw[P][3];
          forward(u, gu);
          sfSync();
          forward(w, gw);
          sfSync();
#pragma unroll
          for (int q = 0; q < P; ++q)
          {
            T const t0 = T(M[q][0]) * gw[q][0] + T(M[q][1]) * gw[q][1] + T(M[q][2]) * gw[q][2];
            T const t1 = T(M[q][1]) * gw[q][0] + T(M[q][3]) * gw[q][1] + T(M[q][4]) * gw[q][2];
            T const t2 = T(M[q][2]) * gw[q][0] + T(M[q][4]) * gw[q][1] + T(M[q][5]) * gw[q][2];
            cq[q] += gu[q][0] * t0 + gu[q][1] * t1 + gu[q][2] * t2;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < P; ++k) // z^T: lane (qx, qy)
      {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < P; ++q)
          s += tB<T>(tab, q, k) * cq[q];
        A0[at(pa, pc, k)] = s;
      }
      sfSync();
#pragma unroll
      for (int j = 0; j < P; ++j) // y^T: lane (qx, k)
      {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < P; ++q)
          s += tB<T>(tab, q, j) * A0[at(pa, q, pc)];
        B0[at(pa, j, pc)] = s;
      }
      sfSync();
#pragma unroll
      for (int i = 0; i < P; ++i) // x^T: lane (j, k)
      {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < P; ++q)
          s += tB<T>(tab, q, i) * B0[at(q, pa, pc)];
        if (live)
          out[(size_t)e * N + loc[i]] = s;
      }
      sfSync();
    }
    else if constexpr (MODE == kSfMetric)
    {
      if (live)
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
          for (int m = 0; m < 6; ++m)
            out[((size_t)e * 6 + m) * N + at(pa, pc, q)] = T(M[q][m]);
    }
    else // kSfDiag: (K_e)_nn = sum_q grad phi_n(q) . M_q grad phi_n(q)
    {
      // all points' metrics to LDS (6 arrays of N values over both regions)
      double * Mall = L.r[0][0];
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int m = 0; m < 6; ++m)
          Mall[m * N + at(pa, pc, q)] = M[q][m];
      sfSync();
#pragma unroll 1
      for (int i = 0; i < P; ++i) // node (i, j, k) = (i, pa, pc)
      {
        double s = 0.;
#pragma unroll 1
        for (int qz = 0; qz < P; ++qz)
#pragma unroll 1
          for (int qy = 0; qy < P; ++qy)
#pragma unroll
            for (int qx = 0; qx < P; ++qx)
            {
              double const gx = tab.D[qx][i] * tab.B[qy][pa] * tab.B[qz][pc];
              double const gy = tab.B[qx][i] * tab.D[qy][pa] * tab.B[qz][pc];
              double const gz = tab.B[qx][i] * tab.B[qy][pa] * tab.D[qz][pc];
              int const q = at(qx, qy, qz);
              s += Mall[q] * gx * gx + Mall[3 * N + q] * gy * gy + Mall[5 * N + q] * gz * gz +
                   2. * (Mall[N + q] * gx * gy + Mall[2 * N + q] * gx * gz + Mall[4 * N + q] * gy * gz);
            }
        if (cur.fl[i])
          s = 1.; // unit diagonal per element on constrained rows
        if (live)
          out[(size_t)e * N + loc[i]] = T(s);
      }
      sfSync();
    }
    cur = nxt;
    if constexpr (NBC == 1 && MODE == kSfApply)
    {
#pragma unroll
      for (int i = 0; i < P; ++i)
        xcur[i][0] = xnxt[i][0];
    }
    bt = bn;
  }
}

// 1D tables of the tensor-product element as compile-time functions (the
// batch kernel's contractions unroll over constant indices, so zero and unit
// entries fold away): Lagrange basis at the nodes -1 (, 0), +1 evaluated at
// the Gauss points (the reference's abscissae, src/qr.cpp:27,38, the 3-point
// one as truncated there), and the collocation derivative Dc[q][p] = L_p'(g_q)
// of the Lagrange polynomials through the Gauss points: for polynomials of
// degree < P, D u = Dc (B u), so the reference gradient at the points follows
// from the values at the points.
__host__ __device__ constexpr double sfGauss(int P, int q)
{
  return P == 3 ? (q == 0 ? -0.774596669241483 : (q == 1 ? 0. : 0.774596669241483))
                : (q == 0 ? -0.5773502691896258 : 0.5773502691896258);
}
__host__ __device__ constexpr double sfNode(int P, int i) { return P == 3 ? (double)(i - 1) : (i == 0 ? -1. : 1.); }
// weight of Gauss point q (src/qr.cpp)
__host__ __device__ constexpr double sfWeight(int P, int q) { return P == 3 ? (q == 1 ? 8. / 9. : 5. / 9.) : 1.; }
// reference local index (include/reffe.hpp:877-992 node orders) of the
// lexicographic node i + P (j + P k); makeTab checks it against the dof points
template <int P>
__host__ __device__ constexpr int sfLocal(int n)
{
  constexpr int q1[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  constexpr int q2[27] = {0,  8,  1,  11, 20, 9,  3,  10, 2,  12, 22, 13, 21, 26,
                          23, 15, 24, 14, 4,  16, 5,  19, 25, 17, 7,  18, 6};
  return P == 3 ? q2[n] : q1[n];
}
template <int P>
__host__ __device__ constexpr double sfB(int q, int i)
{
  double const t = sfGauss(P, q);
  double v = 1.;
  for (int m = 0; m < P; ++m)
    if (m != i)
      v *= (t - sfNode(P, m)) / (sfNode(P, i) - sfNode(P, m));
  return v;
}
template <int P>
__host__ __device__ constexpr double sfDc(int q, int p)
{
  double const t = sfGauss(P, q);
  double s = 0.;
  for (int m = 0; m < P; ++m)
  {
    if (m == p)
      continue;
    double d = 1. / (sfGauss(P, p) - sfGauss(P, m));
    for (int r = 0; r < P; ++r)
      if (r != p && r != m)
        d *= (t - sfGauss(P, r)) / (sfGauss(P, p) - sfGauss(P, r));
    s += d;
  }
  return s;
}
// s + a u with a constant a (zero and one fold away after unrolling; in
// FP32 the table value rounded to FP32)
template <typename T>
__device__ inline T sfMad(double a, T u, T s)
{
  if (a == 0.)
    return s;
  if (a == 1.)
    return s + u;
  return fma(T(a), u, s);
}
// table values at a lane-dependent point index q (selects over constants)
template <int P>
__device__ inline double sfBsel(int q, int i)
{
  double r = sfB<P>(0, i);
#pragma unroll
  for (int t = 1; t < P; ++t)
    r = q == t ? sfB<P>(t, i) : r;
  return r;
}
template <int P>
__device__ inline double sfGsel(int q)
{
  double r = sfGauss(P, 0);
#pragma unroll
  for (int t = 1; t < P; ++t)
    r = q == t ? sfGauss(P, t) : r;
  return r;
}
template <int P>
__device__ inline double sfWsel(int q)
{
  double r = sfWeight(P, 0);
#pragma unroll
  for (int t = 1; t < P; ++t)
    r = q == t ? sfWeight(P, t) : r;
  return r;
}
// values the compiler must have formed here, and memory accesses that stay
// on their side of this point (an empty asm with them as operands)
template <int P>
__device__ inline void sfFence(double (&a)[P], double (&b)[P], double (&c)[P])
{
#pragma unroll
  for (int i = 0; i < P; ++i)
    asm volatile("" : "+v"(a[i]), "+v"(b[i]), "+v"(c[i])::"memory");
}
// in place along direction DIR of a P^3 lexicographic array: w <- A w per
// line, A = B (DC = false) or Dc (DC = true), transposed for TR
template <int P, int DIR, bool TR, bool DC, typename T>
__device__ inline void sfLines(T (&w)[P * P * P])
{
  auto idx = [](int t, int a, int b)
  {
    if constexpr (DIR == 0)
      return t + P * (a + P * b);
    else if constexpr (DIR == 1)
      return a + P * (t + P * b);
    else
      return a + P * (b + P * t);
  };
#pragma unroll
  for (int b = 0; b < P; ++b)
#pragma unroll
    for (int a = 0; a < P; ++a)
    {
      T in[P], r[P];
#pragma unroll
      for (int t = 0; t < P; ++t)
        in[t] = w[idx(t, a, b)];
#pragma unroll
      for (int q = 0; q < P; ++q)
      {
        T s = T(0);
#pragma unroll
        for (int t = 0; t < P; ++t)
        {
          double const m = DC ? (TR ? sfDc<P>(t, q) : sfDc<P>(q, t)) : (TR ? sfB<P>(t, q) : sfB<P>(q, t));
          s = sfMad(m, in[t], s);
        }
        r[q] = s;
      }
#pragma unroll
      for (int q = 0; q < P; ++q)
        w[idx(q, a, b)] = r[q];
    }
}

// Batched apply, one lane per (element, vector): LPE lanes of a wavefront
// share an element and lane c applies the element operator to the vectors
// c, c + LPE, ... of the batch.  The element's point metrics are computed
// once (the lanes take the points in turn) and kept in LDS; the tensor
// contractions of a vector run in the lane's registers (collocation form,
// 5 x 3 P^4 multiply-adds less the folded zeros of the tables), and the
// element rows go to the element buffer ([E][n][nb], summed per dof by the
// gather).
// sum across the two corners of a cube pair in direction DIR (lanes LPE <<
// DIR apart): the nodes of the shared face, low corner's value first (a
// corner without an element holds zeros).  After the three directions every
// corner holds, at each node, the sum over the corners sharing it.
template <int P, int DIR, int LPE, typename T>
__device__ inline void sfCubeSum(T (&o)[P * P * P], int bit)
{
  auto idx = [](int t, int a, int b)
  {
    if constexpr (DIR == 0)
      return t + P * (a + P * b);
    else if constexpr (DIR == 1)
      return a + P * (t + P * b);
    else
      return a + P * (b + P * t);
  };
#pragma unroll
  for (int b = 0; b < P; ++b)
#pragma unroll
    for (int a = 0; a < P; ++a)
    {
      // my face node: t = P - 1 on the low corner, t = 0 on the high one
      T const lo = o[idx(P - 1, a, b)], hi = o[idx(0, a, b)];
      T const mine = bit ? hi : lo;
      T const theirs = __shfl_xor(mine, LPE << DIR);
      T const sum = bit ? theirs + mine : mine + theirs;
      if (bit)
        o[idx(0, a, b)] = sum;
      else
        o[idx(P - 1, a, b)] = sum;
    }
}

template <int P>
struct SfBatchSlot
{
  double M[6][P * P * P]; // metrics of the points (upper triangle, row-major)
  double X[24];           // vertex coordinates
  double kn[P * P * P];   // nodal coefficient values (lexicographic)
  uint32_t id[P * P * P]; // dof ids of the nodes (lexicographic)
};

// a[i] for a lane-dependent i, by selects over constant indices (a table in
// the kernel arguments indexed at run time would be copied to scratch)
template <int P>
__device__ inline double pick(double const (&a)[P], int i)
{
  double r = a[0];
#pragma unroll
  for (int t = 1; t < P; ++t)
    r = i == t ? a[t] : r;
  return r;
}
template <int P>
__device__ inline double pickRow(double const (&a)[P][P], int q, int i)
{
  double r = a[0][i];
#pragma unroll
  for (int t = 1; t < P; ++t)
    r = q == t ? a[t][i] : r;
  return r;
}

// waves per block of the batch kernel: the LDS of a block (EPW element
// slots per wave) stays at that of 4 waves of 8 elements
template <int LPE>
constexpr int sfBatchWaves()
{
  return LPE >= 8 ? kSfWaves : (LPE == 4 ? kSfWaves / 2 : 1);
}

// CUBE: the wave's element slots are the corners (a, b, c) = bits of the
// slot index of 2 x 2 x 2 element cubes (cubeElem[G][8], ~0u for an empty
// corner; cube faces match node for node in the same order).  The nodes a
// cube's elements share are summed across the lanes (x, then y, then z: the
// low corner's value first, so both lanes hold the same bits), and each
// node is written once, by the first corner holding it, as cubeOut[G][n][8]
// says: to y (bit 31, the dof id: every contribution of the dof is in the
// cube) or to the partial buffer [slot][nb], which k_sf_part sums per dof.
// Without CUBE: element e = slot index, element buffer out [E][n][nb].
template <int P, int LPE, bool CUBE>
__global__ void __launch_bounds__(64 * sfBatchWaves<LPE>())
    k_sf_batch(int64_t E, int64_t nslot, int nb, double const * __restrict__ xyz, uint32_t const * __restrict__ conn,
               uint32_t const * __restrict__ emap, uint32_t const * __restrict__ cubeElem,
               uint32_t const * __restrict__ cubeOut, double const * __restrict__ kappa,
               uint8_t const * __restrict__ cons, double const * __restrict__ x, double * __restrict__ out,
               double * __restrict__ part, int * __restrict__ status)
{
  constexpr int N = P * P * P, EPW = 64 / LPE, WPB = sfBatchWaves<LPE>();
  static_assert(!CUBE || EPW % 8 == 0, "whole cubes per wavefront");
  __shared__ SfBatchSlot<P> lds[WPB * EPW];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int const slot = lane / LPE, c = lane % LPE;
  SfBatchSlot<P> & L = lds[wave * EPW + slot];
  auto at = [](int i, int j, int k) { return i + P * (j + P * k); };
  constexpr int NV = (24 + LPE - 1) / LPE; // vertex coordinates staged per lane
  int64_t const ngroup = (nslot + EPW - 1) / EPW;
  int64_t const stride = (int64_t)gridDim.x * WPB;
  for (int64_t grp = (int64_t)blockIdx.x * WPB + wave; grp < ngroup; grp += stride)
  {
    int64_t const gs = grp * EPW + slot; // element slot
    int64_t e = gs;
    if constexpr (CUBE)
      e = gs < nslot ? (int64_t)cubeElem[gs] : (int64_t)0xffffffffu;
    bool const live = e < E;
    // (an opaque copy of the lane's index keeps the compiler from hoisting
    // the lane's point tables out of the loop and holding them in registers)
    int cl = c;
    asm volatile("" : "+v"(cl));
    // ---- 1. dof ids, constraint flags, vertex coordinates and coefficient
    //         values (the last two to LDS, the lanes taking them in turn)
    int64_t const ee = live ? e : E - 1; // a lane past the end reads the last element and stores nothing
    uint32_t id[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
      id[i] = emap[ee * N + sfLocal<P>(i)];
#pragma unroll
    for (int m = 0; m < NV; ++m)
      if (m * LPE + c < 24)
        L.X[m * LPE + c] = xyz[3ull * conn[ee * 8 + (m * LPE + c) / 3] + (m * LPE + c) % 3];
#pragma unroll
    for (int i0 = 0; i0 < N; i0 += LPE)
      if (i0 + c < N)
      {
        uint32_t d = id[0];
#pragma unroll
        for (int i = 0; i < N; ++i) // (constant register indices)
          d = i == i0 + c ? id[i] : d;
        L.kn[i0 + c] = kappa[d];
        L.id[i0 + c] = d;
      }
    uint32_t fl = 0;
    if (cons)
    {
#pragma unroll
      for (int i = 0; i < N; ++i)
        fl |= cons[id[i]] ? (1u << i) : 0u;
    }
    sfSync();
    // ---- 2. metrics of the points c, c + LPE, ... -> LDS
    bool bad = false;
#pragma unroll
    for (int q0 = 0; q0 < N; q0 += LPE)
    {
      int const q = q0 + cl;
      if (q < N)
      {
        int const qx = q % P, qy = (q / P) % P, qz = q / (P * P);
        double bx[P], by[P], bz[P]; // B[qx][.], B[qy][.], B[qz][.]
#pragma unroll
        for (int i = 0; i < P; ++i)
        {
          bx[i] = sfBsel<P>(qx, i);
          by[i] = sfBsel<P>(qy, i);
          bz[i] = sfBsel<P>(qz, i);
        }
        double kq = 0.; // B x B x B interpolation of the coefficient, x first
#pragma unroll
        for (int k = 0; k < P; ++k)
        {
          double t = 0.;
#pragma unroll
          for (int j = 0; j < P; ++j)
          {
            double s = 0.;
#pragma unroll
            for (int i = 0; i < P; ++i)
              s += bx[i] * L.kn[at(i, j, k)];
            t += by[j] * s;
          }
          kq += bz[k] * t;
        }
        double const wq = sfWsel<P>(qx) * sfWsel<P>(qy) * sfWsel<P>(qz);
        double Mq[6];
        bad |= !sfPointMetric(L.X, sfGsel<P>(qx), sfGsel<P>(qy), sfGsel<P>(qz), wq, kq, Mq);
#pragma unroll
        for (int m = 0; m < 6; ++m)
          L.M[m][q] = Mq[m];
      }
    }
    if (live && bad)
      atomicOr(status, 1);
    sfSync();
    // ---- 3. the vectors of the lane (collocation form: values at the
    //         points, gradients by the point-to-point derivative matrix,
    //         its transpose, then the transposed interpolation; all in the
    //         lane's registers with the 1D tables as constants)
    for (int b0 = 0; b0 < nb; b0 += LPE)
    {
      // the metrics are read from LDS where they are used, not hoisted
      // out of this loop (162 values would not fit in registers)
      asm volatile("" ::: "memory");
      int const v = b0 + c;
      int const vc = v < nb ? v : nb - 1;
      // (opaque copies of the constraint bits: their per-node masks are made
      // where they are used, not hoisted out of the loop into scalar registers)
      uint32_t fl1 = fl;
      asm volatile("" : "+v"(fl1));
      double w[N];
#pragma unroll
      for (int i = 0; i < N; ++i)
      {
        double const t = x[L.id[i] * (uint32_t)nb + (uint32_t)vc];
        w[i] = (fl1 >> i) & 1u ? 0. : t;
      }
      sfLines<P, 0, false, false>(w);
      sfLines<P, 1, false, false>(w);
      sfLines<P, 2, false, false>(w);
      double o[N];
#pragma unroll
      for (int n = 0; n < N; ++n)
        o[n] = 0.;
#pragma unroll
      for (int qz = 0; qz < P; ++qz)
#pragma unroll
        for (int qy = 0; qy < P; ++qy)
        {
          double vx[P], vy[P], vz[P];
#pragma unroll
          for (int qx = 0; qx < P; ++qx)
          {
            double gx = 0., gy = 0., gz = 0.;
#pragma unroll
            for (int t = 0; t < P; ++t)
            {
              gx = sfMad(sfDc<P>(qx, t), w[at(t, qy, qz)], gx);
              gy = sfMad(sfDc<P>(qy, t), w[at(qx, t, qz)], gy);
              gz = sfMad(sfDc<P>(qz, t), w[at(qx, qy, t)], gz);
            }
            int const q = at(qx, qy, qz);
            double const m0 = L.M[0][q], m1 = L.M[1][q], m2 = L.M[2][q];
            double const m3 = L.M[3][q], m4 = L.M[4][q], m5 = L.M[5][q];
            vx[qx] = m0 * gx + m1 * gy + m2 * gz;
            vy[qx] = m1 * gx + m3 * gy + m4 * gz;
            vz[qx] = m2 * gx + m4 * gy + m5 * gz;
          }
          // a line of points at a time: the line's products are formed before
          // the next line's metric values are read (the scheduler would
          // otherwise load all 6 P^3 of them at once and run out of registers)
          sfFence<P>(vx, vy, vz);
#pragma unroll
          for (int qx = 0; qx < P; ++qx)
#pragma unroll
            for (int t = 0; t < P; ++t)
            {
              o[at(t, qy, qz)] = sfMad(sfDc<P>(qx, t), vx[qx], o[at(t, qy, qz)]);
              o[at(qx, t, qz)] = sfMad(sfDc<P>(qy, t), vy[qx], o[at(qx, t, qz)]);
              o[at(qx, qy, t)] = sfMad(sfDc<P>(qz, t), vz[qx], o[at(qx, qy, t)]);
            }
        }
      sfLines<P, 0, true, false>(o);
      sfLines<P, 1, true, false>(o);
      sfLines<P, 2, true, false>(o);
      // the unit diagonal per element on constrained rows (the dof ids are
      // read again here rather than kept in registers across the loop)
      asm volatile("" ::: "memory");
      uint32_t fl2 = fl;
      asm volatile("" : "+v"(fl2));
#pragma unroll
      for (int n = 0; n < N; ++n)
        if ((fl2 >> n) & 1u)
          o[n] = x[L.id[n] * (uint32_t)nb + (uint32_t)vc];
      if constexpr (CUBE)
      {
        int const corner = slot & 7;
        if (!live)
        {
#pragma unroll
          for (int n = 0; n < N; ++n)
            o[n] = 0.;
        }
        sfCubeSum<P, 0, LPE>(o, corner & 1);
        sfCubeSum<P, 1, LPE>(o, (corner >> 1) & 1);
        sfCubeSum<P, 2, LPE>(o, (corner >> 2) & 1);
        if (live && v < nb)
        {
          uint32_t const * const co = cubeOut + (gs >> 3) * (N * 8) + corner;
#pragma unroll
          for (int n = 0; n < N; ++n)
          {
            uint32_t const code = co[n * 8];
            if (code == 0xffffffffu)
              continue;
            if (code & 0x80000000u)
              out[(size_t)(code & 0x7fffffffu) * nb + v] = o[n];
            else
              part[(size_t)code * nb + v] = o[n];
          }
        }
      }
      else if (live && v < nb)
      {
        double * const ob = out + (size_t)e * N * nb + v;
#pragma unroll
        for (int n = 0; n < N; ++n)
          ob[(uint32_t)sfLocal<P>(n) * (uint32_t)nb] = o[n];
      }
    }
    sfSync(); // the slot is rewritten by the next group
  }
}

// Apply with the per-point metrics computed once per coefficient field
// (kSfMetric, [E][6][n]: kappa, geometry and weights folded in; the
// "partially assembled" operator): per element only the dof ids, the
// constraint flags, P^3 vector values and 6 P^3 metric values are read, and
// only the tensor contractions are computed.  One vector; waves prefetch the
// next batch's data while the current one computes.
template <typename T, int P>
__global__ void __launch_bounds__(64 * kSfWaves)
    k_sf_apply_pa(int64_t E, SfTab<P> tab, uint32_t const * __restrict__ emap, uint8_t const * __restrict__ cons,
                  T const * __restrict__ met, T const * __restrict__ x, T * __restrict__ out)
{
  constexpr int P2 = P * P, N = P * P * P, EPW = 64 / P2;
  constexpr int SLOTS = (64 + P2 - 1) / P2;
  // three arrays of N values per stage buffer, in T
  struct Lds
  {
    T a[3][N];
    T b[3][N];
  };
  __shared__ Lds lds[kSfWaves * SLOTS];
  int const wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int const slot = lane / P2, pl = lane - slot * P2;
  int const pa = pl % P, pc = pl / P;
  Lds & L = lds[wave * SLOTS + slot];
  auto at = [](int i, int j, int k) { return i + P * (j + P * k); };
  int loc[P];
#pragma unroll
  for (int i = 0; i < P; ++i)
    loc[i] = tab.local[at(i, pa, pc)];
  int64_t const nbatch = (E + EPW - 1) / EPW;
  int64_t const stride = (int64_t)gridDim.x * kSfWaves;
  int64_t bt = (int64_t)blockIdx.x * kSfWaves + wave;
  auto elemOf = [&](int64_t b) -> int64_t { return b * EPW + slot; };
  auto clampE = [&](int64_t b) -> int64_t { return std::min(elemOf(b), E - 1); };
  struct D
  {
    uint32_t id[P];
    T xv[P];
    bool fl[P];
    T m[P][6]; // metric at the lane's z-pencil points (pa, pc, q)
  };
  auto loadIds = [&](D & d, int64_t b)
  {
    int64_t const e = clampE(b);
#pragma unroll
    for (int i = 0; i < P; ++i)
      d.id[i] = emap[e * N + loc[i]];
#pragma unroll
    for (int q = 0; q < P; ++q)
#pragma unroll
      for (int m = 0; m < 6; ++m)
        d.m[q][m] = met[((size_t)e * 6 + m) * N + at(pa, pc, q)];
  };
  auto loadVals = [&](D & d)
  {
#pragma unroll
    for (int i = 0; i < P; ++i)
    {
      d.xv[i] = x[d.id[i]];
      d.fl[i] = cons ? cons[d.id[i]] != 0 : false;
    }
  };
  D cur, nxt;
  loadIds(cur, bt);
  loadVals(cur);
  T * A0 = L.a[0];
  T * A1 = L.a[1];
  T * A2 = L.a[2];
  T * B0 = L.b[0];
  T * B1 = L.b[1];
  T * B2 = L.b[2];
  while (bt < nbatch)
  {
    bool const live = slot < EPW && elemOf(bt) < E;
    int64_t const e = elemOf(bt);
    int64_t const bn = bt + stride;
    loadIds(nxt, bn);
    T u[P];
#pragma unroll
    for (int i = 0; i < P; ++i)
      u[i] = cur.fl[i] ? T(0) : cur.xv[i];
#pragma unroll
    for (int q = 0; q < P; ++q) // x: lane (j, k)
    {
      T sb = T(0), sd = T(0);
#pragma unroll
      for (int i = 0; i < P; ++i)
      {
        sb += tB<T>(tab, q, i) * u[i];
        sd += tD<T>(tab, q, i) * u[i];
      }
      A0[at(q, pa, pc)] = sb;
      A1[at(q, pa, pc)] = sd;
    }
    sfSync();
    loadVals(nxt);
#pragma unroll
    for (int q = 0; q < P; ++q) // y: lane (qx, k)
    {
      T bb = T(0), bd = T(0), db = T(0);
#pragma unroll
      for (int j = 0; j < P; ++j)
      {
        T const ub = A0[at(pa, j, pc)], ud = A1[at(pa, j, pc)];
        bb += tB<T>(tab, q, j) * ub;
        bd += tD<T>(tab, q, j) * ub;
        db += tB<T>(tab, q, j) * ud;
      }
      B0[at(pa, q, pc)] = bb;
      B1[at(pa, q, pc)] = bd;
      B2[at(pa, q, pc)] = db;
    }
    sfSync();
    T v[P][3];
#pragma unroll
    for (int q = 0; q < P; ++q) // z: lane (qx, qy), then the metric
    {
      T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
      for (int k = 0; k < P; ++k)
      {
        gx += tB<T>(tab, q, k) * B2[at(pa, pc, k)];
        gy += tB<T>(tab, q, k) * B1[at(pa, pc, k)];
        gz += tD<T>(tab, q, k) * B0[at(pa, pc, k)];
      }
      T const * M = cur.m[q];
      v[q][0] = M[0] * gx + M[1] * gy + M[2] * gz;
      v[q][1] = M[1] * gx + M[3] * gy + M[4] * gz;
      v[q][2] = M[2] * gx + M[4] * gy + M[5] * gz;
    }
#pragma unroll
    for (int k = 0; k < P; ++k) // z^T
    {
      T s1 = T(0), s2 = T(0), s3 = T(0);
#pragma unroll
      for (int q = 0; q < P; ++q)
      {
        s1 += tB<T>(tab, q, k) * v[q][0];
        s2 += tB<T>(tab, q, k) * v[q][1];
        s3 += tD<T>(tab, q, k) * v[q][2];
      }
      A0[at(pa, pc, k)] = s1;
      A1[at(pa, pc, k)] = s2;
      A2[at(pa, pc, k)] = s3;
    }
    sfSync();
#pragma unroll
    for (int j = 0; j < P; ++j) // y^T: lane (qx, k)
    {
      T sd = T(0), sb = T(0);
#pragma unroll
      for (int q = 0; q < P; ++q)
      {
        sd += tB<T>(tab, q, j) * A0[at(pa, q, pc)];
        sb += tD<T>(tab, q, j) * A1[at(pa, q, pc)] + tB<T>(tab, q, j) * A2[at(pa, q, pc)];
      }
      B0[at(pa, j, pc)] = sd;
      B1[at(pa, j, pc)] = sb;
    }
    sfSync();
#pragma unroll
    for (int i = 0; i < P; ++i) // x^T: lane (j, k)
    {
      T y = T(0);
#pragma unroll
      for (int q = 0; q < P; ++q)
        y += tD<T>(tab, q, i) * B0[at(q, pa, pc)] + tB<T>(tab, q, i) * B1[at(q, pa, pc)];
      if (cur.fl[i])
        y = cur.xv[i]; // unit diagonal per element on constrained rows
      if (live)
        out[(size_t)e * N + loc[i]] = y;
    }
    sfSync();
    cur = nxt;
    bt = bn;
  }
}

// per-dof sum of the element buffer in element order (nb = 1), with the
// epilogue of the operator use: apply y = Ax, residual y = b - Ax, Jacobi
// y = x + omega dinv (b - Ax), inverse y = 1 / Ax (the diagonal)
enum : int
{
  kGApply = 0,
  kGResidual = 1,
  kGJacobi = 2,
  kGInverse = 3,
};

template <typename T, int MODE>
__global__ void __launch_bounds__(256)
    k_sf_gather(int64_t ndofs, uint32_t const * __restrict__ seg, uint32_t const * __restrict__ ids,
                T const * __restrict__ buf, T const * __restrict__ x, T const * __restrict__ b,
                T const * __restrict__ dinv, T omega, T * __restrict__ y)
{
  int64_t const d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndofs)
    return;
  uint32_t const k0 = seg[d], k1 = seg[d + 1];
  T acc = T(0);
  if (k1 > k0)
  {
    constexpr int kMaxC = 8;
    uint32_t id[kMaxC];
#pragma unroll
    for (int j = 0; j < kMaxC; ++j)
      id[j] = ids[min(k0 + j, k1 - 1)];
    T v[kMaxC];
#pragma unroll
    for (int j = 0; j < kMaxC; ++j)
      v[j] = buf[id[j]];
#pragma unroll
    for (int j = 0; j < kMaxC; ++j)
      if (k0 + j < k1)
        acc += v[j];
    for (uint32_t k = k0 + kMaxC; k < k1; ++k)
      acc += buf[ids[k]];
  }
  if constexpr (MODE == kGApply)
    y[d] = acc;
  else if constexpr (MODE == kGResidual)
    y[d] = b[d] - acc;
  else if constexpr (MODE == kGJacobi)
    y[d] = x[d] + omega * dinv[d] * (b[d] - acc);
  else
    y[d] = T(1) / acc;
}

// y[d][b] of the dofs the cubes leave partial: the sum of their partial
// values in cube order
__global__ void k_sf_part(int64_t npd, int nb, uint32_t const * __restrict__ dofs, uint32_t const * __restrict__ seg,
                          uint32_t const * __restrict__ ids, double const * __restrict__ part,
                          double * __restrict__ y)
{
  int64_t const t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npd * nb)
    return;
  int64_t const k = t / nb;
  int const b = (int)(t - k * nb);
  uint32_t const k0 = seg[k], k1 = seg[k + 1];
  constexpr int kMaxC = 8; // a cube-face dof has 2 parts, an edge dof 4, a vertex 8
  uint32_t id[kMaxC];
#pragma unroll
  for (int j = 0; j < kMaxC; ++j)
    id[j] = ids[min(k0 + j, k1 - 1)];
  double v[kMaxC];
#pragma unroll
  for (int j = 0; j < kMaxC; ++j)
    v[j] = part[(size_t)id[j] * nb + b];
  double acc = v[0];
#pragma unroll
  for (int j = 1; j < kMaxC; ++j)
    if (k0 + j < k1)
      acc += v[j];
  for (uint32_t j = k0 + kMaxC; j < k1; ++j)
    acc += part[(size_t)ids[j] * nb + b];
  y[(size_t)dofs[k] * nb + b] = acc;
}

// metrics of k_sf_elem's [E][6][n] layout to the slots of the element
// cubes, [chunk][6][n][64] (element slot innermost)
template <typename T>
__global__ void k_met_slots(int64_t E, int n, uint32_t const * __restrict__ slotOf, T const * __restrict__ met,
                            T * __restrict__ metC)
{
  int64_t const t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t const per = 6 * (int64_t)n;
  if (t >= E * per)
    return;
  int64_t const e = t / per;
  int64_t const r = t - e * per;
  uint32_t const so = slotOf[e];
  metC[((int64_t)(so >> 6) * per + r) * 64 + (so & 63)] = met[t];
}

// the epilogue of a summed operator row d: y = A x (0), b - A x (1),
// x + omega dinv (b - A x) (2)
template <int MODE>
__device__ inline float sfEpilogue(float acc, uint32_t d, float const * __restrict__ x, float const * __restrict__ b,
                                   float const * __restrict__ dinv, float omega)
{
  if constexpr (MODE == kGApply)
    return acc;
  else if constexpr (MODE == kGResidual)
    return b[d] - acc;
  else
    return x[d] + omega * dinv[d] * (b[d] - acc);
}

// One vector through the FP32 operator of stored metrics (the smoother,
// residual and apply of a matrix-free multigrid level), one lane per
// element slot of the element cubes: the slot's dof ids, complete-node bits
// and metrics are read lane-contiguous (chunks of 64 slots), the tensor
// contractions run in registers (collocation form), the cube's shared nodes
// are summed across lanes, and the slot writes the nodes it owns: the
// epilogue's value to y where the cube holds every contribution, else the
// partial value at (cube, lattice node), which k_sf_part_f sums per dof.
template <int P, int MODE>
__global__ void __launch_bounds__(256)
    k_sf_pa_cube(int64_t nChunk, uint32_t const * __restrict__ idC, uint32_t const * __restrict__ cmC,
                 uint8_t const * __restrict__ cons, float const * __restrict__ metC, float const * __restrict__ x,
                 float const * __restrict__ b, float const * __restrict__ dinv, float omega, float * __restrict__ y,
                 float * __restrict__ part)
{
  constexpr int N = P * P * P, LD = 2 * P - 1, LATT = LD * LD * LD;
  int64_t const ck = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (ck >= nChunk) // (whole wavefronts)
    return;
  int const lane = threadIdx.x & 63, corner = lane & 7;
  auto at = [](int i, int j, int k) { return i + P * (j + P * k); };
  uint32_t const cm = cmC[ck * 64 + lane];
  bool const live = (cm >> 31) != 0;
  uint32_t const * const ids = idC + (size_t)ck * N * 64 + lane;
  float const * const mt = metC + (size_t)ck * 6 * N * 64 + lane;
  // loads without branches around them (a slot without an element reads
  // dof 0), so each phase's loads are in flight together
  float w[N];
  uint32_t fl = 0;
  {
    uint32_t d[N];
#pragma unroll
    for (int n = 0; n < N; ++n)
      d[n] = ids[n * 64] & 0x7fffffffu;
#pragma unroll
    for (int n = 0; n < N; ++n)
      w[n] = x[d[n]];
    if (cons) // (uniform)
    {
      uint8_t cn[N];
#pragma unroll
      for (int n = 0; n < N; ++n)
        cn[n] = cons[d[n]];
#pragma unroll
      for (int n = 0; n < N; ++n)
      {
        fl |= cn[n] ? 1u << n : 0u;
        w[n] = cn[n] ? 0.f : w[n];
      }
    }
  }
  sfLines<P, 0, false, false>(w);
  sfLines<P, 1, false, false>(w);
  sfLines<P, 2, false, false>(w);
  float o[N];
#pragma unroll
  for (int n = 0; n < N; ++n)
    o[n] = 0.f;
#pragma unroll
  for (int qz = 0; qz < P; ++qz)
#pragma unroll
    for (int qy = 0; qy < P; ++qy)
#pragma unroll
      for (int qx = 0; qx < P; ++qx)
      {
        float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int t = 0; t < P; ++t)
        {
          gx = sfMad(sfDc<P>(qx, t), w[at(t, qy, qz)], gx);
          gy = sfMad(sfDc<P>(qy, t), w[at(qx, t, qz)], gy);
          gz = sfMad(sfDc<P>(qz, t), w[at(qx, qy, t)], gz);
        }
        int const q = at(qx, qy, qz);
        float const m0 = mt[(0 * N + q) * 64], m1 = mt[(1 * N + q) * 64], m2 = mt[(2 * N + q) * 64];
        float const m3 = mt[(3 * N + q) * 64], m4 = mt[(4 * N + q) * 64], m5 = mt[(5 * N + q) * 64];
        float const vx = m0 * gx + m1 * gy + m2 * gz;
        float const vy = m1 * gx + m3 * gy + m4 * gz;
        float const vz = m2 * gx + m4 * gy + m5 * gz;
#pragma unroll
        for (int t = 0; t < P; ++t)
        {
          o[at(t, qy, qz)] = sfMad(sfDc<P>(qx, t), vx, o[at(t, qy, qz)]);
          o[at(qx, t, qz)] = sfMad(sfDc<P>(qy, t), vy, o[at(qx, t, qz)]);
          o[at(qx, qy, t)] = sfMad(sfDc<P>(qz, t), vz, o[at(qx, qy, t)]);
        }
      }
  sfLines<P, 0, true, false>(o);
  sfLines<P, 1, true, false>(o);
  sfLines<P, 2, true, false>(o);
  // the unit diagonal per element on constrained rows; a slot without an
  // element holds zeros
  if (fl) // (constrained nodes: reload their x values, all at once)
  {
    float xr[N];
#pragma unroll
    for (int n = 0; n < N; ++n)
      xr[n] = x[ids[n * 64] & 0x7fffffffu];
#pragma unroll
    for (int n = 0; n < N; ++n)
      o[n] = (fl >> n) & 1u ? xr[n] : o[n];
  }
#pragma unroll
  for (int n = 0; n < N; ++n)
    o[n] = live ? o[n] : 0.f;
  sfCubeSum<P, 0, 1>(o, corner & 1);
  sfCubeSum<P, 1, 1>(o, (corner >> 1) & 1);
  sfCubeSum<P, 2, 1>(o, (corner >> 2) & 1);
  int64_t const cube = ck * 8 + (lane >> 3);
  // the owned nodes, a third of them at a time: epilogue inputs loaded
  // unconditionally (valid dofs), then predicated stores
  constexpr int G3 = (N + 2) / 3;
#pragma unroll
  for (int n0 = 0; n0 < N; n0 += G3)
  {
    uint32_t iv[G3];
    float ev[G3];
#pragma unroll
    for (int k = 0; k < G3; ++k)
      iv[k] = n0 + k < N ? ids[(n0 + k) * 64] : 0u;
#pragma unroll
    for (int k = 0; k < G3; ++k)
      if (n0 + k < N)
        ev[k] = sfEpilogue<MODE>(o[n0 + k], iv[k] & 0x7fffffffu, x, b, dinv, omega);
#pragma unroll
    for (int k = 0; k < G3; ++k)
    {
      int const n = n0 + k;
      if (n >= N || !(iv[k] & 0x80000000u))
        continue;
      if ((cm >> n) & 1u)
        y[iv[k] & 0x7fffffffu] = ev[k];
      else
      {
        int const X = (corner & 1) * (P - 1) + n % P;
        int const Y = ((corner >> 1) & 1) * (P - 1) + (n / P) % P;
        int const Z = ((corner >> 2) & 1) * (P - 1) + n / (P * P);
        part[cube * LATT + X + LD * (Y + LD * Z)] = o[n];
      }
    }
  }
}

// rows the cubes leave partial: the sum of their partial values in cube
// order, then the epilogue
template <int MODE>
__global__ void k_sf_part_f(int64_t npd, uint32_t const * __restrict__ dofs, uint32_t const * __restrict__ seg,
                            uint32_t const * __restrict__ ids, float const * __restrict__ part,
                            float const * __restrict__ x, float const * __restrict__ b,
                            float const * __restrict__ dinv, float omega, float * __restrict__ y)
{
  int64_t const k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npd)
    return;
  uint32_t const k0 = seg[k], k1 = seg[k + 1];
  constexpr int kMaxC = 8;
  uint32_t id[kMaxC];
#pragma unroll
  for (int j = 0; j < kMaxC; ++j)
    id[j] = ids[min(k0 + j, k1 - 1)];
  float v[kMaxC];
#pragma unroll
  for (int j = 0; j < kMaxC; ++j)
    v[j] = part[id[j]];
  float acc = v[0];
#pragma unroll
  for (int j = 1; j < kMaxC; ++j)
    if (k0 + j < k1)
      acc += v[j];
  for (uint32_t j = k0 + kMaxC; j < k1; ++j)
    acc += part[ids[j]];
  uint32_t const d = dofs[k];
  y[d] = sfEpilogue<MODE>(acc, d, x, b, dinv, omega);
}

template <int P>
SfTab<P> makeTab(int feType)
{
  SfTab<P> t{};
  // 1D Gauss rule of the reference (src/qr.cpp:27,38; the 3-point abscissa
  // as truncated there) and the 1D Lagrange nodes -1 (, 0), +1
  double const a3 = 0.774596669241483, a2 = 0.5773502691896258;
  double const nodes3[3] = {-1., 0., 1.}, nodes2[2] = {-1., 1.};
  double const * nd = P == 3 ? nodes3 : nodes2;
  for (int q = 0; q < P; ++q)
  {
    t.g[q] = P == 3 ? (q == 0 ? -a3 : (q == 1 ? 0. : a3)) : (q == 0 ? -a2 : a2);
    t.w[q] = P == 3 ? (q == 1 ? 8. / 9. : 5. / 9.) : 1.;
    for (int a = 0; a < P; ++a)
    {
      double v = 1., dv = 0.;
      for (int m = 0; m < P; ++m)
      {
        if (m == a)
          continue;
        double const f = (t.g[q] - nd[m]) / (nd[a] - nd[m]);
        double dp = 1. / (nd[a] - nd[m]);
        for (int r = 0; r < P; ++r)
          if (r != a && r != m)
            dp *= (t.g[q] - nd[r]) / (nd[a] - nd[r]);
        dv += dp;
        v *= f;
      }
      t.B[q][a] = v;
      t.D[q][a] = dv;
      t.Bf[q][a] = (float)v;
      t.Df[q][a] = (float)dv;
    }
  }
  // reference local index of each lexicographic node: the node whose
  // dof point (in reference coordinates) is (nd[i], nd[j], nd[k])
  FeInfo const & fe = feInfo(feType);
  double vtx[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                      {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  double pts[27 * 3];
  fe.dofPoints(&vtx[0][0], pts);
  for (int k = 0; k < P; ++k)
    for (int j = 0; j < P; ++j)
      for (int i = 0; i < P; ++i)
      {
        int found = -1;
        for (int L = 0; L < fe.n; ++L)
          if (pts[3 * L] == nd[i] && pts[3 * L + 1] == nd[j] && pts[3 * L + 2] == nd[k])
            found = L;
        if (found < 0 || found != sfLocal<P>(i + P * (j + P * k)))
          throw Error(PXG_E_INVALID, "tensor-product node table");
        t.local[i + P * (j + P * k)] = found;
      }
  return t;
}

template <int P>
SfTab<P> const & tabFor(int feType)
{
  static SfTab<P> const t = makeTab<P>(feType);
  return t;
}

// the context holds a Q1 / Q2 hexahedron with the recommended tensor Gauss rule
int sfOrder(pxg_ctx const * c)
{
  FeTables const & T = c->fe;
  int P = 0;
  if (T.feType == PXG_FE_HEX_Q1)
    P = 2;
  else if (T.feType == PXG_FE_HEX_Q2)
    P = 3;
  else
    return 0;
  Quadrature const q = recommendedQuadrature(T.feType);
  if (T.nq != q.nq || T.xi != q.xi || T.w != q.w)
    return 0;
  return P;
}

// PXG_SF_BATCH=0 keeps FP64 batches on the pencil kernel (for comparisons)
bool sfBatchOn()
{
  static bool const on = []
  {
    char const * e = std::getenv("PXG_SF_BATCH");
    return !(e && e[0] == '0');
  }();
  return on;
}

template <typename T, int P, int MODE, int NBC>
void launchSfP(pxg_ctx * c, int nb, double const * kappa, uint8_t const * cons, T const * x, T const * ybar, T * out,
               hipStream_t s)
{
  int64_t const E = c->nelem;
  constexpr int EPW = 64 / (P * P);
  auto kern = k_sf_elem<T, P, MODE, NBC>;
  int dev = 0, ncu = 0, perCu = 0;
  PXG_HIP(hipGetDevice(&dev));
  PXG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  PXG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kern, 64 * kSfWaves, 0));
  int64_t const nbatch = (E + EPW - 1) / EPW;
  int64_t const need = (nbatch + kSfWaves - 1) / kSfWaves;
  // one resident wave of blocks; each wave loops over its batches
  int64_t const grid = std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ncu * std::max(perCu, 1)));
  kern<<<(unsigned)grid, 64 * kSfWaves, 0, s>>>(E, nb, tabFor<P>(P == 3 ? PXG_FE_HEX_Q2 : PXG_FE_HEX_Q1), c->xyz.p,
                                                c->conn.p, c->elemMap.p, kappa, cons, x, ybar, out, c->status.p);
  PXG_HIP(hipGetLastError());
}

// CUBE: out = y, part = the partial buffer; otherwise out = element buffer
template <int P, int LPE, bool CUBE>
void launchSfBatchP(pxg_ctx * c, int nb, double const * kappa, uint8_t const * cons, double const * x, double * out,
                    double * part, hipStream_t s)
{
  int64_t const E = c->nelem;
  int64_t const nslot = CUBE ? c->nCube * 8 : E;
  constexpr int EPW = 64 / LPE, WPB = sfBatchWaves<LPE>();
  auto kern = k_sf_batch<P, LPE, CUBE>;
  int dev = 0, ncu = 0, perCu = 0;
  PXG_HIP(hipGetDevice(&dev));
  PXG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  PXG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kern, 64 * WPB, 0));
  int64_t const ngroup = (nslot + EPW - 1) / EPW;
  int64_t const need = (ngroup + WPB - 1) / WPB;
  int64_t const grid = std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ncu * std::max(perCu, 1)));
  kern<<<(unsigned)grid, 64 * WPB, 0, s>>>(E, nslot, nb, c->xyz.p, c->conn.p, c->elemMap.p, c->cubeElem.p,
                                           c->cubeOut.p, kappa, cons, x, out, part, c->status.p);
  PXG_HIP(hipGetLastError());
}

// FP64 batches: lanes per element = the batch rounded up to 2, 4 or 8
template <int P, bool CUBE>
void launchSfBatch(pxg_ctx * c, int nb, double const * kappa, uint8_t const * cons, double const * x, double * out,
                   double * part, hipStream_t s)
{
  if (nb > 4)
    launchSfBatchP<P, 8, CUBE>(c, nb, kappa, cons, x, out, part, s);
  else if (nb > 2)
    launchSfBatchP<P, 4, CUBE>(c, nb, kappa, cons, x, out, part, s);
  else
    launchSfBatchP<P, 2, CUBE>(c, nb, kappa, cons, x, out, part, s);
}

// ---------------------------------------------------------------------------
// Element cubes of the batch apply, built once per mesh and dof map on the
// host.  Face neighbours: element f is the +d neighbour of e when f's low
// face in direction d holds e's high face's dofs in the same order.  Cubes
// are formed greedily in element order, e at corner (0, 0, 0), the other
// corners from the neighbour relation; a corner that is missing, taken or
// inconsistent stays empty (every filled pair of corners must be linked, and
// distinct lattice nodes of the cube must hold distinct dofs, else the cube
// keeps e alone).  Each lattice node is written by the first corner holding
// it: to y when the cube holds every contribution of its dof (the dof's
// element count), else to a partial slot; per partial dof its slots in cube
// order.  Returns false when the cubes would leave most wavefront slots
// empty (meshes without the structure).
template <int P>
bool buildCubes(pxg_ctx * c)
{
  constexpr int N = P * P * P, F = P * P, LD = 2 * P - 1;
  int64_t const E = c->nelem, nd = c->ndofs;
  if (nd >= ((int64_t)1 << 31) || E >= (int64_t)0xffffffffLL)
    return false;
  std::vector<uint32_t> em((size_t)E * N);
  c->elemMap.download(em.data(), em.size(), c->stream);
  PXG_HIP(hipStreamSynchronize(c->stream));
  auto lex = [&](int64_t e, int n) { return em[(size_t)e * N + sfLocal<P>(n)]; };
  auto node = [](int d, int t, int a, int b)
  { return d == 0 ? t + P * (a + P * b) : (d == 1 ? a + P * (t + P * b) : a + P * (b + P * t)); };
  using Face = std::array<uint32_t, F>;
  struct FaceHash
  {
    size_t operator()(Face const & f) const
    {
      uint64_t h = 1469598103934665603ull;
      for (uint32_t v : f)
        h = (h ^ v) * 1099511628211ull;
      return (size_t)h;
    }
  };
  auto face = [&](int64_t e, int d, int t)
  {
    Face f;
    for (int b = 0; b < P; ++b)
      for (int a = 0; a < P; ++a)
        f[a + P * b] = lex(e, node(d, t, a, b));
    return f;
  };
  std::vector<uint32_t> nbr((size_t)3 * E, 0xffffffffu);
  for (int d = 0; d < 3; ++d)
  {
    std::unordered_map<Face, uint32_t, FaceHash> low;
    low.reserve((size_t)E * 2);
    for (int64_t e = 0; e < E; ++e)
      low.emplace(face(e, d, 0), (uint32_t)e);
    for (int64_t e = 0; e < E; ++e)
    {
      auto it = low.find(face(e, d, P - 1));
      if (it != low.end() && it->second != (uint32_t)e)
        nbr[(size_t)d * E + e] = it->second;
    }
  }
  std::vector<uint32_t> count((size_t)nd, 0);
  for (uint32_t v : em)
    count[v]++;
  for (uint32_t k : count)
    if (!k) // a dof of no element: the element-buffer gather writes its zero
      return false;
  std::vector<uint8_t> used((size_t)E, 0);
  std::vector<uint32_t> ce, co;
  std::vector<std::pair<uint32_t, uint32_t>> parts; // (dof, slot)
  // one-vector FP32 operator (k_sf_pa_cube): per slot its nodes' dofs
  // (bit 31: the slot writes the node), lanes of a chunk of 64 slots
  // innermost; per slot the complete nodes (bit 31: the slot has an element);
  // partial values at (cube, lattice node)
  constexpr int LATT = LD * LD * LD;
  std::vector<uint32_t> idC, cmC, slotOf((size_t)E, 0);
  std::vector<std::pair<uint32_t, uint32_t>> partsF;
  uint32_t nslot = 0;
  int64_t full = 0;
  std::vector<int> posOf(LD * LD * LD);
  std::unordered_map<uint32_t, int> dofPos;
  for (int64_t e0 = 0; e0 < E; ++e0)
  {
    if (used[e0])
      continue;
    uint32_t sl[8];
    auto nx = [&](uint32_t e, int d) { return e == 0xffffffffu ? e : nbr[(size_t)d * E + e]; };
    sl[0] = (uint32_t)e0;
    sl[1] = nx(sl[0], 0);
    sl[2] = nx(sl[0], 1);
    sl[3] = nx(sl[1], 1);
    sl[4] = nx(sl[0], 2);
    sl[5] = nx(sl[1], 2);
    sl[6] = nx(sl[2], 2);
    sl[7] = nx(sl[3], 2);
    for (int t = 1; t < 8; ++t)
    {
      bool bad = sl[t] != 0xffffffffu && used[sl[t]];
      for (int u = 0; u < t && !bad; ++u)
        bad = sl[u] == sl[t];
      if (bad)
        sl[t] = 0xffffffffu;
    }
    bool ok = true;
    for (int t = 0; t < 8 && ok; ++t)
      for (int d = 0; d < 3 && ok; ++d)
        if (!((t >> d) & 1) && sl[t] != 0xffffffffu && sl[t | (1 << d)] != 0xffffffffu)
          ok = nbr[(size_t)d * E + sl[t]] == sl[t | (1 << d)];
    // lattice consistency: one dof per node, one node per dof
    std::fill(posOf.begin(), posOf.end(), -1);
    dofPos.clear();
    for (int t = 0; t < 8 && ok; ++t)
    {
      if (sl[t] == 0xffffffffu)
        continue;
      for (int n = 0; n < N && ok; ++n)
      {
        int const i = n % P, j = (n / P) % P, k = n / F;
        int const X = (t & 1) * (P - 1) + i, Y = ((t >> 1) & 1) * (P - 1) + j, Z = ((t >> 2) & 1) * (P - 1) + k;
        int const pos = X + LD * (Y + LD * Z);
        uint32_t const dof = lex(sl[t], n);
        if (posOf[pos] < 0)
        {
          auto r = dofPos.emplace(dof, pos);
          ok = r.second;
          posOf[pos] = (int)dof;
        }
        else
          ok = (uint32_t)posOf[pos] == dof;
      }
    }
    if (!ok)
      for (int t = 1; t < 8; ++t)
        sl[t] = 0xffffffffu;
    int filled = 0;
    for (int t = 0; t < 8; ++t)
      if (sl[t] != 0xffffffffu)
      {
        used[sl[t]] = 1;
        ++filled;
      }
    full += filled == 8;
    size_t const g = ce.size() / 8;
    ce.insert(ce.end(), sl, sl + 8);
    co.resize((g + 1) * N * 8, 0xffffffffu);
    size_t const nck = (8 * (g + 1) + 63) / 64;
    idC.resize(nck * N * 64, 0u);
    cmC.resize(nck * 64, 0u);
    for (int t = 0; t < 8; ++t)
    {
      if (sl[t] == 0xffffffffu)
        continue;
      size_t const so = g * 8 + t;
      slotOf[sl[t]] = (uint32_t)so;
      cmC[so] = 0x80000000u;
      for (int n = 0; n < N; ++n)
        idC[((so / 64) * N + n) * 64 + so % 64] = lex(sl[t], n);
    }
    // owner and completeness of each lattice node
    for (int Z = 0; Z < LD; ++Z)
      for (int Y = 0; Y < LD; ++Y)
        for (int X = 0; X < LD; ++X)
        {
          int owner = -1, owned = 0, onode = 0;
          for (int t = 0; t < 8; ++t)
          {
            if (sl[t] == 0xffffffffu)
              continue;
            int const i = X - (t & 1) * (P - 1), j = Y - ((t >> 1) & 1) * (P - 1), k = Z - ((t >> 2) & 1) * (P - 1);
            if (i < 0 || i >= P || j < 0 || j >= P || k < 0 || k >= P)
              continue;
            ++owned;
            if (owner < 0)
            {
              owner = t;
              onode = i + P * (j + P * k);
            }
          }
          if (owner < 0)
            continue;
          uint32_t const dof = lex(sl[owner], onode);
          uint32_t code;
          size_t const so = g * 8 + owner;
          idC[((so / 64) * N + onode) * 64 + so % 64] |= 0x80000000u;
          if ((uint32_t)owned == count[dof])
          {
            code = dof | 0x80000000u;
            cmC[so] |= 1u << onode;
          }
          else
          {
            code = nslot++;
            parts.emplace_back(dof, code);
            partsF.emplace_back(dof, (uint32_t)(g * LATT + (X + LD * (Y + LD * Z))));
          }
          co[(g * N + onode) * 8 + owner] = code;
        }
  }
  int64_t const G = (int64_t)(ce.size() / 8);
  if (8 * G > 2 * E && full * 8 < E) // mostly lone elements: keep the element buffer
    return false;
  // per partial dof its slots in cube order
  auto lists = [](std::vector<std::pair<uint32_t, uint32_t>> & pr, std::vector<uint32_t> & pd,
                  std::vector<uint32_t> & ps, std::vector<uint32_t> & pi)
  {
    std::stable_sort(pr.begin(), pr.end(), [](auto const & a, auto const & b) { return a.first < b.first; });
    pi.reserve(pr.size());
    for (size_t k = 0; k < pr.size(); ++k)
    {
      if (k == 0 || pr[k].first != pr[k - 1].first)
      {
        pd.push_back(pr[k].first);
        ps.push_back((uint32_t)k);
      }
      pi.push_back(pr[k].second);
    }
    ps.push_back((uint32_t)pr.size());
  };
  std::vector<uint32_t> pd, ps, pi, pdF, psF, piF;
  lists(parts, pd, ps, pi);
  lists(partsF, pdF, psF, piF);
  hipStream_t s = c->stream;
  c->cubeElem.upload(ce.data(), ce.size(), s);
  c->cubeOut.upload(co.data(), co.size(), s);
  c->partDof.upload(pd.data(), pd.size(), s);
  c->partSeg.upload(ps.data(), ps.size(), s);
  c->partIds.upload(pi.data(), pi.size(), s);
  c->cubeIdC.upload(idC.data(), idC.size(), s);
  c->cubeCmC.upload(cmC.data(), cmC.size(), s);
  c->cubeSlotOf.upload(slotOf.data(), slotOf.size(), s);
  c->partDofF.upload(pdF.data(), pdF.size(), s);
  c->partSegF.upload(psF.data(), psF.size(), s);
  c->partIdsF.upload(piF.data(), piF.size(), s);
  PXG_HIP(hipStreamSynchronize(s));
  c->nCube = G;
  c->nPartDof = (int64_t)pd.size();
  c->nPart = nslot;
  c->nChunk = (int64_t)(cmC.size() / 64);
  c->nPartDofF = (int64_t)pdF.size();
  return true;
}

template <typename T, int MODE>
void launchSf(pxg_ctx * c, int P, int nb, double const * kappa, uint8_t const * cons, T const * x, T const * ybar,
              T * out, hipStream_t s)
{
  if constexpr (MODE == kSfApply && sizeof(T) == 8)
  {
    if (nb > 1 && sfBatchOn())
    {
      PXG_REQUIRE((double)c->ndofs * nb < 4294967295., PXG_E_INVALID, "batch too large for 32-bit vector offsets");
      if (P == 3)
        launchSfBatch<3, false>(c, nb, kappa, cons, x, out, nullptr, s);
      else
        launchSfBatch<2, false>(c, nb, kappa, cons, x, out, nullptr, s);
      return;
    }
  }
  // vectors per chunk: 1 for a single vector (its values ride with the
  // prefetch), 8 for batches of the apply, 4 for the gradient (two vectors
  // per chunk entry)
  if (P == 3)
  {
    if constexpr (MODE == kSfDiag || MODE == kSfMetric)
      launchSfP<T, 3, MODE, 1>(c, nb, kappa, cons, x, ybar, out, s);
    else if (nb > 1)
      launchSfP<T, 3, MODE, MODE == kSfApply ? 8 : 4>(c, nb, kappa, cons, x, ybar, out, s);
    else
      launchSfP<T, 3, MODE, 1>(c, nb, kappa, cons, x, ybar, out, s);
  }
  else
  {
    if constexpr (MODE == kSfDiag || MODE == kSfMetric)
      launchSfP<T, 2, MODE, 1>(c, nb, kappa, cons, x, ybar, out, s);
    else if (nb > 1)
      launchSfP<T, 2, MODE, MODE == kSfApply ? 8 : 4>(c, nb, kappa, cons, x, ybar, out, s);
    else
      launchSfP<T, 2, MODE, 1>(c, nb, kappa, cons, x, ybar, out, s);
  }
}

template <typename T, int MODE>
void launchGather(pxg_ctx * c, T const * buf, T const * x, T const * b, T const * dinv, double omega, T * y,
                  hipStream_t s)
{
  int64_t const n = c->ndofs;
  k_sf_gather<T, MODE><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, c->dofSeg.p, c->dofContrib.p, buf, x, b, dinv,
                                                                     T(omega), y);
  PXG_HIP(hipGetLastError());
}
} // namespace

bool sfAvailable(pxg_ctx const * c) { return sfOrder(c) != 0; }

void sfPrepare(pxg_ctx * c)
{
  PXG_REQUIRE(c->fe.feType && c->nelem && c->ndofs, PXG_E_STATE, "matrix-free needs set_fe, set_mesh, set_dofmap");
  PXG_REQUIRE(sfOrder(c), PXG_E_INVALID, "sum-factorised operator needs Q1 / Q2 hexahedra with the Gauss rule");
  if (!c->dofSeg.p)
    buildDofAdjacency(c);
}

// PXG_SF_CUBES=0 keeps batches on the element buffer and the per-dof gather
bool sfCubesOn()
{
  static bool const on = []
  {
    char const * e = std::getenv("PXG_SF_CUBES");
    return !(e && e[0] == '0');
  }();
  return on;
}

bool sfApplyCubes(pxg_ctx * c, double const * kappa, int nb, double const * x, double * y, bool useBC,
                  hipStream_t s)
{
  int const P = sfOrder(c);
  if (nb < 2 || !sfBatchOn() || !sfCubesReady(c))
    return false;
  PXG_REQUIRE((double)c->ndofs * nb < 4294967295., PXG_E_INVALID, "batch too large for 32-bit vector offsets");
  c->partBuf.resize(std::max<size_t>(c->partBuf.n, (size_t)c->nPart * nb));
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  if (P == 3)
    launchSfBatch<3, true>(c, nb, kappa, cons, x, y, c->partBuf.p, s);
  else
    launchSfBatch<2, true>(c, nb, kappa, cons, x, y, c->partBuf.p, s);
  int64_t const m = c->nPartDof * nb;
  if (m)
  {
    k_sf_part<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(c->nPartDof, nb, c->partDof.p, c->partSeg.p, c->partIds.p,
                                                          c->partBuf.p, y);
    PXG_HIP(hipGetLastError());
  }
  return true;
}

bool sfCubesReady(pxg_ctx * c)
{
  int const P = sfOrder(c);
  if (!P || !sfCubesOn())
    return false;
  sfPrepare(c);
  if (c->cubeState == 0)
    c->cubeState = (P == 3 ? buildCubes<3>(c) : buildCubes<2>(c)) ? 1 : -1;
  return c->cubeState > 0;
}

void sfMetricSlots(pxg_ctx * c, float const * met, float * metC, hipStream_t s)
{
  PXG_REQUIRE(c->cubeState > 0, PXG_E_STATE, "element cubes not built");
  int const n = c->fe.n;
  PXG_HIP(hipMemsetAsync(metC, 0, sizeof(float) * (size_t)c->nChunk * 6 * n * 64, s));
  int64_t const m = c->nelem * 6 * n;
  k_met_slots<float><<<(unsigned)((m + 255) / 256), 256, 0, s>>>(c->nelem, n, c->cubeSlotOf.p, met, metC);
  PXG_HIP(hipGetLastError());
}

void sfApplyPaCubes(pxg_ctx * c, int mode, float const * metC, float const * x, float const * b, float const * dinv,
                    double omega, float * y, float * part, bool useBC, hipStream_t s)
{
  PXG_REQUIRE(c->cubeState > 0, PXG_E_STATE, "element cubes not built");
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  unsigned const g = (unsigned)((c->nChunk + 3) / 4);
  int64_t const npd = c->nPartDofF;
  unsigned const gp = (unsigned)((npd + 255) / 256);
  auto go = [&](auto pC, auto mC)
  {
    constexpr int P = decltype(pC)::value, MODE = decltype(mC)::value;
    k_sf_pa_cube<P, MODE><<<g, 256, 0, s>>>(c->nChunk, c->cubeIdC.p, c->cubeCmC.p, cons, metC, x, b, dinv,
                                            (float)omega, y, part);
    if (npd)
      k_sf_part_f<MODE><<<gp, 256, 0, s>>>(npd, c->partDofF.p, c->partSegF.p, c->partIdsF.p, part, x, b, dinv,
                                           (float)omega, y);
  };
  auto pick = [&](auto pC)
  {
    switch (mode)
    {
    case kGApply: go(pC, std::integral_constant<int, kGApply>{}); break;
    case kGResidual: go(pC, std::integral_constant<int, kGResidual>{}); break;
    case kGJacobi: go(pC, std::integral_constant<int, kGJacobi>{}); break;
    default: throw Error(PXG_E_INVALID, "cube operator mode");
    }
  };
  if (sfOrder(c) == 3)
    pick(std::integral_constant<int, 3>{});
  else
    pick(std::integral_constant<int, 2>{});
  PXG_HIP(hipGetLastError());
}

template <typename T>
void sfElementApply(pxg_ctx * c, double const * kappa, int nb, T const * x, T * buf, bool useBC, hipStream_t s)
{
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  launchSf<T, kSfApply>(c, sfOrder(c), nb, kappa, cons, x, nullptr, buf, s);
}

void sfElementGrad(pxg_ctx * c, int nb, double const * x, double const * ybar, double * buf, bool useBC,
                   hipStream_t s)
{
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  launchSf<double, kSfGrad>(c, sfOrder(c), nb, nullptr, cons, x, ybar, buf, s);
}

template <typename T>
void sfElementDiag(pxg_ctx * c, double const * kappa, T * buf, hipStream_t s)
{
  uint8_t const * cons = c->hasBC ? c->bcFlag.p : nullptr;
  launchSf<T, kSfDiag>(c, sfOrder(c), 1, kappa, cons, nullptr, nullptr, buf, s);
}

template <typename T>
void sfMetric(pxg_ctx * c, double const * kappa, T * met, hipStream_t s)
{
  launchSf<T, kSfMetric>(c, sfOrder(c), 1, kappa, nullptr, nullptr, nullptr, met, s);
}

template <typename T>
void sfApplyPA(pxg_ctx * c, T const * met, T const * x, T * buf, bool useBC, hipStream_t s)
{
  uint8_t const * cons = (useBC && c->hasBC) ? c->bcFlag.p : nullptr;
  int64_t const E = c->nelem;
  int dev = 0, ncu = 0, perCu = 0;
  PXG_HIP(hipGetDevice(&dev));
  PXG_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  auto go = [&](auto pC)
  {
    constexpr int P = decltype(pC)::value;
    constexpr int EPW = 64 / (P * P);
    auto kern = k_sf_apply_pa<T, P>;
    PXG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kern, 64 * kSfWaves, 0));
    int64_t const nbatch = (E + EPW - 1) / EPW;
    int64_t const need = (nbatch + kSfWaves - 1) / kSfWaves;
    int64_t const grid = std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)ncu * std::max(perCu, 1)));
    kern<<<(unsigned)grid, 64 * kSfWaves, 0, s>>>(E, tabFor<P>(P == 3 ? PXG_FE_HEX_Q2 : PXG_FE_HEX_Q1),
                                                  c->elemMap.p, cons, met, x, buf);
  };
  if (sfOrder(c) == 3)
    go(std::integral_constant<int, 3>{});
  else
    go(std::integral_constant<int, 2>{});
  PXG_HIP(hipGetLastError());
}

template <typename T>
void sfGather(pxg_ctx * c, int mode, T const * buf, T const * x, T const * b, T const * dinv, double omega, T * y,
              hipStream_t s)
{
  switch (mode)
  {
  case kGApply: launchGather<T, kGApply>(c, buf, x, b, dinv, omega, y, s); break;
  case kGResidual: launchGather<T, kGResidual>(c, buf, x, b, dinv, omega, y, s); break;
  case kGJacobi: launchGather<T, kGJacobi>(c, buf, x, b, dinv, omega, y, s); break;
  default: launchGather<T, kGInverse>(c, buf, x, b, dinv, omega, y, s); break;
  }
}

template void sfElementApply<double>(pxg_ctx *, double const *, int, double const *, double *, bool, hipStream_t);
template void sfElementApply<float>(pxg_ctx *, double const *, int, float const *, float *, bool, hipStream_t);
template void sfElementDiag<double>(pxg_ctx *, double const *, double *, hipStream_t);
template void sfElementDiag<float>(pxg_ctx *, double const *, float *, hipStream_t);
template void sfMetric<double>(pxg_ctx *, double const *, double *, hipStream_t);
template void sfMetric<float>(pxg_ctx *, double const *, float *, hipStream_t);
template void sfApplyPA<double>(pxg_ctx *, double const *, double const *, double *, bool, hipStream_t);
template void sfApplyPA<float>(pxg_ctx *, float const *, float const *, float *, bool, hipStream_t);
template void sfGather<double>(pxg_ctx *, int, double const *, double const *, double const *, double const *,
                               double, double *, hipStream_t);
template void sfGather<float>(pxg_ctx *, int, float const *, float const *, float const *, float const *, double,
                              float *, hipStream_t);

} // namespace pxg
