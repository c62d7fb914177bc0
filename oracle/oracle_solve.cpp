// Oracle (test infrastructure only): sparse kernels, MG-PCG, matrix-free operator.
#include "oracle_solve.hpp"

#include <stdexcept>

#include <algorithm>
#include <cmath>

namespace orc
{

void spmv(Csr const & A, double const * x, double * y)
{
  for (i64 r = 0; r < A.nrows; ++r)
  {
    double s = 0.;
    for (i64 k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k)
      s += A.val[k] * x[A.col[k]];
    y[r] = s;
  }
}

void jacobiSweep(Csr const & A, std::vector<double> const & dinv, double omega, double const * b, double * x)
{
  std::vector<double> ax(A.nrows);
  spmv(A, x, ax.data());
  for (i64 r = 0; r < A.nrows; ++r)
    x[r] += omega * dinv[r] * (b[r] - ax[r]);
}

static double dot(i64 n, double const * a, double const * b)
{
  double s = 0.;
  for (i64 i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

static Csr fromTriplets(i64 nrows, std::vector<u32> const & r, std::vector<u32> const & c, std::vector<double> const & v)
{
  Csr A;
  A.nrows = nrows;
  A.rowptr.assign(nrows + 1, 0);
  for (u32 x: r)
    A.rowptr[x + 1]++;
  for (i64 i = 0; i < nrows; ++i)
    A.rowptr[i + 1] += A.rowptr[i];
  A.col.resize(r.size());
  A.val.resize(r.size());
  std::vector<i64> pos(A.rowptr.begin(), A.rowptr.end() - 1);
  for (size_t k = 0; k < r.size(); ++k)
  {
    A.col[pos[r[k]]] = c[k];
    A.val[pos[r[k]]++] = v[k];
  }
  for (i64 i = 0; i < nrows; ++i)
  {
    std::vector<std::pair<u32, double>> row;
    for (i64 k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
      row.push_back({A.col[k], A.val[k]});
    std::stable_sort(row.begin(), row.end(), [](auto const & a, auto const & b) { return a.first < b.first; });
    for (size_t k = 0; k < row.size(); ++k)
    {
      A.col[A.rowptr[i] + k] = row[k].first;
      A.val[A.rowptr[i] + k] = row[k].second;
    }
  }
  return A;
}

static inline double snapHalf(double x)
{
  double const t = std::round(2. * x);
  return std::fabs(2. * x - t) < 1e-8 ? 0.5 * t : x;
}

Csr prolongationHex(
    Mesh const & fine, DofMap const & df, u32 nxf, u32 nyf, u32 nzf, Mesh const & coarse,
    DofMap const & dc, RefFE const & fe)
{
  (void)nzf;
  u32 const nxc = nxf / 2, nyc = nyf / 2;
  int const n = fe.n;
  std::vector<u32> rr, cc;
  std::vector<double> vv;
  std::set<std::pair<u32, u32>> done;
  std::vector<double> vf(3 * fine.npe), vc(3 * coarse.npe), pf(3 * n);
  for (i64 ef = 0; ef < fine.nelem(); ++ef)
  {
    u32 const i = ef % nxf, j = (ef / nxf) % nyf, k = ef / ((i64)nxf * nyf);
    i64 const ec = (i / 2) + (i64)nxc * ((j / 2) + (i64)nyc * (k / 2));
    for (int p = 0; p < fine.npe; ++p)
      for (int d = 0; d < 3; ++d)
        vf[3 * p + d] = fine.coords[3ull * fine.conn[ef * fine.npe + p] + d];
    for (int p = 0; p < coarse.npe; ++p)
      for (int d = 0; d < 3; ++d)
        vc[3 * p + d] = coarse.coords[3ull * coarse.conn[ec * coarse.npe + p] + d];
    fe.dofPts(vf.data(), pf.data());
    // affine map of the coarse parallelepiped: x = mid + J xi
    double mid[3], J[3][3] = {{0}};
    fe.origin(vc.data(), mid);
    double const zero[3] = {0., 0., 0.};
    RefFE const & q1 = refFE(FE_HEX_Q1);
    for (int v = 0; v < 8; ++v)
    {
      double g[3];
      q1.dphi(v, zero, g);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          J[r][c] += vc[3 * v + r] * g[c];
    }
    // invert J (3x3)
    double const det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double I[3][3];
    I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int iF = 0; iF < n; ++iF)
    {
      double xi[3];
      for (int r = 0; r < 3; ++r)
      {
        double s = 0.;
        for (int c = 0; c < 3; ++c)
          s += I[r][c] * (pf[3 * iF + c] - mid[c]);
        xi[r] = snapHalf(s);
      }
      u32 const dofF = df.elemMap[ef * n + iF];
      for (int iC = 0; iC < n; ++iC)
      {
        u32 const dofC = dc.elemMap[ec * n + iC];
        if (done.count({dofF, dofC}))
          continue;
        done.insert({dofF, dofC});
        double const v = fe.phi(iC, xi);
        if (std::fabs(v) > 1e-12)
        {
          rr.push_back(dofF);
          cc.push_back(dofC);
          vv.push_back(v);
        }
      }
    }
  }
  return fromTriplets(df.ndofs, rr, cc, vv);
}

Csr transpose(Csr const & A, i64 ncols)
{
  std::vector<u32> r, c;
  std::vector<double> v;
  for (i64 i = 0; i < A.nrows; ++i)
    for (i64 k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
    {
      r.push_back(A.col[k]);
      c.push_back((u32)i);
      v.push_back(A.val[k]);
    }
  return fromTriplets(ncols, r, c, v);
}

Csr rowNormalised(Csr const & A)
{
  Csr B = A;
  for (i64 i = 0; i < B.nrows; ++i)
  {
    double s = 0.;
    for (i64 k = B.rowptr[i]; k < B.rowptr[i + 1]; ++k)
      s += B.val[k];
    for (i64 k = B.rowptr[i]; k < B.rowptr[i + 1]; ++k)
      B.val[k] /= s;
  }
  return B;
}

void Multigrid::factorCoarse()
{
  coarseL.clear();
  Csr const & A = lv.back().A;
  i64 const n = A.nrows;
  if (n > kDenseCoarse)
    return;
  std::vector<double> L(n * n, 0.);
  for (i64 r = 0; r < n; ++r)
    for (i64 k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k)
      if ((i64)A.col[k] <= r)
        L[r * n + A.col[k]] = A.val[k];
  // right-looking Cholesky, A = L L^T
  for (i64 j = 0; j < n; ++j)
  {
    double d = L[j * n + j];
    if (!(d > 0.))
      throw std::runtime_error("coarsest matrix is not SPD");
    d = std::sqrt(d);
    L[j * n + j] = d;
    for (i64 i = j + 1; i < n; ++i)
      L[i * n + j] /= d;
    for (i64 k = j + 1; k < n; ++k)
    {
      double const lkj = L[k * n + j];
      if (lkj != 0.)
        for (i64 i = k; i < n; ++i)
          L[i * n + k] -= L[i * n + j] * lkj;
    }
  }
  coarseL = std::move(L);
}

void Multigrid::vcycle(int l, double const * b, double * x) const
{
  MgLevel const & L = lv[l];
  i64 const n = L.A.nrows;
  if (l == (int)lv.size() - 1 && !coarseL.empty())
  {
    // forward and back substitution with the Cholesky factor
    for (i64 i = 0; i < n; ++i)
    {
      double t = b[i];
      for (i64 k = 0; k < i; ++k)
        t -= coarseL[i * n + k] * x[k];
      x[i] = t / coarseL[i * n + i];
    }
    for (i64 i = n - 1; i >= 0; --i)
    {
      double t = x[i];
      for (i64 k = i + 1; k < n; ++k)
        t -= coarseL[k * n + i] * x[k];
      x[i] = t / coarseL[i * n + i];
    }
    return;
  }
  if (l == (int)lv.size() - 1)
  {
    // coarsest level: Jacobi-preconditioned CG from zero to coarseRtol
    std::vector<double> r(b, b + n), z(n), p(n), ap(n);
    std::fill(x, x + n, 0.);
    double const bn = std::sqrt(dot(n, b, b));
    if (bn == 0.)
      return;
    for (i64 i = 0; i < n; ++i)
      z[i] = L.dinv[i] * r[i];
    p = z;
    double rz = dot(n, r.data(), z.data());
    for (int it = 0; it < coarseMaxit; ++it)
    {
      spmv(L.A, p.data(), ap.data());
      double const alpha = rz / dot(n, p.data(), ap.data());
      for (i64 i = 0; i < n; ++i)
      {
        x[i] += alpha * p[i];
        r[i] -= alpha * ap[i];
      }
      if (std::sqrt(dot(n, r.data(), r.data())) <= coarseRtol * bn)
        break;
      for (i64 i = 0; i < n; ++i)
        z[i] = L.dinv[i] * r[i];
      double const rz2 = dot(n, r.data(), z.data());
      double const beta = rz2 / rz;
      rz = rz2;
      for (i64 i = 0; i < n; ++i)
        p[i] = z[i] + beta * p[i];
    }
    return;
  }
  std::fill(x, x + n, 0.);
  for (int s = 0; s < nu1; ++s)
    jacobiSweep(L.A, L.dinv, omega, b, x);
  std::vector<double> r(n);
  spmv(L.A, x, r.data());
  for (i64 i = 0; i < n; ++i)
    r[i] = b[i] - r[i];
  i64 const nc = lv[l + 1].A.nrows;
  std::vector<double> bc(nc), ec(nc);
  spmv(L.R, r.data(), bc.data());
  vcycle(l + 1, bc.data(), ec.data());
  std::vector<double> pe(n);
  spmv(L.P, ec.data(), pe.data());
  for (i64 i = 0; i < n; ++i)
    x[i] += pe[i];
  for (int s = 0; s < nu2; ++s)
    jacobiSweep(L.A, L.dinv, omega, b, x);
}

int Multigrid::pcg(double const * b, double * x, double rtol, int maxit, double & relres) const
{
  Csr const & A = lv[0].A;
  i64 const n = A.nrows;
  std::vector<double> r(n), z(n), p(n), ap(n);
  spmv(A, x, ap.data());
  for (i64 i = 0; i < n; ++i)
    r[i] = b[i] - ap[i];
  double const bn = std::sqrt(dot(n, b, b));
  relres = std::sqrt(dot(n, r.data(), r.data())) / bn;
  if (relres <= rtol)
    return 0;
  vcycle(0, r.data(), z.data());
  p = z;
  double rz = dot(n, r.data(), z.data());
  int it = 0;
  while (it < maxit)
  {
    spmv(A, p.data(), ap.data());
    double const alpha = rz / dot(n, p.data(), ap.data());
    for (i64 i = 0; i < n; ++i)
    {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
    }
    ++it;
    relres = std::sqrt(dot(n, r.data(), r.data())) / bn;
    if (relres <= rtol)
      break;
    vcycle(0, r.data(), z.data());
    double const rz2 = dot(n, r.data(), z.data());
    double const beta = rz2 / rz;
    rz = rz2;
    for (i64 i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  return it;
}

static void elemVtx(Mesh const & m, i64 e, std::vector<double> & vtx)
{
  vtx.resize(3 * m.npe);
  for (int p = 0; p < m.npe; ++p)
    for (int k = 0; k < 3; ++k)
      vtx[3 * p + k] = m.coords[3ull * m.conn[e * m.npe + p] + k];
}

void mfApply(
    Mesh const & m, CurFE & cfe, DofMap const & d, double const * kappa,
    uint8_t const * cons, int nb, double const * x, double * y)
{
  int const n = cfe.n, nq = cfe.nq;
  std::fill(y, y + (size_t)d.ndofs * nb, 0.);
  std::vector<double> vtx, Ke(n * n), xe(n * nb);
  for (i64 e = 0; e < m.nelem(); ++e)
  {
    elemVtx(m, e, vtx);
    cfe.reinit(vtx.data());
    u32 const * ids = &d.elemMap[e * n];
    std::fill(Ke.begin(), Ke.end(), 0.);
    for (int q = 0; q < nq; ++q)
    {
      double kq = 0.;
      for (int i = 0; i < n; ++i)
        kq += cfe.phiRef[q * n + i] * kappa[ids[i]];
      double const s = kq * cfe.JxW[q];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
        {
          double const * gi = &cfe.dphi[(q * n + i) * 3];
          double const * gj = &cfe.dphi[(q * n + j) * 3];
          Ke[i * n + j] += s * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
        }
    }
    for (int j = 0; j < n; ++j)
      for (int b = 0; b < nb; ++b)
        xe[j * nb + b] = (cons && cons[ids[j]]) ? 0. : x[(size_t)ids[j] * nb + b];
    for (int i = 0; i < n; ++i)
    {
      bool const ci = cons && cons[ids[i]];
      for (int b = 0; b < nb; ++b)
      {
        double s = 0.;
        if (!ci)
          for (int j = 0; j < n; ++j)
            s += Ke[i * n + j] * xe[j * nb + b];
        else
          s = x[(size_t)ids[i] * nb + b]; // diagonal 1 per element
        y[(size_t)ids[i] * nb + b] += s;
      }
    }
  }
}

void mfGradKappa(
    Mesh const & m, CurFE & cfe, DofMap const & d, uint8_t const * cons, int nb,
    double const * x, double const * ybar, double * gk)
{
  int const n = cfe.n, nq = cfe.nq;
  std::fill(gk, gk + d.ndofs, 0.);
  std::vector<double> vtx;
  for (i64 e = 0; e < m.nelem(); ++e)
  {
    elemVtx(m, e, vtx);
    cfe.reinit(vtx.data());
    u32 const * ids = &d.elemMap[e * n];
    for (int q = 0; q < nq; ++q)
    {
      double s = 0.;
      for (int b = 0; b < nb; ++b)
      {
        double gx[3] = {0, 0, 0}, gy[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j)
        {
          if (cons && cons[ids[j]])
            continue;
          double const xv = x[(size_t)ids[j] * nb + b], yv = ybar[(size_t)ids[j] * nb + b];
          for (int k = 0; k < 3; ++k)
          {
            gx[k] += cfe.dphi[(q * n + j) * 3 + k] * xv;
            gy[k] += cfe.dphi[(q * n + j) * 3 + k] * yv;
          }
        }
        s += gx[0] * gy[0] + gx[1] * gy[1] + gx[2] * gy[2];
      }
      s *= cfe.JxW[q];
      for (int i = 0; i < n; ++i)
        gk[ids[i]] += cfe.phiRef[q * n + i] * s;
    }
  }
}

} // namespace orc
