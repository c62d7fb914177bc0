// The Builder façade (include/proxpde_amd_builder.hpp) driven with the
// reference's call shape, checked against the CPU oracle (oracle/liboracle.so,
// the test suite's checker) and against a known answer of the reference's own
// tests.  Needs a GPU; exit status 0 when every case passes.
//
//   case 1: test/poisson3dhex_q2.cpp (n = 4): AssemblyStiffness(1) +
//           AssemblyRhsAnalytic, u = 0 on LEFT and BOTTOM; the error norm of
//           the solved system against the exact solution is the reference's
//           expected_error 0.0070042836287997.
//   case 2: Q2 hex 3^3, FEVar kappa: AssemblyStiffnessFE(kappa) +
//           AssemblyMassFE(0.5, kappa) + AssemblyRhsAnalytic, non-zero
//           values on all six sides (later BCEss win on shared dofs), against
//           the oracle's Builder on the same data.
//   case 3: Q1 hex 6^3, AssemblyStiffness(2) + AssemblyScalarMass(1), u = g on
//           LEFT and FRONT: matrix bitwise the oracle's.
//   case 4: P2 tet (5 per cube) 2^3 without Dirichlet data, two buildLhs
//           calls accumulated before one closeMatrix, against one oracle
//           build with both terms.
#include "proxpde_amd_builder.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

extern "C" {
void * orc_new(int feType);
void orc_free(void * p);
char const * orc_error(void * p);
int orc_mesh(void * p, int kind, double const * origin, double const * length, int nx, int ny, int nz);
int64_t orc_info(void * p, int what);
int orc_get_elem_map(void * p, uint32_t * out);
int64_t orc_bc_dofs(void * p, int marker, uint32_t * out);
int orc_geometry(void * p, int64_t e0, int64_t e1, double * jac, double * detJ, double * JxW, double * qpt,
                 double * dphi);
int orc_assemble(void * p, int nterms, int const * kinds, double const * coefs, double const * kappa,
                 uint8_t const * bcFlag, double const * bcVal, int rhsKind, double const * rhsParams,
                 double const * qvalues, double pruneEps);
int orc_get_csr(void * p, int64_t * rowptr, uint32_t * col, double * val, double * b);
}

using namespace proxpde_amd;

namespace
{
int failures = 0;

void expect(bool ok, std::string const & what)
{
  std::printf("  %s: %s\n", ok ? "ok  " : "FAIL", what.c_str());
  std::fflush(stdout);
  if (!ok)
    ++failures;
}

struct OracleCsr
{
  std::vector<int64_t> rowptr;
  std::vector<uint32_t> col;
  std::vector<double> val, b;
};

// the oracle's Builder on the space's mesh kind and size; qvalues at the
// oracle's own quadrature points
struct Oracle
{
  void * p;
  int64_t ndofs, nelem, nq;
  Oracle(int feType, int meshKind, int nx, int ny, int nz)
  {
    p = orc_new(feType);
    double const o[3] = {0., 0., 0.}, l[3] = {1., 1., 1.};
    if (!p || orc_mesh(p, meshKind, o, l, nx, ny, nz) != 0)
      throw std::runtime_error("oracle mesh");
    ndofs = orc_info(p, 3);
    nelem = orc_info(p, 1);
    nq = orc_info(p, 5);
  }
  ~Oracle() { orc_free(p); }
  std::vector<double> qvalues(scalarFun_T const & f) const
  {
    int const n = (int)orc_info(p, 4), dim = 3;
    std::vector<double> jac(nelem * nq * 3 * dim), detJ(nelem * nq), JxW(nelem * nq), qpt(nelem * nq * 3),
        dphi(nelem * nq * n * 3), out(nelem * nq);
    orc_geometry(p, 0, nelem, jac.data(), detJ.data(), JxW.data(), qpt.data(), dphi.data());
    for (size_t k = 0; k < out.size(); ++k)
      out[k] = f(Vec3{qpt[3 * k], qpt[3 * k + 1], qpt[3 * k + 2]});
    return out;
  }
  OracleCsr assemble(std::vector<int> const & kinds, std::vector<double> const & coefs, double const * kappa,
                     std::vector<uint8_t> const * flag, std::vector<double> const * val,
                     std::vector<double> const * qv) const
  {
    if (orc_assemble(p, (int)kinds.size(), kinds.data(), coefs.data(), kappa, flag ? flag->data() : nullptr,
                     val ? val->data() : nullptr, qv ? 5 : 0, nullptr, qv ? qv->data() : nullptr, 1e-16) != 0)
      throw std::runtime_error(std::string("oracle assemble: ") + orc_error(p));
    OracleCsr r;
    int64_t const nnz = orc_info(p, 6);
    r.rowptr.resize(ndofs + 1);
    r.col.resize(nnz);
    r.val.resize(nnz);
    r.b.resize(ndofs);
    orc_get_csr(p, r.rowptr.data(), r.col.data(), r.val.data(), r.b.data());
    return r;
  }
};

// same pattern; values within tol of the row's largest oracle entry; b
// within tol of its largest entry (bitwise: tol = 0)
void compare(CsrMatrix const & A, std::vector<double> const & b, OracleCsr const & o, double tol, char const * tag)
{
  bool same = A.rowptr == o.rowptr && A.col.size() == o.col.size();
  for (size_t k = 0; same && k < A.col.size(); ++k)
    same = (uint32_t)A.col[k] == o.col[k];
  expect(same, std::string(tag) + ": pattern equals the oracle's (" + std::to_string(A.val.size()) + " nnz)");
  if (!same)
    return;
  double worst = 0.;
  bool bitwise = true;
  for (int64_t i = 0; i < A.rows; ++i)
  {
    double m = 0.;
    for (int64_t k = o.rowptr[i]; k < o.rowptr[i + 1]; ++k)
      m = std::max(m, std::fabs(o.val[k]));
    for (int64_t k = o.rowptr[i]; k < o.rowptr[i + 1]; ++k)
    {
      bitwise = bitwise && std::memcmp(&A.val[k], &o.val[k], sizeof(double)) == 0;
      worst = std::max(worst, std::fabs(A.val[k] - o.val[k]) / (m > 0. ? m : 1.));
    }
  }
  if (tol == 0.)
    expect(bitwise, std::string(tag) + ": values bitwise the oracle's");
  else
    expect(worst <= tol, std::string(tag) + ": values within " + std::to_string(tol) +
                             " of the row maximum (worst " + std::to_string(worst) + ")");
  double bm = 0., be = 0.;
  for (size_t i = 0; i < b.size(); ++i)
    bm = std::max(bm, std::fabs(o.b[i]));
  for (size_t i = 0; i < b.size(); ++i)
    be = std::max(be, std::fabs(b[i] - o.b[i]));
  double const btol = tol == 0. ? 1e-13 : tol;
  expect(be <= btol * (bm > 0. ? bm : 1.), std::string(tag) + ": b within " + std::to_string(btol) +
                                               " of its maximum (" + std::to_string(be) + ")");
}

// Jacobi-preconditioned CG on the host CSR (the matrix after elimination is
// symmetric positive definite)
std::vector<double> solveCg(CsrMatrix const & A, std::vector<double> const & b)
{
  int64_t const n = A.rows;
  std::vector<double> x(n, 0.), r = b, z(n), p(n), q(n), d(n, 1.);
  for (int64_t i = 0; i < n; ++i)
    d[i] = 1. / A.coeff(i, i);
  auto dot = [&](std::vector<double> const & u, std::vector<double> const & v)
  {
    double s = 0.;
    for (int64_t i = 0; i < n; ++i)
      s += u[i] * v[i];
    return s;
  };
  double const bn = std::sqrt(dot(b, b));
  for (int64_t i = 0; i < n; ++i)
    p[i] = z[i] = d[i] * r[i];
  double rz = dot(r, z);
  for (int it = 0; it < 10 * n && std::sqrt(dot(r, r)) > 1e-15 * bn; ++it)
  {
    for (int64_t i = 0; i < n; ++i)
    {
      double s = 0.;
      for (int64_t k = A.rowptr[i]; k < A.rowptr[i + 1]; ++k)
        s += A.val[k] * p[A.col[k]];
      q[i] = s;
    }
    double const al = rz / dot(p, q);
    for (int64_t i = 0; i < n; ++i)
    {
      x[i] += al * p[i];
      r[i] -= al * q[i];
      z[i] = d[i] * r[i];
    }
    double const rz1 = dot(r, z);
    for (int64_t i = 0; i < n; ++i)
      p[i] = z[i] + (rz1 / rz) * p[i];
    rz = rz1;
  }
  return x;
}

template <typename FE>
std::vector<uint8_t> flags(FE const & fe, std::vector<BCEss<FE>> const & bcs, std::vector<double> & val)
{
  std::vector<uint8_t> f(fe.dof.size, 0);
  val.assign(fe.dof.size, 0.);
  for (auto const & bc: bcs)
    for (size_t k = 0; k < bc.dofs.size(); ++k)
    {
      f[bc.dofs[k]] = 1;
      val[bc.dofs[k]] = bc.values[k];
    }
  return f;
}

template <typename FE>
void checkSpace(FE const & fe, Oracle const & o, std::vector<BCEss<FE>> const & bcs, char const * tag)
{
  std::vector<uint32_t> em(fe.dof.elemMap.size());
  orc_get_elem_map(o.p, em.data());
  expect((int64_t)fe.dof.size == o.ndofs && em == fe.dof.elemMap, std::string(tag) + ": dof map equals the oracle's");
  bool same = true;
  for (auto const & bc: bcs)
  {
    std::vector<uint32_t> od(fe.dof.size);
    int64_t const m = orc_bc_dofs(o.p, bc.marker, od.data());
    od.resize(m);
    same = same && od == bc.dofs;
  }
  expect(same, std::string(tag) + ": BCEss dofs equal the oracle's");
}

void case1()
{
  std::puts("case 1: poisson3dhex_q2, n = 4 (reference test/poisson3dhex_q2.cpp)");
  using Elem_T = Hexahedron;
  using Mesh_T = Mesh<Elem_T>;
  using FESpace_T = FESpace<Mesh_T, LagrangeFE<Elem_T, 2>::RefFE_T, LagrangeFE<Elem_T, 2>::RecommendedQR>;
  scalarFun_T const rhs = [](Vec3 const & p)
  { return 2.5 * M_PI * M_PI * std::sin(0.5 * M_PI * p[0]) * std::sin(1.5 * M_PI * p[1]); };
  scalarFun_T const exactSol = [](Vec3 const & p) { return std::sin(0.5 * M_PI * p[0]) * std::sin(1.5 * M_PI * p[1]); };
  uint const n = 4;
  Mesh_T mesh;
  buildHyperCube(mesh, {0., 0., 0.}, {1., 1., 1.}, {n, n, n});
  FESpace_T feSpace{mesh};
  auto const bcLeft = BCEss{feSpace, side::LEFT, [](Vec3 const &) { return 0.; }};
  auto const bcBottom = BCEss{feSpace, side::BOTTOM, [](Vec3 const &) { return 0.; }};
  auto const bcs = std::vector{bcLeft, bcBottom};
  Builder<StorageType::RowMajor> builder{feSpace.dof.size};
  builder.buildLhs(std::tuple{AssemblyStiffness(1.0, feSpace)}, bcs);
  builder.buildRhs(std::tuple{AssemblyRhsAnalytic(rhs, feSpace)}, bcs);
  builder.closeMatrix();
  auto const sol = solveCg(builder.A, builder.b);
  std::vector<double> exact;
  interpolateAnalyticFunction(exactSol, feSpace, exact);
  double e2 = 0.;
  for (size_t i = 0; i < sol.size(); ++i)
    e2 += (sol[i] - exact[i]) * (sol[i] - exact[i]);
  double const norm = std::sqrt(e2), expected = 0.0070042836287997;
  std::printf("  the norm of the error is %.16g (reference expected_error %.16g)\n", norm, expected);
  expect(std::fabs(norm - expected) <= 1e-10 * expected, "error norm matches the reference's known answer");

  Oracle o(PXG_FE_HEX_Q2, PXH_MESH_CUBE_HEX, n, n, n);
  checkSpace(feSpace, o, bcs, "case 1");
  std::vector<double> val;
  auto const flag = flags(feSpace, bcs, val);
  auto const qv = o.qvalues(rhs);
  auto const ref = o.assemble({PXG_TERM_STIFFNESS}, {1.0}, nullptr, &flag, &val, &qv);
  compare(builder.A, builder.b, ref, 1e-12, "case 1");
}

void case2()
{
  std::puts("case 2: Q2 hex 3^3, AssemblyStiffnessFE(kappa) + AssemblyMassFE(0.5, kappa), all sides");
  using Elem_T = Hexahedron;
  using FESpace_T = FESpace<Mesh<Elem_T>, LagrangeFE<Elem_T, 2>::RefFE_T, LagrangeFE<Elem_T, 2>::RecommendedQR>;
  Mesh<Elem_T> mesh;
  buildHyperCube(mesh, {0., 0., 0.}, {1., 1., 1.}, {3, 3, 3});
  FESpace_T feSpace{mesh};
  FEVar kappa{"kappa", feSpace};
  kappa << [](Vec3 const & p) { return std::exp(std::sin(3. * p[0]) * std::cos(2. * p[1]) + 0.5 * p[2]); };
  scalarFun_T const g = [](Vec3 const & p) { return 1. + p[0] + 2. * p[1] - 0.5 * p[2]; };
  scalarFun_T const f = [](Vec3 const & p) { return std::cos(p[0]) + p[1] * p[2]; };
  std::vector<BCEss<FESpace_T>> bcs;
  for (marker_T m: {side::BOTTOM, side::RIGHT, side::TOP, side::LEFT, side::BACK, side::FRONT})
    bcs.emplace_back(feSpace, m, [&g, m](Vec3 const & p) { return g(p) + 0.01 * m; });
  Builder<StorageType::RowMajor> builder{feSpace.dof.size};
  builder.buildLhs(std::tuple{AssemblyStiffnessFE(kappa, feSpace), AssemblyMassFE(0.5, kappa, feSpace)}, bcs);
  builder.buildRhs(std::tuple{AssemblyRhsAnalytic(f, feSpace)}, bcs);
  builder.closeMatrix();
  Oracle o(PXG_FE_HEX_Q2, PXH_MESH_CUBE_HEX, 3, 3, 3);
  checkSpace(feSpace, o, bcs, "case 2");
  std::vector<double> val;
  auto const flag = flags(feSpace, bcs, val);
  auto const qv = o.qvalues(f);
  auto const ref = o.assemble({PXG_TERM_STIFFNESS_FE, PXG_TERM_MASS_FE}, {1.0, 0.5}, kappa.data.data(), &flag, &val,
                              &qv);
  compare(builder.A, builder.b, ref, 1e-12, "case 2");
}

void case3()
{
  std::puts("case 3: Q1 hex 6^3, AssemblyStiffness(2) + AssemblyScalarMass(1), LEFT and FRONT");
  using Elem_T = Hexahedron;
  using FESpace_T = FESpace<Mesh<Elem_T>, LagrangeFE<Elem_T, 1>::RefFE_T, LagrangeFE<Elem_T, 1>::RecommendedQR>;
  Mesh<Elem_T> mesh;
  buildHyperCube(mesh, {0., 0., 0.}, {1., 1., 1.}, {6, 6, 6});
  FESpace_T feSpace{mesh};
  auto const bcs = std::vector{BCEss{feSpace, side::LEFT, [](Vec3 const & p) { return p[1]; }},
                               BCEss{feSpace, side::FRONT, [](Vec3 const & p) { return 1. - p[0]; }}};
  Builder<StorageType::RowMajor> builder{feSpace.dof.size};
  builder.buildLhs(std::tuple{AssemblyStiffness(2.0, feSpace), AssemblyScalarMass(1.0, feSpace)}, bcs);
  builder.closeMatrix();
  Oracle o(PXG_FE_HEX_Q1, PXH_MESH_CUBE_HEX, 6, 6, 6);
  checkSpace(feSpace, o, bcs, "case 3");
  std::vector<double> val;
  auto const flag = flags(feSpace, bcs, val);
  auto const ref = o.assemble({PXG_TERM_STIFFNESS, PXG_TERM_MASS}, {2.0, 1.0}, nullptr, &flag, &val, nullptr);
  compare(builder.A, builder.b, ref, 0., "case 3");
}

void case4()
{
  std::puts("case 4: P2 tet 2^3 (5 per cube), buildLhs(stiffness) + buildLhs(mass), one closeMatrix");
  using Elem_T = Tetrahedron;
  using FESpace_T = FESpace<Mesh<Elem_T>, LagrangeFE<Elem_T, 2>::RefFE_T, LagrangeFE<Elem_T, 2>::RecommendedQR>;
  Mesh<Elem_T> mesh;
  buildHyperCube(mesh, {0., 0., 0.}, {1., 1., 1.}, {2, 2, 2});
  FESpace_T feSpace{mesh};
  scalarFun_T const f = [](Vec3 const & p) { return 1. + p[0] * p[1]; };
  Builder<StorageType::RowMajor> builder{feSpace.dof.size};
  builder.buildLhs(std::tuple{AssemblyStiffness(1.0, feSpace)});
  builder.buildLhs(std::tuple{AssemblyScalarMass(3.0, feSpace)});
  builder.buildRhs(std::tuple{AssemblyRhsAnalytic(f, feSpace)});
  builder.closeMatrix();
  Oracle o(PXG_FE_TET_P2, PXH_MESH_CUBE_TET5, 2, 2, 2);
  checkSpace(feSpace, o, std::vector<BCEss<FESpace_T>>{}, "case 4");
  auto const qv = o.qvalues(f);
  // one oracle build with both terms: the triplet sums of the two builds
  // equal it up to rounding
  auto const ref = o.assemble({PXG_TERM_STIFFNESS, PXG_TERM_MASS}, {1.0, 3.0}, nullptr, nullptr, nullptr, &qv);
  compare(builder.A, builder.b, ref, 1e-12, "case 4");
}
} // namespace

int main()
{
  try
  {
    case1();
    case2();
    case3();
    case4();
  }
  catch (std::exception const & e)
  {
    std::printf("FAIL: exception: %s\n", e.what());
    return 2;
  }
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "passed", failures);
  return failures ? 1 : 0;
}
