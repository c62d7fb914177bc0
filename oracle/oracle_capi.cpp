// Oracle C-ABI (test infrastructure only): lets the Python tests and the CPU
// baseline drive the CPU restatement.  The product library never calls this.
#include "oracle_fe.hpp"
#include "oracle_solve.hpp"

#include <cstring>
#include <memory>
#include <string>

using namespace orc;

namespace
{
struct Ctx
{
  int feType;
  Mesh mesh;
  DofMap dofs;
  QRule qr;
  bool customQR = false;
  Builder builder;
  Csr A;
  std::vector<uint8_t> bcFlag;
  std::string err;
  u32 nx = 0, ny = 0, nz = 0;
};

template <typename F>
int guard(Ctx * c, F && f)
{
  try
  {
    return f();
  }
  catch (std::exception const & e)
  {
    if (c)
      c->err = e.what();
    return -1;
  }
}
} // namespace

extern "C" {

void * orc_new(int feType)
{
  try
  {
    refFE(feType);
    auto * c = new Ctx;
    c->feType = feType;
    c->qr = recommendedRule(feType);
    return c;
  }
  catch (...)
  {
    return nullptr;
  }
}

void orc_free(void * p) { delete static_cast<Ctx *>(p); }

char const * orc_error(void * p) { return static_cast<Ctx *>(p)->err.c_str(); }

// kind: 1 line, 2 square (quads), 3 cube (hexes), 4 cube (5 tets/cube), 5 cube (6 tets/cube)
int orc_mesh(void * p, int kind, double const * origin, double const * length, int nx, int ny, int nz)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        switch (kind)
        {
        case 1: buildLine(c->mesh, origin, length, nx); break;
        case 2: buildSquare(c->mesh, origin, length, nx, ny); break;
        case 3: buildCubeHex(c->mesh, origin, length, nx, ny, nz); break;
        case 4: buildCubeTet5(c->mesh, origin, length, nx, ny, nz); break;
        case 5: buildCubeTet6(c->mesh, origin, length, nx, ny, nz); break;
        default: throw std::runtime_error("unknown mesh kind");
        }
        c->nx = nx;
        c->ny = ny;
        c->nz = nz;
        if (c->mesh.geo != refFE(c->feType).geo)
          throw std::runtime_error("mesh element does not match the FE");
        c->dofs = buildDofMap(c->mesh, refFE(c->feType));
        return 0;
      });
}

// mesh of a GMSH 2.2 file (geo: 1 line, 2 quad, 3 hex, 4 tet)
int orc_mesh_gmsh(void * p, char const * path, int geo)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        readGmsh(c->mesh, path, geo);
        if (c->mesh.geo != refFE(c->feType).geo)
          throw std::runtime_error("mesh element does not match the FE");
        c->dofs = buildDofMap(c->mesh, refFE(c->feType));
        return 0;
      });
}

// overwrite point coordinates (e.g. perturbed interior vertices)
int orc_set_coords(void * p, double const * xyz)
{
  auto * c = static_cast<Ctx *>(p);
  std::memcpy(c->mesh.coords.data(), xyz, c->mesh.coords.size() * sizeof(double));
  return 0;
}

int orc_set_qr(void * p, int nq, double const * nodes, double const * weights)
{
  auto * c = static_cast<Ctx *>(p);
  int const dim = refFE(c->feType).dim;
  c->qr.dim = dim;
  c->qr.nq = nq;
  c->qr.node.assign(nodes, nodes + nq * dim);
  c->qr.weight.assign(weights, weights + nq);
  c->customQR = true;
  return 0;
}

// threads of the Builder's element loop and close (1: the reference's serial loop)
int orc_set_threads(void * p, int n)
{
  static_cast<Ctx *>(p)->builder.threads = n < 1 ? 1 : n;
  return 0;
}

// 0 npts, 1 nelem, 2 points/elem, 3 ndofs, 4 local dofs, 5 quadrature points,
// 6 nnz (closed matrix), 7 boundary facets
int64_t orc_info(void * p, int what)
{
  auto * c = static_cast<Ctx *>(p);
  switch (what)
  {
  case 0: return c->mesh.npts();
  case 1: return c->mesh.nelem();
  case 2: return c->mesh.npe;
  case 3: return c->dofs.ndofs;
  case 4: return refFE(c->feType).n;
  case 5: return c->qr.nq;
  case 6: return (int64_t)c->A.col.size();
  case 7: return (int64_t)c->mesh.bfacets.size();
  }
  return -1;
}

int orc_get_mesh(void * p, double * coords, uint32_t * conn)
{
  auto * c = static_cast<Ctx *>(p);
  if (coords)
    std::memcpy(coords, c->mesh.coords.data(), c->mesh.coords.size() * sizeof(double));
  if (conn)
    std::memcpy(conn, c->mesh.conn.data(), c->mesh.conn.size() * sizeof(uint32_t));
  return 0;
}

// boundary facets: per facet (elem, side, marker)
int orc_get_bfacets(void * p, int32_t * out)
{
  auto * c = static_cast<Ctx *>(p);
  for (size_t k = 0; k < c->mesh.bfacets.size(); ++k)
  {
    out[3 * k] = (int32_t)c->mesh.bfacets[k].elem;
    out[3 * k + 1] = c->mesh.bfacets[k].side;
    out[3 * k + 2] = c->mesh.bfacets[k].marker;
  }
  return 0;
}

int orc_get_elem_map(void * p, uint32_t * out)
{
  auto * c = static_cast<Ctx *>(p);
  std::memcpy(out, c->dofs.elemMap.data(), c->dofs.elemMap.size() * sizeof(uint32_t));
  return 0;
}

int orc_get_dof_coords(void * p, double * out)
{
  auto * c = static_cast<Ctx *>(p);
  auto v = dofCoords(c->mesh, refFE(c->feType), c->dofs);
  std::memcpy(out, v.data(), v.size() * sizeof(double));
  return 0;
}

int64_t orc_bc_dofs(void * p, int marker, uint32_t * out)
{
  auto * c = static_cast<Ctx *>(p);
  auto v = constrainedDofs(c->mesh, refFE(c->feType), c->dofs, marker);
  if (out)
    std::memcpy(out, v.data(), v.size() * sizeof(uint32_t));
  return (int64_t)v.size();
}

// per element and quadrature point: jac (3 x dim), detJ, JxW, qpoint (3),
// dphi (n x 3); element range [e0, e1)
int orc_geometry(
    void * p, int64_t e0, int64_t e1, double * jac, double * detJ, double * JxW,
    double * qpt, double * dphi)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        RefFE const & fe = refFE(c->feType);
        CurFE cfe(fe, c->qr);
        int const nq = cfe.nq, dim = cfe.dim, n = cfe.n;
        std::vector<double> vtx(3 * c->mesh.npe);
        for (int64_t e = e0; e < e1; ++e)
        {
          for (int k = 0; k < c->mesh.npe; ++k)
            for (int d = 0; d < 3; ++d)
              vtx[3 * k + d] = c->mesh.coords[3ull * c->mesh.conn[e * c->mesh.npe + k] + d];
          cfe.reinit(vtx.data());
          int64_t const o = e - e0;
          std::memcpy(jac + o * nq * 3 * dim, cfe.jac.data(), nq * 3 * dim * sizeof(double));
          std::memcpy(detJ + o * nq, cfe.detJ.data(), nq * sizeof(double));
          std::memcpy(JxW + o * nq, cfe.JxW.data(), nq * sizeof(double));
          std::memcpy(qpt + o * nq * 3, cfe.qpt.data(), nq * 3 * sizeof(double));
          std::memcpy(dphi + o * nq * n * 3, cfe.dphi.data(), nq * n * 3 * sizeof(double));
        }
        return 0;
      });
}

// Builder sequence: init, buildLhs(terms, bcs), buildRhs(rhs, bcs), closeMatrix.
// terms: kinds[k] in {1 stiffness, 2 stiffness FE, 3 mass, 4 mass FE}, coef[k];
// FE terms read `kappa` (nodal values, ndofs).  bcFlag/bcVal: per dof (may be null).
// rhsKind 0 skips buildRhs.  Returns 0 or -1.
int orc_assemble(
    void * p, int nterms, int const * kinds, double const * coefs, double const * kappa,
    uint8_t const * bcFlag, double const * bcVal, int rhsKind, double const * rhsParams,
    double const * qvalues, double pruneEps)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        RefFE const & fe = refFE(c->feType);
        CurFE cfe(fe, c->qr);
        std::vector<Term> terms;
        for (int k = 0; k < nterms; ++k)
          terms.push_back(Term{kinds[k], coefs[k], kappa});
        std::vector<uint8_t> flag;
        std::vector<double> val;
        if (bcFlag)
        {
          flag.assign(bcFlag, bcFlag + c->dofs.ndofs);
          val.assign(bcVal, bcVal + c->dofs.ndofs);
        }
        c->bcFlag = flag;
        c->builder.init(c->dofs.ndofs);
        if (nterms > 0)
          c->builder.buildLhs(c->mesh, cfe, c->dofs, terms, flag, val);
        if (rhsKind != RHS_NONE)
          c->builder.buildRhs(c->mesh, cfe, c->dofs, rhsKind, rhsParams, qvalues, flag);
        c->A = c->builder.close(pruneEps);
        c->builder.trow.clear();
        c->builder.tcol.clear();
        c->builder.tval.clear();
        c->builder.trow.shrink_to_fit();
        c->builder.tcol.shrink_to_fit();
        c->builder.tval.shrink_to_fit();
        return 0;
      });
}

int orc_get_csr(void * p, int64_t * rowptr, uint32_t * col, double * val, double * b)
{
  auto * c = static_cast<Ctx *>(p);
  if (rowptr)
    std::memcpy(rowptr, c->A.rowptr.data(), c->A.rowptr.size() * sizeof(int64_t));
  if (col)
    std::memcpy(col, c->A.col.data(), c->A.col.size() * sizeof(uint32_t));
  if (val)
    std::memcpy(val, c->A.val.data(), c->A.val.size() * sizeof(double));
  if (b)
    std::memcpy(b, c->builder.b.data(), c->builder.b.size() * sizeof(double));
  return 0;
}

// Stokes block system of the reference's test/assemblyvector.cpp,
// test/stokes2dquad.cpp, test/stokes3dhex.cpp and test/stokes3dtet.cpp on the
// mesh of a QUAD_Q2, HEX_Q2 or TET_P2 context: velocity with dim components
// (BLOCK or INTERLEAVED dofs) at [0, dim nU), pressure in the linear space of
// the same geometry (Q1 / P1) at [dim nU, dim nU + nP), both with the
// context's rule.  flag / val: Dirichlet data per dof of the whole system
// (global numbering).  mode 0: vector spaces,
//   buildLhs({AssemblyStiffness{cS, vel}, AssemblyTensorStiffness{cT, vel},
//             AssemblyScalarMass{cM, vel}}, bcsVel)
//   [+ buildLhs(AssemblyDummy{p}, bcsP) when some pressure dof is constrained]
//   + buildCoupling(AssemblyGrad{cG, vel, p}, bcsVel, bcsP)
//   + buildCoupling(AssemblyDiv{cD, p, vel}, bcsP, bcsVel);
// mode 1: the scalar component spaces of assemblyvector.cpp, comp lists {d}
// (no tensor term: it couples the components).  Terms with a zero
// coefficient are left out.  The closed matrix and b replace the context's
// (orc_get_csr).
int orc_stokes(void * p, int interleaved, double cS, double cT, double cM, double cG, double cD,
               uint8_t const * flagIn, double const * valIn, int mode, double pruneEps, int64_t * ndofsOut,
               int64_t * nnzOut)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        int pType = 0;
        switch (c->feType)
        {
        case FE_QUAD_Q2: pType = FE_QUAD_Q1; break;
        case FE_HEX_Q2: pType = FE_HEX_Q1; break;
        case FE_TET_P2: pType = FE_TET_P1; break;
        default: throw std::runtime_error("orc_stokes needs a QUAD_Q2, HEX_Q2 or TET_P2 context");
        }
        if (mode != 0 && cT != 0.)
          throw std::runtime_error("the tensor stiffness has no component-space form");
        RefFE const & fv = refFE(c->feType);
        RefFE const & fp = refFE(pType);
        int const dim = fv.dim;
        CurFE cv(fv, c->qr), cp(fp, c->qr);
        DofMap const & sv = c->dofs;
        DofMap dp = buildDofMap(c->mesh, fp);
        u32 const nU = sv.ndofs, nP = dp.ndofs, off = dim * nU;
        for (auto & id: dp.elemMap)
          id += off;
        dp.ndofs = nP;
        i64 const nd = (i64)off + nP;
        std::vector<uint8_t> flag(flagIn, flagIn + nd);
        std::vector<double> val(valIn, valIn + nd);
        bool pinned = false;
        for (i64 k = off; k < nd; ++k)
          pinned = pinned || flag[k];
        std::vector<Term> dummy; // AssemblyDummy: no term, Ke = 0
        std::vector<int> all;
        for (int d = 0; d < dim; ++d)
          all.push_back(d);
        c->builder.init(nd);
        if (mode == 0)
        {
          DofMap vel = vectorDofMap(sv, dim, interleaved != 0, 0);
          c->builder.buildLhsVector(c->mesh, cv, vel, dim, cS, cT, cM, flag, val);
          if (pinned)
            c->builder.buildLhs(c->mesh, cp, dp, dummy, flag, val);
          c->builder.buildCoupling(c->mesh, cv, cp, vel, dim, dp, 1, COUPLING_GRAD, cG, all, flag, flag, val);
          c->builder.buildCoupling(c->mesh, cp, cv, dp, 1, vel, dim, COUPLING_DIV, cD, all, flag, flag, val);
        }
        else
        {
          std::vector<Term> lhs;
          if (cS != 0.)
            lhs.push_back(Term{TERM_STIFFNESS, cS, nullptr});
          if (cM != 0.)
            lhs.push_back(Term{TERM_MASS, cM, nullptr});
          for (int d = 0; d < dim; ++d)
          {
            DofMap u = sv;
            for (size_t k = 0; k < u.elemMap.size(); ++k)
            {
              u32 const id = sv.elemMap[k];
              u.elemMap[k] = interleaved ? dim * id + d : id + d * nU;
            }
            c->builder.buildLhs(c->mesh, cv, u, lhs, flag, val);
            if (pinned && d == 0)
              c->builder.buildLhs(c->mesh, cp, dp, dummy, flag, val);
            c->builder.buildCoupling(c->mesh, cv, cp, u, 1, dp, 1, COUPLING_GRAD, cG, {d}, flag, flag, val);
            c->builder.buildCoupling(c->mesh, cp, cv, dp, 1, u, 1, COUPLING_DIV, cD, {d}, flag, flag, val);
          }
        }
        c->A = c->builder.close(pruneEps);
        c->builder.trow.clear();
        c->builder.tcol.clear();
        c->builder.tval.clear();
        if (ndofsOut)
          *ndofsOut = nd;
        if (nnzOut)
          *nnzOut = (int64_t)c->A.col.size();
        return 0;
      });
}

// structural pattern (before value pruning) and the element->nnz map;
// returns nnz, arrays filled when non-null (second call)
int64_t orc_pattern(void * p, int64_t * rowptr, uint32_t * col, uint32_t * emap)
{
  auto * c = static_cast<Ctx *>(p);
  std::vector<int64_t> rp;
  std::vector<uint32_t> cl, em;
  structuralPattern(c->dofs, c->mesh.nelem(), rp, cl, em);
  if (rowptr)
    std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int64_t));
  if (col)
    std::memcpy(col, cl.data(), cl.size() * sizeof(uint32_t));
  if (emap)
    std::memcpy(emap, em.data(), em.size() * sizeof(uint32_t));
  return (int64_t)cl.size();
}

} // extern "C"

// ---------------------------------------------------------------------------
// sparse kernels and multigrid-preconditioned CG
extern "C" {

// y = A x with the closed matrix of a context
int orc_spmv(void * p, double const * x, double * y)
{
  auto * c = static_cast<Ctx *>(p);
  spmv(c->A, x, y);
  return 0;
}

// `nsweeps` weighted-Jacobi sweeps on the closed matrix, in place on x
int orc_jacobi(void * p, double omega, int nsweeps, double const * b, double * x)
{
  auto * c = static_cast<Ctx *>(p);
  std::vector<double> dinv(c->A.nrows, 0.);
  for (int64_t r = 0; r < c->A.nrows; ++r)
    for (int64_t k = c->A.rowptr[r]; k < c->A.rowptr[r + 1]; ++k)
      if (c->A.col[k] == (uint32_t)r)
        dinv[r] = 1. / c->A.val[k];
  for (int s = 0; s < nsweeps; ++s)
    jacobiSweep(c->A, dinv, omega, b, x);
  return 0;
}

struct MgCtx
{
  Multigrid mg;
};

// levels: assembled contexts, finest first, structured hex meshes with halved
// element counts per level.  Prolongation rows of constrained fine dofs and
// columns of constrained coarse dofs are removed; R = P^T.
void * orc_mg_new(int nlev, void ** ctxs, double omega, int nu1, int nu2, double coarseRtol)
{
  try
  {
    auto * m = new MgCtx;
    m->mg.omega = omega;
    m->mg.nu1 = nu1;
    m->mg.nu2 = nu2;
    m->mg.coarseRtol = coarseRtol;
    m->mg.lv.resize(nlev);
    for (int l = 0; l < nlev; ++l)
    {
      auto * c = static_cast<Ctx *>(ctxs[l]);
      MgLevel & L = m->mg.lv[l];
      L.A = c->A;
      L.cons = c->bcFlag;
      L.dinv.assign(L.A.nrows, 0.);
      for (int64_t r = 0; r < L.A.nrows; ++r)
        for (int64_t k = L.A.rowptr[r]; k < L.A.rowptr[r + 1]; ++k)
          if (L.A.col[k] == (uint32_t)r)
            L.dinv[r] = 1. / L.A.val[k];
      if (l + 1 < nlev)
      {
        auto * cc = static_cast<Ctx *>(ctxs[l + 1]);
        Csr P = prolongationHex(
            c->mesh, c->dofs, c->nx, c->ny, c->nz, cc->mesh, cc->dofs, refFE(c->feType));
        // drop constrained rows / columns
        std::vector<uint8_t> const & fc = c->bcFlag;
        std::vector<uint8_t> const & ccons = cc->bcFlag;
        Csr Q;
        Q.nrows = P.nrows;
        Q.rowptr.assign(P.nrows + 1, 0);
        for (int64_t r = 0; r < P.nrows; ++r)
        {
          if (fc.empty() || !fc[r])
            for (int64_t k = P.rowptr[r]; k < P.rowptr[r + 1]; ++k)
              if (ccons.empty() || !ccons[P.col[k]])
              {
                Q.col.push_back(P.col[k]);
                Q.val.push_back(P.val[k]);
              }
          Q.rowptr[r + 1] = (int64_t)Q.col.size();
        }
        L.P = Q;
        L.R = transpose(Q, cc->dofs.ndofs);
      }
    }
    m->mg.factorCoarse();
    return m;
  }
  catch (...)
  {
    return nullptr;
  }
}

void orc_mg_free(void * p) { delete static_cast<MgCtx *>(p); }

int orc_mg_vcycle(void * p, double const * b, double * x)
{
  auto * m = static_cast<MgCtx *>(p);
  m->mg.vcycle(0, b, x);
  return 0;
}

int orc_mg_pcg(void * p, double const * b, double * x, double rtol, int maxit, int * iters, double * relres)
{
  auto * m = static_cast<MgCtx *>(p);
  double rr = 0.;
  *iters = m->mg.pcg(b, x, rtol, maxit, rr);
  *relres = rr;
  return 0;
}

// prolongation P_l (from level l+1 to l): returns nnz; arrays filled when non-null
int64_t orc_mg_prolongation(void * p, int l, int64_t * rowptr, uint32_t * col, double * val)
{
  auto * m = static_cast<MgCtx *>(p);
  Csr const & P = m->mg.lv[l].P;
  if (rowptr)
    std::memcpy(rowptr, P.rowptr.data(), P.rowptr.size() * sizeof(int64_t));
  if (col)
    std::memcpy(col, P.col.data(), P.col.size() * sizeof(uint32_t));
  if (val)
    std::memcpy(val, P.val.data(), P.val.size() * sizeof(double));
  return (int64_t)P.col.size();
}

// reference-style transfer operators between two structured hex contexts
// (no constrained-dof removal): which 0 = P, 1 = row-normalised P^T
int64_t orc_transfer(void * pf, void * pc, int which, int64_t * rowptr, uint32_t * col, double * val)
{
  auto * f = static_cast<Ctx *>(pf);
  auto * c = static_cast<Ctx *>(pc);
  Csr P = prolongationHex(f->mesh, f->dofs, f->nx, f->ny, f->nz, c->mesh, c->dofs, refFE(f->feType));
  Csr T = which == 0 ? P : rowNormalised(transpose(P, c->dofs.ndofs));
  if (rowptr)
    std::memcpy(rowptr, T.rowptr.data(), T.rowptr.size() * sizeof(int64_t));
  if (col)
    std::memcpy(col, T.col.data(), T.col.size() * sizeof(uint32_t));
  if (val)
    std::memcpy(val, T.val.data(), T.val.size() * sizeof(double));
  return (int64_t)T.col.size();
}

// matrix-free y = K(kappa) x, nb interleaved vectors; cons may be null
int orc_mf_apply(void * p, double const * kappa, uint8_t const * cons, int nb, double const * x, double * y)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        CurFE cfe(refFE(c->feType), c->qr);
        mfApply(c->mesh, cfe, c->dofs, kappa, cons, nb, x, y);
        return 0;
      });
}

int orc_mf_grad_kappa(
    void * p, uint8_t const * cons, int nb, double const * x, double const * ybar, double * gk)
{
  auto * c = static_cast<Ctx *>(p);
  return guard(
      c,
      [&]
      {
        CurFE cfe(refFE(c->feType), c->qr);
        mfGradKappa(c->mesh, cfe, c->dofs, cons, nb, x, ybar, gk);
        return 0;
      });
}

} // extern "C"
