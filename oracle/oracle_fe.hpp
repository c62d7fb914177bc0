// Oracle: clean-room CPU restatement of the proxpde assembly path.
//
// TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP product path
// is compared against; only tests/, __graft_entry__.smoke() and the CPU-baseline
// leg of bench.py may load it.  The product library never links or calls it.
//
// Each function names the reference file:line whose behaviour it restates.
// Nothing here is copied from the reference; the tables are re-derived from the
// element definitions (vertex order, basis products) and the published Gauss rules.
#pragma once

#include <string>

#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <stdexcept>
#include <vector>

namespace orc
{
using u32 = uint32_t;
using i64 = int64_t;

// ---------------------------------------------------------------------------
// element / FE identifiers (shared numeric values with the product C-ABI)
enum FeType : int
{
  FE_LINE_P1 = 1,
  FE_QUAD_Q1 = 2,
  FE_QUAD_Q2 = 3,
  FE_HEX_Q1 = 4,
  FE_HEX_Q2 = 5,
  FE_TET_P1 = 6,
  FE_TET_P2 = 7,
};

enum GeoKind : int
{
  GEO_LINE = 1,
  GEO_QUAD = 2,
  GEO_HEX = 3,
  GEO_TET = 4,
};

// boundary markers, reference include/mesh.hpp:178-188
enum Side : int
{
  SIDE_NONE = 0,
  SIDE_BOTTOM = 1,
  SIDE_RIGHT = 2,
  SIDE_TOP = 3,
  SIDE_LEFT = 4,
  SIDE_BACK = 5,
  SIDE_FRONT = 6,
};

// ---------------------------------------------------------------------------
// geometric element topology, reference include/geo.hpp:185-232 (Line),
// :380-420 (Quad), :508-535 (Tetrahedron), :626-662 (Hexahedron)
struct GeoTopo
{
  int kind;
  int dim;
  int npts;
  int nfacets, facet_npts;
  std::vector<std::vector<int>> facets;
  int nedges;
  std::vector<std::array<int, 2>> edges;
  int nfaces, face_npts;
  std::vector<std::vector<int>> faces;
};

GeoTopo const & geoTopo(int kind);

// ---------------------------------------------------------------------------
// reference finite elements, reference include/reffe.hpp + src/reffe.cpp
struct RefFE
{
  int type;
  int geo;       // GeoKind
  int dim;       // reference dimension
  int n;         // local dofs
  int ng;        // mapping points (isoparametric: == n)
  int dofPlace[4]; // {element, face, edge, point}
  int dofPerFacet;
  std::vector<std::vector<int>> dofOnFacet;
  // basis value and reference gradient
  double phi(int i, double const * xi) const;
  void dphi(int i, double const * xi, double * g) const;
  // dof points (== mapping points) from the element vertex coordinates
  void dofPts(double const * vtx /*npts x 3*/, double * out /* n x 3 */) const;
  void origin(double const * vtx, double * o) const;
};

RefFE const & refFE(int type);

// ---------------------------------------------------------------------------
// quadrature rule (nodes in reference coordinates, weights)
struct QRule
{
  int dim;
  int nq;
  std::vector<double> node;   // nq x dim
  std::vector<double> weight; // nq
};

// Gauss rules, reference src/qr.cpp:27-307; `recommended` follows include/fe.hpp:30-132
QRule gaussRule(int geo, int nq);
QRule recommendedRule(int feType);

// ---------------------------------------------------------------------------
struct Facet
{
  std::vector<u32> pts;
  u32 elem;
  int side;
  bool boundary;
  int marker;
};

struct Mesh
{
  int geo = 0;
  int npe = 0; // points per element
  std::vector<double> coords; // npts x 3
  std::vector<u32> conn;      // nelem x npe
  std::vector<Facet> bfacets; // boundary facets in reference storage order
  i64 npts() const { return (i64)coords.size() / 3; }
  i64 nelem() const { return npe ? (i64)conn.size() / npe : 0; }
};

// structured builders, reference src/mesh.cpp:17-46 (line), :124-175 (square),
// :355-454 (cube, 5 tets), :456-535 (cube, hex); the 6-tet Kuhn split has no
// reference counterpart (all cubes split along the 0-6 diagonal)
void buildLine(Mesh & m, double const * origin, double const * length, u32 n);
void buildSquare(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny);
void buildCubeHex(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz);
void buildCubeTet5(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz);
void buildCubeTet6(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz);
// buildFacets + markFacetsCube, reference include/mesh.hpp:226-341, :443-481
void buildBoundaryFacets(Mesh & m, double const * origin, double const * length);
// readGMSH, reference include/mesh.hpp:785-1062 (format 2.2 ASCII): volume
// elements of geometry `geo` (quadratic ones through their vertices), used
// points renumbered in file order, boundary facets marked by the first tag of
// the file's facet element on the same vertices (0 if none)
void readGmsh(Mesh & m, std::string const & path, int geo);

// ---------------------------------------------------------------------------
struct DofMap
{
  int n = 0;
  u32 ndofs = 0;
  std::vector<u32> elemMap; // nelem x n
};

// first-visit numbering, reference include/dof.hpp:112-329
DofMap buildDofMap(Mesh const & m, RefFE const & fe);

// dofs constrained by facets carrying `marker`, reference include/bc.hpp:253-295
std::vector<u32> constrainedDofs(Mesh const & m, RefFE const & fe, DofMap const & d, int marker);

// dof coordinates (interpolation points), reference include/fespace.hpp:195-238
std::vector<double> dofCoords(Mesh const & m, RefFE const & fe, DofMap const & d);

// ---------------------------------------------------------------------------
// CurFE::reinit, reference include/curfe.hpp:50-103
struct CurFE
{
  RefFE const * fe;
  QRule qr;
  int n, nq, dim;
  std::vector<double> phiRef;  // nq x n
  std::vector<double> dphiRef; // nq x n x dim
  std::vector<double> mapG;    // nq x ng x dim
  // per element
  std::vector<double> jac;   // nq x 3 x dim
  std::vector<double> detJ;  // nq
  std::vector<double> JxW;   // nq
  std::vector<double> qpt;   // nq x 3
  std::vector<double> dphi;  // nq x n x 3
  std::vector<double> mpts;  // n x 3
  CurFE(RefFE const & f, QRule const & q);
  void reinit(double const * vtx);
};

// ---------------------------------------------------------------------------
// element terms, reference include/assembly_lhs.hpp:36-113 (stiffness),
// :191-283 (mass), include/assembly_base.hpp:10-19 (ScalarCoef)
enum TermKind : int
{
  TERM_STIFFNESS = 1,    // coef * (grad phi_j, grad phi_i)
  TERM_STIFFNESS_FE = 2, // (kappa_h grad phi_j, grad phi_i)
  TERM_MASS = 3,         // coef * (phi_j, phi_i)
  TERM_MASS_FE = 4,      // coef * (kappa_h phi_j, phi_i)
};

struct Term
{
  int kind;
  double coef;
  double const * kappa; // nodal field on the same space (FE terms)
};

// analytic right-hand sides used by the reference tests and the workloads
enum RhsKind : int
{
  RHS_NONE = 0,
  RHS_CONST = 1,      // p0
  RHS_SINSIN = 2,     // p0 * sin(p1 x) * sin(p2 y)
  RHS_SINSINSIN = 3,  // p0 * sin(p1 x) * sin(p2 y) * sin(p3 z)
  RHS_FEVAR1D = 4,    // sin(p1 x) - p0 * cos(p2 x)
  RHS_QVALUES = 5,    // values given per (element, quadrature point)
  RHS_PROJECTION = 6, // AssemblyProjection(p0, f_h): nodal values f_h of the same FE space
};

double evalRhs(int kind, double const * p, double const * x);

struct Csr
{
  i64 nrows = 0;
  std::vector<i64> rowptr;
  std::vector<u32> col;
  std::vector<double> val;
};

// Builder::buildLhs / buildRhs / closeMatrix, reference include/builder.hpp:31-154,
// :279-364, :366-371
struct Builder
{
  i64 ndofs = 0;
  int threads = 1; // > 1: element loop and close in parallel (same results)
  std::vector<double> b;
  std::vector<u32> trow, tcol; // triplets, reference order
  std::vector<double> tval;
  void init(i64 n)
  {
    ndofs = n;
    b.assign(n, 0.0);
    trow.clear();
    tcol.clear();
    tval.clear();
  }
  void buildLhs(
      Mesh const & m, CurFE & cfe, DofMap const & d, std::vector<Term> const & terms,
      std::vector<uint8_t> const & bcFlag, std::vector<double> const & bcVal);
  void buildRhs(
      Mesh const & m, CurFE & cfe, DofMap const & d, int rhsKind, double const * params,
      double const * qvalues, std::vector<uint8_t> const & bcFlag);
  // setFromTriplets (sum duplicates in insertion order, sorted columns) + prune
  Csr close(double pruneEps) const;
  Csr closeThreaded(double pruneEps) const;
  // buildLhs({AssemblyStiffness(cS), AssemblyTensorStiffness(cT),
  // AssemblyScalarMass(cM)}) on a space of `dim` components (reference
  // builder.hpp:31-154, assembly_lhs.hpp:36-75, 115-173, 190-229; a zero
  // coefficient leaves its term out): the map has n*dim columns, local dof i
  // of component d at i + d*n
  void buildLhsVector(
      Mesh const & m, CurFE & cfe, DofMap const & d, int dim, double cS, double cT, double cM,
      std::vector<uint8_t> const & bcFlag, std::vector<double> const & bcVal);
  // buildCoupling(AssemblyGrad / AssemblyDiv, bcsRow, bcsClm), reference
  // builder.hpp:156-277, assembly_coupling.hpp:42-132: rows of the row space
  // (dimR components, map nR*dimR wide), columns of the column space; comps:
  // the physical derivative of each component block (the CompList)
  void buildCoupling(
      Mesh const & m, CurFE & row, CurFE & clm, DofMap const & dr, int dimR, DofMap const & dc, int dimC, int kind,
      double coef, std::vector<int> const & comps, std::vector<uint8_t> const & bcRowFlag,
      std::vector<uint8_t> const & bcClmFlag, std::vector<double> const & bcClmVal);
};

enum CouplingKind : int
{
  COUPLING_GRAD = 1, // coef * (phiP_j, d phiV_i / dx_c): rows velocity, columns pressure
  COUPLING_DIV = 2,  // coef * (d phiV_j / dx_c, phiP_i): rows pressure, columns velocity
};

// vector dof map from a scalar one, reference include/dof.hpp:112-329:
// BLOCK id + d*size, INTERLEAVED id*dim + d, plus the space offset
DofMap vectorDofMap(DofMap const & s, int dim, bool interleaved, u32 offset);

// structural pattern and element->nnz map (position of each (e,i,j) in the
// unpruned pattern); the reference never materialises this map, it is the
// bookkeeping of setFromTriplets made explicit
void structuralPattern(
    DofMap const & d, i64 nelem, std::vector<i64> & rowptr, std::vector<u32> & col,
    std::vector<u32> & emap);

} // namespace orc
