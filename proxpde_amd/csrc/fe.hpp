// Reference finite elements and quadrature rules of the product library.
//
// Node orders, facet-dof tables and Gauss rules follow the reference so that
// dof numbering and element matrices agree with it:
//   RefLineP1        include/reffe.hpp:96-136,  src/reffe.cpp:43-61
//   RefQuadQ1        include/reffe.hpp:443-483, src/reffe.cpp:246-282
//   RefQuadQ2        include/reffe.hpp:535-584, src/reffe.cpp:374-396
//   RefHexahedronQ1  include/reffe.hpp:877-921, src/reffe.cpp:684-759
//   RefHexahedronQ2  include/reffe.hpp:923-992, src/reffe.cpp:762-1003
//   RefTetrahedronP1/P2 include/reffe.hpp:680-776, src/reffe.cpp:505-642
//   GaussQR          src/qr.cpp:27-307 (3-point abscissa truncated as at :38)
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace pxg
{

enum Geo : int
{
  kLine = 1,
  kQuad = 2,
  kHex = 3,
  kTet = 4
};

struct GeoInfo
{
  int geo, dim, npts;
  std::vector<std::vector<int>> facets; // vertex lists, reference side order
  std::vector<std::array<int, 2>> edges;
  std::vector<std::vector<int>> faces;
};

GeoInfo const & geoInfo(int geo);

struct FeInfo
{
  int type, geo, dim, n;
  int dofPlace[4]; // {interior, face, edge, point}
  std::vector<std::vector<int>> dofOnFacet;
  // value and reference gradient of local basis i at reference point xi
  double phi(int i, double const * xi) const;
  void grad(int i, double const * xi, double * g) const;
  // dof points of an element from its vertex coordinates (npts x 3 -> n x 3)
  void dofPoints(double const * vtx, double * out) const;
};

FeInfo const & feInfo(int type);

struct Quadrature
{
  int dim = 0, nq = 0;
  std::vector<double> xi; // nq x dim
  std::vector<double> w;  // nq
};

Quadrature recommendedQuadrature(int feType);

// value / gradient of the vertex (geometric) basis: Q1 for line/quad/hex,
// P1 for tets; the reference's isoparametric maps of the Q2/P2 elements use
// mapping points that are vertex averages, so they coincide with this map.
double geoPhi(int geo, int v, double const * xi);
void geoGrad(int geo, int v, double const * xi, double * g);

} // namespace pxg
