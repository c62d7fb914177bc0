// Oracle (test infrastructure only): reference elements, quadrature, meshes,
// dof numbering, CurFE and Builder restated on the CPU.  See oracle_fe.hpp.
#include "oracle_fe.hpp"

#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <numeric>

namespace orc
{

// ===========================================================================
// topology tables (vertex numbering of each reference entity)
// Line   include/geo.hpp:185-232
// Quad   include/geo.hpp:380-420
// Tet    include/geo.hpp:508-535
// Hex    include/geo.hpp:626-662
static GeoTopo makeTopo(int kind)
{
  GeoTopo t{};
  t.kind = kind;
  switch (kind)
  {
  case GEO_LINE:
    t.dim = 1;
    t.npts = 2;
    t.nfacets = 2;
    t.facet_npts = 1;
    t.facets = {{0}, {1}};
    t.nedges = 1;
    t.edges = {{0, 1}};
    t.nfaces = 0;
    t.face_npts = 0;
    break;
  case GEO_QUAD:
    t.dim = 2;
    t.npts = 4;
    t.nfacets = 4;
    t.facet_npts = 2;
    t.facets = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    t.nedges = 4;
    t.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    t.nfaces = 1;
    t.face_npts = 4;
    t.faces = {{0, 1, 2, 3}};
    break;
  case GEO_TET:
    t.dim = 3;
    t.npts = 4;
    t.nfacets = 4;
    t.facet_npts = 3;
    t.facets = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
    t.nedges = 6;
    t.edges = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    t.nfaces = 4;
    t.face_npts = 3;
    t.faces = t.facets;
    break;
  case GEO_HEX:
    t.dim = 3;
    t.npts = 8;
    t.nfacets = 6;
    t.facet_npts = 4;
    t.facets = {{0, 3, 2, 1}, {3, 0, 4, 7}, {0, 1, 5, 4},
                {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
    t.nedges = 12;
    t.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
               {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    t.nfaces = 6;
    t.face_npts = 4;
    t.faces = t.facets;
    break;
  default:
    throw std::runtime_error("unknown geometry");
  }
  return t;
}

GeoTopo const & geoTopo(int kind)
{
  static GeoTopo const line = makeTopo(GEO_LINE);
  static GeoTopo const quad = makeTopo(GEO_QUAD);
  static GeoTopo const tet = makeTopo(GEO_TET);
  static GeoTopo const hex = makeTopo(GEO_HEX);
  switch (kind)
  {
  case GEO_LINE: return line;
  case GEO_QUAD: return quad;
  case GEO_TET: return tet;
  case GEO_HEX: return hex;
  }
  throw std::runtime_error("unknown geometry");
}

// ===========================================================================
// 1D Lagrange factors, reference src/reffe.cpp:246-253 (Q1), :374-383 (Q2)
static inline double q1f(int a, double x) { return a == 0 ? 0.5 * (1. - x) : 0.5 * (1. + x); }
static inline double q1d(int a, double) { return a == 0 ? -0.5 : 0.5; }
static inline double q2f(int a, double x)
{
  return a == 0 ? 0.5 * x * (x - 1.) : (a == 1 ? 0.5 * x * (x + 1.) : 1. - x * x);
}
static inline double q2d(int a, double x)
{
  return a == 0 ? x - 0.5 : (a == 1 ? x + 0.5 : -2. * x);
}

// tensor index of each local node: which 1D factor per direction
// (0: the -1 end, 1: the +1 end, 2: the midpoint bubble)
static int const quadQ1Idx[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
static int const quadQ2Idx[9][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}, {2, 0},
                                    {1, 2}, {2, 1}, {0, 2}, {2, 2}};
static int const hexQ1Idx[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
// reference src/reffe.cpp:762-799 (node order of RefHexahedronQ2)
static int const hexQ2Idx[27][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1},
    {0, 1, 1}, {2, 0, 0}, {1, 2, 0}, {2, 1, 0}, {0, 2, 0}, {0, 0, 2}, {1, 0, 2},
    {1, 1, 2}, {0, 1, 2}, {2, 0, 1}, {1, 2, 1}, {2, 1, 1}, {0, 2, 1}, {2, 2, 0},
    {0, 2, 2}, {2, 0, 2}, {1, 2, 2}, {2, 1, 2}, {2, 2, 1}, {2, 2, 2}};
// P2 tet edge nodes as ordered barycentric pairs, reference src/reffe.cpp:534-555
static int const tetP2Edge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

static inline double bary(int a, double const * p)
{
  switch (a)
  {
  case 0: return 1. - p[0] - p[1] - p[2];
  case 1: return p[0];
  case 2: return p[1];
  default: return p[2];
  }
}
static inline void dbary(int a, double * g)
{
  g[0] = a == 0 ? -1. : (a == 1 ? 1. : 0.);
  g[1] = a == 0 ? -1. : (a == 2 ? 1. : 0.);
  g[2] = a == 0 ? -1. : (a == 3 ? 1. : 0.);
}

double RefFE::phi(int i, double const * xi) const
{
  switch (type)
  {
  case FE_LINE_P1: return q1f(i, xi[0]);
  case FE_QUAD_Q1: return q1f(quadQ1Idx[i][0], xi[0]) * q1f(quadQ1Idx[i][1], xi[1]);
  case FE_QUAD_Q2: return q2f(quadQ2Idx[i][0], xi[0]) * q2f(quadQ2Idx[i][1], xi[1]);
  case FE_HEX_Q1:
    return q1f(hexQ1Idx[i][0], xi[0]) * q1f(hexQ1Idx[i][1], xi[1]) *
           q1f(hexQ1Idx[i][2], xi[2]);
  case FE_HEX_Q2:
    return q2f(hexQ2Idx[i][0], xi[0]) * q2f(hexQ2Idx[i][1], xi[1]) *
           q2f(hexQ2Idx[i][2], xi[2]);
  case FE_TET_P1: return bary(i, xi);
  case FE_TET_P2:
    if (i < 4)
    {
      double const z = bary(i, xi);
      return z * (2. * z - 1.);
    }
    return 4. * bary(tetP2Edge[i - 4][0], xi) * bary(tetP2Edge[i - 4][1], xi);
  }
  return 0.;
}

void RefFE::dphi(int i, double const * xi, double * g) const
{
  switch (type)
  {
  case FE_LINE_P1: g[0] = q1d(i, xi[0]); return;
  case FE_QUAD_Q1:
  {
    int a = quadQ1Idx[i][0], b = quadQ1Idx[i][1];
    g[0] = q1d(a, xi[0]) * q1f(b, xi[1]);
    g[1] = q1f(a, xi[0]) * q1d(b, xi[1]);
    return;
  }
  case FE_QUAD_Q2:
  {
    int a = quadQ2Idx[i][0], b = quadQ2Idx[i][1];
    g[0] = q2d(a, xi[0]) * q2f(b, xi[1]);
    g[1] = q2f(a, xi[0]) * q2d(b, xi[1]);
    return;
  }
  case FE_HEX_Q1:
  {
    int a = hexQ1Idx[i][0], b = hexQ1Idx[i][1], c = hexQ1Idx[i][2];
    g[0] = q1d(a, xi[0]) * q1f(b, xi[1]) * q1f(c, xi[2]);
    g[1] = q1f(a, xi[0]) * q1d(b, xi[1]) * q1f(c, xi[2]);
    g[2] = q1f(a, xi[0]) * q1f(b, xi[1]) * q1d(c, xi[2]);
    return;
  }
  case FE_HEX_Q2:
  {
    int a = hexQ2Idx[i][0], b = hexQ2Idx[i][1], c = hexQ2Idx[i][2];
    g[0] = q2d(a, xi[0]) * q2f(b, xi[1]) * q2f(c, xi[2]);
    g[1] = q2f(a, xi[0]) * q2d(b, xi[1]) * q2f(c, xi[2]);
    g[2] = q2f(a, xi[0]) * q2f(b, xi[1]) * q2d(c, xi[2]);
    return;
  }
  case FE_TET_P1: dbary(i, g); return;
  case FE_TET_P2:
    if (i < 4)
    {
      double d[3];
      dbary(i, d);
      double const z = bary(i, xi);
      for (int k = 0; k < 3; ++k)
        g[k] = (4. * z - 1.) * d[k];
    }
    else
    {
      int const a = tetP2Edge[i - 4][0], b = tetP2Edge[i - 4][1];
      double da[3], db[3];
      dbary(a, da);
      dbary(b, db);
      double const za = bary(a, xi), zb = bary(b, xi);
      for (int k = 0; k < 3; ++k)
        g[k] = 4. * (za * db[k] + da[k] * zb);
    }
    return;
  }
}

static inline void avg2(double const * v, int a, int b, double * o)
{
  for (int k = 0; k < 3; ++k)
    o[k] = 0.5 * (v[3 * a + k] + v[3 * b + k]);
}
static inline void avg4(double const * v, int a, int b, int c, int d, double * o)
{
  for (int k = 0; k < 3; ++k)
    o[k] = 0.25 * (v[3 * a + k] + v[3 * b + k] + v[3 * c + k] + v[3 * d + k]);
}
static inline void avg8(double const * v, double * o)
{
  for (int k = 0; k < 3; ++k)
  {
    double s = v[k];
    for (int p = 1; p < 8; ++p)
      s += v[3 * p + k];
    o[k] = .125 * s;
  }
}

// dof points, reference include/reffe.hpp:118-125 (line), :466-475 (quad Q1),
// :559-572 (quad Q2), :741-763 (tet P2), :897-911 (hex Q1), :949-980 (hex Q2)
void RefFE::dofPts(double const * v, double * o) const
{
  int const nv = geoTopo(geo).npts;
  for (int p = 0; p < nv && p < n; ++p)
    for (int k = 0; k < 3; ++k)
      o[3 * p + k] = v[3 * p + k];
  switch (type)
  {
  case FE_QUAD_Q2:
    avg2(v, 0, 1, o + 12);
    avg2(v, 1, 2, o + 15);
    avg2(v, 2, 3, o + 18);
    avg2(v, 3, 0, o + 21);
    avg4(v, 0, 1, 2, 3, o + 24);
    break;
  case FE_TET_P2:
    for (int e = 0; e < 6; ++e)
      avg2(v, tetP2Edge[e][0], tetP2Edge[e][1], o + 3 * (4 + e));
    break;
  case FE_HEX_Q2:
  {
    static int const ed[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5},
                                  {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    for (int e = 0; e < 12; ++e)
      avg2(v, ed[e][0], ed[e][1], o + 3 * (8 + e));
    avg4(v, 0, 3, 2, 1, o + 3 * 20);
    avg4(v, 3, 0, 4, 7, o + 3 * 21);
    avg4(v, 0, 1, 5, 4, o + 3 * 22);
    avg4(v, 1, 2, 6, 5, o + 3 * 23);
    avg4(v, 2, 3, 7, 6, o + 3 * 24);
    avg4(v, 4, 5, 6, 7, o + 3 * 25);
    avg8(v, o + 3 * 26);
    break;
  }
  default: break;
  }
}

// element origin used by the affine qpoint formula: midpoint for line/quad/hex,
// first vertex for tets (include/geo.hpp:230-232, :455, :588, :728)
void RefFE::origin(double const * v, double * o) const
{
  switch (geo)
  {
  case GEO_LINE:
    for (int k = 0; k < 3; ++k)
      o[k] = 0.5 * (v[3 + k] + v[k]);
    break;
  case GEO_QUAD:
    for (int k = 0; k < 3; ++k)
      o[k] = 0.25 * (v[k] + v[3 + k] + v[6 + k] + v[9 + k]);
    break;
  case GEO_HEX: avg8(v, o); break;
  case GEO_TET:
    for (int k = 0; k < 3; ++k)
      o[k] = v[k];
    break;
  }
}

static RefFE makeFE(int type)
{
  RefFE f{};
  f.type = type;
  auto set = [&](int geo, int dim, int n, std::array<int, 4> place, int dpf)
  {
    f.geo = geo;
    f.dim = dim;
    f.n = n;
    f.ng = n;
    for (int k = 0; k < 4; ++k)
      f.dofPlace[k] = place[k];
    f.dofPerFacet = dpf;
  };
  switch (type)
  {
  case FE_LINE_P1:
    set(GEO_LINE, 1, 2, {0, 0, 0, 1}, 1);
    f.dofOnFacet = {{0}, {1}};
    break;
  case FE_QUAD_Q1:
    set(GEO_QUAD, 2, 4, {0, 0, 0, 1}, 2);
    f.dofOnFacet = geoTopo(GEO_QUAD).facets;
    break;
  case FE_QUAD_Q2:
    set(GEO_QUAD, 2, 9, {0, 1, 1, 1}, 3);
    f.dofOnFacet = {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};
    break;
  case FE_HEX_Q1:
    set(GEO_HEX, 3, 8, {0, 0, 0, 1}, 4);
    f.dofOnFacet = geoTopo(GEO_HEX).facets;
    break;
  case FE_HEX_Q2:
    set(GEO_HEX, 3, 27, {1, 1, 1, 1}, 9);
    f.dofOnFacet = {{0, 3, 2, 1, 11, 10, 9, 8, 20},   {3, 0, 4, 7, 11, 12, 19, 15, 21},
                    {0, 1, 5, 4, 8, 13, 16, 12, 22},   {1, 2, 6, 5, 9, 14, 17, 13, 23},
                    {2, 3, 7, 6, 10, 15, 18, 14, 24},  {4, 5, 6, 7, 16, 17, 18, 19, 25}};
    break;
  case FE_TET_P1:
    set(GEO_TET, 3, 4, {0, 0, 0, 1}, 3);
    f.dofOnFacet = geoTopo(GEO_TET).facets;
    break;
  case FE_TET_P2:
    set(GEO_TET, 3, 10, {0, 0, 1, 1}, 6);
    f.dofOnFacet = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 8, 7},
                    {0, 3, 2, 7, 9, 6}, {1, 2, 3, 5, 9, 8}};
    break;
  default: throw std::runtime_error("unknown FE type");
  }
  return f;
}

RefFE const & refFE(int type)
{
  static RefFE const tab[8] = {RefFE{},
                               makeFE(FE_LINE_P1),
                               makeFE(FE_QUAD_Q1),
                               makeFE(FE_QUAD_Q2),
                               makeFE(FE_HEX_Q1),
                               makeFE(FE_HEX_Q2),
                               makeFE(FE_TET_P1),
                               makeFE(FE_TET_P2)};
  if (type < 1 || type > 7)
    throw std::runtime_error("unknown FE type");
  return tab[type];
}

// ===========================================================================
// Gauss rules, reference src/qr.cpp.  The 3-point abscissa is kept at the
// reference's truncated value 0.774596669241483 (src/qr.cpp:38).
static double const kS13 = 0.5773502691896258;
static double const kS35 = 0.774596669241483;

QRule gaussRule(int geo, int nq)
{
  QRule r{};
  r.nq = nq;
  auto push = [&](std::initializer_list<double> x, double w)
  {
    for (double v: x)
      r.node.push_back(v);
    r.weight.push_back(w);
  };
  if (geo == GEO_LINE && nq == 2)
  {
    r.dim = 1;
    push({-kS13}, 1.);
    push({kS13}, 1.);
  }
  else if (geo == GEO_QUAD && nq == 4)
  {
    r.dim = 2;
    push({-kS13, -kS13}, 1.);
    push({kS13, -kS13}, 1.);
    push({-kS13, kS13}, 1.);
    push({kS13, kS13}, 1.);
  }
  else if (geo == GEO_QUAD && nq == 9)
  {
    r.dim = 2;
    double const x[3] = {-kS35, 0., kS35};
    double const w1[3] = {5., 8., 5.};
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i)
        push({x[i], x[j]}, (w1[i] * w1[j]) / 81.);
  }
  else if (geo == GEO_HEX && nq == 8)
  {
    r.dim = 3;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i)
          push({i ? kS13 : -kS13, j ? kS13 : -kS13, k ? kS13 : -kS13}, 1.);
  }
  else if (geo == GEO_HEX && nq == 27)
  {
    r.dim = 3;
    double const x[3] = {-kS35, 0., kS35};
    double const w1[3] = {5., 8., 5.};
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
          push({x[i], x[j], x[k]}, (w1[i] * w1[j] * w1[k]) / 729.);
  }
  else if (geo == GEO_TET && nq == 1)
  {
    r.dim = 3;
    push({.25, .25, .25}, 1. / 6);
  }
  else if (geo == GEO_TET && nq == 4)
  {
    r.dim = 3;
    double const a = 0.1381966011250105, b = 0.5854101966249685;
    push({a, a, a}, .25 / 6);
    push({b, a, a}, .25 / 6);
    push({a, b, a}, .25 / 6);
    push({a, a, b}, .25 / 6);
  }
  else if (geo == GEO_TET && nq == 11)
  {
    // reference src/qr.cpp:222-243 (weights as printed there)
    r.dim = 3;
    double const a = 0.0714285714285714, b = 0.7857142857142857;
    double const c = 0.399403576166799, d = 0.100596423833201;
    double const w0 = -0.0131555555555556, w1 = 0.0076222222222222, w2 = 0.024888888888889;
    push({.25, .25, .25}, w0);
    push({a, a, a}, w1);
    push({b, a, a}, w1);
    push({a, b, a}, w1);
    push({a, a, b}, w1);
    push({c, c, d}, w2);
    push({c, d, c}, w2);
    push({d, c, c}, w2);
    push({c, d, d}, w2);
    push({d, c, d}, w2);
    push({d, d, c}, w2);
  }
  else
    throw std::runtime_error("unsupported Gauss rule");
  return r;
}

// include/fe.hpp:30-132
QRule recommendedRule(int feType)
{
  switch (feType)
  {
  case FE_LINE_P1: return gaussRule(GEO_LINE, 2);
  case FE_QUAD_Q1: return gaussRule(GEO_QUAD, 4);
  case FE_QUAD_Q2: return gaussRule(GEO_QUAD, 9);
  case FE_HEX_Q1: return gaussRule(GEO_HEX, 8);
  case FE_HEX_Q2: return gaussRule(GEO_HEX, 27);
  case FE_TET_P1: return gaussRule(GEO_TET, 4);
  case FE_TET_P2: return gaussRule(GEO_TET, 11);
  }
  throw std::runtime_error("unknown FE type");
}

// ===========================================================================
// meshes
void buildLine(Mesh & m, double const * origin, double const * length, u32 n)
{
  m.geo = GEO_LINE;
  m.npe = 2;
  m.coords.assign(3 * (n + 1), 0.);
  double h[3];
  for (int k = 0; k < 3; ++k)
    h[k] = length[k] / n;
  for (u32 p = 0; p <= n; ++p)
    for (int k = 0; k < 3; ++k)
      m.coords[3 * p + k] = origin[k] + p * h[k];
  m.conn.resize(2 * n);
  for (u32 e = 0; e < n; ++e)
  {
    m.conn[2 * e] = e;
    m.conn[2 * e + 1] = e + 1;
  }
  m.bfacets.clear();
  // src/mesh.cpp:40-43: the two end points, LEFT then RIGHT
  m.bfacets.push_back(Facet{{0}, 0, 0, true, SIDE_LEFT});
  m.bfacets.push_back(Facet{{n}, n - 1, 1, true, SIDE_RIGHT});
}

static void gridPoints(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz)
{
  double const h[3] = {length[0] / nx, length[1] / ny, nz ? length[2] / nz : 1.};
  u32 const px = nx + 1, py = ny + 1, pz = nz + 1;
  m.coords.assign(3ull * px * py * pz, 0.);
  for (u32 k = 0; k < pz; ++k)
    for (u32 j = 0; j < py; ++j)
      for (u32 i = 0; i < px; ++i)
      {
        size_t const id = i + (size_t)px * j + (size_t)px * py * k;
        double const p[3] = {i * h[0], j * h[1], nz ? k * h[2] : 0.0};
        for (int c = 0; c < 3; ++c)
          m.coords[3 * id + c] = origin[c] + p[c];
      }
}

void buildSquare(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny)
{
  m.geo = GEO_QUAD;
  m.npe = 4;
  gridPoints(m, origin, length, nx, ny, 0);
  u32 const px = nx + 1;
  m.conn.clear();
  m.conn.reserve(4ull * nx * ny);
  for (u32 j = 0; j < ny; ++j)
    for (u32 i = 0; i < nx; ++i)
    {
      u32 const b = i + j * px;
      m.conn.insert(m.conn.end(), {b, b + 1, b + px + 1, b + px});
    }
  buildBoundaryFacets(m, origin, length);
}

void buildCubeHex(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz)
{
  m.geo = GEO_HEX;
  m.npe = 8;
  gridPoints(m, origin, length, nx, ny, nz);
  u32 const dx = 1, dy = nx + 1, dz = (nx + 1) * (ny + 1);
  m.conn.clear();
  m.conn.reserve(8ull * nx * ny * nz);
  for (u32 k = 0; k < nz; ++k)
    for (u32 j = 0; j < ny; ++j)
      for (u32 i = 0; i < nx; ++i)
      {
        u32 const b = i + j * dy + k * dz;
        m.conn.insert(
            m.conn.end(),
            {b, b + dx, b + dx + dy, b + dy, b + dz, b + dx + dz, b + dx + dy + dz, b + dy + dz});
      }
  buildBoundaryFacets(m, origin, length);
}

void buildCubeTet5(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz)
{
  m.geo = GEO_TET;
  m.npe = 4;
  gridPoints(m, origin, length, nx, ny, nz);
  u32 const dx = 1, dy = nx + 1, dz = (nx + 1) * (ny + 1);
  m.conn.clear();
  m.conn.reserve(20ull * nx * ny * nz);
  for (u32 k = 0; k < nz; ++k)
    for (u32 j = 0; j < ny; ++j)
      for (u32 i = 0; i < nx; ++i)
      {
        u32 const b = i + j * dy + k * dz;
        // cube corners in hex order 0..7
        u32 const c[8] = {b, b + dx, b + dx + dy, b + dy, b + dz, b + dx + dz, b + dx + dy + dz, b + dy + dz};
        // src/mesh.cpp:418-448: split parity alternates with i+j+k
        static int const odd[5][4] = {{1, 3, 4, 6}, {0, 1, 3, 4}, {1, 2, 3, 6}, {1, 4, 5, 6}, {3, 4, 6, 7}};
        static int const even[5][4] = {{0, 2, 5, 7}, {0, 1, 2, 5}, {0, 2, 3, 7}, {0, 4, 5, 7}, {2, 5, 6, 7}};
        auto const & t = ((i + j + k) % 2) ? odd : even;
        for (int r = 0; r < 5; ++r)
          m.conn.insert(m.conn.end(), {c[t[r][0]], c[t[r][1]], c[t[r][2]], c[t[r][3]]});
      }
  buildBoundaryFacets(m, origin, length);
}

void buildCubeTet6(Mesh & m, double const * origin, double const * length, u32 nx, u32 ny, u32 nz)
{
  m.geo = GEO_TET;
  m.npe = 4;
  gridPoints(m, origin, length, nx, ny, nz);
  u32 const dx = 1, dy = nx + 1, dz = (nx + 1) * (ny + 1);
  m.conn.clear();
  m.conn.reserve(24ull * nx * ny * nz);
  // Kuhn split: six tets around the 0-6 diagonal, same in every cube
  static int const kuhn[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                                 {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
  for (u32 k = 0; k < nz; ++k)
    for (u32 j = 0; j < ny; ++j)
      for (u32 i = 0; i < nx; ++i)
      {
        u32 const b = i + j * dy + k * dz;
        u32 const c[8] = {b, b + dx, b + dx + dy, b + dy, b + dz, b + dx + dz, b + dx + dy + dz, b + dy + dz};
        for (int r = 0; r < 6; ++r)
          m.conn.insert(m.conn.end(), {c[kuhn[r][0]], c[kuhn[r][1]], c[kuhn[r][2]], c[kuhn[r][3]]});
      }
  buildBoundaryFacets(m, origin, length);
}

void buildBoundaryFacets(Mesh & m, double const * origin, double const * length)
{
  GeoTopo const & t = geoTopo(m.geo);
  struct Entry
  {
    std::vector<u32> pts;
    u32 elem;
    int side;
    int count;
  };
  std::map<std::vector<u32>, Entry> fmap;
  i64 const ne = m.nelem();
  for (i64 e = 0; e < ne; ++e)
  {
    u32 const * v = &m.conn[e * m.npe];
    for (int s = 0; s < t.nfacets; ++s)
    {
      std::vector<u32> pts(t.facet_npts), key(t.facet_npts);
      for (int p = 0; p < t.facet_npts; ++p)
        pts[p] = key[p] = v[t.facets[s][p]];
      std::sort(key.begin(), key.end());
      auto [it, ins] = fmap.insert({key, Entry{pts, (u32)e, s, 1}});
      if (!ins)
        it->second.count++;
    }
  }
  m.bfacets.clear();
  for (auto const & [key, f]: fmap)
  {
    if (f.count != 1)
      continue;
    Facet fc{f.pts, f.elem, f.side, true, SIDE_NONE};
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (u32 p: fc.pts)
      for (int k = 0; k < 3; ++k)
      {
        mn[k] = std::min(mn[k], m.coords[3 * p + k]);
        mx[k] = std::max(mx[k], m.coords[3 * p + k]);
      }
    if (std::fabs(mx[1] - origin[1]) < 1e-6 * length[1])
      fc.marker = SIDE_BOTTOM;
    else if (std::fabs(mn[0] - origin[0] - length[0]) < 1e-6 * length[0])
      fc.marker = SIDE_RIGHT;
    else if (std::fabs(mn[1] - origin[1] - length[1]) < 1e-6 * length[1])
      fc.marker = SIDE_TOP;
    else if (std::fabs(mx[0] - origin[0]) < 1e-6 * length[0])
      fc.marker = SIDE_LEFT;
    else if (std::fabs(mx[2] - origin[2]) < 1e-6 * length[2])
      fc.marker = SIDE_BACK;
    else if (std::fabs(mn[2] - origin[2] - length[2]) < 1e-6 * length[2])
      fc.marker = SIDE_FRONT;
    m.bfacets.push_back(fc);
  }
}

void readGmsh(Mesh & m, std::string const & path, int geo)
{
  std::ifstream in(path);
  if (!in)
    throw std::runtime_error("cannot open " + path);
  std::vector<std::string> tok;
  for (std::string w; in >> w;)
    tok.push_back(w);
  size_t k = 0;
  auto next = [&]() -> std::string const &
  {
    if (k >= tok.size())
      throw std::runtime_error("gmsh: unexpected end of file");
    return tok[k++];
  };
  if (next() != "$MeshFormat" || next() != "2.2" || next() != "0" || next() != "8" || next() != "$EndMeshFormat")
    throw std::runtime_error("gmsh: not a 2.2 ASCII file");
  GeoTopo const & t = geoTopo(geo);
  // node counts of the element types of the format
  std::map<int, int> nodes{{15, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 4}, {5, 8}, {8, 3}, {9, 6}, {10, 9}, {11, 10}, {12, 27}};
  std::map<int, std::pair<int, int>> volType{{1, {1, 8}}, {2, {3, 10}}, {3, {5, 12}}, {4, {4, 11}}};
  std::map<int, std::pair<int, int>> facetType{{1, {15, 15}}, {2, {1, 8}}, {3, {3, 10}}, {4, {2, 9}}};
  auto const vt = volType.at(geo), ft = facetType.at(geo);
  std::vector<double> pts;
  std::vector<u32> vconn;
  std::vector<std::pair<std::vector<u32>, int>> ffac;
  while (k < tok.size())
  {
    std::string const sec = next();
    if (sec == "$Nodes")
    {
      long const nn = std::stol(next());
      for (long i = 0; i < nn; ++i)
      {
        if (std::stol(next()) != i + 1)
          throw std::runtime_error("gmsh: node ids not consecutive");
        for (int d = 0; d < 3; ++d)
          pts.push_back(std::stod(next()));
      }
      if (next() != "$EndNodes")
        throw std::runtime_error("gmsh: $EndNodes");
    }
    else if (sec == "$Elements")
    {
      long const ne = std::stol(next());
      for (long i = 0; i < ne; ++i)
      {
        next(); // id
        int const type = std::stoi(next());
        int const ntags = std::stoi(next());
        int tag0 = 0;
        for (int j = 0; j < ntags; ++j)
        {
          int const tg = std::stoi(next());
          if (j == 0)
            tag0 = tg;
        }
        std::vector<u32> c(nodes.at(type));
        for (auto & v: c)
          v = (u32)(std::stol(next()) - 1);
        if (type == vt.first || type == vt.second)
          vconn.insert(vconn.end(), c.begin(), c.begin() + t.npts);
        else if (type == ft.first || type == ft.second)
          ffac.push_back({std::vector<u32>(c.begin(), c.begin() + t.facet_npts), tag0});
      }
      if (next() != "$EndElements")
        throw std::runtime_error("gmsh: $EndElements");
    }
    else
    {
      std::string const end = "$End" + sec.substr(1);
      while (next() != end)
      {
      }
    }
  }
  // renumber the points used by volume elements, in file order
  std::vector<char> used(pts.size() / 3, 0);
  for (u32 v: vconn)
    used.at(v) = 1;
  std::vector<u32> id(used.size(), 0xffffffffu);
  m = Mesh{};
  m.geo = geo;
  m.npe = t.npts;
  for (size_t i = 0; i < used.size(); ++i)
    if (used[i])
    {
      id[i] = (u32)(m.coords.size() / 3);
      m.coords.insert(m.coords.end(), {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]});
    }
  for (u32 v: vconn)
    m.conn.push_back(id[v]);
  // boundary facets (seen once), marked from the file's facets
  std::map<std::vector<u32>, int> mark;
  for (auto & f: ffac)
  {
    std::vector<u32> key;
    for (u32 v: f.first)
      key.push_back(id.at(v));
    std::sort(key.begin(), key.end());
    mark[key] = f.second;
  }
  std::map<std::vector<u32>, std::pair<Facet, int>> seen;
  for (i64 e = 0; e < m.nelem(); ++e)
    for (int s = 0; s < t.nfacets; ++s)
    {
      std::vector<u32> p(t.facet_npts);
      for (int j = 0; j < t.facet_npts; ++j)
        p[j] = m.conn[e * m.npe + t.facets[s][j]];
      std::vector<u32> key = p;
      std::sort(key.begin(), key.end());
      auto [it, ins] = seen.insert({key, {Facet{p, (u32)e, s, true, SIDE_NONE}, 0}});
      it->second.second++;
    }
  for (auto & [key, fc]: seen)
    if (fc.second == 1)
    {
      Facet f = fc.first;
      auto it = mark.find(key);
      f.marker = it == mark.end() ? SIDE_NONE : it->second;
      m.bfacets.push_back(f);
    }
}

// ===========================================================================
DofMap buildDofMap(Mesh const & m, RefFE const & fe)
{
  GeoTopo const & t = geoTopo(fe.geo);
  DofMap d;
  d.n = fe.n;
  i64 const ne = m.nelem();
  d.elemMap.assign(ne * fe.n, 0);
  u32 const notSet = 0xffffffffu;
  std::vector<u32> ptMap(m.npts(), notSet);
  std::map<std::set<u32>, u32> edgeDofs, faceDofs;
  u32 size = 0;
  for (i64 e = 0; e < ne; ++e)
  {
    u32 const * v = &m.conn[e * m.npe];
    u32 * row = &d.elemMap[e * fe.n];
    int loc = 0;
    if (fe.dofPlace[3])
      for (int p = 0; p < t.npts; ++p)
      {
        if (ptMap[v[p]] == notSet)
          ptMap[v[p]] = size++;
        row[loc++] = ptMap[v[p]];
      }
    if (fe.dofPlace[2])
      for (int k = 0; k < t.nedges; ++k)
      {
        std::set<u32> key{v[t.edges[k][0]], v[t.edges[k][1]]};
        auto it = edgeDofs.find(key);
        if (it == edgeDofs.end())
          it = edgeDofs.insert({key, size++}).first;
        row[loc++] = it->second;
      }
    if (fe.dofPlace[1])
      for (int k = 0; k < t.nfaces; ++k)
      {
        std::set<u32> key;
        for (int p = 0; p < t.face_npts; ++p)
          key.insert(v[t.faces[k][p]]);
        auto it = faceDofs.find(key);
        if (it == faceDofs.end())
          it = faceDofs.insert({key, size++}).first;
        row[loc++] = it->second;
      }
    if (fe.dofPlace[0])
      row[loc++] = size++;
    if (loc != fe.n)
      throw std::runtime_error("dof count mismatch");
  }
  d.ndofs = size;
  return d;
}

std::vector<u32> constrainedDofs(Mesh const & m, RefFE const & fe, DofMap const & d, int marker)
{
  std::set<u32> s;
  for (auto const & f: m.bfacets)
    if (f.marker == marker)
      for (int k: fe.dofOnFacet[f.side])
        s.insert(d.elemMap[(i64)f.elem * fe.n + k]);
  return {s.begin(), s.end()};
}

std::vector<double> dofCoords(Mesh const & m, RefFE const & fe, DofMap const & d)
{
  std::vector<double> out(3ull * d.ndofs, 0.);
  std::vector<double> vtx(3 * m.npe), pts(3 * fe.n);
  for (i64 e = 0; e < m.nelem(); ++e)
  {
    for (int p = 0; p < m.npe; ++p)
      for (int k = 0; k < 3; ++k)
        vtx[3 * p + k] = m.coords[3ull * m.conn[e * m.npe + p] + k];
    fe.dofPts(vtx.data(), pts.data());
    for (int i = 0; i < fe.n; ++i)
      for (int k = 0; k < 3; ++k)
        out[3ull * d.elemMap[e * fe.n + i] + k] = pts[3 * i + k];
  }
  return out;
}

// ===========================================================================
CurFE::CurFE(RefFE const & f, QRule const & q): fe(&f), qr(q), n(f.n), nq(q.nq), dim(f.dim)
{
  if (q.dim != f.dim)
    throw std::runtime_error("quadrature dimension does not match the element");
  phiRef.assign(nq * n, 0.);
  dphiRef.assign(nq * n * dim, 0.);
  mapG.assign(nq * f.ng * dim, 0.);
  for (int iq = 0; iq < nq; ++iq)
    for (int i = 0; i < n; ++i)
    {
      double const * xi = &qr.node[iq * dim];
      phiRef[iq * n + i] = f.phi(i, xi);
      f.dphi(i, xi, &dphiRef[(iq * n + i) * dim]);
      f.dphi(i, xi, &mapG[(iq * f.ng + i) * dim]); // isoparametric mapping
    }
  jac.assign(nq * 3 * dim, 0.);
  detJ.assign(nq, 0.);
  JxW.assign(nq, 0.);
  qpt.assign(nq * 3, 0.);
  dphi.assign(nq * n * 3, 0.);
  mpts.assign(n * 3, 0.);
}

void CurFE::reinit(double const * vtx)
{
  fe->dofPts(vtx, mpts.data());
  double org[3];
  fe->origin(vtx, org);
  int const ng = fe->ng;
  for (int iq = 0; iq < nq; ++iq)
  {
    double J[3][3] = {{0}};
    for (int g = 0; g < ng; ++g)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < dim; ++c)
          J[r][c] += mpts[3 * g + r] * mapG[(iq * ng + g) * dim + c];
    double A[3][3] = {{0}};
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b)
      {
        double s = 0.;
        for (int r = 0; r < 3; ++r)
          s += J[r][a] * J[r][b];
        A[a][b] = s;
      }
    double det = 0., I[3][3] = {{0}};
    if (dim == 1)
    {
      det = A[0][0];
      I[0][0] = 1. / A[0][0];
    }
    else if (dim == 2)
    {
      det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
      I[0][0] = A[1][1] / det;
      I[0][1] = -A[0][1] / det;
      I[1][0] = -A[1][0] / det;
      I[1][1] = A[0][0] / det;
    }
    else
    {
      double const c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
      double const c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
      double const c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
      det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
      I[0][0] = c00 / det;
      I[1][0] = c01 / det;
      I[2][0] = c02 / det;
      I[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
      I[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
      I[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
      I[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
      I[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
      I[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
    }
    detJ[iq] = std::sqrt(det);
    JxW[iq] = detJ[iq] * qr.weight[iq];
    // J^+ = (J^T J)^{-1} J^T   (dim x 3)
    double Jp[3][3] = {{0}};
    for (int a = 0; a < dim; ++a)
      for (int r = 0; r < 3; ++r)
      {
        double s = 0.;
        for (int b = 0; b < dim; ++b)
          s += I[a][b] * J[r][b];
        Jp[a][r] = s;
      }
    for (int r = 0; r < 3; ++r)
    {
      double s = 0.;
      for (int c = 0; c < dim; ++c)
        s += J[r][c] * qr.node[iq * dim + c];
      qpt[iq * 3 + r] = org[r] + s;
      for (int c = 0; c < dim; ++c)
        jac[(iq * 3 + r) * dim + c] = J[r][c];
    }
    for (int i = 0; i < n; ++i)
      for (int r = 0; r < 3; ++r)
      {
        double s = 0.;
        for (int a = 0; a < dim; ++a)
          s += dphiRef[(iq * n + i) * dim + a] * Jp[a][r];
        dphi[(iq * n + i) * 3 + r] = s;
      }
  }
}

// ===========================================================================
double evalRhs(int kind, double const * p, double const * x)
{
  switch (kind)
  {
  case RHS_CONST: return p[0];
  case RHS_SINSIN: return p[0] * std::sin(p[1] * x[0]) * std::sin(p[2] * x[1]);
  case RHS_SINSINSIN:
    return p[0] * std::sin(p[1] * x[0]) * std::sin(p[2] * x[1]) * std::sin(p[3] * x[2]);
  case RHS_FEVAR1D: return std::sin(p[1] * x[0]) - p[0] * std::cos(p[2] * x[0]);
  }
  return 0.;
}

static void gatherVtx(Mesh const & m, i64 e, std::vector<double> & vtx)
{
  vtx.resize(3 * m.npe);
  for (int p = 0; p < m.npe; ++p)
    for (int k = 0; k < 3; ++k)
      vtx[3 * p + k] = m.coords[3ull * m.conn[e * m.npe + p] + k];
}

// element matrix and right-hand side of element e (buildLhs loop body,
// include/builder.hpp:53-128): terms in call order, per-element Dirichlet
// elimination; Ke is n x n row-major, Fe n values
static void elementLhs(
    Mesh const & m, CurFE & cfe, DofMap const & d, std::vector<Term> const & terms,
    std::vector<uint8_t> const & bcFlag, std::vector<double> const & bcVal, i64 e, double * Ke, double * Fe,
    std::vector<double> & h, std::vector<double> & vtx, std::vector<double> & kloc, std::vector<int> & cons)
{
  int const n = cfe.n, nq = cfe.nq;
  bool const hasBC = !bcFlag.empty();
  gatherVtx(m, e, vtx);
  cfe.reinit(vtx.data());
  u32 const * ids = &d.elemMap[e * n];
  std::fill(Ke, Ke + n * n, 0.);
  std::fill(Fe, Fe + n, 0.);
  for (auto const & t: terms)
  {
    bool const fe = t.kind == TERM_STIFFNESS_FE || t.kind == TERM_MASS_FE;
    if (fe)
      for (int i = 0; i < n; ++i)
        kloc[i] = t.kappa[ids[i]];
    for (int q = 0; q < nq; ++q)
    {
      double kq = 0.;
      if (fe)
        for (int i = 0; i < n; ++i)
          kq += cfe.phiRef[q * n + i] * kloc[i];
      double s = 0.;
      switch (t.kind)
      {
      case TERM_STIFFNESS:
      case TERM_MASS: s = t.coef * cfe.JxW[q]; break;
      case TERM_STIFFNESS_FE: s = kq * cfe.JxW[q]; break;
      case TERM_MASS_FE: s = (t.coef * kq) * cfe.JxW[q]; break;
      }
      bool const stiff = t.kind == TERM_STIFFNESS || t.kind == TERM_STIFFNESS_FE;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
        {
          double v;
          if (stiff)
          {
            double const * gi = &cfe.dphi[(q * n + i) * 3];
            double const * gj = &cfe.dphi[(q * n + j) * 3];
            v = gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2];
          }
          else
            v = cfe.phiRef[q * n + i] * cfe.phiRef[q * n + j];
          Ke[i * n + j] += s * v;
        }
    }
  }
  // Dirichlet elimination, include/builder.hpp:83-124
  if (hasBC)
  {
    bool any = false;
    for (int i = 0; i < n; ++i)
    {
      cons[i] = bcFlag[ids[i]];
      h[i] = cons[i] ? bcVal[ids[i]] : 0.;
      any = any || cons[i];
    }
    if (any)
    {
      for (int i = 0; i < n; ++i)
      {
        double s = 0.;
        for (int j = 0; j < n; ++j)
          s += Ke[i * n + j] * h[j];
        Fe[i] = cons[i] ? 0. : (Fe[i] - s);
      }
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          if (cons[i] || cons[j])
            Ke[i * n + j] = 0.;
      for (int i = 0; i < n; ++i)
        if (cons[i])
        {
          Ke[i * n + i] = 1.0; // BCEss::diag, include/bc.hpp:301
          Fe[i] = h[i];
        }
    }
  }
}

void Builder::buildLhs(
    Mesh const & m, CurFE & cfe, DofMap const & d, std::vector<Term> const & terms,
    std::vector<uint8_t> const & bcFlag, std::vector<double> const & bcVal)
{
  int const n = cfe.n;
  i64 const E = m.nelem();
  size_t const nn = (size_t)n * n, base = trow.size();
  trow.resize(base + (size_t)E * nn);
  tcol.resize(base + (size_t)E * nn);
  tval.resize(base + (size_t)E * nn);
  if (threads <= 1)
  {
    // the reference's serial loop: triplets and b in element order
    std::vector<double> Ke(nn), Fe(n), h(n), vtx, kloc(n);
    std::vector<int> cons(n);
    for (i64 e = 0; e < E; ++e)
    {
      elementLhs(m, cfe, d, terms, bcFlag, bcVal, e, Ke.data(), Fe.data(), h, vtx, kloc, cons);
      u32 const * ids = &d.elemMap[e * n];
      for (int i = 0; i < n; ++i)
      {
        b[ids[i]] += Fe[i];
        for (int j = 0; j < n; ++j)
        {
          size_t const k = base + e * nn + (size_t)i * n + j;
          trow[k] = ids[i];
          tcol[k] = ids[j];
          tval[k] = Ke[i * n + j];
        }
      }
    }
    return;
  }
  // threaded element loop (the CPU baseline on all host cores): every
  // element writes its own triplet range; b is summed afterwards in element
  // order, so the result is bitwise the serial one
  std::vector<double> FeAll((size_t)E * n);
#pragma omp parallel num_threads(threads)
  {
    CurFE local = cfe;
    std::vector<double> Ke(nn), h(n), vtx, kloc(n);
    std::vector<int> cons(n);
#pragma omp for schedule(static)
    for (i64 e = 0; e < E; ++e)
    {
      elementLhs(m, local, d, terms, bcFlag, bcVal, e, Ke.data(), &FeAll[(size_t)e * n], h, vtx, kloc, cons);
      u32 const * ids = &d.elemMap[e * n];
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
        {
          size_t const k = base + e * nn + (size_t)i * n + j;
          trow[k] = ids[i];
          tcol[k] = ids[j];
          tval[k] = Ke[i * n + j];
        }
    }
  }
  for (i64 e = 0; e < E; ++e)
    for (int i = 0; i < n; ++i)
      b[d.elemMap[e * n + i]] += FeAll[(size_t)e * n + i];
}

DofMap vectorDofMap(DofMap const & s, int dim, bool interleaved, u32 offset)
{
  DofMap v;
  v.n = s.n * dim;
  v.ndofs = s.ndofs * dim;
  i64 const ne = s.n ? (i64)s.elemMap.size() / s.n : 0;
  v.elemMap.resize((size_t)ne * v.n);
  for (i64 e = 0; e < ne; ++e)
    for (int d = 0; d < dim; ++d)
      for (int i = 0; i < s.n; ++i)
      {
        u32 const id = s.elemMap[e * s.n + i];
        v.elemMap[e * v.n + i + d * s.n] = offset + (interleaved ? id * dim + d : id + d * s.ndofs);
      }
  return v;
}

void Builder::buildLhsVector(
    Mesh const & m, CurFE & cfe, DofMap const & d, int dim, double cS, double cT, double cM,
    std::vector<uint8_t> const & bcFlag, std::vector<double> const & bcVal)
{
  int const n = cfe.n, nq = cfe.nq, nn = n * dim;
  std::vector<double> Ke(nn * nn), Fe(nn), h(nn), vtx;
  std::vector<int> cons(nn);
  bool const hasBC = !bcFlag.empty();
  for (i64 e = 0; e < m.nelem(); ++e)
  {
    gatherVtx(m, e, vtx);
    cfe.reinit(vtx.data());
    u32 const * ids = &d.elemMap[e * nn];
    std::fill(Ke.begin(), Ke.end(), 0.);
    std::fill(Fe.begin(), Fe.end(), 0.);
    // the terms in the order of the tuple {stiffness, tensor stiffness, mass}
    auto blockStiff = [&](double coef)
    {
      // AssemblyStiffness::build: for q, for d: block(d, d) += coef JxW dphi dphi^T
      for (int q = 0; q < nq; ++q)
      {
        double const s = coef * cfe.JxW[q];
        for (int c = 0; c < dim; ++c)
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
            {
              double const * gi = &cfe.dphi[(q * n + i) * 3];
              double const * gj = &cfe.dphi[(q * n + j) * 3];
              Ke[(i + c * n) * nn + j + c * n] += s * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
            }
      }
    };
    if (cS != 0.)
      blockStiff(cS);
    if (cT != 0.)
    {
      // AssemblyTensorStiffness::build (assembly_lhs.hpp:131-163): per q and
      // component di the block-diagonal term, then for every dj
      // block(di, dj) += coef JxW dphi.col(di) dphi.col(dj)^T
      for (int q = 0; q < nq; ++q)
      {
        double const s = cT * cfe.JxW[q];
        for (int di = 0; di < dim; ++di)
        {
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
            {
              double const * gi = &cfe.dphi[(q * n + i) * 3];
              double const * gj = &cfe.dphi[(q * n + j) * 3];
              Ke[(i + di * n) * nn + j + di * n] += s * (gi[0] * gj[0] + gi[1] * gj[1] + gi[2] * gj[2]);
            }
          for (int dj = 0; dj < dim; ++dj)
            for (int i = 0; i < n; ++i)
              for (int j = 0; j < n; ++j)
                Ke[(i + di * n) * nn + j + dj * n] +=
                    s * (cfe.dphi[(q * n + i) * 3 + di] * cfe.dphi[(q * n + j) * 3 + dj]);
        }
      }
    }
    if (cM != 0.)
    {
      // AssemblyScalarMass::build (assembly_lhs.hpp:207-221): block(d, d) += coef JxW phi phi^T
      for (int q = 0; q < nq; ++q)
      {
        double const s = cM * cfe.JxW[q];
        for (int c = 0; c < dim; ++c)
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
              Ke[(i + c * n) * nn + j + c * n] += s * (cfe.phiRef[q * n + i] * cfe.phiRef[q * n + j]);
      }
    }
    if (hasBC)
    {
      bool any = false;
      for (int i = 0; i < nn; ++i)
      {
        cons[i] = bcFlag[ids[i]];
        h[i] = cons[i] ? bcVal[ids[i]] : 0.;
        any = any || cons[i];
      }
      if (any)
      {
        for (int i = 0; i < nn; ++i)
        {
          double s = 0.;
          for (int j = 0; j < nn; ++j)
            s += Ke[i * nn + j] * h[j];
          Fe[i] = cons[i] ? 0. : (Fe[i] - s);
        }
        for (int i = 0; i < nn; ++i)
          for (int j = 0; j < nn; ++j)
            if (cons[i] || cons[j])
              Ke[i * nn + j] = 0.;
        for (int i = 0; i < nn; ++i)
          if (cons[i])
          {
            Ke[i * nn + i] = 1.0; // BCEss::diag
            Fe[i] = h[i];
          }
      }
    }
    // storage order of builder.hpp:128-152: i, d1, j, d2
    for (int i = 0; i < n; ++i)
      for (int d1 = 0; d1 < dim; ++d1)
      {
        int const r = i + d1 * n;
        b[ids[r]] += Fe[r];
        for (int j = 0; j < n; ++j)
          for (int d2 = 0; d2 < dim; ++d2)
          {
            int const c = j + d2 * n;
            trow.push_back(ids[r]);
            tcol.push_back(ids[c]);
            tval.push_back(Ke[r * nn + c]);
          }
      }
  }
}

void Builder::buildCoupling(
    Mesh const & m, CurFE & row, CurFE & clm, DofMap const & dr, int dimR, DofMap const & dc, int dimC, int kind,
    double coef, std::vector<int> const & comps, std::vector<uint8_t> const & bcRowFlag,
    std::vector<uint8_t> const & bcClmFlag, std::vector<double> const & bcClmVal)
{
  int const nR = row.n, nC = clm.n, nq = row.nq;
  int const NR = nR * dimR, NC = nC * dimC;
  std::vector<double> Ke((size_t)NR * NC), Fe(NR), h(NC), vtx;
  std::vector<int> crow(NR), cclm(NC);
  for (i64 e = 0; e < m.nelem(); ++e)
  {
    gatherVtx(m, e, vtx);
    row.reinit(vtx.data());
    clm.reinit(vtx.data());
    u32 const * ir = &dr.elemMap[e * NR];
    u32 const * ic = &dc.elemMap[e * NC];
    std::fill(Ke.begin(), Ke.end(), 0.);
    // AssemblyGrad: block (d nR, 0) += coef JxW dphiR.col(c) phiC^T
    // AssemblyDiv:  block (0, d nC) += coef JxW phiR dphiC.col(c)^T
    int blk = 0;
    for (int c: comps)
    {
      for (int q = 0; q < nq; ++q)
      {
        double const s = coef * row.JxW[q];
        if (kind == COUPLING_GRAD)
          for (int i = 0; i < nR; ++i)
            for (int j = 0; j < nC; ++j)
              Ke[(size_t)(i + blk * nR) * NC + j] += (s * row.dphi[(q * nR + i) * 3 + c]) * clm.phiRef[q * nC + j];
        else
          for (int i = 0; i < nR; ++i)
            for (int j = 0; j < nC; ++j)
              Ke[(size_t)i * NC + j + blk * nC] += (s * row.phiRef[q * nR + i]) * clm.dphi[(q * nC + j) * 3 + c];
      }
      ++blk;
    }
    // Crow clears constrained rows, Cclm constrained columns; Fe = -Crow Ke h
    for (int i = 0; i < NR; ++i)
      crow[i] = bcRowFlag.empty() ? 0 : bcRowFlag[ir[i]];
    for (int j = 0; j < NC; ++j)
    {
      cclm[j] = bcClmFlag.empty() ? 0 : bcClmFlag[ic[j]];
      h[j] = cclm[j] ? bcClmVal[ic[j]] : 0.;
    }
    for (int i = 0; i < NR; ++i)
    {
      double s = 0.;
      for (int j = 0; j < NC; ++j)
        s += Ke[(size_t)i * NC + j] * h[j];
      Fe[i] = crow[i] ? 0. : -s;
    }
    for (int i = 0; i < NR; ++i)
      for (int j = 0; j < NC; ++j)
        if (crow[i] || cclm[j])
          Ke[(size_t)i * NC + j] = 0.;
    for (int i = 0; i < nR; ++i)
      for (int d1 = 0; d1 < dimR; ++d1)
      {
        int const r = i + d1 * nR;
        b[ir[r]] += Fe[r];
        for (int j = 0; j < nC; ++j)
          for (int d2 = 0; d2 < dimC; ++d2)
          {
            int const c = j + d2 * nC;
            trow.push_back(ir[r]);
            tcol.push_back(ic[c]);
            tval.push_back(Ke[(size_t)r * NC + c]);
          }
      }
  }
}

void Builder::buildRhs(
    Mesh const & m, CurFE & cfe, DofMap const & d, int rhsKind, double const * params,
    double const * qvalues, std::vector<uint8_t> const & bcFlag)
{
  int const n = cfe.n, nq = cfe.nq;
  std::vector<double> Fe(n), vtx;
  for (i64 e = 0; e < m.nelem(); ++e)
  {
    gatherVtx(m, e, vtx);
    cfe.reinit(vtx.data());
    u32 const * ids = &d.elemMap[e * n];
    std::fill(Fe.begin(), Fe.end(), 0.);
    for (int q = 0; q < nq; ++q)
    {
      if (rhsKind == RHS_PROJECTION)
      {
        // AssemblyS2SProjection::build (include/assembly_rhs.hpp:134-148):
        // Fe += coef JxW_q phi_q (phi_q . f_local)
        double fq = 0.;
        for (int j = 0; j < n; ++j)
          fq += cfe.phiRef[q * n + j] * qvalues[ids[j]];
        double const w = params[0] * cfe.JxW[q];
        for (int i = 0; i < n; ++i)
          Fe[i] += (w * cfe.phiRef[q * n + i]) * fq;
        continue;
      }
      double const f = rhsKind == RHS_QVALUES ? qvalues[e * nq + q]
                                              : evalRhs(rhsKind, params, &cfe.qpt[3 * q]);
      for (int i = 0; i < n; ++i)
        Fe[i] += (cfe.JxW[q] * cfe.phiRef[q * n + i]) * f;
    }
    for (int i = 0; i < n; ++i)
      if (!bcFlag.empty() && bcFlag[ids[i]])
        Fe[i] = 0.;
    for (int i = 0; i < n; ++i)
      b[ids[i]] += Fe[i];
  }
}

Csr Builder::close(double pruneEps) const
{
  if (threads > 1)
    return closeThreaded(pruneEps);
  size_t const nt = trow.size();
  // counting sort by row (stable), then stable sort by column inside each row:
  // duplicates end up adjacent in insertion order and are summed left to right
  std::vector<i64> cnt(ndofs + 1, 0);
  for (size_t k = 0; k < nt; ++k)
    cnt[trow[k] + 1]++;
  for (i64 r = 0; r < ndofs; ++r)
    cnt[r + 1] += cnt[r];
  std::vector<size_t> ord(nt);
  {
    std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < nt; ++k)
      ord[pos[trow[k]]++] = k;
  }
  Csr A;
  A.nrows = ndofs;
  A.rowptr.assign(ndofs + 1, 0);
  for (i64 r = 0; r < ndofs; ++r)
  {
    auto b0 = ord.begin() + cnt[r], b1 = ord.begin() + cnt[r + 1];
    std::stable_sort(b0, b1, [&](size_t a, size_t c) { return tcol[a] < tcol[c]; });
    for (auto it = b0; it != b1;)
    {
      u32 const c = tcol[*it];
      double s = tval[*it];
      ++it;
      while (it != b1 && tcol[*it] == c)
      {
        s += tval[*it];
        ++it;
      }
      if (!(std::fabs(s) <= pruneEps)) // Eigen prune(1, eps): drop |v| <= eps
      {
        A.col.push_back(c);
        A.val.push_back(s);
      }
    }
    A.rowptr[r + 1] = (i64)A.col.size();
  }
  return A;
}

// close() with the rows sorted and summed in parallel: the same per-row
// stable sort and left-to-right sums, so the same CSR bit for bit
Csr Builder::closeThreaded(double pruneEps) const
{
  size_t const nt = trow.size();
  std::vector<i64> cnt(ndofs + 1, 0);
  for (size_t k = 0; k < nt; ++k)
    cnt[trow[k] + 1]++;
  for (i64 r = 0; r < ndofs; ++r)
    cnt[r + 1] += cnt[r];
  std::vector<size_t> ord(nt);
  {
    std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < nt; ++k)
      ord[pos[trow[k]]++] = k;
  }
  // per row: the kept (column, sum) pairs at the start of its segment
  std::vector<u32> ucol(nt);
  std::vector<double> uval(nt);
  std::vector<i64> kept(ndofs + 1, 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1024)
  for (i64 r = 0; r < ndofs; ++r)
  {
    auto b0 = ord.begin() + cnt[r], b1 = ord.begin() + cnt[r + 1];
    std::stable_sort(b0, b1, [&](size_t a, size_t c) { return tcol[a] < tcol[c]; });
    i64 u = cnt[r];
    for (auto it = b0; it != b1;)
    {
      u32 const c = tcol[*it];
      double s = tval[*it];
      ++it;
      while (it != b1 && tcol[*it] == c)
      {
        s += tval[*it];
        ++it;
      }
      if (!(std::fabs(s) <= pruneEps))
      {
        ucol[u] = c;
        uval[u] = s;
        ++u;
      }
    }
    kept[r + 1] = u - cnt[r];
  }
  Csr A;
  A.nrows = ndofs;
  A.rowptr.assign(ndofs + 1, 0);
  for (i64 r = 0; r < ndofs; ++r)
    A.rowptr[r + 1] = A.rowptr[r] + kept[r + 1];
  A.col.resize(A.rowptr[ndofs]);
  A.val.resize(A.rowptr[ndofs]);
#pragma omp parallel for num_threads(threads) schedule(static)
  for (i64 r = 0; r < ndofs; ++r)
    for (i64 k = 0; k < kept[r + 1]; ++k)
    {
      A.col[A.rowptr[r] + k] = ucol[cnt[r] + k];
      A.val[A.rowptr[r] + k] = uval[cnt[r] + k];
    }
  return A;
}

void structuralPattern(
    DofMap const & d, i64 nelem, std::vector<i64> & rowptr, std::vector<u32> & col,
    std::vector<u32> & emap)
{
  int const n = d.n;
  std::vector<std::set<u32>> rows(d.ndofs);
  for (i64 e = 0; e < nelem; ++e)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        rows[d.elemMap[e * n + i]].insert(d.elemMap[e * n + j]);
  rowptr.assign(d.ndofs + 1, 0);
  col.clear();
  for (u32 r = 0; r < d.ndofs; ++r)
  {
    col.insert(col.end(), rows[r].begin(), rows[r].end());
    rowptr[r + 1] = (i64)col.size();
  }
  emap.assign((size_t)nelem * n * n, 0);
  for (i64 e = 0; e < nelem; ++e)
    for (int i = 0; i < n; ++i)
    {
      u32 const r = d.elemMap[e * n + i];
      auto b0 = col.begin() + rowptr[r], b1 = col.begin() + rowptr[r + 1];
      for (int j = 0; j < n; ++j)
      {
        auto it = std::lower_bound(b0, b1, d.elemMap[e * n + j]);
        emap[((size_t)e * n + i) * n + j] = (u32)(it - col.begin());
      }
    }
}

} // namespace orc
