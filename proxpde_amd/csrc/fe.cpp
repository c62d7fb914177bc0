// Reference finite elements and quadrature rules (host tabulation for the kernels).
#include "fe.hpp"

#include <stdexcept>

namespace pxg
{

namespace
{
// 1D Lagrange factors on [-1, 1] indexed by node role:
// 0 -> node at -1, 1 -> node at +1, 2 -> node at 0 (quadratic bubble)
struct Lagrange1D
{
  int order;
  double value(int role, double x) const
  {
    if (order == 1)
      return role == 0 ? 0.5 * (1. - x) : 0.5 * (1. + x);
    if (role == 0)
      return 0.5 * x * (x - 1.);
    if (role == 1)
      return 0.5 * x * (x + 1.);
    return 1. - x * x;
  }
  double deriv(int role, double x) const
  {
    if (order == 1)
      return role == 0 ? -0.5 : 0.5;
    if (role == 0)
      return x - 0.5;
    if (role == 1)
      return x + 0.5;
    return -2. * x;
  }
};

// per-node 1D roles (x, y, z); layouts of the reference node orders
constexpr int kQuadQ1[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
constexpr int kQuadQ2[9][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}, {2, 0}, {1, 2}, {2, 1}, {0, 2}, {2, 2}};
constexpr int kHexQ1[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                              {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kHexQ2[27][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, // corners
    {2, 0, 0}, {1, 2, 0}, {2, 1, 0}, {0, 2, 0},                                         // bottom edges
    {0, 0, 2}, {1, 0, 2}, {1, 1, 2}, {0, 1, 2},                                         // vertical edges
    {2, 0, 1}, {1, 2, 1}, {2, 1, 1}, {0, 2, 1},                                         // top edges
    {2, 2, 0}, {0, 2, 2}, {2, 0, 2}, {1, 2, 2}, {2, 1, 2}, {2, 2, 1},                   // faces
    {2, 2, 2}};                                                                          // centre
// P2 tet edge nodes: (first, second) barycentric coordinate
constexpr int kTetEdge[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};

inline double lam(int a, double const * p)
{
  return a == 0 ? 1. - p[0] - p[1] - p[2] : p[a - 1];
}
inline double dlam(int a, int k) { return a == 0 ? -1. : (a - 1 == k ? 1. : 0.); }

GeoInfo makeGeo(int geo)
{
  GeoInfo g{};
  g.geo = geo;
  if (geo == kLine)
  {
    g.dim = 1;
    g.npts = 2;
    g.facets = {{0}, {1}};
    g.edges = {{0, 1}};
  }
  else if (geo == kQuad)
  {
    g.dim = 2;
    g.npts = 4;
    g.facets = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    g.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    g.faces = {{0, 1, 2, 3}};
  }
  else if (geo == kHex)
  {
    g.dim = 3;
    g.npts = 8;
    g.facets = {{0, 3, 2, 1}, {3, 0, 4, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {4, 5, 6, 7}};
    g.edges = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {7, 4}};
    g.faces = g.facets;
  }
  else if (geo == kTet)
  {
    g.dim = 3;
    g.npts = 4;
    g.facets = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
    g.edges = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    g.faces = g.facets;
  }
  else
    throw std::invalid_argument("unknown geometry");
  return g;
}

FeInfo makeFe(int type)
{
  FeInfo f{};
  f.type = type;
  auto init = [&](int geo, int dim, int n, int p0, int p1, int p2, int p3)
  {
    f.geo = geo;
    f.dim = dim;
    f.n = n;
    f.dofPlace[0] = p0;
    f.dofPlace[1] = p1;
    f.dofPlace[2] = p2;
    f.dofPlace[3] = p3;
  };
  switch (type)
  {
  case 1: init(kLine, 1, 2, 0, 0, 0, 1), f.dofOnFacet = {{0}, {1}}; break;
  case 2: init(kQuad, 2, 4, 0, 0, 0, 1), f.dofOnFacet = geoInfo(kQuad).facets; break;
  case 3: init(kQuad, 2, 9, 0, 1, 1, 1), f.dofOnFacet = {{0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}}; break;
  case 4: init(kHex, 3, 8, 0, 0, 0, 1), f.dofOnFacet = geoInfo(kHex).facets; break;
  case 5:
    init(kHex, 3, 27, 1, 1, 1, 1);
    f.dofOnFacet = {{0, 3, 2, 1, 11, 10, 9, 8, 20}, {3, 0, 4, 7, 11, 12, 19, 15, 21},
                    {0, 1, 5, 4, 8, 13, 16, 12, 22}, {1, 2, 6, 5, 9, 14, 17, 13, 23},
                    {2, 3, 7, 6, 10, 15, 18, 14, 24}, {4, 5, 6, 7, 16, 17, 18, 19, 25}};
    break;
  case 6: init(kTet, 3, 4, 0, 0, 0, 1), f.dofOnFacet = geoInfo(kTet).facets; break;
  case 7:
    init(kTet, 3, 10, 0, 0, 1, 1);
    f.dofOnFacet = {{0, 2, 1, 6, 5, 4}, {0, 1, 3, 4, 8, 7}, {0, 3, 2, 7, 9, 6}, {1, 2, 3, 5, 9, 8}};
    break;
  default: throw std::invalid_argument("unknown FE type");
  }
  return f;
}
} // namespace

GeoInfo const & geoInfo(int geo)
{
  static GeoInfo const g[5] = {GeoInfo{}, makeGeo(kLine), makeGeo(kQuad), makeGeo(kHex), makeGeo(kTet)};
  if (geo < 1 || geo > 4)
    throw std::invalid_argument("unknown geometry");
  return g[geo];
}

FeInfo const & feInfo(int type)
{
  static FeInfo const f[8] = {FeInfo{}, makeFe(1), makeFe(2), makeFe(3), makeFe(4), makeFe(5), makeFe(6), makeFe(7)};
  if (type < 1 || type > 7)
    throw std::invalid_argument("unknown FE type");
  return f[type];
}

double FeInfo::phi(int i, double const * xi) const
{
  Lagrange1D const l1{1}, l2{2};
  switch (type)
  {
  case 1: return l1.value(i, xi[0]);
  case 2: return l1.value(kQuadQ1[i][0], xi[0]) * l1.value(kQuadQ1[i][1], xi[1]);
  case 3: return l2.value(kQuadQ2[i][0], xi[0]) * l2.value(kQuadQ2[i][1], xi[1]);
  case 4:
    return l1.value(kHexQ1[i][0], xi[0]) * l1.value(kHexQ1[i][1], xi[1]) * l1.value(kHexQ1[i][2], xi[2]);
  case 5:
    return l2.value(kHexQ2[i][0], xi[0]) * l2.value(kHexQ2[i][1], xi[1]) * l2.value(kHexQ2[i][2], xi[2]);
  case 6: return lam(i, xi);
  case 7:
    if (i < 4)
      return lam(i, xi) * (2. * lam(i, xi) - 1.);
    return 4. * lam(kTetEdge[i - 4][0], xi) * lam(kTetEdge[i - 4][1], xi);
  }
  return 0.;
}

void FeInfo::grad(int i, double const * xi, double * g) const
{
  Lagrange1D const l1{1}, l2{2};
  switch (type)
  {
  case 1: g[0] = l1.deriv(i, xi[0]); return;
  case 2:
  case 3:
  {
    Lagrange1D const & l = type == 2 ? l1 : l2;
    int const * r = type == 2 ? kQuadQ1[i] : kQuadQ2[i];
    g[0] = l.deriv(r[0], xi[0]) * l.value(r[1], xi[1]);
    g[1] = l.value(r[0], xi[0]) * l.deriv(r[1], xi[1]);
    return;
  }
  case 4:
  case 5:
  {
    Lagrange1D const & l = type == 4 ? l1 : l2;
    int const * r = type == 4 ? kHexQ1[i] : kHexQ2[i];
    g[0] = l.deriv(r[0], xi[0]) * l.value(r[1], xi[1]) * l.value(r[2], xi[2]);
    g[1] = l.value(r[0], xi[0]) * l.deriv(r[1], xi[1]) * l.value(r[2], xi[2]);
    g[2] = l.value(r[0], xi[0]) * l.value(r[1], xi[1]) * l.deriv(r[2], xi[2]);
    return;
  }
  case 6:
    for (int k = 0; k < 3; ++k)
      g[k] = dlam(i, k);
    return;
  case 7:
    for (int k = 0; k < 3; ++k)
    {
      if (i < 4)
        g[k] = (4. * lam(i, xi) - 1.) * dlam(i, k);
      else
      {
        int const a = kTetEdge[i - 4][0], b = kTetEdge[i - 4][1];
        g[k] = 4. * (lam(a, xi) * dlam(b, k) + dlam(a, k) * lam(b, xi));
      }
    }
    return;
  }
}

void FeInfo::dofPoints(double const * v, double * out) const
{
  GeoInfo const & g = geoInfo(geo);
  auto mid = [&](std::initializer_list<int> ids, double * o)
  {
    double const s = 1.0 / ids.size();
    for (int k = 0; k < 3; ++k)
    {
      double acc = 0.;
      for (int id: ids)
        acc += v[3 * id + k];
      o[k] = s * acc;
    }
  };
  for (int p = 0; p < g.npts; ++p)
    for (int k = 0; k < 3; ++k)
      out[3 * p + k] = v[3 * p + k];
  if (type == 3)
  {
    for (int e = 0; e < 4; ++e)
      mid({g.edges[e][0], g.edges[e][1]}, out + 3 * (4 + e));
    mid({0, 1, 2, 3}, out + 3 * 8);
  }
  else if (type == 5)
  {
    for (int e = 0; e < 12; ++e)
      mid({g.edges[e][0], g.edges[e][1]}, out + 3 * (8 + e));
    for (int f = 0; f < 6; ++f)
      mid({g.facets[f][0], g.facets[f][1], g.facets[f][2], g.facets[f][3]}, out + 3 * (20 + f));
    mid({0, 1, 2, 3, 4, 5, 6, 7}, out + 3 * 26);
  }
  else if (type == 7)
  {
    for (int e = 0; e < 6; ++e)
      mid({kTetEdge[e][0], kTetEdge[e][1]}, out + 3 * (4 + e));
  }
}

Quadrature recommendedQuadrature(int feType)
{
  // 1D Gauss abscissae as stored by the reference (src/qr.cpp:27,38)
  double const a2 = 0.5773502691896258;
  double const a3 = 0.774596669241483;
  Quadrature q;
  FeInfo const & f = feInfo(feType);
  q.dim = f.dim;
  auto tensor = [&](int npts1)
  {
    double const x2[2] = {-a2, a2}, w2[2] = {1., 1.};
    double const x3[3] = {-a3, 0., a3}, w3[3] = {5., 8., 5.};
    double const * x = npts1 == 2 ? x2 : x3;
    double const * w = npts1 == 2 ? w2 : w3;
    double const wscale = npts1 == 2 ? 1. : (f.dim == 1 ? 9. : (f.dim == 2 ? 81. : 729.));
    int const nz = f.dim > 2 ? npts1 : 1, ny = f.dim > 1 ? npts1 : 1;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < npts1; ++i)
        {
          q.xi.push_back(x[i]);
          double wt = w[i];
          if (f.dim > 1)
          {
            q.xi.push_back(x[j]);
            wt *= w[j];
          }
          if (f.dim > 2)
          {
            q.xi.push_back(x[k]);
            wt *= w[k];
          }
          q.w.push_back(wt / wscale);
        }
  };
  switch (feType)
  {
  case 1:
  case 2:
  case 4: tensor(2); break;
  case 3:
  case 5: tensor(3); break;
  case 6:
  {
    double const a = 0.1381966011250105, b = 0.5854101966249685;
    double const pts[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    for (auto const & p: pts)
    {
      q.xi.insert(q.xi.end(), p, p + 3);
      q.w.push_back(.25 / 6);
    }
    break;
  }
  case 7:
  {
    // GaussQR<Tetrahedron, 11>, src/qr.cpp:222-243 (values as printed there)
    double const a = 0.0714285714285714, b = 0.7857142857142857;
    double const c = 0.399403576166799, d = 0.100596423833201;
    double const pts[11][3] = {{.25, .25, .25}, {a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}, {c, c, d},
                               {c, d, c},       {d, c, c}, {c, d, d}, {d, c, d}, {d, d, c}};
    double const wt[11] = {-0.0131555555555556, 0.0076222222222222, 0.0076222222222222, 0.0076222222222222,
                           0.0076222222222222,  0.024888888888889,  0.024888888888889,  0.024888888888889,
                           0.024888888888889,   0.024888888888889,  0.024888888888889};
    for (int k = 0; k < 11; ++k)
    {
      q.xi.insert(q.xi.end(), pts[k], pts[k] + 3);
      q.w.push_back(wt[k]);
    }
    break;
  }
  default: throw std::invalid_argument("unknown FE type");
  }
  q.nq = (int)q.w.size();
  return q;
}

double geoPhi(int geo, int v, double const * xi)
{
  switch (geo)
  {
  case kLine: return feInfo(1).phi(v, xi);
  case kQuad: return feInfo(2).phi(v, xi);
  case kHex: return feInfo(4).phi(v, xi);
  case kTet: return feInfo(6).phi(v, xi);
  }
  return 0.;
}

void geoGrad(int geo, int v, double const * xi, double * g)
{
  switch (geo)
  {
  case kLine: feInfo(1).grad(v, xi, g); return;
  case kQuad: feInfo(2).grad(v, xi, g); return;
  case kHex: feInfo(4).grad(v, xi, g); return;
  case kTet: feInfo(6).grad(v, xi, g); return;
  }
}

} // namespace pxg
