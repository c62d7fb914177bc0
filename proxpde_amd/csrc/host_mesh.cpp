// Host mesh and dof-map layer: the setup side of the path (runs once per mesh).
//
// Structured builders follow the reference's point and element numbering
// (src/mesh.cpp:17-535); boundary facets are emitted directly from the grid
// structure for lines, quads and hexes and found by facet hashing for tets,
// then marked with the bounding-box rules of markFacetsCube
// (include/mesh.hpp:443-481).  The dof map is the reference's first-visit
// numbering (include/dof.hpp:112-329) with open-addressing hash tables in
// place of std::map<std::set<id>>.
#include "../../include/proxpde_amd.h"
#include "../../include/proxpde_amd_host.h"
#include "fe.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <map>
#include <fstream>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

struct pxh_mesh
{
  int geo = 0;
  int npe = 0;
  std::vector<double> xyz;  // npts x 3
  std::vector<uint32_t> conn; // nelem x npe
  struct BFacet
  {
    uint32_t elem;
    int side;
    int marker;
  };
  std::vector<BFacet> bfacets;
  double origin[3], length[3];
  int64_t npts() const { return (int64_t)xyz.size() / 3; }
  int64_t nelem() const { return npe ? (int64_t)conn.size() / npe : 0; }
};

namespace
{
std::string g_hostError; // message of the last failed host call
using u32 = uint32_t;
using u64 = uint64_t;

// open-addressing hash map from a 128-bit key to a dof id
class KeyMap
{
public:
  explicit KeyMap(size_t expected)
  {
    size_t cap = 16;
    while (cap < 2 * expected)
      cap <<= 1;
    keys_.assign(2 * cap, ~0ull);
    vals_.assign(cap, 0);
    mask_ = cap - 1;
  }
  // returns pointer to the value slot and whether it was inserted
  u32 * find_or_insert(u64 k0, u64 k1, bool & inserted)
  {
    u64 h = (k0 * 0x9E3779B97F4A7C15ull) ^ (k1 + 0x632BE59BD9B4E019ull + (k0 << 6) + (k0 >> 2));
    h ^= h >> 29;
    size_t i = h & mask_;
    while (true)
    {
      if (keys_[2 * i] == ~0ull && keys_[2 * i + 1] == ~0ull)
      {
        keys_[2 * i] = k0;
        keys_[2 * i + 1] = k1;
        inserted = true;
        return &vals_[i];
      }
      if (keys_[2 * i] == k0 && keys_[2 * i + 1] == k1)
      {
        inserted = false;
        return &vals_[i];
      }
      i = (i + 1) & mask_;
    }
  }

private:
  std::vector<u64> keys_;
  std::vector<u32> vals_;
  size_t mask_;
};

// sorted vertex ids packed into a 128-bit key (unused slots = 0xffffffff)
inline void packSorted(u32 * v, int n, u64 & k0, u64 & k1)
{
  std::sort(v, v + n);
  u32 w[4] = {~0u, ~0u, ~0u, ~0u};
  for (int i = 0; i < n; ++i)
    w[i] = v[i];
  k0 = ((u64)w[0] << 32) | w[1];
  k1 = ((u64)w[2] << 32) | w[3];
}

void gridPoints(pxh_mesh & m, int64_t px, int64_t py, int64_t pz, double const * h, bool useZ)
{
  m.xyz.resize(3 * px * py * pz);
  for (int64_t k = 0; k < pz; ++k)
    for (int64_t j = 0; j < py; ++j)
      for (int64_t i = 0; i < px; ++i)
      {
        int64_t const id = i + px * (j + py * k);
        m.xyz[3 * id + 0] = m.origin[0] + i * h[0];
        m.xyz[3 * id + 1] = m.origin[1] + j * h[1];
        m.xyz[3 * id + 2] = m.origin[2] + (useZ ? k * h[2] : 0.0);
      }
}

int markerOf(pxh_mesh const & m, uint32_t e, int side)
{
  pxg::GeoInfo const & g = pxg::geoInfo(m.geo);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int p: g.facets[side])
  {
    double const * x = &m.xyz[3ull * m.conn[(size_t)e * m.npe + p]];
    for (int k = 0; k < 3; ++k)
    {
      lo[k] = std::min(lo[k], x[k]);
      hi[k] = std::max(hi[k], x[k]);
    }
  }
  double const * o = m.origin;
  double const * L = m.length;
  if (std::fabs(hi[1] - o[1]) < 1e-6 * L[1])
    return PXH_SIDE_BOTTOM;
  if (std::fabs(lo[0] - o[0] - L[0]) < 1e-6 * L[0])
    return PXH_SIDE_RIGHT;
  if (std::fabs(lo[1] - o[1] - L[1]) < 1e-6 * L[1])
    return PXH_SIDE_TOP;
  if (std::fabs(hi[0] - o[0]) < 1e-6 * L[0])
    return PXH_SIDE_LEFT;
  if (std::fabs(hi[2] - o[2]) < 1e-6 * L[2])
    return PXH_SIDE_BACK;
  if (std::fabs(lo[2] - o[2] - L[2]) < 1e-6 * L[2])
    return PXH_SIDE_FRONT;
  return 0;
}

void markAll(pxh_mesh & m)
{
  for (auto & f: m.bfacets)
    f.marker = markerOf(m, f.elem, f.side);
}

// boundary facets of an unstructured mesh: facets seen exactly once, in order
// of first visit (element order, then the reference side order)
void boundaryFacets(pxh_mesh & m)
{
  pxg::GeoInfo const & g = pxg::geoInfo(m.geo);
  int64_t const ne = m.nelem();
  int const nf = (int)g.facets.size();
  KeyMap map((size_t)ne * nf);
  std::vector<std::pair<u32, int>> first; // (elem, side) of the first visit
  std::vector<int> count;
  first.reserve(ne * nf / 2 + 16);
  for (int64_t e = 0; e < ne; ++e)
    for (int s = 0; s < nf; ++s)
    {
      u32 v[4];
      int const np = (int)g.facets[s].size();
      for (int p = 0; p < np; ++p)
        v[p] = m.conn[e * m.npe + g.facets[s][p]];
      u64 k0, k1;
      packSorted(v, np, k0, k1);
      bool ins;
      u32 * slot = map.find_or_insert(k0, k1, ins);
      if (ins)
      {
        *slot = (u32)first.size();
        first.push_back({(u32)e, s});
        count.push_back(1);
      }
      else
        count[*slot]++;
    }
  for (size_t k = 0; k < first.size(); ++k)
    if (count[k] == 1)
      m.bfacets.push_back({first[k].first, first[k].second, 0});
}

// ---- GMSH 2.2 ASCII reader (the reference's readGMSH, include/mesh.hpp:785-1062)
// element type codes of the format: linear / quadratic volume types and the
// number of nodes listed for each
struct GmshType
{
  int lin, quad, quadNodes;
};
GmshType gmshTypeOf(int geo)
{
  switch (geo)
  {
  case pxg::kLine: return {1, 8, 3};
  case pxg::kQuad: return {3, 10, 9};
  case pxg::kHex: return {5, 12, 27};
  case pxg::kTet: return {4, 11, 10};
  default: return {-1, -1, 0};
  }
}
// facet geometry of each volume geometry (line facets are points)
GmshType gmshFacetTypeOf(int geo)
{
  switch (geo)
  {
  case pxg::kLine: return {15, 15, 1};
  case pxg::kQuad: return {1, 8, 3};
  case pxg::kHex: return {3, 10, 9};
  case pxg::kTet: return {2, 9, 6};
  default: return {-1, -1, 0};
  }
}
// ridges (edges of 3D meshes, points of 2D meshes) are skipped
GmshType gmshRidgeTypeOf(int geo)
{
  switch (geo)
  {
  case pxg::kQuad: return {15, 15, 1};
  case pxg::kHex:
  case pxg::kTet: return {1, 8, 3};
  default: return {-1, -1, 0};
  }
}
int nodesOfType(int type)
{
  switch (type)
  {
  case 15: return 1;
  case 1: return 2;
  case 2: return 3;
  case 3: return 4;
  case 4: return 4;
  case 5: return 8;
  case 8: return 3;
  case 9: return 6;
  case 10: return 9;
  case 11: return 10;
  case 12: return 27;
  default: return -1;
  }
}

void readGmsh(pxh_mesh & m, char const * path)
{
  std::ifstream in(path);
  if (!in)
    throw std::runtime_error(std::string("mesh file not found: ") + path);
  std::string buf;
  auto expect = [&](char const * tok)
  {
    in >> buf;
    if (buf != tok)
      throw std::runtime_error(std::string("gmsh: expected ") + tok + ", found " + buf);
  };
  expect("$MeshFormat");
  std::string version;
  int filetype = -1, datasize = -1;
  in >> version >> filetype >> datasize;
  if (version != "2.2" || filetype != 0 || datasize != 8)
    throw std::runtime_error("gmsh: only the 2.2 ASCII format is read");
  expect("$EndMeshFormat");
  GmshType const vt = gmshTypeOf(m.geo), ft = gmshFacetTypeOf(m.geo), rt = gmshRidgeTypeOf(m.geo);
  pxg::GeoInfo const & g = pxg::geoInfo(m.geo);
  int const npe = g.npts, nfp = (int)g.facets[0].size();
  std::vector<double> pts;      // file nodes
  std::vector<int> used;        // elements touching each file node
  std::vector<u32> vconn;       // volume connectivity, file node ids
  std::vector<std::pair<std::vector<u32>, int>> ffacets; // file facets (sorted ids, marker)
  while (in >> buf)
  {
    if (buf == "$Nodes")
    {
      int64_t nn = 0;
      in >> nn;
      pts.resize(3 * nn);
      used.assign(nn, 0);
      for (int64_t k = 0; k < nn; ++k)
      {
        int64_t id;
        in >> id >> pts[3 * k] >> pts[3 * k + 1] >> pts[3 * k + 2];
        if (id != k + 1)
          throw std::runtime_error("gmsh: node ids must be consecutive from 1");
      }
      expect("$EndNodes");
    }
    else if (buf == "$Elements")
    {
      int64_t ne = 0;
      in >> ne;
      for (int64_t k = 0; k < ne; ++k)
      {
        int64_t id;
        int type, ntags;
        in >> id >> type >> ntags;
        std::vector<int> tags(ntags);
        for (int t = 0; t < ntags; ++t)
          in >> tags[t];
        int const nn = nodesOfType(type);
        if (nn < 0)
          throw std::runtime_error("gmsh: unknown element type " + std::to_string(type));
        std::vector<u32> c(nn);
        for (int t = 0; t < nn; ++t)
        {
          int64_t v;
          in >> v;
          c[t] = (u32)(v - 1);
        }
        if (type == vt.lin || type == vt.quad)
        {
          // quadratic elements: the vertices come first, the rest is skipped
          for (int t = 0; t < npe; ++t)
          {
            vconn.push_back(c[t]);
            used.at(c[t])++;
          }
        }
        else if (type == ft.lin || type == ft.quad)
        {
          std::vector<u32> f(c.begin(), c.begin() + nfp);
          std::sort(f.begin(), f.end());
          ffacets.push_back({f, ntags > 0 ? tags[0] : 0});
        }
        else if (type == rt.lin || type == rt.quad)
          continue;
        else
          throw std::runtime_error("gmsh: element type " + std::to_string(type) + " does not fit the mesh");
      }
      expect("$EndElements");
    }
    else if (!buf.empty() && buf[0] == '$')
    {
      // any other section ($PhysicalNames, ...) is skipped to its end tag
      std::string const end = "$End" + buf.substr(1);
      while (in >> buf && buf != end)
      {
      }
    }
    else
      throw std::runtime_error("gmsh: unexpected token " + buf);
  }
  // points touched by volume elements, renumbered in file order
  std::vector<int64_t> newId(used.size(), -1);
  m.xyz.clear();
  int64_t cnt = 0;
  for (size_t k = 0; k < used.size(); ++k)
    if (used[k] > 0)
    {
      newId[k] = cnt++;
      m.xyz.insert(m.xyz.end(), {pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]});
    }
  m.npe = npe;
  m.conn.resize(vconn.size());
  for (size_t k = 0; k < vconn.size(); ++k)
    m.conn[k] = (u32)newId[vconn[k]];
  boundaryFacets(m);
  // markers of the boundary facets listed in the file (matched by vertex set)
  std::map<std::vector<u32>, int> fileMarker;
  for (auto & f: ffacets)
  {
    std::vector<u32> key;
    for (u32 v: f.first)
    {
      if (newId.at(v) < 0)
        throw std::runtime_error("gmsh: facet on a node of no element");
      key.push_back((u32)newId[v]);
    }
    std::sort(key.begin(), key.end());
    fileMarker[key] = f.second;
  }
  for (auto & b: m.bfacets)
  {
    std::vector<u32> key;
    for (int p: g.facets[b.side])
      key.push_back(m.conn[(size_t)b.elem * m.npe + p]);
    std::sort(key.begin(), key.end());
    auto it = fileMarker.find(key);
    b.marker = it == fileMarker.end() ? 0 : it->second;
  }
}

} // namespace

extern "C" {

int pxh_mesh_structured(pxh_mesh ** out, int kind, double const * origin, double const * length, int nx, int ny,
                        int nz)
{
  if (!out || !origin || !length || nx <= 0)
    return PXG_E_INVALID;
  try
  {
    auto * m = new pxh_mesh;
    for (int k = 0; k < 3; ++k)
    {
      m->origin[k] = origin[k];
      m->length[k] = length[k];
    }
    switch (kind)
    {
    case PXH_MESH_LINE:
    {
      m->geo = pxg::kLine;
      m->npe = 2;
      double const h[3] = {length[0] / nx, length[1] / nx, length[2] / nx};
      m->xyz.resize(3 * (nx + 1));
      for (int p = 0; p <= nx; ++p)
        for (int k = 0; k < 3; ++k)
          m->xyz[3 * p + k] = origin[k] + p * h[k];
      for (int e = 0; e < nx; ++e)
        m->conn.insert(m->conn.end(), {(u32)e, (u32)e + 1});
      m->bfacets = {{0u, 0, PXH_SIDE_LEFT}, {(u32)nx - 1, 1, PXH_SIDE_RIGHT}};
      break;
    }
    case PXH_MESH_SQUARE:
    {
      if (ny <= 0)
        return delete m, PXG_E_INVALID;
      m->geo = pxg::kQuad;
      m->npe = 4;
      double const h[3] = {length[0] / nx, length[1] / ny, 1.};
      u32 const px = nx + 1;
      gridPoints(*m, px, ny + 1, 1, h, false);
      m->conn.reserve(4ull * nx * ny);
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
          u32 const b = i + j * px;
          m->conn.insert(m->conn.end(), {b, b + 1, b + px + 1, b + px});
          u32 const e = i + j * nx;
          if (j == 0)
            m->bfacets.push_back({e, 0, 0});
          if (i == nx - 1)
            m->bfacets.push_back({e, 1, 0});
          if (j == ny - 1)
            m->bfacets.push_back({e, 2, 0});
          if (i == 0)
            m->bfacets.push_back({e, 3, 0});
        }
      markAll(*m);
      break;
    }
    case PXH_MESH_CUBE_HEX:
    case PXH_MESH_CUBE_TET5:
    case PXH_MESH_CUBE_TET6:
    {
      if (ny <= 0 || nz <= 0)
        return delete m, PXG_E_INVALID;
      double const h[3] = {length[0] / nx, length[1] / ny, length[2] / nz};
      u32 const dx = 1, dy = nx + 1, dz = (u32)(nx + 1) * (ny + 1);
      gridPoints(*m, nx + 1, ny + 1, nz + 1, h, true);
      bool const hex = kind == PXH_MESH_CUBE_HEX;
      m->geo = hex ? pxg::kHex : pxg::kTet;
      m->npe = hex ? 8 : 4;
      int const per = hex ? 1 : (kind == PXH_MESH_CUBE_TET5 ? 5 : 6);
      m->conn.reserve((size_t)nx * ny * nz * per * m->npe);
      // 5-tet split by cube parity (src/mesh.cpp:418-448) and the Kuhn split
      static int const odd5[5][4] = {{1, 3, 4, 6}, {0, 1, 3, 4}, {1, 2, 3, 6}, {1, 4, 5, 6}, {3, 4, 6, 7}};
      static int const even5[5][4] = {{0, 2, 5, 7}, {0, 1, 2, 5}, {0, 2, 3, 7}, {0, 4, 5, 7}, {2, 5, 6, 7}};
      static int const kuhn6[6][4] = {{0, 1, 2, 6}, {0, 2, 3, 6}, {0, 3, 7, 6},
                                      {0, 7, 4, 6}, {0, 4, 5, 6}, {0, 5, 1, 6}};
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i)
          {
            u32 const b = i + j * dy + k * dz;
            u32 const c[8] = {b, b + dx, b + dx + dy, b + dy, b + dz, b + dx + dz, b + dx + dy + dz, b + dy + dz};
            if (hex)
            {
              m->conn.insert(m->conn.end(), c, c + 8);
              u32 const e = i + (u32)nx * (j + (u32)ny * k);
              // sides: 0 z-, 1 x-, 2 y-, 3 x+, 4 y+, 5 z+
              if (k == 0)
                m->bfacets.push_back({e, 0, 0});
              if (i == 0)
                m->bfacets.push_back({e, 1, 0});
              if (j == 0)
                m->bfacets.push_back({e, 2, 0});
              if (i == nx - 1)
                m->bfacets.push_back({e, 3, 0});
              if (j == ny - 1)
                m->bfacets.push_back({e, 4, 0});
              if (k == nz - 1)
                m->bfacets.push_back({e, 5, 0});
            }
            else if (per == 5)
            {
              auto const & t = ((i + j + k) % 2) ? odd5 : even5;
              for (auto const & r: t)
                m->conn.insert(m->conn.end(), {c[r[0]], c[r[1]], c[r[2]], c[r[3]]});
            }
            else
              for (auto const & r: kuhn6)
                m->conn.insert(m->conn.end(), {c[r[0]], c[r[1]], c[r[2]], c[r[3]]});
          }
      if (!hex)
        boundaryFacets(*m);
      markAll(*m);
      break;
    }
    default: delete m; return PXG_E_INVALID;
    }
    *out = m;
    return PXG_OK;
  }
  catch (std::bad_alloc const &)
  {
    return PXG_E_OOM;
  }
  catch (...)
  {
    return PXG_E_INVALID;
  }
}

int pxh_mesh_gmsh(pxh_mesh ** out, char const * path, int geo)
{
  if (!out || !path || geo < pxg::kLine || geo > pxg::kTet)
    return PXG_E_INVALID;
  *out = nullptr;
  auto * m = new (std::nothrow) pxh_mesh;
  if (!m)
    return PXG_E_OOM;
  try
  {
    m->geo = geo;
    for (int k = 0; k < 3; ++k)
      m->origin[k] = 0., m->length[k] = 1.;
    readGmsh(*m, path);
  }
  catch (std::exception const & e)
  {
    g_hostError = e.what();
    delete m;
    return PXG_E_INVALID;
  }
  *out = m;
  return PXG_OK;
}

const char * pxh_last_error(void) { return g_hostError.c_str(); }

int pxh_mesh_destroy(pxh_mesh * m)
{
  delete m;
  return PXG_OK;
}

int pxh_mesh_info(pxh_mesh const * m, int64_t * npts, int64_t * nelem, int * npe, int64_t * nbf)
{
  if (!m)
    return PXG_E_INVALID;
  if (npts)
    *npts = m->npts();
  if (nelem)
    *nelem = m->nelem();
  if (npe)
    *npe = m->npe;
  if (nbf)
    *nbf = (int64_t)m->bfacets.size();
  return PXG_OK;
}

int pxh_mesh_get(pxh_mesh const * m, double * coords, uint32_t * conn)
{
  if (!m)
    return PXG_E_INVALID;
  if (coords)
    std::memcpy(coords, m->xyz.data(), m->xyz.size() * sizeof(double));
  if (conn)
    std::memcpy(conn, m->conn.data(), m->conn.size() * sizeof(uint32_t));
  return PXG_OK;
}

int pxh_mesh_set_coords(pxh_mesh * m, double const * coords)
{
  if (!m || !coords)
    return PXG_E_INVALID;
  std::memcpy(m->xyz.data(), coords, m->xyz.size() * sizeof(double));
  return PXG_OK;
}

int pxh_mesh_bfacets(pxh_mesh const * m, int32_t * out)
{
  if (!m || !out)
    return PXG_E_INVALID;
  for (size_t k = 0; k < m->bfacets.size(); ++k)
  {
    out[3 * k] = (int32_t)m->bfacets[k].elem;
    out[3 * k + 1] = m->bfacets[k].side;
    out[3 * k + 2] = m->bfacets[k].marker;
  }
  return PXG_OK;
}

int pxh_dofmap(pxh_mesh const * m, int feType, int64_t * ndofs, int * nlocal, uint32_t * elemMap)
{
  if (!m)
    return PXG_E_INVALID;
  try
  {
    pxg::FeInfo const & fe = pxg::feInfo(feType);
    if (fe.geo != m->geo)
      return PXG_E_INVALID;
    pxg::GeoInfo const & g = pxg::geoInfo(fe.geo);
    int64_t const ne = m->nelem();
    if (nlocal)
      *nlocal = fe.n;
    std::vector<u32> tmp;
    u32 * map = elemMap;
    if (!map)
    {
      tmp.resize((size_t)ne * fe.n);
      map = tmp.data();
    }
    u32 const unset = ~0u;
    std::vector<u32> ptDof(m->npts(), unset);
    KeyMap edges(fe.dofPlace[2] ? (size_t)ne * g.edges.size() : 1);
    KeyMap faces(fe.dofPlace[1] ? (size_t)ne * g.faces.size() : 1);
    u32 next = 0;
    for (int64_t e = 0; e < ne; ++e)
    {
      u32 const * v = &m->conn[e * m->npe];
      u32 * row = &map[e * fe.n];
      int loc = 0;
      if (fe.dofPlace[3])
        for (int p = 0; p < g.npts; ++p)
        {
          if (ptDof[v[p]] == unset)
            ptDof[v[p]] = next++;
          row[loc++] = ptDof[v[p]];
        }
      if (fe.dofPlace[2])
        for (auto const & ed: g.edges)
        {
          u32 a = v[ed[0]], b = v[ed[1]];
          if (a > b)
            std::swap(a, b);
          bool ins;
          u32 * s = edges.find_or_insert(((u64)a << 32) | b, 0, ins);
          if (ins)
            *s = next++;
          row[loc++] = *s;
        }
      if (fe.dofPlace[1])
        for (auto const & fc: g.faces)
        {
          u32 ids[4];
          int const np = (int)fc.size();
          for (int p = 0; p < np; ++p)
            ids[p] = v[fc[p]];
          u64 k0, k1;
          packSorted(ids, np, k0, k1);
          bool ins;
          u32 * s = faces.find_or_insert(k0, k1, ins);
          if (ins)
            *s = next++;
          row[loc++] = *s;
        }
      if (fe.dofPlace[0])
        row[loc++] = next++;
    }
    if (ndofs)
      *ndofs = next;
    return PXG_OK;
  }
  catch (std::bad_alloc const &)
  {
    return PXG_E_OOM;
  }
  catch (...)
  {
    return PXG_E_INVALID;
  }
}

int pxh_bc_dofs(pxh_mesh const * m, int feType, uint32_t const * elemMap, int marker, uint32_t * dofs,
                int64_t * count)
{
  if (!m || !elemMap || !count)
    return PXG_E_INVALID;
  try
  {
    pxg::FeInfo const & fe = pxg::feInfo(feType);
    std::vector<u32> out;
    for (auto const & f: m->bfacets)
      if (f.marker == marker)
        for (int k: fe.dofOnFacet[f.side])
          out.push_back(elemMap[(size_t)f.elem * fe.n + k]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    if (dofs)
    {
      if (*count < (int64_t)out.size())
        return PXG_E_INVALID;
      std::memcpy(dofs, out.data(), out.size() * sizeof(u32));
    }
    *count = (int64_t)out.size();
    return PXG_OK;
  }
  catch (...)
  {
    return PXG_E_INVALID;
  }
}

int pxh_dof_coords(pxh_mesh const * m, int feType, uint32_t const * elemMap, int64_t ndofs, double * out)
{
  if (!m || !elemMap || !out)
    return PXG_E_INVALID;
  try
  {
    pxg::FeInfo const & fe = pxg::feInfo(feType);
    std::vector<double> vtx(3 * m->npe), pts(3 * fe.n);
    for (int64_t e = 0; e < m->nelem(); ++e)
    {
      for (int p = 0; p < m->npe; ++p)
        for (int k = 0; k < 3; ++k)
          vtx[3 * p + k] = m->xyz[3ull * m->conn[e * m->npe + p] + k];
      fe.dofPoints(vtx.data(), pts.data());
      for (int i = 0; i < fe.n; ++i)
      {
        u32 const d = elemMap[e * fe.n + i];
        if (d >= ndofs)
          return PXG_E_INVALID;
        for (int k = 0; k < 3; ++k)
          out[3ull * d + k] = pts[3 * i + k];
      }
    }
    return PXG_OK;
  }
  catch (...)
  {
    return PXG_E_INVALID;
  }
}

} // extern "C"
