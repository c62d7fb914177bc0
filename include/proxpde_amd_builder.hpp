/* proxpde_amd_builder.hpp -- C++20 façade with the call shape of the
 * reference's Builder (include/builder.hpp:12-385), over the C ABI of
 * proxpde_amd.h and the host layer of proxpde_amd_host.h.
 *
 *   Mesh<Hexahedron> mesh;
 *   buildHyperCube(mesh, {0., 0., 0.}, {1., 1., 1.}, {n, n, n});
 *   FESpace_T feSpace{mesh};                                   // GPU context
 *   auto const bcs = std::vector{BCEss{feSpace, side::LEFT, zero}, ...};
 *   Builder<StorageType::RowMajor> builder{feSpace.dof.size};
 *   builder.buildLhs(std::tuple{AssemblyStiffness(1.0, feSpace)}, bcs);
 *   builder.buildRhs(std::tuple{AssemblyRhsAnalytic(rhs, feSpace)}, bcs);
 *   builder.closeMatrix();                                      // builder.A, builder.b
 *
 * Every call forwards to the device: the assembly terms become pxg_term
 * entries of one pxg_assemble, the FEVar of an FE term is the context's
 * coefficient field, BCEss lists become pxg_set_dirichlet, the analytic
 * right-hand side is evaluated on the host at pxg_get_qpoints() and summed on
 * the device (PXG_RHS_QVALUES), closeMatrix is pxg_close with the reference's
 * prune(1, 1e-16) followed by a copy of the CSR and b to the host.
 * Errors throw proxpde_amd::Error with the library's message.
 *
 * Scope: scalar Lagrange spaces with the recommended quadrature (the element
 * and rule pairs of PXG_FE_*), FP64, row-major storage (the reference's
 * default ClmMajor would need the transposed layout; it is rejected at
 * compile time).  Link with -lproxpde_amd.
 */
#ifndef PROXPDE_AMD_BUILDER_HPP
#define PROXPDE_AMD_BUILDER_HPP

#include "proxpde_amd.h"
#include "proxpde_amd_host.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace proxpde_amd
{

using uint = unsigned int;
using marker_T = int;
using Vec3 = std::array<double, 3>;
using scalarFun_T = std::function<double(Vec3 const &)>;

struct Error: std::runtime_error
{
  int code;
  Error(int c, std::string const & what): std::runtime_error(what), code(c) {}
};

namespace detail
{
inline void check(int rc, pxg_ctx const * ctx)
{
  if (rc != PXG_OK)
    throw Error(rc, pxg_last_error(ctx));
}
inline void checkHost(int rc)
{
  if (rc != PXG_OK)
    throw Error(rc, pxh_last_error());
}
struct CtxDeleter
{
  void operator()(pxg_ctx * c) const { pxg_destroy(c); }
};
struct MeshDeleter
{
  void operator()(pxh_mesh * m) const { pxh_mesh_destroy(m); }
};
} // namespace detail

// sparse_matrix.hpp:15-19
enum class StorageType : uint8_t
{
  RowMajor,
  ClmMajor,
};

// boundary markers of the structured meshes (mesh.hpp side::*)
namespace side
{
inline constexpr marker_T BOTTOM = PXH_SIDE_BOTTOM;
inline constexpr marker_T RIGHT = PXH_SIDE_RIGHT;
inline constexpr marker_T TOP = PXH_SIDE_TOP;
inline constexpr marker_T LEFT = PXH_SIDE_LEFT;
inline constexpr marker_T BACK = PXH_SIDE_BACK;
inline constexpr marker_T FRONT = PXH_SIDE_FRONT;
} // namespace side

// ---- geometry tags (geo.hpp): the structured mesh each one builds
struct Line
{
  static constexpr int meshKind = PXH_MESH_LINE;
};
struct Quad
{
  static constexpr int meshKind = PXH_MESH_SQUARE;
};
struct Hexahedron
{
  static constexpr int meshKind = PXH_MESH_CUBE_HEX;
};
struct Tetrahedron
{
  static constexpr int meshKind = PXH_MESH_CUBE_TET5; // buildCube<Tetrahedron>: 5 tets per cube
};

// Mesh<Elem> (mesh.hpp): points and elements on the host
template <typename Elem>
struct Mesh
{
  using Elem_T = Elem;
  std::unique_ptr<pxh_mesh, detail::MeshDeleter> h;
};

// buildHyperCube (mesh.hpp:386-422)
template <typename Elem>
void buildHyperCube(Mesh<Elem> & mesh, Vec3 const & origin, Vec3 const & length, std::array<uint, 3> const numElems)
{
  pxh_mesh * m = nullptr;
  detail::checkHost(pxh_mesh_structured(&m, Elem::meshKind, origin.data(), length.data(), (int)numElems[0],
                                        (int)numElems[1], (int)numElems[2]));
  mesh.h.reset(m);
}

// ---- reference elements and their recommended rules (fe.hpp LagrangeFE)
template <int FeType, int NumPts>
struct RefFE
{
  static constexpr int feType = FeType;
  struct RecommendedQR
  {
    static constexpr int numPts = NumPts;
  };
};
template <typename Elem, int Order>
struct LagrangeFE;
template <>
struct LagrangeFE<Line, 1>
{
  using RefFE_T = RefFE<PXG_FE_LINE_P1, 2>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};
template <>
struct LagrangeFE<Quad, 1>
{
  using RefFE_T = RefFE<PXG_FE_QUAD_Q1, 4>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};
template <>
struct LagrangeFE<Quad, 2>
{
  using RefFE_T = RefFE<PXG_FE_QUAD_Q2, 9>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};
template <>
struct LagrangeFE<Hexahedron, 1>
{
  using RefFE_T = RefFE<PXG_FE_HEX_Q1, 8>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};
template <>
struct LagrangeFE<Hexahedron, 2>
{
  using RefFE_T = RefFE<PXG_FE_HEX_Q2, 27>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};
template <>
struct LagrangeFE<Tetrahedron, 1>
{
  using RefFE_T = RefFE<PXG_FE_TET_P1, 4>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};
template <>
struct LagrangeFE<Tetrahedron, 2>
{
  using RefFE_T = RefFE<PXG_FE_TET_P2, 11>;
  using RecommendedQR = RefFE_T::RecommendedQR;
};

// FESpace<Mesh, RefFE, QR> (fespace.hpp): the dof map (DOF::elemMap,
// dof.hpp:73-76) and one GPU context holding the mesh, the element and its
// rule, on `device`
template <typename Mesh_, typename RefFE_, typename QR_>
struct FESpace
{
  using Mesh_T = Mesh_;
  using RefFE_T = RefFE_;
  using QR_T = QR_;
  static constexpr uint dim = 1;

  struct DOF
  {
    uint size = 0;
    std::vector<uint32_t> elemMap; // n_elem x numDOFs
    int numDOFs = 0;
  };

  Mesh_T const * mesh = nullptr;
  DOF dof;
  int64_t nElems = 0;
  std::shared_ptr<pxg_ctx> ctx;

  explicit FESpace(Mesh_T const & m, int device = 0): mesh(&m)
  {
    pxh_mesh const * h = m.h.get();
    int64_t npts = 0, nb = 0;
    int npe = 0;
    detail::checkHost(pxh_mesh_info(h, &npts, &nElems, &npe, &nb));
    std::vector<double> coords((size_t)npts * 3);
    std::vector<uint32_t> conn((size_t)nElems * npe);
    detail::checkHost(pxh_mesh_get(h, coords.data(), conn.data()));
    int64_t nd = 0;
    detail::checkHost(pxh_dofmap(h, RefFE_T::feType, &nd, &dof.numDOFs, nullptr));
    dof.size = (uint)nd;
    dof.elemMap.resize((size_t)nElems * dof.numDOFs);
    detail::checkHost(pxh_dofmap(h, RefFE_T::feType, &nd, &dof.numDOFs, dof.elemMap.data()));
    pxg_ctx * c = nullptr;
    detail::check(pxg_create(&c, device), nullptr);
    ctx.reset(c, detail::CtxDeleter{});
    detail::check(pxg_set_fe(c, RefFE_T::feType, 0, nullptr, nullptr), c);
    detail::check(pxg_set_mesh(c, npts, coords.data(), nElems, npe, conn.data()), c);
    detail::check(pxg_set_dofmap(c, nd, dof.elemMap.data()), c);
  }

  // dof coordinates (n_dofs x 3), for interpolating analytic functions
  std::vector<double> dofCoords() const
  {
    std::vector<double> x((size_t)dof.size * 3);
    detail::checkHost(pxh_dof_coords(mesh->h.get(), RefFE_T::feType, dof.elemMap.data(), dof.size, x.data()));
    return x;
  }
};

// interpolateAnalyticFunction (fe.hpp) at the dofs of a Lagrange space
template <typename FESpace_T>
void interpolateAnalyticFunction(scalarFun_T const & f, FESpace_T const & fe, std::vector<double> & v)
{
  auto const x = fe.dofCoords();
  v.resize(fe.dof.size);
  for (size_t i = 0; i < v.size(); ++i)
    v[i] = f(Vec3{x[3 * i], x[3 * i + 1], x[3 * i + 2]});
}

// FEVar (var.hpp:56-130): nodal values on a space
template <typename FESpace_>
struct FEVar
{
  using FESpace_T = FESpace_;
  std::string name;
  FESpace_T const * feSpace = nullptr;
  std::vector<double> data;

  FEVar(std::string n, FESpace_T const & fe): name(std::move(n)), feSpace(&fe), data(fe.dof.size, 0.) {}
  explicit FEVar(FESpace_T const & fe): FEVar("none", fe) {}
  FEVar & operator<<(scalarFun_T const & f)
  {
    interpolateAnalyticFunction(f, *feSpace, data);
    return *this;
  }
};

// BCEss (bc.hpp:31-309): the dofs on a marked boundary and their values
template <typename FESpace_>
struct BCEss
{
  using FESpace_T = FESpace_;
  FESpace_T const * feSpace = nullptr;
  marker_T marker = 0;
  std::vector<uint32_t> dofs;   // sorted, unique
  std::vector<double> values;   // per entry of dofs
  double diag = 1.0;

  BCEss(FESpace_T const & fe, marker_T m, scalarFun_T const & fun = [](Vec3 const &) { return 0.; }):
      feSpace(&fe), marker(m)
  {
    int64_t cnt = fe.dof.size;
    dofs.resize((size_t)cnt);
    detail::checkHost(pxh_bc_dofs(fe.mesh->h.get(), FESpace_T::RefFE_T::feType, fe.dof.elemMap.data(), m,
                                  dofs.data(), &cnt));
    dofs.resize((size_t)cnt);
    *this << fun;
  }
  BCEss & operator<<(scalarFun_T const & f)
  {
    auto const x = feSpace->dofCoords();
    values.resize(dofs.size());
    for (size_t k = 0; k < dofs.size(); ++k)
    {
      size_t const i = dofs[k];
      values[k] = f(Vec3{x[3 * i], x[3 * i + 1], x[3 * i + 2]});
    }
    return *this;
  }
  bool isConstrained(uint32_t id) const { return std::binary_search(dofs.begin(), dofs.end(), id); }
};

// ---- assemblies (assembly_lhs.hpp, assembly_rhs.hpp)
template <typename FESpace_>
struct AssemblyStiffness // assembly_lhs.hpp:36-75
{
  using FESpace_T = FESpace_;
  double coef;
  FESpace_T const * feSpace;
  AssemblyStiffness(double c, FESpace_T const & fe): coef(c), feSpace(&fe) {}
  pxg_term term() const { return {PXG_TERM_STIFFNESS, coef}; }
  std::vector<double> const * field() const { return nullptr; }
};

template <typename FESpace_, typename Coef>
struct AssemblyStiffnessFE // assembly_lhs.hpp:77-113 (kappa at the points, no scalar factor)
{
  using FESpace_T = FESpace_;
  Coef * coef;
  FESpace_T const * feSpace;
  AssemblyStiffnessFE(Coef & c, FESpace_T const & fe): coef(&c), feSpace(&fe) {}
  pxg_term term() const { return {PXG_TERM_STIFFNESS_FE, 1.0}; }
  std::vector<double> const * field() const { return &coef->data; }
};

template <typename FESpace_>
struct AssemblyScalarMass // assembly_lhs.hpp:191-231
{
  using FESpace_T = FESpace_;
  double coef;
  FESpace_T const * feSpace;
  AssemblyScalarMass(double c, FESpace_T const & fe): coef(c), feSpace(&fe) {}
  pxg_term term() const { return {PXG_TERM_MASS, coef}; }
  std::vector<double> const * field() const { return nullptr; }
};
template <typename FESpace_>
using AssemblyMass = AssemblyScalarMass<FESpace_>;

template <typename FESpace_, typename Coef>
struct AssemblyMassFE // assembly_lhs.hpp:233-283 (coef * kappa at the points)
{
  using FESpace_T = FESpace_;
  double coef;
  Coef * coefFE;
  FESpace_T const * feSpace;
  AssemblyMassFE(double c, Coef & cFE, FESpace_T const & fe): coef(c), coefFE(&cFE), feSpace(&fe) {}
  AssemblyMassFE(Coef & cFE, FESpace_T const & fe): coef(1.0), coefFE(&cFE), feSpace(&fe) {}
  pxg_term term() const { return {PXG_TERM_MASS_FE, coef}; }
  std::vector<double> const * field() const { return &coefFE->data; }
};

template <typename FESpace_>
struct AssemblyRhsAnalytic // assembly_rhs.hpp:35-90
{
  using FESpace_T = FESpace_;
  scalarFun_T rhs;
  FESpace_T const * feSpace;
  AssemblyRhsAnalytic(scalarFun_T r, FESpace_T const & fe): rhs(std::move(r)), feSpace(&fe) {}
};

// deduction guides (the reference's CTAD: AssemblyStiffness(1.0, feSpace) etc.)
template <typename F>
AssemblyStiffness(double, F const &) -> AssemblyStiffness<F>;
template <typename F, typename C>
AssemblyStiffnessFE(C &, F const &) -> AssemblyStiffnessFE<F, C>;
template <typename F>
AssemblyScalarMass(double, F const &) -> AssemblyScalarMass<F>;
template <typename F, typename C>
AssemblyMassFE(double, C &, F const &) -> AssemblyMassFE<F, C>;
template <typename F, typename C>
AssemblyMassFE(C &, F const &) -> AssemblyMassFE<F, C>;
template <typename F, typename Fun>
AssemblyRhsAnalytic(Fun, F const &) -> AssemblyRhsAnalytic<F>;
template <typename F, typename Fun>
BCEss(F const &, marker_T, Fun) -> BCEss<F>;
template <typename F>
BCEss(F const &, marker_T) -> BCEss<F>;
template <typename F>
FEVar(std::string, F const &) -> FEVar<F>;
template <typename F>
FEVar(F const &) -> FEVar<F>;

// closed matrix on the host: row-major CSR with 32-bit column indices
struct CsrMatrix
{
  int64_t rows = 0, cols = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;
  int64_t nonZeros() const { return (int64_t)val.size(); }
  double coeff(int64_t i, int64_t j) const
  {
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (col[k] == j)
        return val[k];
    return 0.;
  }
};

// Builder (builder.hpp:12-385)
template <StorageType Storage = StorageType::ClmMajor>
struct Builder
{
  static_assert(Storage == StorageType::RowMajor,
                "the device matrix is row-major CSR: use Builder<StorageType::RowMajor>");

  CsrMatrix A;
  std::vector<double> b;

  Builder() = default;
  explicit Builder(uint const size) { init(size); }
  void init(uint const size)
  {
    A = CsrMatrix{};
    A.rows = A.cols = size;
    b.assign(size, 0.);
  }

  // builder.hpp:31-154: all terms in one pass of the element kernels
  template <typename Assemblies, typename BC = BCEss<typename std::tuple_element_t<0, Assemblies>::FESpace_T>>
  void buildLhs(Assemblies const & assemblies, std::vector<BC> const & bcs = {})
  {
    static_assert(std::tuple_size_v<Assemblies> > 0, "we need at least 1 assembly.");
    auto const & fe = *std::get<0>(assemblies).feSpace;
    pxg_ctx * c = bind(fe);
    setBCs(c, bcs);
    std::vector<pxg_term> terms;
    std::vector<double> const * kappa = nullptr;
    std::apply(
        [&](auto const &... as)
        {
          (
              [&](auto const & a)
              {
                if (a.feSpace != &fe)
                  throw Error(PXG_E_INVALID, "all assemblies of a build must be on the same FESpace");
                terms.push_back(a.term());
                if (auto const * f = a.field())
                {
                  if (kappa && kappa != f)
                    throw Error(PXG_E_INVALID, "one FE coefficient field per build");
                  kappa = f;
                }
              }(as),
              ...);
        },
        assemblies);
    if (kappa)
      detail::check(pxg_set_coefficient(c, kappa->data()), c);
    detail::check(pxg_assemble(c, (int)terms.size(), terms.data(), PXG_FP64), c);
  }

  // builder.hpp:279-364: analytic right-hand sides at the quadrature points
  template <typename Assemblies, typename BC = BCEss<typename std::tuple_element_t<0, Assemblies>::FESpace_T>>
  void buildRhs(Assemblies const & assemblies, std::vector<BC> const & bcs = {})
  {
    static_assert(std::tuple_size_v<Assemblies> > 0, "we need at least 1 assembly.");
    auto const & fe = *std::get<0>(assemblies).feSpace;
    using FE_T = std::remove_cvref_t<decltype(fe)>;
    pxg_ctx * c = bind(fe);
    setBCs(c, bcs);
    int64_t const nq = FE_T::QR_T::numPts;
    std::vector<double> xq((size_t)fe.nElems * nq * 3), fq((size_t)fe.nElems * nq);
    detail::check(pxg_get_qpoints(c, xq.data()), c);
    std::apply(
        [&](auto const &... as)
        {
          (
              [&](auto const & a)
              {
                if (a.feSpace != &fe)
                  throw Error(PXG_E_INVALID, "all assemblies of a build must be on the same FESpace");
                for (size_t k = 0; k < fq.size(); ++k)
                  fq[k] = a.rhs(Vec3{xq[3 * k], xq[3 * k + 1], xq[3 * k + 2]});
                detail::check(pxg_assemble_rhs(c, PXG_RHS_QVALUES, nullptr, fq.data()), c);
              }(as),
              ...);
        },
        assemblies);
  }

  // builder.hpp:366-371: sum, prune(1, 1e-16), then A and b on the host
  void closeMatrix()
  {
    if (!ctx_)
      throw Error(PXG_E_STATE, "closeMatrix before any build");
    pxg_ctx * c = ctx_.get();
    int64_t nnz = 0;
    detail::check(pxg_close(c, 1e-16, &nnz), c);
    A.rowptr.assign((size_t)A.rows + 1, 0);
    A.col.resize((size_t)nnz);
    A.val.resize((size_t)nnz);
    detail::check(pxg_get_csr(c, A.rowptr.data(), A.col.data(), A.val.data(), nullptr, b.data()), c);
  }

  // builder.hpp:373-385
  void clear()
  {
    if (ctx_)
      detail::check(pxg_clear(ctx_.get()), ctx_.get());
    init((uint)A.rows);
  }

  // the device context of the space the builder works on (device-side CSR,
  // SpMV, solvers through the C ABI)
  pxg_ctx * context() const { return ctx_.get(); }

private:
  std::shared_ptr<pxg_ctx> ctx_;
  bool patternBuilt_ = false;
  std::vector<uint32_t> bcDofs_;
  std::vector<double> bcVals_;

  template <typename FE>
  pxg_ctx * bind(FE const & fe)
  {
    if (!ctx_)
      ctx_ = fe.ctx;
    else if (ctx_ != fe.ctx)
      throw Error(PXG_E_INVALID, "a Builder works on one FESpace");
    if ((int64_t)fe.dof.size != A.rows)
      throw Error(PXG_E_INVALID, "Builder size differs from the FESpace's dof count");
    if (!patternBuilt_)
    {
      int64_t nnzS = 0;
      detail::check(pxg_build_pattern(ctx_.get(), &nnzS), ctx_.get());
      patternBuilt_ = true;
    }
    return ctx_.get();
  }

  // the dofs of all bcs in list order (a dof fixed twice takes the later
  // value, as the reference's loop over bcs leaves it), sent when changed
  template <typename BC>
  void setBCs(pxg_ctx * c, std::vector<BC> const & bcs)
  {
    std::vector<uint32_t> d;
    std::vector<double> v;
    for (auto const & bc: bcs)
    {
      d.insert(d.end(), bc.dofs.begin(), bc.dofs.end());
      v.insert(v.end(), bc.values.begin(), bc.values.end());
    }
    if (d != bcDofs_ || v != bcVals_)
    {
      detail::check(pxg_set_dirichlet(c, (int64_t)d.size(), d.data(), v.data()), c);
      bcDofs_ = std::move(d);
      bcVals_ = std::move(v);
    }
  }
};

} // namespace proxpde_amd

#endif // PROXPDE_AMD_BUILDER_HPP
