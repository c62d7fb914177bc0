// C ABI of the assembly library (declarations and reference mapping in
// include/proxpde_amd.h).  Every entry point converts exceptions into status
// codes and records the message for pxg_last_error().
#include "kernels.hpp"

#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

using namespace pxg;

// Second vector of the sweeps of pxg_jacobi_noalloc: one copy per GPU, made
// when this code object is loaded on the device and held until the process
// ends.  In host code the name is a host-side stand-in, not a device
// address: host code uses only the address resolved per device at
// pxg_create (noAllocScratch).
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 22; // 4,194,304 entries, 32 MiB of doubles
static __device__ __align__(256) double scratch[kNoAllocMaxRows];

namespace
{
// message of the last failure that had no context to hold it (pxg_create)
thread_local std::string g_lastError;

// device address of `scratch` per device, resolved once per device by the
// first pxg_create on it (with that device current)
std::mutex g_scratchMu;
std::vector<void *> g_scratch;

void resolveNoAllocScratch(int device)
{
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if ((int)g_scratch.size() > device && g_scratch[device])
    return;
  void * p = nullptr;
  PXG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  PXG_REQUIRE(p, PXG_E_HIP, "no device address for the scratch vector");
  if ((int)g_scratch.size() <= device)
    g_scratch.resize(device + 1, nullptr);
  g_scratch[device] = p;
}

void * noAllocScratch(int device)
{
  std::lock_guard<std::mutex> lk(g_scratchMu);
  return device >= 0 && device < (int)g_scratch.size() ? g_scratch[device] : nullptr;
}

template <typename F>
int guarded(pxg_ctx * c, F && f)
{
  if (!c)
    return PXG_E_INVALID;
  try
  {
    PXG_HIP(hipSetDevice(c->device));
    f();
    c->err.clear();
    return PXG_OK;
  }
  catch (Error const & e)
  {
    c->err = e.what();
    return e.code;
  }
  catch (std::bad_alloc const &)
  {
    c->err = "host allocation failed";
    return PXG_E_OOM;
  }
  catch (std::exception const & e)
  {
    c->err = e.what();
    return PXG_E_INVALID;
  }
}

// Taylor-Hood pairs: velocity FE, pressure FE, velocity components
struct ThPair
{
  int vel, pre, dim;
};
bool thPair(int feType, ThPair & p)
{
  switch (feType)
  {
  case PXG_FE_QUAD_Q2Q1: p = {PXG_FE_QUAD_Q2, PXG_FE_QUAD_Q1, 2}; return true;
  case PXG_FE_HEX_Q2Q1: p = {PXG_FE_HEX_Q2, PXG_FE_HEX_Q1, 3}; return true;
  case PXG_FE_TET_P2P1: p = {PXG_FE_TET_P2, PXG_FE_TET_P1, 3}; return true;
  default: return false;
  }
}

// local basis b of the FE of a context: the FE itself, or for a Taylor-Hood
// pair the velocity basis b % NV of every component, then the pressure basis
struct LocalBasis
{
  FeInfo const * f;
  int i;
};
LocalBasis localBasis(int feType, int b)
{
  ThPair p;
  if (!thPair(feType, p))
    return {&feInfo(feType), b};
  int const nv = feInfo(p.vel).n;
  if (b < p.dim * nv)
    return {&feInfo(p.vel), b % nv};
  return {&feInfo(p.pre), b - p.dim * nv};
}

void buildTables(pxg_ctx * c, int feType, int nq, double const * nodes, double const * weights)
{
  ThPair pair{0, 0, 0};
  bool const mixed = thPair(feType, pair);
  FeInfo fe = feInfo(mixed ? pair.vel : feType);
  if (mixed)
  {
    fe.type = feType;
    fe.n = pair.dim * fe.n + feInfo(pair.pre).n;
  }
  FeTables & T = c->fe;
  Quadrature q;
  if (nq <= 0) // a pair uses the velocity's rule for both spaces, as the reference tests do
    q = recommendedQuadrature(mixed ? pair.vel : feType);
  else
  {
    PXG_REQUIRE(nodes && weights, PXG_E_INVALID, "quadrature nodes and weights required");
    q.dim = fe.dim;
    q.nq = nq;
    q.xi.assign(nodes, nodes + (size_t)nq * fe.dim);
    q.w.assign(weights, weights + nq);
  }
  int const npe = geoInfo(fe.geo).npts;
  T.feType = feType;
  T.geo = fe.geo;
  T.dim = fe.dim;
  T.n = fe.n;
  T.npe = npe;
  T.nq = q.nq;
  T.xi = q.xi;
  T.w = q.w;
  T.phi.assign((size_t)q.nq * fe.n, 0.);
  T.grad.assign((size_t)q.nq * fe.n * fe.dim, 0.);
  T.geoGrad.assign((size_t)q.nq * npe * fe.dim, 0.);
  for (int iq = 0; iq < q.nq; ++iq)
  {
    double const * xi = &q.xi[(size_t)iq * fe.dim];
    for (int i = 0; i < fe.n; ++i)
    {
      LocalBasis const lb = localBasis(feType, i);
      T.phi[iq * fe.n + i] = lb.f->phi(lb.i, xi);
      lb.f->grad(lb.i, xi, &T.grad[((size_t)iq * fe.n + i) * fe.dim]);
    }
    for (int v = 0; v < npe; ++v)
      geoGrad(fe.geo, v, xi, &T.geoGrad[((size_t)iq * npe + v) * fe.dim]);
  }
  hipStream_t s = c->stream;
  T.d_phi.upload(T.phi.data(), T.phi.size(), s);
  T.d_grad.upload(T.grad.data(), T.grad.size(), s);
  T.d_geoGrad.upload(T.geoGrad.data(), T.geoGrad.size(), s);
  T.d_xi.upload(T.xi.data(), T.xi.size(), s);
  T.d_w.upload(T.w.data(), T.w.size(), s);
  std::vector<float> p32(T.phi.begin(), T.phi.end()), g32(T.grad.begin(), T.grad.end());
  T.d_phi32.upload(p32.data(), p32.size(), s);
  T.d_grad32.upload(g32.data(), g32.size(), s);
  PXG_HIP(hipStreamSynchronize(s));
}
// Everything derived from the dof map (pattern, element->nnz and stage maps,
// dof adjacency, values, the closed matrix and its packed copy) is dropped
// when the dof map, the mesh or the FE changes: buffers sized for the old
// element or dof counts must not meet kernels launched for the new ones.
void dropPattern(pxg_ctx * c)
{
  PXG_HIP(hipStreamSynchronize(c->stream));
  c->rowptrS.release();
  c->colS.release();
  c->segPtr.release();
  c->contrib.release();
  c->emapNnz.release();
  c->diagPos.release();
  c->diagC.release();
  c->keep.release();
  c->pos.release();
  c->dofSeg.release();
  c->dofContrib.release();
  c->cubeState = 0;
  c->nCube = c->nPartDof = c->nPart = 0;
  c->cubeElem.release();
  c->cubeOut.release();
  c->partDof.release();
  c->partSeg.release();
  c->partIds.release();
  c->partBuf.release();
  c->nChunk = c->nPartDofF = 0;
  c->cubeIdC.release();
  c->cubeCmC.release();
  c->cubeSlotOf.release();
  c->partDofF.release();
  c->partSegF.release();
  c->partIdsF.release();
  c->stageR0.release();
  c->stageBlk.release();
  c->stageInfo.release();
  c->stageSid.release();
  c->pairPos.release();
  c->valS.release();
  c->valS32.release();
  c->b.release();
  c->rowptr.release();
  c->col.release();
  c->val.release();
  c->val32.release();
  c->dinv.release();
  c->dinv32.release();
  c->blockRow.release();
  c->packBase.release();
  c->packLo.release();
  c->packHi.release();
  c->packVal.release();
  c->packVal32.release();
  c->keBuf.release();
  c->keBuf32.release();
  c->feBuf.release();
  c->geoBuf.release();
  c->mfBuf.release();
  c->nnzS = 0;
  c->nnz = 0;
  c->nStage = 0;
  c->nStreamBlocks = 0;
  c->closed = false;
  c->valsDirty = false;
  c->valsZero = false;
  c->keepFresh = false;
  c->val32Valid = false;
  c->val64Valid = false;
  c->layoutValid = false;
  c->valsSplit = false;
  c->stageValid = false;
  c->stagedFused = false;
  c->packOk = false;
  c->packStamp = c->packValStamp = c->packVal32Stamp = 0;
  c->blockStamp = 0;
  ++c->patStamp;
  ++c->stamp;
}

void dropDofs(pxg_ctx * c)
{
  dropPattern(c);
  c->ndofs = 0;
  c->elemMap.release();
  c->bcFlag.release();
  c->bcVal.release();
  c->kappa.release();
  c->hasBC = false;
  c->hasKappa = false;
  ++c->kappaStamp;
}
} // namespace

extern "C" {

int pxg_create(pxg_ctx ** out, int device)
{
  if (!out)
    return PXG_E_INVALID;
  *out = nullptr;
  pxg_ctx * c = nullptr;
  try
  {
    c = new pxg_ctx;
  }
  catch (...)
  {
    return PXG_E_OOM;
  }
  c->device = device;
  int rc = guarded(c,
                   [&]
                   {
                     PXG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
                     c->ownStream = true;
                     c->status.resize(1);
                     resolveNoAllocScratch(device);
                   });
  if (rc != PXG_OK)
  {
    g_lastError = c->err;
    delete c;
    return rc;
  }
  *out = c;
  return PXG_OK;
}

int pxg_destroy(pxg_ctx * c)
{
  if (!c)
    return PXG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream)
    (void)hipStreamSynchronize(c->stream);
  hipStream_t s = c->stream;
  bool const own = c->ownStream;
  delete c; // buffers free themselves
  if (own && s)
    (void)hipStreamDestroy(s);
  return PXG_OK;
}

const char * pxg_last_error(const pxg_ctx * c) { return c ? c->err.c_str() : g_lastError.c_str(); }

int pxg_set_stream(pxg_ctx * c, void * stream)
{
  return guarded(c,
                 [&]
                 {
                   PXG_HIP(hipStreamSynchronize(c->stream));
                   if (c->ownStream)
                     PXG_HIP(hipStreamDestroy(c->stream));
                   c->stream = (hipStream_t)stream; // NULL: the default stream
                   c->ownStream = false;
                 });
}

int pxg_synchronize(pxg_ctx * c)
{
  return guarded(c, [&] { PXG_HIP(hipStreamSynchronize(c->stream)); });
}

int pxg_get_sizes(pxg_ctx * c, int64_t * ndofs, int64_t * nelem, int * nlocal, int * device)
{
  return guarded(c,
                 [&]
                 {
                   if (ndofs)
                     *ndofs = c->ndofs;
                   if (nelem)
                     *nelem = c->nelem;
                   if (nlocal)
                     *nlocal = c->fe.n;
                   if (device)
                     *device = c->device;
                 });
}

int pxg_profile(pxg_ctx * c, int enable)
{
  return guarded(c,
                 [&]
                 {
                   PXG_HIP(hipStreamSynchronize(c->stream));
                   profFlush(c);
                   c->prof = enable != 0;
                   c->profAcc.clear();
                 });
}

int pxg_profile_get(pxg_ctx * c, const char * name, double * ms, int64_t * count)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(name, PXG_E_INVALID, "null name");
                   PXG_HIP(hipStreamSynchronize(c->stream));
                   profFlush(c);
                   auto it = c->profAcc.find(name);
                   if (ms)
                     *ms = it == c->profAcc.end() ? 0. : it->second.first;
                   if (count)
                     *count = it == c->profAcc.end() ? 0 : it->second.second;
                 });
}

int pxg_set_fe(pxg_ctx * c, int feType, int nq, const double * nodes, const double * weights)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(feType >= PXG_FE_LINE_P1 && feType <= PXG_FE_TET_P2P1, PXG_E_INVALID,
                               "unknown FE type");
                   PXG_REQUIRE(nq >= 0 && nq <= 64, PXG_E_INVALID, "quadrature size must be 0..64");
                   int const oldN = c->fe.n, oldNpe = c->fe.npe, oldType = c->fe.feType;
                   buildTables(c, feType, nq, nodes, weights);
                   // another element shape invalidates the mesh, other local
                   // dofs the dof map; another quadrature rule keeps both
                   if (oldType && c->fe.npe != oldNpe)
                   {
                     dropDofs(c);
                     c->npts = c->nelem = 0;
                     c->npe = 0;
                     c->xyz.release();
                     c->conn.release();
                   }
                   else if (oldType && (c->fe.n != oldN || c->fe.feType != oldType))
                     dropDofs(c);
                 });
}

int pxg_set_mesh(pxg_ctx * c, int64_t npts, const double * coords, int64_t nelem, int npe, const uint32_t * conn)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->fe.feType, PXG_E_STATE, "set_fe before set_mesh");
                   PXG_REQUIRE(npe == c->fe.npe, PXG_E_INVALID, "points per element do not match the FE");
                   PXG_REQUIRE(npts > 0 && nelem > 0 && coords && conn, PXG_E_INVALID, "empty mesh");
                   PXG_REQUIRE(npts < (int64_t)0xffffffffLL, PXG_E_INVALID, "too many points");
                   for (int64_t k = 0; k < nelem * npe; ++k)
                     PXG_REQUIRE(conn[k] < (uint64_t)npts, PXG_E_INVALID, "connectivity out of range");
                   if (c->ndofs && nelem != c->nelem)
                     dropDofs(c); // the dof map has one row per element of the old mesh
                   c->npts = npts;
                   c->nelem = nelem;
                   c->npe = npe;
                   c->xyz.upload(coords, 3 * npts, c->stream);
                   c->conn.upload(conn, nelem * npe, c->stream);
                   PXG_HIP(hipStreamSynchronize(c->stream));
                 });
}

int pxg_set_dofmap(pxg_ctx * c, int64_t ndofs, const uint32_t * emap)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->nelem > 0, PXG_E_STATE, "set_mesh before set_dofmap");
                   PXG_REQUIRE(ndofs > 0 && emap, PXG_E_INVALID, "empty dof map");
                   int64_t const m = c->nelem * c->fe.n;
                   for (int64_t k = 0; k < m; ++k)
                     PXG_REQUIRE(emap[k] < (uint64_t)ndofs, PXG_E_INVALID, "dof id out of range");
                   dropDofs(c);
                   c->ndofs = ndofs;
                   c->elemMap.upload(emap, m, c->stream);
                   c->bcFlag.resize(ndofs);
                   c->bcVal.resize(ndofs);
                   PXG_HIP(hipMemsetAsync(c->bcFlag.p, 0, ndofs, c->stream));
                   PXG_HIP(hipMemsetAsync(c->bcVal.p, 0, ndofs * sizeof(double), c->stream));
                   PXG_HIP(hipStreamSynchronize(c->stream));
                 });
}

int pxg_set_dirichlet(pxg_ctx * c, int64_t nc, const uint32_t * dofs, const double * values)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->ndofs > 0, PXG_E_STATE, "set_dofmap before set_dirichlet");
                   PXG_REQUIRE(nc >= 0 && (nc == 0 || dofs), PXG_E_INVALID, "constrained dofs");
                   std::vector<uint8_t> flag(c->ndofs, 0);
                   std::vector<double> val(c->ndofs, 0.);
                   for (int64_t k = 0; k < nc; ++k)
                   {
                     PXG_REQUIRE(dofs[k] < (uint64_t)c->ndofs, PXG_E_INVALID, "constrained dof out of range");
                     flag[dofs[k]] = 1;
                     val[dofs[k]] = values ? values[k] : 0.;
                   }
                   c->bcFlag.upload(flag.data(), flag.size(), c->stream);
                   c->bcVal.upload(val.data(), val.size(), c->stream);
                   c->hasBC = nc > 0;
                   PXG_HIP(hipStreamSynchronize(c->stream));
                 });
}

int pxg_set_coefficient(pxg_ctx * c, const double * kappa)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->ndofs > 0, PXG_E_STATE, "set_dofmap before set_coefficient");
                   PXG_REQUIRE(kappa, PXG_E_INVALID, "null coefficient");
                   c->kappa.upload(kappa, c->ndofs, c->stream);
                   c->hasKappa = true;
                   ++c->kappaStamp;
                   PXG_HIP(hipStreamSynchronize(c->stream));
                 });
}

int pxg_set_coefficient_device(pxg_ctx * c, const double * kappa)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->ndofs > 0, PXG_E_STATE, "set_dofmap before set_coefficient");
                   PXG_REQUIRE(kappa, PXG_E_INVALID, "null coefficient");
                   c->kappa.resize(c->ndofs);
                   PXG_HIP(hipMemcpyAsync(c->kappa.p, kappa, c->ndofs * sizeof(double), hipMemcpyDeviceToDevice,
                                          c->stream));
                   c->hasKappa = true;
                   ++c->kappaStamp;
                 });
}

int pxg_build_pattern(pxg_ctx * c, int64_t * nnz)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->ndofs > 0, PXG_E_STATE, "set_dofmap before build_pattern");
                   buildPattern(c);
                   if (nnz)
                     *nnz = c->nnzS;
                 });
}

int pxg_assemble(pxg_ctx * c, int nterms, const pxg_term * terms, int precision)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(nterms > 0 && terms, PXG_E_INVALID, "at least one term");
                   assembleLhs(c, nterms, terms, precision);
                 });
}

int pxg_assemble_rhs(pxg_ctx * c, int kind, const double * params, const double * qvalues)
{
  return guarded(c, [&] { assembleRhs(c, kind, params, qvalues); });
}

int pxg_get_qpoints(pxg_ctx * c, double * out)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(out, PXG_E_INVALID, "null output");
                   qpoints(c, out);
                 });
}

int pxg_close(pxg_ctx * c, double eps, int64_t * nnz)
{
  return guarded(c,
                 [&]
                 {
                   closeMatrix(c, eps);
                   if (nnz)
                     *nnz = c->nnz;
                 });
}

int pxg_clear(pxg_ctx * c)
{
  return guarded(c, [&] { clearValues(c); });
}

int pxg_get_csr(pxg_ctx * c, int64_t * rowptr, int32_t * col, double * val, float * val32, double * b)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->closed || (!rowptr && !col && !val && !val32), PXG_E_STATE,
                               "close before get_csr");
                   hipStream_t s = c->stream;
                   if (rowptr)
                     c->rowptr.download(rowptr, c->ndofs + 1, s);
                   if (col)
                     c->col.download(col, c->nnz, s);
                   if (val)
                   {
                     ensureVal64(c);
                     c->val.download(val, c->nnz, s);
                   }
                   if (val32)
                   {
                     ensureVal32(c);
                     c->val32.download(val32, c->nnz, s);
                   }
                   if (b)
                     c->b.download(b, c->ndofs, s);
                   PXG_HIP(hipStreamSynchronize(s));
                 });
}

int pxg_get_pattern(pxg_ctx * c, int64_t * rowptr, int32_t * col, uint32_t * emap)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->nnzS > 0, PXG_E_STATE, "build_pattern first");
                   hipStream_t s = c->stream;
                   if (rowptr)
                     c->rowptrS.download(rowptr, c->ndofs + 1, s);
                   if (col)
                     c->colS.download(col, c->nnzS, s);
                   if (emap)
                     c->emapNnz.download(emap, c->nelem * c->fe.n * c->fe.n, s);
                   PXG_HIP(hipStreamSynchronize(s));
                 });
}

int pxg_get_geometry(pxg_ctx * c, double * detJ, double * JxW)
{
  return guarded(c, [&] { geometryOut(c, detJ, JxW); });
}

int pxg_spmv(pxg_ctx * c, int precision, const void * x, void * y)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->closed, PXG_E_STATE, "close before spmv");
                   PXG_REQUIRE(x && y && x != y, PXG_E_INVALID, "x and y must be distinct device buffers");
                   ensureOperator(c, precision == PXG_FP32 ? PXG_FP32 : PXG_FP64);
                   CsrView const A = viewOf(c);
                   if (precision == PXG_FP32)
                     spmv<float>(A, (float const *)x, (float *)y, c->stream);
                   else
                     spmv<double>(A, (double const *)x, (double *)y, c->stream);
                 });
}

int pxg_jacobi(pxg_ctx * c, int precision, double omega, int sweeps, const void * b, void * x)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->closed, PXG_E_STATE, "close before jacobi");
                   PXG_REQUIRE(b && x && sweeps >= 0, PXG_E_INVALID, "jacobi arguments");
                   ensureOperator(c, precision == PXG_FP32 ? PXG_FP32 : PXG_FP64);
                   CsrView const A = viewOf(c);
                   int64_t const n = c->ndofs;
                   // sweeps alternate between x and a buffer of the context; an
                   // odd count ends with one copy back into x
                   auto run = [&](auto * xs, auto const * bs, auto & tmp)
                   {
                     using T = std::remove_pointer_t<decltype(xs)>;
                     tmp.resize(n);
                     T * cur = xs;
                     T * nxt = tmp.p;
                     for (int k = 0; k < sweeps; ++k)
                     {
                       jacobiStep<T>(A, bs, cur, nxt, omega, c->stream);
                       std::swap(cur, nxt);
                     }
                     if (cur != xs)
                       PXG_HIP(hipMemcpyAsync(xs, cur, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
                   };
                   if (precision == PXG_FP32)
                     run((float *)x, (float const *)b, c->jacTmp32);
                   else
                     run((double *)x, (double const *)b, c->jacTmp);
                   PXG_HIP(hipStreamSynchronize(c->stream));
                 });
}

int pxg_jacobi_noalloc(pxg_ctx * c, int precision, double omega, int sweeps, const void * b, void * x)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->closed, PXG_E_STATE, "close before jacobi");
                   PXG_REQUIRE(b && x && b != x && sweeps >= 0, PXG_E_INVALID, "jacobi arguments");
                   int64_t const n = c->ndofs;
                   PXG_REQUIRE(n <= kNoAllocMaxRows, PXG_E_INVALID,
                               "pxg_jacobi_noalloc takes at most 4194304 rows (its scratch vector)");
                   void * scr = noAllocScratch(c->device);
                   PXG_REQUIRE(scr, PXG_E_STATE, "scratch vector not resolved on this device");
                   // values in the asked precision (a cast into the closed
                   // buffers, no allocation); row blocks and the packed copy
                   // are used only if current, never built here
                   if (precision == PXG_FP32)
                     ensureVal32(c);
                   else
                     ensureVal64(c);
                   CsrView const A = viewOf(c);
                   auto run = [&](auto * xs, auto const * bs)
                   {
                     using T = std::remove_pointer_t<decltype(xs)>;
                     T * cur = xs;
                     T * nxt = static_cast<T *>(scr);
                     for (int k = 0; k < sweeps; ++k)
                     {
                       jacobiStep<T>(A, bs, cur, nxt, omega, c->stream);
                       std::swap(cur, nxt);
                     }
                     if (cur != xs)
                       PXG_HIP(hipMemcpyAsync(xs, cur, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
                   };
                   if (precision == PXG_FP32)
                     run((float *)x, (float const *)b);
                   else
                     run((double *)x, (double const *)b);
                 });
}

int pxg_csr_device(pxg_ctx * c, const int64_t ** rowptr, const int32_t ** col, const double ** val,
                   const float ** val32, int64_t * nnz)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->closed, PXG_E_STATE, "close first");
                   if (rowptr)
                     *rowptr = c->rowptr.p;
                   if (col)
                     *col = c->col.p;
                   if (val)
                   {
                     ensureVal64(c);
                     *val = c->val.p;
                   }
                   if (val32)
                   {
                     ensureVal32(c);
                     *val32 = c->val32.p;
                   }
                   if (nnz)
                     *nnz = c->nnz;
                 });
}

int pxg_matrix_stamp(pxg_ctx * c, uint64_t * stamp)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(stamp, PXG_E_INVALID, "null stamp");
                   *stamp = c->stamp;
                 });
}

int pxg_csr_pack(pxg_ctx * c, int enable, int * active)
{
  return guarded(c,
                 [&]
                 {
                   if (enable >= 0)
                     c->packWanted = enable != 0;
                   if (active)
                     *active = c->closed && c->packWanted && c->packOk && c->packStamp == c->patStamp ? 1 : 0;
                 });
}

int pxg_rhs_device(pxg_ctx * c, double ** b)
{
  return guarded(c,
                 [&]
                 {
                   PXG_REQUIRE(c->b.p && b, PXG_E_STATE, "build_pattern first");
                   *b = c->b.p;
                 });
}

} // extern "C"
