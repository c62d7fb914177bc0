// LibTorch autograd operators over the HIP library (C ABI include/proxpde_amd.h).
//
//   proxpde::kappa_apply(kappa, x, ctx, stream, use_bc) -> y = K(kappa) x
//       forward : matrix-free kappa-stiffness operator (AssemblyStiffnessFE
//                 bilinear form, include/assembly_lhs.hpp:93-109), x [N] or [N, nb]
//       backward: dx = K^T dy = K dy;  dkappa_n = sum_e sum_q phi_n JxW grad(dy).grad(x)
//                 (geometry recomputed, nothing stored but kappa and x)
//
//   proxpde::kappa_solve(kappa, b, ctxs, inject, mg, stream, rtol, mode, maxit, fp32_levels, fine_fp64)
//       forward : re-assembles K(kappa) on every multigrid level (kappa injected to
//                 the coarse dofs through `inject`), closes, and solves K u = b with
//                 MG-PCG (mode 0: FP64; mode 1: FP32 CG refined by FP64 residuals;
//                 mode 2: as 1 with all levels but the coarsest applied
//                 matrix-free, not assembled).
//                 fp32_levels: element matrices of all but the coarsest level in FP32;
//                 the refinement residual then uses the FP64 matrix-free operator.
//                 fine_fp64: the finest level is still assembled in FP64 (its FP32
//                 copy is rounded from it) and the residual uses that matrix.
//       backward: adjoint solve K lambda = du;  db = lambda;
//                 dkappa = -d/dkappa (lambda . K(kappa) u)  (Dirichlet rows masked)
//                 The level matrices are re-assembled only if some level was
//                 closed or cleared since the forward pass (pxg_matrix_stamp).
//
// Handles (pxg_ctx*, pxg_mg*) and the HIP stream are passed as integers.
#include <torch/library.h>
#include <torch/torch.h>

#include "../../include/proxpde_amd.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace
{
using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void check(int rc, pxg_ctx * c, char const * what)
{
  if (rc != PXG_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + pxg_last_error(c));
}

void checkMg(int rc, pxg_mg * m, char const * what)
{
  if (rc != PXG_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + pxg_mg_last_error(m));
}

// sizes of a context (dofs, device)
struct Sizes
{
  int64_t ndofs = 0;
  int device = 0;
};

Sizes sizesOf(pxg_ctx * c)
{
  Sizes z;
  check(pxg_get_sizes(c, &z.ndofs, nullptr, nullptr, &z.device), c, "pxg_get_sizes");
  return z;
}

// a float64 GPU tensor: [n] (vectors) or [n] / [n, nb] with 1 <= nb <= 16
// (right-hand sides); checked before the context is touched
void checkKind(Tensor const & t, char const * name, bool batched)
{
  TORCH_CHECK(t.defined(), name, " is undefined");
  TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float64");
  TORCH_CHECK(t.is_cuda(), name, " must be a GPU tensor");
  TORCH_CHECK(batched ? (t.dim() == 1 || t.dim() == 2) : t.dim() == 1, name, " has the wrong number of dimensions");
  if (t.dim() == 2)
    TORCH_CHECK(t.size(1) >= 1 && t.size(1) <= 16, name, ": 1 to 16 right-hand sides");
}

// ... on the context's device with one row per dof, so no kernel reads past
// the end of the buffer or dereferences memory of another device
void checkDofTensor(Tensor const & t, Sizes const & z, char const * name, bool batched)
{
  checkKind(t, name, batched);
  TORCH_CHECK(t.get_device() == z.device, name, " is on device ", t.get_device(), ", the context on ", z.device);
  TORCH_CHECK(t.size(0) == z.ndofs, name, " has ", t.size(0), " rows, the context ", z.ndofs, " dofs");
}

Tensor applyK(pxg_ctx * c, void * stream, Tensor const & kappa, Tensor const & x, bool useBC)
{
  checkKind(kappa, "kappa", false);
  checkKind(x, "x", true);
  Sizes const z = sizesOf(c);
  checkDofTensor(kappa, z, "kappa", false);
  checkDofTensor(x, z, "x", true);
  TORCH_CHECK(kappa.is_contiguous(), "kappa must be contiguous");
  auto xc = x.contiguous();
  int const nb = xc.dim() == 1 ? 1 : (int)xc.size(1);
  auto y = torch::empty_like(xc);
  check(pxg_set_stream(c, stream), c, "pxg_set_stream");
  check(pxg_mf_apply(c, kappa.data_ptr<double>(), nb, xc.data_ptr<double>(), y.data_ptr<double>(), useBC ? 1 : 0),
        c, "pxg_mf_apply");
  return y;
}

Tensor gradKappa(pxg_ctx * c, void * stream, Tensor const & x, Tensor const & ybar, bool useBC)
{
  checkKind(x, "x", true);
  checkKind(ybar, "ybar", true);
  Sizes const z = sizesOf(c);
  checkDofTensor(x, z, "x", true);
  checkDofTensor(ybar, z, "ybar", true);
  TORCH_CHECK(x.sizes() == ybar.sizes(), "x and the output gradient must have the same shape");
  auto xc = x.contiguous();
  auto yc = ybar.contiguous();
  int const nb = xc.dim() == 1 ? 1 : (int)xc.size(1);
  auto g = torch::empty({xc.size(0)}, xc.options());
  check(pxg_set_stream(c, stream), c, "pxg_set_stream");
  check(pxg_mf_grad_kappa(c, nb, xc.data_ptr<double>(), yc.data_ptr<double>(), g.data_ptr<double>(), useBC ? 1 : 0),
        c, "pxg_mf_grad_kappa");
  return g;
}

class KappaApply: public torch::autograd::Function<KappaApply>
{
public:
  static Tensor forward(AutogradContext * ctx, Tensor kappa, Tensor x, int64_t handle, int64_t stream, bool useBC)
  {
    TORCH_CHECK(kappa.scalar_type() == torch::kFloat64 && x.scalar_type() == torch::kFloat64, "float64 tensors");
    TORCH_CHECK(kappa.is_contiguous() && kappa.dim() == 1, "kappa must be a contiguous vector");
    auto * c = reinterpret_cast<pxg_ctx *>(handle);
    ctx->saved_data["handle"] = handle;
    ctx->saved_data["stream"] = stream;
    ctx->saved_data["use_bc"] = useBC;
    ctx->save_for_backward({kappa, x});
    return applyK(c, reinterpret_cast<void *>(stream), kappa, x, useBC);
  }

  static tensor_list backward(AutogradContext * ctx, tensor_list grads)
  {
    auto saved = ctx->get_saved_variables();
    auto kappa = saved[0], x = saved[1];
    auto * c = reinterpret_cast<pxg_ctx *>(ctx->saved_data["handle"].toInt());
    void * stream = reinterpret_cast<void *>(ctx->saved_data["stream"].toInt());
    bool const useBC = ctx->saved_data["use_bc"].toBool();
    auto gy = grads[0].contiguous();
    Tensor gk, gx;
    if (ctx->needs_input_grad(1))
      gx = applyK(c, stream, kappa, gy, useBC); // K is symmetric
    if (ctx->needs_input_grad(0))
      gk = gradKappa(c, stream, x, gy, useBC);
    return {gk, gx, Tensor(), Tensor(), Tensor()};
  }
};

struct SolveSetup
{
  std::vector<pxg_ctx *> levels;
  std::vector<Tensor> inject;
  pxg_mg * mg;
  void * stream;
  double rtol;
  int64_t mode, maxit;
  bool fp32Levels, fineFp64;
};

std::vector<int64_t> stampsOf(SolveSetup const & S)
{
  std::vector<int64_t> out;
  for (auto * c: S.levels)
  {
    uint64_t st = 0;
    check(pxg_matrix_stamp(c, &st), c, "pxg_matrix_stamp");
    out.push_back((int64_t)st);
  }
  return out;
}

// assemble K(kappa) on all levels (coarse kappa by injection), close them
void assembleLevels(SolveSetup const & S, Tensor const & kappa)
{
  pxg_term const term{PXG_TERM_STIFFNESS_FE, 1.0};
  size_t const nl = S.levels.size();
  for (size_t l = 0; l < nl; ++l)
  {
    pxg_ctx * c = S.levels[l];
    Tensor kl = l == 0 ? kappa : kappa.index_select(0, S.inject[l - 1]).contiguous();
    check(pxg_set_stream(c, S.stream), c, "pxg_set_stream");
    check(pxg_set_coefficient_device(c, kl.data_ptr<double>()), c, "pxg_set_coefficient_device");
    if (l + 1 < nl && S.mode == PXG_FP32_MF)
      continue; // applied matrix-free from kappa (only the coarsest is assembled)
    check(pxg_clear(c), c, "pxg_clear");
    // the coarsest level stays FP64 (FP64 coarse solve)
    int const prec = (S.fp32Levels && l + 1 < nl && !(l == 0 && S.fineFp64)) ? PXG_FP32 : PXG_FP64;
    check(pxg_assemble(c, 1, &term, prec), c, "pxg_assemble");
    int64_t nnz = 0;
    check(pxg_close(c, 1e-16, &nnz), c, "pxg_close");
  }
  checkMg(pxg_mg_update(S.mg), S.mg, "pxg_mg_update");
}

Tensor solveWith(SolveSetup const & S, Tensor const & rhs)
{
  auto b = rhs.contiguous();
  auto u = torch::zeros_like(b);
  int iters = 0;
  double rr = 0.;
  checkMg(pxg_mg_solve(S.mg, (int)S.mode, b.data_ptr<double>(), u.data_ptr<double>(), S.rtol, (int)S.maxit, &iters,
                       &rr),
          S.mg, "pxg_mg_solve");
  return u;
}

SolveSetup setupFrom(AutogradContext * ctx)
{
  SolveSetup S;
  for (auto v: ctx->saved_data["levels"].toIntVector())
    S.levels.push_back(reinterpret_cast<pxg_ctx *>(v));
  S.inject = ctx->saved_data["inject"].toTensorVector();
  S.mg = reinterpret_cast<pxg_mg *>(ctx->saved_data["mg"].toInt());
  S.stream = reinterpret_cast<void *>(ctx->saved_data["stream"].toInt());
  S.rtol = ctx->saved_data["rtol"].toDouble();
  S.mode = ctx->saved_data["mode"].toInt();
  S.maxit = ctx->saved_data["maxit"].toInt();
  S.fp32Levels = ctx->saved_data["fp32"].toBool();
  S.fineFp64 = ctx->saved_data["fine64"].toBool();
  return S;
}

class KappaSolve: public torch::autograd::Function<KappaSolve>
{
public:
  static Tensor forward(AutogradContext * ctx, Tensor kappa, Tensor b, std::vector<int64_t> levels,
                        std::vector<Tensor> inject, int64_t mg, int64_t stream, double rtol, int64_t mode,
                        int64_t maxit, bool fp32Levels, bool fineFp64)
  {
    TORCH_CHECK(kappa.scalar_type() == torch::kFloat64 && b.scalar_type() == torch::kFloat64, "float64 tensors");
    TORCH_CHECK(!levels.empty() && inject.size() + 1 == levels.size(), "one injection map per coarse level");
    checkKind(kappa, "kappa", false);
    checkKind(b, "b", false);
    SolveSetup S;
    for (auto v: levels)
      S.levels.push_back(reinterpret_cast<pxg_ctx *>(v));
    Sizes const z = sizesOf(S.levels[0]);
    checkDofTensor(kappa, z, "kappa", false);
    checkDofTensor(b, z, "b", false);
    for (size_t l = 0; l < inject.size(); ++l)
    {
      Sizes const zl = sizesOf(S.levels[l + 1]);
      auto const & m = inject[l];
      TORCH_CHECK(m.scalar_type() == torch::kInt64 && m.dim() == 1 && m.is_cuda() && m.get_device() == z.device,
                  "injection maps: int64 vectors on the context's device");
      TORCH_CHECK(m.size(0) == zl.ndofs, "injection map ", l, " has ", m.size(0), " entries, level ", l + 1, " ",
                  zl.ndofs, " dofs");
    }
    S.inject = inject;
    S.mg = reinterpret_cast<pxg_mg *>(mg);
    S.stream = reinterpret_cast<void *>(stream);
    S.rtol = rtol;
    S.mode = mode;
    S.maxit = maxit;
    S.fp32Levels = fp32Levels;
    S.fineFp64 = fineFp64;
    auto kc = kappa.contiguous();
    assembleLevels(S, kc);
    Tensor u = solveWith(S, b);
    ctx->saved_data["levels"] = levels;
    ctx->saved_data["inject"] = inject;
    ctx->saved_data["mg"] = mg;
    ctx->saved_data["stream"] = stream;
    ctx->saved_data["rtol"] = rtol;
    ctx->saved_data["mode"] = mode;
    ctx->saved_data["maxit"] = maxit;
    ctx->saved_data["fp32"] = fp32Levels;
    ctx->saved_data["fine64"] = fineFp64;
    ctx->saved_data["stamps"] = stampsOf(S);
    ctx->save_for_backward({kc, u});
    return u;
  }

  static tensor_list backward(AutogradContext * ctx, tensor_list grads)
  {
    auto saved = ctx->get_saved_variables();
    auto kappa = saved[0], u = saved[1];
    SolveSetup S = setupFrom(ctx);
    // the level matrices still hold K(kappa) unless a level was closed or
    // cleared since the forward pass (another solve, a right-hand side)
    if (stampsOf(S) != ctx->saved_data["stamps"].toIntVector())
      assembleLevels(S, kappa);
    checkDofTensor(grads[0], sizesOf(S.levels[0]), "grad_u", false);
    Tensor lambda = solveWith(S, grads[0]);
    Tensor gk;
    if (ctx->needs_input_grad(0))
      gk = -gradKappa(S.levels[0], S.stream, u, lambda, true);
    return {gk, lambda, Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
  }
};

Tensor kappaApply(Tensor kappa, Tensor x, int64_t handle, int64_t stream, bool useBC)
{
  return KappaApply::apply(kappa, x, handle, stream, useBC);
}

Tensor kappaSolve(Tensor kappa, Tensor b, std::vector<int64_t> levels, std::vector<Tensor> inject, int64_t mg,
                  int64_t stream, double rtol, int64_t mode, int64_t maxit, bool fp32Levels, bool fineFp64)
{
  return KappaSolve::apply(kappa, b, levels, inject, mg, stream, rtol, mode, maxit, fp32Levels, fineFp64);
}
} // namespace

TORCH_LIBRARY(proxpde, m)
{
  m.def("kappa_apply(Tensor kappa, Tensor x, int ctx, int stream, bool use_bc) -> Tensor");
  m.def("kappa_solve(Tensor kappa, Tensor b, int[] levels, Tensor[] inject, int mg, int stream, float rtol, "
        "int mode, int maxit, bool fp32_levels, bool fine_fp64) -> Tensor");
}

TORCH_LIBRARY_IMPL(proxpde, CompositeImplicitAutograd, m)
{
  m.impl("kappa_apply", &kappaApply);
  m.impl("kappa_solve", &kappaSolve);
}
