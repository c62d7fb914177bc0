// Oracle (test infrastructure only): sparse kernels, multigrid-preconditioned CG,
// matrix-free operator and its gradients, restated on the CPU.
#pragma once

#include "oracle_fe.hpp"

namespace orc
{

void spmv(Csr const & A, double const * x, double * y);

// one weighted-Jacobi sweep x <- x + omega D^{-1} (b - A x)
void jacobiSweep(Csr const & A, std::vector<double> const & dinv, double omega, double const * b, double * x);

// Prolongation from a coarse to a fine structured mesh of the same FE space:
// P(f, c) = value of coarse basis c at fine dof point f, over the coarse
// element containing the fine element (parent = element index halved per axis).
// Entries with |v| <= 1e-12 are dropped; the first write of a (f, c) pair wins,
// as in reference include/multigrid.hpp:37-70.
Csr prolongationHex(
    Mesh const & fine, DofMap const & df, u32 nxf, u32 nyf, u32 nzf, Mesh const & coarse,
    DofMap const & dc, RefFE const & fe);

// transpose, and the reference restriction (rows of P^T scaled to sum 1,
// include/multigrid.hpp:108-175)
Csr transpose(Csr const & A, i64 ncols);
Csr rowNormalised(Csr const & A);

struct MgLevel
{
  Csr A;
  std::vector<double> dinv;
  std::vector<uint8_t> cons;
  Csr P; // prolongation from the next coarser level (empty on the coarsest)
  Csr R; // P^T
};

struct Multigrid
{
  std::vector<MgLevel> lv; // 0 = finest
  double omega = 2. / 3.;
  int nu1 = 2, nu2 = 2;
  double coarseRtol = 1e-10;
  int coarseMaxit = 10000;
  // coarsest level: dense Cholesky factor (row-major lower triangle) when it
  // has at most kDenseCoarse rows, the direct coarse solve of the reference
  // test (test/multigrid.cpp: SparseLU); larger: Jacobi-CG to coarseRtol
  static constexpr i64 kDenseCoarse = 2048;
  std::vector<double> coarseL;
  void factorCoarse();
  void vcycle(int l, double const * b, double * x) const;
  // preconditioned CG; returns iterations, final relative residual in relres
  int pcg(double const * b, double * x, double rtol, int maxit, double & relres) const;
};

// matrix-free y = K(kappa) x for nb interleaved vectors (layout [ndofs][nb]);
// K is the assembled kappa-stiffness; when `cons` is given the Dirichlet form
// C K C + D_c (D_c = element multiplicity on constrained dofs) is applied.
void mfApply(
    Mesh const & m, CurFE & cfe, DofMap const & d, double const * kappa,
    uint8_t const * cons, int nb, double const * x, double * y);

// d/dkappa_n of sum_b ybar_b . K(kappa) x_b  (geometry recomputed per element)
void mfGradKappa(
    Mesh const & m, CurFE & cfe, DofMap const & d, uint8_t const * cons, int nb,
    double const * x, double const * ybar, double * gk);

} // namespace orc
