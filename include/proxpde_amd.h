/*
 * proxpde_amd: MI355X (gfx950) element assembly and operator-apply library.
 *
 * C ABI.  Plain pointers and sizes; no exceptions cross it; every call returns
 * PXG_OK (0) or a negative status, and pxg_last_error() explains the last one.
 * A context owns its device memory and one HIP stream; it is externally
 * synchronised (one host thread at a time per context).
 *
 * The assembly entry points replace the CPU path of the reference's
 * proxpde::Builder (include/builder.hpp:12-391):
 *   Builder(uint size) / init(size)          -> pxg_create + pxg_set_dofmap
 *   FESpace{mesh} / CurFE<RefFE, QR>         -> pxg_set_fe + pxg_set_mesh
 *   BCEss{feSpace, side, fun}                -> pxg_set_dirichlet
 *   buildLhs(tuple{AssemblyStiffness, ...})  -> pxg_assemble
 *   buildRhs(tuple{AssemblyRhsAnalytic})     -> pxg_assemble_rhs
 *   closeMatrix()                            -> pxg_close
 *   A (row-major Eigen::SparseMatrix), b     -> pxg_get_csr
 * Host-side mesh and dof-map construction (the mirror of the reference's
 * buildHyperCube / DOF::setupElemMap / BCEss::setupDofMap) is declared in
 * proxpde_amd_host.h.
 */
#ifndef PROXPDE_AMD_H
#define PROXPDE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PXG_OK 0
#define PXG_E_INVALID (-1) /* bad argument or shape                         */
#define PXG_E_OOM (-2)     /* device or host allocation failed              */
#define PXG_E_HIP (-3)     /* HIP runtime error                             */
#define PXG_E_NUMERIC (-4) /* non-positive Jacobian, CG breakdown, NaN      */
#define PXG_E_STATE (-5)   /* call out of order (e.g. assemble before mesh) */
#define PXG_E_NOCONV (-6)  /* iterative solve hit its iteration limit       */

/* ---- finite elements (reference include/fe.hpp:30-132) ----------------- */
#define PXG_FE_LINE_P1 1 /* RefLineP1,        GaussQR<Line,2>          */
#define PXG_FE_QUAD_Q1 2 /* RefQuadQ1,        GaussQR<Quad,4>          */
#define PXG_FE_QUAD_Q2 3 /* RefQuadQ2,        GaussQR<Quad,9>          */
#define PXG_FE_HEX_Q1 4  /* RefHexahedronQ1,  GaussQR<Hexahedron,8>    */
#define PXG_FE_HEX_Q2 5  /* RefHexahedronQ2,  GaussQR<Hexahedron,27>   */
#define PXG_FE_TET_P1 6  /* RefTetrahedronP1, GaussQR<Tetrahedron,4>   */
#define PXG_FE_TET_P2 7  /* RefTetrahedronP2, GaussQR<Tetrahedron,11>  */
/* Taylor-Hood pairs as one element: the velocity FESpace<Mesh, Q2 / P2,
 * RecommendedQR, dim> (local i + NV d for component d, include/fespace.hpp)
 * followed by the pressure FESpace<Mesh, Q1 / P1, same rule> (local
 * dim NV + k), as used together by test/assemblyvector.cpp,
 * test/stokes2dquad.cpp, test/stokes3dhex.cpp and test/stokes3dtet.cpp.
 * The dof map row of an element holds the dim NV velocity ids (BLOCK:
 * id + d nU, INTERLEAVED: dim id + d, include/dof.hpp:139-166) and the
 * pressure ids (after the velocity, the FESpace offset).  Terms: STIFFNESS,
 * TENSOR_STIFFNESS, MASS on the velocity block, GRAD and DIV couplings.
 * A constrained pressure dof acts as buildLhs(AssemblyDummy{p}, bcsP) plus
 * bcsP in both couplings.                                                  */
#define PXG_FE_QUAD_Q2Q1 8  /* 2 x 9 + 4 = 22 local dofs, GaussQR<Quad,9>         */
#define PXG_FE_HEX_Q2Q1 9   /* 3 x 27 + 8 = 89 local dofs, GaussQR<Hexahedron,27> */
#define PXG_FE_TET_P2P1 10  /* 3 x 10 + 4 = 34 local dofs, GaussQR<Tetrahedron,11> */

/* ---- assembly terms (reference include/assembly_lhs.hpp) --------------- */
#define PXG_TERM_STIFFNESS 1    /* AssemblyStiffness(coef)          :36-75   */
#define PXG_TERM_STIFFNESS_FE 2 /* AssemblyStiffnessFE(kappa)       :77-113  */
#define PXG_TERM_MASS 3         /* AssemblyMass / ScalarMass(coef)  :191-231 */
#define PXG_TERM_MASS_FE 4      /* AssemblyMassFE(coef, kappa)      :233-283 */
/* couplings of PXG_FE_QUAD_Q2Q1 (reference include/assembly_coupling.hpp and
 * Builder::buildCoupling, include/builder.hpp:156-277), comp list {0, 1}:   */
#define PXG_TERM_GRAD 5 /* AssemblyGrad(coef, vel, p)  :41-86, rows velocity  */
#define PXG_TERM_DIV 6  /* AssemblyDiv(coef, p, vel)   :88-133, rows pressure */
/* AssemblyTensorStiffness(coef, vel), assembly_lhs.hpp:115-173 (velocity)   */
#define PXG_TERM_TENSOR_STIFFNESS 7

/* ---- analytic right-hand sides (AssemblyRhsAnalytic, assembly_rhs.hpp:34-91)
 * The reference takes a host std::function; the closed forms used by its
 * tests and by the workloads are evaluated on the device, any other function
 * goes through PXG_RHS_QVALUES with values at pxg_get_qpoints() points.    */
#define PXG_RHS_NONE 0
#define PXG_RHS_CONST 1      /* p0                                      */
#define PXG_RHS_SINSIN 2     /* p0 sin(p1 x) sin(p2 y)                  */
#define PXG_RHS_SINSINSIN 3  /* p0 sin(p1 x) sin(p2 y) sin(p3 z)        */
#define PXG_RHS_FEVAR1D 4    /* sin(p1 x) - p0 cos(p2 x)                */
#define PXG_RHS_QVALUES 5    /* host values per (element, qpoint)       */
/* AssemblyProjection(p0, f_h) (assembly_rhs.hpp:94-148, 285-311): qvalues
 * holds the n_dofs nodal values f_h in the same FE space;
 * Fe += p0 JxW_q phi_q (phi_q . f_h,local)                               */
#define PXG_RHS_PROJECTION 6

#define PXG_FP64 0
#define PXG_FP32 1
/* pxg_mg_solve only: PXG_FP32 with the levels above the coarsest applied
 * matrix-free (below)                                                     */
#define PXG_FP32_MF 2

typedef struct pxg_ctx pxg_ctx;

typedef struct pxg_term
{
  int kind;    /* PXG_TERM_*                              */
  double coef; /* scalar coefficient (FE terms: multiplier) */
} pxg_term;

/* context ----------------------------------------------------------------- */
int pxg_create(pxg_ctx ** ctx, int device);
int pxg_destroy(pxg_ctx * ctx);
/* message of the last failure on ctx (ctx NULL: the last failed pxg_create) */
const char * pxg_last_error(const pxg_ctx * ctx);
/* run the context's work on an existing HIP stream (NULL: the default
 * stream); the context creates a non-blocking stream of its own at create  */
int pxg_set_stream(pxg_ctx * ctx, void * hip_stream);
int pxg_synchronize(pxg_ctx * ctx);
/* sizes of the context (any pointer may be NULL): dofs of the dof map (0
 * before pxg_set_dofmap), mesh elements, local dofs per element, HIP device.
 * No counterpart in the reference (a Builder is sized by its FESpace); it is
 * for bindings that check the lengths of the buffers they pass.           */
int pxg_get_sizes(pxg_ctx * ctx, int64_t * n_dofs, int64_t * n_elem, int * n_local, int * device);
/* per-kernel timing with HIP events on the context stream: enable resets the
 * totals; names: "geometry", "element", "element32", "scatter", "scatter32",
 * "rhs_scatter", "close", "mf_apply", "mf_gradk", "spmv", "jacobi"            */
int pxg_profile(pxg_ctx * ctx, int enable);
int pxg_profile_get(pxg_ctx * ctx, const char * name, double * total_ms, int64_t * count);

/* FESpace / CurFE: FE type and quadrature rule.  nq = 0 selects the
 * reference's RecommendedQR (of the velocity FE for a Taylor-Hood pair);
 * otherwise nodes (nq x dim) and weights (nq).
 * Replaces CurFE<RefFE,QR>::CurFE (include/curfe.hpp:20-45).             */
int pxg_set_fe(pxg_ctx * ctx, int fe_type, int nq, const double * qr_nodes, const double * qr_weights);

/* mesh: point coordinates (n_pts x 3) and element-to-vertex connectivity
 * (n_elem x npe, reference vertex order, include/geo.hpp).                 */
int pxg_set_mesh(pxg_ctx * ctx, int64_t n_pts, const double * coords, int64_t n_elem, int npe,
                 const uint32_t * conn);

/* DOF::elemMap (include/dof.hpp:73-76): n_elem x n_local dof ids          */
int pxg_set_dofmap(pxg_ctx * ctx, int64_t n_dofs, const uint32_t * elem_map);

/* BCEss (include/bc.hpp:31-309): constrained dofs and their values.
 * Several BCEss merge by later-wins; diag = 1 as BCEss::diag.             */
int pxg_set_dirichlet(pxg_ctx * ctx, int64_t n_c, const uint32_t * dofs, const double * values);

/* nodal coefficient field kappa (FEVar on the same space) for FE terms    */
int pxg_set_coefficient(pxg_ctx * ctx, const double * kappa_host);
/* same from device memory (copied on the context stream)                  */
int pxg_set_coefficient_device(pxg_ctx * ctx, const double * kappa_dev);

/* CSR sparsity pattern and element->nnz map from the dof map (once per mesh):
 * sorted segmented reduction of the E*n*n (row, col) keys.                */
int pxg_build_pattern(pxg_ctx * ctx, int64_t * nnz_structural);

/* Builder::buildLhs (include/builder.hpp:31-154): element matrices of the
 * given terms, per-element Dirichlet elimination, deterministic scatter into
 * the structural CSR values and the lifting into b.  Accumulates, as
 * successive buildLhs calls accumulate triplets.  precision: PXG_FP64 or
 * PXG_FP32 (element kernel and values in float).  Returns without waiting
 * for the GPU: an element with a non-positive Jacobian determinant is
 * reported (PXG_E_NUMERIC) by the next pxg_close, or by the next
 * pxg_assemble if that comes first.  Scalar FE types take STIFFNESS,
 * STIFFNESS_FE, MASS, MASS_FE; the Taylor-Hood pairs take STIFFNESS,
 * TENSOR_STIFFNESS, MASS (velocity block), GRAD and DIV, which together are
 * buildLhs on the velocity plus both buildCoupling calls
 * (include/builder.hpp:156-277); any other combination is PXG_E_INVALID. */
int pxg_assemble(pxg_ctx * ctx, int n_terms, const pxg_term * terms, int precision);

/* Builder::buildRhs with AssemblyRhsAnalytic or AssemblyProjection
 * (include/builder.hpp:279-364).  A Taylor-Hood context takes the velocity
 * forcing of AssemblyAnalyticRhs on its vector space: PXG_RHS_CONST with
 * params[d] the constant of component d, or PXG_RHS_QVALUES with qvalues
 * n_elem x nq x dim (component fastest) at pxg_get_qpoints() points; the
 * pressure rows get nothing (other kinds: PXG_E_INVALID)                  */
int pxg_assemble_rhs(pxg_ctx * ctx, int rhs_kind, const double * params, const double * qvalues);

/* quadrature points x_q of every element (n_elem x nq x 3), for host
 * evaluation of arbitrary right-hand sides (CurFE::qpoint, curfe.hpp:80)  */
int pxg_get_qpoints(pxg_ctx * ctx, double * out);

/* Builder::closeMatrix (include/builder.hpp:366-371): sum the builds since
 * the last close or clear, drop |v| <= eps.  Like the reference, which clears
 * its triplet list here, close uses the builds up: the next pxg_assemble
 * starts a new matrix (b keeps accumulating until pxg_clear), and a close
 * with no build since the last one gives the empty matrix.                  */
int pxg_close(pxg_ctx * ctx, double prune_eps, int64_t * nnz);

/* Builder::clear (include/builder.hpp:381-385): zero values and b          */
int pxg_clear(pxg_ctx * ctx);

/* closed matrix (row-major CSR, int32 column indices like Eigen's default
 * StorageIndex) and b; host buffers, synchronous.  Any pointer may be NULL. */
int pxg_get_csr(pxg_ctx * ctx, int64_t * rowptr, int32_t * col, double * val, float * val32, double * b);

/* structural pattern and element->nnz map (n_elem x n x n)                 */
int pxg_get_pattern(pxg_ctx * ctx, int64_t * rowptr, int32_t * col, uint32_t * emap);

/* per element and quadrature point geometry of CurFE::reinit
 * (include/curfe.hpp:50-103): detJ, JxW (n_elem x nq each)                 */
int pxg_get_geometry(pxg_ctx * ctx, double * detJ, double * JxW);

/* ---- operator apply (device pointers, the context's stream) ------------- */
/* y = A x with the closed matrix (SparseMatrix<RowMajor> * Vec)            */
int pxg_spmv(pxg_ctx * ctx, int precision, const void * x_dev, void * y_dev);
/* sweeps of x <- x + omega D^-1 (b - A x); returns when they are complete */
int pxg_jacobi(pxg_ctx * ctx, int precision, double omega, int sweeps, const void * b_dev, void * x_dev);
/* pxg_jacobi that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * other vector of the sweeps is a device array of 2^22 doubles (32 MiB) per
 * GPU, made when the library's code object loads on that GPU and held until
 * the process ends; a matrix with more than 4,194,304 rows is rejected
 * (PXG_E_INVALID).  The closed matrix is applied as it is: its row blocks
 * and packed copy only if an earlier pxg_spmv / pxg_jacobi made them
 * current, the plain CSR kernels otherwise (same results).  Returns without
 * waiting for the GPU.  Every call on one GPU shares the array and leaves
 * its last contents there: calls that overlap in time on one GPU (other
 * streams, other threads) race.                                            */
int pxg_jacobi_noalloc(pxg_ctx * ctx, int precision, double omega, int sweeps, const void * b_dev, void * x_dev);
/* device pointers to the closed CSR arrays (valid until the next close)    */
int pxg_csr_device(pxg_ctx * ctx, const int64_t ** rowptr, const int32_t ** col, const double ** val,
                   const float ** val32, int64_t * nnz);
/* device pointer to b (n_dofs doubles)                                     */
int pxg_rhs_device(pxg_ctx * ctx, double ** b_dev);
/* number of closes of this context so far: state derived from the closed
 * matrix (a factorisation, a saved solve) is current while it is unchanged.
 * No counterpart in the reference, whose matrices are plain values; it is
 * for callers that cache per-matrix state (autograd solve, multigrid).     */
int pxg_matrix_stamp(pxg_ctx * ctx, uint64_t * stamp);
/* enable 1: pxg_spmv / pxg_jacobi (and the multigrid on this level) read a
 * block-packed copy of the closed matrix, made on first use after each
 * close: one base column per block of <= 2048 entries, 18-bit column
 * offsets, so an entry streams 6.25 (FP32) / 10.25 (FP64) bytes instead of
 * 8 / 12; results are bitwise those of the plain CSR.  Skipped (plain CSR)
 * when some block spans 2^18 columns or more.  enable 0: plain CSR;
 * negative: unchanged.  *active (may be NULL): 1 if the packed copy of the
 * current closed matrix exists (it is made by the first SpMV / Jacobi).
 * No counterpart in the reference (Eigen's SpMV reads its CSR arrays); it
 * is for callers that apply one matrix many times.  pxg_mg_create enables
 * it on its levels.                                                        */
int pxg_csr_pack(pxg_ctx * ctx, int enable, int * active);

/* ---- matrix-free y = K(kappa) x and its gradients ------------------------
 * nb vectors interleaved as [n_dofs][nb]; use_bc applies the Dirichlet form
 * C K C + D_c of the assembled matrix.  Geometry is recomputed per call.
 * Scalar FE types only (PXG_E_INVALID for a Taylor-Hood context).          */
int pxg_mf_apply(pxg_ctx * ctx, const double * kappa_dev, int nb, const double * x_dev, double * y_dev,
                 int use_bc);
/* gk[n] = sum_b d/dkappa_n ( ybar_b . K(kappa) x_b )                         */
int pxg_mf_grad_kappa(pxg_ctx * ctx, int nb, const double * x_dev, const double * ybar_dev, double * gk_dev,
                      int use_bc);

/* ---- geometric multigrid-preconditioned CG ------------------------------ */
typedef struct pxg_mg pxg_mg;
/* levels: closed contexts (finest first) on structured hex meshes whose
 * element counts halve per level; dims: n_lev x 3 element counts.
 * Weighted Jacobi (omega, nu1 pre / nu2 post sweeps), R = P^T, coarsest
 * level solved directly in FP64 (dense inverse of its free-dof block, up to
 * 4096 free dofs; larger: Jacobi-CG to coarse_rtol).  Enables
 * pxg_csr_pack on every level (pxg_csr_pack(level, 0) afterwards turns it
 * off again).                                                              */
int pxg_mg_create(pxg_mg ** mg, int n_lev, pxg_ctx ** levels, const int * dims, double omega, int nu1, int nu2,
                  double coarse_rtol);
int pxg_mg_destroy(pxg_mg * mg);
/* refresh level matrices after a re-assembly (same patterns): refactors
 * the coarsest level now (solve and vcycle also do it when they find the
 * coarsest matrix closed again since the last factorisation); needs the
 * coarsest level closed, the others are checked by the solve              */
int pxg_mg_update(pxg_mg * mg);
/* solve A x = b (device pointers, FP64); mode PXG_FP64: MG-PCG in FP64;
 * PXG_FP32: MG-PCG on FP32 matrices and vectors with an FP64 iterate,
 * refined by replacing the CG residual with the FP64 residual b - A x each
 * time it has fallen a hundredfold (FP64 inner products, FP64 coarsest
 * solve; convergence judged on the FP64 residual).
 * PXG_FP32_MF: as PXG_FP32, but only the coarsest level needs an
 * assembled matrix: every other level's operator is the kappa-stiffness
 * form of AssemblyStiffnessFE with the Dirichlet rows of the assembled form
 * (C K C + D_c), applied matrix-free from the kappa field of that level's
 * context -- FP32 contractions with the per-point metrics computed once per
 * kappa (6 FP32 values per point), FP32 inverse diagonal for the smoother;
 * on level 0 the FP64 sum-factorised operator gives the refinement
 * residual.  Q1 / Q2 hexahedra with their Gauss rule.                      */
int pxg_mg_solve(pxg_mg * mg, int mode, const double * b_dev, double * x_dev, double rtol, int maxit,
                 int * iters, double * relres);
/* one V-cycle x = M^{-1} b (FP64), for testing                            */
int pxg_mg_vcycle(pxg_mg * mg, const double * b_dev, double * x_dev);
/* prolongation P_l (level l+1 -> l) as host CSR; returns nnz via *nnz     */
int pxg_mg_prolongation(pxg_mg * mg, int level, int64_t * rowptr, int32_t * col, double * val, int64_t * nnz);
/* Restrictor of the reference (include/multigrid.hpp:108-175, level l -> l+1):
 * P_l^T with every row scaled to sum one, as host CSR (the cycle applies
 * P^T itself, the Galerkin choice for CG)                                  */
int pxg_mg_restriction(pxg_mg * mg, int level, int64_t * rowptr, int32_t * col, double * val, int64_t * nnz);
const char * pxg_mg_last_error(const pxg_mg * mg);

#ifdef __cplusplus
}
#endif

#endif /* PROXPDE_AMD_H */
