/*
 * gsx.h -- C ABI of libgsx, the MI355X (gfx950) isogeometric Poisson
 * assemble + solve library.
 *
 * Plain pointers and sizes only.  Host pointers are named *_host; device
 * pointers (from gsx_device_alloc or any HIP allocation) are named *_dev.
 * Every entry point returns a gsx_status (0 = success); the message of the
 * last failure of a context is returned by gsx_last_error().
 *
 * Each entry point replaces one step of the G+Smo hot path; the comment
 * names the reference interface (paths relative to the G+Smo source tree).
 * The C++ drop-in classes (gismo_amd/host/gismo) and the Python mirror
 * (gismo_amd) are written on top of these functions.
 */
#ifndef GSX_H
#define GSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GSX_OK = 0,
    GSX_ERR_ARG = 1,           /* invalid argument / unsupported input       */
    GSX_ERR_OOM = 2,           /* device allocation failed                   */
    GSX_ERR_JACOBIAN = 3,      /* Jacobian vanishes or changes sign          */
    GSX_ERR_HIP = 4,           /* HIP runtime error                          */
    GSX_ERR_NOT_CONVERGED = 5, /* solver stopped at max iterations           */
    GSX_ERR_NO_DEVICE = 6,     /* no GPU visible                             */
    GSX_ERR_BUDGET = 7         /* request exceeds the HBM budget             */
} gsx_status;

typedef enum {
    GSX_FP64 = 0,   /* plain FP64 everywhere                                  */
    GSX_MIXED = 1   /* FP64 operator + FP32 inner iterations, FP64 residuals  */
} gsx_precision;

typedef enum {
    GSX_PATTERN_EXPR = 0,    /* gsExprAssembler: skip exact-zero local entries */
    GSX_PATTERN_FULL = 1,    /* gsPoissonAssembler: every free-free pair       */
    GSX_ASM_REBUILD_PATTERN = 0x100, /* flag: recompute the cached pattern     */
    GSX_ASM_CSR_ONLY = 0x200,        /* flag: no structured operator (CSR
                                        solves only, e.g. batched members)    */
    GSX_ASM_KEEP_ELEMENTS = 0x400,   /* flag: the problem keeps this assembly's
                                        element matrices (when they were formed
                                        in one pass) until its next assembly;
                                        gsx_mg_create on the returned matrix
                                        takes them instead of recomputing      */
    GSX_ASM_ELEMENTWISE = 0x800,     /* flag: element matrices + row gather
                                        (summed in element order) also where
                                        the sum-factorised row assembly of a
                                        single 3D patch would run              */
    GSX_ASM_ROWWISE = 0x1000         /* flag: the sum-factorised row assembly
                                        also for degrees 1 and 4 (single 3D
                                        patch; degrees 2, 3 take it anyway)    */
} gsx_pattern_mode;

typedef enum {
    GSX_PREC_NONE = 0,
    GSX_PREC_JACOBI = 1,
    GSX_PREC_MG = 2,
    GSX_PREC_FD = 3      /* fast diagonalization: gsx_cg_solve_fd */
} gsx_precond;

typedef enum {
    GSX_SMOOTH_JACOBI = 0,      /* damped Jacobi (default)                   */
    GSX_SMOOTH_RICHARDSON = 1,  /* x += damping (b - A x)                    */
    GSX_SMOOTH_CHEBYSHEV = 2,   /* Jacobi-preconditioned Chebyshev            */
    GSX_SMOOTH_GAUSS_SEIDEL = 3, /* forward sweeps before, reverse after       */
    GSX_SMOOTH_SCMS = 4          /* subspace corrected mass smoother           */
} gsx_smoother;

typedef struct gsx_ctx gsx_ctx;
typedef struct gsx_fd gsx_fd;
typedef struct gsx_problem gsx_problem;
typedef struct gsx_matrix gsx_matrix;
typedef struct gsx_mg gsx_mg;

/* One tensor-product B-spline patch, geometry and discretisation sharing the
 * basis (isoparametric).  Knot vectors are open with strictly increasing
 * interior knots; degree 1..4 per direction; d = 2 or 3.  Control points are
 * row-major nbasis x d, lexicographic with direction 0 fastest.  kappa and
 * felem (per element, lexicographic) may be NULL.
 * fkind 1: f = d*pi^2*prod_r sin(pi x_r); fkind 0: f = felem[e]. */
typedef struct {
    int32_t d;
    int32_t p[3];
    int32_t nknots[3];
    const double *knots_host[3];
    const double *ctrl_host;
    const double *kappa_host;
    const double *felem_host;
    int32_t fkind;
} gsx_patch_desc;

/* ---- context --------------------------------------------------------- */
/* No reference counterpart: owns the HIP stream, events and the scratch
 * pool of one device. */
gsx_status gsx_ctx_create(int32_t device, gsx_ctx **out);
gsx_status gsx_ctx_destroy(gsx_ctx *ctx);
const char *gsx_last_error(const gsx_ctx *ctx);
gsx_status gsx_ctx_synchronize(gsx_ctx *ctx);
/* The device's architecture name (e.g. "gfx950:sramecc+:xnack-"), NUL-
 * terminated in buf (cap bytes); keys the autotuner's tables. */
gsx_status gsx_ctx_arch(const gsx_ctx *ctx, char *buf, int32_t cap);
/* Kernel timing: when enabled, every launch is bracketed by HIP events on
 * the context's stream; gsx_ctx_kernel_stats returns, for kernel name
 * `name`, the number of launches and the summed milliseconds since the last
 * reset. */
gsx_status gsx_ctx_enable_timing(gsx_ctx *ctx, int32_t enable);
gsx_status gsx_ctx_reset_timing(gsx_ctx *ctx);
/* Restrict timing to the scopes named `name` (NULL or "": all scopes). */
gsx_status gsx_ctx_timing_filter(gsx_ctx *ctx, const char *name);
gsx_status gsx_ctx_kernel_stats(gsx_ctx *ctx, const char *name,
                                int64_t *launches, double *total_ms);
/* Device memory helpers for callers without their own allocator. */
gsx_status gsx_device_alloc(gsx_ctx *ctx, int64_t bytes, void **dev_out);
gsx_status gsx_device_free(gsx_ctx *ctx, void *dev);
gsx_status gsx_copy_to_device(gsx_ctx *ctx, void *dev, const void *host, int64_t bytes);
gsx_status gsx_copy_to_host(gsx_ctx *ctx, void *host, const void *dev, int64_t bytes);
/* Bytes of device memory currently held by the context. */
gsx_status gsx_ctx_memory(gsx_ctx *ctx, int64_t *bytes_in_use);
/* Device bytes the element matrices of one assembly pass may occupy
 * (default 16 GiB).  A single patch larger than this is assembled in slabs
 * of element layers (the result is identical); a multipatch assembly that
 * does not fit fails with GSX_ERR_BUDGET.  No reference counterpart (the
 * reference assembles element by element on the host). */
gsx_status gsx_ctx_set_elem_budget(gsx_ctx *ctx, int64_t bytes);
/* Largest coarsest multigrid level (rows, 1..32768, default 32768) solved
 * exactly through its dense FP64 inverse, one GEMV per coarse solve.  A
 * larger coarsest level is factored as a banded FP64 Cholesky in reverse
 * Cuthill-McKee order (half-bandwidth at most 4096); both are exact solves,
 * like the reference's sparse Cholesky of its coarsest level
 * (src/gsSolver/gsMatrixOp.h:357-370, makeSparseCholeskySolver).  Read when
 * a multigrid is created. */
gsx_status gsx_ctx_set_coarse_dense_max(gsx_ctx *ctx, int64_t rows);

/* ---- host-only building blocks (no device needed) ---------------------- */
/* Gauss-Legendre rule with n nodes on [-1,1], ascending: the values of
 * gsGaussRule::lookupReference (src/gsAssembler/gsGaussRule.hpp:218-547). */
gsx_status gsx_gauss_legendre(int32_t n, double *x_host, double *w_host);
/* The DOF mapper alone (what gsx_problem_create uses): nbasis_dir holds
 * npatch x 3 basis sizes per direction; dofs_out receives the global index
 * of every (patch, local) function; gsDofMapper::finalize numbering
 * (src/gsCore/gsDofMapper.cpp:240-340). */
gsx_status gsx_mapper_build(int32_t d, int32_t npatch, const int32_t *nbasis_dir_host,
                            int32_t ndir, const int32_t *dir_sides_host, int32_t nif,
                            const int32_t *ifaces_host, int32_t *dofs_out_host,
                            int64_t *nfree, int64_t *nglued);

/* ---- problem: patches + DOF mapper ------------------------------------ */
/* Replaces gsMultiPatch/gsMultiBasis + gsFeSpace::setup(bc, dirichlet, 0)
 * building the gsDofMapper (src/gsAssembler/gsExpressions.h:1091-1121,
 * src/gsCore/gsDofMapper.cpp:99-340): interfaces glued (C0), Dirichlet sides
 * eliminated, free DOFs numbered standard-first, then coupled, then
 * eliminated.  dir_sides: ndir pairs (patch, side 1..2d).  ifaces: nif
 * records (p1, s1, p2, s2, m0, f0, m1, f1). */
gsx_status gsx_problem_create(gsx_ctx *ctx, int32_t npatch,
                              const gsx_patch_desc *patches, int32_t ndir,
                              const int32_t *dir_sides_host, int32_t nif,
                              const int32_t *ifaces_host, gsx_problem **out);
gsx_status gsx_problem_destroy(gsx_problem *prob);
/* A problem discretised with the bases of `patches` (degree and knot
 * vectors; control points, kappa and sources are not read) over the
 * geometry map, kappa and source of `geo`: gsCDRAssembler(mp, bases[i], ...)
 * of examples/pMultiGrid_example.cpp:445-460, where a level's basis differs
 * from the geometry's in degree or mesh.  Each patch covers the parameter box
 * of geo's patch; the quadrature follows the discretisation basis (q + 1
 * Gauss points per direction on its elements) and the geometry is evaluated
 * in its own basis at every point.  gsx_assemble takes the element path;
 * geo must outlive the problem.  Boundary data and error norms need the
 * problem's own geometry (GSX_ERR_ARG). */
gsx_status gsx_problem_create_geo(gsx_problem *geo, int32_t npatch, const gsx_patch_desc *patches,
                                  int32_t ndir, const int32_t *dir_sides_host, int32_t nif,
                                  const int32_t *ifaces_host, gsx_problem **out);
/* nbasis: all basis functions (summed over patches, before gluing);
 * nfree: free DOFs; nelem: elements; nglued: glued space dimension. */
gsx_status gsx_problem_info(const gsx_problem *prob, int64_t *nbasis,
                            int64_t *nfree, int64_t *nelem, int64_t *nglued);
/* Global DOF index of each (patch, local basis function), concatenated over
 * patches (gsDofMapper::index, src/gsCore/gsDofMapper.h:325). */
gsx_status gsx_problem_dofs(const gsx_problem *prob, int32_t *dofs_host);

/* ---- boundary data ------------------------------------------------------ */
/* Boundary condition of one patch side (gsBoundaryConditions::
 * addCondition, src/gsPde/gsBoundaryConditions.h).  type: Dirichlet sides
 * must be among the problem's dir_sides (they fix the DOF map); a Dirichlet
 * side without a record is homogeneous.  Neumann sides are boundary sides
 * that are not Dirichlet.  kind and value:
 *   GSX_BCF_ZERO          g = 0;
 *   GSX_BCF_MANUFACTURED  Dirichlet g = prod_r sin(pi x_r); Neumann flux
 *                         grad g (integrand N_i grad g . nv(G));
 *   GSX_BCF_AFFINE        Dirichlet g = value[0] + sum_r value[1+r] x_r;
 *                         Neumann flux (value[1], value[2], value[3]).
 * x are physical coordinates. */
typedef enum { GSX_BC_DIRICHLET = 1, GSX_BC_NEUMANN = 2 } gsx_bc_type;
typedef enum { GSX_BCF_ZERO = 0, GSX_BCF_MANUFACTURED = 1, GSX_BCF_AFFINE = 2 } gsx_bc_kind;
typedef struct {
    int32_t patch;     /* 0-based */
    int32_t side;      /* 1..2d: west, east, south, north, front, back */
    int32_t type;      /* gsx_bc_type */
    int32_t kind;      /* gsx_bc_kind */
    double value[4];
} gsx_bc;
/* Values of the eliminated Dirichlet DOFs (gsFeSpace::setup(bc,
 * dir_values, 0), src/gsAssembler/gsDirichletValues.h:27-60):
 * GSX_DIR_HOMOGENEOUS zero, GSX_DIR_L2PROJECTION the L2 projection of g onto
 * the Dirichlet sides' functions (gsDirichletValuesByL2Projection,
 * gsDirichletValues.h:257-435: p+1 Gauss points along the side, weights
 * w |det J|, solved to rounding level).  Neumann records add the boundary
 * load of gsExprAssembler::assembleBdr(bc.get("Neumann"), u * g_N.tr() *
 * nv(G)) (src/gsAssembler/gsExprAssembler.h:835-895) through
 * gsx_assemble_bdr; every later gsx_assemble moves the fixed columns to the
 * right-hand side (rhs_i -= K_ij g_j, gsExprAssembler.h:553-630).  Replaces
 * any earlier boundary data of the problem. */
/* GSX_DIR_INTERPOLATION: gsAssembler::computeDirichletDofsIntpl
 * (src/gsAssembler/gsAssembler.hpp:314-390, the gsPoissonAssembler default
 * dirichlet::interpolation): per Dirichlet side in problem order, g at the
 * side's anchor grid (Greville abscissae, normal coordinate at the knot
 * vector's end) mapped by the geometry, interpolated by the side's tensor
 * basis; a later side overwrites the functions it shares with an earlier
 * one, homogeneous sides write zeros. */
/* GSX_DIR_L2PROJECTION_INHOM: the projection of gsAssembler::
 * computeDirichletDofsL2Proj (src/gsAssembler/gsAssembler.hpp:392-534, the
 * gsPoissonAssembler path), which skips the homogeneous Dirichlet sides:
 * their functions keep 0 unless an inhomogeneous side shares them. */
typedef enum {
    GSX_DIR_HOMOGENEOUS = 0,
    GSX_DIR_L2PROJECTION = 1,
    GSX_DIR_L2PROJECTION_INHOM = 2,
    GSX_DIR_INTERPOLATION = 3
} gsx_dir_values;
gsx_status gsx_problem_set_bc(gsx_problem *prob, int32_t nbc, const gsx_bc *bc,
                              int32_t dir_values);
/* Host copy of the fixed (eliminated) DOF values, in the order of their
 * global index minus nfree (gsFeSpace::fixedPart); count = nglued - nfree. */
gsx_status gsx_problem_fixed_values(const gsx_problem *prob, double *fixed_host,
                                    int32_t count);
/* rhs_dev (nfree, device) += the Neumann boundary load of the problem's
 * boundary data: gsExprAssembler::assembleBdr(bc.get("Neumann"),
 * u * g_N.tr() * nv(G)) (src/gsAssembler/gsExprAssembler.h:835-895). */
gsx_status gsx_assemble_bdr(gsx_problem *prob, double *rhs_dev);

/* ---- assembly ---------------------------------------------------------- */
/* Replaces gsExprAssembler<T>::assemble(igrad(u,G)*igrad(u,G).tr()*meas(G)
 * [*kappa], u*ff*meas(G)) (src/gsAssembler/gsExprAssembler.h:754-833) and
 * gsPoissonAssembler<T>::assemble() (src/gsAssembler/gsPoissonAssembler.hpp:
 * 41-78): builds the CSR pattern and FP64 values of the free-free stiffness
 * matrix and the load vector on the device.  The matrix is symmetric
 * bitwise, so its CSR arrays are also its CSC arrays (gsSparseMatrix is
 * column-major, src/gsCore/gsForwardDeclarations.h:279-280).
 * precision GSX_MIXED additionally keeps an FP32 copy of the values for the
 * inner iterations.  rhs_dev: nfree doubles (device) or NULL.
 * A single 3D patch of equal degree 2, 3 or 4 in every direction (also
 * degree 1 with GSX_ASM_ROWWISE) whose kappa has no exact zero is assembled row by row by
 * sum factorisation on the FP64 matrix cores, without element matrices (the
 * same sums over the same Gauss points, grouped by point columns instead of
 * elements: values agree with the element path to rounding).  There no
 * entry is dropped by the exact-zero rule: an entry with only exact-zero
 * element contributions needs kappa = 0 on every element they share.  Other
 * problems, GSX_ASM_KEEP_ELEMENTS and GSX_ASM_ELEMENTWISE take the element
 * path. */
gsx_status gsx_assemble(gsx_problem *prob, int32_t pattern_mode,
                        int32_t precision, gsx_matrix **A_out,
                        double *rhs_dev);
gsx_status gsx_matrix_destroy(gsx_matrix *A);
/* Device matrix from host CSR arrays (any square sparse matrix, e.g. one a
 * caller assembled itself): what gsMatrixOp wraps (src/gsSolver/
 * gsMatrixOp.h:64-75).  with_fp32 keeps an FP32 copy for mixed solves. */
gsx_status gsx_matrix_from_csr(gsx_ctx *ctx, int32_t n, const int32_t *rowptr_host,
                               const int32_t *col_host, const double *val_host,
                               int32_t with_fp32, gsx_matrix **out);
/* Return the context's cached device blocks to the driver.  No reference
 * counterpart. */
gsx_status gsx_ctx_release_cache(gsx_ctx *ctx);
gsx_status gsx_matrix_info(const gsx_matrix *A, int64_t *n, int64_t *nnz);
/* Host copies of the CSR arrays (n+1, nnz, nnz). */
gsx_status gsx_matrix_download(const gsx_matrix *A, int32_t *rowptr_host,
                               int32_t *col_host, double *val_host);
/* Device pointers of the CSR arrays (owned by the matrix). */
gsx_status gsx_matrix_device_ptrs(const gsx_matrix *A, const int32_t **rowptr_dev,
                                  const int32_t **col_dev, const double **val_dev);
/* Smallest oriented Jacobian determinant (det J times the sign of the
 * patch's parametrization, taken from its control net at the parametric
 * origin) met at any Gauss point of the last assembly of prob.  The measure
 * is |det J| (meas(G), src/gsAssembler/gsExpressions.h:752-781), so patches
 * of either orientation are valid; a non-positive value (a fold, a
 * degenerate point) makes gsx_assemble fail with GSX_ERR_JACOBIAN. */
gsx_status gsx_problem_min_detj(const gsx_problem *prob, double *min_det);

/* ---- operators ----------------------------------------------------------- */
/* Operator format used by the solvers for A: structured = 1 (default) uses
 * the (2p+1)^d stencil form of single-patch matrices (no column indices,
 * vectors on the padded tensor grid); 0 forces CSR.  Multipatch matrices
 * always use CSR.  No reference counterpart (gsMatrixOp always applies the
 * sparse matrix). */
gsx_status gsx_matrix_set_operator(gsx_matrix *A, int32_t structured);
/* y = A x on the device (gsMatrixOp::apply, src/gsSolver/gsMatrixOp.h:74-75). */
gsx_status gsx_spmv(gsx_matrix *A, const double *x_dev, double *y_dev);

/* ---- solvers --------------------------------------------------------------- */
/* Preconditioned CG: gsConjugateGradient::solve (src/gsSolver/
 * gsConjugateGradient.hpp:20-93, src/gsSolver/gsIterativeSolver.h:114-200)
 * with gsJacobiOp (src/gsSolver/gsSimplePreconditioners.h:146-244, damping
 * tau) or gsMultiGridOp (mg != NULL) as preconditioner.  Stops when
 * ||r||_2/||b||_2 < tol; x_dev holds the initial guess on entry.
 * precision GSX_MIXED: iterations in FP32 with FP64 reliable-update
 * residuals b - A x (the FP64 operator), the stopping test on the FP64
 * residual.  Returns GSX_ERR_NOT_CONVERGED (not an error of the call) when
 * maxit is reached; iters/relres are always written. */
gsx_status gsx_cg_solve(gsx_matrix *A, int32_t precond, double tau,
                        gsx_mg *mg, int32_t precision, const double *b_dev,
                        double *x_dev, double tol, int32_t maxit,
                        int32_t *iters, double *relres);
/* Relative residual history of the last FP64 solve of A (iters+1 values,
 * gsIterativeSolver::solveDetailed, gsIterativeSolver.h:134-165). */
/* Fast diagonalization preconditioner of a single-patch problem
 * (gsPatchPreconditionersCreator::fastDiagonalizationOp(basis, bc, opt,
 * alpha, beta), src/gsSolver/gsPatchPreconditionersCreator.hpp:258-348):
 * the exact inverse of alpha M + beta sum_k M x .. x K_k x .. x M, with K_k,
 * M_k the univariate stiffness and mass of knot vector k on the parameter
 * line (p+1 Gauss points per span, the Dirichlet ends of the problem
 * removed, gsPatchPreconditionersCreator.hpp:27-126), applied as
 * (Q x .. x Q) diag(1/(alpha + sum lambda)) (Q x .. x Q)^T with
 * K_k Q_k = M_k Q_k diag(lambda_k), Q_k^T M_k Q_k = I.  Needs beta > 0 and,
 * without a Dirichlet end in any direction, alpha > 0 (GSX_ERR_ARG). */
gsx_status gsx_fd_create(gsx_problem *prob, double alpha, double beta, gsx_fd **out);
/* Subspace corrected mass smoother of a single-patch problem
 * (gsPatchPreconditionersCreator::subspaceCorrectedMassSmootherOp(basis,
 * bc, opt, sigma, alpha, beta), src/gsSolver/gsPatchPreconditionersCreator.
 * hpp:574-758, tilde spaces :368-512): per direction the space of splines
 * with vanishing odd derivatives at the ends and its complement, the sum
 * over the 2^d subspace combinations of the transfer, the interior mass
 * inverses scaled by 1/(alpha + beta (d - |J|)/(sigma h^2)) and the
 * boundary correction.  Returned in the gsx_fd handle: gsx_fd_apply and
 * gsx_cg_solve_fd use it like the fast diagonalization. */
gsx_status gsx_scms_create(gsx_problem *prob, double sigma, double alpha, double beta, gsx_fd **out);
gsx_status gsx_fd_destroy(gsx_fd *fd);
/* z = FD r over the free DOFs (device vectors, nfree entries; may alias). */
gsx_status gsx_fd_apply(gsx_fd *fd, const double *r, double *z);
/* gsConjugateGradient(A, fastDiagonalizationOp(..)).solve(b, x)
 * (src/gsSolver/gsConjugateGradient.hpp:20-93) on the structured operator of
 * a matrix assembled from the fast diagonalization's problem; GSX_MIXED runs
 * FP32 iterations (FP32 operator and factors) with FP64 reliable updates, as
 * gsx_cg_solve does. */
gsx_status gsx_cg_solve_fd(gsx_matrix *A, gsx_fd *fd, int32_t precision, const double *b, double *x,
                           double tol, int32_t maxit, int32_t *iters, double *relres);
gsx_status gsx_cg_history(gsx_matrix *A, double *hist_host, int32_t cap,
                          int32_t *count);
/* Allocation-free CG: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): gsx_cg_solve's FP64 CSR
 * CG (precond GSX_PREC_NONE or GSX_PREC_JACOBI, damping tau) with its whole
 * workspace in a fixed 64 MiB device array per GPU (held until the process
 * ends): 5 n + 4 nparts + maxit + 1 doubles plus the solver state, nparts
 * about n / 256, each piece rounded up to 256 bytes -- about 1.6 M rows at
 * maxit 1000.  Larger requests return GSX_ERR_ARG.  No hipMalloc, hipFree
 * or allocating library call is made; the caller passes b and x (device,
 * n doubles).  The residual history is not kept (gsx_cg_history does not
 * see it).  Every call on one GPU uses the same array: calls that overlap on
 * one GPU (several host threads or streams at once) race. */
gsx_status gsx_cg_solve_noalloc(gsx_matrix *A, int32_t precond, double tau, const double *b, double *x,
                                double tol, int32_t maxit, int32_t *iters, double *relres);

/* ---- geometric multigrid -------------------------------------------------- */
/* gsMultiGridOp::make + init (src/gsMultiGrid/gsMultiGrid.hpp:19-94) on a
 * hierarchy of nlevels obtained by coarsening every patch's knot vectors
 * (gsGridHierarchy, src/gsMultiGrid/gsGridHierarchy.hpp:27-101; the fine
 * problem needs an even element count per direction on every level):
 * transfers from knot insertion restricted to free DOFs
 * (gsMultiBasis::combineTransferMatrices, src/gsCore/gsMultiBasis.hpp:
 * 137-178), Galerkin coarse operators P^T A P, damped Jacobi smoothing
 * (npre/npost steps, damping), coarse-grid correction damping, exact FP64
 * coarse solve.  precision GSX_MIXED runs the cycle in FP32.  Only
 * V-cycles (ncycles = 1). */
gsx_status gsx_mg_create(gsx_problem *fine, gsx_matrix *A, int32_t nlevels,
                         int32_t npre, int32_t npost, int32_t ncycles,
                         double damping, double cgc_damping, int32_t precision,
                         gsx_mg **out);
/* gsMultiGridOp::make(fineMatrix, transferMatrices) and its init
 * (src/gsMultiGrid/gsMultiGrid.h:136-141, gsMultiGrid.hpp:19-35,70-94):
 * the hierarchy given by the caller's transfers.  Transfer l (0 <= l <
 * ntransfer) prolongates level l to level l + 1: a row-major CSR matrix
 * (host arrays prow[l] of n_{l+1} + 1 entries, pcol[l]/pval[l], columns
 * ascending within a row) of n_{l+1} rows and ncoarse[l] columns, where
 * level ntransfer is A.  The coarse operators P^T A P and the restrictions
 * P^T are formed on the device; smoothing and the exact coarse solve are as
 * in gsx_mg_create.  The subspace corrected mass smoother needs the
 * hierarchy of gsx_mg_create. */
gsx_status gsx_mg_create_transfers(gsx_matrix *A, int32_t ntransfer, const int32_t *ncoarse,
                                   const int32_t *const *prow, const int32_t *const *pcol,
                                   const double *const *pval, int32_t npre, int32_t npost,
                                   int32_t ncycles, double damping, double cgc_damping,
                                   int32_t precision, gsx_mg **out);
/* p-multigrid as examples/pMultiGrid_example.cpp:296-560 sets it up, on the
 * fine problem and its assembled matrix A (nlevels levels, level
 * nlevels - 1 = A).  coarsening[l] (l = 0 .. nlevels - 2) is the transition
 * from level l to level l + 1 (--Coarsening): 'h' a coarser mesh (every other
 * knot, same degree), 'p' a lower degree (same mesh), 'z' both.  projection
 * (--Projection): 1 lowers the degree straight to 1 (one p or z transition),
 * 2 by one.  Level bases keep simple interior knots (the smoothness of
 * gsMultiBasis::degreeIncrease over them).  Operators: p and z levels are
 * assembled on their own basis over the fine geometry (gsx_problem_create_geo);
 * h levels too when coarse_operator = 1, or P^T A P when coarse_operator = 2
 * (--CoarseOperator).  Transfers (--Lumping 1): p and z, L2 projections
 * with lumped mass, P = L_f^-1 M_fc and R = L_c^-1 M_fc^T with
 * M_fc(i, j) = int phi_i psi_j on the finer basis' elements and
 * L(i) = int phi_i (assembleMixedMass/assembleLumpedMass, :803-880);
 * h, the canonical embedding and its transpose.  The coarsest level is
 * solved exactly (dense FP64 inverse); V-cycles with the smoothers of
 * gsx_mg_set_smoother / gsx_mg_set_level_smoother (damped Jacobi by
 * default).  fine must outlive the call. */
gsx_status gsx_pmg_create(gsx_problem *fine, gsx_matrix *A, int32_t nlevels, const char *coarsening,
                          int32_t projection, int32_t coarse_operator, int32_t npre, int32_t npost,
                          double damping, double cgc_damping, int32_t precision, gsx_mg **out);
/* Host copy of the restriction of level lvl (0 <= lvl < nlevels - 1): CSR
 * with one row per free DOF of level lvl, columns of level lvl + 1 (P^T for
 * h transitions, L_c^-1 M_fc^T for p and z); any pointer may be NULL. */
gsx_status gsx_mg_restriction_download(const gsx_mg *mg, int32_t lvl, int32_t *rowptr_host,
                                       int32_t *col_host, double *val_host);
gsx_status gsx_mg_destroy(gsx_mg *mg);
/* Smoother of every level but the coarsest (the sweeps npre/npost of
 * gsx_mg_create).  GSX_SMOOTH_JACOBI: gsJacobiOp::step with the damping of
 * gsx_mg_create (src/gsSolver/gsSimplePreconditioners.h:178-202);
 * GSX_SMOOTH_RICHARDSON: gsRichardsonOp::step, x += damping (b - A x)
 * (same file :33-131); GSX_SMOOTH_CHEBYSHEV (no reference counterpart):
 * npre/npost-step Chebyshev iteration on D^-1 A over [lmax/param, lmax]
 * (param <= 1: 30), lmax = 1.1 x a 20-step power-iteration estimate. */
/* GSX_SMOOTH_GAUSS_SEIDEL: gsGaussSeidelOp (makeGaussSeidelOp, the default
 * smoother of examples/multiGrid_example.cpp:41) with gsMultiGridOp's
 * SymmSmooth: forward sweeps (step) before and reverse sweeps (stepT) after
 * the coarse correction (gsMultiGrid.hpp:166-199,
 * gsSimplePreconditioners.h:310-329, sweeps gsSimplePreconditioners.hpp:
 * 21-68), run exactly in the sequential order by level sets of the
 * triangular dependencies (one launch per set: made for parity, slow on
 * large levels). */
/* GSX_SMOOTH_SCMS: the subspace corrected mass smoother as the reference's
 * multigrid example sets it up (setupSubspaceCorrectedMassSmoother,
 * examples/multiGrid_example.cpp:407-567, conforming interfaces).  For a
 * single patch with Dirichlet conditions on every side:
 * x += tau_l S_l (b - A_l x), S_l the smoother of level l's basis
 * (gsx_scms_create with sigma = param, 0.12 when param <= 0, alpha = 0,
 * beta = 1), tau_l = 1 / lambda with lambda from ten power iterations of
 * S_l A_l (src/gsSolver/gsPreconditioner.h:121-132; the reference starts
 * from a random vector, here v0_i = 1 / (1 + i mod 5)); the first level
 * above nlevels - 5 with more elements than the degree in every direction
 * keeps its tau for all finer levels.  Otherwise (several patches, or sides
 * without a Dirichlet condition) the free DOFs split into the components of
 * gsBoxTopology::allComponents(true): patch interiors, sides (glued sides
 * are one), edges of 3D patches, and all corners together;
 * x += sum_c T_c op_c T_c^T (b - A_l x) with op_c the smoother of the patch
 * interior or 3D face basis (faces: alpha = p/h, beta = h/p) damped per
 * component as above, and the exact inverse of A_l restricted to the other
 * components (at most 2048 DOFs each).  Several components need FP64
 * cycles and a hierarchy built from a problem (gsx_mg_create,
 * gsx_pmg_create). */
gsx_status gsx_mg_set_smoother(gsx_mg *mg, int32_t kind, double param);
/* gsMultiGridOp::setSmoother(lvl, op) (src/gsMultiGrid/gsMultiGrid.h:295):
 * the smoother of level lvl alone (1 <= lvl < nlevels), kind and param as in
 * gsx_mg_set_smoother, damping = the Jacobi/Richardson damping of that
 * level's operator (gsJacobiOp/gsRichardsonOp option Damping).  Levels are
 * set up in ascending order as the reference's example does; a scms level
 * keeps the damping of the first qualifying coarser level. */
gsx_status gsx_mg_set_level_smoother(gsx_mg *mg, int32_t lvl, int32_t kind, double damping, double param);
/* gsMultiGridOp::setOptions (src/gsMultiGrid/gsMultiGrid.hpp:300-330):
 * NumPreSmooth, NumPostSmooth and CoarseGridCorrectionDamping of an
 * existing hierarchy. */
gsx_status gsx_mg_set_cycle(gsx_mg *mg, int32_t npre, int32_t npost, double cgc_damping);
/* The level parameter of level lvl (0 before gsx_mg_set_smoother): the
 * Chebyshev bound lmax, or tau_l of GSX_SMOOTH_SCMS. */
gsx_status gsx_mg_smoother_bound(const gsx_mg *mg, int32_t lvl, double *lmax);
/* z = MG(r): one cycle from a zero start (gsPreconditionerOp::apply,
 * src/gsSolver/gsPreconditioner.h:79-84). */
gsx_status gsx_mg_apply(gsx_mg *mg, const double *r_dev, double *z_dev);
/* Level lvl (0 = coarsest): operator size/nnz and, below the finest level,
 * the nnz of the transfer to the next finer level. */
gsx_status gsx_mg_level_info(const gsx_mg *mg, int32_t lvl, int64_t *n,
                             int64_t *nnz, int64_t *transfer_nnz);
/* Host copies of the level operator (CSR) and of its transfer (CSR, rows =
 * next finer level's free DOFs); any pointer may be NULL. */
gsx_status gsx_mg_level_download(const gsx_mg *mg, int32_t lvl, int32_t *rowptr_host,
                                 int32_t *col_host, double *val_host,
                                 int32_t *prowptr_host, int32_t *pcol_host,
                                 double *pval_host);

/* ---- batched solves ------------------------------------------------------ */
/* Many independent Jacobi-preconditioned (or plain) CG solves advanced
 * together: every CG stage is one launch for all running members.  Each
 * member runs gsConjugateGradient's algorithm and stopping rule
 * (src/gsSolver/gsConjugateGradient.hpp:20-93, gsIterativeSolver.h:114-200);
 * the reference solves such problems one after another and has no batch
 * object.  A member's iterates do not depend on the other members: a solve
 * gives bitwise the same result alone and in any mix.
 * precision GSX_MIXED: FP32 iterations with FP64 reliable updates (as in
 * gsx_cg_solve), for members whose matrix was assembled with GSX_MIXED.
 * capacity: maximum number of members at one time. */
typedef struct gsx_batch gsx_batch;
gsx_status gsx_batch_create(gsx_ctx *ctx, int32_t precision, int32_t capacity,
                            gsx_batch **out);
gsx_status gsx_batch_destroy(gsx_batch *batch);
/* Add the solve A[r, r] x = b for the rows r in [row0, row0 + nrows) of A,
 * which must not couple to rows outside that range (a diagonal block, e.g.
 * one patch of a matrix assembled from disjoint patches, or the whole
 * matrix).  b_dev/x_dev hold nrows doubles; x_dev is the initial guess and
 * receives the solution.  Members may share a matrix (several right-hand
 * sides).  precond: GSX_PREC_NONE or GSX_PREC_JACOBI (damping tau).
 * Returns the member id. */
gsx_status gsx_batch_add(gsx_batch *batch, gsx_matrix *A, int32_t row0,
                         int32_t nrows, int32_t precond, double tau,
                         const double *b_dev, double *x_dev, double tol,
                         int32_t maxit, int32_t *id);
/* The same for a member whose rows are the free functions of one
 * tensor-product patch of degree p (the d-dimensional box m0 x m1 x m2 of
 * free functions in lexicographic order, m2 = 1 in 2D), e.g. a single patch
 * with Dirichlet sides or one of several disjoint patches: its CSR rows must
 * couple only within the box's (2p+1)^d stencil window (else GSX_ERR_ARG).
 * The member then iterates on a structured copy of its block (all stencil
 * slots in FP64; for GSX_MIXED the upper half (half = 1) or all slots
 * (half = 0) in FP32; no column indices), built here from A's values.
 * lanes: threads per row of its SpMV (1, 2 or 4), as gsx_batch_set_tile. */
gsx_status gsx_batch_add_box(gsx_batch *batch, gsx_matrix *A, int32_t row0,
                             int32_t nrows, int32_t d, int32_t p, int32_t m0,
                             int32_t m1, int32_t m2, int32_t lanes, int32_t half,
                             int32_t precond, double tau, const double *b_dev,
                             double *x_dev, double tol, int32_t maxit,
                             int32_t *id);
/* Tile shape of a member's SpMV, set before its first step.  CSR members:
 * lanes per row 8, 16, 32 or 64 (default from the mean row length).  Box
 * members: 1, 2 or 4 threads per row (default 1; a chunk holds 256 / lanes
 * rows and each thread sums every lanes-th stencil line), and for GSX_MIXED
 * the FP32 operator as the upper half of the stencil (half = 1, default) or
 * all slots (half = 0).  The choice changes the summation order, not the
 * operator: a member's iterates depend on its own shape only.  The project's
 * autotuner (gismo_amd/autotune.py) times the shapes per (d, p, rows) class
 * on the device and keeps the table in gismo_amd/tuning/. */
gsx_status gsx_batch_set_tile(gsx_batch *batch, int32_t id, int32_t lanes, int32_t half);
/* Initialise newly added members, run iters CG iterations of every running
 * member (members that finish stop on their own), then the reliable update
 * (mixed).  running: members still iterating afterwards.  The step's
 * launches are captured once as a HIP graph for the current set of members
 * and replayed by every following step (one graph launch per step) until a
 * member is added, finishes or is removed; with timing enabled on the
 * context (gsx_ctx_enable_timing) the launches go out one by one instead. */
gsx_status gsx_batch_step(gsx_batch *batch, int32_t iters, int32_t *running);
/* Members on one structured operator with one tile shape (box members on
 * the same single-patch matrix: several right-hand sides) share their SpMV
 * launch in groups of up to group_max (1..8; default 1, no groups): each
 * coefficient is read once for the group, and the members' inputs are
 * interleaved before the launch (8 values per row, so a coefficient's x
 * values come in two 16-byte loads).  A member's iterates do not depend on
 * its group.  Only members of at least gsx_batch_set_group_rows rows
 * (default 131072) group: below that, separate launches are faster
 * (measured 1.7x-1.8x for 8 right-hand sides of 0.27 M to 0.94 M rows, a
 * loss below about 0.1 M). */
gsx_status gsx_batch_set_group(gsx_batch *batch, int32_t group_max);
gsx_status gsx_batch_set_group_rows(gsx_batch *batch, int64_t rows);
/* Graph capture of the steps on (default) or off; graph launches so far. */
gsx_status gsx_batch_set_graph(gsx_batch *batch, int32_t enable);
gsx_status gsx_batch_graph_launches(const gsx_batch *batch, int64_t *count);
/* status: 0 running, 1 converged, 2 iteration limit, 3 zero right-hand side. */
gsx_status gsx_batch_result(const gsx_batch *batch, int32_t id, int32_t *status,
                            int32_t *iters, double *relres);
/* Relative residual history of a member (iters + 1 values; FP64 members). */
gsx_status gsx_batch_history(const gsx_batch *batch, int32_t id, double *hist_host,
                             int32_t cap, int32_t *count);
/* Release a member's slot and workspace (its x_dev keeps the solution). */
gsx_status gsx_batch_remove(gsx_batch *batch, int32_t id);

/* ---- serving ---------------------------------------------------------- */
/* Continuous-batching server of single-patch requests on one GPU (native
 * threads; no reference counterpart: the reference assembles and solves one
 * problem at a time, src/gsAssembler/gsExprAssembler.h:754-833,
 * src/gsSolver/gsConjugateGradient.hpp:20-93).  An assembler thread admits
 * requests in arrival order while their estimated device bytes fit
 * budget_bytes and assembles them; a solver thread runs every right-hand
 * side as a box member of one gsx_batch (Jacobi-CG, members join and leave
 * between steps of iters_per_step iterations); requests with at least
 * mg_min_free free DOFs whose elements coarsen dyadically to at most
 * mg_max_coarse coarse DOFs go to the multigrid threads (V(2,2), damped
 * Jacobi 0.6, hierarchy built once per request; mg_workers requests at a
 * time).  One context per thread on `device`.  A request's solutions are
 * bitwise the same alone and in any mix. */
typedef struct gsx_server gsx_server;
typedef struct {
    int32_t device;
    int32_t precision;        /* GSX_FP64 or GSX_MIXED                          */
    int64_t budget_bytes;     /* estimated device bytes of admitted requests     */
    int32_t iters_per_step;   /* batched CG iterations per step                  */
    int32_t capacity;         /* batch members (right-hand sides) at one time    */
    int64_t mg_min_free;      /* multigrid lane from this many free DOFs; 0: off */
    int64_t mg_max_coarse;    /* coarsest-level free DOFs at most (0: 4096)      */
    int64_t elem_budget;      /* multigrid lane element-matrix bytes (0: 16 GiB) */
    int32_t group_max;        /* gsx_batch_set_group of the batch lane (0: 8)    */
    int32_t mg_workers;       /* multigrid-lane threads, one context each (0: 1) */
    int64_t reserve_bytes;    /* device memory kept free beside the admitted
                                 requests (runtime scratch; 0: 16 GiB)           */
    int64_t group_rows;       /* gsx_batch_set_group_rows (0: 131072)            */
} gsx_server_opts;
/* One patch with homogeneous Dirichlet conditions on every side; nrhs
 * right-hand sides of nfree values (rhs_host NULL or rhs_host[0] NULL: the
 * assembled load vector for the first; the others must be given), copied at
 * submission; x_host (or an entry) NULL: the solution is not returned,
 * otherwise x_host[k] receives nfree values and must stay valid until the
 * request's result has been polled. */
typedef struct {
    const gsx_patch_desc *patch;
    int32_t nrhs;
    const double *const *rhs_host;
    double *const *x_host;
    double tol;
    int32_t maxit;
    /* optional (NULL), nrhs entries each, written when the request finishes:
     * iterations, relative residual and status (1, 2) per right-hand side */
    int32_t *iters_host;
    double *relres_host;
    int32_t *status_host;
} gsx_server_request;
enum { GSX_SERVER_FAILED = 3 };
typedef struct {
    int64_t rid;
    int32_t status;           /* 1 converged, 2 a right-hand side reached maxit,
                                 GSX_SERVER_FAILED (error says why)              */
    int32_t lane;             /* 0 batched CG, 1 multigrid                        */
    int32_t nrhs;
    int32_t iters_max;
    int64_t nbasis, nfree;
    double relres_max;
    double t_submit, t_admit, t_assembled, t_done;   /* server clock, seconds    */
    char error[160];
} gsx_server_result;
gsx_status gsx_server_create(const gsx_server_opts *opts, gsx_server **out);
/* Tile shape (gsx_batch_set_tile) of members of degree p in d dimensions
 * with rows in [rows_lo, rows_hi); first matching rule wins; default 1, 1. */
gsx_status gsx_server_set_tile(gsx_server *server, int32_t d, int32_t p,
                               int64_t rows_lo, int64_t rows_hi, int32_t lanes,
                               int32_t half);
gsx_status gsx_server_submit(gsx_server *server, const gsx_server_request *req,
                             int64_t *rid);
/* Up to max finished requests since the last poll (n of them in out). */
gsx_status gsx_server_poll(gsx_server *server, int32_t max, gsx_server_result *out,
                           int32_t *n);
gsx_status gsx_server_status(gsx_server *server, int64_t *outstanding,
                             int64_t *used_bytes, double *clock);
/* Message of a failure of the server itself (GSX_ERR_HIP from poll/status). */
const char *gsx_server_error(const gsx_server *server);
/* Wait until every submitted request has finished (GSX_ERR_NOT_CONVERGED on
 * timeout). */
gsx_status gsx_server_drain(gsx_server *server, double timeout_s);
/* Stops the threads; requests not finished are dropped and released. */
gsx_status gsx_server_destroy(gsx_server *server);

/* ---- reporting ---------------------------------------------------------- */
/* L2 error of the discrete solution (free coefficients, eliminated ones at
 * the problem's fixed values) against u = prod_r sin(pi x_r) with the
 * assembly quadrature
 * (gsExprEvaluator::integral, src/gsAssembler/gsExprEvaluator.h:382-450). */
gsx_status gsx_l2_error(gsx_problem *prob, const double *sol_dev,
                        double *err);
/* Both error norms of examples/poisson2_example.cpp:173-176: the L2 error
 * and the H1 seminorm error |u_h - u|_H1 (igrad(u_ex) - igrad(u_sol, G)),
 * each integrated with the assembly quadrature; either output may be NULL. */
gsx_status gsx_error_norms(gsx_problem *prob, const double *sol_dev,
                           double *l2, double *h1semi);

#ifdef __cplusplus
}
#endif
#endif
