/*
 * gsx_oracle.h -- CPU FP64 restatement of G+Smo's isogeometric Poisson
 * assemble + solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP product path (libgsx).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; nothing in the product links it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the G+Smo source tree).  Plain C99, single-threaded, so
 * the accumulation order equals the reference's single-threaded order
 * (patch by patch, elements lexicographic with direction 0 fastest).
 *
 * Pinning (tests/test_oracle.py, tests/test_lib_cpu.py):
 *   oc_gauss_legendre   -- reference 30-digit tables (golden fixture), exactness
 *                          on monomials, numpy leggauss.
 *   oc_knots_uniform,
 *   oc_find_span        -- span search cases of the reference's knot vector
 *                          tests, uniform spacing by construction.
 *   oc_bspline_ders     -- brute-force Cox-de Boor in exact rationals,
 *                          partition of unity, sum of derivatives = 0.
 *   oc_element_geometry -- scipy BSpline evaluation of the same patch.
 *   oc_mapper_build     -- hand-counted small multipatch cases, closed-form
 *                          free-DOF counts of the glued grids, glued p=1
 *                          grid = single patch.
 *   oc_assemble         -- Kronecker closed form of affine patches (matrix and
 *                          load), isoparametric identities (constant in the
 *                          kernel, energy identity, cross term, load sum),
 *                          symmetry, EXPR pattern = FULL pattern.
 *   oc_assemble_rows    -- rows of oc_assemble.
 *   oc_csr_spmv         -- inside oc_cg_solve, pinned by its checks.
 *   oc_cg_solve         -- 1D FD Poisson with known solution, residual
 *                          history against a numpy CG on tiny systems.
 *   oc_transfer_1d,
 *   oc_transfer_multipatch -- refinement identity (coarse spline = fine
 *                          spline at sample points), binomial subdivision masks.
 *   oc_mg_*             -- Galerkin identity P^T A_f P = A_c (nested spaces),
 *                          two-grid cycle against numpy, MG-PCG convergence.
 *   oc_l2_error         -- convergence order p+1 on the manufactured solution,
 *                          ||u||_L2 of the zero solution in closed form.
 *   oc_h1_error         -- convergence order p, |u|_H1 of the zero solution
 *                          in closed form (pi/sqrt(2) in 2D, pi sqrt(3/8) in 3D).
 */
#ifndef GSX_ORACLE_H
#define GSX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OC_MAXP 8

/* One tensor-product B-spline patch (geometry and discretisation share the
 * basis: isoparametric).  Control points are row-major nbasis x d in
 * lexicographic order with direction 0 fastest (gsTensorBasis numbering,
 * gsTensorBasis.hpp:203-236). */
typedef struct {
    int d;                  /* parametric = physical dimension, 2 or 3 */
    int p[3];               /* degree per direction */
    int nknots[3];          /* knot vector length per direction */
    const double *knots[3]; /* open knot vectors */
    const double *ctrl;     /* control points */
    const double *kappa;    /* per-element coefficient (NULL: 1) */
    const double *felem;    /* per-element constant source (fkind 0) */
    int fkind;              /* 0: felem; 1: d*pi^2*prod sin(pi x_r) */
} oc_patch;

typedef struct {
    int n;          /* rows = cols = free DOFs */
    int64_t nnz;
    int *rowptr;    /* n+1 */
    int *col;       /* nnz, sorted per row */
    double *val;    /* nnz */
} oc_csr;

/* --- quadrature, knots, basis ------------------------------------------ */
void oc_gauss_legendre(int n, double *x, double *w);
void oc_knots_uniform(int nelem, int p, double a, double b, double *knots);
int  oc_find_span(const double *knots, int nknots, int p, double u);
void oc_bspline_ders(const double *knots, int p, int span, double u,
                     int nder, double *out);
int  oc_patch_nbasis(const oc_patch *P, int dir);
int  oc_patch_nelem(const oc_patch *P, int dir);

/* --- geometry at the Gauss points of one element ------------------------ */
/* Returns physical points (nq x d), Jacobians (nq x d x d, J[r][a] =
 * dx_r/dxi_a), measures |det J| (nq) and weights (nq). */
int  oc_element_geometry(const oc_patch *P, int elem, double *xq,
                         double *jac, double *meas, double *wq);

/* --- DOF mapper --------------------------------------------------------- */
/* dir_sides: ndir pairs (patch, side) with side 1..2d (west,east,south,north,
 * front,back).  ifaces: nif records of 8 ints (p1,s1,p2,s2,m0,f0,m1,f1):
 * for the j-th tangential direction of patch p1 (increasing order), mj is the
 * matching direction of patch p2 and fj=1 if its orientation is reversed.
 * dofs_out: sum of patch basis sizes; global index per (patch, local).
 * Returns the number of free DOFs, or -1 on error. */
int  oc_mapper_build(int npatch, const oc_patch *patches, int ndir,
                     const int *dir_sides, int nif, const int *ifaces,
                     int *dofs_out);

/* --- assembly ----------------------------------------------------------- */
/* mode 0: gsExprAssembler semantics (exact-zero local entries skipped);
 * mode 1: gsPoissonAssembler semantics (every free-free pair inserted).
 * rhs_out: nfree.  Returns a malloc'd CSR (free with oc_csr_free). */
oc_csr *oc_assemble(int npatch, const oc_patch *patches, const int *dofs,
                    int nfree, int mode, double *rhs_out);
void oc_csr_free(oc_csr *A);
int  oc_assemble_rows(const oc_patch *P, const int *dofs, int nfree, int nrows,
                      const int *rows, int maxc, int *counts, int *cols,
                      double *vals, double *rhs);
void oc_csr_spmv(const oc_csr *A, const double *x, double *y);

/* --- solvers ------------------------------------------------------------ */
typedef struct oc_mg oc_mg;

/* precond: 0 none, 1 Jacobi(tau), 2 multigrid (mg != NULL), 3 Jacobi with
 * Eigen's rule for zero diagonal entries (1 there).
 * x: in/out (initial guess).  hist (optional, maxit+1): error history.
 * Returns iterations; *err_out = final ||r||/||b||. */
int  oc_cg_solve(const oc_csr *A, const double *b, double *x, double tol,
                 int maxit, int precond, double tau, const oc_mg *mg,
                 double *err_out, double *hist);

/* 1D uniform-refinement transfer (midpoint insertion) for an open knot
 * vector: fine (nf x nc) dense, row-major.  Returns nf. */
int  oc_transfer_1d(const double *knots, int nknots, int p, double *T);

/* Multigrid over nlev levels; A_fine and transfers P_l (fine_l x coarse_l,
 * CSR) for l = 0..nlev-2 (P_l maps level l to level l+1).  Galerkin coarse
 * operators are formed inside. */
oc_mg *oc_mg_create(const oc_csr *A_fine, int nlev, const oc_csr *const *P,
                    int npre, int npost, double damping, double cgc_damping);
void oc_mg_apply(const oc_mg *mg, const double *r, double *z);
/* smoother: 0 Jacobi (default), 1 Richardson, 2 Chebyshev with interval
 * [lmax / param, lmax]; lmax per level from 20 power iterations on D^-1 A;
 * 3 Gauss-Seidel (forward sweeps before, reverse sweeps after the coarse
 * correction). */
void oc_mg_set_smoother(oc_mg *mg, int kind, double param);
/* one Gauss-Seidel sweep (reverse != 0: rows descending), x updated in place */
void oc_gs_sweep(const oc_csr *A, const double *b, double *x, int reverse);
double oc_mg_lmax(const oc_mg *mg, int lvl);
const oc_csr *oc_mg_matrix(const oc_mg *mg, int lvl);
void oc_mg_free(oc_mg *mg);

/* Transfer between two mapped multipatch spaces (coarse -> fine) following
 * gsMultiBasis::combineTransferMatrices; patches of the fine space are the
 * coarse ones uniformly refined once in every direction. */
oc_csr *oc_transfer_multipatch(int npatch, const oc_patch *coarse,
                               const int *dofs_c, int nfree_c,
                               const int *dofs_f, int nfree_f);

/* --- boundary data ------------------------------------------------------- */
/* type 1 Dirichlet, 2 Neumann; kind 0 zero, 1 manufactured (g = prod
 * sin(pi x_r), Neumann flux grad g), 2 affine (g = value[0] + value[1+r]
 * x_r, flux (value[1], value[2], value[3])). */
typedef struct { int patch, side, type, kind; double value[4]; } oc_bc;
/* Fixed values (nfix = eliminated DOFs) by the L2 projection over the
 * Dirichlet records (every Dirichlet side, homogeneous ones with kind 0). */
int  oc_dirichlet_l2(int npatch, const oc_patch *patches, const int *dofs, int nfree, int nfix,
                     int nbc, const oc_bc *bcs, double *fixed_out);
/* rhs (nfree) += Neumann boundary load of the Neumann records. */
void oc_neumann(int npatch, const oc_patch *patches, const int *dofs, int nfree, int nbc,
                const oc_bc *bcs, double *rhs);
/* oc_assemble with the fixed values moved to the rhs (fixed NULL: zero). */
oc_csr *oc_assemble_bc(int npatch, const oc_patch *patches, const int *dofs,
                       int nfree, int mode, const double *fixed, double *rhs_out);
/* The same assembly on a team of nthreads OpenMP threads, element-strided
 * with the push under a lock, as the reference threads it
 * (gsExprAssembler.h:761-796).  CPU baseline timing only; fixed values are
 * not applied.  Built with OpenMP only in libgsx_cpuref.so. */
oc_csr *oc_assemble_par(int npatch, const oc_patch *patches, const int *dofs, int nfree, int mode,
                        double *rhs_out, int nthreads);

/* --- reporting ---------------------------------------------------------- */
/* L2 error of the discrete solution (free coefficients, eliminated = 0)
 * against u = prod sin(pi x_r), with the assembly quadrature. */
double oc_l2_error(int npatch, const oc_patch *patches, const int *dofs,
                   int nfree, const double *sol);
/* H1 seminorm error against the same u (examples/poisson2_example.cpp:173-176). */
double oc_h1_error(int npatch, const oc_patch *patches, const int *dofs,
                   int nfree, const double *sol);
/* both with the eliminated coefficients at `fixed` (NULL: zero) */
double oc_l2_error_bc(int npatch, const oc_patch *patches, const int *dofs,
                      int nfree, const double *sol, const double *fixed);
double oc_h1_error_bc(int npatch, const oc_patch *patches, const int *dofs,
                      int nfree, const double *sol, const double *fixed);

#ifdef __cplusplus
}
#endif
#endif
