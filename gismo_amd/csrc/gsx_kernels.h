// Kernel launchers of libgsx (defined in gsx_assembly.hip, gsx_solve.hip).
#pragma once

#include "gsx_common.h"

namespace gsx {

// sum-factorised single-patch 3D assembly (gsx_sfasm.hip): the patch's
// basis sizes, free box [flo, fhi] (m = its extent), Gauss points per
// direction g, offsets of the 1D tables, tiles of 2 x 2 rows in (y, z)
struct SfBox {
    int nb[3], nel[3], flo[3], fhi[3], m[3], g[3];
    int64_t toff[3];
    int nty, ntz;
};
bool sf_supported(int d, int p);
bool sf_mfma(int p);   // degrees with the matrix-core variant (the default row assembly)
int64_t sf_table_doubles(int p, const int *nel);
// doubles of the per-Gauss-point D (6 per point) of the matrix-core variant
// (degrees 2, 3), 0 for the others
int64_t sf_dq_doubles(int p, const int *nel);
// tables, element load vectors + det check (+ D), stencil rows (FP64, all
// slots of free rows whose column is in the patch); the tiling (nty, ntz)
// is chosen here
hipError_t launch_sf_assemble(int p, const PatchDev *patches, const GaussTab &gt, double *tab, SfBox bx,
                              double *Fbuf, double *detmin, double *Dq, double *s64, int64_t nelem, hipStream_t s);
// CSR values, FP32 half stencil (may be null), rhs; zeroes the slots of
// columns outside the free box; err |= 1 on a CSR row length mismatch
hipError_t launch_sf_finish(int p, const SfBox &bx, int nfree, const int *rowptr, double *s64, float *s32,
                            double *val, double *rhs, const double *Fbuf, const int *dofs, const double *fixed,
                            int *err, hipStream_t s);

// assembly
hipError_t launch_elem(int d, int p, const PatchDev *patches, const int *gpatch,
                       const int64_t *gprefix, int ng, int64_t nelem, const GaussTab &gt,
                       double *Kbuf, double *Fbuf, double *detmin, hipStream_t s);
// elements [e0, e1) of the group, written from Kbuf/Fbuf offset 0
hipError_t launch_elem_range(int d, int p, const PatchDev *patches, const int *gpatch,
                             const int64_t *gprefix, int ng, int64_t e0, int64_t e1, const GaussTab &gt,
                             double *Kbuf, double *Fbuf, double *detmin, hipStream_t s);
hipError_t launch_l2err(int d, int p, const PatchDev *patches, const int *gpatch,
                        const int64_t *gprefix, int ng, int64_t nelem, const GaussTab &gt,
                        const double *sol, int nfree, double *out, double *out_h1, const double *fixed,
                        hipStream_t s);
hipError_t launch_rowcount(const PatchDev *patches, const int *occ_ptr, const int *occ_patch,
                           const int64_t *occ_local, int nfree, int *counts, int *err,
                           hipStream_t s);
hipError_t launch_rowfill(const PatchDev *patches, const int *occ_ptr, const int *occ_patch,
                          const int64_t *occ_local, int nfree, const int *rowptr, int *col,
                          hipStream_t s);
hipError_t launch_scan(const int *counts, int n, int *rowptr, long long *bsum, long long *total,
                       hipStream_t s);
int gather_maxrow_bucket(int maxrow);
// boundary element terms (k_bdry): per item, the boundary-layer mass matrix
// M ((p+1)^(2(d-1)), Dirichlet items only) and load F ((p+1)^(d-1))
hipError_t launch_bdry(int d, int p, const PatchDev *patches, const BdryItem *items, int nitems,
                       const GaussTab &gt, double *M, double *F, hipStream_t s);
hipError_t launch_gather(const PatchDev *patches, const int *occ_ptr, const int *occ_patch,
                         const int64_t *occ_local, int nfree, const int *rowptr, const int *col,
                         const double *Kbuf, const double *Fbuf, double *val, double *rhs,
                         unsigned char *keep, int maxrow, const double *fixed, hipStream_t s);
hipError_t launch_rowfill_patch(int d, int p, const int *flo, const int *fhi, const int *nb,
                                const int *nel, const int *padded, int nfree, const int *f2g,
                                const int *rowptr, int *col, hipStream_t s);
// rows [r0, r1) of the free box; Kbuf/Fbuf hold the elements from global
// element id elbase on
hipError_t launch_gather_patch(int d, int p, const int *flo, const int *fhi, const int *nb,
                               const int *nel, const int *padded, int nfree, int r0, int r1, int64_t elbase,
                               const int *f2g, const int *rowptr, const double *Kbuf, const double *Fbuf,
                               double *val, double *rhs, unsigned char *keep, double *s64,
                               float *s32, const int *dofs, const double *fixed, hipStream_t s);
hipError_t launch_keep(const int *rowptr, const int *col, const double *val,
                       const unsigned char *keep, int n, int64_t nnz, int *counts, int *nrowptr,
                       int *ncol, double *nval, long long *bsum, long long *total,
                       unsigned long long *dropped, int phase, hipStream_t s);
hipError_t launch_min(const double *v, int64_t n, double *out, hipStream_t s);
hipError_t launch_sum(const double *v, int64_t n, double *out, hipStream_t s);

// solve
int spmv_lpr(double avg_row);
int spmv_blocks(int n, int lpr);
hipError_t launch_spmv_f64(int lpr, int n, const int *rp, const int *ci, const double *v,
                           const double *x, double *y, const double *b, double *part, int mode,
                           const CGState *st, hipStream_t s);
hipError_t launch_spmv_f32(int lpr, int n, const int *rp, const int *ci, const float *v,
                           const float *x, float *y, const float *b, double *part, int mode,
                           const CGState *st, hipStream_t s);
hipError_t launch_cg_update_f64(int n, double *x, double *r, const double *p, const double *t,
                                double *z, const double *diag, double *prr, double *prz,
                                const CGState *st, hipStream_t s);
hipError_t launch_cg_update_f32(int n, float *x, float *r, const float *p, const float *t,
                                float *z, const float *diag, double *prr, double *prz,
                                const CGState *st, hipStream_t s);
hipError_t launch_cg_xr_f64(int n, double *x, double *r, const double *p, const double *t,
                            double *prr, const CGState *st, hipStream_t s);
hipError_t launch_cg_pupdate_f64(int n, double *p, const double *z, const CGState *st,
                                 hipStream_t s);
hipError_t launch_cg_pupdate_f32(int n, float *p, const float *z, const CGState *st,
                                 hipStream_t s);
hipError_t launch_dot_f64(int n, const double *a, const double *b, double *part,
                          const CGState *st, hipStream_t s);
hipError_t launch_dot_f32(int n, const float *a, const float *b, double *part,
                          const CGState *st, hipStream_t s);
hipError_t launch_jacobi_apply_f64(int n, const double *r, double *z, const double *diag,
                                   double tau, double *prz, const CGState *st, hipStream_t s);
hipError_t launch_jacobi_apply_f32(int n, const float *r, float *z, const float *diag,
                                   double tau, double *prz, const CGState *st, hipStream_t s);
hipError_t launch_jacobi_step_f64(int lpr, int n, const int *rp, const int *ci, const double *v,
                                  const double *diag, const double *b, const double *xin,
                                  double *xout, double tau, int x_is_zero, hipStream_t s);
// FP32 values, FP64 vectors (the finest level of a mixed multigrid cycle)
hipError_t launch_jacobi_step_v32(int lpr, int n, const int *rp, const int *ci, const float *v,
                                  const double *diag, const double *b, const double *xin,
                                  double *xout, double tau, int x_is_zero, hipStream_t s);
hipError_t launch_residual_v32(int lpr, int n, const int *rp, const int *ci, const float *v, const double *x,
                               const double *b, double *r, hipStream_t s);
hipError_t launch_restrict_v32(int n, const int *rp, const int *ci, const float *v, const double *x, float *y,
                               hipStream_t s);
hipError_t launch_prolong_add_v32(int n, const int *rp, const int *ci, const float *v, const float *x,
                                  double *y, double alpha, hipStream_t s);
hipError_t launch_jacobi_step_f32(int lpr, int n, const int *rp, const int *ci, const float *v,
                                  const float *diag, const float *b, const float *xin,
                                  float *xout, float tau, int x_is_zero, hipStream_t s);
hipError_t launch_diag_f64(int n, const int *rp, const int *ci, const double *v, double *diag,
                           hipStream_t s);
hipError_t launch_diag_f32(int n, const int *rp, const int *ci, const float *v, float *diag,
                           hipStream_t s);
hipError_t launch_f64_to_f32(int64_t n, const double *a, float *b, hipStream_t s);
hipError_t launch_accum_correction(int n, double *x, float *xc, hipStream_t s);
hipError_t launch_axpby(int n, double a, const double *x, double b, double *y, hipStream_t s);
hipError_t launch_cg_scalar(int step, const double *p0, int np0, const double *p1, int np1,
                            CGState *st, double *hist, hipStream_t s);
hipError_t launch_cg_reliable(const double *p0, int np, CGState *st, double *hist,
                              hipStream_t s);

// multigrid (gsx_mg.hip)
hipError_t launch_transfer(int phase, int nfree_f, const int *occ_ptr, const int *occ_patch,
                           const int64_t *occ_local, const int *fine_nb, const void *t1d, const int *trp,
                           const int *tcol, const double *tval, const int *cdofs_off, const int *cdofs,
                           const int *coarse_nb, int d, int nfree_c, int *counts, int *err,
                           const int *rowptr, int *col, double *val, hipStream_t s);
hipError_t launch_transpose(int phase, int n, const int *rp, const int *ci, const double *v, int *counts,
                            const int *trp, int *cursor, int *tci, double *tv, int nt, hipStream_t s);
int ap_width(int d);
hipError_t launch_ap(int d, int n, const int *arp, const int *aci, const double *av, const int *prp,
                     const int *pci, const double *pv, int *bcnt, int *bci, double *bv, int *err,
                     hipStream_t s);
hipError_t launch_rap(int d, int nc, const int *rrp, const int *rci, const double *rv, const int *bcnt,
                      const int *bci, const double *bv, const int *crp, const int *cci, double *cval,
                      int *err, hipStream_t s);
// pattern of R B (count pass with cnt, fill pass with crp/cci); widths as launch_rap
hipError_t launch_rap_pattern(int d, int nc, const int *rrp, const int *rci, const int *bcnt, const int *bci,
                              int *cnt, const int *crp, int *cci, int *err, hipStream_t s);
// element-level Galerkin coarsening (k_coarsen): supported (d, p)
bool coarsen_supported(int d, int p);
hipError_t launch_coarsen(int d, int p, const PatchDev *cpatches, const PatchDev *fpatches, const int *gpatch,
                          const int64_t *gprefix, int ng, int64_t nelem_c, const double *T1, const int64_t *toff,
                          const double *Kf, double *Kc, hipStream_t s);
// banded Cholesky of an SPD CSR matrix of n rows, rows ordered by pos (row
// i of A is row pos[i] of the factor), half-bandwidth b in that order: Lb of
// n (b + 1) doubles; err[0] = 1 on a pivot <= 0.  Solve x = A^-1 rhs with a
// work vector of n doubles.
hipError_t launch_band_factor(int n, int b, const int *rp, const int *ci, const double *v, const int *pos,
                              double *Lb, int *err, hipStream_t s);
hipError_t launch_band_solve(int n, int b, const int *pos, const double *Lb, const double *rhs, double *x,
                             double *work, hipStream_t s);
// dense inverse of an SPD CSR matrix (n rows) into D (npad x npad, padded
// with the identity) by symmetric Gauss-Jordan sweeps; npad from
// dense_inverse_npad(n), work of dense_inverse_work_doubles(npad) doubles
int dense_inverse_npad(int n);
int64_t dense_inverse_work_doubles(int npad);
hipError_t launch_dense_inverse(int n, int npad, const int *rp, const int *ci, const double *v,
                                double *D, double *work, hipStream_t s);
hipError_t launch_dense_gemv_f64(int n, int npad, const double *M, const double *x, double *y,
                                 hipStream_t s);
hipError_t launch_prolong_add_f64(int n, const int *rp, const int *ci, const double *v,
                                  const double *x, double *y, double alpha, hipStream_t s);
hipError_t launch_prolong_add_f32(int n, const int *rp, const int *ci, const float *v,
                                  const float *x, float *y, float alpha, hipStream_t s);
hipError_t launch_restrict_f64(int n, const int *rp, const int *ci, const double *v, const double *x,
                               double *y, hipStream_t s);
hipError_t launch_restrict_f32(int n, const int *rp, const int *ci, const float *v, const float *x,
                               float *y, hipStream_t s);
hipError_t launch_residual_f64(int lpr, int n, const int *rp, const int *ci, const double *v,
                               const double *x, const double *b, double *r, hipStream_t s);
hipError_t launch_residual_f32(int lpr, int n, const int *rp, const int *ci, const float *v,
                               const float *x, const float *b, float *r, hipStream_t s);
hipError_t launch_conv_f64_f32(int64_t n, const double *a, float *b, hipStream_t s);
hipError_t launch_conv_f32_f64(int64_t n, const float *a, double *b, hipStream_t s);
hipError_t launch_smooth_update_f64(int n, double *x, double *d, const double *r, const double *diag, double c1,
                                    double c2, int first, hipStream_t s);
hipError_t launch_smooth_update_f32(int n, float *x, float *d, const float *r, const float *diag, float c1, float c2,
                                    int first, hipStream_t s);
hipError_t launch_div_sq(int n, double *w, const double *diag, double *part, hipStream_t s);

// structured operator + grid-layout CG (gsx_stencil.hip)
int sspmv_blocks(int nfree);
int vec_blocks(int64_t n);
hipError_t launch_fill_f64(double *a, int64_t n, double v, hipStream_t s);
hipError_t launch_fill_f32(float *a, int64_t n, float v, hipStream_t s);
hipError_t launch_csr_to_stencil(int d, int p, int nfree, const int *rp, const int *ci,
                                 const double *v, const int *f2g, int nbx, int nby, double *s64,
                                 float *s32, hipStream_t s);
// centre slot of the structured operator (ns slots) onto the grid
hipError_t launch_stencil_diag(int nfree, int ns, const int *f2g, const double *s64, double *diag_grid,
                              hipStream_t s);
hipError_t launch_scatter_f64_to_grid(int nfree, const int *f2g, const double *src, double *dst,
                                      hipStream_t s);
hipError_t launch_gather_f64_from_grid(int nfree, const int *f2g, const double *src, double *dst,
                                       hipStream_t s);
// MODE 1 with st: skipped once the solve stopped or paused at iteration k
hipError_t launch_sspmv_f64(int d, int p, int mode, int nfree, const int *f2g, int sy, int sz,
                            const double *sv, const double *x, double *y, const double *b,
                            double *part, CGState *st, int k, hipStream_t s);
// FP32 operator in the half-stored layout (stencil_half); m0 x m1 x m2 = the
// free box (rows lexicographic), m0 m1 m2 == nfree.
hipError_t launch_sspmv_f32(int d, int p, int mode, int nfree, const int *f2g, int sy, int sz, int m0, int m1,
                            int m2, const float *sv, const float *x, float *y, const float *b,
                            double *part, CGState *st, int k, hipStream_t s);
hipError_t launch_cg2_update_f64(int n, double *x, double *r, const double *p, const double *t,
                                 double *z, const double *diag, double tau, const double *pap,
                                 int npap, double *prr, double *prz, CGState *st, int k, hipStream_t s);
hipError_t launch_cg2_update_f32(int n, float *x, float *r, const float *p, const float *t, float *z,
                                 const float *diag, double tau, const double *pap, int npap,
                                 double *prr, double *prz, CGState *st, int k, hipStream_t s);
hipError_t launch_cg2_direction_f64(int n, double *p, const double *z, const double *prr,
                                    const double *prz, int np, CGState *st, int k, double *hist,
                                    hipStream_t s);
hipError_t launch_cg2_direction_f32_mixed(int n, float *p, const float *z, const double *prr,
                                          const double *prz, int np, CGState *st, int k,
                                          hipStream_t s);
// Gauss-Seidel: update the cnt rows listed in rows (one level set)
hipError_t launch_gs_rows_f64(int cnt, const int *rows, const int *rp, const int *ci, const double *va,
                              const double *b, double *x, hipStream_t s);
hipError_t launch_gs_rows_v32(int cnt, const int *rows, const int *rp, const int *ci, const float *va,
                              const double *b, double *x, hipStream_t s);
hipError_t launch_gs_rows_f32(int cnt, const int *rows, const int *rp, const int *ci, const float *va,
                              const float *b, float *x, hipStream_t s);
hipError_t launch_cg2_dot_f64(int n, const double *a, const double *b, double *part, const CGState *st,
                              hipStream_t s);
hipError_t launch_cg2_dot_f32(int n, const float *a, const float *b, double *part, const CGState *st,
                              hipStream_t s);
// fast diagonalization: one mode product y_a = sum_i M[i*m + a] x_i along
// direction dir of the m0 x m1 x m2 box (element (a,b,c) of x at
// xo + a + b xs1 + c xs2, of y at yo + a + b ys1 + c ys2); with l0 != null
// the result is divided by alpha + l0[a] + l1[b] (+ l2[c])
hipError_t launch_fd_mode_f64(int m0, int m1, int m2, int dir, const double *M, const double *x, int64_t xo,
                              int64_t xs1, int64_t xs2, double *y, int64_t yo, int64_t ys1, int64_t ys2,
                              const double *l0, const double *l1, const double *l2, double alpha,
                              const CGState *st, hipStream_t s);
hipError_t launch_fd_mode_f32(int m0, int m1, int m2, int dir, const float *M, const float *x, int64_t xo,
                              int64_t xs1, int64_t xs2, float *y, int64_t yo, int64_t ys1, int64_t ys2,
                              const double *l0, const double *l1, const double *l2, double alpha,
                              const CGState *st, hipStream_t s);
// rectangular mode product (output box o0 x o1 x o2, m_in inputs along
// dir), optional scaling by 1/(alpha + l0[a] + l1[b] + l2[c]) (null: 0)
// and accumulation into y
hipError_t launch_tensor_mode_f64(int o0, int o1, int o2, int dir, int m_in, const double *M, const double *x,
                                  int64_t xo, int64_t xs1, int64_t xs2, double *y, int64_t yo, int64_t ys1,
                                  int64_t ys2, int scale, double alpha, const double *l0, const double *l1,
                                  const double *l2, int accumulate, const CGState *st, hipStream_t s);
hipError_t launch_tensor_mode_f32(int o0, int o1, int o2, int dir, int m_in, const float *M, const float *x,
                                  int64_t xo, int64_t xs1, int64_t xs2, float *y, int64_t yo, int64_t ys1,
                                  int64_t ys2, int scale, double alpha, const double *l0, const double *l1,
                                  const double *l2, int accumulate, const CGState *st, hipStream_t s);
hipError_t launch_cg2_precond_f64(int n, const double *r, double *z, const double *diag,
                                  double tau, double *prz, hipStream_t s);
hipError_t launch_cg2_precond_f32(int n, const float *r, float *z, const float *diag, double tau,
                                  double *prz, hipStream_t s);
hipError_t launch_set_rz(const double *parts, int np, CGState *st, int slot, hipStream_t s);
hipError_t launch_grid_f64_to_f32(int64_t n, const double *a, float *b, hipStream_t s);
hipError_t launch_accum_grid(int64_t n, double *x, float *xc, hipStream_t s);

// batched CG (gsx_batch.hip)
int batch_chunk_rows(int lpr);
hipError_t launch_batch_check(int row0, int nrows, const int *rp, const int *ci, int *flag, hipStream_t s);
int64_t batch_box_slots(int d, int p, bool half);
hipError_t launch_batch_box_build(int d, int p, int row0, int n, int m0, int m1, bool half, const int *rp,
                                  const int *ci, const double *val, double *s64, float *s32, int *flag,
                                  hipStream_t s);
hipError_t launch_batch_box_s32(int d, int p, int n, bool half, const double *s64, float *s32, hipStream_t s);
hipError_t launch_batch_init(int nchunks, const BChunk *chunks, const BSeg *segs, BState *st, double *p0,
                             double *p1, double *p2, int mixed, hipStream_t s);
hipError_t launch_batch_iteration(int nchunks, const BChunk *chunks, int nsingle, int nmulti,
                                  const BSChunk *schunks, const BGroup *groups, const BSeg *segs, BState *st,
                                  double *p0, double *p1, double *p2, int mixed, double delta2, hipStream_t s);
hipError_t launch_batch_reliable(int nchunks, const BChunk *chunks, const BSeg *segs, BState *st, double *p0,
                                 double *p2, hipStream_t s);

// assembly with the geometry of another problem (gsx_pmg.hip): elements
// [e0, e1) of patch ip of the discretisation patches ups, geometry patch ip
// of gps; K (may be null) at kmat_off + e NB^2 - kshift, f (may be null) at
// fvec_off + e NB - fshift with f = 1 when unit_f (the lumped mass)
hipError_t launch_geo_elem(int d, const PatchDev *ups, const PatchDev *gps, int ip, int64_t e0, int64_t e1,
                           const GaussTab &gt, double *Kbuf, double *Fbuf, double *detmin, int64_t kshift,
                           int64_t fshift, int unit_f, hipStream_t s);
// mixed mass element matrices (U elements x U functions x V functions) of patch ip at moff
hipError_t launch_geo_mixed(int d, const PatchDev *ups, const PatchDev *vps, const PatchDev *gps, int ip,
                            int64_t nelem, const GaussTab &gt, double *Mbuf, int64_t moff, hipStream_t s);
// mixed mass pattern (rows: U's free DOFs, columns: V's free DOFs): phase 0
// counts candidates (counts), 1 writes them at cptr, sorts and dedups
// (counts = kept), 2 compacts into (rowptr, col)
hipError_t launch_mix_pattern(int phase, int d, const PatchDev *ups, const PatchDev *vps, const int *occ_ptr,
                              const int *occ_patch, const int64_t *occ_local, int nrows, int nfree_v,
                              const int *cptr, int *cand, int *counts, const int *rowptr, int *col, hipStream_t s);
// values of the mixed mass: one wavefront per row when every row has at
// most 1024 entries (maxrow), else one thread per row (the same sums)
hipError_t launch_mix_gather(int d, const PatchDev *ups, const PatchDev *vps, const int *occ_ptr, const int *occ_patch,
                             const int64_t *occ_local, int nrows, int nfree_v, const int *rowptr, const int *col,
                             const double *Mbuf, const int64_t *moff, double *val, int maxrow, hipStream_t s);
// per free DOF: sum of the element vector entries of its functions
hipError_t launch_vec_gather(int d, const PatchDev *ups, const int *occ_ptr, const int *occ_patch,
                             const int64_t *occ_local, int nrows, const double *Fbuf, double *out, hipStream_t s);
// out[i] = in[idx[i]]; x[idx[i]] += scale z[i] (distinct indices)
hipError_t launch_gather_idx(int n, const int *idx, const double *in, double *out, hipStream_t s);
hipError_t launch_scatter_add_idx(int n, const int *idx, const double *z, double scale, double *x, hipStream_t s);
// val(row r) *= 1 / L(r); err[0] = 1 on L(r) == 0
hipError_t launch_scale_rows(int nrows, const int *rowptr, const double *L, double *val, int *err, hipStream_t s);

}  // namespace gsx
