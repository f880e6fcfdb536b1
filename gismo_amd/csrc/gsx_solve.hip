// Solver kernels of libgsx (gfx950): CSR SpMV, the vector steps of
// preconditioned CG with device-resident scalars, Jacobi smoothing and
// mixed-precision helpers.
//
// Reference path (G+Smo source tree):
//   gsMatrixOp::apply            src/gsSolver/gsMatrixOp.h:74-75
//   gsConjugateGradient          src/gsSolver/gsConjugateGradient.hpp:20-93
//   gsIterativeSolver::solve     src/gsSolver/gsIterativeSolver.h:114-200
//   gsJacobiOp::apply/step       src/gsSolver/gsSimplePreconditioners.h:178-202
//
// Every reduction is two-stage (per-block partials, then one block summing
// them in a fixed order), so results are bitwise reproducible run to run.
#include "gsx_common.h"
#include "gsx_kernels.h"

namespace gsx {

template <int LPR, typename T>
__device__ inline T group_sum(T v)
{
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LPR);
    return v;
}

__device__ inline double block_sum256(double v, double *s)
{
    // wavefront shuffles, then the 4 wave sums in wave order (fixed order)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// SpMV, LPR lanes per row, 256 threads per block.
//   MODE 0: y = A x
//   MODE 1: y = A x and partial[block] = sum_rows x_i y_i
//   MODE 2: y = b - A x and partial[block] = sum_rows y_i^2
// Skipped entirely when state->done != 0 (state may be null).
// ---------------------------------------------------------------------------
template <int LPR, typename V, typename T, int MODE>
__global__ __launch_bounds__(256) void k_spmv(int n, const int *__restrict__ rowptr,
                                              const int *__restrict__ col,
                                              const V *__restrict__ val, const T *__restrict__ x,
                                              T *__restrict__ y, const T *__restrict__ b,
                                              double *__restrict__ partial,
                                              const CGState *__restrict__ state)
{
    __shared__ double s[256];
    if (state && (state->done || state->frozen)) return;
    constexpr int RPB = 256 / LPR;
    const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    const int r = blockIdx.x * RPB + g;
    T acc = 0;
    if (r < n) {
        const int k0 = rowptr[r], k1 = rowptr[r + 1];
        for (int k = k0 + lane; k < k1; k += LPR) acc += (T)val[k] * x[col[k]];
    }
    acc = group_sum<LPR, T>(acc);
    double contrib = 0.0;
    if (r < n && lane == 0) {
        if (MODE == 2) {
            const T v = b[r] - acc;
            y[r] = v;
            contrib = (double)v * (double)v;
        } else {
            y[r] = acc;
            if (MODE == 1) contrib = (double)x[r] * (double)acc;
        }
    }
    if (MODE != 0) {
        const double tot = block_sum256(contrib, s);
        if (threadIdx.x == 0) partial[blockIdx.x] = tot;
    }
}

// ---------------------------------------------------------------------------
// CG vector updates.
// ---------------------------------------------------------------------------
// x += alpha p; r -= alpha t; z = tau r / diag; partial rr, rz
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_cg_update(int n, TX *__restrict__ x, T *__restrict__ r,
                                                   const T *__restrict__ p, const T *__restrict__ t,
                                                   T *__restrict__ z, const T *__restrict__ diag,
                                                   double *__restrict__ part_rr,
                                                   double *__restrict__ part_rz,
                                                   const CGState *__restrict__ st)
{
    __shared__ double s[256];
    if (st->done || st->frozen) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < n) {
        const T alpha = (T)st->alpha;
        x[i] += (TX)(alpha * p[i]);
        const T ri = r[i] - alpha * t[i];
        r[i] = ri;
        const T zi = (T)st->tau * ri / diag[i];
        z[i] = zi;
        rr = (double)ri * (double)ri;
        rz = (double)ri * (double)zi;
    }
    const double a = block_sum256(rr, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = a;
    __syncthreads();
    const double c = block_sum256(rz, s);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = c;
}

// un-preconditioned variant: z = r
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_cg_update_id(int n, TX *__restrict__ x, T *__restrict__ r,
                                                      const T *__restrict__ p, const T *__restrict__ t,
                                                      double *__restrict__ part_rr,
                                                      const CGState *__restrict__ st)
{
    __shared__ double s[256];
    if (st->done || st->frozen) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        const T alpha = (T)st->alpha;
        x[i] += (TX)(alpha * p[i]);
        const T ri = r[i] - alpha * t[i];
        r[i] = ri;
        rr = (double)ri * (double)ri;
    }
    const double a = block_sum256(rr, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = a;
}

// x += alpha p; r -= alpha t; partial rr (preconditioner applied separately)
template <typename T>
__global__ __launch_bounds__(256) void k_cg_xr(int n, T *__restrict__ x, T *__restrict__ r,
                                               const T *__restrict__ p, const T *__restrict__ t,
                                               double *__restrict__ part_rr,
                                               const CGState *__restrict__ st)
{
    __shared__ double s[256];
    if (st->done) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        const T alpha = (T)st->alpha;
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * t[i];
        r[i] = ri;
        rr = (double)ri * (double)ri;
    }
    const double a = block_sum256(rr, s);
    if (threadIdx.x == 0) part_rr[blockIdx.x] = a;
}

// p = z + beta p
template <typename T>
__global__ __launch_bounds__(256) void k_cg_pupdate(int n, T *__restrict__ p, const T *__restrict__ z,
                                                    const CGState *__restrict__ st)
{
    if (st->done || st->frozen) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = z[i] + (T)st->beta * p[i];
}

// partial dot(a, b)
template <typename T>
__global__ __launch_bounds__(256) void k_dot(int n, const T *__restrict__ a, const T *__restrict__ b,
                                             double *__restrict__ part, const CGState *__restrict__ st)
{
    __shared__ double s[256];
    if (st && st->done) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const double v = (i < n) ? (double)a[i] * (double)b[i] : 0.0;
    const double tot = block_sum256(v, s);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// z = tau r / diag, partial rz
template <typename T>
__global__ __launch_bounds__(256) void k_jacobi_apply(int n, const T *__restrict__ r, T *__restrict__ z,
                                                      const T *__restrict__ diag, double tau,
                                                      double *__restrict__ part_rz,
                                                      const CGState *__restrict__ st)
{
    __shared__ double s[256];
    if (st && st->done) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const T zi = (T)tau * r[i] / diag[i];
        z[i] = zi;
        v = (double)r[i] * (double)zi;
    }
    if (part_rz) {
        const double tot = block_sum256(v, s);
        if (threadIdx.x == 0) part_rz[blockIdx.x] = tot;
    }
}

// Damped Jacobi smoothing step x += tau (b - A x) / diag, fused with the
// SpMV (gsJacobiOp::step).  Reads x_in, writes x_out (x_in != x_out).
template <int LPR, typename V, typename T>
__global__ __launch_bounds__(256) void k_jacobi_step(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ col,
                                                     const V *__restrict__ val,
                                                     const T *__restrict__ diag,
                                                     const T *__restrict__ b, const T *__restrict__ xin,
                                                     T *__restrict__ xout, T tau, int x_is_zero)
{
    // products and sums in FP64 whatever the storage type: an FP32 row sum
    // of a stiffness row cancels to a relative error of eps * cond(A), which
    // costs a multigrid preconditioner outer iterations
    constexpr int RPB = 256 / LPR;
    const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    const int r = blockIdx.x * RPB + g;
    double acc = 0.0;
    if (r < n && !x_is_zero) {
        const int k0 = rowptr[r], k1 = rowptr[r + 1];
        for (int k = k0 + lane; k < k1; k += LPR) acc += (double)val[k] * (double)xin[col[k]];
    }
    acc = group_sum<LPR, double>(acc);
    if (r < n && lane == 0) {
        const double xo = x_is_zero ? 0.0 : (double)xin[r];
        xout[r] = (T)(xo + (double)tau * ((double)b[r] - acc) / (double)diag[r]);
    }
}

// extract diagonal (value type V -> vector type T)
template <typename V, typename T>
__global__ void k_diag(int n, const int *__restrict__ rowptr, const int *__restrict__ col,
                       const V *__restrict__ val, T *__restrict__ diag)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    T d = 0;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
        if (col[k] == r) { d = (T)val[k]; break; }
    diag[r] = d;
}

template <typename S, typename D>
__global__ void k_convert(int64_t n, const S *__restrict__ s, D *__restrict__ d)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (D)s[i];
}

// x64 += x32; x32 = 0
__global__ void k_accum_correction(int n, double *__restrict__ x, float *__restrict__ xc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] += (double)xc[i];
        xc[i] = 0.0f;
    }
}

__global__ void k_axpby(int n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}

// ---------------------------------------------------------------------------
// Scalar steps (one block of 256): sum partials in a fixed order, update the
// CG state.
// ---------------------------------------------------------------------------
__device__ inline double sum_parts(const double *__restrict__ p, int np, double *s)
{
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) a += p[i];
    return block_sum256(a, s);
}

// step 0: init -- p0 = ||b||^2 parts, p1 = ||r0||^2 parts (r0 = b - A x0)
// step 1: rz init -- p0 = (r,z) parts
// step 2: alpha -- p0 = (p, A p) parts
// step 3: beta -- p0 = ||r||^2 parts, p1 = (r,z) parts; counts the iteration
// step 4: beta (mixed, FP32 residual) -- like 3, records convergence
//         candidate instead of convergence
__global__ __launch_bounds__(256) void k_cg_scalar(int step, const double *__restrict__ p0,
                                                   int np0, const double *__restrict__ p1, int np1,
                                                   CGState *__restrict__ st,
                                                   double *__restrict__ hist)
{
    __shared__ double s[256];
    if (step != 0 && st->done) return;
    const double a = sum_parts(p0, np0, s);
    __syncthreads();
    const double b = p1 ? sum_parts(p1, np1, s) : 0.0;
    if (threadIdx.x != 0) return;
    switch (step) {
    case 0: {
        st->bnorm = sqrt(a);
        st->it = 0;
        st->done = 0;
        if (st->bnorm == 0.0) {
            st->err = 0.0;
            st->done = 3; // zero rhs: x = 0 is the solution
        } else {
            st->rr = b;
            st->err = sqrt(b) / st->bnorm;
            if (st->err < st->tol) st->done = 1;
        }
        st->rr_ref = b;
        if (hist) hist[0] = st->err;
        break;
    }
    case 1:
        st->rz = a;
        break;
    case 2:
        // mixed: an update is due since the last iteration -> wait for it
        st->frozen = st->need_update;
        if (st->frozen) break;
        st->pap = a;
        st->alpha = st->rz / a;
        break;
    case 3: {
        st->it += 1;
        st->rr = a;
        st->err = sqrt(a) / st->bnorm;
        if (hist) hist[st->it] = st->err;
        if (st->err < st->tol) {
            st->done = 1;
        } else {
            st->rz_new = b;
            st->beta = b / st->rz;
            st->rz = b;
            if (st->it >= st->maxit) st->done = 2;
        }
        break;
    }
    case 4: {
        if (st->frozen) break;
        st->it += 1;
        st->rr = a;
        st->err = sqrt(a) / st->bnorm;
        st->rz_new = b;
        st->beta = b / st->rz;
        st->rz = b;
        // reliable update due (done by the host at the end of the chunk):
        // the residual dropped below delta^2 of the reference, or the
        // recurrence claims convergence.  The iteration itself completes.
        if (a < st->rel_since_update * st->rr_ref || st->err < st->tol) st->need_update = 1;
        if (st->it >= st->maxit) st->done = 2;
        break;
    }
    case 5: {
        // residual check of a preconditioned iteration whose preconditioner
        // is applied afterwards by the host (multigrid)
        st->it += 1;
        st->rr = a;
        st->err = sqrt(a) / st->bnorm;
        if (hist) hist[st->it] = st->err;
        if (st->err < st->tol) st->done = 1;
        else if (st->it >= st->maxit) st->done = 2;
        break;
    }
    case 6:
        st->beta = a / st->rz;
        st->rz = a;
        break;
    default:
        break;
    }
}

// After a reliable update: p5 = ||r64||^2 parts; set err, decide.
__global__ __launch_bounds__(256) void k_cg_reliable(const double *__restrict__ p0, int np,
                                                     CGState *__restrict__ st,
                                                     double *__restrict__ hist)
{
    __shared__ double s[256];
    const double a = sum_parts(p0, np, s);
    if (threadIdx.x != 0) return;
    st->rr = a;
    st->rr_ref = a;
    st->err = sqrt(a) / st->bnorm;
    if (hist) hist[st->it] = st->err;
    st->need_update = 0;
    st->flag_next = 0;
    st->frozen = 0;
    st->done = (st->err < st->tol) ? 1 : (st->it >= st->maxit ? 2 : 0);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int spmv_lpr(double avg_row)
{
    if (avg_row > 96) return 64;
    if (avg_row > 40) return 32;
    if (avg_row > 16) return 16;
    return 8;
}

int spmv_blocks(int n, int lpr) { return (n + 256 / lpr - 1) / (256 / lpr); }

template <typename V, typename T, int MODE>
static hipError_t spmv_dispatch(int lpr, int n, const int *rowptr, const int *col, const V *val,
                                const T *x, T *y, const T *b, double *partial,
                                const CGState *st, hipStream_t s)
{
    const dim3 grid(spmv_blocks(n, lpr)), block(256);
    switch (lpr) {
    case 64: hipLaunchKernelGGL((k_spmv<64, V, T, MODE>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, st); break;
    case 32: hipLaunchKernelGGL((k_spmv<32, V, T, MODE>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, st); break;
    case 16: hipLaunchKernelGGL((k_spmv<16, V, T, MODE>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, st); break;
    case 8: hipLaunchKernelGGL((k_spmv<8, V, T, MODE>), grid, block, 0, s, n, rowptr, col, val, x, y, b, partial, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_spmv_f64(int lpr, int n, const int *rp, const int *ci, const double *v,
                           const double *x, double *y, const double *b, double *part, int mode,
                           const CGState *st, hipStream_t s)
{
    if (mode == 0) return spmv_dispatch<double, double, 0>(lpr, n, rp, ci, v, x, y, b, part, st, s);
    if (mode == 1) return spmv_dispatch<double, double, 1>(lpr, n, rp, ci, v, x, y, b, part, st, s);
    return spmv_dispatch<double, double, 2>(lpr, n, rp, ci, v, x, y, b, part, st, s);
}

hipError_t launch_spmv_f32(int lpr, int n, const int *rp, const int *ci, const float *v,
                           const float *x, float *y, const float *b, double *part, int mode,
                           const CGState *st, hipStream_t s)
{
    if (mode == 0) return spmv_dispatch<float, float, 0>(lpr, n, rp, ci, v, x, y, b, part, st, s);
    if (mode == 1) return spmv_dispatch<float, float, 1>(lpr, n, rp, ci, v, x, y, b, part, st, s);
    return spmv_dispatch<float, float, 2>(lpr, n, rp, ci, v, x, y, b, part, st, s);
}

static inline dim3 vgrid(int n) { return dim3((n + 255) / 256); }

hipError_t launch_cg_update_f64(int n, double *x, double *r, const double *p, const double *t,
                                double *z, const double *diag, double *prr, double *prz,
                                const CGState *st, hipStream_t s)
{
    if (diag)
        hipLaunchKernelGGL((k_cg_update<double, double>), vgrid(n), dim3(256), 0, s, n, x, r, p, t, z, diag, prr, prz, st);
    else
        hipLaunchKernelGGL((k_cg_update_id<double, double>), vgrid(n), dim3(256), 0, s, n, x, r, p, t, prr, st);
    return hipGetLastError();
}

hipError_t launch_cg_update_f32(int n, float *x, float *r, const float *p, const float *t,
                                float *z, const float *diag, double *prr, double *prz,
                                const CGState *st, hipStream_t s)
{
    if (diag)
        hipLaunchKernelGGL((k_cg_update<float, float>), vgrid(n), dim3(256), 0, s, n, x, r, p, t, z, diag, prr, prz, st);
    else
        hipLaunchKernelGGL((k_cg_update_id<float, float>), vgrid(n), dim3(256), 0, s, n, x, r, p, t, prr, st);
    return hipGetLastError();
}

hipError_t launch_cg_xr_f64(int n, double *x, double *r, const double *p, const double *t,
                            double *prr, const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_xr<double>), vgrid(n), dim3(256), 0, s, n, x, r, p, t, prr, st);
    return hipGetLastError();
}

hipError_t launch_cg_pupdate_f64(int n, double *p, const double *z, const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_pupdate<double>), vgrid(n), dim3(256), 0, s, n, p, z, st);
    return hipGetLastError();
}

hipError_t launch_cg_pupdate_f32(int n, float *p, const float *z, const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg_pupdate<float>), vgrid(n), dim3(256), 0, s, n, p, z, st);
    return hipGetLastError();
}

hipError_t launch_dot_f64(int n, const double *a, const double *b, double *part,
                          const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_dot<double>), vgrid(n), dim3(256), 0, s, n, a, b, part, st);
    return hipGetLastError();
}

hipError_t launch_dot_f32(int n, const float *a, const float *b, double *part,
                          const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_dot<float>), vgrid(n), dim3(256), 0, s, n, a, b, part, st);
    return hipGetLastError();
}

hipError_t launch_jacobi_apply_f64(int n, const double *r, double *z, const double *diag, double tau,
                                   double *prz, const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_jacobi_apply<double>), vgrid(n), dim3(256), 0, s, n, r, z, diag, tau, prz, st);
    return hipGetLastError();
}

hipError_t launch_jacobi_apply_f32(int n, const float *r, float *z, const float *diag, double tau,
                                   double *prz, const CGState *st, hipStream_t s)
{
    hipLaunchKernelGGL((k_jacobi_apply<float>), vgrid(n), dim3(256), 0, s, n, r, z, diag, tau, prz, st);
    return hipGetLastError();
}

hipError_t launch_jacobi_step_f64(int lpr, int n, const int *rp, const int *ci, const double *v,
                                  const double *diag, const double *b, const double *xin,
                                  double *xout, double tau, int x_is_zero, hipStream_t s)
{
    const dim3 grid(spmv_blocks(n, lpr)), block(256);
    switch (lpr) {
    case 64: hipLaunchKernelGGL((k_jacobi_step<64, double, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 32: hipLaunchKernelGGL((k_jacobi_step<32, double, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 16: hipLaunchKernelGGL((k_jacobi_step<16, double, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 8: hipLaunchKernelGGL((k_jacobi_step<8, double, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_jacobi_step_v32(int lpr, int n, const int *rp, const int *ci, const float *v,
                                  const double *diag, const double *b, const double *xin,
                                  double *xout, double tau, int x_is_zero, hipStream_t s)
{
    const dim3 grid(spmv_blocks(n, lpr)), block(256);
    switch (lpr) {
    case 64: hipLaunchKernelGGL((k_jacobi_step<64, float, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 32: hipLaunchKernelGGL((k_jacobi_step<32, float, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 16: hipLaunchKernelGGL((k_jacobi_step<16, float, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 8: hipLaunchKernelGGL((k_jacobi_step<8, float, double>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_jacobi_step_f32(int lpr, int n, const int *rp, const int *ci, const float *v,
                                  const float *diag, const float *b, const float *xin,
                                  float *xout, float tau, int x_is_zero, hipStream_t s)
{
    const dim3 grid(spmv_blocks(n, lpr)), block(256);
    switch (lpr) {
    case 64: hipLaunchKernelGGL((k_jacobi_step<64, float, float>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 32: hipLaunchKernelGGL((k_jacobi_step<32, float, float>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 16: hipLaunchKernelGGL((k_jacobi_step<16, float, float>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    case 8: hipLaunchKernelGGL((k_jacobi_step<8, float, float>), grid, block, 0, s, n, rp, ci, v, diag, b, xin, xout, tau, x_is_zero); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_diag_f64(int n, const int *rp, const int *ci, const double *v, double *diag,
                           hipStream_t s)
{
    hipLaunchKernelGGL((k_diag<double, double>), dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, v, diag);
    return hipGetLastError();
}

hipError_t launch_diag_f32(int n, const int *rp, const int *ci, const float *v, float *diag,
                           hipStream_t s)
{
    hipLaunchKernelGGL((k_diag<float, float>), dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, v, diag);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(int64_t n, const double *a, float *b, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, b);
    return hipGetLastError();
}

hipError_t launch_accum_correction(int n, double *x, float *xc, hipStream_t s)
{
    hipLaunchKernelGGL(k_accum_correction, dim3((n + 255) / 256), dim3(256), 0, s, n, x, xc);
    return hipGetLastError();
}

hipError_t launch_axpby(int n, double a, const double *x, double b, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpby, dim3((n + 255) / 256), dim3(256), 0, s, n, a, x, b, y);
    return hipGetLastError();
}

hipError_t launch_cg_scalar(int step, const double *p0, int np0, const double *p1, int np1,
                            CGState *st, double *hist, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_scalar, dim3(1), dim3(256), 0, s, step, p0, np0, p1, np1, st, hist);
    return hipGetLastError();
}

hipError_t launch_cg_reliable(const double *p0, int np, CGState *st, double *hist, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_reliable, dim3(1), dim3(256), 0, s, p0, np, st, hist);
    return hipGetLastError();
}

}  // namespace gsx
