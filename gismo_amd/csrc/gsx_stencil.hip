// Structured (stencil) operator of single-patch problems and the grid-layout
// CG kernels of libgsx (gfx950).
//
// On one tensor-product patch with open knots the row of free function i
// couples exactly the functions j with |j_k - i_k| <= p in every direction,
// so the row is a (2p+1)^d stencil.  The stencil operator stores, for every
// free row, all (2p+1)^d coefficients (zero where the neighbour is outside
// the patch or eliminated) in [stencil slot][row] order and no column
// indices.  Vectors live on the full tensor grid of basis functions with
// zeros at the eliminated functions, so neighbour indices are the row's grid
// index plus a constant offset.  Same products and summation order per row
// as the CSR SpMV (columns ascending = grid order).
//
// CG (gsConjugateGradient::step, src/gsSolver/gsConjugateGradient.hpp:55-93)
// in three launches per iteration: SpMV + partial (p, Ap); update
// (x, r, z = tau r / diag, partial (r,r), (r,z)); direction update
// (p = z + beta p).  Every block of a launch sums the previous launch's
// partials itself (bounded counts, fixed order) and derives alpha or beta,
// so there is no single-block reduction kernel in the loop; block 0 records
// the scalars.  The (r,z) of iteration k lives in slot k & 1.
#include "gsx_common.h"
#include "gsx_kernels.h"

namespace gsx {

constexpr int kMaxPartials = 2048;

// block sum of 256 threads: wavefront shuffles, then the 4 wave sums in
// wave order (fixed order: bitwise reproducible)
__device__ inline double bsum256(double v, double *s)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    return r;
}

__device__ inline double sum_partials(const double *__restrict__ p, int np, double *s)
{
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) a += p[i];
    return bsum256(a, s);
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, int64_t n, T v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// CSR (free numbering, sorted columns) -> stencil slots.  One thread per row.
template <int D, int P>
__global__ void k_csr_to_stencil(int nfree, const int *__restrict__ rowptr,
                                 const int *__restrict__ col, const double *__restrict__ val,
                                 const int *__restrict__ free2grid, int nbx, int nby,
                                 double *__restrict__ s64, float *__restrict__ s32)
{
    constexpr int W = 2 * P + 1;
    constexpr int NS = (D == 3) ? W * W * W : W * W;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nfree) return;
    const int g = free2grid[r];
    const int gx = g % nbx, gy = (g / nbx) % nby, gz = g / (nbx * nby);
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const int h = free2grid[col[k]];
        const int dx = h % nbx - gx, dy = (h / nbx) % nby - gy;
        const int dz = (D == 3) ? h / (nbx * nby) - gz : 0;
        const int slot = (D == 3) ? ((dz + P) * W + (dy + P)) * W + (dx + P) : (dy + P) * W + (dx + P);
        if (s64) s64[stencil_index(slot, r, NS)] = val[k];
        if (s32 && slot >= NS / 2) s32[stencil_index(slot - NS / 2, r, stencil_half(NS))] = (float)val[k];
    }
}

// diagonal (centre slot) of the structured operator onto the grid
__global__ void k_stencil_diag(int nfree, int ns, int centre, const int *__restrict__ free2grid,
                               const double *__restrict__ s64, double *__restrict__ dst)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nfree) dst[free2grid[r]] = s64[stencil_index(centre, r, ns)];
}

template <typename S, typename Dt>
__global__ void k_scatter_to_grid(int nfree, const int *__restrict__ free2grid,
                                  const S *__restrict__ src, Dt *__restrict__ dst)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nfree) dst[free2grid[r]] = (Dt)src[r];
}

template <typename S>
__global__ void k_gather_from_grid(int nfree, const int *__restrict__ free2grid,
                                   const S *__restrict__ src, double *__restrict__ dst)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nfree) dst[r] = (double)src[free2grid[r]];
}

// ---------------------------------------------------------------------------
// Stencil SpMV.  MODE 0: y = A x; 1: y = A x, partial (x, y); 2: y = b - A x,
// partial (y, y).  Vectors on the grid; rows = free functions.
// ---------------------------------------------------------------------------
template <int D, int P, typename V, typename T, int MODE>
__global__ __launch_bounds__(256) void k_sspmv(int nfree, const int *__restrict__ free2grid,
                                               int sy, int sz, const V *__restrict__ sval,
                                               const T *__restrict__ x, T *__restrict__ y,
                                               const T *__restrict__ b,
                                               double *__restrict__ partial,
                                               CGState *__restrict__ st, int k)
{
    __shared__ double s[4];
    if (st && (st->done || (st->flag_next && st->flag_next <= k))) return;
    constexpr int W = 2 * P + 1;
    constexpr int NS = (D == 3) ? W * W * W : W * W;
    static_assert(kStencilTile == 256, "one block applies one row tile");
    double contrib = 0.0;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    for (int r = blk * 256 + threadIdx.x; r < nfree; r += gridDim.x * 256) {
        const int g = free2grid[r];
        const T *xb = x + g;
        const V *vb = sval + stencil_index(0, r, NS);
        T acc = 0;
        if (D == 3) {
#pragma unroll 1
            for (int dz = -P; dz <= P; ++dz) {
#pragma unroll
                for (int dy = -P; dy <= P; ++dy) {
                    const int base = dz * sz + dy * sy;
                    const int slot0 = ((dz + P) * W + (dy + P)) * W;
#pragma unroll
                    for (int dx = -P; dx <= P; ++dx)
                        acc += (T)vb[(slot0 + dx + P) * kStencilTile] * xb[base + dx];
                }
            }
        } else {
#pragma unroll
            for (int dy = -P; dy <= P; ++dy) {
                const int base = dy * sy;
                const int slot0 = (dy + P) * W;
#pragma unroll
                for (int dx = -P; dx <= P; ++dx)
                    acc += (T)vb[(slot0 + dx + P) * kStencilTile] * xb[base + dx];
            }
        }
        if (MODE == 2) {
            const T v = b[g] - acc;
            y[g] = v;
            contrib += (double)v * (double)v;
        } else {
            y[g] = acc;
            if (MODE == 1) contrib += (double)x[g] * (double)acc;
        }
    }
    if (MODE != 0) {
        const double tot = bsum256(contrib, s);
        if (threadIdx.x == 0) partial[blockIdx.x] = tot;
    }
}

// The same SpMV on the half-stored FP32 operator (stencil_half layout): an
// upper slot is read from the row itself, a lower slot from the neighbour row
// at that offset (free box m0 x m1 x m2, rows lexicographic), whose mirrored
// upper slot holds the same coefficient since the operator is bitwise
// symmetric.  Neighbours outside the free box contribute 0 (x is 0 there).
//
// Order: each upper x-line of the stencil, (dz, dy) > 0 lexicographically
// (the centre line from dx = 0), is taken together with its mirror line
// (-dz, -dy).  Both read the same slot plane of the operator, the mirror one
// at the rows shifted by the offset, so a block reads the plane at the same
// point of its loop as the blocks that own those rows (up to 3 z-planes back,
// on the same XCD through xcd_block) and the second read hits the L2.  With
// the lexicographic order the z-lower mirrors were read at the start of the
// loop and their owners read them at its end: each was fetched from HBM twice.
template <int D, int P, typename T, int MODE>
__global__ __launch_bounds__(256) void k_sspmv_half(int nfree, const int *__restrict__ free2grid,
                                                    int sy, int sz, int m0, int m1, int m2,
                                                    const float *__restrict__ sval,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    const T *__restrict__ b,
                                                    double *__restrict__ partial,
                                                    CGState *__restrict__ st, int k)
{
    __shared__ double s[4];
    if (st && (st->done || (st->flag_next && st->flag_next <= k))) return;
    constexpr int W = 2 * P + 1;
    constexpr int NS = (D == 3) ? W * W * W : W * W;
    constexpr int C = NS / 2, NH = stencil_half(NS);
    // full upper x-lines after the centre one: dz = 0, dy = 1..P, then (3D)
    // dz = 1..P, dy = -P..P
    constexpr int NL = (D == 3) ? P + P * W : P;
    double contrib = 0.0;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    for (int r = blk * 256 + threadIdx.x; r < nfree; r += gridDim.x * 256) {
        const int g = free2grid[r];
        const int ia = r % m0, ib = (r / m0) % m1, ic = (D == 3) ? r / (m0 * m1) : 0;
        const T *xb = x + g;
        const float *vb = sval + stencil_index(0, r, NH);
        T acc = 0;
        // centre line: own slots dx = 0..P (half slots 0..P), mirrors dx = -P..-1
#pragma unroll
        for (int dx = 0; dx <= P; ++dx) acc += (T)vb[dx * kStencilTile] * xb[dx];
#pragma unroll
        for (int dx = -P; dx < 0; ++dx)
            acc += (T)(ia + dx >= 0 ? sval[stencil_index(-dx, r + dx, NH)] : 0.0f) * xb[dx];
#pragma unroll 1
        for (int u = 0; u < NL; ++u) {
            const int dz = (D == 2 || u < P) ? 0 : 1 + (u - P) / W;
            const int dy = (D == 2 || u < P) ? u + 1 : (u - P) % W - P;
            const int h0 = (D == 3 ? ((dz + P) * W + (dy + P)) * W : (dy + P) * W) - C;   // (dz, dy, -P)
            const int base = dz * sz + dy * sy;
#pragma unroll
            for (int dx = -P; dx <= P; ++dx) acc += (T)vb[(h0 + dx + P) * kStencilTile] * xb[base + dx];
            const bool okyz = ib - dy >= 0 && ib - dy < m1 && ic - dz >= 0;
            const int rm = r - m0 * (dy + m1 * dz);
#pragma unroll
            for (int dx = -P; dx <= P; ++dx) {
                const bool ok = okyz && ia + dx >= 0 && ia + dx < m0;
                acc += (T)(ok ? sval[stencil_index(h0 - dx + P, rm + dx, NH)] : 0.0f) * xb[dx - base];
            }
        }
        if (MODE == 2) {
            const T v = b[g] - acc;
            y[g] = v;
            contrib += (double)v * (double)v;
        } else {
            y[g] = acc;
            if (MODE == 1) contrib += (double)x[g] * (double)acc;
        }
    }
    if (MODE != 0) {
        const double tot = bsum256(contrib, s);
        if (threadIdx.x == 0) partial[blockIdx.x] = tot;
    }
}

// ---------------------------------------------------------------------------
// grid-layout CG vector kernels (n = grid size; diag = 1 at eliminated
// functions so z stays 0 there)
// ---------------------------------------------------------------------------
// x += alpha p; r -= alpha t; z = tau r / diag; partials (r, r), (r, z).
// Every block derives alpha = rz_k / (p, A p) from the SpMV's partials
// itself (a bounded count, summed in a fixed order).
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_cg2_update(int n, TX *__restrict__ x, T *__restrict__ r,
                                                    const T *__restrict__ p, const T *__restrict__ t,
                                                    T *__restrict__ z, const T *__restrict__ diag,
                                                    double tau, const double *__restrict__ pap_parts,
                                                    int npap, double *__restrict__ part_rr,
                                                    double *__restrict__ part_rz,
                                                    CGState *__restrict__ st, int k)
{
    __shared__ double s[4];
    if (st->done || (st->flag_next && st->flag_next <= k)) return;
    const double pap = sum_partials(pap_parts, npap, s);
    const double alpha_d = st->rzs[k & 1] / pap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->pap = pap;
        st->alpha = alpha_d;
    }
    const T alpha = (T)alpha_d;
    const T ta = (T)tau;
    double rr = 0.0, rz = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        x[i] += (TX)(alpha * p[i]);
        const T ri = r[i] - alpha * t[i];
        r[i] = ri;
        const T zi = diag ? ta * ri / diag[i] : ri;
        z[i] = zi;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    const double a = bsum256(rr, s);
    const double c = bsum256(rz, s);
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = a;
        part_rz[blockIdx.x] = c;
    }
}

// Stop test and p = z + beta p.  MIXED == 0: reference stopping rule (err <
// tol stops, before the direction update).  MIXED == 1: FP32 recurrence of
// the mixed solver: never stops on its own residual, flags a reliable
// update instead and pauses from iteration k + 1 on until it has run.
template <typename T, int MIXED>
__global__ __launch_bounds__(256) void k_cg2_direction(int n, T *__restrict__ p, const T *__restrict__ z,
                                                       const double *__restrict__ part_rr,
                                                       const double *__restrict__ part_rz, int np,
                                                       CGState *__restrict__ st, int k,
                                                       double *__restrict__ hist)
{
    __shared__ double s[4];
    // flag_next is only ever set to k + 1 below, so the blocks of this
    // launch read the same decision whether or not block 0 has written it
    if (st->done || (st->flag_next && st->flag_next <= k)) return;
    const double rr = sum_partials(part_rr, np, s);
    const double rzn = sum_partials(part_rz, np, s);
    const double err = sqrt(rr) / st->bnorm;
    const bool conv = !MIXED && err < st->tol;
    const double beta = rzn / st->rzs[k & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->it = k + 1;
        st->rr = rr;
        st->err = err;
        if (hist) hist[k + 1] = err;
        if (conv) {
            st->done = 1;
        } else {
            st->beta = beta;
            st->rzs[(k + 1) & 1] = rzn;
            if (MIXED && (rr < st->rel_since_update * st->rr_ref || err < st->tol)) {
                st->need_update = 1;
                st->flag_next = k + 1;
            }
            if (k + 1 >= st->maxit) st->done = 2;
        }
    }
    if (conv) return;
    const T tb = (T)beta;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = z[i] + tb * p[i];
}

// z = tau r / diag (or z = r), partial (r, z)
template <typename T>
__global__ __launch_bounds__(256) void k_cg2_precond(int n, const T *__restrict__ r, T *__restrict__ z,
                                                     const T *__restrict__ diag, double tau,
                                                     double *__restrict__ part_rz)
{
    __shared__ double s[4];
    double rz = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const T zi = diag ? (T)tau * r[i] / diag[i] : r[i];
        z[i] = zi;
        rz += (double)r[i] * (double)zi;
    }
    const double c = bsum256(rz, s);
    if (threadIdx.x == 0) part_rz[blockIdx.x] = c;
}

// partial (a, b) in the grid solvers' block layout (a preconditioner that
// is not fused into the update writes z, then this gives the (r, z) partials)
template <typename T>
__global__ __launch_bounds__(256) void k_cg2_dot(int n, const T *__restrict__ a, const T *__restrict__ b,
                                                 double *__restrict__ part, const CGState *__restrict__ st)
{
    __shared__ double s[4];
    if (st && st->done) return;
    double v = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) v += (double)a[i] * (double)b[i];
    const double c = bsum256(v, s);
    if (threadIdx.x == 0) part[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_set_rz(const double *__restrict__ parts, int np, CGState *__restrict__ st,
                                                int slot)
{
    __shared__ double s[4];
    const double a = sum_partials(parts, np, s);
    if (threadIdx.x == 0) st->rzs[slot & 1] = a;
}

template <typename S, typename Dt>
__global__ void k_copy_convert(int64_t n, const S *__restrict__ a, Dt *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (Dt)a[i];
}

// x64 += xc32; xc32 = 0 (grid)
__global__ void k_accum_grid(int64_t n, double *__restrict__ x, float *__restrict__ xc)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] += (double)xc[i];
        xc[i] = 0.0f;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int cap)
{
    int64_t b = (n + 255) / 256;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned)b;
}

int sspmv_blocks(int nfree) { return (int)blocks_for(nfree, kMaxPartials); }
int vec_blocks(int64_t n) { return (int)blocks_for(n, 1024); }

template <typename T>
hipError_t launch_fill_t(T *a, int64_t n, T v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_fill<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *a, int64_t n, double v, hipStream_t s) { return launch_fill_t<double>(a, n, v, s); }
hipError_t launch_fill_f32(float *a, int64_t n, float v, hipStream_t s) { return launch_fill_t<float>(a, n, v, s); }

hipError_t launch_csr_to_stencil(int d, int p, int nfree, const int *rp, const int *ci,
                                 const double *v, const int *f2g, int nbx, int nby, double *s64,
                                 float *s32, hipStream_t s)
{
    const dim3 grid((nfree + 255) / 256), blk(256);
#define GSX_C(D, P) if (d == D && p == P) { hipLaunchKernelGGL((k_csr_to_stencil<D, P>), grid, blk, 0, s, nfree, rp, ci, v, f2g, nbx, nby, s64, s32); return hipGetLastError(); }
    GSX_C(2, 1) GSX_C(2, 2) GSX_C(2, 3) GSX_C(2, 4)
    GSX_C(3, 1) GSX_C(3, 2) GSX_C(3, 3) GSX_C(3, 4)
#undef GSX_C
    return hipErrorInvalidValue;
}

hipError_t launch_stencil_diag(int nfree, int ns, const int *f2g, const double *s64, double *diag_grid,
                              hipStream_t s)
{
    if (nfree <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stencil_diag, dim3((nfree + 255) / 256), dim3(256), 0, s, nfree, ns, (ns - 1) / 2, f2g, s64,
                       diag_grid);
    return hipGetLastError();
}

hipError_t launch_scatter_f64_to_grid(int nfree, const int *f2g, const double *src, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL((k_scatter_to_grid<double, double>), dim3((nfree + 255) / 256), dim3(256), 0, s, nfree, f2g, src, dst);
    return hipGetLastError();
}

hipError_t launch_gather_f64_from_grid(int nfree, const int *f2g, const double *src, double *dst, hipStream_t s)
{
    hipLaunchKernelGGL((k_gather_from_grid<double>), dim3((nfree + 255) / 256), dim3(256), 0, s, nfree, f2g, src, dst);
    return hipGetLastError();
}

template <typename V, typename T>
static hipError_t sspmv_dispatch(int d, int p, int mode, int nfree, const int *f2g, int sy, int sz,
                                 const V *sv, const T *x, T *y, const T *b, double *part,
                                 CGState *st, int k, hipStream_t s)
{
    const dim3 grid(sspmv_blocks(nfree)), blk(256);
#define GSX_S(D, P, M) if (d == D && p == P && mode == M) { hipLaunchKernelGGL((k_sspmv<D, P, V, T, M>), grid, blk, 0, s, nfree, f2g, sy, sz, sv, x, y, b, part, st, k); return hipGetLastError(); }
#define GSX_S3(D, P) GSX_S(D, P, 0) GSX_S(D, P, 1) GSX_S(D, P, 2)
    GSX_S3(2, 1) GSX_S3(2, 2) GSX_S3(2, 3) GSX_S3(2, 4)
    GSX_S3(3, 1) GSX_S3(3, 2) GSX_S3(3, 3) GSX_S3(3, 4)
#undef GSX_S3
#undef GSX_S
    return hipErrorInvalidValue;
}

hipError_t launch_sspmv_f64(int d, int p, int mode, int nfree, const int *f2g, int sy, int sz,
                            const double *sv, const double *x, double *y, const double *b,
                            double *part, CGState *st, int k, hipStream_t s)
{
    return sspmv_dispatch<double, double>(d, p, mode, nfree, f2g, sy, sz, sv, x, y, b, part, st, k, s);
}

hipError_t launch_sspmv_f32(int d, int p, int mode, int nfree, const int *f2g, int sy, int sz, int m0, int m1,
                            int m2, const float *sv, const float *x, float *y, const float *b,
                            double *part, CGState *st, int k, hipStream_t s)
{
    if ((int64_t)m0 * m1 * m2 != nfree) return hipErrorInvalidValue;
    const dim3 grid(sspmv_blocks(nfree)), blk(256);
#define GSX_S(D, P, M) if (d == D && p == P && mode == M) { hipLaunchKernelGGL((k_sspmv_half<D, P, float, M>), grid, blk, 0, s, nfree, f2g, sy, sz, m0, m1, m2, sv, x, y, b, part, st, k); return hipGetLastError(); }
#define GSX_S3(D, P) GSX_S(D, P, 0) GSX_S(D, P, 1) GSX_S(D, P, 2)
    GSX_S3(2, 1) GSX_S3(2, 2) GSX_S3(2, 3) GSX_S3(2, 4)
    GSX_S3(3, 1) GSX_S3(3, 2) GSX_S3(3, 3) GSX_S3(3, 4)
#undef GSX_S3
#undef GSX_S
    return hipErrorInvalidValue;
}

hipError_t launch_cg2_update_f64(int n, double *x, double *r, const double *p, const double *t,
                                 double *z, const double *diag, double tau, const double *pap,
                                 int npap, double *prr, double *prz, CGState *st, int k, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_update<double, double>), dim3(vec_blocks(n)), dim3(256), 0, s, n, x, r, p, t, z,
                       diag, tau, pap, npap, prr, prz, st, k);
    return hipGetLastError();
}

hipError_t launch_cg2_update_f32(int n, float *x, float *r, const float *p, const float *t, float *z,
                                 const float *diag, double tau, const double *pap, int npap,
                                 double *prr, double *prz, CGState *st, int k, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_update<float, float>), dim3(vec_blocks(n)), dim3(256), 0, s, n, x, r, p, t, z,
                       diag, tau, pap, npap, prr, prz, st, k);
    return hipGetLastError();
}

hipError_t launch_cg2_direction_f64(int n, double *p, const double *z, const double *prr,
                                    const double *prz, int np, CGState *st, int k, double *hist,
                                    hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_direction<double, 0>), dim3(vec_blocks(n)), dim3(256), 0, s, n, p, z, prr, prz,
                       np, st, k, hist);
    return hipGetLastError();
}

hipError_t launch_cg2_direction_f32_mixed(int n, float *p, const float *z, const double *prr,
                                          const double *prz, int np, CGState *st, int k,
                                          hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_direction<float, 1>), dim3(vec_blocks(n)), dim3(256), 0, s, n, p, z, prr, prz,
                       np, st, k, nullptr);
    return hipGetLastError();
}

hipError_t launch_cg2_dot_f64(int n, const double *a, const double *b, double *part, const CGState *st,
                              hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_dot<double>), dim3(vec_blocks(n)), dim3(256), 0, s, n, a, b, part, st);
    return hipGetLastError();
}

hipError_t launch_cg2_dot_f32(int n, const float *a, const float *b, double *part, const CGState *st,
                              hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_dot<float>), dim3(vec_blocks(n)), dim3(256), 0, s, n, a, b, part, st);
    return hipGetLastError();
}

hipError_t launch_cg2_precond_f64(int n, const double *r, double *z, const double *diag, double tau,
                                  double *prz, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_precond<double>), dim3(vec_blocks(n)), dim3(256), 0, s, n, r, z, diag, tau, prz);
    return hipGetLastError();
}

hipError_t launch_cg2_precond_f32(int n, const float *r, float *z, const float *diag, double tau,
                                  double *prz, hipStream_t s)
{
    hipLaunchKernelGGL((k_cg2_precond<float>), dim3(vec_blocks(n)), dim3(256), 0, s, n, r, z, diag, tau, prz);
    return hipGetLastError();
}

hipError_t launch_set_rz(const double *parts, int np, CGState *st, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_rz, dim3(1), dim3(256), 0, s, parts, np, st, slot);
    return hipGetLastError();
}

hipError_t launch_grid_f64_to_f32(int64_t n, const double *a, float *b, hipStream_t s)
{
    hipLaunchKernelGGL((k_copy_convert<double, float>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, b);
    return hipGetLastError();
}

hipError_t launch_accum_grid(int64_t n, double *x, float *xc, hipStream_t s)
{
    hipLaunchKernelGGL(k_accum_grid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, xc);
    return hipGetLastError();
}

}  // namespace gsx
