// Exact FP64 coarse solve by a banded Cholesky factorisation (gfx950).
//
// The coarsest multigrid operator is sparse and, after a reverse
// Cuthill-McKee ordering of its rows, banded with half-bandwidth b of a few
// hundred.  A = L L^T is factored in that band (no fill outside it):
// n b^2 flops instead of the O(n^3) of a dense inverse.  The reference
// solves its coarsest level with makeSparseCholeskySolver
// (src/gsSolver/gsMatrixOp.h:357-370, a sparse LL^T); this is the same
// exact FP64 solve in another storage.
//
// Storage: the lower band row by row, Lb[i (b+1) + t] = L(i, i - b + t) for
// t = 0..b (t = b: the diagonal); entries left of column 0 are zero.
// Factorisation: blocked right-looking over panels of 64 columns; per panel
// one block factors the 64 x 64 diagonal block in LDS and solves the rows
// below it inside the band, then the trailing window is updated.  Solve:
// one block runs the forward and the backward substitution panel by panel
// (the band products in parallel, the 64-row triangles by one wavefront).
#include "gsx_common.h"
#include "gsx_kernels.h"

namespace gsx {

namespace {

constexpr int kBP = 64;   // panel width

__device__ inline int64_t bidx(int i, int j, int b) { return (int64_t)i * (b + 1) + (j - i + b); }

// Lb = 0, then the permuted lower triangle of A: row i of A is row pos[i]
__global__ void k_band_zero(int64_t total, double *__restrict__ Lb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) Lb[i] = 0.0;
}

__global__ void k_band_fill(int n, int b, const int *__restrict__ rp, const int *__restrict__ ci,
                            const double *__restrict__ v, const int *__restrict__ pos, double *__restrict__ Lb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int pi = pos[i];
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const int pj = pos[ci[k]];
        if (pj <= pi && pi - pj <= b) Lb[bidx(pi, pj, b)] = v[k];
    }
}

// Panel k0: Cholesky of the diagonal block in LDS (column by column, every
// thread computing the pivot's square root itself), then the rows below it
// in the band: x L11^T = a row by row.  err[0] = 1 on a pivot <= 0.
__global__ __launch_bounds__(256) void k_band_panel(int n, int b, int k0, double *__restrict__ Lb, int *__restrict__ err)
{
    __shared__ double D[kBP][kBP + 1];
    const int kb = min(kBP, n - k0);
    const int t = threadIdx.x;
    for (int e = t; e < kb * kb; e += blockDim.x) {
        const int r = e / kb, c = e % kb;
        D[r][c] = (c <= r && r - c <= b) ? Lb[bidx(k0 + r, k0 + c, b)] : 0.0;
    }
    __syncthreads();
    for (int s = 0; s < kb; ++s) {
        const double piv = D[s][s];
        if (!(piv > 0.0)) {
            if (t == 0) err[0] = 1;
            return;
        }
        const double d = sqrt(piv);
        // new values of the trailing entries from the column s (unchanged
        // in this step), written after the barrier (fixed trip count: the
        // values stay in registers)
        constexpr int PER = (kBP * kBP + 255) / 256;
        double nv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = t + 256 * q;
            double x = 0.0;
            if (e < kb * kb) {
                const int r = e / kb, c = e % kb;
                x = D[r][c];
                if (c == s && r == s) x = d;
                else if (c == s && r > s) x = D[r][s] / d;
                else if (c > s && r >= c) x = D[r][c] - (D[r][s] / d) * (D[c][s] / d);
            }
            nv[q] = x;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = t + 256 * q;
            if (e < kb * kb) D[e / kb][e % kb] = nv[q];
        }
        __syncthreads();
    }
    for (int e = t; e < kb * kb; e += blockDim.x) {
        const int r = e / kb, c = e % kb;
        if (c <= r && r - c <= b) Lb[bidx(k0 + r, k0 + c, b)] = D[r][c];
    }
    // rows below the panel inside the band: i in [k0 + kb, min(n, k0 + kb + b))
    // (row i reaches panel column k0 + kb - 1 while i - (k0 + kb - 1) <= b)
    const int i1 = min(n, k0 + kb + b);
    for (int i = k0 + kb + t; i < i1; i += blockDim.x) {
        const int c0 = max(0, i - b - k0);   // first panel column inside row i's band
        // x L11^T = a in place: x_q (q < c) are already stored in the band
        for (int c = c0; c < kb; ++c) {
            double a = Lb[bidx(i, k0 + c, b)];
            for (int q = c0; q < c; ++q) a -= Lb[bidx(i, k0 + q, b)] * D[c][q];
            Lb[bidx(i, k0 + c, b)] = a / D[c][c];
        }
    }
}

// Trailing update of panel k0: A(i, j) -= sum_c L(i, k0 + c) L(j, k0 + c)
// for k0 + kb <= j <= i < min(n, k0 + kb + b); 32 x 32 tiles of the lower
// triangle of that window, the panel columns of the tile's rows in LDS.
__global__ __launch_bounds__(256) void k_band_update(int n, int b, int k0, double *__restrict__ Lb)
{
    __shared__ double Li[32][kBP + 1], Lj[32][kBP + 1];
    const int kb = min(kBP, n - k0);
    const int w0 = k0 + kb, w1 = min(n, k0 + kb + b);
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (tj > ti) return;
    const int i0 = w0 + 32 * ti, j0 = w0 + 32 * tj;
    if (i0 >= w1) return;
    const int t = threadIdx.x;
    for (int e = t; e < 32 * kb; e += blockDim.x) {
        const int r = e / kb, c = e % kb;
        const int gi = i0 + r, gj = j0 + r, col = k0 + c;
        Li[r][c] = (gi < w1 && gi - col <= b) ? Lb[bidx(gi, col, b)] : 0.0;
        Lj[r][c] = (gj < w1 && gj - col <= b) ? Lb[bidx(gj, col, b)] : 0.0;
    }
    __syncthreads();
    for (int e = t; e < 32 * 32; e += blockDim.x) {
        const int r = e / 32, c = e % 32;
        const int gi = i0 + r, gj = j0 + c;
        if (gi >= w1 || gj > gi || gi - gj > b) continue;
        double s = 0.0;
        for (int q = 0; q < kb; ++q) s += Li[r][q] * Lj[c][q];
        Lb[bidx(gi, gj, b)] -= s;
    }
}

// x = A^-1 rhs through the factor: y = P rhs, L z = y, L^T w = z, x = P^T w.
// One block of 1024 threads; y (n doubles) is its work vector.
__global__ __launch_bounds__(1024) void k_band_solve(int n, int b, const int *__restrict__ pos,
                                                     const double *__restrict__ Lb, const double *__restrict__ rhs,
                                                     double *__restrict__ x, double *__restrict__ y)
{
    __shared__ double part[kBP];
    const int t = threadIdx.x, w = t / kWave, lane = t % kWave;
    constexpr int NW = 1024 / kWave;
    for (int i = t; i < n; i += blockDim.x) y[pos[i]] = rhs[i];
    __syncthreads();
    // forward: rows of panel k0 minus the band left of the panel, then the
    // panel's triangle by wavefront 0 (lane r holds row k0 + r)
    for (int k0 = 0; k0 < n; k0 += kBP) {
        const int kb = min(kBP, n - k0);
        for (int r = w; r < kb; r += NW) {
            const int i = k0 + r, j0 = max(0, i - b);
            double s = 0.0;
            for (int j = j0 + lane; j < k0; j += kWave) s += Lb[bidx(i, j, b)] * y[j];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
            if (lane == 0) part[r] = y[i] - s;
        }
        __syncthreads();
        if (w == 0) {
            double v = lane < kb ? part[lane] : 0.0;
            for (int s = 0; s < kb; ++s) {
                const double ys = __shfl(v, s, kWave) / Lb[bidx(k0 + s, k0 + s, b)];
                if (lane == s) v = ys;
                else if (lane > s && lane < kb && k0 + lane - (k0 + s) <= b) v -= Lb[bidx(k0 + lane, k0 + s, b)] * ys;
            }
            if (lane < kb) y[k0 + lane] = v;
        }
        __syncthreads();
    }
    // backward: L^T w = z, panels from the last; rows below the panel enter
    // through column entries L(j, i), j in (panel end, i + b]
    const int np = (n + kBP - 1) / kBP;
    for (int pb = np - 1; pb >= 0; --pb) {
        const int k0 = pb * kBP, kb = min(kBP, n - k0), kend = k0 + kb;
        for (int r = w; r < kb; r += NW) {
            const int i = k0 + r, j1 = min(n, i + b + 1);
            double s = 0.0;
            for (int j = kend + lane; j < j1; j += kWave) s += Lb[bidx(j, i, b)] * y[j];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
            if (lane == 0) part[r] = y[i] - s;
        }
        __syncthreads();
        if (w == 0) {
            double v = lane < kb ? part[lane] : 0.0;
            for (int s = kb - 1; s >= 0; --s) {
                const double ys = __shfl(v, s, kWave) / Lb[bidx(k0 + s, k0 + s, b)];
                if (lane == s) v = ys;
                else if (lane < s && s - lane <= b) v -= Lb[bidx(k0 + s, k0 + lane, b)] * ys;
            }
            if (lane < kb) y[k0 + lane] = v;
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += blockDim.x) x[i] = y[pos[i]];
}

}  // namespace

hipError_t launch_band_factor(int n, int b, const int *rp, const int *ci, const double *v, const int *pos,
                              double *Lb, int *err, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t total = (int64_t)n * (b + 1);
    hipLaunchKernelGGL(k_band_zero, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, total, Lb);
    hipLaunchKernelGGL(k_band_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, b, rp, ci, v, pos, Lb);
    for (int k0 = 0; k0 < n; k0 += kBP) {
        hipLaunchKernelGGL(k_band_panel, dim3(1), dim3(256), 0, s, n, b, k0, Lb, err);
        const int kb = std::min(kBP, n - k0);
        const int m = std::min(n, k0 + kb + b) - (k0 + kb);
        if (m > 0) {
            const int nt = (m + 31) / 32;
            hipLaunchKernelGGL(k_band_update, dim3(nt, nt), dim3(256), 0, s, n, b, k0, Lb);
        }
    }
    return hipGetLastError();
}

hipError_t launch_band_solve(int n, int b, const int *pos, const double *Lb, const double *rhs, double *x,
                             double *work, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_band_solve, dim3(1), dim3(1024), 0, s, n, b, pos, Lb, rhs, x, work);
    return hipGetLastError();
}

}  // namespace gsx
