// Fast diagonalization preconditioner on the device
// (gsPatchPreconditionersCreator::fastDiagonalizationOp,
// src/gsSolver/gsPatchPreconditionersCreator.hpp:258-348):
//   z = (Q_{d-1} x ... x Q_0) diag(1 / (alpha + sum_k lambda_k)) (Q_{d-1} x ... x Q_0)^T r
// over the free box of a single patch (direction 0 fastest).  Each factor is
// a mode product along one direction: y_a = sum_i M[i*m + a] x_i with
// M = Q_k (to apply Q_k^T) or M = Q_k^T (to apply Q_k).  One thread per
// output: along direction 0 the lanes share the input line (broadcast) and
// read consecutive columns of M; along directions 1, 2 they read
// consecutive inputs and share the row of M.
#include "gsx_common.h"
#include "gsx_kernels.h"

namespace gsx {

template <typename T>
__global__ __launch_bounds__(256) void k_fd_mode(int m0, int m1, int m2, int dir, const T *__restrict__ M,
                                                 const T *__restrict__ x, int64_t xo, int64_t xs1, int64_t xs2,
                                                 T *__restrict__ y, int64_t yo, int64_t ys1, int64_t ys2,
                                                 const double *__restrict__ l0, const double *__restrict__ l1,
                                                 const double *__restrict__ l2, double alpha,
                                                 const CGState *__restrict__ st)
{
    if (st && st->done) return;
    const int64_t n = (int64_t)m0 * m1 * m2;
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    const int a = (int)(o % m0), b = (int)((o / m0) % m1), c = (int)(o / ((int64_t)m0 * m1));
    int m, j;
    int64_t stride, base;
    if (dir == 0) {
        m = m0; j = a; stride = 1; base = xo + b * xs1 + c * xs2;
    } else if (dir == 1) {
        m = m1; j = b; stride = xs1; base = xo + a + c * xs2;
    } else {
        m = m2; j = c; stride = xs2; base = xo + a + b * xs1;
    }
    T acc = 0;
    for (int i = 0; i < m; ++i) acc += M[(int64_t)i * m + j] * x[base + i * stride];
    if (l0) {
        const double den = alpha + l0[a] + l1[b] + (l2 ? l2[c] : 0.0);
        acc = (T)((double)acc / den);
    }
    y[yo + a + b * ys1 + c * ys2] = acc;
}

// Slab form: the block stages the 2D slab of its fixed third index (c for
// directions 0 and 1, b for direction 2; rows x m0 values) in LDS and each
// thread computes one output of that slab from it: along direction 0
// y[c][b][j] = sum_i M[i][j] x[c][b][i] (the lanes share the slab row and
// read consecutive columns of M), along directions 1 and 2
// y[..][j][a] = sum_i M[i][j] x[..][i][a] (the lanes read consecutive slab
// entries and share the column of M).
template <typename T, int DIR>
__global__ __launch_bounds__(256) void k_fd_slab(int m0, int m1, int m2, const T *__restrict__ M,
                                                 const T *__restrict__ x, int64_t xo, int64_t xs1, int64_t xs2,
                                                 T *__restrict__ y, int64_t yo, int64_t ys1, int64_t ys2,
                                                 const double *__restrict__ l0, const double *__restrict__ l1,
                                                 const double *__restrict__ l2, double alpha,
                                                 const CGState *__restrict__ st)
{
    extern __shared__ double fd_smem[];
    T *sX = reinterpret_cast<T *>(fd_smem);
    if (st && st->done) return;
    const int fixed = blockIdx.y;
    const int rows = DIR == 2 ? m2 : m1;
    const int64_t base = xo + (DIR == 2 ? (int64_t)fixed * xs1 : (int64_t)fixed * xs2);
    const int64_t rstride = DIR == 2 ? xs2 : xs1;
    const int n = rows * m0;
    for (int idx = threadIdx.x; idx < n; idx += 256) {
        const int r = idx / m0, a = idx % m0;
        sX[idx] = x[base + a + r * rstride];
    }
    __syncthreads();
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= n) return;
    const int r = idx / m0, q = idx % m0;
    T acc = 0;
    if (DIR == 0) {
        const T *xr = sX + r * m0;
        for (int i = 0; i < m0; ++i) acc += M[(int64_t)i * m0 + q] * xr[i];
    } else {
        for (int i = 0; i < rows; ++i) acc += M[(int64_t)i * rows + r] * sX[i * m0 + q];
    }
    const int a = q, b = DIR == 2 ? fixed : r, c = DIR == 2 ? r : fixed;
    if (l0) acc = (T)((double)acc / (alpha + l0[a] + l1[b] + (l2 ? l2[c] : 0.0)));
    y[yo + a + (int64_t)b * ys1 + (int64_t)c * ys2] = acc;
}

template <typename T>
hipError_t launch_fd_mode_t(int m0, int m1, int m2, int dir, const T *M, const T *x, int64_t xo, int64_t xs1,
                            int64_t xs2, T *y, int64_t yo, int64_t ys1, int64_t ys2, const double *l0,
                            const double *l1, const double *l2, double alpha, const CGState *st, hipStream_t s)
{
    const int64_t n = (int64_t)m0 * m1 * m2;
    if (n <= 0) return hipSuccess;
    if (dir < 0 || dir > 2) return hipErrorInvalidValue;
    const int rows = dir == 2 ? m2 : m1;
    const size_t lds = sizeof(T) * (size_t)rows * m0;
    if (lds <= 60 * 1024) {
        const dim3 grid((unsigned)((rows * m0 + 255) / 256), (unsigned)(dir == 2 ? m1 : m2));
        if (dir == 0)
            hipLaunchKernelGGL((k_fd_slab<T, 0>), grid, dim3(256), lds, s, m0, m1, m2, M, x, xo, xs1, xs2, y, yo, ys1,
                               ys2, l0, l1, l2, alpha, st);
        else if (dir == 1)
            hipLaunchKernelGGL((k_fd_slab<T, 1>), grid, dim3(256), lds, s, m0, m1, m2, M, x, xo, xs1, xs2, y, yo, ys1,
                               ys2, l0, l1, l2, alpha, st);
        else
            hipLaunchKernelGGL((k_fd_slab<T, 2>), grid, dim3(256), lds, s, m0, m1, m2, M, x, xo, xs1, xs2, y, yo, ys1,
                               ys2, l0, l1, l2, alpha, st);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_fd_mode<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m0, m1, m2, dir, M, x,
                       xo, xs1, xs2, y, yo, ys1, ys2, l0, l1, l2, alpha, st);
    return hipGetLastError();
}

hipError_t launch_fd_mode_f64(int m0, int m1, int m2, int dir, const double *M, const double *x, int64_t xo,
                              int64_t xs1, int64_t xs2, double *y, int64_t yo, int64_t ys1, int64_t ys2,
                              const double *l0, const double *l1, const double *l2, double alpha,
                              const CGState *st, hipStream_t s)
{
    return launch_fd_mode_t<double>(m0, m1, m2, dir, M, x, xo, xs1, xs2, y, yo, ys1, ys2, l0, l1, l2, alpha, st, s);
}

hipError_t launch_fd_mode_f32(int m0, int m1, int m2, int dir, const float *M, const float *x, int64_t xo,
                              int64_t xs1, int64_t xs2, float *y, int64_t yo, int64_t ys1, int64_t ys2,
                              const double *l0, const double *l1, const double *l2, double alpha,
                              const CGState *st, hipStream_t s)
{
    return launch_fd_mode_t<float>(m0, m1, m2, dir, M, x, xo, xs1, xs2, y, yo, ys1, ys2, l0, l1, l2, alpha, st, s);
}

// Rectangular mode product of the subspace corrected mass smoother: output
// box o0 x o1 x o2, along direction dir the input has m_in entries and
// y_a = sum_{i < m_in} M[i*o_dir + a] x_i; optionally divided by
// alpha + l0[a] + l1[b] + l2[c] (null arrays count 0) and added to y.
template <typename T>
__global__ __launch_bounds__(256) void k_tensor_mode(int o0, int o1, int o2, int dir, int m_in,
                                                     const T *__restrict__ M, const T *__restrict__ x, int64_t xo,
                                                     int64_t xs1, int64_t xs2, T *__restrict__ y, int64_t yo,
                                                     int64_t ys1, int64_t ys2, int scale, double alpha,
                                                     const double *__restrict__ l0, const double *__restrict__ l1,
                                                     const double *__restrict__ l2, int accumulate,
                                                     const CGState *__restrict__ st)
{
    if (st && st->done) return;
    const int64_t n = (int64_t)o0 * o1 * o2;
    const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (o >= n) return;
    const int a = (int)(o % o0), b = (int)((o / o0) % o1), c = (int)(o / ((int64_t)o0 * o1));
    const int od = dir == 0 ? o0 : (dir == 1 ? o1 : o2);
    const int j = dir == 0 ? a : (dir == 1 ? b : c);
    const int64_t stride = dir == 0 ? 1 : (dir == 1 ? xs1 : xs2);
    const int64_t base = xo + (dir == 0 ? 0 : a) + (dir == 1 ? 0 : b * xs1) + (dir == 2 ? 0 : c * xs2);
    double acc = 0.0;
    for (int i = 0; i < m_in; ++i) acc += (double)M[(int64_t)i * od + j] * (double)x[base + i * stride];
    if (scale) acc /= alpha + (l0 ? l0[a] : 0.0) + (l1 ? l1[b] : 0.0) + (l2 ? l2[c] : 0.0);
    const int64_t at = yo + a + b * ys1 + c * ys2;
    y[at] = accumulate ? (T)((double)y[at] + acc) : (T)acc;
}

template <typename T>
hipError_t launch_tensor_mode_t(int o0, int o1, int o2, int dir, int m_in, const T *M, const T *x, int64_t xo,
                                int64_t xs1, int64_t xs2, T *y, int64_t yo, int64_t ys1, int64_t ys2, int scale,
                                double alpha, const double *l0, const double *l1, const double *l2, int accumulate,
                                const CGState *st, hipStream_t s)
{
    const int64_t n = (int64_t)o0 * o1 * o2;
    if (n <= 0) return hipSuccess;
    if (dir < 0 || dir > 2 || m_in <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_tensor_mode<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, o0, o1, o2, dir, m_in,
                       M, x, xo, xs1, xs2, y, yo, ys1, ys2, scale, alpha, l0, l1, l2, accumulate, st);
    return hipGetLastError();
}

hipError_t launch_tensor_mode_f64(int o0, int o1, int o2, int dir, int m_in, const double *M, const double *x,
                                  int64_t xo, int64_t xs1, int64_t xs2, double *y, int64_t yo, int64_t ys1,
                                  int64_t ys2, int scale, double alpha, const double *l0, const double *l1,
                                  const double *l2, int accumulate, const CGState *st, hipStream_t s)
{
    return launch_tensor_mode_t<double>(o0, o1, o2, dir, m_in, M, x, xo, xs1, xs2, y, yo, ys1, ys2, scale, alpha, l0,
                                        l1, l2, accumulate, st, s);
}

hipError_t launch_tensor_mode_f32(int o0, int o1, int o2, int dir, int m_in, const float *M, const float *x,
                                  int64_t xo, int64_t xs1, int64_t xs2, float *y, int64_t yo, int64_t ys1,
                                  int64_t ys2, int scale, double alpha, const double *l0, const double *l1,
                                  const double *l2, int accumulate, const CGState *st, hipStream_t s)
{
    return launch_tensor_mode_t<float>(o0, o1, o2, dir, m_in, M, x, xo, xs1, xs2, y, yo, ys1, ys2, scale, alpha, l0,
                                       l1, l2, accumulate, st, s);
}

}  // namespace gsx
