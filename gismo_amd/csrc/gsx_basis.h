// Device basis helpers shared by the assembly kernels of libgsx (gfx950).
#pragma once

#include "gsx_common.h"

namespace gsx {

// ---------------------------------------------------------------------------
// 1D basis: values and first derivatives of the P+1 non-zero B-splines at u
// in knot span `span` -- the triangular table of values and knot
// differences, then the k = 1 step of the derivative recursion, in the
// operation order of the reference (a = 1/den, d = a*num, times p).
// ---------------------------------------------------------------------------
template <int P>
__device__ inline void deboor_ders(const double *__restrict__ U, int span, double u,
                                   double *__restrict__ N, double *__restrict__ dN)
{
    double tab[P + 1][P + 1];
    double left[P + 1], right[P + 1];
    tab[0][0] = 1.0;
#pragma unroll
    for (int j = 1; j <= P; ++j) {
        left[j] = u - U[span + 1 - j];
        right[j] = U[span + j] - u;
        double carry = 0.0;
#pragma unroll
        for (int r = 0; r < j; ++r) {
            tab[j][r] = right[r + 1] + left[j - r];
            const double t = tab[r][j - 1] / tab[j][r];
            tab[r][j] = carry + right[r + 1] * t;
            carry = left[j - r] * t;
        }
        tab[j][j] = carry;
    }
#pragma unroll
    for (int r = 0; r <= P; ++r) {
        N[r] = tab[r][P];
        double d = 0.0;
        if (r >= 1) {
            const double a0 = 1.0 / tab[P][r - 1];
            d = a0 * tab[r - 1][P - 1];
        }
        if (r <= P - 1) {
            const double a1 = -1.0 / tab[P][r];
            d += a1 * tab[r][P - 1];
        }
        dN[r] = d * (double)P;
    }
}

__device__ inline double source_value(int fkind, int d, const double *x, const double *felem,
                                      int64_t e)
{
    if (fkind == 1) {
        const double pi = 3.14159265358979323846;
        double v = (double)d * pi * pi;
        for (int r = 0; r < d; ++r) v *= sin(pi * x[r]);
        return v;
    }
    return felem ? felem[e] : 0.0;
}

// Tables of the 1D basis at the mapped Gauss nodes of one element, for all
// directions: sV[k][q][i], sD[k][q][i], mapped weights sW[k][q].
template <int D, int P>
__device__ inline void basis_tables(const PatchDev &pd, const int *ei, const GaussTab &gt,
                                    double (*sV)[P + 1][P + 1], double (*sD)[P + 1][P + 1],
                                    double (*sW)[P + 1])
{
    constexpr int NQ1 = P + 1;
    const int t = threadIdx.x;
    if (t < D * NQ1) {
        const int k = t / NQ1, q = t % NQ1;
        const double *U = pd.knots[k];
        const int span = P + ei[k];
        const double a = U[span], b = U[span + 1];
        const double h = (b - a) / 2.0;
        const double u = h * (gt.x[NQ1][q] + 1.0) + a;
        sW[k][q] = (0.0 == h ? 0.5 : h) * gt.w[NQ1][q];
        double N[P + 1], dN[P + 1];
        deboor_ders<P>(U, span, u, N, dN);
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            sV[k][q][i] = N[i];
            sD[k][q][i] = dN[i];
        }
    }
}

}  // namespace gsx
