// Host-side dense linear algebra and 1D B-spline matrices for the fast
// diagonalization preconditioner (gsPatchPreconditionersCreator::
// fastDiagonalizationOp, src/gsSolver/gsPatchPreconditionersCreator.hpp:
// 258-348).  Plain C++ (no HIP): the library and the CPU tests both use it.
//
//   bspline_1d_matrices  univariate mass and stiffness of one knot vector on
//                        the parameter line (assembleMass/assembleStiffness,
//                        gsPatchPreconditionersCreator.hpp:70-96), Gauss rule
//                        with p+1 points per knot span
//   gen_sym_eig          K q = lambda M q with Q^T M Q = I, eigenvalues
//                        ascending (the role of Eigen's
//                        GeneralizedSelfAdjointEigenSolver): Cholesky of M,
//                        C = L^-1 K L^-T, C = V diag(lambda) V^T by
//                        Householder tridiagonalisation and the implicit QL
//                        iteration with Wilkinson-type shifts (Golub & Van
//                        Loan, Matrix Computations, 4th ed., 8.3), Q = L^-T V
//   lu_solve, greville,  collocation at the anchors for Dirichlet values by
//   find_span            interpolation (gsAssembler::
//                        computeDirichletDofsIntpl)
#pragma once

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace gsx {
namespace dense {

// In-place Cholesky factor (lower triangle, row-major n x n); false if A is
// not positive definite.
inline bool cholesky(int n, std::vector<double> &A)
{
    for (int j = 0; j < n; ++j) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        A[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double t = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = t / ljj;
        }
        for (int i = 0; i < j; ++i) A[(size_t)i * n + j] = 0.0;
    }
    return true;
}

// Householder reduction of the symmetric A (row-major, destroyed) to the
// tridiagonal (d, e) with A = Z T Z^T; e[i] couples i and i+1, e[n-1] = 0.
inline void tridiagonalize(int n, std::vector<double> &A, std::vector<double> &d, std::vector<double> &e,
                           std::vector<double> &Z)
{
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), w(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;   // length of x = A[k+1.., k]
        double xn = 0.0;
        for (int i = 0; i < m; ++i) xn += A[(size_t)(k + 1 + i) * n + k] * A[(size_t)(k + 1 + i) * n + k];
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const double x0 = A[(size_t)(k + 1) * n + k];
        const double alpha = x0 > 0 ? -xn : xn;
        for (int i = 0; i < m; ++i) v[i] = A[(size_t)(k + 1 + i) * n + k];
        v[0] -= alpha;
        double vv = 0.0;
        for (int i = 0; i < m; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        // p = beta A_sub v ; w = p - (beta/2)(p.v) v ; A_sub -= v w^T + w v^T
        double pv = 0.0;
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += A[(size_t)(k + 1 + i) * n + (k + 1 + j)] * v[j];
            p[i] = beta * s;
            pv += p[i] * v[i];
        }
        const double K = 0.5 * beta * pv;
        for (int i = 0; i < m; ++i) w[i] = p[i] - K * v[i];
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j)
                A[(size_t)(k + 1 + i) * n + (k + 1 + j)] -= v[i] * w[j] + w[i] * v[j];
        A[(size_t)(k + 1) * n + k] = alpha;
        A[(size_t)k * n + (k + 1)] = alpha;
        for (int i = 1; i < m; ++i) {
            A[(size_t)(k + 1 + i) * n + k] = 0.0;
            A[(size_t)k * n + (k + 1 + i)] = 0.0;
        }
        // Z = Z H
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += Z[(size_t)r * n + (k + 1 + i)] * v[i];
            s *= beta;
            for (int i = 0; i < m; ++i) Z[(size_t)r * n + (k + 1 + i)] -= s * v[i];
        }
    }
    d.assign(n, 0.0);
    e.assign(n, 0.0);
    for (int i = 0; i < n; ++i) d[i] = A[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; ++i) e[i] = A[(size_t)(i + 1) * n + i];
}

// Implicit QL iteration on the tridiagonal (d, e), rotations accumulated into
// the columns of Z.  Returns false if an eigenvalue takes more than 60 sweeps.
inline bool tridiagonal_ql(int n, std::vector<double> &d, std::vector<double> &e, std::vector<double> &Z)
{
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        int m;
        do {
            for (m = l; m + 1 < n; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-16 * dd) break;
            }
            if (m != l) {
                if (++iter > 60) return false;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                bool underflow = false;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        underflow = true;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; ++k) {
                        f = Z[(size_t)k * n + (i + 1)];
                        Z[(size_t)k * n + (i + 1)] = s * Z[(size_t)k * n + i] + c * f;
                        Z[(size_t)k * n + i] = c * Z[(size_t)k * n + i] - s * f;
                    }
                }
                if (underflow) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return true;
}

// Symmetric eigen-decomposition: A = V diag(w) V^T, w ascending; column k of
// V (V[i*n+k]) is the eigenvector of w[k].  A is destroyed.
inline bool sym_eig(int n, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V)
{
    std::vector<double> e, Z;
    tridiagonalize(n, A, w, e, Z);
    if (!tridiagonal_ql(n, w, e, Z)) return false;
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] < w[b]; });
    std::vector<double> ws(n);
    V.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        ws[k] = w[ord[k]];
        for (int i = 0; i < n; ++i) V[(size_t)i * n + k] = Z[(size_t)i * n + ord[k]];
    }
    w.swap(ws);
    return true;
}

// K q = lambda M q (K symmetric, M SPD; row-major n x n): lambda ascending,
// Q[i*n+k] = component i of eigenvector k, Q^T M Q = I, Q^T K Q = diag.
inline bool gen_sym_eig(int n, const std::vector<double> &K, const std::vector<double> &M, std::vector<double> &lam,
                        std::vector<double> &Q)
{
    std::vector<double> L = M;
    if (!cholesky(n, L)) return false;
    // C = L^-1 K L^-T: forward solves on the columns, then on the rows
    std::vector<double> X = K;   // X = L^-1 K
    for (int c = 0; c < n; ++c)
        for (int i = 0; i < n; ++i) {
            double s = X[(size_t)i * n + c];
            for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * X[(size_t)k * n + c];
            X[(size_t)i * n + c] = s / L[(size_t)i * n + i];
        }
    std::vector<double> C((size_t)n * n);   // C = X L^-T = (L^-1 X^T)^T
    for (int r = 0; r < n; ++r)
        for (int i = 0; i < n; ++i) {
            double s = X[(size_t)r * n + i];
            for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * C[(size_t)r * n + k];
            C[(size_t)r * n + i] = s / L[(size_t)i * n + i];
        }
    for (int i = 0; i < n; ++i)   // symmetrise the rounding
        for (int j = 0; j < i; ++j) {
            const double a = 0.5 * (C[(size_t)i * n + j] + C[(size_t)j * n + i]);
            C[(size_t)i * n + j] = C[(size_t)j * n + i] = a;
        }
    std::vector<double> V;
    if (!sym_eig(n, C, lam, V)) return false;
    // Q = L^-T V: back substitution per column
    Q.assign((size_t)n * n, 0.0);
    for (int c = 0; c < n; ++c)
        for (int i = n - 1; i >= 0; --i) {
            double s = V[(size_t)i * n + c];
            for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * Q[(size_t)k * n + c];
            Q[(size_t)i * n + c] = s / L[(size_t)i * n + i];
        }
    return true;
}

// Solve A X = B in place (A n x n row-major, destroyed; B n x nrhs
// row-major, overwritten by X) by Gaussian elimination with partial
// pivoting; false if A is singular.
inline bool lu_solve(int n, std::vector<double> &A, std::vector<double> &B, int nrhs)
{
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)i * n + k]) > std::fabs(A[(size_t)piv * n + k])) piv = i;
        if (A[(size_t)piv * n + k] == 0.0) return false;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)piv * n + j]);
            for (int j = 0; j < nrhs; ++j) std::swap(B[(size_t)k * nrhs + j], B[(size_t)piv * nrhs + j]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A[(size_t)i * n + k] / A[(size_t)k * n + k];
            if (f == 0.0) continue;
            for (int j = k; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
            for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] -= f * B[(size_t)k * nrhs + j];
        }
    }
    for (int k = n - 1; k >= 0; --k)
        for (int j = 0; j < nrhs; ++j) {
            double v = B[(size_t)k * nrhs + j];
            for (int i = k + 1; i < n; ++i) v -= A[(size_t)k * n + i] * B[(size_t)i * nrhs + j];
            B[(size_t)k * nrhs + j] = v / A[(size_t)k * n + k];
        }
    return true;
}

// Greville abscissae (the anchors of gsBSplineBasis) of an open knot vector.
inline std::vector<double> greville(const std::vector<double> &U, int p)
{
    const int n = (int)U.size() - p - 1;
    std::vector<double> g(n);
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = 1; k <= p; ++k) s += U[i + k];
        g[i] = s / p;
    }
    return g;
}

// Knot span of u in the open knot vector U (the last nonempty span at the
// right end).
inline int find_span(const std::vector<double> &U, int p, double u)
{
    const int n = (int)U.size() - p - 1;
    if (u >= U[n]) {
        int s = n - 1;
        while (s > p && !(U[s + 1] > U[s])) --s;
        return s;
    }
    int s = p;
    while (s + 1 < n && U[s + 1] <= u) ++s;
    return s;
}

// B-spline values and first derivatives of the p+1 functions active on
// [U[span], U[span+1]) at u (Cox-de Boor triangle, The NURBS Book A2.2/A2.3).
inline void bspline_ders(const std::vector<double> &U, int p, int span, double u, double *N, double *dN)
{
    std::vector<double> left(p + 1), right(p + 1);
    std::vector<std::vector<double>> ndu(p + 1, std::vector<double>(p + 1, 0.0));
    ndu[0][0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[span + 1 - j];
        right[j] = U[span + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            ndu[j][r] = right[r + 1] + left[j - r];   // lower triangle: knot differences
            const double tmp = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        ndu[j][j] = saved;
    }
    for (int j = 0; j <= p; ++j) N[j] = ndu[j][p];
    for (int r = 0; r <= p; ++r) {
        // first derivative: p (N_{r,p-1}/(U[..]) - N_{r+1,p-1}/(..))
        double der = 0.0;
        if (r >= 1) der += ndu[r - 1][p - 1] / ndu[p][r - 1];
        if (r <= p - 1) der -= ndu[r][p - 1] / ndu[p][r];
        dN[r] = p * der;
    }
}

// Values and derivatives up to order nder (<= p) of the p+1 functions active
// on span at u (The NURBS Book A2.3): ders[k][j] = d^k N_{span-p+j} / du^k.
inline void bspline_ders_all(const std::vector<double> &U, int p, int span, double u, int nder,
                             std::vector<std::vector<double>> &ders)
{
    std::vector<double> left(p + 1), right(p + 1);
    std::vector<std::vector<double>> ndu(p + 1, std::vector<double>(p + 1, 0.0)), a(2, std::vector<double>(p + 1));
    ndu[0][0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[span + 1 - j];
        right[j] = U[span + j] - u;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            ndu[j][r] = right[r + 1] + left[j - r];
            const double tmp = ndu[r][j - 1] / ndu[j][r];
            ndu[r][j] = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        ndu[j][j] = saved;
    }
    ders.assign(nder + 1, std::vector<double>(p + 1, 0.0));
    for (int j = 0; j <= p; ++j) ders[0][j] = ndu[j][p];
    for (int r = 0; r <= p; ++r) {
        int s1 = 0, s2 = 1;
        a[0][0] = 1.0;
        for (int k = 1; k <= nder; ++k) {
            double d = 0.0;
            const int rk = r - k, pk = p - k;
            if (r >= k) {
                a[s2][0] = a[s1][0] / ndu[pk + 1][rk];
                d = a[s2][0] * ndu[rk][pk];
            }
            const int j1 = rk >= -1 ? 1 : -rk, j2 = (r - 1 <= pk) ? k - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                a[s2][j] = (a[s1][j] - a[s1][j - 1]) / ndu[pk + 1][rk + j];
                d += a[s2][j] * ndu[rk + j][pk];
            }
            if (r <= pk) {
                a[s2][k] = -a[s1][k - 1] / ndu[pk + 1][r];
                d += a[s2][k] * ndu[r][pk];
            }
            ders[k][r] = d;
            std::swap(s1, s2);
        }
    }
    double f = p;
    for (int k = 1; k <= nder; ++k) {
        for (int j = 0; j <= p; ++j) ders[k][j] *= f;
        f *= (p - k);
    }
}

// One side's boundary space of the subspace corrected mass smoother
// (tildeSpaceBasis_oneside, src/gsSolver/gsPatchPreconditionersCreator.hpp:
// 368-423): the p - b functions at the end u (b = 1 if that end is
// eliminated) after skipping the one always in the interior space; the
// matrix D of their odd derivatives of order 1..p-1 at u, scaled by h^i;
// tilde = an orthonormal basis of the null space of D (the
// (p+1)/2 - b directions of smallest singular value), compl = the rest.
// Columns of tilde/compl are vectors over those p - b functions (row-major
// (p-b) x cols).  The null space is taken from the eigenvectors of D^T D
// (ascending eigenvalues), which spans the same subspaces as the SVD.
inline bool tilde_oneside(const std::vector<double> &U, int p, bool left, bool dir, double h,
                          std::vector<double> &tilde, int &nt, std::vector<double> &compl_, int &nc)
{
    const int b = dir ? 1 : 0, m = p - b;
    tilde.clear();
    compl_.clear();
    nt = nc = 0;
    if (m <= 0) return true;
    const double u = left ? U.front() : U.back();
    const int span = find_span(U, p, u);
    std::vector<std::vector<double>> ders;
    bspline_ders_all(U, p, span, u, p - 1, ders);
    std::vector<double> D((size_t)m * m, 0.0);
    const int offset = left ? b : 1;
    for (int i = 1; i < p; i += 2)
        for (int j = 0; j < m; ++j) D[(size_t)(i - b) * m + j] = std::pow(h, i) * ders[i][j + offset];
    std::vector<double> DtD((size_t)m * m, 0.0), w, V;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += D[(size_t)k * m + i] * D[(size_t)k * m + j];
            DtD[(size_t)i * m + j] = s;
        }
    if (!sym_eig(m, DtD, w, V)) return false;
    nt = (p + 1) / 2 - b;
    if (nt < 0) nt = 0;
    nc = m - nt;
    tilde.assign((size_t)m * nt, 0.0);
    compl_.assign((size_t)m * nc, 0.0);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < nt; ++j) tilde[(size_t)i * nt + j] = V[(size_t)i * m + j];
        for (int j = 0; j < nc; ++j) compl_[(size_t)i * nc + j] = V[(size_t)i * m + nt + j];
    }
    return true;
}

// Univariate mass M_ij = int N_i N_j and stiffness K_ij = int N_i' N_j' on
// the parameter line of the open knot vector U (degree p), with the Gauss
// rule (gx, gw) on [-1, 1] mapped to every nonempty span; n = |U| - p - 1.
inline void bspline_1d_matrices(const std::vector<double> &U, int p, const std::vector<double> &gx,
                                const std::vector<double> &gw, std::vector<double> &K, std::vector<double> &M)
{
    const int n = (int)U.size() - p - 1;
    K.assign((size_t)n * n, 0.0);
    M.assign((size_t)n * n, 0.0);
    std::vector<double> N(p + 1), dN(p + 1);
    for (int span = p; span < n; ++span) {
        const double a = U[span], b = U[span + 1];
        if (!(b > a)) continue;
        const double h = 0.5 * (b - a);
        for (size_t q = 0; q < gx.size(); ++q) {
            const double u = a + h * (gx[q] + 1.0), wq = h * gw[q];
            bspline_ders(U, p, span, u, N.data(), dN.data());
            for (int i = 0; i <= p; ++i)
                for (int j = 0; j <= p; ++j) {
                    const size_t at = (size_t)(span - p + i) * n + (span - p + j);
                    M[at] += wq * N[i] * N[j];
                    K[at] += wq * dN[i] * dN[j];
                }
        }
    }
}

}  // namespace dense
}  // namespace gsx
