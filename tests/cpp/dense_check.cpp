// Host checks of gismo_amd/csrc/gsx_dense.h (no GPU): symmetric and
// generalized eigen-decompositions against their defining identities, the
// univariate B-spline mass/stiffness against partition of unity, and the
// one-sided tilde spaces (orthonormal; projector printed for comparison).
// Prints "name value" lines; the generalized eigenvalues of a case are
// printed for comparison with scipy.  Exits nonzero on a failed check.
#include "gsx_dense.h"

#include <cstdio>
#include <random>

using namespace gsx::dense;

static int failures = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                             \
        }                                                           \
    } while (0)

static std::vector<double> uniform_knots(int p, int nel, double a, double b)
{
    std::vector<double> U;
    for (int i = 0; i <= p; ++i) U.push_back(a);
    for (int i = 1; i < nel; ++i) U.push_back(a + (b - a) * i / nel);
    for (int i = 0; i <= p; ++i) U.push_back(b);
    return U;
}

static void gauss(int n, std::vector<double> &x, std::vector<double> &w)
{
    // Newton on the Legendre polynomial from the Chebyshev guesses
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; ++k) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

int main()
{
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    double worst_res = 0.0, worst_orth = 0.0;
    for (int n : {1, 2, 3, 5, 8, 17, 40, 65}) {
        std::vector<double> A((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j <= i; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i] = U(rng);
        std::vector<double> A0 = A, w, V;
        CHECK(sym_eig(n, A, w, V));
        double an = 0.0;
        for (double v : A0) an = std::max(an, std::fabs(v));
        for (int k = 0; k < n; ++k) {
            if (k) CHECK(w[k] >= w[k - 1]);
            for (int i = 0; i < n; ++i) {
                double s = 0.0;
                for (int j = 0; j < n; ++j) s += A0[(size_t)i * n + j] * V[(size_t)j * n + k];
                worst_res = std::max(worst_res, std::fabs(s - w[k] * V[(size_t)i * n + k]) / (an * n));
            }
            for (int l = 0; l < n; ++l) {
                double s = 0.0;
                for (int i = 0; i < n; ++i) s += V[(size_t)i * n + k] * V[(size_t)i * n + l];
                worst_orth = std::max(worst_orth, std::fabs(s - (k == l ? 1.0 : 0.0)));
            }
        }
    }
    CHECK(worst_res < 1e-13);
    CHECK(worst_orth < 1e-12);
    std::printf("sym_eig residual %.3e orthogonality %.3e\n", worst_res, worst_orth);

    // univariate matrices: partition of unity and the generalized problem
    for (int p = 1; p <= 4; ++p) {
        const int nel = 9;
        std::vector<double> Uk = uniform_knots(p, nel, 0.25, 1.75), gx, gw, K, M;
        gauss(p + 1, gx, gw);
        bspline_1d_matrices(Uk, p, gx, gw, K, M);
        const int n = (int)Uk.size() - p - 1;
        double msum = 0.0, krow = 0.0;
        for (int i = 0; i < n; ++i) {
            double r = 0.0;
            for (int j = 0; j < n; ++j) {
                msum += M[(size_t)i * n + j];
                r += K[(size_t)i * n + j];
            }
            krow = std::max(krow, std::fabs(r));
        }
        CHECK(std::fabs(msum - 1.5) < 1e-13);
        CHECK(krow < 1e-12);
        // interior block (both ends eliminated)
        const int m = n - 2;
        std::vector<double> Ki((size_t)m * m), Mi((size_t)m * m), lam, Q;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                Ki[(size_t)i * m + j] = K[(size_t)(i + 1) * n + (j + 1)];
                Mi[(size_t)i * m + j] = M[(size_t)(i + 1) * n + (j + 1)];
            }
        CHECK(gen_sym_eig(m, Ki, Mi, lam, Q));
        double em = 0.0, ek = 0.0, kn = 0.0;
        for (double v : Ki) kn = std::max(kn, std::fabs(v));
        for (int a = 0; a < m; ++a)
            for (int b = 0; b < m; ++b) {
                double sm = 0.0, sk = 0.0;
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < m; ++j) {
                        sm += Q[(size_t)i * m + a] * Mi[(size_t)i * m + j] * Q[(size_t)j * m + b];
                        sk += Q[(size_t)i * m + a] * Ki[(size_t)i * m + j] * Q[(size_t)j * m + b];
                    }
                em = std::max(em, std::fabs(sm - (a == b ? 1.0 : 0.0)));
                ek = std::max(ek, std::fabs(sk - (a == b ? lam[a] : 0.0)) / std::max(1.0, lam[m - 1]));
            }
        CHECK(em < 1e-11);
        CHECK(ek < 1e-11);
        CHECK(lam[0] > 0.0);
        std::printf("gen_eig p %d QtMQ %.3e QtKQ %.3e lambda", p, em, ek);
        for (double l : lam) std::printf(" %.17g", l);
        std::printf("\n");
    }
    // tilde spaces: orthonormal columns; the projector onto the tilde space
    // is printed for comparison with the oracle's SVD
    for (int p = 1; p <= 5; ++p)
        for (int left = 0; left < 2; ++left)
            for (int dir = 0; dir < 2; ++dir) {
                std::vector<double> Uk = uniform_knots(p, 9, 0.25, 1.75), T, Cm;
                int nt = 0, nc = 0;
                CHECK(tilde_oneside(Uk, p, left == 1, dir == 1, 1.5 / 9, T, nt, Cm, nc));
                const int m = nt + nc;
                CHECK(m == std::max(0, p - dir));
                double eo = 0.0;
                for (int a = 0; a < m; ++a)
                    for (int b = 0; b < m; ++b) {
                        double s = 0.0;
                        for (int i = 0; i < m; ++i) {
                            const double va = a < nt ? T[(size_t)i * nt + a] : Cm[(size_t)i * nc + a - nt];
                            const double vb = b < nt ? T[(size_t)i * nt + b] : Cm[(size_t)i * nc + b - nt];
                            s += va * vb;
                        }
                        eo = std::max(eo, std::fabs(s - (a == b ? 1.0 : 0.0)));
                    }
                CHECK(eo < 1e-12);
                std::printf("tilde p %d left %d dir %d nt %d proj", p, left, dir, nt);
                for (int i = 0; i < m; ++i)
                    for (int j = 0; j < m; ++j) {
                        double s = 0.0;
                        for (int a = 0; a < nt; ++a) s += T[(size_t)i * nt + a] * T[(size_t)j * nt + a];
                        std::printf(" %.17g", s);
                    }
                std::printf("\n");
            }
    std::printf("failures %d\n", failures);
    return failures ? 1 : 0;
}
