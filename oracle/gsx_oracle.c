/*
 * gsx_oracle.c -- CPU FP64 restatement of G+Smo's isogeometric Poisson
 * assemble + solve path.  TEST INFRASTRUCTURE ONLY (the parity checker);
 * the product path never links or calls it.  See gsx_oracle.h for what pins
 * each function.  Citations are relative to the G+Smo source tree.
 */
#include "gsx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P1MAX (OC_MAXP + 1)

/* ------------------------------------------------------------------------
 * Gauss-Legendre rule on [-1,1], nodes ascending.
 * Restates the role of gsGaussRule::lookupReference/computeReference
 * (src/gsAssembler/gsGaussRule.hpp:95-121,218-547): the reference tabulates
 * 30-digit nodes; here they are computed by Newton's method on P_n in long
 * double, which rounds to the same doubles (checked against the tables).
 * ---------------------------------------------------------------------- */
void oc_gauss_legendre(int n, double *x, double *w)
{
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int i = 0; i < n; ++i) {
        /* i-th root from the top; stored ascending at n-1-i */
        long double z = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double dp = 0.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = 0.0L;
            for (int k = 1; k <= n; ++k) {
                long double p2 = p1;
                p1 = p0;
                p0 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p2) / (long double)k;
            }
            dp = (long double)n * (z * p0 - p1) / (z * z - 1.0L);
            long double dz = p0 / dp;
            z -= dz;
            if (fabsl(dz) < 1e-30L) break;
        }
        /* recompute the derivative at the converged root */
        {
            long double p0 = 1.0L, p1 = 0.0L;
            for (int k = 1; k <= n; ++k) {
                long double p2 = p1;
                p1 = p0;
                p0 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p2) / (long double)k;
            }
            dp = (long double)n * (z * p0 - p1) / (z * z - 1.0L);
        }
        long double wi = 2.0L / ((1.0L - z * z) * dp * dp);
        if (n % 2 == 1 && i == n / 2) z = 0.0L;
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)wi;
    }
    /* exact symmetry of the reference tables */
    for (int i = 0; i < n / 2; ++i) {
        x[i] = -x[n - 1 - i];
        w[i] = w[n - 1 - i];
    }
}

/* Open uniform knot vector on [a,b] with nelem elements, degree p.
 * Restates gsKnotVector::initUniform (src/gsNurbs/gsKnotVector.hpp:643-688):
 * interior knot i is first + i*h with h = (last-first)/(interior+1). */
void oc_knots_uniform(int nelem, int p, double a, double b, double *knots)
{
    const double h = (b - a) / (double)nelem;
    int k = 0;
    for (int i = 0; i <= p; ++i) knots[k++] = a;
    for (int i = 1; i < nelem; ++i) knots[k++] = a + (double)i * h;
    for (int i = 0; i <= p; ++i) knots[k++] = b;
}

/* Span index (last appearance of the unique knot that starts the element
 * containing u).  Restates gsKnotVector::iFind / uFind
 * (src/gsNurbs/gsKnotVector.hpp:747-791): the domain end belongs to the last
 * element. */
int oc_find_span(const double *knots, int nknots, int p, double u)
{
    const int lo = p, hi = nknots - p - 1; /* domain [knots[lo], knots[hi]] */
    if (u >= knots[hi]) {
        /* last element: last knot strictly below the domain end */
        int s = hi - 1;
        while (s > lo && knots[s] == knots[hi]) --s;
        return s;
    }
    /* upper_bound over the domain knots, minus one */
    int s = lo;
    while (s + 1 < hi && knots[s + 1] <= u) ++s;
    return s;
}

/* Values and derivatives up to nder of the p+1 non-zero B-splines at u in
 * knot span `span`; out is (nder+1) x (p+1) row-major.
 * Restates gsTensorBSplineBasis<1,T>::evalAllDers_into
 * (src/gsNurbs/gsBSplineBasis.hpp:862-1040): the triangular table of
 * basis values and knot differences, then the two-row derivative recursion,
 * then the factor p!/(p-k)!. */
void oc_bspline_ders(const double *U, int p, int span, double u, int nder,
                     double *out)
{
    double tab[P1MAX][P1MAX]; /* tab[r][j]: value of N_{span-j+r, j}; tab[j][r], r<j: knot differences */
    double left[P1MAX], right[P1MAX], arow[2][P1MAX];
    const int p1 = p + 1;

    tab[0][0] = 1.0;
    for (int j = 1; j <= p; ++j) {
        left[j] = u - U[span + 1 - j];
        right[j] = U[span + j] - u;
        double carry = 0.0;
        for (int r = 0; r < j; ++r) {
            tab[j][r] = right[r + 1] + left[j - r];
            const double t = tab[r][j - 1] / tab[j][r];
            tab[r][j] = carry + right[r + 1] * t;
            carry = left[j - r] * t;
        }
        tab[j][j] = carry;
    }
    for (int j = 0; j <= p; ++j) out[j] = tab[j][p];

    for (int r = 0; r <= p; ++r) {
        int cur = 0, nxt = 1;
        arow[0][0] = 1.0;
        for (int k = 1; k <= nder; ++k) {
            double dsum = 0.0;
            const int rk = r - k, pk = p - k;
            if (r >= k) {
                arow[nxt][0] = arow[cur][0] / tab[pk + 1][rk];
                dsum = arow[nxt][0] * tab[rk][pk];
            }
            const int j1 = (rk >= -1) ? 1 : -rk;
            const int j2 = (r - 1 <= pk) ? k - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                arow[nxt][j] = (arow[cur][j] - arow[cur][j - 1]) / tab[pk + 1][rk + j];
                dsum += arow[nxt][j] * tab[rk + j][pk];
            }
            if (r <= pk) {
                arow[nxt][k] = -arow[cur][k - 1] / tab[pk + 1][r];
                dsum += arow[nxt][k] * tab[r][pk];
            }
            out[k * p1 + r] = dsum;
            cur ^= 1;
            nxt ^= 1;
        }
    }
    int fac = p;
    for (int k = 1; k <= nder; ++k) {
        for (int j = 0; j <= p; ++j) out[k * p1 + j] *= (double)fac;
        fac *= p - k;
    }
}

/* ------------------------------------------------------------------------
 * Patch helpers: unique-knot elements (gsTensorDomainIterator,
 * src/gsTensor/gsTensorDomainIterator.h:35-221, lexicographic, dir 0
 * fastest).
 * ---------------------------------------------------------------------- */
int oc_patch_nbasis(const oc_patch *P, int dir)
{
    return P->nknots[dir] - P->p[dir] - 1;
}

/* element e along dir: lower break, upper break, span */
typedef struct { int n; double *lo, *hi; int *span; } elem1d;

static void elem1d_build(const oc_patch *P, int dir, elem1d *E)
{
    const double *U = P->knots[dir];
    const int p = P->p[dir], nk = P->nknots[dir];
    const int lo = p, hi = nk - p - 1;
    E->n = 0;
    E->lo = (double *)malloc(sizeof(double) * nk);
    E->hi = (double *)malloc(sizeof(double) * nk);
    E->span = (int *)malloc(sizeof(int) * nk);
    for (int i = lo; i < hi; ++i) {
        if (U[i + 1] > U[i]) {
            E->lo[E->n] = U[i];
            E->hi[E->n] = U[i + 1];
            E->span[E->n] = i; /* last appearance of U[i] */
            E->n++;
        }
    }
}

static void elem1d_free(elem1d *E)
{
    free(E->lo); free(E->hi); free(E->span);
}

int oc_patch_nelem(const oc_patch *P, int dir)
{
    elem1d E;
    elem1d_build(P, dir, &E);
    int n = E.n;
    elem1d_free(&E);
    return n;
}

/* ------------------------------------------------------------------------
 * Per-element evaluation: tensor basis values/gradients at the Gauss points
 * (gsTensorBasis::evalAllDers_into, src/gsTensor/gsTensorBasis.hpp:634-729:
 * value = N_0*N_1*..., dN/dxi_k = N'_k * prod_{i!=k} N_i in direction
 * order), quadrature mapTo (src/gsAssembler/gsQuadRule.hpp:24-117: node =
 * h(xi+1)+a, weight h*w, tensor weight w_0*w_1*...), geometry map
 * (gsFunction::computeAuxiliaryData, src/gsCore/gsFunction.hpp:604-760:
 * measure |det J|, cramer inverse).
 * ---------------------------------------------------------------------- */
typedef struct {
    int d, nb, nq, p1[3], nq1[3];
    int act[512];           /* patch-local index of each active function */
    double N[512 * 512];    /* nb x nq */
    double dN[512 * 512 * 3]; /* (nb*d + k) x nq layout: [b][k][q] */
    double w[512];          /* tensor weights */
    double x[512 * 3];      /* physical points */
    double J[512 * 9];      /* J[q][r][a] */
    double Jinv[512 * 9];   /* Jinv[q][a][c] */
    double meas[512];
} elemdata;

static double gauss_x[32][32], gauss_w[32][32];
static int gauss_init[32];

static void gauss_get(int n, const double **x, const double **w)
{
    if (!gauss_init[n]) {
        oc_gauss_legendre(n, gauss_x[n], gauss_w[n]);
        gauss_init[n] = 1;
    }
    *x = gauss_x[n];
    *w = gauss_w[n];
}

/* Number of Gauss nodes per direction: floor(quA*p + quB + 0.5) with
 * quA = quB = 1 (src/gsAssembler/gsQuadrature.h:152-171). */
static int nodes_per_dir(int p) { return (int)(1.0 * p + 1 + 0.5); }

static void elem_eval(const oc_patch *P, elem1d *E, int elem, elemdata *D)
{
    const int d = P->d;
    int ei[3] = {0, 0, 0};
    {
        int r = elem;
        for (int k = 0; k < d; ++k) { ei[k] = r % E[k].n; r /= E[k].n; }
    }
    double v1[3][32][P1MAX], d1[3][32][P1MAX], wq1[3][32];
    D->d = d;
    D->nb = 1;
    D->nq = 1;
    for (int k = 0; k < d; ++k) {
        const int p = P->p[k];
        const int nq = nodes_per_dir(p);
        const double *gx, *gw;
        gauss_get(nq, &gx, &gw);
        const double a = E[k].lo[ei[k]], b = E[k].hi[ei[k]];
        const double h = (b - a) / 2.0;
        for (int q = 0; q < nq; ++q) {
            const double u = h * (gx[q] + 1.0) + a;
            wq1[k][q] = (0.0 == h ? 0.5 : h) * gw[q];
            double tmp[2 * P1MAX];
            oc_bspline_ders(P->knots[k], p, E[k].span[ei[k]], u, 1, tmp);
            for (int j = 0; j <= p; ++j) {
                v1[k][q][j] = tmp[j];
                d1[k][q][j] = tmp[(p + 1) + j];
            }
        }
        D->p1[k] = p + 1;
        D->nq1[k] = nq;
        D->nb *= p + 1;
        D->nq *= nq;
    }
    const int nb = D->nb, nq = D->nq;
    /* actives (lexicographic, dir 0 fastest) */
    {
        int nbd[3] = {1, 1, 1};
        for (int k = 0; k < d; ++k) nbd[k] = oc_patch_nbasis(P, k);
        for (int b = 0; b < nb; ++b) {
            int r = b, g = 0, str = 1;
            for (int k = 0; k < d; ++k) {
                const int bk = r % D->p1[k];
                r /= D->p1[k];
                g += (E[k].span[ei[k]] - P->p[k] + bk) * str;
                str *= nbd[k];
            }
            D->act[b] = g;
        }
    }
    /* tensor points, weights, basis */
    for (int q = 0; q < nq; ++q) {
        int qi[3] = {0, 0, 0};
        int r = q;
        for (int k = 0; k < d; ++k) { qi[k] = r % D->nq1[k]; r /= D->nq1[k]; }
        double wt = wq1[0][qi[0]];
        for (int k = 1; k < d; ++k) wt *= wq1[k][qi[k]];
        D->w[q] = wt;
        for (int b = 0; b < nb; ++b) {
            int bi[3] = {0, 0, 0};
            int rb = b;
            for (int k = 0; k < d; ++k) { bi[k] = rb % D->p1[k]; rb /= D->p1[k]; }
            double v = v1[0][qi[0]][bi[0]];
            for (int k = 1; k < d; ++k) v *= v1[k][qi[k]][bi[k]];
            D->N[b * nq + q] = v;
            for (int kk = 0; kk < d; ++kk) {
                double g = d1[kk][qi[kk]][bi[kk]];
                for (int i = 0; i < d; ++i)
                    if (i != kk) g *= v1[i][qi[i]][bi[i]];
                D->dN[(b * d + kk) * nq + q] = g;
            }
        }
    }
    /* geometry: x = sum_b c_b N_b ; J[r][a] = sum_b c_b[r] dN_a(b) */
    for (int q = 0; q < nq; ++q) {
        double xx[3] = {0, 0, 0}, J[3][3] = {{0}};
        for (int b = 0; b < nb; ++b) {
            const double *c = P->ctrl + (size_t)D->act[b] * d;
            for (int r = 0; r < d; ++r) {
                xx[r] += c[r] * D->N[b * nq + q];
                for (int a = 0; a < d; ++a) J[r][a] += c[r] * D->dN[(b * d + a) * nq + q];
            }
        }
        double det, inv[3][3];
        if (d == 2) {
            det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            inv[0][0] = J[1][1] / det;  inv[0][1] = -J[0][1] / det;
            inv[1][0] = -J[1][0] / det; inv[1][1] = J[0][0] / det;
        } else {
            const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
            const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
            const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
            det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
            inv[0][0] = c00 / det;
            inv[1][0] = c01 / det;
            inv[2][0] = c02 / det;
            inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
        for (int r = 0; r < d; ++r) {
            D->x[q * 3 + r] = xx[r];
            for (int a = 0; a < d; ++a) {
                D->J[q * 9 + r * 3 + a] = J[r][a];
                D->Jinv[q * 9 + r * 3 + a] = inv[r][a];
            }
        }
        D->meas[q] = fabs(det);
    }
}

static elemdata *g_ed = NULL;

static elemdata *ed_get(void)
{
    if (!g_ed) g_ed = (elemdata *)malloc(sizeof(elemdata));
    return g_ed;
}

int oc_element_geometry(const oc_patch *P, int elem, double *xq, double *jac,
                        double *meas, double *wq)
{
    elem1d E[3];
    for (int k = 0; k < P->d; ++k) elem1d_build(P, k, &E[k]);
    elemdata *D = ed_get();
    elem_eval(P, E, elem, D);
    const int d = P->d;
    for (int q = 0; q < D->nq; ++q) {
        for (int r = 0; r < d; ++r) {
            xq[q * d + r] = D->x[q * 3 + r];
            for (int a = 0; a < d; ++a) jac[(q * d + r) * d + a] = D->J[q * 9 + r * 3 + a];
        }
        meas[q] = D->meas[q];
        wq[q] = D->w[q];
    }
    for (int k = 0; k < P->d; ++k) elem1d_free(&E[k]);
    return D->nq;
}

/* ------------------------------------------------------------------------
 * DOF mapper: literal restatement of gsDofMapper (src/gsCore/gsDofMapper.cpp:
 * matchDof :99-157, eliminateDof :219-238, finalizeComp :271-340,
 * replaceDofGlobally/mergeDofsGlobally :445-478) with one component, the
 * interfaces matched before the Dirichlet sides are eliminated (gsFeSpace::
 * setup, src/gsAssembler/gsExpressions.h:1091-1121).
 * ---------------------------------------------------------------------- */
typedef struct {
    int *dofs;      /* concatenated per patch */
    int *offset;
    int total;
    int numFree;
    int numCpld;    /* next coupling id (starts at 1) */
    int curElimId;  /* next elimination id (starts at -1) */
} omapper;

static void om_replace(omapper *M, int oldv, int newv)
{
    for (int i = 0; i < M->total; ++i)
        if (M->dofs[i] == oldv) M->dofs[i] = newv;
}

static void om_merge(omapper *M, int a, int b)
{
    if (a != b) {
        if (a < b) { int t = a; a = b; b = t; }
        om_replace(M, a, b);
    }
}

static void om_match(omapper *M, int u, int i, int v, int j)
{
    int *pd1 = &M->dofs[M->offset[u] + i];
    int *pd2 = &M->dofs[M->offset[v] + j];
    int d1 = *pd1, d2 = *pd2;
    if (d1 > d2) {
        int t = d1; d1 = d2; d2 = t;
        int *tp = pd1; pd1 = pd2; pd2 = tp;
        t = u; u = v; v = t;
        t = i; i = j; j = t;
    }
    if (d1 < 0) {
        if (d2 < 0) om_merge(M, d1, d2);
        else if (d2 == 0) *pd2 = d1;
        else om_replace(M, d2, d1);
    } else if (d1 == 0) {
        if (d2 == 0) {
            *pd1 = *pd2 = M->numCpld++;
            if (u == v && i == j) return;
        } else {
            *pd1 = d2;
        }
    } else {
        om_merge(M, d1, d2);
    }
    if ((d1 != d2 && (d1 >= 0 || d2 >= 0)) || (d1 == 0 && d2 == 0)) --M->numFree;
}

static void om_eliminate(omapper *M, int k, int i)
{
    int *pd = &M->dofs[M->offset[k] + i];
    const int old = *pd;
    if (old == 0) {
        --M->numFree;
        *pd = M->curElimId--;
    } else if (old > 0) {
        --M->numFree;
        om_replace(M, old, M->curElimId--);
    }
}

/* Boundary functions of `side` in lexicographic order of the remaining
 * directions (gsTensorBasis::boundaryOffset -> coefSlice,
 * src/gsTensor/gsTensorBasis.hpp:304-311). */
static int side_dofs(const oc_patch *P, int side, int *out)
{
    const int d = P->d, dir = (side - 1) / 2, par = (side - 1) % 2;
    int nb[3] = {1, 1, 1};
    for (int k = 0; k < d; ++k) nb[k] = oc_patch_nbasis(P, k);
    const int fix = par ? nb[dir] - 1 : 0;
    int cnt = 0;
    int tot = 1;
    for (int k = 0; k < d; ++k) if (k != dir) tot *= nb[k];
    for (int t = 0; t < tot; ++t) {
        int idx[3] = {0, 0, 0}, r = t;
        for (int k = 0; k < d; ++k) {
            if (k == dir) { idx[k] = fix; continue; }
            idx[k] = r % nb[k];
            r /= nb[k];
        }
        out[cnt++] = idx[0] + nb[0] * (idx[1] + nb[1] * idx[2]);
    }
    return cnt;
}

int oc_mapper_build(int npatch, const oc_patch *patches, int ndir,
                    const int *dir_sides, int nif, const int *ifaces,
                    int *dofs_out)
{
    omapper M;
    M.offset = (int *)malloc(sizeof(int) * (npatch + 1));
    M.total = 0;
    for (int k = 0; k < npatch; ++k) {
        M.offset[k] = M.total;
        int s = 1;
        for (int j = 0; j < patches[k].d; ++j) s *= oc_patch_nbasis(&patches[k], j);
        M.total += s;
    }
    M.offset[npatch] = M.total;
    M.dofs = (int *)calloc(M.total, sizeof(int));
    M.numFree = M.total;
    M.numCpld = 1;
    M.curElimId = -1;

    int *bufA = (int *)malloc(sizeof(int) * M.total);
    /* interfaces: pair boundary functions through the orientation map
     * (gsTensorBasis::matchWith, src/gsTensor/gsTensorBasis.hpp:973-1030) */
    for (int f = 0; f < nif; ++f) {
        const int *I = ifaces + 8 * f;
        const int p1 = I[0], s1 = I[1], p2 = I[2], s2 = I[3];
        const oc_patch *A = &patches[p1], *B = &patches[p2];
        const int d = A->d;
        const int dir1 = (s1 - 1) / 2, par1 = (s1 - 1) % 2;
        const int dir2 = (s2 - 1) / 2, par2 = (s2 - 1) % 2;
        int na[3] = {1, 1, 1}, nbb[3] = {1, 1, 1};
        for (int k = 0; k < d; ++k) {
            na[k] = oc_patch_nbasis(A, k);
            nbb[k] = oc_patch_nbasis(B, k);
        }
        const int cnt = side_dofs(A, s1, bufA);
        for (int t = 0; t < cnt; ++t) {
            const int la = bufA[t];
            int ia[3] = {la % na[0], (la / na[0]) % na[1], d == 3 ? la / (na[0] * na[1]) : 0};
            int ib[3] = {0, 0, 0};
            ib[dir2] = par2 ? nbb[dir2] - 1 : 0;
            int j = 0;
            for (int k = 0; k < d; ++k) {
                if (k == dir1) continue;
                const int m = I[4 + 2 * j], flip = I[5 + 2 * j];
                ib[m] = flip ? (nbb[m] - 1 - ia[k]) : ia[k];
                ++j;
            }
            (void)par1;
            const int lb = ib[0] + nbb[0] * (ib[1] + nbb[1] * ib[2]);
            om_match(&M, p1, la, p2, lb);
        }
    }
    /* Dirichlet sides (gsDofMapper::markBoundary, gsDofMapper.cpp:188-192) */
    for (int s = 0; s < ndir; ++s) {
        const int k = dir_sides[2 * s], side = dir_sides[2 * s + 1];
        const int cnt = side_dofs(&patches[k], side, bufA);
        for (int t = 0; t < cnt; ++t) om_eliminate(&M, k, bufA[t]);
    }
    free(bufA);

    /* finalizeComp */
    const int nfree = M.numFree;
    int ncpl_ids = M.numCpld - 1;
    int *cplmap = (int *)malloc(sizeof(int) * (ncpl_ids + 1));
    for (int i = 0; i <= ncpl_ids; ++i) cplmap[i] = -1;
    int nelim_ids = -M.curElimId;
    int *elimmap = (int *)malloc(sizeof(int) * (nelim_ids + 1));
    for (int i = 0; i <= nelim_ids; ++i) elimmap[i] = -1;
    int curFree = 0;
    int curElim = nfree;
    int curCpl = 0;
    for (int i = 0; i < M.total; ++i) if (M.dofs[i] == 0) ++curCpl;
    for (int i = 0; i < M.total; ++i) {
        const int t = M.dofs[i];
        if (t == 0) dofs_out[i] = curFree++;
        else if (t < 0) {
            const int id = -t - 1;
            if (elimmap[id] < 0) elimmap[id] = curElim++;
            dofs_out[i] = elimmap[id];
        } else {
            const int id = t - 1;
            if (cplmap[id] < 0) cplmap[id] = curCpl++;
            dofs_out[i] = cplmap[id];
        }
    }
    free(cplmap);
    free(elimmap);
    free(M.dofs);
    free(M.offset);
    return nfree;
}

/* ------------------------------------------------------------------------
 * Sparse matrix built by sorted insertion, like Eigen's coeffRef into a
 * reserved column (gsSparseMatrix, src/gsMatrix/gsSparseMatrix.h:139-448;
 * initSystem reserve gsExprAssembler.h:290-341).
 * ---------------------------------------------------------------------- */
typedef struct { int n, cap; int *c; double *v; } orow;

static void orow_add(orow *R, int col, double val)
{
    /* binary search */
    int lo = 0, hi = R->n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (R->c[mid] < col) lo = mid + 1; else hi = mid;
    }
    if (lo < R->n && R->c[lo] == col) { R->v[lo] += val; return; }
    if (R->n == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 16;
        R->c = (int *)realloc(R->c, sizeof(int) * R->cap);
        R->v = (double *)realloc(R->v, sizeof(double) * R->cap);
    }
    memmove(R->c + lo + 1, R->c + lo, sizeof(int) * (R->n - lo));
    memmove(R->v + lo + 1, R->v + lo, sizeof(double) * (R->n - lo));
    R->c[lo] = col;
    R->v[lo] = 0.0 + val; /* coeffRef inserts 0 then += */
    R->n++;
}

static oc_csr *rows_to_csr(orow *rows, int n)
{
    oc_csr *A = (oc_csr *)malloc(sizeof(oc_csr));
    A->n = n;
    A->rowptr = (int *)malloc(sizeof(int) * (n + 1));
    int64_t nnz = 0;
    for (int i = 0; i < n; ++i) { A->rowptr[i] = (int)nnz; nnz += rows[i].n; }
    A->rowptr[n] = (int)nnz;
    A->nnz = nnz;
    A->col = (int *)malloc(sizeof(int) * (nnz ? nnz : 1));
    A->val = (double *)malloc(sizeof(double) * (nnz ? nnz : 1));
    for (int i = 0; i < n; ++i) {
        memcpy(A->col + A->rowptr[i], rows[i].c, sizeof(int) * rows[i].n);
        memcpy(A->val + A->rowptr[i], rows[i].v, sizeof(double) * rows[i].n);
        free(rows[i].c);
        free(rows[i].v);
    }
    return A;
}

void oc_csr_free(oc_csr *A)
{
    if (!A) return;
    free(A->rowptr); free(A->col); free(A->val); free(A);
}

void oc_csr_spmv(const oc_csr *A, const double *x, double *y)
{
    for (int i = 0; i < A->n; ++i) {
        double s = 0.0;
        for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) s += A->val[k] * x[A->col[k]];
        y[i] = s;
    }
}

static double source_f(const oc_patch *P, int elem, const double *x)
{
    if (P->fkind == 1) {
        const double pi = 3.14159265358979323846;
        double v = (double)P->d * pi * pi;
        for (int r = 0; r < P->d; ++r) v *= sin(pi * x[r]);
        return v;
    }
    return P->felem ? P->felem[elem] : 0.0;
}

/* Assembly of a(u,v) = int kappa grad u . grad v, l(v) = int f v.
 * Restates gsExprAssembler::assemble (src/gsAssembler/gsExprAssembler.h:
 * 754-833) with the expressions igrad(u,G)*igrad(u,G).tr()*meas(G) and
 * u*ff*meas(G): per point eval = (G G^T)*meas [*kappa], quadrature
 * lm = w_0 eval(0) + sum_k w_k eval(k) (:497-505), push (:553-630) with the
 * exact-zero skip (:599) in mode 0; mode 1 inserts every free pair like
 * gsSparseSystem::push (src/gsAssembler/gsSparseSystem.h:972-1010). */
oc_csr *oc_assemble(int npatch, const oc_patch *patches, const int *dofs,
                    int nfree, int mode, double *rhs_out)
{
    return oc_assemble_bc(npatch, patches, dofs, nfree, mode, NULL, rhs_out);
}

/* Local matrix lm (nb x nb) and load lr of element e: per point eval =
 * (G G^T)*meas [*kappa], quadrature lm = w_0 eval(0) + sum_k w_k eval(k)
 * (gsExprAssembler.h:497-505).  G is scratch of nb x 3. */
static void elem_local(const oc_patch *P, elem1d *E, int e, elemdata *D, double *lm, double *lr, double *G)
{
    const int d = P->d;
    elem_eval(P, E, e, D);
    const int nb = D->nb, nq = D->nq;
    const double kap = P->kappa ? P->kappa[e] : 1.0;
    for (int q = 0; q < nq; ++q) {
        /* G = grad(u) * jacInv(G): G[b][c] = sum_a dN_a(b) Jinv[a][c] */
        const double *Ji = D->Jinv + q * 9;
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < d; ++c) {
                double s = 0.0;
                for (int a = 0; a < d; ++a) s += D->dN[(b * d + a) * nq + q] * Ji[a * 3 + c];
                G[b * 3 + c] = s;
            }
        const double wq = D->w[q], ms = D->meas[q];
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j) {
                double s = 0.0;
                for (int c = 0; c < d; ++c) s += G[i * 3 + c] * G[j * 3 + c];
                double ev = s * ms;
                if (P->kappa) ev *= kap;
                if (q == 0) lm[i * nb + j] = wq * ev;
                else lm[i * nb + j] += wq * ev;
            }
        const double fx = source_f(P, e, D->x + q * 3);
        for (int i = 0; i < nb; ++i) {
            const double ev = (D->N[i * nq + q] * fx) * ms;
            if (q == 0) lr[i] = wq * ev;
            else lr[i] += wq * ev;
        }
    }
}

/* push (gsExprAssembler.h:553-630) of one element's lm/lr */
static void elem_push(const elemdata *D, const int *dofs, int off, int nfree, int mode, const double *fixed,
                      const double *lm, const double *lr, orow *rows, double *rhs_out)
{
    const int nb = D->nb;
    for (int i = 0; i < nb; ++i) {
        const int ii = dofs[off + D->act[i]];
        if (ii >= nfree) continue;
        for (int j = 0; j < nb; ++j) {
            const double v = lm[i * nb + j];
            if (mode == 0 && v == 0.0) continue;
            const int jj = dofs[off + D->act[j]];
            if (jj < nfree) orow_add(&rows[ii], jj, v);
            /* eliminated column: rhs_i -= K_ij g_j (:613-619);
             * homogeneous values (fixed NULL) leave the rhs as is */
            else if (fixed && v != 0.0) rhs_out[ii] -= v * fixed[jj - nfree];
        }
        rhs_out[ii] += lr[i];
    }
}

oc_csr *oc_assemble_bc(int npatch, const oc_patch *patches, const int *dofs,
                       int nfree, int mode, const double *fixed, double *rhs_out)
{
    orow *rows = (orow *)calloc(nfree, sizeof(orow));
    for (int i = 0; i < nfree; ++i) rhs_out[i] = 0.0;
    elemdata *D = ed_get();
    static double lm[512 * 512];
    static double lr[512];
    static double G[512 * 3];
    int off = 0;
    for (int pi = 0; pi < npatch; ++pi) {
        const oc_patch *P = &patches[pi];
        const int d = P->d;
        elem1d E[3];
        int nel = 1, nbas = 1;
        for (int k = 0; k < d; ++k) {
            elem1d_build(P, k, &E[k]);
            nel *= E[k].n;
            nbas *= oc_patch_nbasis(P, k);
        }
        for (int e = 0; e < nel; ++e) {
            elem_local(P, E, e, D, lm, lr, G);
            elem_push(D, dofs, off, nfree, mode, fixed, lm, lr, rows, rhs_out);
        }
        for (int k = 0; k < d; ++k) elem1d_free(&E[k]);
        off += nbas;
    }
    return rows_to_csr(rows, nfree);
}

/* The reference's threaded assembly (gsExprAssembler.h:761-796): a team of
 * nthreads threads, thread t takes elements t, t + nt, ... of each patch
 * with its own evaluation buffers, and the push runs under one lock (omp
 * critical, :607).  Used only as the CPU baseline's timing; the order in
 * which elements reach the matrix depends on the schedule, so values agree
 * with oc_assemble to rounding.  Without OpenMP it runs on one thread. */
oc_csr *oc_assemble_par(int npatch, const oc_patch *patches, const int *dofs, int nfree, int mode,
                        double *rhs_out, int nthreads)
{
    orow *rows = (orow *)calloc(nfree, sizeof(orow));
    for (int i = 0; i < nfree; ++i) rhs_out[i] = 0.0;
    for (int n = 1; n < 32; ++n) {   /* Gauss tables before the team starts */
        const double *x, *w;
        gauss_get(n, &x, &w);
    }
    int off = 0;
    for (int pi = 0; pi < npatch; ++pi) {
        const oc_patch *P = &patches[pi];
        const int d = P->d;
        int nel = 1, nbas = 1;
        for (int k = 0; k < d; ++k) {
            nel *= oc_patch_nelem(P, k);
            nbas *= oc_patch_nbasis(P, k);
        }
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
        {
            int tid = 0, nt = 1;
#ifdef _OPENMP
            tid = omp_get_thread_num();
            nt = omp_get_num_threads();
#endif
            elemdata *D = (elemdata *)malloc(sizeof(elemdata));
            double *lm = (double *)malloc(sizeof(double) * 512 * 512);
            double *lr = (double *)malloc(sizeof(double) * 512);
            double *G = (double *)malloc(sizeof(double) * 512 * 3);
            elem1d E[3];
            for (int k = 0; k < d; ++k) elem1d_build(P, k, &E[k]);
            for (int e = tid; e < nel; e += nt) {
                elem_local(P, E, e, D, lm, lr, G);
#pragma omp critical
                elem_push(D, dofs, off, nfree, mode, NULL, lm, lr, rows, rhs_out);
            }
            for (int k = 0; k < d; ++k) elem1d_free(&E[k]);
            free(D);
            free(lm);
            free(lr);
            free(G);
        }
        off += nbas;
    }
    return rows_to_csr(rows, nfree);
}


/* Rows `rows` (free DOF indices) of the single-patch matrix and load vector,
 * computed from the elements in their supports only, with the same element
 * order and arithmetic as oc_assemble (mode 1).  Used to check sampled rows
 * of problems too large for a full CPU assembly.  Output per requested row:
 * up to maxc (column, value) pairs sorted by column; counts[r] entries. */
int oc_assemble_rows(const oc_patch *P, const int *dofs, int nfree, int nrows,
                     const int *rows, int maxc, int *counts, int *cols,
                     double *vals, double *rhs)
{
    const int d = P->d;
    elem1d E[3];
    int nbd[3] = {1, 1, 1}, nbas = 1;
    for (int k = 0; k < d; ++k) {
        elem1d_build(P, k, &E[k]);
        nbd[k] = oc_patch_nbasis(P, k);
        nbas *= nbd[k];
    }
    int *local_of = (int *)malloc(sizeof(int) * nfree);
    for (int i = 0; i < nfree; ++i) local_of[i] = -1;
    for (int l = 0; l < nbas; ++l) if (dofs[l] < nfree) local_of[dofs[l]] = l;
    elemdata *D = ed_get();
    static double G[512 * 3];
    for (int t = 0; t < nrows; ++t) {
        const int r = rows[t];
        orow R = {0, 0, NULL, NULL};
        double racc = 0.0;
        const int l = local_of[r];
        int lc[3] = {l % nbd[0], (l / nbd[0]) % nbd[1], d == 3 ? l / (nbd[0] * nbd[1]) : 0};
        int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
        for (int k = 0; k < d; ++k) {
            /* uniform-span assumption: element e has actives e..e+p */
            elo[k] = lc[k] - P->p[k] < 0 ? 0 : lc[k] - P->p[k];
            ehi[k] = lc[k] < E[k].n - 1 ? lc[k] : E[k].n - 1;
        }
        for (int ez = elo[2]; ez <= ehi[2]; ++ez)
            for (int ey = elo[1]; ey <= ehi[1]; ++ey)
                for (int ex = elo[0]; ex <= ehi[0]; ++ex) {
                    const int e = ex + E[0].n * (ey + E[1].n * ez);
                    elem_eval(P, E, e, D);
                    const int nb = D->nb, nq = D->nq;
                    int i = -1;
                    for (int b = 0; b < nb; ++b) if (D->act[b] == l) i = b;
                    const double kap = P->kappa ? P->kappa[e] : 1.0;
                    double lrow[512], lr = 0.0;
                    for (int q = 0; q < nq; ++q) {
                        const double *Ji = D->Jinv + q * 9;
                        for (int b = 0; b < nb; ++b)
                            for (int c = 0; c < d; ++c) {
                                double s = 0.0;
                                for (int a = 0; a < d; ++a) s += D->dN[(b * d + a) * nq + q] * Ji[a * 3 + c];
                                G[b * 3 + c] = s;
                            }
                        const double wq = D->w[q], ms = D->meas[q];
                        for (int j = 0; j < nb; ++j) {
                            double s = 0.0;
                            for (int c = 0; c < d; ++c) s += G[i * 3 + c] * G[j * 3 + c];
                            double ev = s * ms;
                            if (P->kappa) ev *= kap;
                            if (q == 0) lrow[j] = wq * ev; else lrow[j] += wq * ev;
                        }
                        const double fx = source_f(P, e, D->x + q * 3);
                        const double ev = (D->N[i * nq + q] * fx) * ms;
                        if (q == 0) lr = wq * ev; else lr += wq * ev;
                    }
                    for (int j = 0; j < nb; ++j) {
                        const int jj = dofs[D->act[j]];
                        if (jj < nfree) orow_add(&R, jj, lrow[j]);
                    }
                    racc += lr;
                }
        const int n = R.n < maxc ? R.n : maxc;
        counts[t] = R.n;
        for (int k = 0; k < n; ++k) {
            cols[(size_t)t * maxc + k] = R.c[k];
            vals[(size_t)t * maxc + k] = R.v[k];
        }
        rhs[t] = racc;
        free(R.c);
        free(R.v);
    }
    free(local_of);
    for (int k = 0; k < d; ++k) elem1d_free(&E[k]);
    return 0;
}

/* ------------------------------------------------------------------------
 * Solvers.
 * ---------------------------------------------------------------------- */
static double dot(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double diag_of(const oc_csr *A, int i)
{
    for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k)
        if (A->col[k] == i) return A->val[k];
    return 0.0;
}

/* Preconditioned CG: restates gsConjugateGradient::initIteration/step
 * (src/gsSolver/gsConjugateGradient.hpp:20-93) inside gsIterativeSolver::
 * solve (src/gsSolver/gsIterativeSolver.h:114-124,172-200).  The Jacobi
 * apply is x = tau*r/diag(A) (src/gsSolver/gsSimplePreconditioners.h:
 * 190-202). */
int oc_cg_solve(const oc_csr *A, const double *b, double *x, double tol,
                int maxit, int precond, double tau, const oc_mg *mg,
                double *err_out, double *hist)
{
    const int n = A->n;
    double *r = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n);
    double *p = (double *)malloc(sizeof(double) * n);
    double *t = (double *)malloc(sizeof(double) * n);
    double *dinv = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) dinv[i] = diag_of(A, i);
    /* precond 3: Eigen's DiagonalPreconditioner (external/gsEigen/src/
     * IterativeLinearSolvers/BasicPreconditioners.h:66-77) takes 1 where the
     * diagonal entry is zero or missing */
    if (precond == 3)
        for (int i = 0; i < n; ++i)
            if (dinv[i] == 0.0) dinv[i] = 1.0;
    int it = 0;
    double err = 0.0;
    const double bnorm = sqrt(dot(b, b, n));
    if (bnorm == 0.0) {
        for (int i = 0; i < n; ++i) x[i] = 0.0;
        if (err_out) *err_out = 0.0;
        if (hist) hist[0] = 0.0;
        goto done;
    }
    oc_csr_spmv(A, x, t);
    for (int i = 0; i < n; ++i) r[i] = b[i] - t[i];
    err = sqrt(dot(r, r, n)) / bnorm;
    if (hist) hist[0] = err;
    if (err < tol) goto done;
#define APPLY_PREC(in, out)                                                  \
    do {                                                                     \
        if (precond == 1 || precond == 3) for (int i = 0; i < n; ++i) out[i] = tau * in[i] / dinv[i]; \
        else if (precond == 2) oc_mg_apply(mg, in, out);                     \
        else memcpy(out, in, sizeof(double) * n);                            \
    } while (0)
    APPLY_PREC(r, p);
    double absnew = dot(r, p, n);
    while (it < maxit) {
        ++it;
        oc_csr_spmv(A, p, t);
        const double alpha = absnew / dot(p, t, n);
        for (int i = 0; i < n; ++i) x[i] += alpha * p[i];
        for (int i = 0; i < n; ++i) r[i] -= alpha * t[i];
        err = sqrt(dot(r, r, n)) / bnorm;
        if (hist) hist[it] = err;
        if (err < tol) break;
        APPLY_PREC(r, t);
        const double absold = absnew;
        absnew = dot(r, t, n);
        const double beta = absnew / absold;
        for (int i = 0; i < n; ++i) p[i] = t[i] + beta * p[i];
    }
#undef APPLY_PREC
done:
    if (err_out) *err_out = err;
    free(r); free(z); free(p); free(t); free(dinv);
    return it;
}

/* 1D transfer for one uniform refinement (midpoints, gsKnotVector::
 * getUniformRefinementKnots, src/gsNurbs/gsKnotVector.hpp:1048-1062),
 * obtained by knot insertion applied to the identity coefficients
 * (gsTensorBSplineBasis<1,T>::refine_withTransfer -> gsBoehmRefine,
 * src/gsNurbs/gsBSplineBasis.hpp:1054-1061, src/gsNurbs/gsBoehm.hpp:161-226);
 * here by single-knot insertions (Boehm) one midpoint at a time. */
int oc_transfer_1d(const double *knots, int nknots, int p, double *Tout)
{
    const int nc = nknots - p - 1;
    /* unique midpoints */
    double mids[4096];
    int nm = 0;
    for (int i = p; i < nknots - p - 1; ++i)
        if (knots[i + 1] > knots[i]) mids[nm++] = knots[i] + 1.0 * ((knots[i + 1] - knots[i]) / 2.0);
    const int nf = nc + nm;
    double *U = (double *)malloc(sizeof(double) * (nknots + nm));
    memcpy(U, knots, sizeof(double) * nknots);
    int nk = nknots;
    /* coefficient matrix C: (current n) x nc, starts at identity */
    double *C = (double *)calloc((size_t)nf * nc, sizeof(double));
    double *Cn = (double *)calloc((size_t)nf * nc, sizeof(double));
    for (int i = 0; i < nc; ++i) C[(size_t)i * nc + i] = 1.0;
    int ncur = nc;
    for (int m = 0; m < nm; ++m) {
        const double u = mids[m];
        /* span k: U[k] <= u < U[k+1] */
        int k = p;
        while (k + 1 < nk && U[k + 1] <= u) ++k;
        for (int i = 0; i <= ncur; ++i) {
            double *dst = Cn + (size_t)i * nc;
            if (i <= k - p) {
                memcpy(dst, C + (size_t)i * nc, sizeof(double) * nc);
            } else if (i >= k + 1) {
                memcpy(dst, C + (size_t)(i - 1) * nc, sizeof(double) * nc);
            } else {
                const double a = (u - U[i]) / (U[i + p] - U[i]);
                for (int c = 0; c < nc; ++c)
                    dst[c] = (1.0 - a) * C[(size_t)(i - 1) * nc + c] + a * C[(size_t)i * nc + c];
            }
        }
        ++ncur;
        double *tmp = C; C = Cn; Cn = tmp;
        /* insert u into U after position k */
        memmove(U + k + 2, U + k + 1, sizeof(double) * (nk - k - 1));
        U[k + 1] = u;
        ++nk;
    }
    memcpy(Tout, C, sizeof(double) * (size_t)nf * nc);
    free(U); free(C); free(Cn);
    return nf;
}

/* Transfer between two mapped spaces: gsMultiBasis::combineTransferMatrices
 * (src/gsCore/gsMultiBasis.hpp:137-178): tensor (Kronecker) products of the
 * 1D transfers (gsTensorBasis::uniformRefine_withTransfer,
 * src/gsTensor/gsTensorBasis.hpp:861-880), kept where both indices are
 * free, duplicates resolved by take_first. */
oc_csr *oc_transfer_multipatch(int npatch, const oc_patch *coarse,
                               const int *dofs_c, int nfree_c,
                               const int *dofs_f, int nfree_f)
{
    orow *rows = (orow *)calloc(nfree_f, sizeof(orow));
    int offc = 0, offf = 0;
    for (int pi = 0; pi < npatch; ++pi) {
        const oc_patch *P = &coarse[pi];
        const int d = P->d;
        double *T[3] = {NULL, NULL, NULL};
        int nf[3] = {1, 1, 1}, nc[3] = {1, 1, 1};
        for (int k = 0; k < d; ++k) {
            nc[k] = oc_patch_nbasis(P, k);
            T[k] = (double *)malloc(sizeof(double) * (size_t)(2 * nc[k] + 8) * nc[k]);
            nf[k] = oc_transfer_1d(P->knots[k], P->nknots[k], P->p[k], T[k]);
        }
        const int totf = nf[0] * nf[1] * nf[2];
        const int totc = nc[0] * nc[1] * nc[2];
        for (int fi = 0; fi < totf; ++fi) {
            const int fx = fi % nf[0], fy = (fi / nf[0]) % nf[1], fz = fi / (nf[0] * nf[1]);
            const int gf = dofs_f[offf + fi];
            if (gf >= nfree_f) continue;
            for (int ci = 0; ci < totc; ++ci) {
                const int cx = ci % nc[0], cy = (ci / nc[0]) % nc[1], cz = ci / (nc[0] * nc[1]);
                double v = T[0][(size_t)fx * nc[0] + cx];
                if (v == 0.0) continue;
                v *= T[1][(size_t)fy * nc[1] + cy];
                if (v == 0.0) continue;
                if (d == 3) {
                    v *= T[2][(size_t)fz * nc[2] + cz];
                    if (v == 0.0) continue;
                }
                const int gc = dofs_c[offc + ci];
                if (gc >= nfree_c) continue;
                /* take_first: keep the first value seen for (gf,gc) */
                orow *R = &rows[gf];
                int lo = 0, hi = R->n;
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if (R->c[mid] < gc) lo = mid + 1; else hi = mid;
                }
                if (lo < R->n && R->c[lo] == gc) continue;
                orow_add(R, gc, v);
            }
        }
        for (int k = 0; k < d; ++k) free(T[k]);
        offc += totc;
        offf += totf;
    }
    return rows_to_csr(rows, nfree_f);
}

/* ------------------------------------------------------------------------
 * Multigrid: gsMultiGridOp (src/gsMultiGrid/gsMultiGrid.hpp): Galerkin
 * coarse operators A_c = P^T A_f P (init :70-94), V/W cycle
 * multiGridStep (:148-194) with damped Jacobi smoothing steps
 * x += tau (b - A x)/diag(A) (gsJacobiOp::step, gsSimplePreconditioners.h:
 * 178-186; stepT = step for symmetric smoothers, gsPreconditioner.h:79-84),
 * residual A x - b, correction x -= damping * P c, and an exact coarse solve
 * (dense Cholesky here; the reference uses a sparse LU/Cholesky).
 * ---------------------------------------------------------------------- */
struct oc_mg {
    int nlev, npre, npost;
    double tau, cgc;
    int smoother;       /* 0 Jacobi, 1 Richardson, 2 Chebyshev (Jacobi-preconditioned) */
    double cheb_ratio;  /* Chebyshev interval [lmax / ratio, lmax] */
    double *lmax;       /* per level: 1.1 x power-iteration estimate of lambda_max(D^-1 A) */
    oc_csr **A;   /* per level (owned except finest) */
    oc_csr **P;   /* copies of transfers */
    double **dinv;
    double *chol; /* coarse dense Cholesky factor (lower) */
    int nc;
};

static oc_csr *csr_copy(const oc_csr *S)
{
    oc_csr *A = (oc_csr *)malloc(sizeof(oc_csr));
    A->n = S->n;
    A->nnz = S->nnz;
    A->rowptr = (int *)malloc(sizeof(int) * (S->n + 1));
    memcpy(A->rowptr, S->rowptr, sizeof(int) * (S->n + 1));
    A->col = (int *)malloc(sizeof(int) * (S->nnz ? S->nnz : 1));
    A->val = (double *)malloc(sizeof(double) * (S->nnz ? S->nnz : 1));
    memcpy(A->col, S->col, sizeof(int) * S->nnz);
    memcpy(A->val, S->val, sizeof(double) * S->nnz);
    return A;
}

/* P^T A P with P (nf x nc) */
static oc_csr *galerkin(const oc_csr *A, const oc_csr *P, int nc)
{
    const int nf = A->n;
    /* AP: nf x nc, rowwise dense accumulator */
    double *acc = (double *)calloc(nc, sizeof(double));
    char *mark = (char *)calloc(nc, 1);
    int *list = (int *)malloc(sizeof(int) * nc);
    orow *apr = (orow *)calloc(nf, sizeof(orow));
    for (int i = 0; i < nf; ++i) {
        int nl = 0;
        for (int k = A->rowptr[i]; k < A->rowptr[i + 1]; ++k) {
            const int j = A->col[k];
            const double a = A->val[k];
            for (int m = P->rowptr[j]; m < P->rowptr[j + 1]; ++m) {
                const int c = P->col[m];
                if (!mark[c]) { mark[c] = 1; list[nl++] = c; acc[c] = 0.0; }
                acc[c] += a * P->val[m];
            }
        }
        for (int t = 0; t < nl; ++t) { orow_add(&apr[i], list[t], acc[list[t]]); mark[list[t]] = 0; }
    }
    /* P^T (AP): coarse rows */
    orow *cr = (orow *)calloc(nc, sizeof(orow));
    for (int i = 0; i < nf; ++i) {
        for (int m = P->rowptr[i]; m < P->rowptr[i + 1]; ++m) {
            const int c = P->col[m];
            const double pv = P->val[m];
            for (int t = 0; t < apr[i].n; ++t) orow_add(&cr[c], apr[i].c[t], pv * apr[i].v[t]);
        }
        free(apr[i].c); free(apr[i].v);
    }
    free(apr); free(acc); free(mark); free(list);
    return rows_to_csr(cr, nc);
}

oc_mg *oc_mg_create(const oc_csr *A_fine, int nlev, const oc_csr *const *P,
                    int npre, int npost, double damping, double cgc_damping)
{
    oc_mg *mg = (oc_mg *)calloc(1, sizeof(oc_mg));
    mg->nlev = nlev;
    mg->npre = npre;
    mg->npost = npost;
    mg->tau = damping;
    mg->cgc = cgc_damping;
    mg->A = (oc_csr **)calloc(nlev, sizeof(oc_csr *));
    mg->P = (oc_csr **)calloc(nlev, sizeof(oc_csr *));
    mg->dinv = (double **)calloc(nlev, sizeof(double *));
    mg->lmax = (double *)calloc(nlev, sizeof(double));
    mg->cheb_ratio = 30.0;
    mg->A[nlev - 1] = csr_copy(A_fine);
    for (int l = nlev - 2; l >= 0; --l) {
        mg->P[l] = csr_copy(P[l]);
        /* coarse size = max col + 1 over P rows is not reliable; use the
         * number of columns implied by transfers: scan */
        int nc = 0;
        for (int64_t k = 0; k < P[l]->nnz; ++k) if (P[l]->col[k] + 1 > nc) nc = P[l]->col[k] + 1;
        mg->A[l] = galerkin(mg->A[l + 1], P[l], nc);
    }
    for (int l = 0; l < nlev; ++l) {
        const int n = mg->A[l]->n;
        mg->dinv[l] = (double *)malloc(sizeof(double) * n);
        for (int i = 0; i < n; ++i) mg->dinv[l][i] = diag_of(mg->A[l], i);
    }
    /* coarse dense Cholesky */
    const int n = mg->A[0]->n;
    mg->nc = n;
    double *L = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int k = mg->A[0]->rowptr[i]; k < mg->A[0]->rowptr[i + 1]; ++k)
            L[(size_t)i * n + mg->A[0]->col[k]] = mg->A[0]->val[k];
    for (int j = 0; j < n; ++j) {
        double s = L[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        const double djj = sqrt(s);
        L[(size_t)j * n + j] = djj;
        for (int i = j + 1; i < n; ++i) {
            double t = L[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = t / djj;
        }
    }
    mg->chol = L;
    return mg;
}

static void coarse_solve(const oc_mg *mg, const double *b, double *x)
{
    const int n = mg->nc;
    const double *L = mg->chol;
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * x[k];
        x[i] = s / L[(size_t)i * n + i];
    }
}

static void jacobi_step(const oc_mg *mg, int l, const double *b, double *x, double *tmp)
{
    const oc_csr *A = mg->A[l];
    oc_csr_spmv(A, x, tmp);
    for (int i = 0; i < A->n; ++i) x[i] += mg->tau * (b[i] - tmp[i]) / mg->dinv[l][i];
}

/* lambda_max(D^-1 A) by power iteration: v0_i = 1 / (1 + i mod 5),
 * normalised; 20 times w = D^-1 A v, lambda = ||w||, v = w / lambda. */
static double power_lmax(const oc_csr *A, const double *dg)
{
    const int n = A->n;
    double *v = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * n);
    double nv = 0.0;
    for (int i = 0; i < n; ++i) { v[i] = 1.0 / (1.0 + (double)(i % 5)); nv += v[i] * v[i]; }
    nv = sqrt(nv);
    for (int i = 0; i < n; ++i) v[i] /= nv;
    double lam = 0.0;
    for (int it = 0; it < 20; ++it) {
        oc_csr_spmv(A, v, w);
        double s = 0.0;
        for (int i = 0; i < n; ++i) { w[i] /= dg[i]; s += w[i] * w[i]; }
        lam = sqrt(s);
        for (int i = 0; i < n; ++i) v[i] = w[i] / lam;
    }
    free(v); free(w);
    return lam;
}

/* Smoothing sweeps of level l.  Jacobi: x += tau D^-1 (b - A x)
 * (gsJacobiOp::step, src/gsSolver/gsSimplePreconditioners.h:178-202);
 * Richardson: x += tau (b - A x) (gsRichardsonOp::step, same file :60-90);
 * Chebyshev: m-step Jacobi-preconditioned Chebyshev iteration on
 * [lmax / ratio, lmax] (no reference counterpart):
 *   theta = (l1 + l0) / 2, delta = (l1 - l0) / 2, sigma = theta / delta,
 *   rho_1 = 1 / sigma, d_1 = D^-1 r_0 / theta,
 *   rho_{k+1} = 1 / (2 sigma - rho_k),
 *   d_{k+1} = rho_{k+1} rho_k d_k + (2 rho_{k+1} / delta) D^-1 r_k,
 *   x_{k+1} = x_k + d_{k+1}. */
/* One Gauss-Seidel sweep, rows ascending (internal::gaussSeidelSweep) or
 * descending (internal::reverseGaussSeidelSweep), src/gsSolver/
 * gsSimplePreconditioners.hpp:21-68: x_i += (f_i - sum_j a_ij x_j) / a_ii
 * with the row's sum in storage order over the current x. */
void oc_gs_sweep(const oc_csr *A, const double *b, double *x, int reverse)
{
    const int n = A->n;
    for (int k = 0; k < n; ++k) {
        const int i = reverse ? n - 1 - k : k;
        double diag = 0.0, sum = 0.0;
        for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; ++j) {
            sum += A->val[j] * x[A->col[j]];
            if (A->col[j] == i) diag = A->val[j];
        }
        x[i] += (b[i] - sum) / diag;
    }
}

/* post != 0: post-smoothing, which gsMultiGridOp runs with stepT
 * (SymmSmooth, gsMultiGrid.hpp:192-199): the reverse sweep for
 * gsGaussSeidelOp<forward> (gsSimplePreconditioners.h:318-329). */
static void smooth(const oc_mg *mg, int l, const double *b, double *x, int m, double *tmp, double *d, int post)
{
    const oc_csr *A = mg->A[l];
    const int n = A->n;
    if (mg->smoother == 3) {
        for (int s = 0; s < m; ++s) oc_gs_sweep(A, b, x, post);
        return;
    }
    if (mg->smoother == 0) {
        for (int s = 0; s < m; ++s) jacobi_step(mg, l, b, x, tmp);
        return;
    }
    if (mg->smoother == 1) {
        for (int s = 0; s < m; ++s) {
            oc_csr_spmv(A, x, tmp);
            for (int i = 0; i < n; ++i) x[i] += mg->tau * (b[i] - tmp[i]);
        }
        return;
    }
    const double l1 = mg->lmax[l], l0 = l1 / mg->cheb_ratio;
    const double theta = 0.5 * (l1 + l0), delta = 0.5 * (l1 - l0), sigma = theta / delta;
    double rho = 1.0 / sigma;
    for (int s = 0; s < m; ++s) {
        oc_csr_spmv(A, x, tmp);
        double c1, c2;
        if (s == 0) { c1 = 0.0; c2 = 1.0 / theta; }
        else {
            const double rn = 1.0 / (2.0 * sigma - rho);
            c1 = rn * rho;
            c2 = 2.0 * rn / delta;
            rho = rn;
        }
        for (int i = 0; i < n; ++i) {
            const double ri = (b[i] - tmp[i]) / mg->dinv[l][i];
            d[i] = (s == 0 ? 0.0 : c1 * d[i]) + c2 * ri;
            x[i] += d[i];
        }
    }
}

void oc_mg_set_smoother(oc_mg *mg, int kind, double param)
{
    mg->smoother = kind;
    if (kind == 2) {
        mg->cheb_ratio = param > 1.0 ? param : 30.0;
        for (int l = 1; l < mg->nlev; ++l) mg->lmax[l] = 1.1 * power_lmax(mg->A[l], mg->dinv[l]);
    }
}

double oc_mg_lmax(const oc_mg *mg, int lvl) { return mg->lmax[lvl]; }

static void mg_step(const oc_mg *mg, int l, const double *b, double *x)
{
    if (l == 0) { coarse_solve(mg, b, x); return; }
    const int nf = mg->A[l]->n, nc = mg->A[l - 1]->n;
    double *tmp = (double *)malloc(sizeof(double) * nf);
    double *res = (double *)malloc(sizeof(double) * nf);
    double *cr = (double *)calloc(nc, sizeof(double));
    double *cc = (double *)calloc(nc, sizeof(double));
    double *dd = (double *)calloc(nf, sizeof(double));
    smooth(mg, l, b, x, mg->npre, tmp, dd, 0);
    oc_csr_spmv(mg->A[l], x, res);
    for (int i = 0; i < nf; ++i) res[i] -= b[i];
    const oc_csr *P = mg->P[l - 1];
    for (int i = 0; i < nf; ++i)
        for (int m = P->rowptr[i]; m < P->rowptr[i + 1]; ++m) cr[P->col[m]] += P->val[m] * res[i];
    mg_step(mg, l - 1, cr, cc);
    for (int i = 0; i < nf; ++i) {
        double s = 0.0;
        for (int m = P->rowptr[i]; m < P->rowptr[i + 1]; ++m) s += P->val[m] * cc[P->col[m]];
        x[i] -= mg->cgc * s;
    }
    smooth(mg, l, b, x, mg->npost, tmp, dd, 1);
    free(tmp); free(res); free(cr); free(cc); free(dd);
}

/* gsPreconditionerOp::apply: x = 0, one step (src/gsSolver/gsPreconditioner.h:79-84) */
void oc_mg_apply(const oc_mg *mg, const double *r, double *z)
{
    const int n = mg->A[mg->nlev - 1]->n;
    for (int i = 0; i < n; ++i) z[i] = 0.0;
    mg_step(mg, mg->nlev - 1, r, z);
}

const oc_csr *oc_mg_matrix(const oc_mg *mg, int lvl) { return mg->A[lvl]; }

void oc_mg_free(oc_mg *mg)
{
    if (!mg) return;
    for (int l = 0; l < mg->nlev; ++l) {
        oc_csr_free(mg->A[l]);
        if (mg->P[l]) oc_csr_free(mg->P[l]);
        free(mg->dinv[l]);
    }
    free(mg->A); free(mg->P); free(mg->dinv); free(mg->lmax); free(mg->chol); free(mg);
}

/* L2 error against u = prod sin(pi x_r): gsExprEvaluator::integral of
 * (u_h - u)^2 meas (src/gsAssembler/gsExprEvaluator.h:382-450,
 * examples/poisson2_example.cpp:173-176). */
/* H1 seminorm error |u_h - u|_{H1} against u = prod sin(pi x_r) with the
 * assembly quadrature: the second term of the error printed by
 * examples/poisson2_example.cpp:173-176 (igrad(u_ex) - igrad(u_sol, G)),
 * integrated as gsExprEvaluator::integral (src/gsAssembler/gsExprEvaluator.h:
 * 382-450); physical gradients as in the stiffness (gsExpressions.h:
 * 3263-3297). */
double oc_h1_error(int npatch, const oc_patch *patches, const int *dofs,
                   int nfree, const double *sol)
{
    return oc_h1_error_bc(npatch, patches, dofs, nfree, sol, NULL);
}

double oc_h1_error_bc(int npatch, const oc_patch *patches, const int *dofs,
                      int nfree, const double *sol, const double *fixed)
{
    const double pi = 3.14159265358979323846;
    elemdata *D = ed_get();
    double total = 0.0;
    int off = 0;
    for (int pi_ = 0; pi_ < npatch; ++pi_) {
        const oc_patch *P = &patches[pi_];
        const int d = P->d;
        elem1d E[3];
        int nel = 1, nbas = 1;
        for (int k = 0; k < d; ++k) {
            elem1d_build(P, k, &E[k]);
            nel *= E[k].n;
            nbas *= oc_patch_nbasis(P, k);
        }
        for (int e = 0; e < nel; ++e) {
            elem_eval(P, E, e, D);
            double esum = 0.0;
            for (int q = 0; q < D->nq; ++q) {
                double gp[3] = {0, 0, 0};   /* parametric gradient of u_h */
                for (int b = 0; b < D->nb; ++b) {
                    const int g = dofs[off + D->act[b]];
                    if (g >= nfree && !fixed) continue;
                    const double cb = g < nfree ? sol[g] : fixed[g - nfree];
                    for (int a = 0; a < d; ++a) gp[a] += cb * D->dN[(b * d + a) * D->nq + q];
                }
                const double *Ji = D->Jinv + q * 9;
                double diff2 = 0.0;
                for (int c = 0; c < d; ++c) {
                    double gh = 0.0;
                    for (int a = 0; a < d; ++a) gh += gp[a] * Ji[a * 3 + c];
                    double ge = pi * cos(pi * D->x[q * 3 + c]);
                    for (int r = 0; r < d; ++r)
                        if (r != c) ge *= sin(pi * D->x[q * 3 + r]);
                    diff2 += (gh - ge) * (gh - ge);
                }
                esum += D->w[q] * (diff2 * D->meas[q]);
            }
            total += esum;
        }
        for (int k = 0; k < d; ++k) elem1d_free(&E[k]);
        off += nbas;
    }
    return sqrt(total);
}

double oc_l2_error(int npatch, const oc_patch *patches, const int *dofs,
                   int nfree, const double *sol)
{
    return oc_l2_error_bc(npatch, patches, dofs, nfree, sol, NULL);
}

double oc_l2_error_bc(int npatch, const oc_patch *patches, const int *dofs,
                      int nfree, const double *sol, const double *fixed)
{
    const double pi = 3.14159265358979323846;
    elemdata *D = ed_get();
    double total = 0.0;
    int off = 0;
    for (int pi_ = 0; pi_ < npatch; ++pi_) {
        const oc_patch *P = &patches[pi_];
        elem1d E[3];
        int nel = 1, nbas = 1;
        for (int k = 0; k < P->d; ++k) {
            elem1d_build(P, k, &E[k]);
            nel *= E[k].n;
            nbas *= oc_patch_nbasis(P, k);
        }
        for (int e = 0; e < nel; ++e) {
            elem_eval(P, E, e, D);
            double esum = 0.0;
            for (int q = 0; q < D->nq; ++q) {
                double uh = 0.0;
                for (int b = 0; b < D->nb; ++b) {
                    const int g = dofs[off + D->act[b]];
                    if (g < nfree) uh += sol[g] * D->N[b * D->nq + q];
                    else if (fixed) uh += fixed[g - nfree] * D->N[b * D->nq + q];
                }
                double ue = 1.0;
                for (int r = 0; r < P->d; ++r) ue *= sin(pi * D->x[q * 3 + r]);
                const double diff = uh - ue;
                esum += D->w[q] * (diff * diff * D->meas[q]);
            }
            total += esum;
        }
        for (int k = 0; k < P->d; ++k) elem1d_free(&E[k]);
        off += nbas;
    }
    return sqrt(total);
}


/* ------------------------------------------------------------------------
 * Boundary data (gsBoundaryConditions with the poisson2_example call
 * sequence).  Quadrature on one side element: gsGaussRule(basis, 1, 1, dir)
 * (src/gsAssembler/gsGaussRule.hpp:23-60): p+1 Gauss nodes per tangential
 * direction, one node on the side in direction dir whose weight 2 is scaled
 * by 0.5 for the zero-length interval in gsQuadRule::mapTo
 * (src/gsAssembler/gsQuadRule.h:177-201).  Points lexicographic, direction 0
 * fastest.  Side elements in the order of the side's domain iterator
 * (lexicographic over the tangential directions).
 * ---------------------------------------------------------------------- */
static double bc_g(const oc_bc *b, int d, const double *x)
{
    if (b->kind == 1) {
        const double pi = 3.14159265358979323846;
        double v = 1.0;
        for (int r = 0; r < d; ++r) v *= sin(pi * x[r]);
        return v;
    }
    if (b->kind == 2) {
        double v = b->value[0];
        for (int r = 0; r < d; ++r) v += b->value[1 + r] * x[r];
        return v;
    }
    return 0.0;
}

static void bc_grad(const oc_bc *b, int d, const double *x, double *f)
{
    const double pi = 3.14159265358979323846;
    for (int c = 0; c < d; ++c) {
        if (b->kind == 1) {
            double v = pi * cos(pi * x[c]);
            for (int r = 0; r < d; ++r)
                if (r != c) v *= sin(pi * x[r]);
            f[c] = v;
        } else if (b->kind == 2) {
            f[c] = b->value[1 + c];
        } else {
            f[c] = 0.0;
        }
    }
}

typedef struct {
    int nb, nq;
    int act[512];
    double N[512 * 32];     /* [b][q] */
    double x[32 * 3], J[32 * 9], det[32], w[32];
} bdata;

static void bdry_eval(const oc_patch *P, elem1d *E, int dir, int par, const int *ei, bdata *B)
{
    const int d = P->d;
    double v1[3][32][P1MAX], d1[3][32][P1MAX], wq1[3][32];
    int p1[3] = {1, 1, 1}, nq1[3] = {1, 1, 1};
    B->nb = 1;
    B->nq = 1;
    for (int k = 0; k < d; ++k) {
        const int p = P->p[k];
        const double a = E[k].lo[ei[k]], b = E[k].hi[ei[k]];
        p1[k] = p + 1;
        if (k == dir) {
            nq1[k] = 1;
            double tmp[2 * P1MAX];
            oc_bspline_ders(P->knots[k], p, E[k].span[ei[k]], par ? b : a, 1, tmp);
            for (int j = 0; j <= p; ++j) {
                v1[k][0][j] = tmp[j];
                d1[k][0][j] = tmp[(p + 1) + j];
            }
            wq1[k][0] = 0.5 * 2.0;
        } else {
            const int nq = nodes_per_dir(p);
            const double *gx, *gw;
            gauss_get(nq, &gx, &gw);
            const double h = (b - a) / 2.0;
            nq1[k] = nq;
            for (int q = 0; q < nq; ++q) {
                const double u = h * (gx[q] + 1.0) + a;
                wq1[k][q] = (0.0 == h ? 0.5 : h) * gw[q];
                double tmp[2 * P1MAX];
                oc_bspline_ders(P->knots[k], p, E[k].span[ei[k]], u, 1, tmp);
                for (int j = 0; j <= p; ++j) {
                    v1[k][q][j] = tmp[j];
                    d1[k][q][j] = tmp[(p + 1) + j];
                }
            }
        }
        B->nb *= p1[k];
        B->nq *= nq1[k];
    }
    int nbd[3] = {1, 1, 1};
    for (int k = 0; k < d; ++k) nbd[k] = oc_patch_nbasis(P, k);
    for (int bb = 0; bb < B->nb; ++bb) {
        int r = bb, g = 0, str = 1;
        for (int k = 0; k < d; ++k) {
            const int bk = r % p1[k];
            r /= p1[k];
            g += (E[k].span[ei[k]] - P->p[k] + bk) * str;
            str *= nbd[k];
        }
        B->act[bb] = g;
    }
    for (int q = 0; q < B->nq; ++q) {
        int qi[3] = {0, 0, 0}, r = q;
        for (int k = 0; k < d; ++k) { qi[k] = r % nq1[k]; r /= nq1[k]; }
        double wt = wq1[0][qi[0]];
        for (int k = 1; k < d; ++k) wt *= wq1[k][qi[k]];
        B->w[q] = wt;
        double xx[3] = {0, 0, 0}, J[3][3] = {{0}};
        for (int bb = 0; bb < B->nb; ++bb) {
            int bi[3] = {0, 0, 0}, rb = bb;
            for (int k = 0; k < d; ++k) { bi[k] = rb % p1[k]; rb /= p1[k]; }
            double v = v1[0][qi[0]][bi[0]];
            for (int k = 1; k < d; ++k) v *= v1[k][qi[k]][bi[k]];
            B->N[bb * 32 + q] = v;
            const double *c = P->ctrl + (size_t)B->act[bb] * d;
            for (int kk = 0; kk < d; ++kk) {
                double g = d1[kk][qi[kk]][bi[kk]];
                for (int i = 0; i < d; ++i)
                    if (i != kk) g *= v1[i][qi[i]][bi[i]];
                for (int rr = 0; rr < d; ++rr) J[rr][kk] += c[rr] * g;
            }
            for (int rr = 0; rr < d; ++rr) xx[rr] += c[rr] * v;
        }
        double det;
        if (d == 2) det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        else det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                   J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        B->det[q] = det;
        for (int rr = 0; rr < d; ++rr) {
            B->x[q * 3 + rr] = xx[rr];
            for (int a = 0; a < d; ++a) B->J[q * 9 + rr * 3 + a] = J[rr][a];
        }
    }
}

/* visit the elements along `side` of patch P in domain-iterator order */
typedef void (*side_visit)(const oc_patch *P, elem1d *E, int dir, int par, const int *ei, void *ctx);
static void for_side_elements(const oc_patch *P, int side, side_visit f, void *ctx)
{
    const int d = P->d, dir = (side - 1) / 2, par = (side - 1) % 2;
    elem1d E[3];
    int cnt = 1;
    for (int k = 0; k < d; ++k) {
        elem1d_build(P, k, &E[k]);
        if (k != dir) cnt *= E[k].n;
    }
    for (int t = 0; t < cnt; ++t) {
        int ei[3] = {0, 0, 0}, r = t;
        for (int k = 0; k < d; ++k) {
            if (k == dir) { ei[k] = par ? E[k].n - 1 : 0; continue; }
            ei[k] = r % E[k].n;
            r /= E[k].n;
        }
        f(P, E, dir, par, ei, ctx);
    }
    for (int k = 0; k < d; ++k) elem1d_free(&E[k]);
}

typedef struct {
    const oc_bc *bc;
    const int *dofs;     /* of this patch (offset applied) */
    int nfree;
    orow *rows;          /* boundary-index rows (Dirichlet) */
    double *rhs;
} bctx;

static bdata *g_bd = NULL;

/* gsDirichletValuesByL2Projection (src/gsAssembler/gsDirichletValues.h:
 * 257-435): weight_k = w_k * measure_k with the map data's measure |det J|
 * (computeMap without a side, src/gsCore/gsFunction.hpp:702-716); entries
 * (weight N_i) N_j and rhs (weight N_i) g over the eliminated actives. */
static void dir_visit(const oc_patch *P, elem1d *E, int dir, int par, const int *ei, void *vctx)
{
    bctx *C = (bctx *)vctx;
    bdata *B = g_bd;
    bdry_eval(P, E, dir, par, ei, B);
    for (int q = 0; q < B->nq; ++q) {
        const double wk = B->w[q] * fabs(B->det[q]);
        const double g = bc_g(C->bc, P->d, B->x + q * 3);
        for (int i = 0; i < B->nb; ++i) {
            const int ii = C->dofs[B->act[i]];
            if (ii < C->nfree) continue;
            for (int j = 0; j < B->nb; ++j) {
                const int jj = C->dofs[B->act[j]];
                if (jj < C->nfree) continue;
                orow_add(&C->rows[ii - C->nfree], jj - C->nfree, wk * B->N[i * 32 + q] * B->N[j * 32 + q]);
            }
            C->rhs[ii - C->nfree] += wk * B->N[i * 32 + q] * g;
        }
    }
}

int oc_dirichlet_l2(int npatch, const oc_patch *patches, const int *dofs, int nfree, int nfix,
                    int nbc, const oc_bc *bcs, double *fixed_out)
{
    if (!g_bd) g_bd = (bdata *)malloc(sizeof(bdata));
    orow *rows = (orow *)calloc(nfix > 0 ? nfix : 1, sizeof(orow));
    double *rhs = (double *)calloc(nfix > 0 ? nfix : 1, sizeof(double));
    int *off = (int *)calloc(npatch + 1, sizeof(int));
    for (int k = 0; k < npatch; ++k) {
        int nb = 1;
        for (int r = 0; r < patches[k].d; ++r) nb *= oc_patch_nbasis(&patches[k], r);
        off[k + 1] = off[k] + nb;
    }
    for (int b = 0; b < nbc; ++b) {
        if (bcs[b].type != 1) continue;
        bctx C = {&bcs[b], dofs + off[bcs[b].patch], nfree, rows, rhs};
        for_side_elements(&patches[bcs[b].patch], bcs[b].side, dir_visit, &C);
    }
    oc_csr *M = rows_to_csr(rows, nfix);
    free(rows);
    for (int i = 0; i < nfix; ++i) fixed_out[i] = 0.0;
    double err = 0.0;
    /* gsSparseSolver::CGDiagonal with Eigen's default tolerance (machine
     * epsilon): Jacobi-preconditioned CG to rounding level.  Functions on
     * no projected side (gsAssembler's variant skips homogeneous sides) have
     * zero or no rows; Eigen's preconditioner takes 1 there and, from
     * x = 0 with a zero right-hand side in those rows, they stay 0. */
    oc_cg_solve(M, rhs, fixed_out, 1e-15, 2 * nfix + 100, 3, 1.0, NULL, &err, NULL);
    oc_csr_free(M);
    free(rhs);
    free(off);
    return 0;
}

/* gsExprAssembler::assembleBdr(bc.get("Neumann"), u * g_N.tr() * nv(G))
 * (src/gsAssembler/gsExprAssembler.h:835-895): per point N_i (g_N . nv),
 * nv the outer normal of gsFunction::computeMap (src/gsCore/gsFunction.hpp:
 * 613-680), quadrature lm = sum_k w_k eval(k), pushed into the free rows. */
static void neu_visit(const oc_patch *P, elem1d *E, int dir, int par, const int *ei, void *vctx)
{
    bctx *C = (bctx *)vctx;
    bdata *B = g_bd;
    const int d = P->d;
    bdry_eval(P, E, dir, par, ei, B);
    double dsg = 0.0;
    for (int q = 0; q < B->nq; ++q)
        if (fabs(B->det[q]) >= 1e-7) { dsg = B->det[q] < 0 ? -1.0 : 1.0; break; }
    if (dsg == 0.0) dsg = 1.0;
    const int s = C->bc->side;
    const double so = ((s + (s + 1) / 2) % 2) ? 1.0 : -1.0;
    static double lr[512];
    for (int i = 0; i < B->nb; ++i) lr[i] = 0.0;
    for (int q = 0; q < B->nq; ++q) {
        const double *J = B->J + q * 9;   /* J[r][a] = dx_r/dxi_a; jacT[a][r] = J[r][a] */
        double nv[3] = {0, 0, 0}, alt = so * dsg;
        for (int c = 0; c < d; ++c) {
            double m;
            if (d == 2) {
                m = J[(1 - c) * 3 + (1 - dir)];
            } else {
                int ra[2], cb[2], na = 0, nb = 0;
                for (int a = 0; a < 3; ++a) if (a != dir) ra[na++] = a;
                for (int b = 0; b < 3; ++b) if (b != c) cb[nb++] = b;
                m = J[cb[0] * 3 + ra[0]] * J[cb[1] * 3 + ra[1]] - J[cb[1] * 3 + ra[0]] * J[cb[0] * 3 + ra[1]];
            }
            nv[c] = alt * m;
            alt = -alt;
        }
        double f[3];
        bc_grad(C->bc, d, B->x + q * 3, f);
        double gn = 0.0;
        for (int c = 0; c < d; ++c) gn += f[c] * nv[c];
        for (int i = 0; i < B->nb; ++i) lr[i] += B->w[q] * (B->N[i * 32 + q] * gn);
    }
    for (int i = 0; i < B->nb; ++i) {
        const int ii = C->dofs[B->act[i]];
        if (ii < C->nfree) C->rhs[ii] += lr[i];
    }
}

void oc_neumann(int npatch, const oc_patch *patches, const int *dofs, int nfree, int nbc,
                const oc_bc *bcs, double *rhs)
{
    if (!g_bd) g_bd = (bdata *)malloc(sizeof(bdata));
    int *off = (int *)calloc(npatch + 1, sizeof(int));
    for (int k = 0; k < npatch; ++k) {
        int nb = 1;
        for (int r = 0; r < patches[k].d; ++r) nb *= oc_patch_nbasis(&patches[k], r);
        off[k + 1] = off[k] + nb;
    }
    for (int b = 0; b < nbc; ++b) {
        if (bcs[b].type != 2) continue;
        bctx C = {&bcs[b], dofs + off[bcs[b].patch], nfree, NULL, rhs};
        for_side_elements(&patches[bcs[b].patch], bcs[b].side, neu_visit, &C);
    }
    free(off);
}
