// Assembly with the geometry in its own basis, and the mass-matrix transfers
// of p-multigrid (gfx950).
//
// A level of a p-multigrid hierarchy discretises with a basis U (degree q,
// possibly a coarser mesh) while the geometry map stays the fine patch
// (degree p, fine mesh): gsCDRAssembler(mp, bases[i], ...) in
// examples/pMultiGrid_example.cpp:445-460.  The quadrature follows U (q + 1
// Gauss points per direction on U's elements, gsQuadrature::get with
// quA = quB = 1, src/gsAssembler/gsQuadrature.h:24-60); at every point the
// geometry's knot span is found by search, so the map is evaluated exactly
// whatever its degree and mesh.
//
// Transfers between two levels of different degree are L2 projections with
// lumped mass (pMultiGrid_example.cpp:380-396, assembleLumpedMass :803-826,
// assembleMixedMass :855-880):
//   M_fc(i, j) = int phi_i psi_j dx  (fine basis i, coarse basis j, on the
//                fine basis' elements and Gauss points),
//   L(i)       = int phi_i dx        (each basis with its own quadrature),
//   P = L_f^-1 M_fc,   R = L_c^-1 M_fc^T.
// Element matrices are written per element and summed per row by the row's
// owner in element order, as the assembly gather does: no atomics, the same
// bits on every run.
#include "gsx_common.h"
#include "gsx_kernels.h"
#include "gsx_basis.h"

namespace gsx {

namespace {

constexpr int kQ1 = kMaxQ1;   // Gauss nodes per direction (degree <= 4)

__device__ inline void deboor_any(int p, const double *U, int span, double u, double *N, double *dN)
{
    switch (p) {
    case 1: deboor_ders<1>(U, span, u, N, dN); break;
    case 2: deboor_ders<2>(U, span, u, N, dN); break;
    case 3: deboor_ders<3>(U, span, u, N, dN); break;
    default: deboor_ders<4>(U, span, u, N, dN); break;
    }
}

// Knot span of u: the largest s in [p, nb - 1] with U[s] <= u
// (gsKnotVector::iFind, the span to the right of an interior knot).
__device__ inline int span_of(const double *U, int p, int nb, double u)
{
    int lo = p, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (U[mid] <= u) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Element of basis V (index along one direction) that contains element e of
// basis U: V's breakpoints are a subset of U's.
__device__ inline int v_elem_of(const PatchDev &U, const PatchDev &V, int k, int e)
{
    const double *Uk = U.knots[k];
    const double um = 0.5 * (Uk[U.p[k] + e] + Uk[U.p[k] + e + 1]);
    return span_of(V.knots[k], V.p[k], V.nb[k], um) - V.p[k];
}

// 1D tables of one element of U at its q + 1 Gauss nodes per direction: U's
// values and derivatives, the mapped weights, and the geometry's values,
// derivatives, first active function and element per node.  Optionally V's
// values on the element of V containing it.
struct Tab1D {
    double uN[3][kQ1][kQ1], udN[3][kQ1][kQ1], w[3][kQ1];
    double gN[3][kQ1][kQ1], gdN[3][kQ1][kQ1];
    int gfirst[3][kQ1], gelem[3][kQ1];
    double vN[3][kQ1][kQ1];
};

template <int D>
__device__ inline void tables_1d(const PatchDev &U, const PatchDev &G, const PatchDev *V, const int *ei,
                                 const GaussTab &gt, Tab1D &T)
{
    const int t = threadIdx.x;
    if (t < D * kQ1) {
        const int k = t / kQ1, q = t % kQ1;
        const int nq = U.p[k] + 1;
        if (q < nq) {
            const double *Uk = U.knots[k];
            const int span = U.p[k] + ei[k];
            const double a = Uk[span], b = Uk[span + 1];
            const double h = (b - a) / 2.0;
            const double u = h * (gt.x[nq][q] + 1.0) + a;
            T.w[k][q] = (0.0 == h ? 0.5 : h) * gt.w[nq][q];
            // (fixed-size loops with a guard keep N, dN in registers)
            double N[kQ1], dN[kQ1];
            deboor_any(U.p[k], Uk, span, u, N, dN);
#pragma unroll
            for (int i = 0; i < kQ1; ++i)
                if (i < nq) {
                    T.uN[k][q][i] = N[i];
                    T.udN[k][q][i] = dN[i];
                }
            const int gs = span_of(G.knots[k], G.p[k], G.nb[k], u);
            deboor_any(G.p[k], G.knots[k], gs, u, N, dN);
#pragma unroll
            for (int i = 0; i < kQ1; ++i)
                if (i <= G.p[k]) {
                    T.gN[k][q][i] = N[i];
                    T.gdN[k][q][i] = dN[i];
                }
            T.gfirst[k][q] = gs - G.p[k];
            T.gelem[k][q] = gs - G.p[k];
            if (V) {
                const int vs = v_elem_of(U, *V, k, ei[k]) + V->p[k];
                deboor_any(V->p[k], V->knots[k], vs, u, N, dN);
#pragma unroll
                for (int i = 0; i < kQ1; ++i)
                    if (i <= V->p[k]) T.vN[k][q][i] = N[i];
            }
        }
    }
}

// Geometry at Gauss point (qx, qy, qz): x, J (J[i][j] = dx_i / du_j), det
// and the geometry element (for kappa and per-element sources).
template <int D>
__device__ inline double geo_point(const PatchDev &G, const Tab1D &T, const int *qi, double *x, double (*J)[3],
                                   int64_t &gel)
{
    for (int i = 0; i < D; ++i) {
        x[i] = 0.0;
        for (int j = 0; j < D; ++j) J[i][j] = 0.0;
    }
    const int n0 = G.p[0] + 1, n1 = G.p[1] + 1, n2 = D == 3 ? G.p[2] + 1 : 1;
    for (int cz = 0; cz < n2; ++cz)
        for (int cy = 0; cy < n1; ++cy)
            for (int cx = 0; cx < n0; ++cx) {
                const double v0 = T.gN[0][qi[0]][cx], v1 = T.gN[1][qi[1]][cy];
                const double d0 = T.gdN[0][qi[0]][cx], d1 = T.gdN[1][qi[1]][cy];
                const double v2 = D == 3 ? T.gN[2][qi[2]][cz] : 1.0, d2 = D == 3 ? T.gdN[2][qi[2]][cz] : 0.0;
                int64_t m = (T.gfirst[0][qi[0]] + cx) +
                            (int64_t)G.nb[0] * (T.gfirst[1][qi[1]] + cy);
                if (D == 3) m += (int64_t)G.nb[0] * G.nb[1] * (T.gfirst[2][qi[2]] + cz);
                const double v = v0 * v1 * v2;
                const double g[3] = {d0 * v1 * v2, v0 * d1 * v2, v0 * v1 * d2};
                for (int i = 0; i < D; ++i) {
                    const double cc = G.ctrl[m * D + i];
                    x[i] += cc * v;
                    for (int j = 0; j < D; ++j) J[i][j] += cc * g[j];
                }
            }
    gel = T.gelem[0][qi[0]] + (int64_t)G.nel[0] * T.gelem[1][qi[1]];
    if (D == 3) gel += (int64_t)G.nel[0] * G.nel[1] * T.gelem[2][qi[2]];
    if (D == 2) return J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// (J^-1 J^-T) as its upper triangle (xx, xy, xz, yy, yz, zz / xx, xy, yy)
template <int D>
__device__ inline void inv_metric(const double (*J)[3], double det, double *m)
{
    if (D == 2) {
        const double i00 = J[1][1] / det, i01 = -J[0][1] / det, i10 = -J[1][0] / det, i11 = J[0][0] / det;
        m[0] = i00 * i00 + i01 * i01;
        m[1] = i00 * i10 + i01 * i11;
        m[2] = i10 * i10 + i11 * i11;
        return;
    }
    double I[3][3];
    I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    // grad_x phi = J^-T grad_u phi, so grad_x phi_a . grad_x phi_b =
    // grad_u phi_a^T (J^-1 J^-T) grad_u phi_b
    int o = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) {
            m[o++] = I[a][0] * I[b][0] + I[a][1] * I[b][1] + I[a][2] * I[b][2];
        }
}

constexpr int kMaxPts = kQ1 * kQ1 * kQ1;

// Element stiffness K_e = sum_q w_q |det J| kappa grad(phi)^T grad(phi) and
// load f_e = sum_q w_q |det J| f(x_q) phi (f = 1 when unit_f: the lumped
// mass) of elements [e0, e0 + gridDim.x) of patch ip of U, geometry from
// patch ip of G.  K written at kmat_off + e * NB^2 - kshift (upper value
// mirrored), f at fvec_off + e * NB - fshift; oriented min det per element.
template <int D>
__global__ __launch_bounds__(256) void k_geo_elem(const PatchDev *__restrict__ ups, const PatchDev *__restrict__ gps,
                                                  int ip, int64_t e0, GaussTab gt, double *__restrict__ Kbuf,
                                                  double *__restrict__ Fbuf, double *__restrict__ detmin,
                                                  int64_t kshift, int64_t fshift, int unit_f)
{
    __shared__ Tab1D T;
    __shared__ double sM[kMaxPts][6];
    __shared__ double sF[kMaxPts];
    __shared__ double sDet[kMaxPts];
    const PatchDev &U = ups[ip];
    const PatchDev &G = gps[ip];
    const int64_t e = e0 + blockIdx.x;
    int ei[3] = {0, 0, 0};
    ei[0] = (int)(e % U.nel[0]);
    ei[1] = (int)((e / U.nel[0]) % U.nel[1]);
    if (D == 3) ei[2] = (int)(e / ((int64_t)U.nel[0] * U.nel[1]));
    tables_1d<D>(U, G, nullptr, ei, gt, T);
    __syncthreads();
    const int q0 = U.p[0] + 1, q1 = U.p[1] + 1, q2 = D == 3 ? U.p[2] + 1 : 1;
    const int nq = q0 * q1 * q2;
    const int t = threadIdx.x;
    for (int q = t; q < nq; q += blockDim.x) {
        const int qi[3] = {q % q0, (q / q0) % q1, q / (q0 * q1)};
        double x[3], J[3][3];
        int64_t gel;
        const double det = geo_point<D>(G, T, qi, x, J, gel);
        const double w = T.w[0][qi[0]] * T.w[1][qi[1]] * (D == 3 ? T.w[2][qi[2]] : 1.0);
        const double meas = fabs(det);
        const double kap = G.kappa ? G.kappa[gel] : 1.0;
        double m[6];
        inv_metric<D>(J, det, m);
        const int nm = D == 3 ? 6 : 3;
        for (int i = 0; i < nm; ++i) sM[q][i] = w * (meas * kap) * m[i];
        const double f = unit_f ? 1.0 : source_value(G.fkind, D, x, G.felem, gel);
        sF[q] = w * (f * meas);
        sDet[q] = G.orient * det;
    }
    __syncthreads();
    const int n0 = q0, n1 = q1, n2 = q2;   // functions per direction = nodes
    const int NB = n0 * n1 * n2;
    if (Kbuf) {
        double *Ke = Kbuf + U.kmat_off + e * (int64_t)NB * NB - kshift;
        for (int idx = t; idx < NB * NB; idx += blockDim.x) {
            const int a = idx / NB, b = idx % NB;
            if (a > b) continue;
            const int ax = a % n0, ay = (a / n0) % n1, az = a / (n0 * n1);
            const int bx = b % n0, by = (b / n0) % n1, bz = b / (n0 * n1);
            double acc = 0.0;
            for (int q = 0; q < nq; ++q) {
                const int qx = q % q0, qy = (q / q0) % q1, qz = q / (q0 * q1);
                const double az_v = D == 3 ? T.uN[2][qz][az] : 1.0, az_d = D == 3 ? T.udN[2][qz][az] : 0.0;
                const double bz_v = D == 3 ? T.uN[2][qz][bz] : 1.0, bz_d = D == 3 ? T.udN[2][qz][bz] : 0.0;
                const double ga[3] = {T.udN[0][qx][ax] * T.uN[1][qy][ay] * az_v,
                                      T.uN[0][qx][ax] * T.udN[1][qy][ay] * az_v,
                                      T.uN[0][qx][ax] * T.uN[1][qy][ay] * az_d};
                const double gb[3] = {T.udN[0][qx][bx] * T.uN[1][qy][by] * bz_v,
                                      T.uN[0][qx][bx] * T.udN[1][qy][by] * bz_v,
                                      T.uN[0][qx][bx] * T.uN[1][qy][by] * bz_d};
                double s;
                if (D == 2) {
                    s = ga[0] * (sM[q][0] * gb[0] + sM[q][1] * gb[1]) + ga[1] * (sM[q][1] * gb[0] + sM[q][2] * gb[1]);
                } else {
                    s = ga[0] * (sM[q][0] * gb[0] + sM[q][1] * gb[1] + sM[q][2] * gb[2]) +
                        ga[1] * (sM[q][1] * gb[0] + sM[q][3] * gb[1] + sM[q][4] * gb[2]) +
                        ga[2] * (sM[q][2] * gb[0] + sM[q][4] * gb[1] + sM[q][5] * gb[2]);
                }
                acc += s;
            }
            Ke[(int64_t)a * NB + b] = acc;
            Ke[(int64_t)b * NB + a] = acc;
        }
    }
    if (Fbuf) {
        for (int a = t; a < NB; a += blockDim.x) {
            const int ax = a % n0, ay = (a / n0) % n1, az = a / (n0 * n1);
            double acc = 0.0;
            for (int q = 0; q < nq; ++q) {
                const int qx = q % q0, qy = (q / q0) % q1, qz = q / (q0 * q1);
                const double v = T.uN[0][qx][ax] * T.uN[1][qy][ay] * (D == 3 ? T.uN[2][qz][az] : 1.0);
                acc += sF[q] * v;
            }
            Fbuf[U.fvec_off + e * NB + a - fshift] = acc;
        }
    }
    if (detmin && t == 0) {
        double m = sDet[0];
        for (int q = 1; q < nq; ++q) m = fmin(m, sDet[q]);
        detmin[U.elem_off + e] = m;
    }
}

// Mixed mass of one element of U (the finer level): M_e(a, b) =
// sum_q w_q |det J| phi_a psi_b, psi the functions of V on V's element
// containing it; written at moff + e * NBU * NBV.
template <int D>
__global__ __launch_bounds__(256) void k_geo_mixed(const PatchDev *__restrict__ ups, const PatchDev *__restrict__ vps,
                                                   const PatchDev *__restrict__ gps, int ip, GaussTab gt,
                                                   double *__restrict__ Mbuf, int64_t moff)
{
    __shared__ Tab1D T;
    __shared__ double sW[kMaxPts];
    const PatchDev &U = ups[ip];
    const PatchDev &V = vps[ip];
    const PatchDev &G = gps[ip];
    const int64_t e = blockIdx.x;
    int ei[3] = {0, 0, 0};
    ei[0] = (int)(e % U.nel[0]);
    ei[1] = (int)((e / U.nel[0]) % U.nel[1]);
    if (D == 3) ei[2] = (int)(e / ((int64_t)U.nel[0] * U.nel[1]));
    tables_1d<D>(U, G, &V, ei, gt, T);
    __syncthreads();
    const int q0 = U.p[0] + 1, q1 = U.p[1] + 1, q2 = D == 3 ? U.p[2] + 1 : 1;
    const int nq = q0 * q1 * q2;
    const int t = threadIdx.x;
    for (int q = t; q < nq; q += blockDim.x) {
        const int qi[3] = {q % q0, (q / q0) % q1, q / (q0 * q1)};
        double x[3], J[3][3];
        int64_t gel;
        const double det = geo_point<D>(G, T, qi, x, J, gel);
        sW[q] = T.w[0][qi[0]] * T.w[1][qi[1]] * (D == 3 ? T.w[2][qi[2]] : 1.0) * fabs(det);
    }
    __syncthreads();
    const int v0 = V.p[0] + 1, v1 = V.p[1] + 1, v2 = D == 3 ? V.p[2] + 1 : 1;
    const int NBU = nq, NBV = v0 * v1 * v2;
    double *Me = Mbuf + moff + e * (int64_t)NBU * NBV;
    for (int idx = t; idx < NBU * NBV; idx += blockDim.x) {
        const int a = idx / NBV, b = idx % NBV;
        const int ax = a % q0, ay = (a / q0) % q1, az = a / (q0 * q1);
        const int bx = b % v0, by = (b / v0) % v1, bz = b / (v0 * v1);
        double acc = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int qx = q % q0, qy = (q / q0) % q1, qz = q / (q0 * q1);
            const double pa = T.uN[0][qx][ax] * T.uN[1][qy][ay] * (D == 3 ? T.uN[2][qz][az] : 1.0);
            const double pb = T.vN[0][qx][bx] * T.vN[1][qy][by] * (D == 3 ? T.vN[2][qz][bz] : 1.0);
            acc += sW[q] * (pa * pb);
        }
        Me[idx] = acc;
    }
}

__device__ inline void coords_of(const PatchDev &pd, int64_t l, int *c)
{
    c[0] = (int)(l % pd.nb[0]);
    const int64_t r = l / pd.nb[0];
    c[1] = (int)(r % pd.nb[1]);
    c[2] = (int)(r / pd.nb[1]);
}

// Element box of U function l (elements e in [lo, hi] per direction) and the
// range of V functions that meet those elements.
template <int D>
__device__ inline void mix_ranges(const PatchDev &U, const PatchDev &V, const int *l, int *elo, int *ehi, int *vlo,
                                  int *vhi)
{
    for (int k = 0; k < 3; ++k) {
        elo[k] = ehi[k] = vlo[k] = vhi[k] = 0;
        if (k >= D) continue;
        elo[k] = max(0, l[k] - U.p[k]);
        ehi[k] = min(U.nel[k] - 1, l[k]);
        vlo[k] = v_elem_of(U, V, k, elo[k]);
        vhi[k] = min(V.nb[k] - 1, v_elem_of(U, V, k, ehi[k]) + V.p[k]);
    }
}

// Candidate columns of row r of M (V's free functions meeting the row's
// support, with repetitions across the row's patches): count or write.
template <int D, bool WRITE>
__global__ void k_mix_cand(const PatchDev *__restrict__ ups, const PatchDev *__restrict__ vps,
                           const int *__restrict__ occ_ptr, const int *__restrict__ occ_patch,
                           const int64_t *__restrict__ occ_local, int nrows, int nfree_v,
                           const int *__restrict__ cptr, int *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int cnt = 0;
    int *dst = WRITE ? out + cptr[r] : nullptr;
    for (int o = occ_ptr[r]; o < occ_ptr[r + 1]; ++o) {
        const PatchDev &U = ups[occ_patch[o]];
        const PatchDev &V = vps[occ_patch[o]];
        int l[3], elo[3], ehi[3], vlo[3], vhi[3];
        coords_of(U, occ_local[o], l);
        mix_ranges<D>(U, V, l, elo, ehi, vlo, vhi);
        for (int z = vlo[2]; z <= vhi[2]; ++z)
            for (int y = vlo[1]; y <= vhi[1]; ++y)
                for (int x = vlo[0]; x <= vhi[0]; ++x) {
                    const int64_t m = x + (int64_t)V.nb[0] * (y + (int64_t)V.nb[1] * z);
                    const int g = V.dofs[m];
                    if (g >= nfree_v) continue;
                    if (WRITE) dst[cnt] = g;
                    ++cnt;
                }
    }
    if (!WRITE) out[r] = cnt;
}

// Sort each row's candidates and drop repetitions; ucount[r] = kept.
__global__ void k_mix_unique(int nrows, const int *__restrict__ cptr, int *__restrict__ cand,
                             int *__restrict__ ucount)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int *c = cand + cptr[r];
    const int n = cptr[r + 1] - cptr[r];
    for (int i = 1; i < n; ++i) {
        const int v = c[i];
        int j = i - 1;
        while (j >= 0 && c[j] > v) {
            c[j + 1] = c[j];
            --j;
        }
        c[j + 1] = v;
    }
    int u = 0;
    for (int i = 0; i < n; ++i)
        if (i == 0 || c[i] != c[i - 1]) c[u++] = c[i];
    ucount[r] = u;
}

__global__ void k_mix_compact(int nrows, const int *__restrict__ cptr, const int *__restrict__ cand,
                              const int *__restrict__ rowptr, int *__restrict__ col)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int n = rowptr[r + 1] - rowptr[r];
    for (int i = 0; i < n; ++i) col[rowptr[r] + i] = cand[cptr[r] + i];
}

// Row r of M: every element of every patch of the row, in element order,
// adds its row a of M_e into the row's columns (binary search).
template <int D>
__global__ void k_mix_gather(const PatchDev *__restrict__ ups, const PatchDev *__restrict__ vps,
                             const int *__restrict__ occ_ptr, const int *__restrict__ occ_patch,
                             const int64_t *__restrict__ occ_local, int nrows, int nfree_v,
                             const int *__restrict__ rowptr, const int *__restrict__ col,
                             const double *__restrict__ Mbuf, const int64_t *__restrict__ moff,
                             double *__restrict__ val)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int rb = rowptr[r], len = rowptr[r + 1] - rb;
    for (int i = 0; i < len; ++i) val[rb + i] = 0.0;
    for (int o = occ_ptr[r]; o < occ_ptr[r + 1]; ++o) {
        const int ip = occ_patch[o];
        const PatchDev &U = ups[ip];
        const PatchDev &V = vps[ip];
        int l[3], elo[3], ehi[3], vlo[3], vhi[3];
        coords_of(U, occ_local[o], l);
        mix_ranges<D>(U, V, l, elo, ehi, vlo, vhi);
        const int u0 = U.p[0] + 1, u1 = U.p[1] + 1, u2 = D == 3 ? U.p[2] + 1 : 1;
        const int w0 = V.p[0] + 1, w1 = V.p[1] + 1, w2 = D == 3 ? V.p[2] + 1 : 1;
        const int NBU = u0 * u1 * u2, NBV = w0 * w1 * w2;
        for (int ez = elo[2]; ez <= ehi[2]; ++ez)
            for (int ey = elo[1]; ey <= ehi[1]; ++ey)
                for (int ex = elo[0]; ex <= ehi[0]; ++ex) {
                    const int64_t el = ex + (int64_t)U.nel[0] * (ey + (int64_t)U.nel[1] * ez);
                    const int a = (l[0] - ex) + u0 * ((l[1] - ey) + u1 * (l[2] - ez));
                    const int f0 = v_elem_of(U, V, 0, ex), f1 = v_elem_of(U, V, 1, ey);
                    const int f2 = D == 3 ? v_elem_of(U, V, 2, ez) : 0;
                    const double *Me = Mbuf + moff[ip] + el * (int64_t)NBU * NBV + (int64_t)a * NBV;
                    for (int b = 0; b < NBV; ++b) {
                        const int bx = b % w0, by = (b / w0) % w1, bz = b / (w0 * w1);
                        const int64_t m = (f0 + bx) + (int64_t)V.nb[0] * ((f1 + by) + (int64_t)V.nb[1] * (f2 + bz));
                        const int g = V.dofs[m];
                        if (g >= nfree_v) continue;
                        int lo = 0, hi = len - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (col[rb + mid] < g) lo = mid + 1; else hi = mid;
                        }
                        val[rb + lo] += Me[b];
                    }
                }
    }
}

// The same with one wavefront per row (4 rows per block): the row's columns
// and sums in LDS, the lanes taking the coarse functions of each element
// (distinct columns), the elements one after the other; each entry sums its
// elements in the order of k_mix_gather, so the values are the same bits.
constexpr int kMixRowMax = 1024;
template <int D>
__global__ __launch_bounds__(256) void k_mix_gather_w(const PatchDev *__restrict__ ups, const PatchDev *__restrict__ vps,
                                                      const int *__restrict__ occ_ptr, const int *__restrict__ occ_patch,
                                                      const int64_t *__restrict__ occ_local, int nrows, int nfree_v,
                                                      const int *__restrict__ rowptr, const int *__restrict__ col,
                                                      const double *__restrict__ Mbuf, const int64_t *__restrict__ moff,
                                                      double *__restrict__ val)
{
    __shared__ int sCol[4][kMixRowMax];
    __shared__ double sAcc[4][kMixRowMax];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int r = blockIdx.x * 4 + w;
    if (r >= nrows) return;
    const int rb = rowptr[r], len = rowptr[r + 1] - rb;
    for (int k = lane; k < len; k += kWave) {
        sCol[w][k] = col[rb + k];
        sAcc[w][k] = 0.0;
    }
    __builtin_amdgcn_wave_barrier();
    for (int o = occ_ptr[r]; o < occ_ptr[r + 1]; ++o) {
        const int ip = occ_patch[o];
        const PatchDev &U = ups[ip];
        const PatchDev &V = vps[ip];
        int l[3], elo[3], ehi[3], vlo[3], vhi[3];
        coords_of(U, occ_local[o], l);
        mix_ranges<D>(U, V, l, elo, ehi, vlo, vhi);
        const int u0 = U.p[0] + 1, u1 = U.p[1] + 1, u2 = D == 3 ? U.p[2] + 1 : 1;
        const int w0 = V.p[0] + 1, w1 = V.p[1] + 1, w2 = D == 3 ? V.p[2] + 1 : 1;
        const int NBU = u0 * u1 * u2, NBV = w0 * w1 * w2;
        for (int ez = elo[2]; ez <= ehi[2]; ++ez)
            for (int ey = elo[1]; ey <= ehi[1]; ++ey)
                for (int ex = elo[0]; ex <= ehi[0]; ++ex) {
                    const int64_t el = ex + (int64_t)U.nel[0] * (ey + (int64_t)U.nel[1] * ez);
                    const int a = (l[0] - ex) + u0 * ((l[1] - ey) + u1 * (l[2] - ez));
                    const int f0 = v_elem_of(U, V, 0, ex), f1 = v_elem_of(U, V, 1, ey);
                    const int f2 = D == 3 ? v_elem_of(U, V, 2, ez) : 0;
                    const double *Me = Mbuf + moff[ip] + el * (int64_t)NBU * NBV + (int64_t)a * NBV;
                    for (int b = lane; b < NBV; b += kWave) {
                        const int bx = b % w0, by = (b / w0) % w1, bz = b / (w0 * w1);
                        const int64_t m = (f0 + bx) + (int64_t)V.nb[0] * ((f1 + by) + (int64_t)V.nb[1] * (f2 + bz));
                        const int g = V.dofs[m];
                        if (g >= nfree_v) continue;
                        int lo = 0, hi = len - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (sCol[w][mid] < g) lo = mid + 1; else hi = mid;
                        }
                        sAcc[w][lo] += Me[b];
                    }
                    __builtin_amdgcn_wave_barrier();
                }
    }
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < len; k += kWave) val[rb + k] = sAcc[w][k];
}

// Row r of an element vector field: sum over the row's patches and elements
// in element order (the load part of the assembly gather, without
// eliminated columns).
template <int D>
__global__ void k_vec_gather(const PatchDev *__restrict__ ups, const int *__restrict__ occ_ptr,
                             const int *__restrict__ occ_patch, const int64_t *__restrict__ occ_local, int nrows,
                             const double *__restrict__ Fbuf, double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double acc = 0.0;
    for (int o = occ_ptr[r]; o < occ_ptr[r + 1]; ++o) {
        const PatchDev &U = ups[occ_patch[o]];
        int l[3];
        coords_of(U, occ_local[o], l);
        int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            elo[k] = max(0, l[k] - U.p[k]);
            ehi[k] = min(U.nel[k] - 1, l[k]);
        }
        const int u0 = U.p[0] + 1, u1 = U.p[1] + 1, u2 = D == 3 ? U.p[2] + 1 : 1;
        const int NB = u0 * u1 * u2;
        for (int ez = elo[2]; ez <= ehi[2]; ++ez)
            for (int ey = elo[1]; ey <= ehi[1]; ++ey)
                for (int ex = elo[0]; ex <= ehi[0]; ++ex) {
                    const int64_t el = ex + (int64_t)U.nel[0] * (ey + (int64_t)U.nel[1] * ez);
                    const int a = (l[0] - ex) + u0 * ((l[1] - ey) + u1 * (l[2] - ez));
                    acc += Fbuf[U.fvec_off + el * NB + a];
                }
    }
    out[r] = acc;
}

// val(row r) *= 1 / L(r) (Eigen's diagonal inverse times a sparse matrix)
__global__ void k_scale_rows(int nrows, const int *__restrict__ rowptr, const double *__restrict__ L,
                             double *__restrict__ val, int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const double l = L[r];
    if (!(l != 0.0)) {
        err[0] = 1;
        return;
    }
    const double inv = 1.0 / l;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = inv * val[k];
}

__global__ void k_gather_idx(int n, const int *__restrict__ idx, const double *__restrict__ in,
                             double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

// x[idx[i]] += scale * z[i]; the indices of one call are distinct
__global__ void k_scatter_add_idx(int n, const int *__restrict__ idx, const double *__restrict__ z, double scale,
                                  double *__restrict__ x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[idx[i]] += scale * z[i];
}

inline dim3 grid1(int64_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

hipError_t launch_gather_idx(int n, const int *idx, const double *in, double *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_idx, grid1(n, 256), dim3(256), 0, s, n, idx, in, out);
    return hipGetLastError();
}

hipError_t launch_scatter_add_idx(int n, const int *idx, const double *z, double scale, double *x, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_add_idx, grid1(n, 256), dim3(256), 0, s, n, idx, z, scale, x);
    return hipGetLastError();
}

hipError_t launch_geo_elem(int d, const PatchDev *ups, const PatchDev *gps, int ip, int64_t e0, int64_t e1,
                           const GaussTab &gt, double *Kbuf, double *Fbuf, double *detmin, int64_t kshift,
                           int64_t fshift, int unit_f, hipStream_t s)
{
    if (e1 <= e0) return hipSuccess;
    const dim3 grid((unsigned)(e1 - e0)), blk(256);
    if (d == 2)
        hipLaunchKernelGGL(k_geo_elem<2>, grid, blk, 0, s, ups, gps, ip, e0, gt, Kbuf, Fbuf, detmin, kshift, fshift,
                           unit_f);
    else
        hipLaunchKernelGGL(k_geo_elem<3>, grid, blk, 0, s, ups, gps, ip, e0, gt, Kbuf, Fbuf, detmin, kshift, fshift,
                           unit_f);
    return hipGetLastError();
}

hipError_t launch_geo_mixed(int d, const PatchDev *ups, const PatchDev *vps, const PatchDev *gps, int ip,
                            int64_t nelem, const GaussTab &gt, double *Mbuf, int64_t moff, hipStream_t s)
{
    if (nelem <= 0) return hipSuccess;
    const dim3 grid((unsigned)nelem), blk(256);
    if (d == 2)
        hipLaunchKernelGGL(k_geo_mixed<2>, grid, blk, 0, s, ups, vps, gps, ip, gt, Mbuf, moff);
    else
        hipLaunchKernelGGL(k_geo_mixed<3>, grid, blk, 0, s, ups, vps, gps, ip, gt, Mbuf, moff);
    return hipGetLastError();
}

hipError_t launch_mix_pattern(int phase, int d, const PatchDev *ups, const PatchDev *vps, const int *occ_ptr,
                              const int *occ_patch, const int64_t *occ_local, int nrows, int nfree_v,
                              const int *cptr, int *cand, int *counts, const int *rowptr, int *col, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    const dim3 g = grid1(nrows, 128), b(128);
    if (phase == 0) {
        if (d == 2)
            hipLaunchKernelGGL((k_mix_cand<2, false>), g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows,
                               nfree_v, nullptr, counts);
        else
            hipLaunchKernelGGL((k_mix_cand<3, false>), g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows,
                               nfree_v, nullptr, counts);
    } else if (phase == 1) {
        if (d == 2)
            hipLaunchKernelGGL((k_mix_cand<2, true>), g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows,
                               nfree_v, cptr, cand);
        else
            hipLaunchKernelGGL((k_mix_cand<3, true>), g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows,
                               nfree_v, cptr, cand);
        hipLaunchKernelGGL(k_mix_unique, g, b, 0, s, nrows, cptr, cand, counts);
    } else {
        hipLaunchKernelGGL(k_mix_compact, g, b, 0, s, nrows, cptr, cand, rowptr, col);
    }
    return hipGetLastError();
}

hipError_t launch_mix_gather(int d, const PatchDev *ups, const PatchDev *vps, const int *occ_ptr, const int *occ_patch,
                             const int64_t *occ_local, int nrows, int nfree_v, const int *rowptr, const int *col,
                             const double *Mbuf, const int64_t *moff, double *val, int maxrow, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    if (maxrow <= kMixRowMax) {
        const dim3 g = grid1(nrows, 4), b(256);
        if (d == 2)
            hipLaunchKernelGGL(k_mix_gather_w<2>, g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows, nfree_v,
                               rowptr, col, Mbuf, moff, val);
        else
            hipLaunchKernelGGL(k_mix_gather_w<3>, g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows, nfree_v,
                               rowptr, col, Mbuf, moff, val);
        return hipGetLastError();
    }
    const dim3 g = grid1(nrows, 128), b(128);
    if (d == 2)
        hipLaunchKernelGGL(k_mix_gather<2>, g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows, nfree_v,
                           rowptr, col, Mbuf, moff, val);
    else
        hipLaunchKernelGGL(k_mix_gather<3>, g, b, 0, s, ups, vps, occ_ptr, occ_patch, occ_local, nrows, nfree_v,
                           rowptr, col, Mbuf, moff, val);
    return hipGetLastError();
}

hipError_t launch_vec_gather(int d, const PatchDev *ups, const int *occ_ptr, const int *occ_patch,
                             const int64_t *occ_local, int nrows, const double *Fbuf, double *out, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    const dim3 g = grid1(nrows, 128), b(128);
    if (d == 2)
        hipLaunchKernelGGL(k_vec_gather<2>, g, b, 0, s, ups, occ_ptr, occ_patch, occ_local, nrows, Fbuf, out);
    else
        hipLaunchKernelGGL(k_vec_gather<3>, g, b, 0, s, ups, occ_ptr, occ_patch, occ_local, nrows, Fbuf, out);
    return hipGetLastError();
}

hipError_t launch_scale_rows(int nrows, const int *rowptr, const double *L, double *val, int *err, hipStream_t s)
{
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale_rows, grid1(nrows, 256), dim3(256), 0, s, nrows, rowptr, L, val, err);
    return hipGetLastError();
}

}  // namespace gsx
