// Assembly kernels of libgsx (gfx950): B-spline basis at Gauss points,
// geometry map, element stiffness/load, CSR pattern and the deterministic
// row-owner gather into CSR.
//
// Reference path (G+Smo source tree):
//   basis      gsTensorBSplineBasis<1,T>::evalAllDers_into
//              src/gsNurbs/gsBSplineBasis.hpp:862-1040,
//              gsTensorBasis::evalAllDers_into src/gsTensor/gsTensorBasis.hpp:634-729
//   quadrature gsQuadRule::mapTo src/gsAssembler/gsQuadRule.hpp:24-117
//   geometry   gsFunction::computeAuxiliaryData src/gsCore/gsFunction.hpp:604-760
//   element    igrad(u,G)*igrad(u,G).tr()*meas(G), u*ff*meas(G)
//              src/gsAssembler/gsExpressions.h:2737-3890, quadrature
//              src/gsAssembler/gsExprAssembler.h:497-505
//   scatter    gsExprAssembler::_eval::push src/gsAssembler/gsExprAssembler.h:553-630
#include "gsx_common.h"
#include "gsx_kernels.h"
#include "gsx_basis.h"

namespace gsx {

__device__ inline int find_patch(const int64_t *__restrict__ prefix, int np, int64_t e)
{
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// 2D element kernel (direct contraction; n_b <= 25).  One element per
// block.  K_e(i,j) = sum_q w_q * ((G_i . G_j) * meas_q [* kappa]) with
// G = grad(N) J^{-1}; f_e(i) = sum_q w_q * ((N_i f(x_q)) * meas_q).
// Writes the full symmetric K_e (upper value mirrored) and f_e.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(128) void k_elem2d(const PatchDev *__restrict__ patches,
                                                const int *__restrict__ gpatch,
                                                const int64_t *__restrict__ gprefix, int ng,
                                                GaussTab gt, double *__restrict__ Kbuf,
                                                double *__restrict__ Fbuf,
                                                double *__restrict__ detmin, ElemRange rg)
{
    constexpr int NB1 = P + 1, NB = NB1 * NB1, NQ1 = P + 1, NQ = NQ1 * NQ1;
    __shared__ double sV[2][NQ1][NB1], sD[2][NQ1][NB1], sW[2][NQ1];
    __shared__ double sC[NB][2];
    __shared__ double sJi[NQ][4], sMeas[NQ], sWq[NQ], sF[NQ];
    __shared__ double sG[NQ][2][NB];
    __shared__ double sDet[NQ];

    const int64_t E = rg.e0 + blockIdx.x;
    const int gi = find_patch(gprefix, ng, E);
    const PatchDev &pd = patches[gpatch[gi]];
    const int64_t e = E - gprefix[gi];
    int ei[3];
    ei[0] = (int)(e % pd.nel[0]);
    ei[1] = (int)(e / pd.nel[0]);
    ei[2] = 0;
    const int t = threadIdx.x;

    basis_tables<2, P>(pd, ei, gt, sV, sD, sW);
    if (t < NB * 2) {
        const int b = t >> 1, r = t & 1;
        const int bx = b % NB1, by = b / NB1;
        const int m = (ei[0] + bx) + pd.nb[0] * (ei[1] + by);
        sC[b][r] = pd.ctrl[(int64_t)m * 2 + r];
    }
    __syncthreads();

    const double kap = pd.kappa ? pd.kappa[e] : 1.0;
    if (t < NQ) {
        const int qx = t % NQ1, qy = t / NQ1;
        double x0 = 0, x1 = 0, J00 = 0, J01 = 0, J10 = 0, J11 = 0;
        for (int b = 0; b < NB; ++b) {
            const int bx = b % NB1, by = b / NB1;
            const double v = sV[0][qx][bx] * sV[1][qy][by];
            const double g0 = sD[0][qx][bx] * sV[1][qy][by];
            const double g1 = sD[1][qy][by] * sV[0][qx][bx];
            x0 += sC[b][0] * v;
            x1 += sC[b][1] * v;
            J00 += sC[b][0] * g0;
            J01 += sC[b][0] * g1;
            J10 += sC[b][1] * g0;
            J11 += sC[b][1] * g1;
        }
        const double det = J00 * J11 - J01 * J10;
        sJi[t][0] = J11 / det;
        sJi[t][1] = -J01 / det;
        sJi[t][2] = -J10 / det;
        sJi[t][3] = J00 / det;
        sMeas[t] = fabs(det);
        sDet[t] = det;
        sWq[t] = sW[0][qx] * sW[1][qy];
        const double x[2] = {x0, x1};
        sF[t] = source_value(pd.fkind, 2, x, pd.felem, e);
    }
    __syncthreads();
    for (int idx = t; idx < NQ * NB; idx += blockDim.x) {
        const int q = idx / NB, b = idx % NB;
        const int qx = q % NQ1, qy = q / NQ1, bx = b % NB1, by = b / NB1;
        const double g0 = sD[0][qx][bx] * sV[1][qy][by];
        const double g1 = sD[1][qy][by] * sV[0][qx][bx];
        sG[q][0][b] = g0 * sJi[q][0] + g1 * sJi[q][2];
        sG[q][1][b] = g0 * sJi[q][1] + g1 * sJi[q][3];
    }
    __syncthreads();

    double *Ke = Kbuf + pd.kmat_off + e * (int64_t)(NB * NB) - rg.kshift;
    for (int idx = t; idx < NB * NB; idx += blockDim.x) {
        const int i = idx / NB, j = idx % NB;
        if (i > j) continue;
        double acc = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const double s = sG[q][0][i] * sG[q][0][j] + sG[q][1][i] * sG[q][1][j];
            double ev = s * sMeas[q];
            if (pd.kappa) ev *= kap;
            acc += sWq[q] * ev;
        }
        Ke[i * NB + j] = acc;
        Ke[j * NB + i] = acc;
    }
    if (t < NB) {
        const int bx = t % NB1, by = t / NB1;
        double acc = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const int qx = q % NQ1, qy = q / NQ1;
            const double v = sV[0][qx][bx] * sV[1][qy][by];
            acc += sWq[q] * ((v * sF[q]) * sMeas[q]);
        }
        Fbuf[pd.fvec_off + e * NB + t - rg.fshift] = acc;
    }
    if (t == 0) {
        // oriented determinant: positive at every point of a valid patch of
        // either orientation (the measure is |det J|)
        double m = pd.orient * sDet[0];
        for (int q = 1; q < NQ; ++q) m = fmin(m, pd.orient * sDet[q]);
        detmin[pd.elem_off + e] = m;
    }
}

// ---------------------------------------------------------------------------
// 3D element kernel, sum-factorised.  One element per 256-thread block.
//   D_ab(q) = w_q |det J_q| kappa (J^{-1} J^{-T})_ab
//   K(i,j)  = sum_ab sum_qz Z^{ab}_{iz jz}(qz) sum_qy Y^{ab}_{iy jy}(qy)
//                 sum_qx X^{ab}_{ix jx}(qx) D_ab(qx,qy,qz)
// with X^{ab}_{ij}(q) = phi^{[a==0]}_i(q) phi^{[b==0]}_j(q) (phi^1 = the
// derivative), likewise Y (direction 1) and Z (direction 2).  This is the
// same bilinear form as the direct contraction grad(N_i) D grad(N_j), in
// (P+1)^5-type stages instead of (P+1)^9.
// ---------------------------------------------------------------------------
template <int P>
struct Elem3Cfg {
    static constexpr int NB1 = P + 1, NB = NB1 * NB1 * NB1;
    static constexpr int NQ1 = P + 1, NQ = NQ1 * NQ1 * NQ1;
    static constexpr int S = (256 / NQ) >= 8 ? 8 : (256 / NQ);   // geometry split
    static constexpr int NC = NB1 * NB1 * NB1 * NB1;               // (ix,jx,iy,jy) combos
    static constexpr int MC = (NC + 255) / 256;                    // combos per thread
};

template <int P>
__global__ __launch_bounds__(256) void k_elem3d(const PatchDev *__restrict__ patches,
                                                const int *__restrict__ gpatch,
                                                const int64_t *__restrict__ gprefix, int ng,
                                                GaussTab gt, double *__restrict__ Kbuf,
                                                double *__restrict__ Fbuf,
                                                double *__restrict__ detmin, ElemRange rg)
{
    using C = Elem3Cfg<P>;
    constexpr int NB1 = C::NB1, NB = C::NB, NQ1 = C::NQ1, NQ = C::NQ, S = C::S;
    constexpr int NC = C::NC, MC = C::MC;
    __shared__ double sV[3][NQ1][NB1], sD[3][NQ1][NB1], sW[3][NQ1];
    __shared__ double sC[NB][3];
    __shared__ double sPart[S][NQ][12];
    __shared__ double sDab[6][NQ];
    __shared__ double sWq[NQ], sF[NQ], sMeas[NQ], sDet[NQ];
    __shared__ double sT1[NB1 * NB1][NQ1 * NQ1];

    const int64_t E = rg.e0 + blockIdx.x;
    const int gi = find_patch(gprefix, ng, E);
    const PatchDev &pd = patches[gpatch[gi]];
    const int64_t e = E - gprefix[gi];
    int ei[3];
    ei[0] = (int)(e % pd.nel[0]);
    ei[1] = (int)((e / pd.nel[0]) % pd.nel[1]);
    ei[2] = (int)(e / ((int64_t)pd.nel[0] * pd.nel[1]));
    const int t = threadIdx.x;

    basis_tables<3, P>(pd, ei, gt, sV, sD, sW);
    for (int idx = t; idx < NB * 3; idx += 256) {
        const int b = idx / 3, r = idx % 3;
        const int bx = b % NB1, by = (b / NB1) % NB1, bz = b / (NB1 * NB1);
        const int64_t m = (int64_t)(ei[0] + bx) +
                          (int64_t)pd.nb[0] * ((ei[1] + by) + (int64_t)pd.nb[1] * (ei[2] + bz));
        sC[b][r] = pd.ctrl[m * 3 + r];
    }
    __syncthreads();

    // geometry: partial sums over basis functions, S parts per point
    if (t < S * NQ) {
        const int q = t % NQ, part = t / NQ;
        const int qx = q % NQ1, qy = (q / NQ1) % NQ1, qz = q / (NQ1 * NQ1);
        double acc[12];
#pragma unroll
        for (int c = 0; c < 12; ++c) acc[c] = 0.0;
        const int b0 = part * NB / S, b1 = (part + 1) * NB / S;
        for (int b = b0; b < b1; ++b) {
            const int bx = b % NB1, by = (b / NB1) % NB1, bz = b / (NB1 * NB1);
            const double vx = sV[0][qx][bx], vy = sV[1][qy][by], vz = sV[2][qz][bz];
            const double v = vx * vy * vz;
            const double g0 = sD[0][qx][bx] * vy * vz;
            const double g1 = sD[1][qy][by] * vx * vz;
            const double g2 = sD[2][qz][bz] * vx * vy;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double c = sC[b][r];
                acc[r] += c * v;
                acc[3 + 3 * r + 0] += c * g0;
                acc[3 + 3 * r + 1] += c * g1;
                acc[3 + 3 * r + 2] += c * g2;
            }
        }
#pragma unroll
        for (int c = 0; c < 12; ++c) sPart[part][q][c] = acc[c];
    }
    __syncthreads();
    const double kap = pd.kappa ? pd.kappa[e] : 1.0;
    if (t < NQ) {
        double g[12];
#pragma unroll
        for (int c = 0; c < 12; ++c) g[c] = sPart[0][t][c];
        for (int part = 1; part < S; ++part)
#pragma unroll
            for (int c = 0; c < 12; ++c) g[c] += sPart[part][t][c];
        const double J00 = g[3], J01 = g[4], J02 = g[5];
        const double J10 = g[6], J11 = g[7], J12 = g[8];
        const double J20 = g[9], J21 = g[10], J22 = g[11];
        const double c00 = J11 * J22 - J12 * J21;
        const double c01 = J12 * J20 - J10 * J22;
        const double c02 = J10 * J21 - J11 * J20;
        const double det = J00 * c00 + J01 * c01 + J02 * c02;
        double inv[3][3];
        inv[0][0] = c00 / det;
        inv[1][0] = c01 / det;
        inv[2][0] = c02 / det;
        inv[0][1] = (J02 * J21 - J01 * J22) / det;
        inv[1][1] = (J00 * J22 - J02 * J20) / det;
        inv[2][1] = (J01 * J20 - J00 * J21) / det;
        inv[0][2] = (J01 * J12 - J02 * J11) / det;
        inv[1][2] = (J02 * J10 - J00 * J12) / det;
        inv[2][2] = (J00 * J11 - J01 * J10) / det;
        const int qx = t % NQ1, qy = (t / NQ1) % NQ1, qz = t / (NQ1 * NQ1);
        const double w = sW[0][qx] * sW[1][qy] * sW[2][qz];
        const double meas = fabs(det);
        double cq = w * meas;
        if (pd.kappa) cq *= kap;
        int s = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = a; b < 3; ++b) {
                const double m = inv[a][0] * inv[b][0] + inv[a][1] * inv[b][1] + inv[a][2] * inv[b][2];
                sDab[s++][t] = cq * m;
            }
        sWq[t] = w;
        sMeas[t] = meas;
        sDet[t] = det;
        const double x[3] = {g[0], g[1], g[2]};
        sF[t] = source_value(pd.fkind, 3, x, pd.felem, e);
    }
    __syncthreads();

    // load vector
    if (t < NB) {
        const int bx = t % NB1, by = (t / NB1) % NB1, bz = t / (NB1 * NB1);
        double acc = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const int qx = q % NQ1, qy = (q / NQ1) % NQ1, qz = q / (NQ1 * NQ1);
            const double v = sV[0][qx][bx] * sV[1][qy][by] * sV[2][qz][bz];
            acc += sWq[q] * ((v * sF[q]) * sMeas[q]);
        }
        Fbuf[pd.fvec_off + e * NB + t - rg.fshift] = acc;
    }
    if (t == 0) {
        // oriented determinant: positive at every point of a valid patch of
        // either orientation (the measure is |det J|)
        double m = pd.orient * sDet[0];
        for (int q = 1; q < NQ; ++q) m = fmin(m, pd.orient * sDet[q]);
        detmin[pd.elem_off + e] = m;
    }

    // stiffness
    double acc[MC][NB1][NB1];
#pragma unroll
    for (int mc = 0; mc < MC; ++mc)
#pragma unroll
        for (int i = 0; i < NB1; ++i)
#pragma unroll
            for (int j = 0; j < NB1; ++j) acc[mc][i][j] = 0.0;

    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) {
            const int sab = (a <= b) ? (a == 0 ? b : (a == 1 ? 2 + b : 5)) : (b == 0 ? a : (b == 1 ? 2 + a : 5));
            // stage 1: T1[ix jx][qy qz] = sum_qx X_{ix jx}(qx) D_ab(qx,qy,qz)
            for (int idx = t; idx < NB1 * NB1 * NQ1 * NQ1; idx += 256) {
                const int ixjx = idx / (NQ1 * NQ1), qyz = idx % (NQ1 * NQ1);
                const int ix = ixjx / NB1, jx = ixjx % NB1;
                double s = 0.0;
#pragma unroll
                for (int qx = 0; qx < NQ1; ++qx) {
                    const double fi = (a == 0) ? sD[0][qx][ix] : sV[0][qx][ix];
                    const double fj = (b == 0) ? sD[0][qx][jx] : sV[0][qx][jx];
                    s += (fi * fj) * sDab[sab][qx + NQ1 * qyz];
                }
                sT1[ixjx][qyz] = s;
            }
            __syncthreads();
            // stages 2+3 in registers
#pragma unroll
            for (int mc = 0; mc < MC; ++mc) {
                const int c = t + 256 * mc;
                if (c < NC) {
                    const int ixjx = c / (NB1 * NB1), iyjy = c % (NB1 * NB1);
                    const int iy = iyjy / NB1, jy = iyjy % NB1;
                    double T2[NQ1];
#pragma unroll
                    for (int qz = 0; qz < NQ1; ++qz) T2[qz] = 0.0;
#pragma unroll
                    for (int qy = 0; qy < NQ1; ++qy) {
                        const double fi = (a == 1) ? sD[1][qy][iy] : sV[1][qy][iy];
                        const double fj = (b == 1) ? sD[1][qy][jy] : sV[1][qy][jy];
                        const double m = fi * fj;
#pragma unroll
                        for (int qz = 0; qz < NQ1; ++qz) T2[qz] += m * sT1[ixjx][qy + NQ1 * qz];
                    }
#pragma unroll
                    for (int qz = 0; qz < NQ1; ++qz) {
                        double u[NB1];
#pragma unroll
                        for (int jz = 0; jz < NB1; ++jz)
                            u[jz] = ((b == 2) ? sD[2][qz][jz] : sV[2][qz][jz]) * T2[qz];
#pragma unroll
                        for (int iz = 0; iz < NB1; ++iz) {
                            const double fi = (a == 2) ? sD[2][qz][iz] : sV[2][qz][iz];
                            // entries iz > jz are the mirror thread's (see the store)
#pragma unroll
                            for (int jz = iz; jz < NB1; ++jz) acc[mc][iz][jz] += fi * u[jz];
                        }
                    }
                }
            }
            __syncthreads();
        }
    }

    // Thread (u, v) = ((ix, iy), (jx, jy)) holds the entries (i, j) with
    // i = (u, iz), j = (v, jz), iz <= jz; the pair with iz > jz is thread
    // (v, u)'s (jz, iz).  Each unordered pair is written once to both (i, j)
    // and (j, i) (bitwise symmetric): for iz == jz by the thread with u <= v.
    double *Ke = Kbuf + pd.kmat_off + e * (int64_t)(NB * NB) - rg.kshift;
#pragma unroll
    for (int mc = 0; mc < MC; ++mc) {
        const int c = t + 256 * mc;
        if (c < NC) {
            const int ixjx = c / (NB1 * NB1), iyjy = c % (NB1 * NB1);
            const int ix = ixjx / NB1, jx = ixjx % NB1, iy = iyjy / NB1, jy = iyjy % NB1;
            const bool uv = ix + NB1 * iy <= jx + NB1 * jy;
#pragma unroll
            for (int iz = 0; iz < NB1; ++iz)
#pragma unroll
                for (int jz = iz; jz < NB1; ++jz) {
                    if (jz == iz && !uv) continue;
                    const int i = ix + NB1 * (iy + NB1 * iz);
                    const int j = jx + NB1 * (jy + NB1 * jz);
                    Ke[i * NB + j] = acc[mc][iz][jz];
                    Ke[j * NB + i] = acc[mc][iz][jz];
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Sparsity pattern.  Row r (a free DOF) is the union, over its occurrences
// (patch, local index l), of the free DOFs m with |m_k - l_k| <= p_k in every
// direction (the functions sharing an element with l).  Columns sorted
// ascending, as Eigen keeps them after makeCompressed().
// ---------------------------------------------------------------------------
constexpr int kMaxMultiCols = 768;

__device__ inline void local_coords(const PatchDev &pd, int64_t l, int *c)
{
    c[0] = (int)(l % pd.nb[0]);
    const int64_t r = l / pd.nb[0];
    c[1] = (int)(r % pd.nb[1]);
    c[2] = (int)(r / pd.nb[1]);
}

template <bool WRITE>
__device__ int row_columns(const PatchDev *__restrict__ patches, const int *__restrict__ occ_ptr,
                           const int *__restrict__ occ_patch, const int64_t *__restrict__ occ_local,
                           int nfree, int r, int *__restrict__ out)
{
    const int o0 = occ_ptr[r], o1 = occ_ptr[r + 1];
    if (o1 - o0 == 1) {
        const PatchDev &pd = patches[occ_patch[o0]];
        int l[3];
        local_coords(pd, occ_local[o0], l);
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int k = 0; k < pd.d; ++k) {
            lo[k] = max(0, l[k] - pd.p[k]);
            hi[k] = min(pd.nb[k] - 1, l[k] + pd.p[k]);
        }
        int cnt = 0, last = -1;
        bool sorted = true;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const int64_t m = x + (int64_t)pd.nb[0] * (y + (int64_t)pd.nb[1] * z);
                    const int g = pd.dofs[m];
                    if (g < nfree) {
                        if (WRITE) out[cnt] = g;
                        if (g < last) sorted = false;
                        last = g;
                        ++cnt;
                    }
                }
        if (WRITE && !sorted) {
            for (int i = 1; i < cnt; ++i) {
                const int v = out[i];
                int j = i - 1;
                while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; --j; }
                out[j + 1] = v;
            }
        }
        return cnt;
    }
    int buf[kMaxMultiCols];
    int cnt = 0;
    for (int o = o0; o < o1; ++o) {
        const PatchDev &pd = patches[occ_patch[o]];
        int l[3];
        local_coords(pd, occ_local[o], l);
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int k = 0; k < pd.d; ++k) {
            lo[k] = max(0, l[k] - pd.p[k]);
            hi[k] = min(pd.nb[k] - 1, l[k] + pd.p[k]);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    const int64_t m = x + (int64_t)pd.nb[0] * (y + (int64_t)pd.nb[1] * z);
                    const int g = pd.dofs[m];
                    if (g >= nfree) continue;
                    // sorted insert with dedup
                    int lo2 = 0, hi2 = cnt;
                    while (lo2 < hi2) {
                        const int mid = (lo2 + hi2) >> 1;
                        if (buf[mid] < g) lo2 = mid + 1; else hi2 = mid;
                    }
                    if (lo2 < cnt && buf[lo2] == g) continue;
                    if (cnt >= kMaxMultiCols) return -1;
                    for (int i = cnt; i > lo2; --i) buf[i] = buf[i - 1];
                    buf[lo2] = g;
                    ++cnt;
                }
    }
    if (WRITE)
        for (int i = 0; i < cnt; ++i) out[i] = buf[i];
    return cnt;
}

__global__ void k_rowcount(const PatchDev *__restrict__ patches, const int *__restrict__ occ_ptr,
                           const int *__restrict__ occ_patch, const int64_t *__restrict__ occ_local,
                           int nfree, int *__restrict__ counts, int *__restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nfree) return;
    const int c = row_columns<false>(patches, occ_ptr, occ_patch, occ_local, nfree, r, nullptr);
    if (c < 0) {
        err[0] = 1;
        counts[r] = 0;
    } else {
        counts[r] = c;
    }
}

__global__ void k_rowfill(const PatchDev *__restrict__ patches, const int *__restrict__ occ_ptr,
                          const int *__restrict__ occ_patch, const int64_t *__restrict__ occ_local,
                          int nfree, const int *__restrict__ rowptr, int *__restrict__ col)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nfree) return;
    row_columns<true>(patches, occ_ptr, occ_patch, occ_local, nfree, r, col + rowptr[r]);
}

// ---------------------------------------------------------------------------
// Exclusive scan of int counts into int rowptr (three passes, deterministic).
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(kScanBlock) void k_scan_local(const int *__restrict__ in, int n,
                                                           int *__restrict__ out,
                                                           long long *__restrict__ bsum)
{
    __shared__ long long s[kScanBlock];
    const int i = blockIdx.x * kScanBlock + threadIdx.x;
    long long v = (i < n) ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        long long a = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += a;
        __syncthreads();
    }
    if (i < n) out[i] = (int)(s[threadIdx.x] - v);
    if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

__global__ void k_scan_blocks(long long *__restrict__ bsum, int nb, long long *__restrict__ total)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        long long acc = 0;
        for (int b = 0; b < nb; ++b) {
            const long long v = bsum[b];
            bsum[b] = acc;
            acc += v;
        }
        total[0] = acc;
    }
}

__global__ void k_scan_add(int *__restrict__ out, int n, const long long *__restrict__ bsum,
                           const long long *__restrict__ total)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int)(out[i] + bsum[i / kScanBlock]);
    if (i == 0) out[n] = (int)total[0];
}

// ---------------------------------------------------------------------------
// Gather: one wavefront owns one CSR row and sums every element
// contribution into it in the reference's order (patches ascending, elements
// lexicographic) -- atomic-free and deterministic.  Within one element the
// lanes hit distinct columns.  With keep_flags != nullptr an entry is marked
// when at least one contribution was non-zero (gsExprAssembler skips exact
// zeros, gsExprAssembler.h:599).
// ---------------------------------------------------------------------------
template <int MAXROW>
__global__ __launch_bounds__(256) void k_gather(const PatchDev *__restrict__ patches,
                                                const int *__restrict__ occ_ptr,
                                                const int *__restrict__ occ_patch,
                                                const int64_t *__restrict__ occ_local, int nfree,
                                                const int *__restrict__ rowptr,
                                                const int *__restrict__ col,
                                                const double *__restrict__ Kbuf,
                                                const double *__restrict__ Fbuf,
                                                double *__restrict__ val, double *__restrict__ rhs,
                                                unsigned char *__restrict__ keep_flags,
                                                const double *__restrict__ fixed)
{
    __shared__ int sCol[4][MAXROW];
    __shared__ double sAcc[4][MAXROW];
    __shared__ unsigned char sNz[4][MAXROW];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int r = blockIdx.x * 4 + w;
    if (r >= nfree) return;
    const int rb = rowptr[r], len = rowptr[r + 1] - rb;
    for (int k = lane; k < len; k += kWave) {
        sCol[w][k] = col[rb + k];
        sAcc[w][k] = 0.0;
        sNz[w][k] = 0;
    }
    __builtin_amdgcn_wave_barrier();
    double racc = 0.0;
    const int o0 = occ_ptr[r], o1 = occ_ptr[r + 1];
    for (int o = o0; o < o1; ++o) {
        const PatchDev &pd = patches[occ_patch[o]];
        int l[3];
        local_coords(pd, occ_local[o], l);
        int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0}, nb1[3] = {1, 1, 1};
        for (int k = 0; k < pd.d; ++k) {
            elo[k] = max(0, l[k] - pd.p[k]);
            ehi[k] = min(pd.nel[k] - 1, l[k]);
            nb1[k] = pd.p[k] + 1;
        }
        const int NB = nb1[0] * nb1[1] * nb1[2];
        const int ne0 = ehi[0] - elo[0] + 1, ne1 = ehi[1] - elo[1] + 1;
        const int ne = ne0 * ne1 * (ehi[2] - elo[2] + 1);
        // Rows of short elements (NB < 64) are read EPL elements per wave
        // load (lane = sub-element * NB + b); the sub-elements are then applied
        // one after the other, so every entry still sums its elements in order.
        const int EPL = NB >= kWave ? 1 : kWave / NB;
        // element t's load entry F_e(lb), added to the rhs after its
        // eliminated columns, element by element
        auto load = [&](int t) {
            const int ex = elo[0] + t % ne0, ey = elo[1] + (t / ne0) % ne1, ez = elo[2] + t / (ne0 * ne1);
            const int64_t el = ex + (int64_t)pd.nel[0] * (ey + (int64_t)pd.nel[1] * ez);
            const int lb = (l[0] - ex) + nb1[0] * ((l[1] - ey) + nb1[1] * (l[2] - ez));
            return Fbuf[pd.fvec_off + el * NB + lb];
        };
        for (int t0 = 0; t0 < ne; t0 += EPL) {
            const int nsub = min(EPL, ne - t0);
            double elim1 = 0.0;   // EPL == 1: eliminated columns over the b chunks
            for (int b0 = 0; b0 < NB; b0 += kWave) {
                const int sub = NB >= kWave ? 0 : lane / NB;
                const int b = NB >= kWave ? b0 + lane : lane % NB;
                const int t = t0 + sub;
                int pos = -1;
                double v = 0.0, el_v = 0.0;
                if (sub < nsub && b < NB) {
                    const int ex = elo[0] + t % ne0, ey = elo[1] + (t / ne0) % ne1, ez = elo[2] + t / (ne0 * ne1);
                    const int64_t el = ex + (int64_t)pd.nel[0] * (ey + (int64_t)pd.nel[1] * ez);
                    const int lb = (l[0] - ex) + nb1[0] * ((l[1] - ey) + nb1[1] * (l[2] - ez));
                    const double kv = Kbuf[pd.kmat_off + el * NB * NB + (int64_t)lb * NB + b];
                    const int bx = b % nb1[0], by = (b / nb1[0]) % nb1[1], bz = b / (nb1[0] * nb1[1]);
                    const int64_t m = (ex + bx) + (int64_t)pd.nb[0] * ((ey + by) + (int64_t)pd.nb[1] * (ez + bz));
                    const int g = pd.dofs[m];
                    if (g < nfree) {
                        int lo = 0, hi = len - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (sCol[w][mid] < g) lo = mid + 1; else hi = mid;
                        }
                        pos = lo;
                        v = kv;
                    } else if (fixed) {
                        // eliminated column: K_ij g_j moves to the rhs
                        el_v = kv * fixed[g - nfree];
                    }
                }
                for (int s2 = 0; s2 < nsub; ++s2) {
                    if (sub == s2 && pos >= 0) {
                        sAcc[w][pos] += v;
                        sNz[w][pos] |= (v != 0.0);
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (fixed) {
                        double e = sub == s2 ? el_v : 0.0;
#pragma unroll
                        for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, kWave);
                        if (EPL == 1) elim1 += e;
                        else racc -= e;
                    }
                    if (EPL > 1 && Fbuf) racc += load(t0 + s2);
                }
            }
            if (EPL == 1) {
                if (fixed) racc -= elim1;
                if (Fbuf) racc += load(t0);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < len; k += kWave) {
        val[rb + k] = sAcc[w][k];
        if (keep_flags) keep_flags[rb + k] = sNz[w][k];
    }
    if (rhs && lane == 0) rhs[r] = racc;
}

// ---------------------------------------------------------------------------
// Single-patch gather.  On one patch whose free functions form the box
// [flo, fhi] (Dirichlet sides eliminated), row i and column j = i + o with
// |o_k| <= P meet in the elements e_k in [max(0, max(i_k, j_k) - P),
// min(nel_k - 1, min(i_k, j_k))]; the contribution of element e is
// K_e(i - e, j - e).  The wavefront of row i sums, element by element in
// lexicographic order (the reference's order), every K_e row of i into the
// (2P+1)^D stencil slots (slot = o + P per direction: no search), then
// writes the CSR values of the row (window clipped to the free box,
// lexicographic = ascending column) and the structured operator's slots.
// ---------------------------------------------------------------------------
struct PatchBox {
    int flo[3], fhi[3];   // free functions per direction (inclusive)
    int nb[3], nel[3];
    int padded[3];        // padded grid dims (structured operator)
};

// Rows per block of the single-patch gather: the block keeps its rows'
// stencil slots in LDS so that the slot-major structured operator is written
// in runs of R consecutive rows (whole cache lines) instead of 8-byte
// pieces a matrix-width apart.
template <int D, int P>
struct GatherCfg {
    static constexpr int W = 2 * P + 1;
    static constexpr int NS = (D == 3) ? W * W * W : W * W;
    static constexpr int R = (NS * 16 * 9 <= 56 * 1024) ? 16 : 8;
};

template <int D, int P>
__global__ __launch_bounds__(256) void k_gather_patch(PatchBox bx, int nfree, int r0, int r1, int64_t elbase,
                                                      const int *__restrict__ f2g,
                                                      const int *__restrict__ rowptr,
                                                      const double *__restrict__ Kbuf,
                                                      const double *__restrict__ Fbuf,
                                                      double *__restrict__ val,
                                                      double *__restrict__ rhs,
                                                      unsigned char *__restrict__ keep,
                                                      double *__restrict__ s64,
                                                      float *__restrict__ s32,
                                                      const int *__restrict__ dofs,
                                                      const double *__restrict__ fixed)
{
    constexpr int W = 2 * P + 1;
    constexpr int NS = GatherCfg<D, P>::NS;
    constexpr int R = GatherCfg<D, P>::R;
    constexpr int NB1 = P + 1;
    constexpr int NB = (D == 3) ? NB1 * NB1 * NB1 : NB1 * NB1;
    __shared__ double sAcc[R][NS];
    __shared__ unsigned char sNz[R][NS];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int rbase = r0 + blockIdx.x * R;
    for (int k = threadIdx.x; k < R * NS; k += 256) {
        sAcc[k / NS][k % NS] = 0.0;
        sNz[k / NS][k % NS] = 0;
    }
    __syncthreads();
    for (int rr = w; rr < R; rr += 4) {
        const int r = rbase + rr;
        if (r >= r1) break;
        const int g = f2g[r];
        int ic[3];
        ic[0] = g % bx.padded[0] - P;
        ic[1] = (g / bx.padded[0]) % bx.padded[1] - P;
        ic[2] = (D == 3) ? g / (bx.padded[0] * bx.padded[1]) - P : 0;
        int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            elo[k] = max(0, ic[k] - P);
            ehi[k] = min(bx.nel[k] - 1, ic[k]);
        }
        double racc = 0.0;
        const int ne0 = ehi[0] - elo[0] + 1, ne1 = ehi[1] - elo[1] + 1, ne2 = ehi[2] - elo[2] + 1;
        const int ne = ne0 * ne1 * ne2;
        // element t of the row's lexicographic element range; interior rows
        // (P+1 elements per direction) decode with constant divisors
        const bool interior = ne0 == NB1 && ne1 == NB1;
        auto decode = [&](int t, int &ex, int &ey, int &ez) {
            if (interior) {
                ex = elo[0] + t % NB1;
                ey = elo[1] + (t / NB1) % NB1;
                ez = elo[2] + t / (NB1 * NB1);
            } else {
                ex = elo[0] + t % ne0;
                ey = elo[1] + (t / ne0) % ne1;
                ez = elo[2] + t / (ne0 * ne1);
            }
        };
        // element rows in flight per lane; with more than one b chunk per
        // element (NB > 64) one element at a time keeps the per-column
        // element order
        constexpr int U = (NB > kWave) ? 1 : 16;
        for (int t0 = 0; t0 < ne; t0 += U) {
            double elimv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) elimv[u] = 0.0;
            for (int b0 = 0; b0 < NB; b0 += kWave) {
                const int b = b0 + lane;
                const int bxx = b % NB1, byy = (b / NB1) % NB1, bzz = (D == 3) ? b / (NB1 * NB1) : 0;
                double v[U], el_u[U];
                int slot[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int t = t0 + u;
                    slot[u] = -1;
                    v[u] = 0.0;
                    el_u[u] = 0.0;
                    if (t < ne && b < NB) {
                        int ex, ey, ez;
                        decode(t, ex, ey, ez);
                        const int jx = ex + bxx, jy = ey + byy, jz = ez + bzz;
                        bool fr = jx >= bx.flo[0] && jx <= bx.fhi[0] && jy >= bx.flo[1] && jy <= bx.fhi[1];
                        if (D == 3) fr = fr && jz >= bx.flo[2] && jz <= bx.fhi[2];
                        const int64_t el = ex + (int64_t)bx.nel[0] * (ey + (int64_t)bx.nel[1] * ez) - elbase;
                        const int la = (ic[0] - ex) + NB1 * ((ic[1] - ey) + NB1 * (ic[2] - ez));
                        if (fr) {
                            v[u] = Kbuf[el * (NB * NB) + (int64_t)la * NB + b];
                            slot[u] = (D == 3) ? ((jz - ic[2] + P) * W + (jy - ic[1] + P)) * W + (jx - ic[0] + P)
                                               : (jy - ic[1] + P) * W + (jx - ic[0] + P);
                        } else if (fixed) {
                            // eliminated column: K_ij g_j moves to the rhs
                            const int64_t m = jx + (int64_t)bx.nb[0] * (jy + (int64_t)bx.nb[1] * jz);
                            el_u[u] = Kbuf[el * (NB * NB) + (int64_t)la * NB + b] * fixed[dofs[m] - nfree];
                        }
                    }
                }
                if (fixed) {
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        double e = el_u[u];
#pragma unroll
                        for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, kWave);
                        elimv[u] += e;
                    }
                }
                // element order is kept: the slots of one element are
                // distinct, successive elements are applied one after the other
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (slot[u] >= 0) {
                        sAcc[rr][slot[u]] += v[u];
                        sNz[rr][slot[u]] |= (v[u] != 0.0);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            }
            for (int u = 0; u < U && t0 + u < ne; ++u) {
                const int t = t0 + u;
                int ex, ey, ez;
                decode(t, ex, ey, ez);
                const int64_t el = ex + (int64_t)bx.nel[0] * (ey + (int64_t)bx.nel[1] * ez) - elbase;
                const int la = (ic[0] - ex) + NB1 * ((ic[1] - ey) + NB1 * (ic[2] - ez));
                if (fixed) racc -= elimv[u];
                racc += Fbuf[el * NB + la];
            }
        }
        __builtin_amdgcn_wave_barrier();
        // CSR row: window clipped to the free box, lexicographic
        int lo[3] = {0, 0, 0}, wd[3] = {1, 1, 1};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            lo[k] = max(bx.flo[k], ic[k] - P);
            wd[k] = min(bx.fhi[k], ic[k] + P) - lo[k] + 1;
        }
        const int rb = rowptr[r], len = rowptr[r + 1] - rb;
        for (int q = lane; q < len; q += kWave) {
            const int qx = q % wd[0], qy = (q / wd[0]) % wd[1], qz = (D == 3) ? q / (wd[0] * wd[1]) : 0;
            const int slot = (D == 3)
                ? ((lo[2] + qz - ic[2] + P) * W + (lo[1] + qy - ic[1] + P)) * W + (lo[0] + qx - ic[0] + P)
                : (lo[1] + qy - ic[1] + P) * W + (lo[0] + qx - ic[0] + P);
            val[rb + q] = sAcc[rr][slot];
            if (keep) keep[rb + q] = sNz[rr][slot];
        }
        if (lane == 0) rhs[r] = racc;
    }
    __syncthreads();
    // structured operator (tiled slot-major), R consecutive rows per slot
    const int nr = min(R, r1 - rbase);
    if (s64)
        for (int k = threadIdx.x; k < R * NS; k += 256) {
            const int slot = k / R, rr = k % R;
            if (rr < nr) s64[stencil_index(slot, rbase + rr, NS)] = sAcc[rr][slot];
        }
    if (s32) {
        // upper half only (stencil_half layout)
        constexpr int C = NS / 2, NH = stencil_half(NS);
        for (int k = threadIdx.x; k < R * NH; k += 256) {
            const int h = k / R, rr = k % R;
            if (rr < nr) s32[stencil_index(h, rbase + rr, NH)] = (float)sAcc[rr][C + h];
        }
    }
}

// Single-patch pattern: row r's columns are the free functions of the
// window [i - P, i + P] clipped to the free box, lexicographic (= ascending
// free index).  One wavefront per row, lanes write consecutive columns.
template <int D, int P>
__global__ __launch_bounds__(256) void k_rowfill_patch(PatchBox bx, int nfree, const int *__restrict__ f2g,
                                                       const int *__restrict__ rowptr,
                                                       int *__restrict__ col)
{
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int r = blockIdx.x * 4 + w;
    if (r >= nfree) return;
    const int g = f2g[r];
    int ic[3];
    ic[0] = g % bx.padded[0] - P;
    ic[1] = (g / bx.padded[0]) % bx.padded[1] - P;
    ic[2] = (D == 3) ? g / (bx.padded[0] * bx.padded[1]) - P : 0;
    int lo[3] = {0, 0, 0}, wd[3] = {1, 1, 1}, fw[3] = {1, 1, 1};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        lo[k] = max(bx.flo[k], ic[k] - P);
        wd[k] = min(bx.fhi[k], ic[k] + P) - lo[k] + 1;
        fw[k] = bx.fhi[k] - bx.flo[k] + 1;
    }
    const int rb = rowptr[r], len = rowptr[r + 1] - rb;
    for (int q = lane; q < len; q += kWave) {
        const int qx = q % wd[0], qy = (q / wd[0]) % wd[1], qz = (D == 3) ? q / (wd[0] * wd[1]) : 0;
        const int jx = lo[0] + qx - bx.flo[0], jy = lo[1] + qy - bx.flo[1];
        const int jz = (D == 3) ? lo[2] + qz - bx.flo[2] : 0;
        col[rb + q] = jx + fw[0] * (jy + fw[1] * jz);
    }
}

hipError_t launch_rowfill_patch(int d, int p, const int *flo, const int *fhi, const int *nb,
                                const int *nel, const int *padded, int nfree, const int *f2g,
                                const int *rowptr, int *col, hipStream_t s)
{
    if (nfree <= 0) return hipSuccess;
    PatchBox bx;
    for (int k = 0; k < 3; ++k) {
        bx.flo[k] = flo[k];
        bx.fhi[k] = fhi[k];
        bx.nb[k] = nb[k];
        bx.nel[k] = nel[k];
        bx.padded[k] = padded[k];
    }
    const dim3 grid((nfree + 3) / 4), blk(256);
#define GSX_F(D, P) if (d == D && p == P) { hipLaunchKernelGGL((k_rowfill_patch<D, P>), grid, blk, 0, s, bx, nfree, f2g, rowptr, col); return hipGetLastError(); }
    GSX_F(2, 1) GSX_F(2, 2) GSX_F(2, 3) GSX_F(2, 4)
    GSX_F(3, 1) GSX_F(3, 2) GSX_F(3, 3) GSX_F(3, 4)
#undef GSX_F
    return hipErrorInvalidValue;
}

hipError_t launch_gather_patch(int d, int p, const int *flo, const int *fhi, const int *nb,
                               const int *nel, const int *padded, int nfree, int r0, int r1, int64_t elbase,
                               const int *f2g, const int *rowptr, const double *Kbuf, const double *Fbuf,
                               double *val, double *rhs, unsigned char *keep, double *s64,
                               float *s32, const int *dofs, const double *fixed, hipStream_t s)
{
    if (r1 <= r0) return hipSuccess;
    PatchBox bx;
    for (int k = 0; k < 3; ++k) {
        bx.flo[k] = flo[k];
        bx.fhi[k] = fhi[k];
        bx.nb[k] = nb[k];
        bx.nel[k] = nel[k];
        bx.padded[k] = padded[k];
    }
#define GSX_G(D, P) if (d == D && p == P) { const dim3 grid((r1 - r0 + GatherCfg<D, P>::R - 1) / GatherCfg<D, P>::R), blk(256); hipLaunchKernelGGL((k_gather_patch<D, P>), grid, blk, 0, s, bx, nfree, r0, r1, elbase, f2g, rowptr, Kbuf, Fbuf, val, rhs, keep, s64, s32, dofs, fixed); return hipGetLastError(); }
    GSX_G(2, 1) GSX_G(2, 2) GSX_G(2, 3) GSX_G(2, 4)
    GSX_G(3, 1) GSX_G(3, 2) GSX_G(3, 3) GSX_G(3, 4)
#undef GSX_G
    return hipErrorInvalidValue;
}

// Compaction after the exact-zero rule removed entries: new counts per row.
__global__ void k_keep_count(const int *__restrict__ rowptr, const unsigned char *__restrict__ keep,
                             int n, int *__restrict__ counts)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int c = 0;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) c += keep[k];
    counts[r] = c;
}

__global__ void k_keep_copy(const int *__restrict__ rowptr, const int *__restrict__ col,
                            const double *__restrict__ val, const unsigned char *__restrict__ keep,
                            int n, const int *__restrict__ nrowptr, int *__restrict__ ncol,
                            double *__restrict__ nval)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int o = nrowptr[r];
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
        if (keep[k]) {
            ncol[o] = col[k];
            nval[o] = val[k];
            ++o;
        }
}

__global__ void k_count_zero_flags(const unsigned char *__restrict__ keep, int64_t nnz,
                                   unsigned long long *__restrict__ dropped)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int mine = (i < nnz && keep[i] == 0) ? 1 : 0;
    // wave-level count then one atomic per wave (integer: order-free)
    const unsigned long long bal = __ballot(mine);
    if ((threadIdx.x % kWave) == 0 && bal) atomicAdd(dropped, (unsigned long long)__popcll(bal));
}

// min over v[0..n): stage 1 writes one minimum per block into part[],
// stage 2 (one block) reduces part[0..nb).
__global__ void k_min_stage(const double *__restrict__ v, int64_t n, double *__restrict__ out)
{
    __shared__ double s[256];
    double m = 1e300;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmin(m, v[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------------------
// L2 error of u_h against prod sin(pi x_r): per element sum over the Gauss
// points of w |det J| (u_h - u)^2 (gsExprEvaluator::integral,
// src/gsAssembler/gsExprEvaluator.h:382-450).
// ---------------------------------------------------------------------------
template <int D, int P>
__global__ __launch_bounds__(128) void k_l2err(const PatchDev *__restrict__ patches,
                                               const int *__restrict__ gpatch,
                                               const int64_t *__restrict__ gprefix, int ng,
                                               GaussTab gt, const double *__restrict__ sol,
                                               int nfree, double *__restrict__ out,
                                               double *__restrict__ out_h1,
                                               const double *__restrict__ fixed)
{
    constexpr int NB1 = P + 1, NQ1 = P + 1;
    constexpr int NB = (D == 2) ? NB1 * NB1 : NB1 * NB1 * NB1;
    constexpr int NQ = NB;
    __shared__ double sV[D][NQ1][NB1], sD[D][NQ1][NB1], sW[D][NQ1];
    __shared__ double sC[NB][D], sU[NB];
    __shared__ double sE[128], sH[128];
    const int64_t E = blockIdx.x;
    const int gi = find_patch(gprefix, ng, E);
    const PatchDev &pd = patches[gpatch[gi]];
    const int64_t e = E - gprefix[gi];
    int ei[3] = {0, 0, 0};
    ei[0] = (int)(e % pd.nel[0]);
    ei[1] = (int)((e / pd.nel[0]) % pd.nel[1]);
    if (D == 3) ei[2] = (int)(e / ((int64_t)pd.nel[0] * pd.nel[1]));
    const int t = threadIdx.x;
    basis_tables<D, P>(pd, ei, gt, sV, sD, sW);
    for (int b = t; b < NB; b += 128) {
        const int bx = b % NB1, by = (b / NB1) % NB1, bz = (D == 3) ? b / (NB1 * NB1) : 0;
        const int64_t m = (ei[0] + bx) + (int64_t)pd.nb[0] * ((ei[1] + by) + (int64_t)pd.nb[1] * (ei[2] + bz));
        for (int r = 0; r < D; ++r) sC[b][r] = pd.ctrl[m * D + r];
        const int g = pd.dofs[m];
        sU[b] = (g < nfree) ? sol[g] : (fixed ? fixed[g - nfree] : 0.0);
    }
    __syncthreads();
    double local = 0.0, local_h1 = 0.0;
    for (int q = t; q < NQ; q += 128) {
        const int qx = q % NQ1, qy = (q / NQ1) % NQ1, qz = (D == 3) ? q / (NQ1 * NQ1) : 0;
        double x[3] = {0, 0, 0}, J[3][3] = {{0}}, uh = 0.0, gp[3] = {0, 0, 0};
        for (int b = 0; b < NB; ++b) {
            const int bx = b % NB1, by = (b / NB1) % NB1, bz = (D == 3) ? b / (NB1 * NB1) : 0;
            double v = sV[0][qx][bx] * sV[1][qy][by];
            double g0 = sD[0][qx][bx] * sV[1][qy][by];
            double g1 = sD[1][qy][by] * sV[0][qx][bx];
            double g2 = 0.0;
            if (D == 3) {
                g2 = sD[2][qz][bz] * sV[0][qx][bx] * sV[1][qy][by];
                v *= sV[2][qz][bz];
                g0 *= sV[2][qz][bz];
                g1 *= sV[2][qz][bz];
            }
            uh += sU[b] * v;
            gp[0] += sU[b] * g0;
            gp[1] += sU[b] * g1;
            if (D == 3) gp[2] += sU[b] * g2;
            for (int r = 0; r < D; ++r) {
                x[r] += sC[b][r] * v;
                J[r][0] += sC[b][r] * g0;
                J[r][1] += sC[b][r] * g1;
                if (D == 3) J[r][2] += sC[b][r] * g2;
            }
        }
        double det;
        if (D == 2) det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        else det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                   J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        double w = sW[0][qx] * sW[1][qy];
        if (D == 3) w *= sW[2][qz];
        double ue = 1.0;
        const double pi = 3.14159265358979323846;
        for (int r = 0; r < D; ++r) ue *= sin(pi * x[r]);
        const double diff = uh - ue;
        local += w * (diff * diff * fabs(det));
        if (out_h1) {
            // physical gradient: grad u_h = J^{-T} grad_xi u_h, inv[a][c] = dxi_a/dx_c
            double inv[3][3] = {{0}};
            if (D == 2) {
                inv[0][0] = J[1][1] / det;
                inv[0][1] = -J[0][1] / det;
                inv[1][0] = -J[1][0] / det;
                inv[1][1] = J[0][0] / det;
            } else {
                inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
                inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
                inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
                inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
                inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
                inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
                inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
                inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
                inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
            }
            double d2 = 0.0;
            for (int c = 0; c < D; ++c) {
                double gh = 0.0;
                for (int a = 0; a < D; ++a) gh += gp[a] * inv[a][c];
                double ge = pi * cos(pi * x[c]);
                for (int r = 0; r < D; ++r)
                    if (r != c) ge *= sin(pi * x[r]);
                d2 += (gh - ge) * (gh - ge);
            }
            local_h1 += w * (d2 * fabs(det));
        }
    }
    sE[t] = local;
    sH[t] = local_h1;
    __syncthreads();
    if (t == 0) {
        double s = 0.0, h = 0.0;
        for (int i = 0; i < 128; ++i) {
            s += sE[i];
            h += sH[i];
        }
        out[pd.elem_off + e] = s;
        if (out_h1) out_h1[pd.elem_off + e] = h;
    }
}

// ---------------------------------------------------------------------------
// Boundary element terms.  Quadrature of gsGaussRule(basis, 1, 1, dir)
// (src/gsAssembler/gsGaussRule.hpp:23-60, mapped by gsQuadRule::mapTo,
// gsQuadRule.h:177-201): p+1 Gauss nodes along each tangential direction,
// one node of weight 1 on the side across it.  Per element, over the
// (p+1)^(d-1) functions of the boundary layer (the others vanish on the
// side):
//   Dirichlet (L2 projection, gsDirichletValuesByL2Projection,
//   src/gsAssembler/gsDirichletValues.h:257-435): M_ij = sum_q (w_q m_q N_i)
//   N_j and F_i = sum_q (w_q m_q N_i) g(x_q), where m_q = |det J| (the map
//   data carries no side there, so the measure is the volume Jacobian's);
//   Neumann (gsExprAssembler::assembleBdr with u * g_N.tr() * nv(G),
//   src/gsAssembler/gsExprAssembler.h:835-895): F_i = sum_q (w_q N_i)
//   (g_N(x_q) . nv_q), nv the outer normal of gsFunction::computeMap
//   (src/gsCore/gsFunction.hpp:613-680): component c is alt_c times the
//   minor of J^T without row dir and column c, alt alternating from
//   sideOrientation(side) times the sign of det J (first node of the element
//   with |det J| >= 1e-7).
// ---------------------------------------------------------------------------
__device__ inline double bc_value(int kind, const double *val, int d, const double *x)
{
    if (kind == 1) {
        const double pi = 3.14159265358979323846;
        double v = 1.0;
        for (int r = 0; r < d; ++r) v *= sin(pi * x[r]);
        return v;
    }
    if (kind == 2) {
        double v = val[0];
        for (int r = 0; r < d; ++r) v += val[1 + r] * x[r];
        return v;
    }
    return 0.0;
}

__device__ inline void bc_flux(int kind, const double *val, int d, const double *x, double *f)
{
    const double pi = 3.14159265358979323846;
    for (int c = 0; c < d; ++c) {
        if (kind == 1) {
            double v = pi * cos(pi * x[c]);
            for (int r = 0; r < d; ++r)
                if (r != c) v *= sin(pi * x[r]);
            f[c] = v;
        } else if (kind == 2) {
            f[c] = val[1 + c];
        } else {
            f[c] = 0.0;
        }
    }
}

template <int D, int P>
__global__ __launch_bounds__(64) void k_bdry(const PatchDev *__restrict__ patches,
                                             const BdryItem *__restrict__ items, GaussTab gt,
                                             double *__restrict__ Mout, double *__restrict__ Fout)
{
    constexpr int NB1 = P + 1;
    constexpr int NB = (D == 3) ? NB1 * NB1 * NB1 : NB1 * NB1;
    constexpr int NBB = (D == 3) ? NB1 * NB1 : NB1;
    __shared__ double sV[D][NB1][NB1], sD[D][NB1][NB1], sW[D][NB1];
    __shared__ double sC[NB][D];
    __shared__ double sNq[NBB][NBB], sWq[NBB], sGq[NBB], sDet[NBB];
    __shared__ double sJ[NBB][D][D];
    const BdryItem it = items[blockIdx.x];
    const PatchDev &pd = patches[it.patch];
    const int dir = (it.side - 1) / 2, par = (it.side - 1) % 2;
    const int t = threadIdx.x;
    if (t < D * NB1) {
        const int k = t / NB1, q = t % NB1;
        const double *U = pd.knots[k];
        const int span = P + it.ei[k];
        double u, w;
        if (k == dir) {
            u = par ? U[span + 1] : U[span];
            w = 1.0;
        } else {
            const double a = U[span], b = U[span + 1];
            const double h = (b - a) / 2.0;
            u = h * (gt.x[NB1][q] + 1.0) + a;
            w = (0.0 == h ? 0.5 : h) * gt.w[NB1][q];
        }
        double N[P + 1], dN[P + 1];
        deboor_ders<P>(U, span, u, N, dN);
#pragma unroll
        for (int i = 0; i <= P; ++i) {
            sV[k][q][i] = N[i];
            sD[k][q][i] = dN[i];
        }
        sW[k][q] = w;
    }
    for (int b = t; b < NB; b += 64) {
        const int bx = b % NB1, by = (b / NB1) % NB1, bz = (D == 3) ? b / (NB1 * NB1) : 0;
        const int64_t m = (it.ei[0] + bx) + (int64_t)pd.nb[0] * ((it.ei[1] + by) + (int64_t)pd.nb[1] * (it.ei[2] + bz));
        for (int r = 0; r < D; ++r) sC[b][r] = pd.ctrl[m * D + r];
    }
    __syncthreads();
    // tangential directions and their node index of boundary point q
    int tdir[2] = {0, 0}, nt = 0;
    for (int k = 0; k < D; ++k)
        if (k != dir) tdir[nt++] = k;
    double x[3] = {0, 0, 0};
    if (t < NBB) {
        int qd[3] = {0, 0, 0};
        qd[tdir[0]] = t % NB1;
        if (D == 3) qd[tdir[1]] = t / NB1;
        double J[3][3] = {{0}};
        for (int b = 0; b < NB; ++b) {
            const int bi[3] = {b % NB1, (b / NB1) % NB1, (D == 3) ? b / (NB1 * NB1) : 0};
            double v = 1.0, g[3] = {1.0, 1.0, 1.0};
            for (int k = 0; k < D; ++k) {
                v *= sV[k][qd[k]][bi[k]];
                for (int c = 0; c < D; ++c) g[c] *= (c == k) ? sD[k][qd[k]][bi[k]] : sV[k][qd[k]][bi[k]];
            }
            for (int r = 0; r < D; ++r) {
                x[r] += sC[b][r] * v;
                for (int c = 0; c < D; ++c) J[r][c] += sC[b][r] * g[c];
            }
        }
        double det;
        if (D == 2) det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        else det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
                   J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        sDet[t] = det;
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) sJ[t][r][c] = J[r][c];
        double w = sW[tdir[0]][qd[tdir[0]]];
        if (D == 3) w *= sW[tdir[1]][qd[tdir[1]]];
        sWq[t] = (it.type == 1) ? w * fabs(det) : w;
        if (it.type == 1) sGq[t] = bc_value(it.kind, it.value, D, x);
        // boundary-layer functions: index in dir fixed at 0 (par 0) or P
        for (int jb = 0; jb < NBB; ++jb) {
            int bi[3];
            bi[dir] = par ? P : 0;
            bi[tdir[0]] = jb % NB1;
            if (D == 3) bi[tdir[1]] = jb / NB1;
            double v = 1.0;
            for (int k = 0; k < D; ++k) v *= sV[k][qd[k]][bi[k]];
            sNq[t][jb] = v;
        }
    }
    __syncthreads();
    if (it.type == 2 && t < NBB) {
        // outer normal (unnormalized): sign of det J from the element's first
        // node with |det J| >= 1e-7 (else the patch orientation)
        double dsg = 0.0;
        for (int q = 0; q < NBB; ++q)
            if (fabs(sDet[q]) >= 1e-7) {
                dsg = sDet[q] < 0 ? -1.0 : 1.0;
                break;
            }
        if (dsg == 0.0) dsg = pd.orient < 0 ? -1.0 : 1.0;
        const int s = it.side;
        const double so = ((s + (s + 1) / 2) % 2) ? 1.0 : -1.0;
        double alt = so * dsg, nv[3] = {0, 0, 0};
        // jacT[a][b] = J[b][a] (row: parametric direction, column: coordinate)
        for (int c = 0; c < D; ++c) {
            double mdet;
            if (D == 2) {
                const int a = 1 - dir, b = 1 - c;
                mdet = sJ[t][b][a];
            } else {
                int ra[2], cb[2], na = 0, nb = 0;
                for (int a = 0; a < 3; ++a)
                    if (a != dir) ra[na++] = a;
                for (int b = 0; b < 3; ++b)
                    if (b != c) cb[nb++] = b;
                mdet = sJ[t][cb[0]][ra[0]] * sJ[t][cb[1]][ra[1]] - sJ[t][cb[1]][ra[0]] * sJ[t][cb[0]][ra[1]];
            }
            nv[c] = alt * mdet;
            alt = -alt;
        }
        double f[3];
        bc_flux(it.kind, it.value, D, x, f);
        double g = 0.0;
        for (int c = 0; c < D; ++c) g += f[c] * nv[c];
        sGq[t] = g;
    }
    __syncthreads();
    if (it.type == 1)
        for (int idx = t; idx < NBB * NBB; idx += 64) {
            const int i = idx / NBB, j = idx % NBB;
            double sacc = 0.0;
            for (int q = 0; q < NBB; ++q) sacc += (sWq[q] * sNq[q][i]) * sNq[q][j];
            Mout[(int64_t)blockIdx.x * NBB * NBB + idx] = sacc;
        }
    for (int i = t; i < NBB; i += 64) {
        double sacc = 0.0;
        for (int q = 0; q < NBB; ++q) sacc += (sWq[q] * sNq[q][i]) * sGq[q];
        Fout[(int64_t)blockIdx.x * NBB + i] = sacc;
    }
}

__global__ void k_sum_serial(const double *__restrict__ v, int64_t n, double *__restrict__ out)
{
    // deterministic two-level sum: 256 strided partials then a fixed tree
    __shared__ double s[256];
    double a = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) a += v[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_elem(int d, int p, const PatchDev *patches, const int *gpatch,
                       const int64_t *gprefix, int ng, int64_t nelem, const GaussTab &gt,
                       double *Kbuf, double *Fbuf, double *detmin, hipStream_t s)
{
    return launch_elem_range(d, p, patches, gpatch, gprefix, ng, 0, nelem, gt, Kbuf, Fbuf, detmin, s);
}

hipError_t launch_elem_range(int d, int p, const PatchDev *patches, const int *gpatch,
                             const int64_t *gprefix, int ng, int64_t e0, int64_t e1, const GaussTab &gt,
                             double *Kbuf, double *Fbuf, double *detmin, hipStream_t s)
{
    const int64_t nelem = e1 - e0;
    if (nelem <= 0) return hipSuccess;
    const int nb = d == 3 ? (p + 1) * (p + 1) * (p + 1) : (p + 1) * (p + 1);
    ElemRange rg;
    rg.e0 = e0;
    rg.kshift = e0 * (int64_t)nb * nb;
    rg.fshift = e0 * (int64_t)nb;
    const dim3 grid((unsigned)nelem);
#define GSX_E2(P) case P: hipLaunchKernelGGL(k_elem2d<P>, grid, dim3(128), 0, s, patches, gpatch, gprefix, ng, gt, Kbuf, Fbuf, detmin, rg); break;
#define GSX_E3(P) case P: hipLaunchKernelGGL(k_elem3d<P>, grid, dim3(256), 0, s, patches, gpatch, gprefix, ng, gt, Kbuf, Fbuf, detmin, rg); break;
    if (d == 2) {
        switch (p) { GSX_E2(1) GSX_E2(2) GSX_E2(3) GSX_E2(4) default: return hipErrorInvalidValue; }
    } else {
        switch (p) { GSX_E3(1) GSX_E3(2) GSX_E3(3) GSX_E3(4) default: return hipErrorInvalidValue; }
    }
#undef GSX_E2
#undef GSX_E3
    return hipGetLastError();
}

hipError_t launch_l2err(int d, int p, const PatchDev *patches, const int *gpatch,
                        const int64_t *gprefix, int ng, int64_t nelem, const GaussTab &gt,
                        const double *sol, int nfree, double *out, double *out_h1, const double *fixed,
                        hipStream_t s)
{
    if (nelem <= 0) return hipSuccess;
    const dim3 grid((unsigned)nelem);
#define GSX_L(D, P) if (d == D && p == P) { hipLaunchKernelGGL((k_l2err<D, P>), grid, dim3(128), 0, s, patches, gpatch, gprefix, ng, gt, sol, nfree, out, out_h1, fixed); return hipGetLastError(); }
    GSX_L(2, 1) GSX_L(2, 2) GSX_L(2, 3) GSX_L(2, 4)
    GSX_L(3, 1) GSX_L(3, 2) GSX_L(3, 3) GSX_L(3, 4)
#undef GSX_L
    return hipErrorInvalidValue;
}

hipError_t launch_rowcount(const PatchDev *patches, const int *occ_ptr, const int *occ_patch,
                           const int64_t *occ_local, int nfree, int *counts, int *err,
                           hipStream_t s)
{
    if (nfree <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rowcount, dim3((nfree + 127) / 128), dim3(128), 0, s, patches, occ_ptr,
                       occ_patch, occ_local, nfree, counts, err);
    return hipGetLastError();
}

hipError_t launch_rowfill(const PatchDev *patches, const int *occ_ptr, const int *occ_patch,
                          const int64_t *occ_local, int nfree, const int *rowptr, int *col,
                          hipStream_t s)
{
    if (nfree <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rowfill, dim3((nfree + 127) / 128), dim3(128), 0, s, patches, occ_ptr,
                       occ_patch, occ_local, nfree, rowptr, col);
    return hipGetLastError();
}

hipError_t launch_scan(const int *counts, int n, int *rowptr, long long *bsum, long long *total,
                       hipStream_t s)
{
    const int nb = (n + kScanBlock - 1) / kScanBlock;
    if (nb > 0)
        hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanBlock), 0, s, counts, n, rowptr, bsum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(64), 0, s, bsum, nb, total);
    hipLaunchKernelGGL(k_scan_add, dim3((n + 1 + 255) / 256), dim3(256), 0, s, rowptr, n, bsum, total);
    return hipGetLastError();
}

int gather_maxrow_bucket(int maxrow)
{
    if (maxrow <= 128) return 128;
    if (maxrow <= 512) return 512;
    if (maxrow <= 1024) return 1024;
    return -1;
}

hipError_t launch_bdry(int d, int p, const PatchDev *patches, const BdryItem *items, int nitems,
                       const GaussTab &gt, double *M, double *F, hipStream_t s)
{
    if (nitems <= 0) return hipSuccess;
#define GSX_B(D, P) if (d == D && p == P) { hipLaunchKernelGGL((k_bdry<D, P>), dim3(nitems), dim3(64), 0, s, patches, items, gt, M, F); return hipGetLastError(); }
    GSX_B(2, 1) GSX_B(2, 2) GSX_B(2, 3) GSX_B(2, 4)
    GSX_B(3, 1) GSX_B(3, 2) GSX_B(3, 3) GSX_B(3, 4)
#undef GSX_B
    return hipErrorInvalidValue;
}

hipError_t launch_gather(const PatchDev *patches, const int *occ_ptr, const int *occ_patch,
                         const int64_t *occ_local, int nfree, const int *rowptr, const int *col,
                         const double *Kbuf, const double *Fbuf, double *val, double *rhs,
                         unsigned char *keep, int maxrow, const double *fixed, hipStream_t s)
{
    if (nfree <= 0) return hipSuccess;
    const dim3 grid((nfree + 3) / 4), block(256);
    switch (gather_maxrow_bucket(maxrow)) {
    case 128: hipLaunchKernelGGL(k_gather<128>, grid, block, 0, s, patches, occ_ptr, occ_patch, occ_local, nfree, rowptr, col, Kbuf, Fbuf, val, rhs, keep, fixed); break;
    case 512: hipLaunchKernelGGL(k_gather<512>, grid, block, 0, s, patches, occ_ptr, occ_patch, occ_local, nfree, rowptr, col, Kbuf, Fbuf, val, rhs, keep, fixed); break;
    case 1024: hipLaunchKernelGGL(k_gather<1024>, grid, block, 0, s, patches, occ_ptr, occ_patch, occ_local, nfree, rowptr, col, Kbuf, Fbuf, val, rhs, keep, fixed); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_keep(const int *rowptr, const int *col, const double *val,
                       const unsigned char *keep, int n, int64_t nnz, int *counts, int *nrowptr,
                       int *ncol, double *nval, long long *bsum, long long *total,
                       unsigned long long *dropped, int phase, hipStream_t s)
{
    if (phase == 0) {
        if (nnz > 0)
            hipLaunchKernelGGL(k_count_zero_flags, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s,
                               keep, nnz, dropped);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_keep_count, dim3((n + 127) / 128), dim3(128), 0, s, rowptr, keep, n, counts);
    hipError_t e = launch_scan(counts, n, nrowptr, bsum, total, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_keep_copy, dim3((n + 127) / 128), dim3(128), 0, s, rowptr, col, val, keep, n,
                       nrowptr, ncol, nval);
    return hipGetLastError();
}

// out must hold 1 + 256 doubles (the last 256 are block minima)
hipError_t launch_min(const double *v, int64_t n, double *out, hipStream_t s)
{
    int64_t nb = (n + 255) / 256;
    if (nb > 256) nb = 256;
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(k_min_stage, dim3((unsigned)nb), dim3(256), 0, s, v, n, out + 1);
    hipLaunchKernelGGL(k_min_stage, dim3(1), dim3(256), 0, s, out + 1, nb, out);
    return hipGetLastError();
}

hipError_t launch_sum(const double *v, int64_t n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_serial, dim3(1), dim3(256), 0, s, v, n, out);
    return hipGetLastError();
}

}  // namespace gsx
