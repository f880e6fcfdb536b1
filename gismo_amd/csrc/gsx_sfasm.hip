// Sum-factorised assembly of single 3D patches (libgsx, gfx950): the
// stiffness matrix row by row, without element matrices.
//
// Reference: the bilinear form igrad(u,G)*igrad(u,G).tr()*meas(G) [* kappa]
// of gsExprAssembler (src/gsAssembler/gsExpressions.h:2737-3890) summed over
// the elements and Gauss points of the patch (gsExprAssembler.h:754-833),
// and the load vector u*ff*meas(G) (gsExpressions.h:946-970) pushed into the
// rhs (gsExprAssembler.h:553-630).
//
// For a tensor-product basis the entry of row i = (ix, iy, iz) and column
// j = i + o is
//   A(i, j) = sum_ab sum_g D^ab(g) X^ab(ix, ox, gx) Y^ab(iy, oy, gy) Z^ab(iz, oz, gz)
// where g = (gx, gy, gz) runs over the Gauss points of the patch (g_k =
// e_k (p+1) + q_k), D^ab(g) = w_g kappa_e |det J_g| (J^-1 J^-T)_ab, and
// X^ab(i, o, g) = B^[a==0](i, g) B^[b==0](i+o, g) with B^0 the 1D basis value
// and B^1 its derivative (Y, Z likewise for directions 1, 2).  This is the
// element-by-element sum grad(N_i) D grad(N_j) regrouped: the same terms,
// summed point column by point column instead of element by element.
//
// A block owns a 2 x 2 tile of basis rows in (y, z) and walks the elements
// along x.  Per Gauss point column gx it computes the geometry of its (y, z)
// support by sum factorisation over the control points, then D, then
// contracts z, then y (folding the nine (a, b) pairs into the four classes
// of their x flags), then x into the accumulators of the p + 1 rows
// ix = e .. e + p that element e touches.  Row e is complete after element e
// and leaves the registers.  Only the upper half of the stencil (o >= 0
// lexicographically in (z, y, x)) is accumulated; each value is stored at
// (i, o) and at its mirror (i + o, -o), so the FP64 stencil is bitwise
// symmetric.  All basis rows are computed (eliminated ones too) so that every
// mirror exists; only free rows are stored.
#include "gsx_common.h"
#include "gsx_kernels.h"
#include "gsx_basis.h"

namespace gsx {

template <int P>
struct SfCfg {
    static constexpr int NB1 = P + 1, W = 2 * P + 1, NS = W * W * W;
    static constexpr int TY = 2, TZ = 2, NT = TY * TZ;
    static constexpr int NEY = TY + P, NGY = NEY * NB1, NCY = NEY + P;   // y support: elements, points, ctrl
    static constexpr int NEZ = TZ + P, NGZ = NEZ * NB1, NCZ = NEZ + P;
    static constexpr int NYZ = W * P + P + 1;          // upper (oy, oz) pairs: oz > 0, or oz = 0 and oy >= 0
    static constexpr int KPR = (NYZ - 1) * W + NB1;    // upper (oy, oz, ox) per row
    static constexpr int NCOMBO = NT * KPR;
    static constexpr int CPT = (NCOMBO + 255) / 256;   // accumulator combos per thread
};

// (oy, oz) of upper pair u: u <= P: (u, 0); else oz = 1 + (u - P - 1) / W,
// oy = (u - P - 1) % W - P.
template <int P>
__device__ inline void sf_pair(int u, int &oy, int &oz)
{
    constexpr int W = 2 * P + 1;
    if (u <= P) {
        oy = u;
        oz = 0;
    } else {
        oz = 1 + (u - P - 1) / W;
        oy = (u - P - 1) % W - P;
    }
}

// 1D tables of the patch's Gauss points, per direction k at offset off_k:
// [g][m] values, then [g][m] derivatives (G_k x (p+1) each), then [g]
// mapped weights, for g = e (p+1) + q (the quantities of basis_tables).
template <int P>
__global__ void k_sf_tables(const PatchDev *__restrict__ patches, GaussTab gt, double *__restrict__ tab, int g0,
                            int g1, int g2)
{
    constexpr int NB1 = P + 1;
    const PatchDev &pd = patches[0];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int G[3] = {g0, g1, g2};
    int k = 0;
    int64_t off = 0;
    while (k < 3 && t >= G[k]) {
        off += (int64_t)G[k] * (2 * NB1 + 1);
        t -= G[k];
        ++k;
    }
    if (k >= 3) return;
    const int e = t / NB1, q = t % NB1;
    const double *U = pd.knots[k];
    const int span = P + e;
    const double a = U[span], b = U[span + 1];
    const double h = (b - a) / 2.0;
    const double u = h * (gt.x[NB1][q] + 1.0) + a;
    double N[NB1], dN[NB1];
    deboor_ders<P>(U, span, u, N, dN);
#pragma unroll
    for (int m = 0; m < NB1; ++m) {
        tab[off + (int64_t)t * NB1 + m] = N[m];
        tab[off + (int64_t)G[k] * NB1 + (int64_t)t * NB1 + m] = dN[m];
    }
    tab[off + 2 * (int64_t)G[k] * NB1 + t] = (0.0 == h ? 0.5 : h) * gt.w[NB1][q];
}

template <int P>
__global__ __launch_bounds__(256) void k_sfasm3d(const PatchDev *__restrict__ patches,
                                                 const double *__restrict__ tab, SfBox bx,
                                                 double *__restrict__ s64)
{
    using C = SfCfg<P>;
    constexpr int NB1 = C::NB1, W = C::W, NS = C::NS, TY = C::TY, TZ = C::TZ, NT = C::NT;
    constexpr int NGY = C::NGY, NGZ = C::NGZ, NCY = C::NCY, NCZ = C::NCZ, NYZ = C::NYZ, KPR = C::KPR;
    constexpr int NCOMBO = C::NCOMBO, CPT = C::CPT;
    __shared__ double sVy[NGY][NB1], sDy[NGY][NB1], sWy[NGY];
    __shared__ double sVz[NGZ][NB1], sDz[NGZ][NB1], sWz[NGZ];
    __shared__ double sVx[NB1][NB1], sDx[NB1][NB1], sWx[NB1];
    __shared__ double sC[NB1][NCZ][NCY][3];
    __shared__ double sK[C::NEZ][C::NEY];
    __shared__ double sGX[NCZ][NCY][6];
    __shared__ double sGY[NCZ][NGY][9];
    __shared__ double sD[6][NGZ][NGY];
    __shared__ double sT1[8][NGY][TZ][NB1];
    __shared__ double sT3[4][NT][NYZ];

    const PatchDev &pd = patches[0];
    const int t = threadIdx.x;
    const int ty = blockIdx.x % bx.nty, tz = blockIdx.x / bx.nty;
    const int Y0 = ty * TY, Z0 = tz * TZ;
    const int nbx = bx.nb[0], nby = bx.nb[1], nbz = bx.nb[2];
    const int nelx = bx.nel[0], nely = bx.nel[1], nelz = bx.nel[2];
    const int ey_lo = max(0, Y0 - P), ey_hi = min(nely - 1, Y0 + TY - 1);
    const int ez_lo = max(0, Z0 - P), ez_hi = min(nelz - 1, Z0 + TZ - 1);
    const int ney = ey_hi - ey_lo + 1, nez = ez_hi - ez_lo + 1;
    const int ngy = ney * NB1, ngz = nez * NB1, ncy = ney + P, ncz = nez + P;
    const double *tx = tab + bx.toff[0], *tyb = tab + bx.toff[1], *tzb = tab + bx.toff[2];
    const int Gx = bx.g[0], Gy = bx.g[1], Gz = bx.g[2];

    for (int k = t; k < ngy * NB1; k += 256) {
        const int gy = k / NB1, m = k % NB1;
        const int g = ey_lo * NB1 + gy;
        sVy[gy][m] = tyb[(int64_t)g * NB1 + m];
        sDy[gy][m] = tyb[(int64_t)Gy * NB1 + (int64_t)g * NB1 + m];
    }
    for (int k = t; k < ngz * NB1; k += 256) {
        const int gz = k / NB1, m = k % NB1;
        const int g = ez_lo * NB1 + gz;
        sVz[gz][m] = tzb[(int64_t)g * NB1 + m];
        sDz[gz][m] = tzb[(int64_t)Gz * NB1 + (int64_t)g * NB1 + m];
    }
    for (int k = t; k < ngy; k += 256) sWy[k] = tyb[2 * (int64_t)Gy * NB1 + ey_lo * NB1 + k];
    for (int k = t; k < ngz; k += 256) sWz[k] = tzb[2 * (int64_t)Gz * NB1 + ez_lo * NB1 + k];

    // accumulators: combo c = t + 256 ci -> row (iyl, izl) = c / KPR, pair
    // and ox from c % KPR; acc[ci][r] belongs to row ix = e + r
    double acc[CPT][NB1];
    int cu[CPT], cox[CPT], crow[CPT];
#pragma unroll
    for (int ci = 0; ci < CPT; ++ci) {
        const int c = t + 256 * ci;
        crow[ci] = -1;
        cu[ci] = 0;
        cox[ci] = 0;
        if (c < NCOMBO) {
            crow[ci] = c / KPR;
            const int k = c % KPR;
            if (k < NB1) {
                cu[ci] = 0;
                cox[ci] = k;
            } else {
                cu[ci] = 1 + (k - NB1) / W;
                cox[ci] = (k - NB1) % W - P;
            }
        }
#pragma unroll
        for (int r = 0; r < NB1; ++r) acc[ci][r] = 0.0;
    }

    // store entry (i, o) of the upper half and its mirror
    auto store = [&](int ix, int iy, int iz, int ox, int oy, int oz, double v) {
        if (ix >= nbx || iy >= nby || iz >= nbz) return;
        const int slot = ((oz + P) * W + (oy + P)) * W + (ox + P);
        const bool fi = ix >= bx.flo[0] && ix <= bx.fhi[0] && iy >= bx.flo[1] && iy <= bx.fhi[1] &&
                        iz >= bx.flo[2] && iz <= bx.fhi[2];
        if (fi) {
            const int ri = (ix - bx.flo[0]) + bx.m[0] * ((iy - bx.flo[1]) + bx.m[1] * (iz - bx.flo[2]));
            s64[stencil_index(slot, ri, NS)] = v;
        }
        if (slot == NS / 2) return;
        const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
        const bool fj = jx >= bx.flo[0] && jx <= bx.fhi[0] && jy >= bx.flo[1] && jy <= bx.fhi[1] &&
                        jz >= bx.flo[2] && jz <= bx.fhi[2];
        if (fj) {
            const int rj = (jx - bx.flo[0]) + bx.m[0] * ((jy - bx.flo[1]) + bx.m[1] * (jz - bx.flo[2]));
            s64[stencil_index(NS - 1 - slot, rj, NS)] = v;
        }
    };

    for (int ex = 0; ex < nelx; ++ex) {
        __syncthreads();   // previous element's readers of sVx/sC are done
        for (int k = t; k < NB1 * NB1; k += 256) {
            const int q = k / NB1, m = k % NB1;
            const int g = ex * NB1 + q;
            sVx[q][m] = tx[(int64_t)g * NB1 + m];
            sDx[q][m] = tx[(int64_t)Gx * NB1 + (int64_t)g * NB1 + m];
        }
        if (t < NB1) sWx[t] = tx[2 * (int64_t)Gx * NB1 + ex * NB1 + t];
        for (int k = t; k < NB1 * ncz * ncy * 3; k += 256) {
            const int c = k % 3, by = (k / 3) % ncy, bz = (k / (3 * ncy)) % ncz, bxl = k / (3 * ncy * ncz);
            const int64_t m = (int64_t)(ex + bxl) + (int64_t)nbx * ((ey_lo + by) + (int64_t)nby * (ez_lo + bz));
            sC[bxl][bz][by][c] = pd.ctrl[m * 3 + c];
        }
        for (int k = t; k < nez * ney; k += 256) {
            const int eyl = k % ney, ezl = k / ney;
            sK[ezl][eyl] = pd.kappa ? pd.kappa[ex + (int64_t)nelx * ((ey_lo + eyl) + (int64_t)nely * (ez_lo + ezl))]
                                    : 1.0;
        }
        __syncthreads();
        for (int qx = 0; qx < NB1; ++qx) {
            // A: x contraction of the control points (value and derivative)
            for (int k = t; k < ncz * ncy * 6; k += 256) {
                const int c = k % 3, f = (k / 3) % 2, by = (k / 6) % ncy, bz = k / (6 * ncy);
                double s = 0.0;
#pragma unroll
                for (int bxl = 0; bxl < NB1; ++bxl) s += sC[bxl][bz][by][c] * (f ? sDx[qx][bxl] : sVx[qx][bxl]);
                sGX[bz][by][f * 3 + c] = s;
            }
            __syncthreads();
            // B: y contraction: (xv yv), (xd yv), (xv yd) per coordinate
            for (int k = t; k < ncz * ngy * 9; k += 256) {
                const int c = k % 3, cb = (k / 3) % 3, gy = (k / 9) % ngy, bz = k / (9 * ngy);
                const int eyl = gy / NB1;
                const int fx = cb == 1, fy = cb == 2;
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < NB1; ++m)
                    s += sGX[bz][eyl + m][fx * 3 + c] * (fy ? sDy[gy][m] : sVy[gy][m]);
                sGY[bz][gy][cb * 3 + c] = s;
            }
            __syncthreads();
            // C: z contraction -> Jacobian at every (gy, gz), then D
            for (int k = t; k < ngy * ngz; k += 256) {
                const int gy = k % ngy, gz = k / ngy;
                const int ezl = gz / NB1, eyl = gy / NB1;
                double J[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double j0 = 0.0, j1 = 0.0, j2 = 0.0;
#pragma unroll
                    for (int m = 0; m < NB1; ++m) {
                        const double vz = sVz[gz][m], dz = sDz[gz][m];
                        j0 += sGY[ezl + m][gy][3 + c] * vz;
                        j1 += sGY[ezl + m][gy][6 + c] * vz;
                        j2 += sGY[ezl + m][gy][c] * dz;
                    }
                    J[c][0] = j0;
                    J[c][1] = j1;
                    J[c][2] = j2;
                }
                const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
                const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
                const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
                const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
                double inv[3][3];
                inv[0][0] = c00 / det;
                inv[1][0] = c01 / det;
                inv[2][0] = c02 / det;
                inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
                inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
                inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
                inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
                inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
                inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
                const double w = sWx[qx] * sWy[gy] * sWz[gz];
                double cq = w * fabs(det);
                if (pd.kappa) cq *= sK[ezl][eyl];
                int s = 0;
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = a; b < 3; ++b) {
                        const double m = inv[a][0] * inv[b][0] + inv[a][1] * inv[b][1] + inv[a][2] * inv[b][2];
                        sD[s++][gz][gy] = cq * m;
                    }
            }
            __syncthreads();
            // D: z contraction of D for oz >= 0.  Arrays: (a,b) = 00, 01 (= 10),
            // 11 with z flags (0,0); 20, 21 with (1,0); 02, 12 with (0,1); 22.
            for (int k = t; k < 8 * ngy * TZ * NB1; k += 256) {
                const int oz = k % NB1, izl = (k / NB1) % TZ, gy = (k / (NB1 * TZ)) % ngy, arr = k / (NB1 * TZ * ngy);
                const int iz = Z0 + izl;
                const int sidx = arr == 0 ? 0 : arr == 1 ? 1 : arr == 2 ? 3 : (arr == 3 || arr == 5) ? 2
                                 : (arr == 4 || arr == 6) ? 4 : 5;
                const int fa = arr == 3 || arr == 4 || arr == 7, fb = arr == 5 || arr == 6 || arr == 7;
                double s = 0.0;
                if (iz < nbz) {
                    const int e0 = max(ez_lo, iz + oz - P), e1 = min(ez_hi, iz);
                    for (int ez = e0; ez <= e1; ++ez)
#pragma unroll
                        for (int q = 0; q < NB1; ++q) {
                            const int gz = (ez - ez_lo) * NB1 + q;
                            const double ba = fa ? sDz[gz][iz - ez] : sVz[gz][iz - ez];
                            const double bb = fb ? sDz[gz][iz + oz - ez] : sVz[gz][iz + oz - ez];
                            s += (ba * bb) * sD[sidx][gz][gy];
                        }
                }
                sT1[arr][gy][izl][oz] = s;
            }
            __syncthreads();
            // E: y contraction, (a, b) folded into the x-flag classes
            // 0: (0,0); 1: (0,1),(0,2); 2: (1,0),(2,0); 3: (1,1),(1,2),(2,1),(2,2)
            for (int k = t; k < 4 * NT * NYZ; k += 256) {
                const int u = k % NYZ, trow = (k / NYZ) % NT, cls = k / (NYZ * NT);
                const int iyl = trow % TY, izl = trow / TY;
                const int iy = Y0 + iyl, iz = Z0 + izl;
                int oy, oz;
                sf_pair<P>(u, oy, oz);
                double s = 0.0;
                if (iy < nby && iz < nbz) {
                    const int e0 = max(ey_lo, max(iy, iy + oy) - P), e1 = min(ey_hi, min(iy, iy + oy));
                    const int npair = cls == 0 ? 1 : cls == 3 ? 4 : 2;
                    for (int pp = 0; pp < npair; ++pp) {
                        int a, b;
                        if (cls == 0) { a = 0; b = 0; }
                        else if (cls == 1) { a = 0; b = 1 + pp; }
                        else if (cls == 2) { a = 1 + pp; b = 0; }
                        else { a = 1 + (pp >> 1); b = 1 + (pp & 1); }
                        const int arr = (a == 0 && b == 0) ? 0 : (a + b == 1) ? 1 : (a == 1 && b == 1) ? 2
                                        : (a == 2 && b == 0) ? 3 : (a == 2 && b == 1) ? 4 : (a == 0 && b == 2) ? 5
                                        : (a == 1 && b == 2) ? 6 : 7;
                        const int fa = a == 1, fb = b == 1;
                        for (int ey = e0; ey <= e1; ++ey)
#pragma unroll
                            for (int q = 0; q < NB1; ++q) {
                                const int gy = (ey - ey_lo) * NB1 + q;
                                const double ba = fa ? sDy[gy][iy - ey] : sVy[gy][iy - ey];
                                const double bb = fb ? sDy[gy][iy + oy - ey] : sVy[gy][iy + oy - ey];
                                s += (ba * bb) * sT1[arr][gy][izl][oz];
                            }
                    }
                }
                sT3[cls][trow][u] = s;
            }
            __syncthreads();
            // F: x contraction into the rows e .. e + P
#pragma unroll
            for (int ci = 0; ci < CPT; ++ci) {
                if (crow[ci] < 0) continue;
                const double t0 = sT3[0][crow[ci]][cu[ci]], t1 = sT3[1][crow[ci]][cu[ci]];
                const double t2 = sT3[2][crow[ci]][cu[ci]], t3 = sT3[3][crow[ci]][cu[ci]];
                const int ox = cox[ci];
#pragma unroll
                for (int r = 0; r < NB1; ++r) {
                    const int r2 = r + ox;
                    if (r2 < 0 || r2 > P) continue;
                    const double va = sVx[qx][r], da = sDx[qx][r], vb = sVx[qx][r2], db = sDx[qx][r2];
                    acc[ci][r] += (((da * db) * t0 + (da * vb) * t1) + (va * db) * t2) + (va * vb) * t3;
                }
            }
        }
        // row ix = ex is complete (all rows at the last element)
        const int nrow = ex == nelx - 1 ? NB1 : 1;
#pragma unroll
        for (int ci = 0; ci < CPT; ++ci) {
            if (crow[ci] < 0) continue;
            const int iyl = crow[ci] % TY, izl = crow[ci] / TY;
            int oy, oz;
            sf_pair<P>(cu[ci], oy, oz);
            for (int r = 0; r < nrow; ++r) store(ex + r, Y0 + iyl, Z0 + izl, cox[ci], oy, oz, acc[ci][r]);
#pragma unroll
            for (int r = 0; r < P; ++r) acc[ci][r] = acc[ci][r + 1];
            acc[ci][P] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// The same row assembly with the three contractions on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64; operand layout: A[i][k] in lane i + 16 k, B[k][j]
// in lane j + 16 k, result [i][j] in lane j + 16 (i % 4), element i / 4).
// D comes precomputed per Gauss point (k_load3d).  A block of 8 waves owns a
// 4 x 4 tile of basis rows in (y, z); per Gauss point column gx:
//   z:  T1_arr[gy][c]   = sum_gz D_s(arr)[gz][gy] Wz_f(arr)[gz][c]        (c = (izl, oz >= 0))
//   y:  T3_cls[e][c]    = sum_(a,b) in cls sum_gy Wy_f(a,b)[e][gy] T1_arr(a,b)[gy][c]   (e = (iyl, oy))
//   x:  acc[(ox, r)][(e, c)] += sum_cls X_cls[(ox, r)] T3_cls[e][c]
// with Wz, Wy the products of the 1D basis values/derivatives of row and
// column (x-independent, built once per block) and X_cls the same products
// along x for the current point.  The x result rows are i = 4 r + g with
// ox = 4 t + g - p: element r of a lane's result register is the ring
// position (row ix = e + r), so moving to the next element shifts the four
// elements of each register -- no data crosses lanes.
// ---------------------------------------------------------------------------
typedef double sf_d4 __attribute__((ext_vector_type(4)));

template <int P>
struct SfmCfg {
    static constexpr int NB1 = P + 1, W = 2 * P + 1, NS = W * W * W;
    // row tile in (y, z): 4 x 4 up to degree 3; 2 x 3 at degree 4 (the z
    // stage's 16 columns hold TZ (p + 1) <= 16, and the LDS arrays fit)
    static constexpr int TY = P == 4 ? 2 : 4, TZ = P == 4 ? 3 : 4;
    static constexpr int NGY = (TY + P) * NB1, NGZ = (TZ + P) * NB1;
    static constexpr int KZ = (NGZ + 3) / 4 * 4;       // z contraction depth (gz)
    static constexpr int MY = (NGY + 15) / 16 * 16;    // z stage rows (gy)
    static constexpr int KY = (NGY + 3) / 4 * 4;       // y contraction depth (gy)
    static constexpr int NE = TY * W;                  // y stage rows (iyl, oy)
    static constexpr int ME = (NE + 15) / 16 * 16;
    static constexpr int NC = TZ * NB1;                // columns (izl, oz >= 0)
    static constexpr int MTF = (W + 3) / 4;            // x stage row tiles (ox)
    static constexpr int NTF = NE;                     // x stage column tiles (one per e)
    static constexpr int NWAVE = 8;
    static constexpr int FPW = (MTF * NTF + NWAVE - 1) / NWAVE;
    static_assert(NC <= 16 && P <= 4, "one column tile; the ring holds at most five rows");
};

template <int P>
__global__ __launch_bounds__(512) void k_sfasm3d_mfma(const double *__restrict__ tab,
                                                      const double *__restrict__ Dq, SfBox bx,
                                                      double *__restrict__ s64)
{
    using C = SfmCfg<P>;
    constexpr int NB1 = C::NB1, W = C::W, NS = C::NS, TY = C::TY, TZ = C::TZ;
    constexpr int KZ = C::KZ, MY = C::MY, KY = C::KY, NE = C::NE, ME = C::ME, NC = C::NC;
    constexpr int MTF = C::MTF, NTF = C::NTF, FPW = C::FPW;
    __shared__ double sWz[4][KZ][16];
    __shared__ double sWy[4][ME][KY];
    __shared__ double sD[6][KZ][MY];
    __shared__ double sT1[8][KY][16];
    __shared__ double sT3[5][ME][16];
    __shared__ double sVx[NB1][NB1], sDx[NB1][NB1];

    const int t = threadIdx.x, w = t / kWave, l = t % kWave;
    const int ty = blockIdx.x % bx.nty, tz = blockIdx.x / bx.nty;
    const int Y0 = ty * TY, Z0 = tz * TZ;
    const int nbx = bx.nb[0], nby = bx.nb[1], nbz = bx.nb[2];
    const int nelx = bx.nel[0], nely = bx.nel[1], nelz = bx.nel[2];
    const int ey_lo = max(0, Y0 - P), ey_hi = min(nely - 1, Y0 + TY - 1);
    const int ez_lo = max(0, Z0 - P), ez_hi = min(nelz - 1, Z0 + TZ - 1);
    const int ngy = (ey_hi - ey_lo + 1) * NB1, ngz = (ez_hi - ez_lo + 1) * NB1;
    const double *tx = tab + bx.toff[0], *tyb = tab + bx.toff[1], *tzb = tab + bx.toff[2];
    const int Gx = bx.g[0], Gy = bx.g[1], Gz = bx.g[2];

    // Wz[f][gz][c]: Bz^fa(iz, gz) Bz^fb(iz + oz, gz), c = izl (p+1) + oz
    for (int k = t; k < 4 * KZ * 16; k += 512) {
        const int c = k % 16, gz = (k / 16) % KZ, f = k / (16 * KZ);
        double v = 0.0;
        if (c < NC && gz < ngz) {
            const int izl = c / NB1, oz = c % NB1;
            const int iz = Z0 + izl, jz = iz + oz, ez = ez_lo + gz / NB1;
            const int m1 = iz - ez, m2 = jz - ez;
            if (iz < nbz && jz < nbz && m1 >= 0 && m1 <= P && m2 >= 0 && m2 <= P) {
                const int64_t g = (int64_t)ez_lo * NB1 + gz;
                const double a = (f & 2) ? tzb[(int64_t)Gz * NB1 + g * NB1 + m1] : tzb[g * NB1 + m1];
                const double b = (f & 1) ? tzb[(int64_t)Gz * NB1 + g * NB1 + m2] : tzb[g * NB1 + m2];
                v = a * b;
            }
        }
        sWz[f][gz][c] = v;
    }
    // Wy[f][e][gy]: By^fa(iy, gy) By^fb(iy + oy, gy), e = iyl (2p+1) + oy + p
    for (int k = t; k < 4 * ME * KY; k += 512) {
        const int gy = k % KY, e = (k / KY) % ME, f = k / (KY * ME);
        double v = 0.0;
        if (e < NE && gy < ngy) {
            const int iyl = e / W, oy = e % W - P;
            const int iy = Y0 + iyl, jy = iy + oy, ey = ey_lo + gy / NB1;
            const int m1 = iy - ey, m2 = jy - ey;
            if (iy < nby && jy >= 0 && jy < nby && m1 >= 0 && m1 <= P && m2 >= 0 && m2 <= P) {
                const int64_t g = (int64_t)ey_lo * NB1 + gy;
                const double a = (f & 2) ? tyb[(int64_t)Gy * NB1 + g * NB1 + m1] : tyb[g * NB1 + m1];
                const double b = (f & 1) ? tyb[(int64_t)Gy * NB1 + g * NB1 + m2] : tyb[g * NB1 + m2];
                v = a * b;
            }
        }
        sWy[f][e][gy] = v;
    }

    // ring positions 0..3 in the four elements of acc; at degree 4 the fifth
    // (row ex + 4, first touched by element ex) accumulates in element 0 of
    // acc4 and moves into element 3 of acc when the ring shifts
    // (only ox <= 0, the x tiles 0 and 1, reach position 4: fi = w + 8 u < 2 NTF)
    constexpr int NA4 = P == 4 ? (2 * NTF + 7) / 8 : 1;
    sf_d4 acc[FPW], acc4[NA4];
#pragma unroll
    for (int u = 0; u < FPW; ++u) acc[u] = sf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int u = 0; u < NA4; ++u) acc4[u] = sf_d4{0.0, 0.0, 0.0, 0.0};
    const int li = l % 16, lk = l / 16;

    // D of point column gc for the tile's (gz, gy) window, zero-padded:
    // each thread holds its DPT entries in registers one column ahead
    // (degree 4 keeps no registers to spare: D goes straight to LDS there)
    constexpr bool PREFETCH = P < 4;
    constexpr int DN = 6 * KZ * MY, DPT = PREFETCH ? (DN + 511) / 512 : 1;
    double dpre[DPT];
    auto d_direct = [&](int64_t gc) {
        for (int k = t; k < DN; k += 512) {
            double v = 0.0;
            if (gc < (int64_t)Gx) {
                const int gy = k % MY, gz = (k / MY) % KZ, sc = k / (MY * KZ);
                if (gy < ngy && gz < ngz) v = Dq[((gc * 6 + sc) * Gz + ez_lo * NB1 + gz) * Gy + ey_lo * NB1 + gy];
            }
            (&sD[0][0][0])[k] = v;
        }
    };
    auto d_load = [&](int64_t gc) {
#pragma unroll
        for (int u = 0; u < DPT; ++u) {
            const int k = t + 512 * u;
            double v = 0.0;
            if (k < DN && gc < (int64_t)Gx) {
                const int gy = k % MY, gz = (k / MY) % KZ, sc = k / (MY * KZ);
                if (gy < ngy && gz < ngz) v = Dq[((gc * 6 + sc) * Gz + ez_lo * NB1 + gz) * Gy + ey_lo * NB1 + gy];
            }
            dpre[u] = v;
        }
    };
    auto d_store = [&]() {
#pragma unroll
        for (int u = 0; u < DPT; ++u) {
            const int k = t + 512 * u;
            if (k < DN) (&sD[0][0][0])[k] = dpre[u];
        }
    };
    auto x_tables = [&](int ex) {
        if (t < NB1 * NB1) {
            const int q = t / NB1, m = t % NB1;
            const int64_t g = (int64_t)ex * NB1 + q;
            sVx[q][m] = tx[g * NB1 + m];
            sDx[q][m] = tx[(int64_t)Gx * NB1 + g * NB1 + m];
        }
    };
    if constexpr (PREFETCH) {
        d_load(0);
        d_store();
        d_load(1);
    } else {
        d_direct(0);
    }
    x_tables(0);
    __syncthreads();

    // y-stage work units: (class, row tile, pair subset); class 3 in two halves
    // (planes 3 and 4 of sT3, summed when the x stage reads them)
    constexpr int NMT = ME / 16, NUNIT = 5 * NMT;
    for (int ex = 0; ex < nelx; ++ex) {
        for (int qx = 0; qx < NB1; ++qx) {
            const int64_t gc = (int64_t)ex * NB1 + qx;
            // z: arrays (a,b) = 00, 01 (= 10), 11 [z flags 00]; 20, 21 [10]; 02, 12 [01]; 22 [11];
            // the wave's row tiles interleaved (independent accumulators)
            {
                constexpr int NT_D = MY / 16;
                const int arr = w;
                const int sidx = arr == 0 ? 0 : arr == 1 ? 1 : arr == 2 ? 3 : (arr == 3 || arr == 5) ? 2
                                 : (arr == 4 || arr == 6) ? 4 : 5;
                const int f = ((arr == 3 || arr == 4 || arr == 7) ? 2 : 0) + ((arr == 5 || arr == 6 || arr == 7) ? 1 : 0);
                sf_d4 a4[NT_D];
#pragma unroll
                for (int mt = 0; mt < NT_D; ++mt) a4[mt] = sf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < KZ / 4; ++ks) {
                    const double bv = sWz[f][4 * ks + lk][li];
#pragma unroll
                    for (int mt = 0; mt < NT_D; ++mt) {
                        const double av = sD[sidx][4 * ks + lk][16 * mt + li];
                        a4[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, a4[mt], 0, 0, 0);
                    }
                }
#pragma unroll
                for (int mt = 0; mt < NT_D; ++mt)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int gy = 16 * mt + 4 * q + lk;
                        if (gy < KY) sT1[arr][gy][li] = a4[mt][q];
                    }
            }
            __syncthreads();
            // sD is free: the next column's D goes in, the one after into registers
            if constexpr (PREFETCH) {
                d_store();
                d_load(gc + 2);
            } else {
                d_direct(gc + 1);
            }
            if (qx == 0) x_tables(ex);   // the previous element's x stage is done
            // y: x-flag classes 0: (0,0); 1: (0,1),(0,2); 2: (1,0),(2,0); 3: (1,1),(1,2) | (2,1),(2,2)
            for (int unit = w; unit < NUNIT; unit += 8) {
                const int plane = (unit < 4 * NMT) ? 1 + unit / NMT : 0;   // units 0..: planes 1,2,3,4 then 0
                const int mt = unit % NMT;
                const int npair = plane == 0 ? 1 : 2;
                sf_d4 a4 = sf_d4{0.0, 0.0, 0.0, 0.0}, b4 = sf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int pp = 0; pp < 2; ++pp) {
                    if (pp >= npair) break;
                    int a, b;
                    if (plane == 0) { a = 0; b = 0; }
                    else if (plane == 1) { a = 0; b = 1 + pp; }
                    else if (plane == 2) { a = 1 + pp; b = 0; }
                    else if (plane == 3) { a = 1; b = 1 + pp; }
                    else { a = 2; b = 1 + pp; }
                    const int arr = (a == 0 && b == 0) ? 0 : (a + b == 1) ? 1 : (a == 1 && b == 1) ? 2
                                    : (a == 2 && b == 0) ? 3 : (a == 2 && b == 1) ? 4 : (a == 0 && b == 2) ? 5
                                    : (a == 1 && b == 2) ? 6 : 7;
                    const int f = (a == 1 ? 2 : 0) + (b == 1 ? 1 : 0);
#pragma unroll
                    for (int ks = 0; ks < KY / 4; ++ks) {
                        const double av = sWy[f][16 * mt + li][4 * ks + lk];
                        const double bv = sT1[arr][4 * ks + lk][li];
                        if (ks & 1) b4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, b4, 0, 0, 0);
                        else a4 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, a4, 0, 0, 0);
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) sT3[plane][16 * mt + 4 * q + lk][li] = a4[q] + b4[q];
            }
            __syncthreads();
            // x: rows i = 4 r + g (ox = 4 t + g - p), depth = the four classes
            double xa[MTF], xa4[MTF];
#pragma unroll
            for (int tt = 0; tt < MTF; ++tt) {
                const int r = li / 4, g = li % 4, ox = 4 * tt + g - P, r2 = r + ox, cls = lk;
                const bool fa = cls == 0 || cls == 1, fb = cls == 0 || cls == 2;
                double v = 0.0;
                if (r <= P && ox <= P && r2 >= 0 && r2 <= P)
                    v = (fa ? sDx[qx][r] : sVx[qx][r]) * (fb ? sDx[qx][r2] : sVx[qx][r2]);
                xa[tt] = v;
                // degree 4: rows g of the second operand hold ring position 4
                const int r4 = P, s2 = r4 + ox;
                double v4 = 0.0;
                if (P == 4 && li < 4 && ox <= P && s2 >= 0 && s2 <= P)
                    v4 = (fa ? sDx[qx][r4] : sVx[qx][r4]) * (fb ? sDx[qx][s2] : sVx[qx][s2]);
                xa4[tt] = v4;
            }
#pragma unroll
            for (int u = 0; u < FPW; ++u) {
                const int fi = w + 8 * u;
                if (fi < MTF * NTF) {
                    const int tt = fi / NTF, n = fi % NTF;
                    const double bv = lk == 3 ? sT3[3][n][li] + sT3[4][n][li] : sT3[lk][n][li];
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[tt], bv, acc[u], 0, 0, 0);
                    if constexpr (P == 4)
                        if (u < NA4 && tt < 2)
                            acc4[u < NA4 ? u : 0] =
                                __builtin_amdgcn_mfma_f64_16x16x4f64(xa4[tt], bv, acc4[u < NA4 ? u : 0], 0, 0, 0);
                }
            }
        }
        // rows complete after this element: ring position 0 (all at the last element)
        const int nrow = ex == nelx - 1 ? NB1 : 1;
#pragma unroll
        for (int u = 0; u < FPW; ++u) {
            const int fi = w + 8 * u;
            if (fi < MTF * NTF) {
                const int tt = fi / NTF, n = fi % NTF;
                const int ox = 4 * tt + lk - P;
                const int iyl = n / W, oy = n % W - P, izl = li / NB1, oz = li % NB1;
                const bool upper = oz > 0 || oy > 0 || (oy == 0 && ox >= 0);
                if (li < NC && ox <= P && upper) {
                    const int iy = Y0 + iyl, iz = Z0 + izl;
                    for (int r = 0; r < nrow; ++r) {
                        const int ix = ex + r;
                        if (ix >= nbx || iy >= nby || iz >= nbz) continue;
                        double v;
                        if constexpr (P == 4) v = r < 4 ? acc[u][r] : (u < NA4 && tt < 2 ? acc4[u < NA4 ? u : 0][0] : 0.0);
                        else v = acc[u][r];
                        const int slot = ((oz + P) * W + (oy + P)) * W + (ox + P);
                        const bool fi_free = ix >= bx.flo[0] && ix <= bx.fhi[0] && iy >= bx.flo[1] &&
                                             iy <= bx.fhi[1] && iz >= bx.flo[2] && iz <= bx.fhi[2];
                        if (fi_free) {
                            const int ri = (ix - bx.flo[0]) + bx.m[0] * ((iy - bx.flo[1]) + bx.m[1] * (iz - bx.flo[2]));
                            s64[stencil_index(slot, ri, NS)] = v;
                        }
                        if (slot != NS / 2) {
                            const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
                            const bool fj = jx >= bx.flo[0] && jx <= bx.fhi[0] && jy >= bx.flo[1] &&
                                            jy <= bx.fhi[1] && jz >= bx.flo[2] && jz <= bx.fhi[2];
                            if (fj) {
                                const int rj = (jx - bx.flo[0]) +
                                               bx.m[0] * ((jy - bx.flo[1]) + bx.m[1] * (jz - bx.flo[2]));
                                s64[stencil_index(NS - 1 - slot, rj, NS)] = v;
                            }
                        }
                    }
                }
                if constexpr (P == 4) {
                    const bool has4 = u < NA4 && tt < 2;
                    acc[u] = sf_d4{acc[u][1], acc[u][2], acc[u][3], has4 ? acc4[u < NA4 ? u : 0][0] : 0.0};
                    if (has4) acc4[u < NA4 ? u : 0] = sf_d4{0.0, 0.0, 0.0, 0.0};
                } else {
                    acc[u] = sf_d4{acc[u][1], acc[u][2], acc[u][3], 0.0};
                }
            }
        }
    }
}

// Element load vectors and the oriented minimum of det J (the load part of
// k_elem3d): the geometry by sum factorisation over the element's control
// points, f at the mapped points, F_e(i) = sum_q w_q ((N_i f) |det J|).
// One block per element; F is stored [i][e].
template <int P>
__global__ __launch_bounds__(128) void k_load3d(const PatchDev *__restrict__ patches, GaussTab gt,
                                                double *__restrict__ Fbuf, double *__restrict__ detmin,
                                                double *__restrict__ Dq, int Gy, int Gz)
{
    constexpr int NB1 = P + 1, NB = NB1 * NB1 * NB1, NQ = NB;
    __shared__ double sV[3][NB1][NB1], sD[3][NB1][NB1], sW[3][NB1];
    __shared__ double sC[NB1][NB1][NB1][3];   // [bz][by][bx][c]
    __shared__ double sGX[NB1][NB1][NB1][6];  // [bz][by][qx][f c]
    __shared__ double sGY[NB1][NB1][NB1][9];  // [bz][qy][qx][cb c]
    __shared__ double sF[NQ], sMeas[NQ], sWq[NQ], sDet[NQ];
    const PatchDev &pd = patches[0];
    const int64_t e = blockIdx.x;
    int ei[3];
    ei[0] = (int)(e % pd.nel[0]);
    ei[1] = (int)((e / pd.nel[0]) % pd.nel[1]);
    ei[2] = (int)(e / ((int64_t)pd.nel[0] * pd.nel[1]));
    const int t = threadIdx.x;
    basis_tables<3, P>(pd, ei, gt, sV, sD, sW);
    for (int k = t; k < NB * 3; k += 128) {
        const int c = k % 3, b = k / 3;
        const int bx = b % NB1, by = (b / NB1) % NB1, bz = b / (NB1 * NB1);
        const int64_t m = (int64_t)(ei[0] + bx) + (int64_t)pd.nb[0] * ((ei[1] + by) + (int64_t)pd.nb[1] * (ei[2] + bz));
        sC[bz][by][bx][c] = pd.ctrl[m * 3 + c];
    }
    __syncthreads();
    for (int k = t; k < NB1 * NB1 * NB1 * 6; k += 128) {
        const int c = k % 3, f = (k / 3) % 2, qx = (k / 6) % NB1, by = (k / (6 * NB1)) % NB1, bz = k / (6 * NB1 * NB1);
        double s = 0.0;
#pragma unroll
        for (int bx = 0; bx < NB1; ++bx) s += sC[bz][by][bx][c] * (f ? sD[0][qx][bx] : sV[0][qx][bx]);
        sGX[bz][by][qx][f * 3 + c] = s;
    }
    __syncthreads();
    for (int k = t; k < NB1 * NB1 * NB1 * 9; k += 128) {
        const int c = k % 3, cb = (k / 3) % 3, qx = (k / 9) % NB1, qy = (k / (9 * NB1)) % NB1, bz = k / (9 * NB1 * NB1);
        const int fx = cb == 1, fy = cb == 2;
        double s = 0.0;
#pragma unroll
        for (int by = 0; by < NB1; ++by) s += sGX[bz][by][qx][fx * 3 + c] * (fy ? sD[1][qy][by] : sV[1][qy][by]);
        sGY[bz][qy][qx][cb * 3 + c] = s;
    }
    __syncthreads();
    for (int q = t; q < NQ; q += 128) {
        const int qx = q % NB1, qy = (q / NB1) % NB1, qz = q / (NB1 * NB1);
        double x[3], J[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double p0 = 0.0, j0 = 0.0, j1 = 0.0, j2 = 0.0;
#pragma unroll
            for (int bz = 0; bz < NB1; ++bz) {
                const double vz = sV[2][qz][bz], dz = sD[2][qz][bz];
                p0 += sGY[bz][qy][qx][c] * vz;
                j0 += sGY[bz][qy][qx][3 + c] * vz;
                j1 += sGY[bz][qy][qx][6 + c] * vz;
                j2 += sGY[bz][qy][qx][c] * dz;
            }
            x[c] = p0;
            J[c][0] = j0;
            J[c][1] = j1;
            J[c][2] = j2;
        }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        sWq[q] = sW[0][qx] * sW[1][qy] * sW[2][qz];
        sMeas[q] = fabs(det);
        sDet[q] = det;
        sF[q] = source_value(pd.fkind, 3, x, pd.felem, e);
        if (Dq) {
            // D_ab = w |det J| kappa (J^-1 J^-T)_ab at the Gauss point, stored
            // as [gx][ab][gz][gy] (gx = e_x (p+1) + q_x, likewise y, z)
            double inv[3][3];
            inv[0][0] = c00 / det;
            inv[1][0] = c01 / det;
            inv[2][0] = c02 / det;
            inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
            inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
            inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
            inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
            inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
            inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
            double cq = sWq[q] * fabs(det);
            if (pd.kappa) cq *= pd.kappa[e];
            const int64_t gx = (int64_t)ei[0] * NB1 + qx, gy = (int64_t)ei[1] * NB1 + qy, gz = (int64_t)ei[2] * NB1 + qz;
            int sidx = 0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = a; b < 3; ++b) {
                    const double m = inv[a][0] * inv[b][0] + inv[a][1] * inv[b][1] + inv[a][2] * inv[b][2];
                    Dq[((gx * 6 + sidx) * Gz + gz) * Gy + gy] = cq * m;
                    ++sidx;
                }
        }
    }
    __syncthreads();
    for (int i = t; i < NB; i += 128) {
        const int bx = i % NB1, by = (i / NB1) % NB1, bz = i / (NB1 * NB1);
        double a = 0.0;
        for (int q = 0; q < NQ; ++q) {
            const int qx = q % NB1, qy = (q / NB1) % NB1, qz = q / (NB1 * NB1);
            const double v = sV[0][qx][bx] * sV[1][qy][by] * sV[2][qz][bz];
            a += sWq[q] * ((v * sF[q]) * sMeas[q]);
        }
        Fbuf[(int64_t)i * gridDim.x + e] = a;   // [local function][element]: coalesced in the finish
    }
    if (t == 0) {
        double m = pd.orient * sDet[0];
        for (int q = 1; q < NQ; ++q) m = fmin(m, pd.orient * sDet[q]);
        detmin[e] = m;
    }
}

template <int P> struct SfFinish { static constexpr int R = P >= 4 ? 8 : 16; };

// Free rows from the stencil: CSR values (free columns in slot order =
// ascending columns), the FP32 upper half, the rhs (element load vectors
// summed in element order, minus K_ij g_j over the eliminated columns), and
// the slots of columns outside the free box set to 0 in the FP64 stencil.
// A block takes R consecutive rows (within one stencil tile): their slots are
// read slot by slot (R consecutive rows each) into LDS, and the CSR values of
// the R rows, one contiguous range, are written entry by entry from there.
template <int P>
__global__ __launch_bounds__(256) void k_sf_finish3d(SfBox bx, int nfree, const int *__restrict__ rowptr,
                                                     double *__restrict__ s64, float *__restrict__ s32,
                                                     double *__restrict__ val, double *__restrict__ rhs,
                                                     const double *__restrict__ Fbuf, int64_t nelem,
                                                     const int *__restrict__ dofs,
                                                     const double *__restrict__ fixed, int *__restrict__ err)
{
    constexpr int NB1 = P + 1, W = 2 * P + 1, NS = W * W * W;
    constexpr int C = NS / 2, NH = stencil_half(NS);
    constexpr int R = SfFinish<P>::R, LD = R + 1;   // LD: odd stride in doubles against bank conflicts
    static_assert(kStencilTile % R == 0, "a block's rows lie in one tile");
    __shared__ double L[NS * LD];
    __shared__ int sBase[R + 1], sLo[R][3], sN[R][3];
    const int r0 = blockIdx.x * R, t = threadIdx.x;
    const int nr = min(R, nfree - r0);
    if (t <= nr) sBase[t] = rowptr[r0 + t];
    if (t < nr) {
        const int r = r0 + t;
        const int i3[3] = {bx.flo[0] + r % bx.m[0], bx.flo[1] + (r / bx.m[0]) % bx.m[1],
                           bx.flo[2] + r / (bx.m[0] * bx.m[1])};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int lo = max(-P, bx.flo[k] - i3[k]), hi = min(P, bx.fhi[k] - i3[k]);
            sLo[t][k] = lo;
            sN[t][k] = hi - lo + 1;
        }
    }
    for (int q = t; q < NS * R; q += 256) {
        const int s = q / R, rr = q % R;
        if (rr < nr) L[s * LD + rr] = s64[stencil_index(s, r0 + rr, NS)];
    }
    __syncthreads();
    if (t < nr && sN[t][0] * sN[t][1] * sN[t][2] != sBase[t + 1] - sBase[t]) atomicOr(err, 1);
    // CSR values: entry k of the block's range -> (row, position) -> slot
    const int b0 = sBase[0], tot = sBase[nr] - b0;
    int rr = 0;
    for (int k = t; k < tot; k += 256) {
        while (rr + 1 < nr && sBase[rr + 1] - b0 <= k) ++rr;
        const int pos = k - (sBase[rr] - b0);
        const int nx = sN[rr][0], ny = sN[rr][1];
        if (pos >= nx * ny * sN[rr][2]) continue;   // row length mismatch, flagged above
        const int ox = sLo[rr][0] + pos % nx, oy = sLo[rr][1] + (pos / nx) % ny, oz = sLo[rr][2] + pos / (nx * ny);
        val[b0 + k] = L[(((oz + P) * W + (oy + P)) * W + (ox + P)) * LD + rr];
    }
    // FP32 upper half; FP64 slots of columns outside the free box to 0
    for (int q = t; q < NS * R; q += 256) {
        const int s = q / R, rw = q % R;
        if (rw >= nr) continue;
        const int ox = s % W - P - sLo[rw][0], oy = (s / W) % W - P - sLo[rw][1], oz = s / (W * W) - P - sLo[rw][2];
        const bool fr = ox >= 0 && ox < sN[rw][0] && oy >= 0 && oy < sN[rw][1] && oz >= 0 && oz < sN[rw][2];
        if (!fr) s64[stencil_index(s, r0 + rw, NS)] = 0.0;
        if (s32 && s >= C) s32[stencil_index(s - C, r0 + rw, NH)] = fr ? (float)L[s * LD + rw] : 0.0f;
    }
    // rhs: thread (row, c, b) sums the loads of the elements (ex, iy - P + b,
    // iz - P + c) touching the row over ex; the row then sums its (c, b)
    // partials in order.  Likewise the eliminated columns' K_ij g_j, by slot
    // classes.
    constexpr int NYZ = NB1 * NB1;
    static_assert(R * NYZ <= 256, "one thread per (row, c, b)");
    __shared__ double sF[R][NYZ], sC[R][NYZ];
    if (t < R * NYZ) {
        const int rw = t % R, j = t / R, c = j / NB1, b = j % NB1;   // consecutive lanes: consecutive rows
        double f = 0.0, g = 0.0;
        if (rw < nr) {
            const int r = r0 + rw;
            const int ix = bx.flo[0] + r % bx.m[0], iy = bx.flo[1] + (r / bx.m[0]) % bx.m[1],
                      iz = bx.flo[2] + r / (bx.m[0] * bx.m[1]);
            const int ey = iy - P + b, ez = iz - P + c;
            if (ey >= 0 && ey < bx.nel[1] && ez >= 0 && ez < bx.nel[2]) {
                const int64_t eyz = (int64_t)bx.nel[0] * (ey + (int64_t)bx.nel[1] * ez);
                const int lyz = NB1 * ((P - b) + NB1 * (P - c));
#pragma unroll
                for (int a = 0; a <= P; ++a) {
                    const int ex = ix - P + a;
                    if (ex >= 0 && ex < bx.nel[0]) f += Fbuf[(int64_t)(lyz + P - a) * nelem + eyz + ex];
                }
            }
            if (fixed) {
                for (int s = j; s < NS; s += NYZ) {
                    const int ox = s % W - P, oy = (s / W) % W - P, oz = s / (W * W) - P;
                    const int jx = ix + ox, jy = iy + oy, jz = iz + oz;
                    const bool inb = jx >= 0 && jx < bx.nb[0] && jy >= 0 && jy < bx.nb[1] && jz >= 0 && jz < bx.nb[2];
                    const bool fr = jx >= bx.flo[0] && jx <= bx.fhi[0] && jy >= bx.flo[1] && jy <= bx.fhi[1] &&
                                    jz >= bx.flo[2] && jz <= bx.fhi[2];
                    if (inb && !fr) {
                        const int64_t m = jx + (int64_t)bx.nb[0] * (jy + (int64_t)bx.nb[1] * jz);
                        g += L[s * LD + rw] * fixed[dofs[m] - nfree];
                    }
                }
            }
        }
        sF[rw][j] = f;
        sC[rw][j] = g;
    }
    __syncthreads();
    if (t < nr) {
        double racc = 0.0, corr = 0.0;
        for (int j = 0; j < NYZ; ++j) {
            racc += sF[t][j];
            corr += sC[t][j];
        }
        rhs[r0 + t] = racc - corr;
    }
}

bool sf_supported(int d, int p) { return d == 3 && p >= 1 && p <= 4; }

int64_t sf_table_doubles(int p, const int *nel)
{
    int64_t n = 0;
    for (int k = 0; k < 3; ++k) n += (int64_t)nel[k] * (p + 1) * (2 * (p + 1) + 1);
    return n;
}

bool sf_mfma(int p) { return p >= 2 && p <= 4; }

int64_t sf_dq_doubles(int p, const int *nel)
{
    if (!sf_mfma(p)) return 0;
    return (int64_t)nel[0] * (p + 1) * 6 * ((int64_t)nel[1] * (p + 1)) * ((int64_t)nel[2] * (p + 1));
}

hipError_t launch_sf_assemble(int p, const PatchDev *patches, const GaussTab &gt, double *tab, SfBox bx,
                              double *Fbuf, double *detmin, double *Dq, double *s64, int64_t nelem, hipStream_t s)
{
    const int G = bx.g[0] + bx.g[1] + bx.g[2];
#define GSX_SF(PP)                                                                                              \
    if (p == PP) {                                                                                              \
        hipLaunchKernelGGL((k_sf_tables<PP>), dim3((G + 255) / 256), dim3(256), 0, s, patches, gt, tab, bx.g[0], \
                           bx.g[1], bx.g[2]);                                                                   \
        hipLaunchKernelGGL((k_load3d<PP>), dim3((unsigned)nelem), dim3(128), 0, s, patches, gt, Fbuf, detmin,   \
                           sf_mfma(PP) ? Dq : nullptr, bx.g[1], bx.g[2]);                                       \
        if (sf_mfma(PP)) {                                                                                      \
            bx.nty = (bx.nb[1] + SfmCfg<(PP > 1 ? PP : 2)>::TY - 1) / SfmCfg<(PP > 1 ? PP : 2)>::TY;            \
            bx.ntz = (bx.nb[2] + SfmCfg<(PP > 1 ? PP : 2)>::TZ - 1) / SfmCfg<(PP > 1 ? PP : 2)>::TZ;            \
            hipLaunchKernelGGL((k_sfasm3d_mfma<(PP > 1 ? PP : 2)>), dim3(bx.nty * bx.ntz), dim3(512), 0, s, tab, \
                               Dq, bx, s64);                                                                    \
        } else {                                                                                                \
            bx.nty = (bx.nb[1] + 1) / 2;                                                                        \
            bx.ntz = (bx.nb[2] + 1) / 2;                                                                        \
            hipLaunchKernelGGL((k_sfasm3d<PP>), dim3(bx.nty * bx.ntz), dim3(256), 0, s, patches, tab, bx, s64); \
        }                                                                                                       \
        return hipGetLastError();                                                                               \
    }
    GSX_SF(1) GSX_SF(2) GSX_SF(3) GSX_SF(4)
#undef GSX_SF
    return hipErrorInvalidValue;
}

hipError_t launch_sf_finish(int p, const SfBox &bx, int nfree, const int *rowptr, double *s64, float *s32,
                            double *val, double *rhs, const double *Fbuf, const int *dofs, const double *fixed,
                            int *err, hipStream_t s)
{
    if (nfree <= 0) return hipSuccess;
    const int64_t nelem = (int64_t)bx.nel[0] * bx.nel[1] * bx.nel[2];
#define GSX_SFF(PP)                                                                                             \
    if (p == PP) {                                                                                              \
        const dim3 grid((nfree + SfFinish<PP>::R - 1) / SfFinish<PP>::R), blk(256);                             \
        hipLaunchKernelGGL((k_sf_finish3d<PP>), grid, blk, 0, s, bx, nfree, rowptr, s64, s32, val, rhs, Fbuf,   \
                           nelem, dofs, fixed, err);                                                            \
        return hipGetLastError();                                                                               \
    }
    GSX_SFF(1) GSX_SFF(2) GSX_SFF(3) GSX_SFF(4)
#undef GSX_SFF
    return hipErrorInvalidValue;
}

}  // namespace gsx
