// Geometric multigrid kernels of libgsx (gfx950): transfer matrices,
// Galerkin coarse operators, dense coarse inverse, V-cycle vector steps.
//
// Reference path (G+Smo source tree):
//   gsMultiGridOp::init (Galerkin A_c = P^T A P)   src/gsMultiGrid/gsMultiGrid.hpp:70-94
//   gsMultiGridOp::multiGridStep (V/W cycle)        src/gsMultiGrid/gsMultiGrid.hpp:148-194
//   gsMultiBasis::combineTransferMatrices           src/gsCore/gsMultiBasis.hpp:137-178
//   gsTensorBasis::uniformRefine_withTransfer       src/gsTensor/gsTensorBasis.hpp:861-880
//   coarse direct solve (SparseLU / Cholesky)       src/gsSolver/gsMatrixOp.h:306-370
#include "gsx_common.h"
#include "gsx_kernels.h"

namespace gsx {

// 1D transfer of one (patch, direction): CSR rows = fine 1D index.
struct T1D {
    int rp_off;     // offset of this matrix's row pointer array (nf + 1 ints)
    int ent_off;    // offset of its first entry
    int nf, nc;
};

// ---------------------------------------------------------------------------
// Transfer P (fine free x coarse free).  Row f takes the tensor products of
// the 1D transfer rows of its first occurrence (patch, local index);
// interface functions get the same entries from every patch they belong to
// (conforming interfaces), which is what take_first keeps.  Columns sorted.
// ---------------------------------------------------------------------------
constexpr int kMaxPRow = 128;

template <bool WRITE>
__device__ int transfer_row(int f, const int *__restrict__ occ_ptr, const int *__restrict__ occ_patch,
                            const int64_t *__restrict__ occ_local, const int *__restrict__ fine_nb,
                            const T1D *__restrict__ t1d, const int *__restrict__ trp,
                            const int *__restrict__ tcol, const double *__restrict__ tval,
                            const int *__restrict__ cdofs_off, const int *__restrict__ cdofs,
                            const int *__restrict__ coarse_nb, int d, int nfree_c,
                            int *__restrict__ ocol, double *__restrict__ oval)
{
    const int o = occ_ptr[f];
    const int k = occ_patch[o];
    const int64_t l = occ_local[o];
    const int nbx = fine_nb[3 * k], nby = fine_nb[3 * k + 1];
    int lc[3] = {(int)(l % nbx), (int)((l / nbx) % nby), d == 3 ? (int)(l / ((int64_t)nbx * nby)) : 0};
    int b0[3], e0[3];
    for (int j = 0; j < 3; ++j) {
        if (j < d) {
            const T1D &T = t1d[3 * k + j];
            b0[j] = trp[T.rp_off + lc[j]];
            e0[j] = trp[T.rp_off + lc[j] + 1];
        } else {
            b0[j] = 0;
            e0[j] = 1;
        }
    }
    const int cnx = coarse_nb[3 * k], cny = coarse_nb[3 * k + 1];
    int cnt = 0;
    // entries go straight to the row's own output slots (sorted there by
    // insertion); no per-thread arrays, so the kernels need no scratch
    for (int iz = b0[2]; iz < e0[2]; ++iz)
        for (int iy = b0[1]; iy < e0[1]; ++iy)
            for (int ix = b0[0]; ix < e0[0]; ++ix) {
                const T1D &Tx = t1d[3 * k], &Ty = t1d[3 * k + 1];
                double v = tval[Tx.ent_off + ix];
                const int cx = tcol[Tx.ent_off + ix];
                v *= tval[Ty.ent_off + iy];
                const int cy = tcol[Ty.ent_off + iy];
                int cz = 0;
                if (d == 3) {
                    const T1D &Tz = t1d[3 * k + 2];
                    v *= tval[Tz.ent_off + iz];
                    cz = tcol[Tz.ent_off + iz];
                }
                if (v == 0.0) continue;
                const int g = cdofs[cdofs_off[k] + cx + cnx * (cy + cny * cz)];
                if (g >= nfree_c) continue;
                if (cnt >= kMaxPRow) return -1;
                if (WRITE) {   // insertion by column
                    int pos = cnt;
                    while (pos > 0 && ocol[pos - 1] > g) {
                        ocol[pos] = ocol[pos - 1];
                        oval[pos] = oval[pos - 1];
                        --pos;
                    }
                    ocol[pos] = g;
                    oval[pos] = v;
                }
                ++cnt;
            }
    return cnt;
}

__global__ void k_transfer_count(int nfree_f, const int *occ_ptr, const int *occ_patch, const int64_t *occ_local,
                                 const int *fine_nb, const T1D *t1d, const int *trp, const int *tcol,
                                 const double *tval, const int *cdofs_off, const int *cdofs,
                                 const int *coarse_nb, int d, int nfree_c, int *counts, int *err)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfree_f) return;
    const int c = transfer_row<false>(f, occ_ptr, occ_patch, occ_local, fine_nb, t1d, trp, tcol, tval,
                                      cdofs_off, cdofs, coarse_nb, d, nfree_c, nullptr, nullptr);
    if (c < 0) { err[0] = 1; counts[f] = 0; } else counts[f] = c;
}

__global__ void k_transfer_fill(int nfree_f, const int *occ_ptr, const int *occ_patch, const int64_t *occ_local,
                                const int *fine_nb, const T1D *t1d, const int *trp, const int *tcol,
                                const double *tval, const int *cdofs_off, const int *cdofs,
                                const int *coarse_nb, int d, int nfree_c, const int *rowptr, int *col,
                                double *val)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfree_f) return;
    transfer_row<true>(f, occ_ptr, occ_patch, occ_local, fine_nb, t1d, trp, tcol, tval, cdofs_off, cdofs,
                       coarse_nb, d, nfree_c, col + rowptr[f], val + rowptr[f]);
}

// ---------------------------------------------------------------------------
// CSR transpose (R = P^T): counts by integer atomics, then each row sorted
// by column so the result does not depend on the atomic order.
// ---------------------------------------------------------------------------
__global__ void k_tr_count(int n, const int *__restrict__ rp, const int *__restrict__ ci, int *__restrict__ counts)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) atomicAdd(&counts[ci[k]], 1);
}

__global__ void k_tr_fill(int n, const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ v,
                          const int *__restrict__ trp, int *__restrict__ cursor, int *__restrict__ tci,
                          double *__restrict__ tv)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
        const int c = ci[k];
        const int pos = trp[c] + atomicAdd(&cursor[c], 1);
        tci[pos] = r;
        tv[pos] = v[k];
    }
}

// Sort the entries of every row by column (the transpose is filled through
// atomic cursors, so its in-row order is arbitrary).  One wavefront per row:
// the row goes to LDS and every entry's position is the count of smaller
// columns (columns are unique within a row).  Rows of kSortCap to kSortLong
// entries (restrictions of p-multigrid) are ranked the same way by a whole
// block (k_sort_rows_long); longer ones are sorted by insertion in place.
constexpr int kSortCap = 512;
constexpr int kSortLong = 4096;
__global__ __launch_bounds__(256) void k_sort_rows(int n, const int *__restrict__ rp, int *__restrict__ ci,
                                                   double *__restrict__ v)
{
    __shared__ int sc[4][kSortCap];
    __shared__ double sv[4][kSortCap];
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int r = blockIdx.x * 4 + w;
    if (r >= n) return;
    const int a = rp[r], len = rp[r + 1] - a;
    if (len > kSortCap) {
        if (len <= kSortLong || lane != 0) return;
        for (int i = a + 1; i < a + len; ++i) {
            const int c = ci[i];
            const double x = v[i];
            int j = i - 1;
            while (j >= a && ci[j] > c) {
                ci[j + 1] = ci[j];
                v[j + 1] = v[j];
                --j;
            }
            ci[j + 1] = c;
            v[j + 1] = x;
        }
        return;
    }
    for (int k = lane; k < len; k += kWave) {
        sc[w][k] = ci[a + k];
        sv[w][k] = v[a + k];
    }
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < len; k += kWave) {
        const int key = sc[w][k];
        int rank = 0;
        for (int j = 0; j < len; ++j) rank += sc[w][j] < key;
        ci[a + rank] = key;
        v[a + rank] = sv[w][k];
    }
}

__global__ __launch_bounds__(256) void k_sort_rows_long(int n, const int *__restrict__ rp, int *__restrict__ ci,
                                                        double *__restrict__ v)
{
    __shared__ int sc[kSortLong];
    __shared__ double sv[kSortLong];
    const int r = blockIdx.x;
    if (r >= n) return;
    const int a = rp[r], len = rp[r + 1] - a;
    if (len <= kSortCap || len > kSortLong) return;
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
        sc[k] = ci[a + k];
        sv[k] = v[a + k];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < len; k += blockDim.x) {
        const int c = sc[k];
        int pos = 0;
        for (int m = 0; m < len; ++m) pos += sc[m] < c;
        ci[a + pos] = c;
        v[a + pos] = sv[k];
    }
}

// ---------------------------------------------------------------------------
// Galerkin product A_c = P^T (A P).
// Stage 1: B = A P, one thread per fine row, ELL rows of width WB with
// columns sorted; products accumulated in (A column, P column) order.
// Stage 2: A_c = R B into the coarse pattern (known: the coarse level's
// element-sharing pattern), one thread per coarse row.
// ---------------------------------------------------------------------------
template <int WB>
__global__ void k_ap(int n, const int *__restrict__ arp, const int *__restrict__ aci, const double *__restrict__ av,
                     const int *__restrict__ prp, const int *__restrict__ pci, const double *__restrict__ pv,
                     int *__restrict__ bcnt, int *__restrict__ bci, double *__restrict__ bv, int *__restrict__ err)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    int cc[WB];
    double cv[WB];
    int cnt = 0;
    for (int k = arp[f]; k < arp[f + 1]; ++k) {
        const int f2 = aci[k];
        const double a = av[k];
        for (int m = prp[f2]; m < prp[f2 + 1]; ++m) {
            const int c = pci[m];
            const double x = a * pv[m];
            int lo = 0, hi = cnt;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cc[mid] < c) lo = mid + 1; else hi = mid;
            }
            if (lo < cnt && cc[lo] == c) {
                cv[lo] += x;
            } else {
                if (cnt >= WB) { err[0] = 1; bcnt[f] = 0; return; }
                for (int i = cnt; i > lo; --i) { cc[i] = cc[i - 1]; cv[i] = cv[i - 1]; }
                cc[lo] = c;
                cv[lo] = x;
                ++cnt;
            }
        }
    }
    bcnt[f] = cnt;
    for (int i = 0; i < cnt; ++i) {
        bci[(int64_t)f * WB + i] = cc[i];
        bv[(int64_t)f * WB + i] = cv[i];
    }
}

template <int WB, int MAXC>
__global__ void k_rap(int nc, const int *__restrict__ rrp, const int *__restrict__ rci, const double *__restrict__ rv,
                      const int *__restrict__ bcnt, const int *__restrict__ bci, const double *__restrict__ bv,
                      const int *__restrict__ crp, const int *__restrict__ cci, double *__restrict__ cval,
                      int *__restrict__ err)
{
    const int c1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (c1 >= nc) return;
    const int a = crp[c1], len = crp[c1 + 1] - a;
    if (len > MAXC) { err[0] = 2; return; }
    double acc[MAXC];
    for (int i = 0; i < len; ++i) acc[i] = 0.0;
    for (int k = rrp[c1]; k < rrp[c1 + 1]; ++k) {
        const int f1 = rci[k];
        const double p = rv[k];
        const int nb = bcnt[f1];
        for (int m = 0; m < nb; ++m) {
            const int c2 = bci[(int64_t)f1 * WB + m];
            int lo = 0, hi = len - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cci[a + mid] < c2) lo = mid + 1; else hi = mid;
            }
            if (cci[a + lo] != c2) { err[0] = 3; continue; }
            acc[lo] += p * bv[(int64_t)f1 * WB + m];
        }
    }
    for (int i = 0; i < len; ++i) cval[a + i] = acc[i];
}

// Pattern of A_c = R B when the coarse pattern is not known in advance
// (transfers supplied by the caller): per coarse row the sorted union of
// the columns of the B rows that R's row references.  cnt != nullptr:
// write the row lengths; otherwise write the columns at crp[c].
template <int WB, int MAXC>
__global__ void k_rap_pattern(int nc, const int *__restrict__ rrp, const int *__restrict__ rci,
                              const int *__restrict__ bcnt, const int *__restrict__ bci, int *__restrict__ cnt,
                              const int *__restrict__ crp, int *__restrict__ cci, int *__restrict__ err)
{
    const int c1 = blockIdx.x * blockDim.x + threadIdx.x;
    if (c1 >= nc) return;
    int cols[MAXC];
    int len = 0;
    for (int k = rrp[c1]; k < rrp[c1 + 1]; ++k) {
        const int f1 = rci[k];
        const int nb = bcnt[f1];
        for (int m = 0; m < nb; ++m) {
            const int c2 = bci[(int64_t)f1 * WB + m];
            int lo = 0, hi = len;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cols[mid] < c2) lo = mid + 1; else hi = mid;
            }
            if (lo < len && cols[lo] == c2) continue;
            if (len >= MAXC) {
                err[0] = 4;
                if (cnt) cnt[c1] = 0;
                return;
            }
            for (int i = len; i > lo; --i) cols[i] = cols[i - 1];
            cols[lo] = c2;
            ++len;
        }
    }
    if (cnt) {
        cnt[c1] = len;
    } else {
        const int a = crp[c1];
        for (int i = 0; i < len; ++i) cci[a + i] = cols[i];
    }
}

// ---------------------------------------------------------------------------
// Element-level Galerkin product.  Coarse element E of a dyadically coarsened
// patch gets K_E = sum over its 2^D child elements e (lexicographic) of
// T_e^T K_e T_e, where T_e = kron of the 1D blocks T[m][a] = coefficient of
// coarse function E_k + a in fine function e_k + m (the knot-insertion
// transfer restricted to the two spans).  Gathered with the coarse DOF map
// this is P^T A P restricted to the free functions (the product that
// gsMultiGridOp::init forms with the global transfer,
// src/gsMultiGrid/gsMultiGrid.hpp:70-94), without forming A P.
// One block per coarse element; the child matrix is transformed in place in
// LDS, one line of P+1 entries along one index digit per thread and pass
// (2 D passes: row digits, then column digits).
// ---------------------------------------------------------------------------
template <int D, int P>
__global__ __launch_bounds__(256) void k_coarsen(const PatchDev *__restrict__ cpatches,
                                                 const PatchDev *__restrict__ fpatches,
                                                 const int *__restrict__ gpatch,
                                                 const int64_t *__restrict__ gprefix, int ng,
                                                 const double *__restrict__ T1,
                                                 const int64_t *__restrict__ toff,
                                                 const double *__restrict__ Kf, double *__restrict__ Kc)
{
    constexpr int NB1 = P + 1;
    constexpr int NB = (D == 3) ? NB1 * NB1 * NB1 : NB1 * NB1;
    constexpr int N2 = NB * NB;
    constexpr int PER = (N2 + 255) / 256;
    constexpr int NLINE = N2 / NB1;
    __shared__ double sK[N2];
    __shared__ double sT[3][NB1][NB1];
    const int tid = threadIdx.x;
    const int64_t E = blockIdx.x;
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (gprefix[mid] <= E) lo = mid; else hi = mid - 1;
    }
    const int pk = gpatch[lo];
    const PatchDev &cp = cpatches[pk];
    const PatchDev &fp = fpatches[pk];
    const int64_t e = E - gprefix[lo];
    int Ei[3];
    Ei[0] = (int)(e % cp.nel[0]);
    Ei[1] = (int)((e / cp.nel[0]) % cp.nel[1]);
    Ei[2] = (D == 3) ? (int)(e / ((int64_t)cp.nel[0] * cp.nel[1])) : 0;
    double acc[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) acc[j] = 0.0;
    for (int child = 0; child < (1 << D); ++child) {
        int ei[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) ei[k] = 2 * Ei[k] + ((child >> k) & 1);
        const int64_t fe = ei[0] + (int64_t)fp.nel[0] * (ei[1] + (int64_t)fp.nel[1] * ei[2]);
        const double *K = Kf + fp.kmat_off + fe * N2;
        for (int i = tid; i < N2; i += 256) sK[i] = K[i];
        for (int i = tid; i < D * NB1 * NB1; i += 256) {
            const int k = i / (NB1 * NB1), q = i % (NB1 * NB1);
            sT[k][q / NB1][q % NB1] = T1[toff[3 * pk + k] + (int64_t)ei[k] * NB1 * NB1 + q];
        }
        __syncthreads();
#pragma unroll 1
        for (int pass = 0; pass < 2 * D; ++pass) {
            const int side = pass / D, k = pass % D;
            int stride = 1;
            for (int j = 0; j < k; ++j) stride *= NB1;
            for (int ln = tid; ln < NLINE; ln += 256) {
                // line ln: the other NB - 1 digits of the transformed index
                // and the full other index
                const int other = ln % NB, rest = ln / NB;   // rest in [0, NB / NB1)
                const int lowd = rest % stride, highd = rest / stride;
                const int base = lowd + highd * stride * NB1;
                double v[NB1], w[NB1];
#pragma unroll
                for (int m = 0; m < NB1; ++m) {
                    const int idx = base + m * stride;
                    v[m] = side == 0 ? sK[idx * NB + other] : sK[other * NB + idx];
                }
#pragma unroll
                for (int a = 0; a < NB1; ++a) {
                    double sacc = 0.0;
#pragma unroll
                    for (int m = 0; m < NB1; ++m) sacc += sT[k][m][a] * v[m];
                    w[a] = sacc;
                }
#pragma unroll
                for (int a = 0; a < NB1; ++a) {
                    const int idx = base + a * stride;
                    if (side == 0) sK[idx * NB + other] = w[a]; else sK[other * NB + idx] = w[a];
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = tid + j * 256;
            if (i < N2) acc[j] += sK[i];
        }
        __syncthreads();
    }
    double *Ko = Kc + cp.kmat_off + e * N2;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = tid + j * 256;
        if (i < N2) Ko[i] = acc[j];
    }
}

// ---------------------------------------------------------------------------
// Dense coarse inverse by blocked Gauss-Jordan (SPD, no pivoting), block 64,
// in its symmetric (sweep) form.
// ---------------------------------------------------------------------------
constexpr int kGJ = 64;

// dense copy of the coarsest operator, padded with the identity: fill, then
// one thread per row scatters its CSR entries
__global__ void k_dense_fill(int n, int npad, double *__restrict__ D)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)npad * npad) return;
    const int r = (int)(i / npad), c = (int)(i % npad);
    D[i] = (r == c && r >= n) ? 1.0 : 0.0;
}

__global__ void k_dense_scatter(int n, int npad, const int *__restrict__ rp, const int *__restrict__ ci,
                                const double *__restrict__ v, double *__restrict__ D)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int k = rp[r]; k < rp[r + 1]; ++k) D[(int64_t)r * npad + ci[k]] = v[k];
}

// invert the pivot block (kb) in LDS; result to Pinv (64x64, row-major).
// In-place Gauss-Jordan over pairs of pivots: step p sweeps rows/columns
// K = {p, p + 1} at once with the 2 x 2 inverse Q of M[K][K] (positive
// determinant: the block is SPD),
//   M'[K][K] = Q,  M'[K][c] = Q M[K][c],  M'[r][K] = -M[r][K] Q,
//   M'[r][c] = M[r][c] - M[r][K] Q M[K][c],
// which equals the two single sweeps in exact arithmetic and halves the
// sequential steps.  Every entry is mapped from the old values (kept in
// registers between the two barriers); 1024 threads, 4 entries each.
constexpr int kPivThreads = 1024;
__global__ __launch_bounds__(kPivThreads) void k_gj_pivot(double *__restrict__ D, int npad, int kb,
                                                          double *__restrict__ Pinv)
{
    __shared__ double M[kGJ][kGJ + 1];
    constexpr int PER = kGJ * kGJ / kPivThreads;
    const int t = threadIdx.x;
    for (int i = t; i < kGJ * kGJ; i += kPivThreads) M[i / kGJ][i % kGJ] = D[(int64_t)(kb + i / kGJ) * npad + kb + i % kGJ];
    __syncthreads();
    for (int p = 0; p < kGJ; p += 2) {
        const double a = M[p][p], b = M[p][p + 1], c = M[p + 1][p], d = M[p + 1][p + 1];
        const double id = 1.0 / (a * d - b * c);
        const double q00 = d * id, q01 = -b * id, q10 = -c * id, q11 = a * id;
        double nv[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = t + kPivThreads * j, r = i / kGJ, col = i % kGJ;
            const bool rin = (r >> 1) == (p >> 1), cin = (col >> 1) == (p >> 1);
            const double m0 = M[p][col], m1 = M[p + 1][col];
            double x;
            if (rin) {
                const double e0 = (r == p) ? q00 : q10, e1 = (r == p) ? q01 : q11;
                x = cin ? ((col == p) ? e0 : e1) : e0 * m0 + e1 * m1;
            } else {
                const double u0 = M[r][p] * q00 + M[r][p + 1] * q10;
                const double u1 = M[r][p] * q01 + M[r][p + 1] * q11;
                x = cin ? -((col == p) ? u0 : u1) : M[r][col] - (u0 * m0 + u1 * m1);
            }
            nv[j] = x;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int i = t + kPivThreads * j;
            M[i / kGJ][i % kGJ] = nv[j];
        }
        __syncthreads();
    }
    for (int i = t; i < kGJ * kGJ; i += kPivThreads) Pinv[i] = M[i / kGJ][i % kGJ];
}

// Symmetric sweep (Gauss-Jordan in the form that keeps every intermediate
// matrix symmetric).  Sweeping pivot block k of M:
//   M'[k][k] = -B^-1,  M'[i][k] = M[i][k] B^-1,  M'[k][j] = B^-1 M[k][j],
//   M'[i][j] = M[i][j] - M[i][k] B^-1 M[k][j]            (B = M[k][k]);
// after all blocks M = -A^-1.  Only the tiles I <= J are kept: a tile below
// the diagonal is read as the transpose of its mirror.

// W[J] = B^-1 M[k][J] for every column tile J != k; one block per tile
__global__ __launch_bounds__(256) void k_gj_rowpanel(const double *__restrict__ D, int npad, int kb,
                                                     const double *__restrict__ Pinv, double *__restrict__ W)
{
    __shared__ double sP[kGJ][kGJ + 1];
    __shared__ double sR[kGJ][kGJ + 1];
    const int t = threadIdx.x, j0 = blockIdx.x * kGJ;
    if (j0 == kb) return;
    for (int i = t; i < kGJ * kGJ; i += 256) {
        const int r = i / kGJ, c = i % kGJ;
        sP[r][c] = Pinv[i];
        // M[k][J][r][c]: upper tile D[k][J], or the transpose of D[J][k]
        // (read along D's rows, transposed into LDS)
        if (j0 > kb) sR[r][c] = D[(int64_t)(kb + r) * npad + j0 + c];
        else sR[c][r] = D[(int64_t)(j0 + r) * npad + kb + c];
    }
    __syncthreads();
    for (int i = t; i < kGJ * kGJ; i += 256) {
        const int r = i / kGJ, c = i % kGJ;
        double s = 0.0;
        for (int m = 0; m < kGJ; ++m) s += sP[r][m] * sR[m][c];
        W[(int64_t)r * npad + j0 + c] = s;
    }
}

// D[I][J] -= M[I][k] W[J] for the upper tiles I <= J, I, J != k.  FP64
// matrix cores: each wavefront forms a 32x32 quadrant of the 64x64 tile as
// 2x2 blocks of v_mfma_f64_16x16x4f64 over the 64-deep panel.  Operand
// layout of that instruction: A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4 (the rows of the
// four result registers interleave).
typedef double gj_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_gj_update(double *__restrict__ D, int npad, int kb,
                                                   const double *__restrict__ W)
{
    const int I = blockIdx.y * kGJ, J = blockIdx.x * kGJ;
    if (I == kb || J == kb || I > J) return;
    __shared__ double sC[kGJ][kGJ + 1];
    __shared__ double sW[kGJ][kGJ + 1];
    const int t = threadIdx.x;
    for (int i = t; i < kGJ * kGJ; i += 256) {
        const int r = i / kGJ, c = i % kGJ;
        // M[I][k]: upper tile D[I][k], or the transpose of D[k][I] (read
        // along D's rows, transposed into LDS)
        if (I < kb) sC[r][c] = D[(int64_t)(I + r) * npad + kb + c];
        else sC[c][r] = D[(int64_t)(kb + r) * npad + I + c];
        sW[r][c] = W[(int64_t)r * npad + J + c];
    }
    __syncthreads();
    const int w = t / kWave, lane = t % kWave;
    const int r0 = (w / 2) * 32, c0 = (w % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    gj_d4 acc[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[u][v] = gj_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int ks = 0; ks < kGJ / 4; ++ks) {
        const int m = 4 * ks + lk;
        const double a0 = sC[r0 + li][m], a1 = sC[r0 + 16 + li][m];
        const double b0 = sW[m][c0 + li], b1 = sW[m][c0 + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = I + r0 + 16 * u + lk + 4 * q, col = J + c0 + 16 * v + li;
                D[(int64_t)row * npad + col] -= acc[u][v][q];
            }
}

// panels of the swept block: M'[I][k] = W[I]^T (I < k), M'[k][J] = W[J]
// (J > k), M'[k][k] = -B^-1
__global__ __launch_bounds__(256) void k_gj_panels(double *__restrict__ D, int npad, int kb,
                                                   const double *__restrict__ W, const double *__restrict__ Pinv)
{
    const int I = blockIdx.x * kGJ;
    const int t = threadIdx.x;
    if (I < kb) {
        // D[I][k] = W[I]^T through LDS, rows read and written contiguously
        __shared__ double sT[kGJ][kGJ + 1];
        for (int i = t; i < kGJ * kGJ; i += 256) sT[i / kGJ][i % kGJ] = W[(int64_t)(i / kGJ) * npad + I + i % kGJ];
        __syncthreads();
        for (int i = t; i < kGJ * kGJ; i += 256) {
            const int r = i / kGJ, c = i % kGJ;
            D[(int64_t)(I + r) * npad + kb + c] = sT[c][r];
        }
        return;
    }
    for (int i = t; i < kGJ * kGJ; i += 256) {
        const int r = i / kGJ, c = i % kGJ;
        if (I == kb) D[(int64_t)(kb + r) * npad + kb + c] = -Pinv[i];
        else D[(int64_t)(kb + r) * npad + I + c] = W[(int64_t)r * npad + I + c];
    }
}

// A^-1 = -M: negate the upper tiles and mirror them into the lower ones
__global__ __launch_bounds__(256) void k_gj_finish(double *__restrict__ D, int npad)
{
    const int I = blockIdx.y * kGJ, J = blockIdx.x * kGJ;
    if (I > J) return;
    __shared__ double sT[kGJ][kGJ + 1];
    const int t = threadIdx.x;
    for (int i = t; i < kGJ * kGJ; i += 256) sT[i / kGJ][i % kGJ] = -D[(int64_t)(I + i / kGJ) * npad + J + i % kGJ];
    __syncthreads();
    for (int i = t; i < kGJ * kGJ; i += 256) {
        const int r = i / kGJ, c = i % kGJ;
        D[(int64_t)(I + r) * npad + J + c] = sT[r][c];
        if (I < J) D[(int64_t)(J + r) * npad + I + c] = sT[c][r];
    }
}

// Two-level sweep for large coarse matrices: outer pivot blocks of 256.
// Sweeping the 256-block K at once equals its four 64-sweeps one after the
// other, but touches each tile of D once instead of four times: B^-1 by the
// 64-sweep on a 256 x 256 copy, W = B^-1 M[K][.] (row panel), the rank-256
// update M[I][J] -= M[I][K] W[J], then the swept panels.
constexpr int kGJB = 256;

// the block K of D as a full symmetric 256 x 256 matrix (a tile below the
// diagonal is the transpose of its kept mirror)
__global__ void k_gjb_take_block(const double *__restrict__ D, int npad, int kb, double *__restrict__ B)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kGJB * kGJB) return;
    const int r = i / kGJB, c = i % kGJB;
    B[i] = (r / kGJ <= c / kGJ) ? D[(int64_t)(kb + r) * npad + kb + c] : D[(int64_t)(kb + c) * npad + kb + r];
}

// MODE 0 (row panel): W[64 ri + .][J + .] = sum_c Binv[64 ri + .][64 c + .] M[kb + 64 c + .][J + .]
//   (grid: column tiles x 4 row tiles ri; tiles J inside the block skipped)
// MODE 1 (update): D[I][J] -= sum_c M[I][kb + 64 c + .] W[64 c + .][J]
//   (grid: column tiles x row tiles, upper tiles outside the block)
// Each 64 x 64 output tile accumulates the four 64-deep chunks on the FP64
// matrix cores (the operand layout of k_gj_update) and is written once.
template <int MODE>
__global__ __launch_bounds__(256) void k_gjb_gemm(double *__restrict__ D, int npad, int kb,
                                                  const double *__restrict__ Binv, double *__restrict__ W)
{
    const int J = blockIdx.x * kGJ, I = blockIdx.y * kGJ;
    const bool inJ = J >= kb && J < kb + kGJB;
    if (MODE == 0) {
        if (inJ) return;
    } else {
        const bool inI = I >= kb && I < kb + kGJB;
        if (inI || inJ || I > J) return;
    }
    __shared__ double sA[kGJ][kGJ + 1];
    __shared__ double sB[kGJ][kGJ + 1];
    const int t = threadIdx.x;
    const int w = t / kWave, lane = t % kWave;
    const int r0 = (w / 2) * 32, c0 = (w % 2) * 32;
    const int li = lane % 16, lk = lane / 16;
    gj_d4 acc[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) acc[u][v] = gj_d4{0.0, 0.0, 0.0, 0.0};
    // the next 64-deep chunk is fetched into registers while the current
    // one is multiplied from LDS (16 + 16 values per thread)
    constexpr int NQ = kGJ * kGJ / 256;
    double ra[NQ], rb[NQ];
    auto fetch = [&](int kc) {
        const int kcol = kb + kc * kGJ;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int i = t + 256 * q, r = i / kGJ, c = i % kGJ;
            if (MODE == 0) {
                ra[q] = Binv[(int64_t)(I + r) * kGJB + kc * kGJ + c];
                // M[kcol + r][J + c]: the upper tile D[kcol][J], or (stored
                // transposed below) D[J][kcol]
                rb[q] = J > kb ? D[(int64_t)(kcol + r) * npad + J + c] : D[(int64_t)(J + r) * npad + kcol + c];
            } else {
                // M[I + r][kcol + c]: the upper tile D[I][kcol], or (stored
                // transposed below) D[kcol][I]
                ra[q] = I < kb ? D[(int64_t)(I + r) * npad + kcol + c] : D[(int64_t)(kcol + r) * npad + I + c];
                rb[q] = W[(int64_t)(kc * kGJ + r) * npad + J + c];
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int i = t + 256 * q, r = i / kGJ, c = i % kGJ;
            if (MODE == 0) {
                sA[r][c] = ra[q];
                if (J > kb) sB[r][c] = rb[q];
                else sB[c][r] = rb[q];
            } else {
                if (I < kb) sA[r][c] = ra[q];
                else sA[c][r] = ra[q];
                sB[r][c] = rb[q];
            }
        }
    };
    fetch(0);
    for (int kc = 0; kc < kGJB / kGJ; ++kc) {
        stash();
        __syncthreads();
        if (kc + 1 < kGJB / kGJ) fetch(kc + 1);
#pragma unroll 4
        for (int ks = 0; ks < kGJ / 4; ++ks) {
            const int m = 4 * ks + lk;
            const double a0 = sA[r0 + li][m], a1 = sA[r0 + 16 + li][m];
            const double b0 = sB[m][c0 + li], b1 = sB[m][c0 + 16 + li];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = r0 + 16 * u + lk + 4 * q, col = J + c0 + 16 * v + li;
                if (MODE == 0) W[(int64_t)(I + row) * npad + col] = acc[u][v][q];
                else D[(int64_t)(I + row) * npad + col] -= acc[u][v][q];
            }
}

// swept panels of block K: M'[I][K] = W[I]^T (I before K), M'[K][J] = W[J]
// (J after K), M'[K][K] = -B^-1 (its upper tiles); grid: column tiles x 4
// row tiles of the block
__global__ __launch_bounds__(256) void k_gjb_panels(double *__restrict__ D, int npad, int kb,
                                                    const double *__restrict__ W, const double *__restrict__ Binv)
{
    const int I = blockIdx.x * kGJ, ri = blockIdx.y, rb = kb + ri * kGJ;
    const int t = threadIdx.x;
    if (I < kb) {
        __shared__ double sT[kGJ][kGJ + 1];
        for (int i = t; i < kGJ * kGJ; i += 256)
            sT[i / kGJ][i % kGJ] = W[(int64_t)(ri * kGJ + i / kGJ) * npad + I + i % kGJ];
        __syncthreads();
        for (int i = t; i < kGJ * kGJ; i += 256) {
            const int r = i / kGJ, c = i % kGJ;
            D[(int64_t)(I + r) * npad + rb + c] = sT[c][r];
        }
        return;
    }
    if (I < kb + kGJB) {
        const int ci = (I - kb) / kGJ;
        if (ri > ci) return;
        for (int i = t; i < kGJ * kGJ; i += 256) {
            const int r = i / kGJ, c = i % kGJ;
            D[(int64_t)(rb + r) * npad + I + c] = -Binv[(int64_t)(ri * kGJ + r) * kGJB + ci * kGJ + c];
        }
        return;
    }
    for (int i = t; i < kGJ * kGJ; i += 256) {
        const int r = i / kGJ, c = i % kGJ;
        D[(int64_t)(rb + r) * npad + I + c] = W[(int64_t)(ri * kGJ + r) * npad + I + c];
    }
}

// y = M x, dense n x n (row stride npad), one wavefront per row
template <typename T>
__global__ __launch_bounds__(256) void k_dense_gemv(int n, int npad, const T *__restrict__ M,
                                                    const T *__restrict__ x, T *__restrict__ y)
{
    const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int r = blockIdx.x * 4 + w;
    if (r >= n) return;
    const T *row = M + (int64_t)r * npad;
    T acc = 0;
    for (int c = lane; c < n; c += kWave) acc += row[c] * x[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (lane == 0) y[r] = acc;
}

// y += alpha * (P x) (prolongation), one thread per fine row (rows short);
// summed in FP64, stored in the fine vector's type T
template <typename V, typename TX, typename T>
__global__ void k_prolong_add(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                              const V *__restrict__ v, const TX *__restrict__ x, T *__restrict__ y, T alpha)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) s += (double)v[k] * (double)x[ci[k]];
    y[r] = (T)((double)y[r] + (double)alpha * s);
}

// r = b - A x (CSR), one wavefront-group per row
template <int LPR, typename V, typename T>
__global__ __launch_bounds__(256) void k_residual(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const V *__restrict__ v, const T *__restrict__ x,
                                                  const T *__restrict__ b, T *__restrict__ r)
{
    constexpr int RPB = 256 / LPR;
    const int g = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    const int row = blockIdx.x * RPB + g;
    double acc = 0.0;   // FP64 sums whatever the storage type (see k_jacobi_step)
    if (row < n)
        for (int k = rp[row] + lane; k < rp[row + 1]; k += LPR) acc += (double)v[k] * (double)x[ci[k]];
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (row < n && lane == 0) r[row] = (T)((double)b[row] - acc);
}

// y = R x (restriction), kRestrictLanes lanes per coarse row (consecutive
// lanes read consecutive entries of the row); lane partials summed in FP64
// by butterfly shuffles (a fixed order), stored in the coarse vector's type TY
constexpr int kRestrictLanes = 8;
template <typename V, typename T, typename TY>
__global__ __launch_bounds__(256) void k_restrict(int n, const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const V *__restrict__ v, const T *__restrict__ x,
                                                  TY *__restrict__ y)
{
    constexpr int L = kRestrictLanes;
    const int r = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / L), lane = threadIdx.x % L;
    if (r >= n) return;   // the row's whole group leaves together
    double s = 0.0;
    for (int k = rp[r] + lane; k < rp[r + 1]; k += L) s += (double)v[k] * (double)x[ci[k]];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) y[r] = (TY)s;
}

template <typename S, typename Dd>
__global__ void k_conv(int64_t n, const S *__restrict__ a, Dd *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (Dd)a[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 g1(int64_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

hipError_t launch_transfer(int phase, int nfree_f, const int *occ_ptr, const int *occ_patch,
                           const int64_t *occ_local, const int *fine_nb, const void *t1d, const int *trp,
                           const int *tcol, const double *tval, const int *cdofs_off, const int *cdofs,
                           const int *coarse_nb, int d, int nfree_c, int *counts, int *err,
                           const int *rowptr, int *col, double *val, hipStream_t s)
{
    if (nfree_f <= 0) return hipSuccess;
    if (phase == 0)
        hipLaunchKernelGGL(k_transfer_count, g1(nfree_f, 128), dim3(128), 0, s, nfree_f, occ_ptr, occ_patch, occ_local,
                           fine_nb, (const T1D *)t1d, trp, tcol, tval, cdofs_off, cdofs, coarse_nb, d, nfree_c,
                           counts, err);
    else
        hipLaunchKernelGGL(k_transfer_fill, g1(nfree_f, 128), dim3(128), 0, s, nfree_f, occ_ptr, occ_patch, occ_local,
                           fine_nb, (const T1D *)t1d, trp, tcol, tval, cdofs_off, cdofs, coarse_nb, d, nfree_c,
                           rowptr, col, val);
    return hipGetLastError();
}

hipError_t launch_transpose(int phase, int n, const int *rp, const int *ci, const double *v, int *counts,
                            const int *trp, int *cursor, int *tci, double *tv, int nt, hipStream_t s)
{
    if (phase == 0) {
        if (n > 0) hipLaunchKernelGGL(k_tr_count, g1(n), dim3(256), 0, s, n, rp, ci, counts);
    } else if (phase == 1) {
        if (n > 0) hipLaunchKernelGGL(k_tr_fill, g1(n), dim3(256), 0, s, n, rp, ci, v, trp, cursor, tci, tv);
    } else {
        if (nt > 0) {
            hipLaunchKernelGGL(k_sort_rows, g1(nt, 4), dim3(256), 0, s, nt, trp, tci, tv);
            hipLaunchKernelGGL(k_sort_rows_long, dim3((unsigned)nt), dim3(256), 0, s, nt, trp, tci, tv);
        }
    }
    return hipGetLastError();
}

int ap_width(int d) { return d == 2 ? 128 : 512; }

hipError_t launch_ap(int d, int n, const int *arp, const int *aci, const double *av, const int *prp,
                     const int *pci, const double *pv, int *bcnt, int *bci, double *bv, int *err, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (d == 2)
        hipLaunchKernelGGL((k_ap<128>), g1(n, 128), dim3(128), 0, s, n, arp, aci, av, prp, pci, pv, bcnt, bci, bv, err);
    else
        hipLaunchKernelGGL((k_ap<512>), g1(n, 64), dim3(64), 0, s, n, arp, aci, av, prp, pci, pv, bcnt, bci, bv, err);
    return hipGetLastError();
}

hipError_t launch_rap(int d, int nc, const int *rrp, const int *rci, const double *rv, const int *bcnt,
                      const int *bci, const double *bv, const int *crp, const int *cci, double *cval, int *err,
                      hipStream_t s)
{
    if (nc <= 0) return hipSuccess;
    if (d == 2)
        hipLaunchKernelGGL((k_rap<128, 128>), g1(nc, 128), dim3(128), 0, s, nc, rrp, rci, rv, bcnt, bci, bv, crp, cci, cval, err);
    else
        hipLaunchKernelGGL((k_rap<512, 1024>), g1(nc, 64), dim3(64), 0, s, nc, rrp, rci, rv, bcnt, bci, bv, crp, cci, cval, err);
    return hipGetLastError();
}

hipError_t launch_rap_pattern(int d, int nc, const int *rrp, const int *rci, const int *bcnt, const int *bci,
                              int *cnt, const int *crp, int *cci, int *err, hipStream_t s)
{
    if (nc <= 0) return hipSuccess;
    if (d == 2)
        hipLaunchKernelGGL((k_rap_pattern<128, 128>), g1(nc, 128), dim3(128), 0, s, nc, rrp, rci, bcnt, bci, cnt, crp,
                           cci, err);
    else
        hipLaunchKernelGGL((k_rap_pattern<512, 1024>), g1(nc, 64), dim3(64), 0, s, nc, rrp, rci, bcnt, bci, cnt, crp,
                           cci, err);
    return hipGetLastError();
}

bool coarsen_supported(int d, int p) { return (d == 2 && p >= 1 && p <= 4) || (d == 3 && p >= 1 && p <= 3); }

hipError_t launch_coarsen(int d, int p, const PatchDev *cpatches, const PatchDev *fpatches, const int *gpatch,
                          const int64_t *gprefix, int ng, int64_t nelem_c, const double *T1, const int64_t *toff,
                          const double *Kf, double *Kc, hipStream_t s)
{
    if (nelem_c <= 0) return hipSuccess;
    const dim3 grid((unsigned)nelem_c), blk(256);
#define GSX_K(D, P) if (d == D && p == P) { hipLaunchKernelGGL((k_coarsen<D, P>), grid, blk, 0, s, cpatches, fpatches, gpatch, gprefix, ng, T1, toff, Kf, Kc); return hipGetLastError(); }
    GSX_K(2, 1) GSX_K(2, 2) GSX_K(2, 3) GSX_K(2, 4)
    GSX_K(3, 1) GSX_K(3, 2) GSX_K(3, 3)
#undef GSX_K
    return hipErrorInvalidValue;
}

// the 64-block symmetric sweep of the npad x npad matrix D (M = -A^-1 on
// the upper tiles)
static void gj_sweep64(double *D, int npad, double *W, double *Pinv, hipStream_t s)
{
    const int nb = npad / kGJ;
    for (int k = 0; k < nb; ++k) {
        const int kb = k * kGJ;
        hipLaunchKernelGGL(k_gj_pivot, dim3(1), dim3(kPivThreads), 0, s, D, npad, kb, Pinv);
        hipLaunchKernelGGL(k_gj_rowpanel, dim3(nb), dim3(256), 0, s, D, npad, kb, Pinv, W);
        hipLaunchKernelGGL(k_gj_update, dim3(nb, nb), dim3(256), 0, s, D, npad, kb, W);
        hipLaunchKernelGGL(k_gj_panels, dim3(nb), dim3(256), 0, s, D, npad, kb, W, Pinv);
    }
}

int dense_inverse_npad(int n)
{
    if (n <= kGJB) return n <= kGJ ? kGJ : (n + kGJ - 1) / kGJ * kGJ;
    return (n + kGJB - 1) / kGJB * kGJB;
}

int64_t dense_inverse_work_doubles(int npad)
{
    if (npad <= kGJB) return (int64_t)kGJ * npad + kGJ * kGJ;
    return (int64_t)kGJB * npad + kGJB * kGJB + kGJ * kGJB + kGJ * kGJ;
}

hipError_t launch_dense_inverse(int n, int npad, const int *rp, const int *ci, const double *v, double *D,
                                double *work, hipStream_t s)
{
    hipLaunchKernelGGL(k_dense_fill, g1((int64_t)npad * npad), dim3(256), 0, s, n, npad, D);
    if (n > 0) hipLaunchKernelGGL(k_dense_scatter, g1(n), dim3(256), 0, s, n, npad, rp, ci, v, D);
    const int nb = npad / kGJ;
    if (npad <= kGJB) {
        gj_sweep64(D, npad, work, work + (int64_t)kGJ * npad, s);
    } else {
        double *W = work, *Bb = W + (int64_t)kGJB * npad, *Wi = Bb + kGJB * kGJB, *Pi = Wi + kGJ * kGJB;
        const int nbb = kGJB / kGJ;
        for (int kb = 0; kb < npad; kb += kGJB) {
            // B^-1 of the pivot block on its own copy
            hipLaunchKernelGGL(k_gjb_take_block, g1(kGJB * kGJB), dim3(256), 0, s, D, npad, kb, Bb);
            gj_sweep64(Bb, kGJB, Wi, Pi, s);
            hipLaunchKernelGGL(k_gj_finish, dim3(nbb, nbb), dim3(256), 0, s, Bb, kGJB);
            hipLaunchKernelGGL(k_gjb_gemm<0>, dim3(nb, nbb), dim3(256), 0, s, D, npad, kb, Bb, W);
            hipLaunchKernelGGL(k_gjb_gemm<1>, dim3(nb, nb), dim3(256), 0, s, D, npad, kb, Bb, W);
            hipLaunchKernelGGL(k_gjb_panels, dim3(nb, nbb), dim3(256), 0, s, D, npad, kb, W, Bb);
        }
    }
    hipLaunchKernelGGL(k_gj_finish, dim3(nb, nb), dim3(256), 0, s, D, npad);
    return hipGetLastError();
}

hipError_t launch_dense_gemv_f64(int n, int npad, const double *M, const double *x, double *y, hipStream_t s)
{
    hipLaunchKernelGGL((k_dense_gemv<double>), g1(n, 4), dim3(256), 0, s, n, npad, M, x, y);
    return hipGetLastError();
}

hipError_t launch_prolong_add_f64(int n, const int *rp, const int *ci, const double *v, const double *x,
                                  double *y, double alpha, hipStream_t s)
{
    hipLaunchKernelGGL((k_prolong_add<double, double, double>), g1(n), dim3(256), 0, s, n, rp, ci, v, x, y, alpha);
    return hipGetLastError();
}

hipError_t launch_prolong_add_f32(int n, const int *rp, const int *ci, const float *v, const float *x,
                                  float *y, float alpha, hipStream_t s)
{
    hipLaunchKernelGGL((k_prolong_add<float, float, float>), g1(n), dim3(256), 0, s, n, rp, ci, v, x, y, alpha);
    return hipGetLastError();
}

hipError_t launch_restrict_f64(int n, const int *rp, const int *ci, const double *v, const double *x,
                               double *y, hipStream_t s)
{
    hipLaunchKernelGGL((k_restrict<double, double, double>), g1((int64_t)n * kRestrictLanes), dim3(256), 0, s, n, rp, ci, v, x, y);
    return hipGetLastError();
}

hipError_t launch_restrict_f32(int n, const int *rp, const int *ci, const float *v, const float *x, float *y,
                               hipStream_t s)
{
    hipLaunchKernelGGL((k_restrict<float, float, float>), g1((int64_t)n * kRestrictLanes), dim3(256), 0, s, n, rp, ci, v, x, y);
    return hipGetLastError();
}

hipError_t launch_prolong_add_v32(int n, const int *rp, const int *ci, const float *v, const float *x,
                                  double *y, double alpha, hipStream_t s)
{
    hipLaunchKernelGGL((k_prolong_add<float, float, double>), g1(n), dim3(256), 0, s, n, rp, ci, v, x, y, alpha);
    return hipGetLastError();
}

hipError_t launch_restrict_v32(int n, const int *rp, const int *ci, const float *v, const double *x, float *y,
                               hipStream_t s)
{
    hipLaunchKernelGGL((k_restrict<float, double, float>), g1((int64_t)n * kRestrictLanes), dim3(256), 0, s, n, rp, ci, v, x, y);
    return hipGetLastError();
}

template <typename V, typename T>
static hipError_t residual_dispatch(int lpr, int n, const int *rp, const int *ci, const V *v, const T *x,
                                    const T *b, T *r, hipStream_t s)
{
    const dim3 grid((n + 256 / lpr - 1) / (256 / lpr)), blk(256);
    switch (lpr) {
    case 64: hipLaunchKernelGGL((k_residual<64, V, T>), grid, blk, 0, s, n, rp, ci, v, x, b, r); break;
    case 32: hipLaunchKernelGGL((k_residual<32, V, T>), grid, blk, 0, s, n, rp, ci, v, x, b, r); break;
    case 16: hipLaunchKernelGGL((k_residual<16, V, T>), grid, blk, 0, s, n, rp, ci, v, x, b, r); break;
    default: hipLaunchKernelGGL((k_residual<8, V, T>), grid, blk, 0, s, n, rp, ci, v, x, b, r); break;
    }
    return hipGetLastError();
}

hipError_t launch_residual_f64(int lpr, int n, const int *rp, const int *ci, const double *v, const double *x,
                               const double *b, double *r, hipStream_t s)
{
    return residual_dispatch<double, double>(lpr, n, rp, ci, v, x, b, r, s);
}

hipError_t launch_residual_v32(int lpr, int n, const int *rp, const int *ci, const float *v, const double *x,
                               const double *b, double *r, hipStream_t s)
{
    return residual_dispatch<float, double>(lpr, n, rp, ci, v, x, b, r, s);
}

hipError_t launch_residual_f32(int lpr, int n, const int *rp, const int *ci, const float *v, const float *x,
                               const float *b, float *r, hipStream_t s)
{
    return residual_dispatch<float, float>(lpr, n, rp, ci, v, x, b, r, s);
}

hipError_t launch_conv_f64_f32(int64_t n, const double *a, float *b, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL((k_conv<double, float>), g1(n), dim3(256), 0, s, n, a, b);
    return hipGetLastError();
}

hipError_t launch_conv_f32_f64(int64_t n, const float *a, double *b, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL((k_conv<float, double>), g1(n), dim3(256), 0, s, n, a, b);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Polynomial smoothers: d = c1 d + c2 D^-1 r (d = c2 D^-1 r on the first
// step; D = 1 when diag is null), x += d.  Richardson: c1 = 0, c2 = tau,
// no diagonal; Chebyshev: the three-term recurrence coefficients.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_smooth_update(int n, T *__restrict__ x, T *__restrict__ d, const T *__restrict__ r,
                                const T *__restrict__ diag, T c1, T c2, int first)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T ri = diag ? r[i] / diag[i] : r[i];
    const T di = first ? c2 * ri : c1 * d[i] + c2 * ri;
    d[i] = di;
    x[i] += di;
}

// power iteration step: w /= diag, partial sum of w^2 per block
__global__ __launch_bounds__(256) void k_div_sq(int n, double *__restrict__ w, const double *__restrict__ diag,
                                                double *__restrict__ part)
{
    __shared__ double s[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double wi = w[i] / diag[i];
        w[i] = wi;
        v = wi * wi;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

hipError_t launch_smooth_update_f64(int n, double *x, double *d, const double *r, const double *diag, double c1,
                                    double c2, int first, hipStream_t s)
{
    hipLaunchKernelGGL((k_smooth_update<double>), dim3((n + 255) / 256), dim3(256), 0, s, n, x, d, r, diag, c1, c2,
                       first);
    return hipGetLastError();
}

hipError_t launch_smooth_update_f32(int n, float *x, float *d, const float *r, const float *diag, float c1, float c2,
                                    int first, hipStream_t s)
{
    hipLaunchKernelGGL((k_smooth_update<float>), dim3((n + 255) / 256), dim3(256), 0, s, n, x, d, r, diag, c1, c2,
                       first);
    return hipGetLastError();
}

hipError_t launch_div_sq(int n, double *w, const double *diag, double *part, hipStream_t s)
{
    hipLaunchKernelGGL(k_div_sq, dim3((n + 255) / 256), dim3(256), 0, s, n, w, diag, part);
    return hipGetLastError();
}

}  // namespace gsx

namespace gsx {

// ---------------------------------------------------------------------------
// Gauss-Seidel sweep by level sets (gsGaussSeidelOp, internal::
// gaussSeidelSweep / reverseGaussSeidelSweep, src/gsSolver/
// gsSimplePreconditioners.hpp:21-68).  A sweep visits the rows one after
// the other; rows of one level set have no entry coupling them and every
// row's earlier neighbours (later ones for the reverse sweep) lie in
// earlier sets, so updating a set at once reads exactly the values the
// sequential sweep reads.  One thread per row of the set, the row's sum in
// storage order.
// ---------------------------------------------------------------------------
template <typename V, typename X>
__global__ __launch_bounds__(256) void k_gs_rows(int cnt, const int *__restrict__ rows, const int *__restrict__ rp,
                                                 const int *__restrict__ ci, const V *__restrict__ va,
                                                 const X *__restrict__ b, X *__restrict__ x)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= cnt) return;
    const int i = rows[k];
    double diag = 0.0, sum = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        const double a = (double)va[j];
        sum += a * (double)x[c];
        if (c == i) diag = a;
    }
    x[i] = (X)((double)x[i] + ((double)b[i] - sum) / diag);
}

template <typename V, typename X>
static hipError_t launch_gs_t(int cnt, const int *rows, const int *rp, const int *ci, const V *va, const X *b, X *x,
                              hipStream_t s)
{
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_gs_rows<V, X>), dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, rows, rp, ci, va, b, x);
    return hipGetLastError();
}

hipError_t launch_gs_rows_f64(int cnt, const int *rows, const int *rp, const int *ci, const double *va,
                              const double *b, double *x, hipStream_t s)
{
    return launch_gs_t<double, double>(cnt, rows, rp, ci, va, b, x, s);
}

hipError_t launch_gs_rows_v32(int cnt, const int *rows, const int *rp, const int *ci, const float *va,
                              const double *b, double *x, hipStream_t s)
{
    return launch_gs_t<float, double>(cnt, rows, rp, ci, va, b, x, s);
}

hipError_t launch_gs_rows_f32(int cnt, const int *rows, const int *rp, const int *ci, const float *va,
                              const float *b, float *x, hipStream_t s)
{
    return launch_gs_t<float, float>(cnt, rows, rp, ci, va, b, x, s);
}

}  // namespace gsx
