// Batched preconditioned CG of libgsx (gfx950): many independent solves
// advanced together, one launch per CG stage for all of them.
//
// Reference path (G+Smo source tree): every member runs the algorithm of
//   gsConjugateGradient::initIteration/step  src/gsSolver/gsConjugateGradient.hpp:20-93
//   gsIterativeSolver::solve (stopping rule)  src/gsSolver/gsIterativeSolver.h:114-200
//   gsJacobiOp::apply                        src/gsSolver/gsSimplePreconditioners.h:178-202
// The reference has no batching: it runs the solves one after another.
//
// Layout.  A member ("segment") is one CG solve: rows [row0, row0 + nrows)
// of a CSR matrix and its own vectors (base pointers indexed by row and by
// column).  Its rows are cut into chunks of (256 / lpr) * 64 rows (CSR) or
// 256 / tpr rows (box members), one thread block per chunk, and chunks
// never straddle two members.  Blocks map to chunks through xcd_block, so
// each XCD takes a contiguous range of chunks and a member's neighbouring
// rows share an L2.  Every
// reduction is per chunk (block partial) and then per member (the member's
// chunk partials summed in chunk order by every block that needs the
// scalar), so a member's arithmetic does not depend on which other members
// share the launch: a solve gives bitwise the same iterates alone and in
// any mix.
//
// Iteration k of a member (three launches for all members):
//   spmv:       t = A p, chunk partial (p, t)
//   update:     alpha = rz_k / (p, A p); x += alpha p; r -= alpha t;
//               z = M r; chunk partials (r, r), (r, z)
//   direction:  stop test on (r, r); beta = rz_{k+1} / rz_k; p = z + beta p
// rz_k lives in slot k & 1 of the member's state, so the blocks of the
// direction kernel read rz_k while the member's first block writes rz_{k+1}.
// The iteration count is published through it_next and taken over by the
// next spmv launch, for the same reason.
//
// Mixed precision (FP32 inner iterations, FP64 reliable updates): the FP32
// recurrence never declares convergence itself.  When its residual has
// dropped by delta^2 since the last reliable update (or claims the
// tolerance), the member is flagged and its FP32 iterations pause (from
// iteration flag_next on) until the next reliable-update pass, which
// forms x64 += xc32, r64 = b - A64 x64 in FP64, decides convergence on
// ||r64|| / ||b||, and restarts the FP32 recurrence from fp32(r64) with the
// same direction (unless the recurrence had parted from the true residual:
// kDirectionReset).
#include "gsx_common.h"
#include "gsx_kernels.h"

namespace gsx {

namespace {

__device__ inline double bsum256(double v, double *s)
{
    // wavefront shuffles, then the 4 wave sums in wave order (fixed order)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    return r;
}

// sum of a member's chunk partials, in chunk order, by the whole block
__device__ inline double member_sum(const double *__restrict__ part, int c0, int nc, double *s)
{
    double v = 0.0;
    for (int k = threadIdx.x; k < nc; k += 256) v += part[c0 + k];
    return bsum256(v, s);
}

template <int LPR, typename V, typename T>
__device__ inline void chunk_spmv(const BSeg &g, int lo, int hi, const V *__restrict__ val,
                                  const T *__restrict__ x, T *__restrict__ y, double &dot)
{
    constexpr int RPB = 256 / LPR;
    const int grp = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    for (int r = lo + grp; r < hi; r += RPB) {
        const int k0 = g.rp[r], k1 = g.rp[r + 1];
        T acc = 0;
        for (int k = k0 + lane; k < k1; k += LPR) acc += (T)val[k] * x[g.ci[k]];
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
        if (lane == 0) {
            y[r] = acc;
            dot += (double)x[r] * (double)acc;
        }
    }
}

template <typename V, typename T>
__device__ inline void chunk_spmv_any(const BSeg &g, int lo, int hi, const V *val, const T *x, T *y,
                                      double &dot)
{
    switch (g.lpr) {
    case 64: chunk_spmv<64, V, T>(g, lo, hi, val, x, y, dot); break;
    case 32: chunk_spmv<32, V, T>(g, lo, hi, val, x, y, dot); break;
    case 16: chunk_spmv<16, V, T>(g, lo, hi, val, x, y, dot); break;
    default: chunk_spmv<8, V, T>(g, lo, hi, val, x, y, dot); break;
    }
}

// r = b - A x (FP64), partial (r, r)
template <int LPR>
__device__ inline void chunk_residual(const BSeg &g, int lo, int hi, double &rr)
{
    constexpr int RPB = 256 / LPR;
    const int grp = threadIdx.x / LPR, lane = threadIdx.x % LPR;
    for (int r = lo + grp; r < hi; r += RPB) {
        const int k0 = g.rp[r], k1 = g.rp[r + 1];
        double acc = 0;
        for (int k = k0 + lane; k < k1; k += LPR) acc += g.v64[k] * g.x[g.ci[k]];
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
        if (lane == 0) {
            const double v = g.b[r] - acc;
            g.r[r] = v;
            rr += v * v;
        }
    }
}

__device__ inline void chunk_residual_any(const BSeg &g, int lo, int hi, double &rr)
{
    switch (g.lpr) {
    case 64: chunk_residual<64>(g, lo, hi, rr); break;
    case 32: chunk_residual<32>(g, lo, hi, rr); break;
    case 16: chunk_residual<16>(g, lo, hi, rr); break;
    default: chunk_residual<8>(g, lo, hi, rr); break;
    }
}

// ---------------------------------------------------------------------------
// Structured members: one thread per row, a chunk is one stencil tile of the
// member (256 rows from row0).  Vectors are indexed by row; the neighbour at
// offset (dx, dy, dz) of member row l is row l + dx + m0 (dy + m1 dz) when it
// lies in the box.  Slots whose neighbour is outside the box hold 0, and the
// vector is then read at the row itself (finite), so the product is 0.
// ---------------------------------------------------------------------------
template <int D, int P>
struct BoxGeo {
    static constexpr int W = 2 * P + 1;
    static constexpr int NS = (D == 3) ? W * W * W : W * W;
    static constexpr int C = NS / 2, NH = stencil_half(NS);
};

// Tile shape (chosen per member): tpr threads per row (1, 2 or 4; a chunk
// holds 256 / tpr rows), each summing every tpr-th stencil line (x-line of
// 2p + 1 slots) of the row; the partial sums are added by butterfly
// shuffles within the row's group, so the result depends on tpr but not on
// the other members.

// sum over the row's tpr lanes (groups are aligned: xor stays inside)
__device__ inline double row_sum(double v, int tpr)
{
    for (int off = tpr >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline float row_sum(float v, int tpr)
{
    for (int off = tpr >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The x values of a group's nv vectors at a column: separate arrays (one
// member, or the generic path), or the group's interleaved copy (kGroupMax
// values per row: one coefficient's x values come in two 16-byte loads in
// FP32 instead of nv scattered ones).  Both add c * x to acc[j] in the same
// order, so a member's sums are the same bits either way.
template <typename T>
struct XSep {
    const T *const *xs;
    template <int KMAX>
    __device__ inline void fma(T c, int col, T *acc, int nv) const
    {
#pragma unroll
        for (int j = 0; j < KMAX; ++j)
            if (j < nv) acc[j] += c * xs[j][col];
    }
    __device__ inline T at(int j, int r) const { return xs[j][r]; }
};
template <typename T>
struct XIl {
    const T *X;
    int row0;
    template <int KMAX>
    __device__ inline void fma(T c, int col, T *acc, int nv) const
    {
        static_assert(KMAX == kGroupMax, "interleaved rows hold kGroupMax values");
        const T *q = X + (int64_t)(col - row0) * kGroupMax;
        T v[kGroupMax];
        if constexpr (sizeof(T) == 4) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            const f4 a = *reinterpret_cast<const f4 *>(q), b = *reinterpret_cast<const f4 *>(q + 4);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        } else {
#pragma unroll
            for (int j = 0; j < kGroupMax; ++j) v[j] = q[j];
        }
        (void)nv;   // slots >= nv hold stale values; their sums are not read
#pragma unroll
        for (int j = 0; j < KMAX; ++j) acc[j] += c * v[j];
    }
    __device__ inline T at(int j, int r) const { return X[(int64_t)(r - row0) * kGroupMax + j]; }
};

// All slots of the row (sv: the FP64 stencil or the full FP32 one), line by
// line in lexicographic (= column) order, applied to nv vectors at once
// (acc[j] for xs[j]): each coefficient is read once for all of them, and
// each vector's sum runs in the same order whatever nv is.
template <int D, int P, typename T, int KMAX, typename XA>
__device__ inline void box_full_rows(const BSeg &g, const T *__restrict__ sv, int r, int sub, int tpr,
                                     const XA &xa, int nv, T *acc)
{
    using G = BoxGeo<D, P>;
    constexpr int NLINES = (D == 3) ? G::W * G::W : G::W;
    const int m0 = g.bm[0], m1 = g.bm[1], m2 = g.bm[2];
    const int l = r - g.row0;
    const int ia = l % m0, ib = (l / m0) % m1, ic = (D == 3) ? l / (m0 * m1) : 0;
    const T *vb = sv + stencil_index(0, l, G::NS);
#pragma unroll 1
    for (int li = sub; li < NLINES; li += tpr) {
        const int dz = (D == 3) ? li / G::W - P : 0, dy = li % G::W - P;
        const bool okyz = ib + dy >= 0 && ib + dy < m1 && ic + dz >= 0 && ic + dz < m2;
        const int rb = r + m0 * (dy + m1 * dz);
#pragma unroll
        for (int dx = -P; dx <= P; ++dx) {
            const bool ok = okyz && ia + dx >= 0 && ia + dx < m0;
            const T c = vb[(li * G::W + dx + P) * kStencilTile];
            const int col = ok ? rb + dx : r;
            xa.template fma<KMAX>(c, col, acc, nv);
        }
    }
}

// FP32 upper half: the centre line (lane 0 of the row) and each upper line
// with its mirror line (the order of the single-patch kernel k_sspmv_half),
// applied to nv vectors at once
template <int D, int P, int KMAX, typename XA>
__device__ inline void box_half_rows(const BSeg &g, int r, int sub, int tpr, const XA &xa, int nv, float *acc)
{
    using G = BoxGeo<D, P>;
    constexpr int W = G::W, C = G::C, NH = G::NH;
    constexpr int NL = (D == 3) ? P + P * W : P;
    const int m0 = g.bm[0], m1 = g.bm[1], m2 = g.bm[2];
    const int l = r - g.row0;
    const int ia = l % m0, ib = (l / m0) % m1, ic = (D == 3) ? l / (m0 * m1) : 0;
    const float *sv = g.s32;
    const float *vb = sv + stencil_index(0, l, NH);
    auto add = [&](float c, int col) { xa.template fma<KMAX>(c, col, acc, nv); };
    if (sub == 0) {
#pragma unroll
        for (int dx = 0; dx <= P; ++dx) add(vb[dx * kStencilTile], ia + dx < m0 ? r + dx : r);
#pragma unroll
        for (int dx = -P; dx < 0; ++dx)
            if (ia + dx >= 0) add(sv[stencil_index(-dx, l + dx, NH)], r + dx);
    }
#pragma unroll 1
    for (int u = sub; u < NL; u += tpr) {
        const int dz = (D == 2 || u < P) ? 0 : 1 + (u - P) / W;
        const int dy = (D == 2 || u < P) ? u + 1 : (u - P) % W - P;
        const int h0 = (D == 3 ? ((dz + P) * W + (dy + P)) * W : (dy + P) * W) - C;   // (dz, dy, -P)
        const int off = m0 * (dy + m1 * dz);
        const bool upyz = ib + dy >= 0 && ib + dy < m1 && ic + dz < m2;
#pragma unroll
        for (int dx = -P; dx <= P; ++dx) {
            const bool ok = upyz && ia + dx >= 0 && ia + dx < m0;
            add(vb[(h0 + dx + P) * kStencilTile], ok ? r + off + dx : r);
        }
        const bool okyz = ib - dy >= 0 && ib - dy < m1 && ic - dz >= 0;
#pragma unroll
        for (int dx = -P; dx <= P; ++dx) {
            const bool ok = okyz && ia + dx >= 0 && ia + dx < m0;
            if (ok) add(sv[stencil_index(h0 - dx + P, l - off + dx, NH)], r - off + dx);
        }
    }
}

// FP64 residual r = b - A x of one box member, partial (r, r)
template <int D, int P>
__device__ inline void box_resid64(const BSeg &g, int lo, int hi, double &dot)
{
    const int tpr = g.tpr, r = lo + (int)threadIdx.x / tpr, sub = (int)threadIdx.x % tpr;
    if (r >= hi) return;   // the row's whole group leaves together
    const double *xs[1] = {g.x};
    double acc[1] = {0.0};
    box_full_rows<D, P, double, 1>(g, g.s64, r, sub, tpr, XSep<double>{xs}, 1, acc);
    const double a = row_sum(acc[0], tpr);
    if (sub != 0) return;
    const double v = g.b[r] - a;
    g.r[r] = v;
    dot += v * v;
}

// y_j = A x_j for the nv members of a group (FP64: all slots; FP32: the half
// or full copy), partials (x_j, y_j)
template <int D, int P, typename T, int KMAX, typename XA>
__device__ inline void box_apply_group(const BSeg &g, int lo, int hi, const XA &xa, T *const *ys, int nv,
                                       double *dot)
{
    const int tpr = g.tpr, r = lo + (int)threadIdx.x / tpr, sub = (int)threadIdx.x % tpr;
    if (r >= hi) return;
    T acc[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) acc[j] = 0;
    if constexpr (sizeof(T) == 8) box_full_rows<D, P, double, KMAX>(g, g.s64, r, sub, tpr, xa, nv, acc);
    else if (g.half) box_half_rows<D, P, KMAX>(g, r, sub, tpr, xa, nv, acc);
    else box_full_rows<D, P, float, KMAX>(g, g.s32, r, sub, tpr, xa, nv, acc);
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        if (j < nv) acc[j] = row_sum(acc[j], tpr);
    if (sub != 0) return;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        if (j < nv) {
            ys[j][r] = acc[j];
            dot[j] += (double)xa.at(j, r) * (double)acc[j];
        }
}

#define GSX_BOX_CASES(MACRO) \
    MACRO(2, 1) MACRO(2, 2) MACRO(2, 3) MACRO(2, 4) MACRO(3, 1) MACRO(3, 2) MACRO(3, 3) MACRO(3, 4)

__device__ inline void box_resid64_any(const BSeg &g, int lo, int hi, double &dot)
{
    switch (g.bd * 8 + g.bp) {
#define GSX_C(D, P) \
    case D * 8 + P: box_resid64<D, P>(g, lo, hi, dot); break;
        GSX_BOX_CASES(GSX_C)
#undef GSX_C
    default: break;
    }
}

template <typename T, int KMAX, typename XA>
__device__ inline void box_apply_group_any(const BSeg &g, int lo, int hi, const XA &xa, T *const *ys,
                                           int nv, double *dot)
{
    switch (g.bd * 8 + g.bp) {
#define GSX_C(D, P) \
    case D * 8 + P: box_apply_group<D, P, T, KMAX>(g, lo, hi, xa, ys, nv, dot); break;
        GSX_BOX_CASES(GSX_C)
#undef GSX_C
    default: break;
    }
}

// r = b - A x (FP64) of a chunk of either kind
__device__ inline void chunk_residual_member(const BSeg &g, int lo, int hi, double &rr)
{
    if (g.bd) box_resid64_any(g, lo, hi, rr);
    else chunk_residual_any(g, lo, hi, rr);
}

}  // namespace

// ---------------------------------------------------------------------------
// Initialisation of members in phase 0 (three launches):
//   I1: r = b - A x (FP64), partials (r, r) and (b, b)
//   I2: bnorm, err, done; z = M r (or the FP32 copies), p = z; partial (r, z)
//   I3: rz_0, phase 1
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_b_init1(const BChunk *__restrict__ chunks,
                                                 const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                 double *__restrict__ p0, double *__restrict__ p1)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    if (st[ch.seg].phase != 0) return;
    const BSeg g = segs[ch.seg];
    double rr = 0.0, bb = 0.0;
    chunk_residual_member(g, ch.lo, ch.hi, rr);
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) bb += g.b[i] * g.b[i];
    const double a = bsum256(rr, s), c = bsum256(bb, s);
    if (threadIdx.x == 0) {
        p0[cb] = a;
        p1[cb] = c;
    }
}

__global__ __launch_bounds__(256) void k_b_init2(const BChunk *__restrict__ chunks,
                                                 const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                 const double *__restrict__ p0, const double *__restrict__ p1,
                                                 double *__restrict__ p2, int mixed)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    BState &S = st[ch.seg];
    if (S.phase != 0) return;
    const BSeg g = segs[ch.seg];
    const double rr = member_sum(p0, g.c0, g.nc, s), bb = member_sum(p1, g.c0, g.nc, s);
    const double bnorm = sqrt(bb);
    int done = 0;
    double err = 0.0;
    if (bnorm == 0.0) done = 3;   // zero rhs: x = 0 is the solution
    else {
        err = sqrt(rr) / bnorm;
        if (err < g.tol) done = 1;
    }
    double rz = 0.0;
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) {
        if (done == 3) g.x[i] = 0.0;
        if (done) continue;
        if (mixed) {
            const float r32 = (float)g.r[i];
            const float z32 = g.d32 ? (float)g.tau * r32 / g.d32[i] : r32;
            g.r32[i] = r32;
            g.z32[i] = z32;
            g.p32[i] = z32;
            g.xc32[i] = 0.0f;
            rz += (double)r32 * (double)z32;
        } else {
            const double ri = g.r[i];
            const double zi = g.d64 ? g.tau * ri / g.d64[i] : ri;
            g.z[i] = zi;
            g.p[i] = zi;
            rz += ri * zi;
        }
    }
    const double a = bsum256(rz, s);
    if (threadIdx.x == 0) p2[cb] = a;
    if (cb == g.c0 && threadIdx.x == 0) {
        S.bnorm = bnorm;
        S.rr = rr;
        S.rr_ref = rr;
        S.err = err;
        S.done = done;
        S.it = 0;
        S.it_next = 0;
        S.need_update = 0;
        S.stop = 0;
        S.pending = 0;
        S.flag_next = 0;
        if (g.hist) g.hist[0] = err;
    }
}

__global__ __launch_bounds__(256) void k_b_init3(const BChunk *__restrict__ chunks,
                                                 const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                 const double *__restrict__ p2)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    BState &S = st[ch.seg];
    if (S.phase != 0) return;
    const BSeg g = segs[ch.seg];
    if (cb != g.c0) return;
    const double rz = member_sum(p2, g.c0, g.nc, s);
    if (threadIdx.x == 0) {
        S.rzs[0] = rz;
        S.phase = 1;
    }
}

// ---------------------------------------------------------------------------
// One CG iteration (three launches).
// ---------------------------------------------------------------------------
// KMAX = 1: groups of one member (the SpMV of most batches); KMAX =
// kGroupMax: groups of several members, which hold KMAX accumulators per
// row (a separate launch, so that the single-member launch keeps its
// register budget and occupancy).
// Interleave the SpMV inputs of each group's active members (the same set
// k_b_spmv forms: the state it reads is not written between the two
// launches): X[(r - xrow0) kGroupMax + j] = p_j[r] over the chunk's rows.
template <typename T>
__global__ __launch_bounds__(256) void k_b_interleave(const BSChunk *__restrict__ schunks,
                                                      const BGroup *__restrict__ groups,
                                                      const BSeg *__restrict__ segs, const BState *__restrict__ st)
{
    const BSChunk ch = schunks[blockIdx.x];
    const BGroup G = groups[ch.grp];
    if (!G.xi) return;
    const T *src[kGroupMax];
    int na = 0;
    for (int j = 0; j < G.k; ++j) {
        const BState &S = st[G.seg[j]];
        if (S.phase != 1 || S.done || S.stop) continue;
        if (S.flag_next && S.flag_next <= S.it_next) continue;
        const BSeg &m = segs[G.seg[j]];
        if constexpr (sizeof(T) == 4) src[na++] = m.p32;
        else src[na++] = m.p;
    }
    if (na == 0) return;
    T *X = static_cast<T *>(G.xi);
    const int n = (ch.hi - ch.lo) * kGroupMax;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int r = ch.lo + t / kGroupMax, j = t % kGroupMax;
        if (j < na) X[(int64_t)(r - G.xrow0) * kGroupMax + j] = src[j][r];
    }
}

template <typename T, int KMAX>
__global__ __launch_bounds__(256) void k_b_spmv(const BSChunk *__restrict__ schunks,
                                                const BGroup *__restrict__ groups,
                                                const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                double *__restrict__ p0)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical group chunk
    const BSChunk ch = schunks[cb];
    const BGroup G = groups[ch.grp];
    // the group's members running this iteration (the same set in every
    // thread: the state fields read here are not written by this launch)
    int act[KMAX];
    int na = 0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        if (j >= G.k) break;
        BState &S = st[G.seg[j]];
        if (S.phase != 1 || S.done || S.stop) continue;
        // the iteration number advances even for a paused member, so that
        // the update and direction launches of this iteration see it paused
        if (ch.q == 0 && threadIdx.x == 0) S.it = S.it_next;
        if (S.flag_next && S.flag_next <= S.it_next) continue;
        act[na++] = j;
    }
    if (na == 0) return;
    const BSeg g = segs[G.seg[act[0]]];
    double dot[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) dot[j] = 0.0;
    if (!g.bd) {   // a CSR member (groups of one)
        if constexpr (sizeof(T) == 4) chunk_spmv_any<float, float>(g, ch.lo, ch.hi, g.v32, g.p32, g.t32, dot[0]);
        else chunk_spmv_any<double, double>(g, ch.lo, ch.hi, g.v64, g.p, g.t, dot[0]);
    } else {
        const T *xs[KMAX];
        T *ys[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const int sid = G.seg[act[j < na ? j : 0]];
            if constexpr (sizeof(T) == 4) {
                xs[j] = segs[sid].p32;
                ys[j] = segs[sid].t32;
            } else {
                xs[j] = segs[sid].p;
                ys[j] = segs[sid].t;
            }
        }
        if constexpr (KMAX > 1) {
            // the inputs of the active members, interleaved by k_b_interleave
            if (G.xi) {
                box_apply_group_any<T, KMAX>(g, ch.lo, ch.hi, XIl<T>{static_cast<const T *>(G.xi), G.xrow0}, ys, na,
                                             dot);
            } else {
                box_apply_group_any<T, KMAX>(g, ch.lo, ch.hi, XSep<T>{xs}, ys, na, dot);
            }
        } else {
            box_apply_group_any<T, KMAX>(g, ch.lo, ch.hi, XSep<T>{xs}, ys, na, dot);
        }
    }
    for (int j = 0; j < na; ++j) {
        const double a = bsum256(dot[j], s);
        if (threadIdx.x == 0) p0[segs[G.seg[act[j]]].c0 + ch.q] = a;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_update(const BChunk *__restrict__ chunks,
                                                  const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                  const double *__restrict__ p0, double *__restrict__ p1,
                                                  double *__restrict__ p2)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    const BState &S = st[ch.seg];
    if (S.phase != 1 || S.done || S.stop || (S.flag_next && S.flag_next <= S.it)) return;
    const BSeg g = segs[ch.seg];
    const double pap = member_sum(p0, g.c0, g.nc, s);
    const double alpha = S.rzs[S.it & 1] / pap;
    double rr = 0.0, rz = 0.0;
    if constexpr (sizeof(T) == 4) {
        const float a = (float)alpha, tau = (float)g.tau;
        for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) {
            g.xc32[i] += a * g.p32[i];
            const float ri = g.r32[i] - a * g.t32[i];
            g.r32[i] = ri;
            const float zi = g.d32 ? tau * ri / g.d32[i] : ri;
            g.z32[i] = zi;
            rr += (double)ri * (double)ri;
            rz += (double)ri * (double)zi;
        }
    } else {
        for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) {
            g.x[i] += alpha * g.p[i];
            const double ri = g.r[i] - alpha * g.t[i];
            g.r[i] = ri;
            const double zi = g.d64 ? g.tau * ri / g.d64[i] : ri;
            g.z[i] = zi;
            rr += ri * ri;
            rz += ri * zi;
        }
    }
    const double a = bsum256(rr, s), c = bsum256(rz, s);
    if (threadIdx.x == 0) {
        p1[cb] = a;
        p2[cb] = c;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_direction(const BChunk *__restrict__ chunks,
                                                     const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                     const double *__restrict__ p1,
                                                     const double *__restrict__ p2, double delta2)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    BState &S = st[ch.seg];
    // flag_next is only set to it + 1 below: every block of this launch
    // takes the same decision whether or not the first block has written it
    if (S.phase != 1 || S.done || S.stop || (S.flag_next && S.flag_next <= S.it)) return;
    const BSeg g = segs[ch.seg];
    const double rr = member_sum(p1, g.c0, g.nc, s), rz = member_sum(p2, g.c0, g.nc, s);
    const int k = S.it;
    const double err = sqrt(rr) / S.bnorm;
    const bool mixed = sizeof(T) == 4;
    const bool conv = !mixed && err < g.tol;
    const bool last = k + 1 >= g.maxit;
    const double beta = rz / S.rzs[k & 1];
    if (!conv) {
        if constexpr (sizeof(T) == 4) {
            const float b = (float)beta;
            for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) g.p32[i] = g.z32[i] + b * g.p32[i];
        } else {
            for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) g.p[i] = g.z[i] + beta * g.p[i];
        }
    }
    if (cb == g.c0 && threadIdx.x == 0) {
        S.it_next = k + 1;
        S.rr = rr;
        S.err = err;
        S.rzs[(k + 1) & 1] = rz;
        if (g.hist && !mixed) g.hist[k + 1] = err;
        if (mixed && (rr < delta2 * S.rr_ref || err < g.tol)) {
            S.need_update = 1;
            S.flag_next = k + 1;
        }
        if (conv) {
            S.done = 1;
        } else if (last) {
            S.stop = 1;
            if (mixed) S.need_update = 1;
            else S.done = 2;
        }
    }
}

// ---------------------------------------------------------------------------
// Reliable update of flagged mixed members (three launches):
//   R1: x64 += xc32, xc32 = 0
//   R2: r64 = b - A64 x64, partial (r, r)
//   R3: decide; restart r32 = fp32(r64), z32 = M r32, partial (r32, z32)
//   R4: rz into the current slot, clear the flag
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_b_rel1(const BChunk *__restrict__ chunks,
                                                const BSeg *__restrict__ segs, BState *__restrict__ st)
{
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    const BState &S = st[ch.seg];
    if (S.phase != 1 || S.done || !S.need_update) return;
    const BSeg g = segs[ch.seg];
    for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) {
        g.x[i] += (double)g.xc32[i];
        g.xc32[i] = 0.0f;
    }
}

__global__ __launch_bounds__(256) void k_b_rel2(const BChunk *__restrict__ chunks,
                                                const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                double *__restrict__ p0)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    const BState &S = st[ch.seg];
    if (S.phase != 1 || S.done || !S.need_update) return;
    const BSeg g = segs[ch.seg];
    double rr = 0.0;
    chunk_residual_member(g, ch.lo, ch.hi, rr);
    const double a = bsum256(rr, s);
    if (threadIdx.x == 0) p0[cb] = a;
}

__global__ __launch_bounds__(256) void k_b_rel3(const BChunk *__restrict__ chunks,
                                                const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                const double *__restrict__ p0, double *__restrict__ p2)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    BState &S = st[ch.seg];
    if (S.phase != 1 || S.done || !S.need_update) return;
    const BSeg g = segs[ch.seg];
    const double rr = member_sum(p0, g.c0, g.nc, s);
    const double err = sqrt(rr) / S.bnorm;
    const int decided = err < g.tol ? 1 : (S.stop ? 2 : 0);
    // S.rr still holds the FP32 recurrence's residual (written by the
    // direction step; this launch leaves it to R4): a recurrence that has
    // parted from the true residual leaves no usable direction, which
    // restarts from z
    const bool reset = S.rr < kDirectionReset * rr;
    double rz = 0.0;
    if (!decided) {
        const float tau = (float)g.tau;
        for (int i = ch.lo + threadIdx.x; i < ch.hi; i += 256) {
            const float r32 = (float)g.r[i];
            const float z32 = g.d32 ? tau * r32 / g.d32[i] : r32;
            g.r32[i] = r32;
            g.z32[i] = z32;
            if (reset) g.p32[i] = z32;
            rz += (double)r32 * (double)z32;
        }
    }
    const double a = bsum256(rz, s);
    if (threadIdx.x == 0) p2[cb] = a;
    if (cb == g.c0 && threadIdx.x == 0) {
        S.rr_ref = rr;
        S.err = err;
        if (g.hist) g.hist[S.it_next] = err;
        S.pending = decided;
    }
}

__global__ __launch_bounds__(256) void k_b_rel4(const BChunk *__restrict__ chunks,
                                                const BSeg *__restrict__ segs, BState *__restrict__ st,
                                                const double *__restrict__ p2)
{
    __shared__ double s[256];
    const int cb = xcd_block(blockIdx.x, gridDim.x);   // logical chunk
    const BChunk ch = chunks[cb];
    BState &S = st[ch.seg];
    if (S.phase != 1 || S.done || !S.need_update) return;
    const BSeg g = segs[ch.seg];
    if (cb != g.c0) return;
    const double rz = member_sum(p2, g.c0, g.nc, s);
    if (threadIdx.x == 0) {
        S.rr = S.rr_ref;
        if (S.pending) S.done = S.pending;
        else S.rzs[S.it_next & 1] = rz;
        S.pending = 0;
        S.need_update = 0;
        S.flag_next = 0;
    }
}

// flag[0] = 1 when a row in [row0, row0 + nrows) has a column outside it
// (every writer stores the same value)
__global__ __launch_bounds__(256) void k_b_check(int row0, int nrows, const int *__restrict__ rp,
                                                 const int *__restrict__ ci, int *__restrict__ flag)
{
    const int r = row0 + blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
    if (r >= row0 + nrows) return;
    bool bad = false;
    for (int k = rp[r] + lane; k < rp[r + 1]; k += 64) bad |= ci[k] < row0 || ci[k] >= row0 + nrows;
    if (bad) flag[0] = 1;
}

// Structured operator of a box member from its CSR rows: every entry of
// member row l must couple to a row of the box within the stencil window
// (else flag[0] = 1); the FP64 stencil takes all slots, the FP32 one the
// upper half.  Both arrays are zero beforehand (slots without an entry).
template <int D, int P>
__global__ __launch_bounds__(256) void k_b_box_build(int row0, int n, int m0, int m1, int half, const int *__restrict__ rp,
                                                     const int *__restrict__ ci, const double *__restrict__ val,
                                                     double *__restrict__ s64, float *__restrict__ s32,
                                                     int *__restrict__ flag)
{
    using G = BoxGeo<D, P>;
    const int l = blockIdx.x * 256 + threadIdx.x;
    if (l >= n) return;
    const int ia = l % m0, ib = (l / m0) % m1, ic = l / (m0 * m1);
    bool bad = false;
    for (int k = rp[row0 + l]; k < rp[row0 + l + 1]; ++k) {
        const int j = ci[k] - row0;
        if (j < 0 || j >= n) {
            bad = true;
            continue;
        }
        const int dx = j % m0 - ia, dy = (j / m0) % m1 - ib, dz = j / (m0 * m1) - ic;
        if (dx < -P || dx > P || dy < -P || dy > P || dz < (D == 3 ? -P : 0) || dz > (D == 3 ? P : 0)) {
            bad = true;
            continue;
        }
        const int slot = ((D == 3 ? (dz + P) * G::W : 0) + (dy + P)) * G::W + (dx + P);
        s64[stencil_index(slot, l, G::NS)] = val[k];
        if (s32 && half && slot >= G::C) s32[stencil_index(slot - G::C, l, G::NH)] = (float)val[k];
        if (s32 && !half) s32[stencil_index(slot, l, G::NS)] = (float)val[k];
    }
    if (bad) flag[0] = 1;
}

// FP32 copy of a box member's FP64 stencil: the upper half (half = 1) or
// all ns slots, same tile layout
__global__ __launch_bounds__(256) void k_b_s32_from_s64(int64_t n32, int ns, int ns32, int c,
                                                        const double *__restrict__ s64, float *__restrict__ s32)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n32) return;
    const int64_t tile = i / ((int64_t)ns32 * kStencilTile);
    const int slot = (int)((i / kStencilTile) % ns32), row = (int)(i % kStencilTile);
    s32[i] = (float)s64[(tile * ns + slot + c) * kStencilTile + row];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
int batch_chunk_rows(int lpr) { return (256 / lpr) * 64; }

hipError_t launch_batch_check(int row0, int nrows, const int *rp, const int *ci, int *flag, hipStream_t s)
{
    hipLaunchKernelGGL(k_b_check, dim3((nrows + 3) / 4), dim3(256), 0, s, row0, nrows, rp, ci, flag);
    return hipGetLastError();
}

int64_t batch_box_slots(int d, int p, bool half)
{
    const int w = 2 * p + 1, ns = d == 3 ? w * w * w : w * w;
    return half ? stencil_half(ns) : ns;
}

hipError_t launch_batch_box_build(int d, int p, int row0, int n, int m0, int m1, bool half, const int *rp,
                                  const int *ci, const double *val, double *s64, float *s32, int *flag,
                                  hipStream_t s)
{
    const dim3 grid((n + 255) / 256), blk(256);
#define GSX_C(D, P)                                                                                             \
    if (d == D && p == P) {                                                                                     \
        hipLaunchKernelGGL((k_b_box_build<D, P>), grid, blk, 0, s, row0, n, m0, m1, half ? 1 : 0, rp, ci, val,  \
                           s64, s32, flag);                                                                     \
        return hipGetLastError();                                                                               \
    }
    GSX_BOX_CASES(GSX_C)
#undef GSX_C
    return hipErrorInvalidValue;
}

hipError_t launch_batch_box_s32(int d, int p, int n, bool half, const double *s64, float *s32, hipStream_t s)
{
    const int ns = (int)batch_box_slots(d, p, false), ns32 = half ? (int)batch_box_slots(d, p, true) : ns;
    const int64_t n32 = stencil_size(n, ns32);
    hipLaunchKernelGGL(k_b_s32_from_s64, dim3((unsigned)((n32 + 255) / 256)), dim3(256), 0, s, n32, ns, ns32,
                       half ? ns / 2 : 0, s64, s32);
    return hipGetLastError();
}

hipError_t launch_batch_init(int nchunks, const BChunk *chunks, const BSeg *segs, BState *st, double *p0,
                             double *p1, double *p2, int mixed, hipStream_t s)
{
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_b_init1, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p0, p1);
    hipLaunchKernelGGL(k_b_init2, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p0, p1, p2, mixed);
    hipLaunchKernelGGL(k_b_init3, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p2);
    return hipGetLastError();
}

hipError_t launch_batch_iteration(int nchunks, const BChunk *chunks, int nsingle, int nmulti,
                                  const BSChunk *schunks, const BGroup *groups, const BSeg *segs, BState *st,
                                  double *p0, double *p1, double *p2, int mixed, double delta2, hipStream_t s)
{
    if (nchunks <= 0) return hipSuccess;
    // schunks: the single-member groups' chunks, then the others'
    if (mixed) {
        if (nsingle)
            hipLaunchKernelGGL((k_b_spmv<float, 1>), dim3(nsingle), dim3(256), 0, s, schunks, groups, segs, st, p0);
        if (nmulti) {
            hipLaunchKernelGGL(k_b_interleave<float>, dim3(nmulti), dim3(256), 0, s, schunks + nsingle, groups, segs,
                               st);
            hipLaunchKernelGGL((k_b_spmv<float, kGroupMax>), dim3(nmulti), dim3(256), 0, s, schunks + nsingle,
                               groups, segs, st, p0);
        }
        hipLaunchKernelGGL(k_b_update<float>, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p0, p1, p2);
        hipLaunchKernelGGL(k_b_direction<float>, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p1, p2,
                           delta2);
    } else {
        if (nsingle)
            hipLaunchKernelGGL((k_b_spmv<double, 1>), dim3(nsingle), dim3(256), 0, s, schunks, groups, segs, st, p0);
        if (nmulti) {
            hipLaunchKernelGGL(k_b_interleave<double>, dim3(nmulti), dim3(256), 0, s, schunks + nsingle, groups, segs,
                               st);
            hipLaunchKernelGGL((k_b_spmv<double, kGroupMax>), dim3(nmulti), dim3(256), 0, s, schunks + nsingle,
                               groups, segs, st, p0);
        }
        hipLaunchKernelGGL(k_b_update<double>, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p0, p1, p2);
        hipLaunchKernelGGL(k_b_direction<double>, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p1, p2,
                           delta2);
    }
    return hipGetLastError();
}

hipError_t launch_batch_reliable(int nchunks, const BChunk *chunks, const BSeg *segs, BState *st, double *p0,
                                 double *p2, hipStream_t s)
{
    if (nchunks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_b_rel1, dim3(nchunks), dim3(256), 0, s, chunks, segs, st);
    hipLaunchKernelGGL(k_b_rel2, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p0);
    hipLaunchKernelGGL(k_b_rel3, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p0, p2);
    hipLaunchKernelGGL(k_b_rel4, dim3(nchunks), dim3(256), 0, s, chunks, segs, st, p2);
    return hipGetLastError();
}

}  // namespace gsx
