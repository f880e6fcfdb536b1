// Shared device-side types of libgsx (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gsx {

constexpr int kMaxDeg = 4;         // degrees 1..4 on the device
constexpr int kMaxQ1 = kMaxDeg + 1; // Gauss nodes per direction (quA=quB=1)
constexpr int kWave = 64;

// One patch as the kernels see it.  Elements are lexicographic (direction 0
// fastest); element e_k along direction k covers knots [p+e_k, p+e_k+1] and
// its active functions are e_k..e_k+p (open knots, no repeated interior
// knots -- checked on the host).
struct PatchDev {
    int d;
    int p[3];
    int nb[3];          // basis functions per direction
    int nel[3];         // elements per direction
    int fkind;          // 1: d*pi^2*prod sin(pi x); 0: felem
    int orient;         // +1 / -1: sign of det J of a valid parametrization
    const double *knots[3];
    const double *ctrl;     // patch control points (nb x d)
    const double *kappa;    // per element or nullptr
    const double *felem;    // per element or nullptr
    const int *dofs;        // global DOF index per local basis function
    int64_t nelem;
    int64_t kmat_off;       // element matrices: doubles offset of element 0
    int64_t fvec_off;       // element load vectors: doubles offset
    int64_t elem_off;       // global element id of element 0
};

// Gauss-Legendre nodes/weights for n = 1..kMaxQ1 points, ascending.
struct GaussTab {
    double x[kMaxQ1 + 1][kMaxQ1];
    double w[kMaxQ1 + 1][kMaxQ1];
};

// Structured operator layout: rows in tiles of kStencilTile; a tile's
// coefficients are stored slot by slot ([tile][slot][row in tile]), so the
// block that applies a tile streams one contiguous region.  The last tile is
// padded to full size (padding rows are never read).
constexpr int kStencilTile = 256;
__host__ __device__ inline int64_t stencil_index(int slot, int r, int ns)
{
    return ((int64_t)(r / kStencilTile) * ns + slot) * kStencilTile + (r % kStencilTile);
}
inline int64_t stencil_size(int n, int ns)
{
    return (int64_t)((n + kStencilTile - 1) / kStencilTile) * ns * kStencilTile;
}
// The FP32 copy of the (bitwise symmetric) structured operator keeps only
// the slots of offsets lexicographically >= 0: slot s >= ns / 2 is stored as
// s - ns / 2 in a layout of stencil_half(ns) slots; a lower slot s of row r
// is slot ns - 1 - s of the row at the opposite offset.
__host__ __device__ constexpr int stencil_half(int ns) { return ns - ns / 2; }

// Mixed-precision CG: a reliable update replaces the FP32 recurrence's
// residual r_rec by the true FP64 residual r and keeps the search
// direction, which is consistent as long as r_rec and r agree (they do up
// to FP32 rounding while the recurrence is above its attainable accuracy).
// Once the recurrence has run into its floor they part: the fused update
// r - alpha t leaves a rounding residue (a few percent of r on tiny systems
// solved in one step) and the kept direction then makes the iteration
// diverge.  When (r_rec, r_rec) < kDirectionReset (r, r) the direction
// restarts from the preconditioned residual instead.
constexpr double kDirectionReset = 0.25;

// Device-resident scalars of one CG solve.
struct CGState {
    double rz;        // (r, z) of the current iterate
    double pap;       // (p, A p)
    double alpha;
    double beta;
    double rr;        // ||r||^2
    double err;       // ||r|| / ||b||
    double bnorm;
    double tol;
    double tau;       // Jacobi damping
    double rz_new;
    double rel_since_update; // mixed: ||r32||^2 relative tracking
    double rr_ref;    // mixed: ||r||^2 at the last reliable update
    double rzs[2];    // (r, z) by iteration parity (grid-layout solver)
    int it;           // completed iterations
    int done;         // 1 once err < tol (later launches exit at once)
    int maxit;
    int need_update;  // mixed: a reliable update is due
    // mixed: the FP32 recurrence stops once an update is due, until the
    // update has run (it would otherwise iterate on a residual at rounding
    // level, where (r, z) can vanish).  Grid solver: iterations k >=
    // flag_next are skipped (0 = none).  CSR solver: frozen, set by the
    // alpha step when an update is due.
    int flag_next;
    int frozen;
};

// One boundary element: the element of `patch` adjacent to side `side`
// (1..2d) at element index ei, the condition's type (1 Dirichlet, 2
// Neumann), data kind (0 zero, 1 manufactured, 2 affine) and values.
struct BdryItem {
    int patch, side, type, kind;
    int ei[3];
    double value[4];
};

// Element range of one element-kernel launch: elements [e0, e0 + grid) of
// the launch's patch group; element matrices and load vectors are written
// shifted down by kshift / fshift doubles (a slab of a larger patch).
struct ElemRange {
    int64_t e0;
    int64_t kshift;
    int64_t fshift;
};

// Workgroups are dispatched to the 8 XCDs round-robin (block b to XCD b % 8),
// each XCD with its own L2.  Logical block of physical block b: XCD x takes
// the contiguous range [start(x), start(x) + count(x)) of logical blocks, so
// neighbouring rows, which read the same x planes, share an L2.  A bijection
// of [0, nb).
__device__ inline int xcd_block(int b, int nb)
{
    const int x = b & 7, j = b >> 3;
    int start = 0;
    for (int y = 0; y < x; ++y) start += (nb - y + 7) >> 3;
    return start + j;
}

// One member of a batched CG solve: rows [row0, row0 + nrows) of a CSR
// matrix whose columns stay inside that range.  Vector pointers are bases
// indexed by row/column number (the member's own arrays shifted by -row0).
struct BSeg {
    const int *rp;
    const int *ci;
    const double *v64;
    const float *v32;
    const double *b;
    double *x, *r, *z, *p, *t;
    float *r32, *z32, *p32, *t32, *xc32;
    const double *d64;   // Jacobi diagonal (null: no preconditioner)
    const float *d32;
    double *hist;        // relative residual per iteration (maxit + 1) or null
    double tau;
    double tol;
    int maxit;
    int lpr;             // lanes per row of the SpMV
    int c0, nc;          // first chunk and chunk count
    // Structured member (bd > 0): the rows are the free functions of one
    // tensor-product patch, a bm[0] x bm[1] x bm[2] box in lexicographic
    // order starting at row0, coupled through the (2 bp + 1)^bd stencil.
    // Coefficients in stencil layout over the member's rows: s64 every
    // slot, s32 the upper half (zero where the neighbour is outside the box).
    int bd, bp, row0;
    int bm[3];
    int tpr;             // threads per row (1, 2, 4); chunks of 256 / tpr rows
    int half;            // FP32 operator: 1 upper half of the stencil, 0 all slots
    const double *s64;
    const float *s32;
};

// Members that iterate on one structured operator with one tile shape (the
// right-hand sides of one problem) share their SpMV launches: a group chunk
// covers rows [lo, hi) of every member of the group, which is chunk q of
// each of them.  Other members form groups of one.
constexpr int kGroupMax = 8;
struct BGroup {
    int k;
    int seg[kGroupMax];
    // groups of several box members: their SpMV inputs interleaved,
    // kGroupMax values per row (row - xrow0), refreshed before each SpMV
    void *xi;
    int xrow0;
};
struct BSChunk {
    int grp;
    int lo, hi;
    int q;
};

struct BChunk {
    int seg;
    int lo, hi;          // rows [lo, hi)
};

struct BState {
    double rzs[2];       // (r, z) of iteration k in slot k & 1
    double bnorm, rr, rr_ref, err;
    int it, it_next;     // completed iterations (it_next published by the direction step)
    int phase;           // 0: to initialise, 1: iterating
    int done;            // 1 converged, 2 iteration limit, 3 zero rhs
    int need_update;     // mixed: reliable update due
    int stop;            // mixed: iteration limit reached, decide at the next update
    int pending;         // mixed: decision of the running update
    int flag_next;       // mixed: iterations k >= flag_next wait for the update (0 = none)
};

}  // namespace gsx
