"""Synthetic isogeometric Poisson problems (inputs of the pipeline).

Every random quantity comes from a counter-based generator (SplitMix64
finaliser keyed by (seed, stream, index)), so a problem is a pure function of
its parameters and any entity (control point, element) can be regenerated
independently.

A *patch* is a tensor-product B-spline patch whose geometry uses the same
basis as the discretisation (isoparametric).  Control points are stored
row-major (nbasis x d) in lexicographic order, direction 0 fastest, the
numbering of gsTensorBasis (src/gsTensor/gsTensorBasis.hpp:203-236).

Sides are numbered like G+Smo's boxSide (src/gsCore/gsBoundary.h): 1 west
(u=0), 2 east (u=1), 3 south (v=0), 4 north (v=1), 5 front (w=0), 6 back
(w=1).  An interface record is (p1, s1, p2, s2, m0, f0, m1, f1): for the j-th
tangential direction of patch p1 (increasing order) mj is the matching
direction of patch p2 and fj = 1 if its orientation is reversed.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

_M64 = (1 << 64) - 1
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, stream, idx):
    """U[0,1) doubles for integer counters idx (array)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _mix(np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & _M64))
        v = _mix(key + (idx + np.uint64(1)) * _GOLDEN)
    return (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal01(seed, stream, idx):
    """Standard normals by Box-Muller over two uniform streams."""
    u1 = 1.0 - uniform01(seed, 2 * stream + 1000, idx)
    u2 = uniform01(seed, 2 * stream + 1001, idx)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def knots_uniform(nelem, p, a=0.0, b=1.0):
    """Open uniform knots, interior knot i = a + i*h, h = (b-a)/nelem
    (gsKnotVector::initUniform, src/gsNurbs/gsKnotVector.hpp:643-688)."""
    h = (b - a) / float(nelem)
    inner = [a + float(i) * h for i in range(1, nelem)]
    return np.array([a] * (p + 1) + inner + [b] * (p + 1), dtype=np.float64)


def greville(knots, p):
    n = len(knots) - p - 1
    return np.array([np.sum(knots[i + 1:i + p + 1]) / p for i in range(n)])


@dataclass
class Patch:
    d: int
    p: List[int]
    knots: List[np.ndarray]
    ctrl: np.ndarray                 # (nbasis, d)
    kappa: Optional[np.ndarray] = None   # (nelem,) or None
    felem: Optional[np.ndarray] = None   # (nelem,) when fkind == 0
    fkind: int = 1                   # 1: d*pi^2*prod sin(pi x_r); 0: felem

    def nbasis_dir(self):
        return [len(self.knots[k]) - self.p[k] - 1 for k in range(self.d)]

    def nelem_dir(self):
        return [len(np.unique(self.knots[k])) - 1 for k in range(self.d)]

    @property
    def nbasis(self):
        return int(np.prod(self.nbasis_dir()))

    @property
    def nelem(self):
        return int(np.prod(self.nelem_dir()))


@dataclass
class Problem:
    patches: List[Patch]
    dir_sides: List[tuple] = field(default_factory=list)
    ifaces: List[tuple] = field(default_factory=list)
    name: str = ""

    @property
    def d(self):
        return self.patches[0].d

    @property
    def nbasis_total(self):
        return sum(P.nbasis for P in self.patches)


def box_patch(d, p, n, lo=None, hi=None):
    """Affine box [lo,hi] as a degree-p patch with n elements per direction:
    control points at the Greville abscissae (exact reproduction of the
    identity map)."""
    lo = [0.0] * d if lo is None else list(lo)
    hi = [1.0] * d if hi is None else list(hi)
    pp = [p] * d if np.isscalar(p) else list(p)
    nn = [n] * d if np.isscalar(n) else list(n)
    knots = [knots_uniform(nn[k], pp[k]) for k in range(d)]
    gr = [greville(knots[k], pp[k]) for k in range(d)]
    grids = np.meshgrid(*[lo[k] + (hi[k] - lo[k]) * gr[k] for k in range(d)], indexing="ij")
    # lexicographic, direction 0 fastest
    ctrl = np.stack([g.transpose(*reversed(range(d))).reshape(-1) for g in grids], axis=1)
    return Patch(d=d, p=pp, knots=knots, ctrl=np.ascontiguousarray(ctrl))


def all_sides(d, patch=0):
    return [(patch, s) for s in range(1, 2 * d + 1)]


def perturb_interior(P, amp, seed, stream0=1):
    """Add U(-amp, amp) to every coordinate of the interior control points."""
    nb = P.nbasis_dir()
    idx = np.arange(P.nbasis)
    coords = []
    r = idx.copy()
    for k in range(P.d):
        coords.append(r % nb[k])
        r //= nb[k]
    interior = np.ones(P.nbasis, dtype=bool)
    for k in range(P.d):
        interior &= (coords[k] > 0) & (coords[k] < nb[k] - 1)
    ctrl = P.ctrl.copy()
    for c in range(P.d):
        u = uniform01(seed, stream0 + c, idx)
        ctrl[:, c] += np.where(interior, (2.0 * u - 1.0) * amp, 0.0)
    P.ctrl = np.ascontiguousarray(ctrl)
    return P


def lognormal_kappa(nelem, sigma, seed, stream=10):
    return np.exp(sigma * normal01(seed, stream, np.arange(nelem)))


def zero_kappa_box(P, lo, hi):
    """Set kappa = 0 on the elements whose multi-index lies in [lo, hi]
    (inclusive, per direction): a void in the coefficient.  Their element
    matrices are exact zeros, so gsExprAssembler's exact-zero skip
    (src/gsAssembler/gsExprAssembler.h:599) leaves out every entry that only
    they contribute to."""
    ne = P.nelem_dir()
    if P.kappa is None:
        P.kappa = np.ones(P.nelem)
    idx = np.arange(P.nelem)
    inside = np.ones(P.nelem, dtype=bool)
    r = idx.copy()
    for k in range(P.d):
        c = r % ne[k]
        r //= ne[k]
        inside &= (c >= lo[k]) & (c <= hi[k])
    P.kappa = np.where(inside, 0.0, P.kappa)
    return P


def noisy_box_fast(d, p, n, noise, sigma, seed):
    """Degree-p patch on the unit box with n elements per direction, interior
    control points moved by U(+-noise/n), lognormal(0, sigma) kappa and
    U(-1, 1) sources per element, drawn from numpy's PCG64 stream `seed`
    (the serving workload: fast enough to generate requests at their
    arrival rate)."""
    rng = np.random.default_rng(seed)
    kn = knots_uniform(n, p)
    g = greville(kn, p)
    nb = g.size
    # C-ordered (z, y, x, coordinate): direction 0 is the fastest index
    ctrl = np.empty((nb,) * d + (d,))
    for k in range(d):
        shape = [1] * d
        shape[d - 1 - k] = nb
        ctrl[..., k] = g.reshape(shape)
    inner = (slice(1, nb - 1),) * d
    ctrl[inner] += rng.uniform(-noise / n, noise / n, (nb - 2,) * d + (d,))
    P = Patch(d=d, p=[p] * d, knots=[kn] * d, ctrl=ctrl.reshape(-1, d))
    P.kappa = np.exp(sigma * rng.standard_normal(P.nelem))
    P.felem = rng.uniform(-1.0, 1.0, P.nelem)
    P.fkind = 0
    return P


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def poisson_square(n=64, p=2):
    """Unit square, open uniform knots, f = 2 pi^2 sin sin, homogeneous
    Dirichlet on all sides (64x64, p=2: 4356 basis functions, 4096 free)."""
    P = box_patch(2, p, n)
    P.fkind = 1
    return Problem([P], all_sides(2), [], name=f"square-p{p}-n{n}")


def poisson_noisy_cube(n=64, p=3, seed=20261202, noise=0.15, sigma=0.5, with_kappa=True):
    """Unit cube, degree p, n^3 elements, interior control points perturbed
    by U(+-noise*h), per-element kappa = exp(N(0, sigma^2)), f = 3 pi^2 sin
    sin sin, homogeneous Dirichlet on all sides."""
    P = box_patch(3, p, n)
    perturb_interior(P, noise / n, seed)
    if with_kappa:
        P.kappa = lognormal_kappa(P.nelem, sigma, seed)
    P.fkind = 1
    return Problem([P], all_sides(3), [], name=f"cube-p{p}-n{n}")


def square_grid(nx=4, ny=4, n=16, p=3, r=1.0):
    """nx x ny unit squares (patch index i*ny + j at [i,i+1] x [j,j+1], the
    numbering of gsNurbsCreator::BSplineSquareGrid,
    src/gsNurbs/gsNurbsCreator.hpp:691-705), C0-glued interfaces,
    homogeneous Dirichlet on the outer boundary, f = 2 pi^2 sin sin."""
    patches = []
    for i in range(nx):
        for j in range(ny):
            P = box_patch(2, p, n, lo=[r * i, r * j], hi=[r * (i + 1), r * (j + 1)])
            P.fkind = 1
            patches.append(P)
    dir_sides, ifaces = [], []
    for i in range(nx):
        for j in range(ny):
            k = i * ny + j
            if i == 0:
                dir_sides.append((k, 1))
            if i == nx - 1:
                dir_sides.append((k, 2))
            if j == 0:
                dir_sides.append((k, 3))
            if j == ny - 1:
                dir_sides.append((k, 4))
            if i + 1 < nx:   # east of k = west of k+ny; tangential dir 1 -> 1
                ifaces.append((k, 2, k + ny, 1, 1, 0, 0, 0))
            if j + 1 < ny:   # north of k = south of k+1; tangential dir 0 -> 0
                ifaces.append((k, 4, k + 1, 3, 0, 0, 0, 0))
    return Problem(patches, dir_sides, ifaces, name=f"grid{nx}x{ny}-p{p}-n{n}")


def two_cubes(n=4, p=2):
    """[0,2] x [0,1]^2 as two glued unit cubes, n elements per direction
    each (patch 0's east side on patch 1's west side, tangential directions
    1 and 2 unchanged), homogeneous Dirichlet on the outer boundary,
    f = 3 pi^2 sin sin sin."""
    a = box_patch(3, p, n)
    b = box_patch(3, p, n, lo=[1.0, 0.0, 0.0], hi=[2.0, 1.0, 1.0])
    for P in (a, b):
        P.fkind = 1
    dir_sides = [(0, s) for s in (1, 3, 4, 5, 6)] + [(1, s) for s in (2, 3, 4, 5, 6)]
    return Problem([a, b], dir_sides, [(0, 2, 1, 1, 1, 0, 2, 0)], name=f"two-cubes-p{p}-n{n}")


def batch_problems(count=512, seed=7, n2d=(8, 96), n3d=(8, 32), degrees=(2, 3, 4),
                   noise=0.1, sigma=0.5):
    """Independent single-patch problems: the first half 2D, the rest 3D,
    elements per direction uniform in the ranges, degree uniform in degrees,
    interior control-point noise U(+-noise*h), lognormal kappa, per-element
    random source f_e ~ U(-1,1)."""
    out = []
    for i in range(count):
        s = seed * 1000003 + i
        d = 2 if i < count // 2 else 3
        lo, hi = n2d if d == 2 else n3d
        n = int(lo + np.floor(uniform01(s, 50, [0])[0] * (hi - lo + 1)))
        p = int(degrees[int(np.floor(uniform01(s, 51, [0])[0] * len(degrees)))])
        P = box_patch(d, p, n)
        perturb_interior(P, noise / n, s)
        P.kappa = lognormal_kappa(P.nelem, sigma, s)
        P.felem = 2.0 * uniform01(s, 52, np.arange(P.nelem)) - 1.0
        P.fkind = 0
        out.append(Problem([P], all_sides(d), [], name=f"batch{i}-d{d}-p{p}-n{n}"))
    return out
