"""Pins of the CPU oracle (runs without a GPU).

Each check compares the oracle with something independent of it: the
reference's own tabulated Gauss rule and knot-span cases, exact rational
Cox-de Boor, scipy's B-spline evaluator, Kronecker closed forms of affine
patches, quadrature-level identities of isoparametric maps (row sums, energy
of the coordinate functions), numpy brute force for CG and two-grid cycles,
and the convergence rate of the discrete solution.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.interpolate import BSpline

import oracle as O
from gismo_amd import problems
from cases import zero_block_problems

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- quadrature
def test_gauss_matches_reference_tables():
    tab = json.load(open(os.path.join(GOLDEN, "gauss_legendre_reference.json")))["rules"]
    for n_str, r in tab.items():
        n = int(n_str)
        x, w = O.gauss_legendre(n)
        xr = np.array([float(s) for s in r["x"]])
        wr = np.array([float(s) for s in r["w"]])
        assert np.max(np.abs(x - xr)) <= 2.3e-16, n
        assert np.max(np.abs(w - wr)) <= 2.3e-16 * max(1.0, wr.max()), n


@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_exactness_and_numpy(n):
    x, w = O.gauss_legendre(n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, xn, atol=1e-14) and np.allclose(w, wn, atol=1e-14)
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-14
    # degree 2n is not integrated exactly
    assert abs(np.dot(w, x ** (2 * n)) - 2.0 / (2 * n + 1)) > 1e-12


# ---------------------------------------------------------------- knots/basis
def test_find_span_reference_cases():
    # gsKnotVector(0,1,3,3,1) (unittests/gsKnotVectors_test.cpp:262-278):
    # multSum of uFind = last appearance + 1
    kv = np.array([0, 0, 0, .25, .5, .75, 1, 1, 1], float)
    for u, ms in ((0.0, 3), (0.49, 4), (0.5, 5), (0.55, 5), (1.0, 6)):
        assert O.find_span(kv, 2, u) == ms - 1
    # knot vector with ghost knots and a double interior knot (:283-300)
    kv2 = np.array([-0.1, 0.0, 0.0, 0.2, 0.3, 0.3, 0.5, 0.7, 1.0, 1.1, 1.2])
    for u, idx in ((0.0, 2), (0.49, 5), (0.5, 6), (0.55, 6), (0.99, 7), (1.0, 7)):
        assert O.find_span(kv2, 2, u) == idx
    # testFindSpan (:183-211): knot-aligned and mid-element points
    for p, n in ((1, 3), (2, 5), (3, 4), (4, 6)):
        kv3 = problems.knots_uniform(n, p)
        for k in range(len(kv3) - p - 1):
            if kv3[k + 1] != kv3[k] and k >= p:
                assert O.find_span(kv3, p, kv3[k]) == k
                assert O.find_span(kv3, p, kv3[k] / 2 + kv3[k + 1] / 2) == k
        last = len(kv3) - p - 1
        assert O.find_span(kv3, p, kv3[last]) == last - 1


def _cox_de_boor(U, i, p, u):
    """Exact rational B-spline N_{i,p}(u) (right-open intervals)."""
    if p == 0:
        return Fraction(1) if U[i] <= u < U[i + 1] else Fraction(0)
    a = Fraction(0)
    if U[i + p] != U[i]:
        a = (u - U[i]) / (U[i + p] - U[i]) * _cox_de_boor(U, i, p - 1, u)
    b = Fraction(0)
    if U[i + p + 1] != U[i + 1]:
        b = (U[i + p + 1] - u) / (U[i + p + 1] - U[i + 1]) * _cox_de_boor(U, i + 1, p - 1, u)
    return a + b


def _cox_de_boor_der(U, i, p, u):
    a = Fraction(0)
    if U[i + p] != U[i]:
        a = Fraction(p) / (U[i + p] - U[i]) * _cox_de_boor(U, i, p - 1, u)
    b = Fraction(0)
    if U[i + p + 1] != U[i + 1]:
        b = Fraction(p) / (U[i + p + 1] - U[i + 1]) * _cox_de_boor(U, i + 1, p - 1, u)
    return a - b


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_bspline_ders_exact_rationals(p):
    # non-uniform open knot vector with a double interior knot
    U = [Fraction(0)] * (p + 1) + [Fraction(1, 5), Fraction(1, 2), Fraction(1, 2), Fraction(7, 10)] + \
        [Fraction(1)] * (p + 1)
    Uf = np.array([float(v) for v in U])
    rng = np.random.default_rng(p)
    pts = [Fraction(int(k), 97) for k in rng.integers(0, 97, size=12)]
    for u in pts:
        span = O.find_span(Uf, p, float(u))
        out = O.bspline_ders(Uf, p, span, float(u), 1)
        for r in range(p + 1):
            i = span - p + r
            v = _cox_de_boor(U, i, p, u)
            dv = _cox_de_boor_der(U, i, p, u)
            assert abs(out[0, r] - float(v)) < 1e-15
            assert abs(out[1, r] - float(dv)) < 1e-13 * max(1.0, abs(float(dv)))
        assert abs(out[0].sum() - 1.0) < 1e-15       # partition of unity
        assert abs(out[1].sum()) < 1e-12             # derivative of 1


# ---------------------------------------------------------------- geometry
def _tensor_eval(P, xi):
    """Independent evaluation of x(xi) and J(xi) with scipy's BSpline."""
    d = P.d
    nb = P.nbasis_dir()
    vals, ders = [], []
    for k in range(d):
        t, p = P.knots[k], P.p[k]
        B = np.zeros(nb[k])
        D = np.zeros(nb[k])
        for i in range(nb[k]):
            c = np.zeros(nb[k])
            c[i] = 1.0
            s = BSpline(t, c, p, extrapolate=False)
            B[i] = np.nan_to_num(s(xi[k]))
            D[i] = np.nan_to_num(s.derivative()(xi[k]))
        vals.append(B)
        ders.append(D)
    C = P.ctrl.reshape(tuple(reversed(nb)) + (d,))
    # contract in lexicographic order (dir 0 fastest = last numpy axis)
    x = np.zeros(d)
    J = np.zeros((d, d))
    if d == 2:
        x = np.einsum("j,i,jir->r", vals[1], vals[0], C)
        J[:, 0] = np.einsum("j,i,jir->r", vals[1], ders[0], C)
        J[:, 1] = np.einsum("j,i,jir->r", ders[1], vals[0], C)
    else:
        x = np.einsum("k,j,i,kjir->r", vals[2], vals[1], vals[0], C)
        J[:, 0] = np.einsum("k,j,i,kjir->r", vals[2], vals[1], ders[0], C)
        J[:, 1] = np.einsum("k,j,i,kjir->r", vals[2], ders[1], vals[0], C)
        J[:, 2] = np.einsum("k,j,i,kjir->r", ders[2], vals[1], vals[0], C)
    return x, J


@pytest.mark.parametrize("d,p,n", [(2, 2, 4), (3, 3, 3)])
def test_element_geometry_vs_scipy(d, p, n):
    prob = problems.poisson_noisy_cube(n=n, p=p, seed=3) if d == 3 else None
    if d == 2:
        P = problems.box_patch(2, p, n)
        problems.perturb_interior(P, 0.2 / n, 4)
    else:
        P = prob.patches[0]
    nel = P.nelem
    gx, gw = np.polynomial.legendre.leggauss(p + 1)
    for e in (0, nel // 2, nel - 1):
        xq, jac, meas, wq = O.element_geometry(P, e)
        # element corners
        ei = []
        r = e
        for k in range(d):
            ei.append(r % n)
            r //= n
        for q in range(len(wq)):
            qi = []
            rr = q
            for k in range(d):
                qi.append(rr % (p + 1))
                rr //= (p + 1)
            xi = [(ei[k] + (gx[qi[k]] + 1) / 2) / n for k in range(d)]
            x, J = _tensor_eval(P, xi)
            assert np.allclose(xq[q], x, atol=1e-14)
            assert np.allclose(jac[q], J, atol=1e-12)
            assert abs(meas[q] - abs(np.linalg.det(J))) < 1e-12
            assert abs(wq[q] - np.prod([gw[qi[k]] / (2 * n) for k in range(d)])) < 1e-16


# ---------------------------------------------------------------- assembly
def _1d_matrices(knots, p, a=0.0, b=1.0):
    """Exact (20-point Gauss per element) 1D stiffness and mass matrices."""
    nb = len(knots) - p - 1
    br = np.unique(knots)
    gx, gw = np.polynomial.legendre.leggauss(20)
    A = np.zeros((nb, nb))
    M = np.zeros((nb, nb))
    splines = []
    for i in range(nb):
        c = np.zeros(nb)
        c[i] = 1.0
        splines.append(BSpline(knots, c, p))
    for e in range(len(br) - 1):
        lo, hi = br[e], br[e + 1]
        u = (hi - lo) / 2 * (gx + 1) + lo
        w = (hi - lo) / 2 * gw
        V = np.array([s(u) for s in splines])
        D = np.array([s.derivative()(u) for s in splines])
        A += (D * w) @ D.T
        M += (V * w) @ V.T
    # map [0,1] -> [a,b]
    L = b - a
    return A / L, M * L


def _interior(nb):
    return np.arange(1, nb - 1)


@pytest.mark.parametrize("p,n", [(1, 5), (2, 8), (3, 6), (4, 5)])
def test_square_matches_kronecker(p, n):
    prob = problems.poisson_square(n, p)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    K, b = O.assemble(prob.patches, dofs, nfree, 0)
    kn = prob.patches[0].knots[0]
    A1, M1 = _1d_matrices(kn, p)
    Kk = np.kron(M1, A1) + np.kron(A1, M1)
    I = _interior(len(kn) - p - 1)
    idx = (I[None, :] + len(kn) * 0 + (len(kn) - p - 1) * I[:, None]).reshape(-1)
    Kk = Kk[np.ix_(idx, idx)]
    assert np.max(np.abs(K.toarray() - Kk)) <= 1e-13 * np.abs(Kk).max()
    # rhs: separable (p+1)-point Gauss of N_i sin(pi x)
    gx, gw = np.polynomial.legendre.leggauss(p + 1)
    br = np.unique(kn)
    g = np.zeros(len(kn) - p - 1)
    for e in range(len(br) - 1):
        lo, hi = br[e], br[e + 1]
        u = (hi - lo) / 2 * (gx + 1) + lo
        w = (hi - lo) / 2 * gw
        for i in range(len(g)):
            c = np.zeros(len(g))
            c[i] = 1
            g[i] += np.sum(w * BSpline(kn, c, p)(u) * np.sin(np.pi * u))
    bk = 2 * np.pi ** 2 * np.kron(g, g)[idx]
    assert np.max(np.abs(b - bk)) <= 1e-13 * np.abs(bk).max()


def test_cube_matches_kronecker():
    p, n = 2, 4
    prob = problems.poisson_noisy_cube(n=n, p=p, seed=1, noise=0.0, with_kappa=False)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    K, _ = O.assemble(prob.patches, dofs, nfree, 1)
    kn = prob.patches[0].knots[0]
    A1, M1 = _1d_matrices(kn, p)
    Kk = np.kron(M1, np.kron(M1, A1)) + np.kron(M1, np.kron(A1, M1)) + np.kron(A1, np.kron(M1, M1))
    nb = len(kn) - p - 1
    I = _interior(nb)
    idx = (I[None, None, :] + nb * I[None, :, None] + nb * nb * I[:, None, None]).reshape(-1)
    Kk = Kk[np.ix_(idx, idx)]
    assert np.max(np.abs(K.toarray() - Kk)) <= 1e-13 * np.abs(Kk).max()


def test_elementwise_kappa_affine():
    # K = sum_e kappa_e (My_e (x) Ax_e + Ay_e (x) Mx_e) on the unit square
    p, n = 2, 4
    P = problems.box_patch(2, p, n)
    P.kappa = np.linspace(0.5, 2.0, P.nelem)
    prob = problems.Problem([P], [], [])   # no Dirichlet: the full matrix
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    K, _ = O.assemble(prob.patches, dofs, nfree, 1)
    kn = P.knots[0]
    nb = len(kn) - p - 1
    Kk = np.zeros((nb * nb, nb * nb))
    for ey in range(n):
        for ex in range(n):
            kx = np.array(kn, float)
            # element matrices via restriction of the 1D matrices to one element
            Ax, Mx = _1d_elem(kn, p, ex)
            Ay, My = _1d_elem(kn, p, ey)
            Ke = np.kron(My, Ax) + np.kron(Ay, Mx)
            act = [(ex + i) + nb * (ey + j) for j in range(p + 1) for i in range(p + 1)]
            Kk[np.ix_(act, act)] += P.kappa[ex + n * ey] * Ke
    assert np.max(np.abs(K.toarray() - Kk)) <= 1e-13 * np.abs(Kk).max()


def _1d_elem(knots, p, e):
    br = np.unique(knots)
    lo, hi = br[e], br[e + 1]
    gx, gw = np.polynomial.legendre.leggauss(20)
    u = (hi - lo) / 2 * (gx + 1) + lo
    w = (hi - lo) / 2 * gw
    nb = len(knots) - p - 1
    V, D = [], []
    for i in range(e, e + p + 1):
        c = np.zeros(nb)
        c[i] = 1
        s = BSpline(knots, c, p)
        V.append(s(u))
        D.append(s.derivative()(u))
    V, D = np.array(V), np.array(D)
    return (D * w) @ D.T, (V * w) @ V.T


@pytest.mark.parametrize("d,p,n", [(2, 3, 5), (3, 2, 3), (3, 3, 3)])
def test_isoparametric_identities(d, p, n):
    """Unconstrained matrix of a noisy patch with kappa: K 1 = 0,
    c_x^T K c_x = sum_e kappa_e sum_q w_q |J_q|, c_x^T K c_y = 0, and the
    load vector of a per-element constant f sums to sum_e f_e |Omega_e|_h."""
    P = problems.box_patch(d, p, n)
    problems.perturb_interior(P, 0.15 / n, 77)
    P.kappa = problems.lognormal_kappa(P.nelem, 0.5, 77)
    P.felem = np.linspace(-1, 1, P.nelem)
    P.fkind = 0
    prob = problems.Problem([P], [], [])
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    assert nfree == P.nbasis and np.array_equal(dofs, np.arange(P.nbasis))
    K, b = O.assemble(prob.patches, dofs, nfree, 1)
    diag = np.abs(K.diagonal()).max()
    assert np.abs(K @ np.ones(nfree)).max() <= 1e-13 * diag
    vol = np.zeros(P.nelem)
    for e in range(P.nelem):
        _, _, meas, wq = O.element_geometry(P, e)
        vol[e] = np.sum(wq * meas)
    cx, cy = P.ctrl[:, 0], P.ctrl[:, 1]
    energy = cx @ (K @ cx)
    assert abs(energy - np.sum(P.kappa * vol)) <= 1e-12 * energy
    assert abs(cx @ (K @ cy)) <= 1e-12 * energy
    assert abs(b.sum() - np.sum(P.felem * vol)) <= 1e-12 * np.abs(P.felem * vol).sum()
    assert (K != K.T).nnz == 0


def test_expr_pattern_equals_full_pattern_on_configs():
    # no element-local entry is exactly zero on these shapes
    for prob in (problems.poisson_square(64, 2), problems.poisson_noisy_cube(6, 3),
                 problems.square_grid(2, 2, 4, 3)):
        dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        A0, _ = O.assemble(prob.patches, dofs, nfree, 0)
        A1, _ = O.assemble(prob.patches, dofs, nfree, 1)
        assert np.array_equal(A0.indptr, A1.indptr) and np.array_equal(A0.indices, A1.indices)


def _support_pattern(prob, dofs, nfree, skip_zero_kappa):
    """Closed-form sparsity pattern: every free pair (i, j) of functions
    active on one element (open uniform knots: element e of a direction has
    the functions e..e+p), over all elements, or only over those with a
    non-zero coefficient.  Independent of the oracle's assembly loop."""
    rows, cols = [], []
    off = 0
    for P in prob.patches:
        ne, nb = P.nelem_dir(), P.nbasis_dir()
        for e in range(P.nelem):
            if skip_zero_kappa and P.kappa is not None and P.kappa[e] == 0.0:
                continue
            c = []
            r = e
            for k in range(P.d):
                c.append(r % ne[k])
                r //= ne[k]
            loc = np.zeros(1, dtype=np.int64)
            stride = 1
            for k in range(P.d):
                loc = (loc[None, :] + stride * (c[k] + np.arange(P.p[k] + 1))[:, None]).reshape(-1)
                stride *= nb[k]
            g = dofs[off + loc]
            g = g[g < nfree]
            rows.append(np.repeat(g, g.size))
            cols.append(np.tile(g, g.size))
        off += P.nbasis
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    M = sp.csr_matrix((np.ones(r.size), (r, c)), shape=(nfree, nfree))
    M.sort_indices()
    return M.indptr, M.indices


@pytest.mark.parametrize("prob", zero_block_problems(), ids=lambda p: p.name)
def test_exact_zero_rule_drops_void_entries(prob):
    """Mode 0 (gsExprAssembler) keeps exactly the pairs that meet on an
    element with kappa != 0; mode 1 (gsPoissonAssembler) keeps every pair
    that meets on any element.  The entries only mode 1 has are exact zeros;
    the common entries and the load vectors are bitwise equal (skipping a
    zero addend changes no sum)."""
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    A0, b0 = O.assemble(prob.patches, dofs, nfree, 0)
    A1, b1 = O.assemble(prob.patches, dofs, nfree, 1)
    ip0, ix0 = _support_pattern(prob, dofs, nfree, True)
    ip1, ix1 = _support_pattern(prob, dofs, nfree, False)
    assert np.array_equal(A0.indptr, ip0) and np.array_equal(A0.indices, ix0)
    assert np.array_equal(A1.indptr, ip1) and np.array_equal(A1.indices, ix1)
    assert A0.nnz < A1.nnz
    assert np.array_equal(b0, b1)
    D1 = A1.toarray()
    mask = sp.csr_matrix((np.ones(A0.nnz), A0.indices, A0.indptr), shape=A0.shape).toarray() > 0
    assert np.all(D1[~mask] == 0.0)
    assert np.array_equal(A0.toarray()[mask], D1[mask])
    assert np.all(np.diff(A0.indptr) > 0)
    np.linalg.cholesky(A0.toarray())   # still SPD


def test_assemble_rows_matches_full():
    prob = problems.poisson_noisy_cube(n=5, p=3, seed=8)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    A, b = O.assemble(prob.patches, dofs, nfree, 1)
    rows = np.array([0, 7, nfree // 2, nfree - 1])
    got, rhs = O.assemble_rows(prob.patches[0], dofs, nfree, rows)
    for t, r in enumerate(rows):
        c, v = got[t]
        s, e = A.indptr[r], A.indptr[r + 1]
        assert np.array_equal(c, A.indices[s:e]) and np.array_equal(v, A.data[s:e])
        assert rhs[t] == b[r]


# ---------------------------------------------------------------- mapper
def test_mapper_two_by_two_p1_by_hand():
    # 2x2 grid of unit squares, one linear element each: 16 local functions,
    # 9 glued, the centre is the only free DOF (coupled, index 0)
    prob = problems.square_grid(2, 2, 1, 1)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    assert nfree == 1
    # patch k = i*ny + j sits at [i,i+1]x[j,j+1]: the centre (1,1) is the
    # local corner (1,1), (1,0), (0,1), (0,0) of patches 0..3
    centre = [3, 1, 2, 0]
    for k in range(4):
        assert dofs[4 * k + centre[k]] == 0
    assert len(set(dofs.tolist())) == 9


def test_mapper_two_by_two_p1_n2_by_hand():
    # each patch 3x3 functions: one standard interior DOF per patch (0..3 in
    # patch order), then coupled DOFs by first appearance in the scan
    prob = problems.square_grid(2, 2, 2, 1)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    assert nfree == 9
    loc = lambda i, j: i + 3 * j
    for k in range(4):
        assert dofs[9 * k + loc(1, 1)] == k
    # patch 0 = [0,1]^2: its east edge midpoint (2,1) is met first at local 5,
    # its north edge midpoint (1,2) at local 7, the centre (2,2) at local 8
    assert dofs[loc(2, 1)] == 4 and dofs[loc(1, 2)] == 5 and dofs[loc(2, 2)] == 6


@pytest.mark.parametrize("nx,ny,n,p", [(4, 4, 16, 3), (4, 4, 32, 3), (3, 2, 5, 2), (1, 1, 64, 2)])
def test_mapper_free_counts(nx, ny, n, p):
    prob = problems.square_grid(nx, ny, n, p)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    Nx, Ny = nx * (n + p) - (nx - 1), ny * (n + p) - (ny - 1)
    assert nfree == (Nx - 2) * (Ny - 2)
    assert len(set(dofs.tolist())) == Nx * Ny


def test_glued_p1_equals_single_patch():
    # C0 gluing of linear patches reproduces the single-patch space
    n = 3
    left = problems.box_patch(2, 1, n, lo=[0, 0], hi=[1, 1])
    right = problems.box_patch(2, 1, n, lo=[1, 0], hi=[2, 1])
    for P in (left, right):
        P.fkind = 1
    glued = problems.Problem([left, right], [(0, 1), (0, 3), (0, 4), (1, 2), (1, 3), (1, 4)],
                             [(0, 2, 1, 1, 1, 0, 0, 0)])
    dg, nfg = O.mapper(glued.patches, glued.dir_sides, glued.ifaces)
    Kg, bg = O.assemble(glued.patches, dg, nfg, 1)
    single = problems.box_patch(2, 1, [2 * n, n], lo=[0, 0], hi=[2, 1])
    single.fkind = 1
    sp_ = problems.Problem([single], problems.all_sides(2), [])
    ds, nfs = O.mapper(sp_.patches, sp_.dir_sides, sp_.ifaces)
    Ks, bs = O.assemble(sp_.patches, ds, nfs, 1)
    assert nfg == nfs
    # permutation glued index -> single index via physical coordinates
    pos_g = {}
    for k, P in enumerate(glued.patches):
        for l in range(P.nbasis):
            g = dg[k * P.nbasis + l]
            if g < nfg:
                pos_g[g] = tuple(np.round(P.ctrl[l], 12))
    pos_s = {tuple(np.round(single.ctrl[l], 12)): ds[l] for l in range(single.nbasis) if ds[l] < nfs}
    perm = np.array([pos_s[pos_g[g]] for g in range(nfg)])
    Kp = Ks.toarray()[np.ix_(perm, perm)]
    assert np.max(np.abs(Kg.toarray() - Kp)) <= 1e-13 * np.abs(Kp).max()
    assert np.max(np.abs(bg - bs[perm])) <= 1e-13 * np.abs(bs).max()


# ---------------------------------------------------------------- solvers
def test_cg_fd_poisson_1d():
    n = 30
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    b = np.ones(n)
    xe = np.linalg.solve(A.toarray(), b)
    for pc in ("none", "jacobi"):
        x, it, err = O.cg_solve(A, b, tol=1e-12, maxit=100, precond=pc)
        assert it <= n // 2 + 1      # symmetric tridiagonal, b even: n/2 distinct modes
        assert np.linalg.norm(x - xe) <= 1e-10 * np.linalg.norm(xe)


def _numpy_pcg(A, b, tol, maxit, dinv, tau):
    x = np.zeros_like(b)
    r = b - A @ x
    bn = np.linalg.norm(b)
    hist = [np.linalg.norm(r) / bn]
    z = tau * r * dinv
    p = z.copy()
    rz = r @ z
    it = 0
    while it < maxit:
        it += 1
        t = A @ p
        a = rz / (p @ t)
        x = x + a * p
        r = r - a * t
        hist.append(np.linalg.norm(r) / bn)
        if hist[-1] < tol:
            break
        z = tau * r * dinv
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return x, it, np.array(hist)


def test_cg_history_matches_numpy():
    prob = problems.poisson_noisy_cube(n=4, p=2, seed=2)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    A, b = O.assemble(prob.patches, dofs, nfree, 0)
    x, it, err, hist = O.cg_solve(A, b, tol=1e-10, maxit=500, tau=0.8, history=True)
    xn, itn, hn = _numpy_pcg(A, b, 1e-10, 500, 1 / A.diagonal(), 0.8)
    assert it == itn
    assert np.allclose(hist, hn, rtol=1e-9, atol=0)
    assert np.linalg.norm(x - xn) <= 1e-9 * np.linalg.norm(xn)


# ---------------------------------------------------------------- multigrid
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_transfer_1d(p):
    kn = problems.knots_uniform(8, p)
    T = O.transfer_1d(kn, p)
    fine = problems.knots_uniform(16, p)
    nc, nf = len(kn) - p - 1, len(fine) - p - 1
    assert T.shape == (nf, nc)
    # refinement identity N^c_c(u) = sum_f T[f,c] N^f_f(u)
    u = np.linspace(0, 1, 41)[:-1] + 0.0123
    for c in range(nc):
        cc = np.zeros(nc)
        cc[c] = 1
        lhs = BSpline(kn, cc, p)(u)
        rhs = BSpline(fine, T[:, c], p)(u)
        assert np.max(np.abs(lhs - rhs)) < 1e-14
    # interior columns: binomial subdivision mask / 2^p
    from math import comb
    mask = np.array([comb(p + 1, k) for k in range(p + 2)]) / 2.0 ** p
    c = nc // 2
    col = T[:, c]
    nzr = np.nonzero(col)[0]
    assert np.allclose(col[nzr[0]:nzr[-1] + 1], mask, atol=0, rtol=4e-16)


def test_galerkin_identity_multipatch():
    coarse = problems.square_grid(2, 2, 2, 3)
    fine = problems.square_grid(2, 2, 4, 3)
    dc, nc = O.mapper(coarse.patches, coarse.dir_sides, coarse.ifaces)
    df, nf = O.mapper(fine.patches, fine.dir_sides, fine.ifaces)
    Ac, _ = O.assemble(coarse.patches, dc, nc, 1)
    Af, _ = O.assemble(fine.patches, df, nf, 1)
    P = O.transfer_multipatch(coarse.patches, dc, nc, df, nf)
    G = (P.T @ Af @ P).toarray()
    assert np.max(np.abs(G - Ac.toarray())) <= 1e-12 * np.abs(Ac).max()
    mg = O.Multigrid(Af, [P], npre=2, npost=2, damping=0.6)
    Gm = mg.matrix(0).toarray()
    assert np.max(np.abs(Gm - Ac.toarray())) <= 1e-12 * np.abs(Ac).max()


def test_two_grid_cycle_matches_numpy():
    coarse = problems.square_grid(2, 2, 2, 2)
    fine = problems.square_grid(2, 2, 4, 2)
    dc, nc = O.mapper(coarse.patches, coarse.dir_sides, coarse.ifaces)
    df, nf = O.mapper(fine.patches, fine.dir_sides, fine.ifaces)
    Af, bf = O.assemble(fine.patches, df, nf, 1)
    P = O.transfer_multipatch(coarse.patches, dc, nc, df, nf)
    mg = O.Multigrid(Af, [P], npre=2, npost=2, damping=0.6)
    z = mg.apply(bf)
    # numpy V(2,2) two-grid with exact coarse solve
    A = Af.toarray()
    Pd = P.toarray()
    Ac = Pd.T @ A @ Pd
    D = np.diag(A)
    x = np.zeros(nf)
    for _ in range(2):
        x = x + 0.6 * (bf - A @ x) / D
    res = A @ x - bf
    c = np.linalg.solve(Ac, Pd.T @ res)
    x = x - Pd @ c
    for _ in range(2):
        x = x + 0.6 * (bf - A @ x) / D
    assert np.linalg.norm(z - x) <= 1e-11 * np.linalg.norm(x)


def test_mg_pcg_converges_fast():
    levels = [problems.square_grid(2, 2, n, 3) for n in (4, 8, 16)]
    maps = [O.mapper(p.patches, p.dir_sides, p.ifaces) for p in levels]
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 1)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(2)]
    mg = O.Multigrid(Af, Ps, npre=2, npost=2, damping=0.6)
    x, it, err = O.cg_solve(Af, bf, tol=1e-8, maxit=200, precond="mg", mg=mg)
    xj, itj, errj = O.cg_solve(Af, bf, tol=1e-8, maxit=2000, precond="jacobi")
    assert err < 1e-8 and it < itj / 2
    assert np.linalg.norm(x - xj) <= 1e-6 * np.linalg.norm(xj)


# ---------------------------------------------------------------- L2 error
def test_l2_convergence_rate():
    errs = []
    for n in (4, 8, 16):
        prob = problems.poisson_square(n, 2)
        dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        A, b = O.assemble(prob.patches, dofs, nfree, 0)
        x = np.linalg.solve(A.toarray(), b)
        errs.append(O.l2_error(prob.patches, dofs, nfree, x))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 2.8), rates   # p+1 = 3 (the reference asserts > 1.85)


@pytest.mark.parametrize("d,p,n", [(2, 3, 16), (3, 3, 6)])
def test_error_norms_of_zero_solution_closed_form(d, p, n):
    """With u_h = 0 the errors are the norms of u = prod sin(pi x_r):
    ||u||_L2 = 2^(-d/2), |u|_H1 = pi sqrt(d) 2^(-d/2) (up to the Gauss rule's
    error on sin^2, tiny at these sizes)."""
    P = problems.box_patch(d, p, n)
    prob = problems.Problem([P], problems.all_sides(d), [])
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    z = np.zeros(nfree)
    assert O.l2_error(prob.patches, dofs, nfree, z) == pytest.approx(2.0 ** (-d / 2), rel=1e-7)
    assert O.h1_error(prob.patches, dofs, nfree, z) == pytest.approx(np.pi * np.sqrt(d) * 2.0 ** (-d / 2), rel=1e-7)


def test_h1_convergence_rate():
    errs = []
    for n in (4, 8, 16):
        prob = problems.poisson_square(n, 2)
        dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        A, b = O.assemble(prob.patches, dofs, nfree, 0)
        x = np.linalg.solve(A.toarray(), b)
        errs.append(O.h1_error(prob.patches, dofs, nfree, x))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 1.9) and np.all(rates < 2.3), rates   # order p = 2


def _two_grid_setup():
    coarse = problems.square_grid(2, 2, 2, 2)
    fine = problems.square_grid(2, 2, 4, 2)
    dc, nc = O.mapper(coarse.patches, coarse.dir_sides, coarse.ifaces)
    df, nf = O.mapper(fine.patches, fine.dir_sides, fine.ifaces)
    Af, bf = O.assemble(fine.patches, df, nf, 1)
    P = O.transfer_multipatch(coarse.patches, dc, nc, df, nf)
    return Af, bf, P


def _numpy_two_grid(A, b, Pd, smooth):
    Ac = Pd.T @ A @ Pd
    x = smooth(np.zeros(len(b)))
    c = np.linalg.solve(Ac, Pd.T @ (A @ x - b))
    return smooth(x - Pd @ c)


def test_two_grid_richardson_and_chebyshev_match_numpy():
    """Two-grid cycles with Richardson and Chebyshev smoothing against numpy
    restatements; the Chebyshev bound against the largest eigenvalue of
    D^-1 A (power iteration: within 5 %, times the 1.1 safety factor)."""
    Af, bf, P = _two_grid_setup()
    A, Pd, D = Af.toarray(), P.toarray(), np.diag(Af.toarray())
    mg = O.Multigrid(Af, [P], npre=2, npost=2, damping=0.3 / np.abs(D).max(), smoother="richardson")
    tau = 0.3 / np.abs(D).max()

    def rich(x):
        for _ in range(2):
            x = x + tau * (bf - A @ x)
        return x
    z = mg.apply(bf)
    ref = _numpy_two_grid(A, bf, Pd, rich)
    assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)

    for m in (1, 3):
        mg = O.Multigrid(Af, [P], npre=m, npost=m, smoother="chebyshev", smoother_param=20.0)
        lmax = mg.lmax(1)
        ev = np.max(np.linalg.eigvals(A / D[:, None]).real)
        assert 1.1 * ev * 0.95 <= lmax <= 1.1 * ev * 1.0000001, (lmax, 1.1 * ev)
        l1, l0 = lmax, lmax / 20.0
        theta, delta = (l1 + l0) / 2, (l1 - l0) / 2
        sigma = theta / delta

        def cheb(x, m=m):
            rho = 1.0 / sigma
            d = (bf - A @ x) / D / theta
            x = x + d
            for _ in range(m - 1):
                rn = 1.0 / (2 * sigma - rho)
                d = rn * rho * d + 2 * rn / delta * (bf - A @ x) / D
                x = x + d
                rho = rn
            return x
        z = mg.apply(bf)
        ref = _numpy_two_grid(A, bf, Pd, cheb)
        assert np.linalg.norm(z - ref) <= 1e-11 * np.linalg.norm(ref)


@pytest.mark.parametrize("m,ratio", [(1, 30.0), (2, 30.0), (3, 30.0), (3, 4.0), (5, 10.0)])
def test_chebyshev_error_propagator_is_the_chebyshev_polynomial(m, ratio):
    """The m-step Chebyshev smoother on [l1/ratio, l1] damps the error by
    p_m(D^-1 A) with p_m(lam) = T_m((theta - lam)/delta) / T_m(theta/delta)
    (theta, delta: midpoint and half-width of the interval).  Checked through
    a two-grid cycle whose spectrum is known in closed form: A = D^1/2 C D^1/2
    with C a symmetric circulant with unit diagonal (D^-1 A is similar to C,
    eigenvalues 1 + 2 sum_j c_j cos(2 pi j k / n)) and the coarse space
    D^-1/2 1 (C's eigenvector of k = 0).  Then, in the coordinates
    D^1/2 x, the cycle's error propagator is p_m(C) (I - q0 q0^T) p_m(C) and
    the cycle from x = 0 returns z = D^-1/2 (I - E) D^1/2 A^-1 b."""
    n = 48
    c = np.zeros(n)
    c[0], c[1], c[2] = 1.0, -0.31, 0.07
    c[-1], c[-2] = c[1], c[2]
    C = np.array([np.roll(c, i) for i in range(n)])
    k = np.arange(n)
    lam_closed = 1.0 + 2 * (c[1] * np.cos(2 * np.pi * k / n) + c[2] * np.cos(4 * np.pi * k / n))
    lam, Q = np.linalg.eigh(C)
    assert np.allclose(np.sort(lam_closed), lam, atol=1e-14)
    rng = np.random.default_rng(5)
    dg = rng.uniform(0.5, 2.0, n)
    A = np.sqrt(dg)[:, None] * C * np.sqrt(dg)[None, :]
    P = (1.0 / np.sqrt(dg))[:, None]
    mg = O.Multigrid(sp.csr_matrix(A), [sp.csr_matrix(P)], npre=m, npost=m, smoother="chebyshev",
                     smoother_param=ratio)
    l1 = mg.lmax(1)
    l0 = l1 / ratio
    theta, delta = (l1 + l0) / 2, (l1 - l0) / 2
    Tm = np.polynomial.chebyshev.Chebyshev.basis(m)
    pm = Tm((theta - lam) / delta) / Tm(theta / delta)
    pC = (Q * pm) @ Q.T
    q0 = np.ones(n) / np.sqrt(n)
    E = pC @ (np.eye(n) - np.outer(q0, q0)) @ pC
    b = rng.standard_normal(n)
    xs = np.linalg.solve(A, b)
    want = (1.0 / np.sqrt(dg)) * ((np.eye(n) - E) @ (np.sqrt(dg) * xs))
    z = mg.apply(b)
    assert np.linalg.norm(z - want) <= 1e-12 * np.linalg.norm(want)
    # a dropped recurrence term (d_k without its rho_k rho_{k-1} d_{k-1}
    # part) is a different polynomial for m >= 2: the check separates them
    if m >= 2:
        assert np.abs(pm).max() < 1.0


@pytest.mark.parametrize("smoother,npre,npost", [("jacobi", 2, 2), ("chebyshev", 3, 3),
                                                  ("gauss_seidel", 1, 1), ("jacobi", 1, 3)])
def test_vcycle_is_symmetric_and_contracting(smoother, npre, npost):
    """Three-level V(nu, nu) cycle with a symmetric smoother pair (damped
    Jacobi, Chebyshev, forward Gauss-Seidel before and reverse after) is a
    symmetric preconditioner B, and I - BA is a contraction in the energy
    norm: the eigenvalues of BA lie in (0, 1].  An unbalanced cycle V(1, 3)
    is not symmetric -- the check can tell."""
    levels = [problems.square_grid(2, 2, nn, 2) for nn in (1, 2, 4)]
    maps = [O.mapper(p.patches, p.dir_sides, p.ifaces) for p in levels]
    Af, _ = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 1)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(2)]
    kw = dict(damping=0.6) if smoother == "jacobi" else dict(smoother=smoother)
    mg = O.Multigrid(Af, Ps, npre=npre, npost=npost, **kw)
    n = Af.shape[0]
    B = np.array([mg.apply(np.eye(n)[i]) for i in range(n)]).T
    asym = np.abs(B - B.T).max() / np.abs(B).max()
    if npre != npost:
        assert asym > 1e-6
        return
    assert asym <= 1e-12
    ev = np.linalg.eigvals(B @ Af.toarray())
    assert np.abs(ev.imag).max() <= 1e-10
    assert ev.real.min() > 0.0 and ev.real.max() <= 1.0 + 1e-10, (ev.real.min(), ev.real.max())


def test_chebyshev_mg_pcg_beats_jacobi_mg_in_3d():
    """Cube p=3: Chebyshev(3) smoothing needs fewer MG-PCG iterations than
    three damped Jacobi sweeps."""
    levels = [problems.box_patch(3, 3, n) for n in (2, 4, 8)]
    probs = [problems.Problem([P], problems.all_sides(3), []) for P in levels]
    maps = [O.mapper(p.patches, p.dir_sides, p.ifaces) for p in probs]
    Af, bf = O.assemble(probs[-1].patches, maps[-1][0], maps[-1][1], 1)
    Ps = [O.transfer_multipatch(probs[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(2)]
    its = {}
    for name, kw in (("jacobi", dict(damping=0.6)), ("chebyshev", dict(smoother="chebyshev"))):
        mg = O.Multigrid(Af, Ps, npre=3, npost=3, **kw)
        x, it, err = O.cg_solve(Af, bf, tol=1e-10, maxit=500, precond="mg", mg=mg)
        assert err < 1e-10
        its[name] = it
    assert its["chebyshev"] < its["jacobi"], its


# ---------------------------------------------------------------------------
# Boundary data: Dirichlet values by L2 projection, elimination, Neumann load
# ---------------------------------------------------------------------------
AFFINE = (0.3, 1.1, -0.7, 0.5)


def _bc_patch(d, p, n, dir_sides, seed, lo=None, hi=None, fkind=0, curved=True):
    """Degree-p box patch whose control points (boundary included) are moved
    by U(+-0.1 h): curved sides, a non-affine map."""
    P = problems.box_patch(d, p, n, lo, hi)
    if curved:
        rng = np.random.default_rng(seed)
        P.ctrl = P.ctrl + rng.uniform(-0.1 / n, 0.1 / n, P.ctrl.shape)
    P.fkind = fkind
    if fkind == 0:
        P.felem = np.zeros(P.nelem)
    return problems.Problem([P], [(0, s) for s in dir_sides], [], name=f"bc-d{d}-p{p}-n{n}")


def _affine_at_ctrl(prob, dofs, nfree, lo_fixed):
    """c + a . C_m at every basis function: the coefficients of the affine
    function in the isoparametric space (x = sum C_m N_m, sum N_m = 1)."""
    P = prob.patches[0]
    vals = AFFINE[0] + P.ctrl @ np.array(AFFINE[1:1 + P.d])
    nfix = P.nbasis - nfree
    free = np.zeros(nfree)
    fixed = np.zeros(nfix)
    for m, g in enumerate(dofs):
        if g < nfree:
            free[g] = vals[m]
        else:
            fixed[g - nfree] = vals[m]
    return free, fixed


@pytest.mark.parametrize("d,p", [(2, 2), (2, 3), (3, 2)])
def test_dirichlet_projection_reproduces_affine_data(d, p):
    """The boundary L2 projection of an affine g (which lies in the trace of
    the isoparametric space) returns its coefficients c + a . C_m exactly."""
    sides = list(range(1, 2 * d + 1))
    prob = _bc_patch(d, p, 4, sides, seed=11)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = prob.patches[0].nbasis - nfree
    bcs = [(0, s, 1, 2, AFFINE) for s in sides]
    fixed = O.dirichlet_values(prob.patches, dofs, nfree, nfix, bcs)
    _, expect = _affine_at_ctrl(prob, dofs, nfree, 0)
    assert np.abs(fixed - expect).max() <= 1e-12


@pytest.mark.parametrize("d,p,skip", [(2, 3, [3]), (3, 2, [3, 6])])
def test_projection_skipping_homogeneous_sides(d, p, skip):
    """gsAssembler::computeDirichletDofsL2Proj (src/gsAssembler/
    gsAssembler.hpp:393-534) projects the inhomogeneous sides only: with an
    affine g on those and the homogeneous sides left out, every function
    with a trace on a projected side (the shared edges and corners
    included) gets g's coefficient c + a . C_m exactly, and the functions
    living on the skipped sides alone (zero rows of the projection matrix)
    stay 0."""
    sides = list(range(1, 2 * d + 1))
    prob = _bc_patch(d, p, 4, sides, seed=13)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = prob.patches[0].nbasis - nfree
    bcs = [(0, s, 1, 2, AFFINE) for s in sides if s not in skip]
    fixed = O.dirichlet_values(prob.patches, dofs, nfree, nfix, bcs)
    _, expect = _affine_at_ctrl(prob, dofs, nfree, 0)
    P = prob.patches[0]
    nb = [P.knots[r].size - P.p[r] - 1 for r in range(d)]
    on_projected = np.zeros(nfix, bool)
    for m, g in enumerate(dofs):
        idx = np.unravel_index(m, nb[::-1])[::-1]   # direction 0 fastest
        for s in sides:
            if s in skip:
                continue
            r, last = (s - 1) // 2, (s - 1) % 2
            if idx[r] == (nb[r] - 1 if last else 0):
                on_projected[g - nfree] = True
    assert on_projected.any() and (~on_projected).any()
    assert np.abs(fixed[on_projected] - expect[on_projected]).max() <= 1e-12
    assert np.all(fixed[~on_projected] == 0.0)


@pytest.mark.parametrize("d,p,dir_sides", [(2, 2, [1, 3]), (2, 3, [2, 4]), (3, 2, [1, 3, 6])])
def test_mixed_boundary_patch_test(d, p, dir_sides):
    """Patch test: u = c + a . x solves -div grad u = 0; with u on the
    Dirichlet sides (projected) and the flux a . nv on the Neumann sides,
    the discrete solution is u itself on a curved, non-affine patch -- pins
    the elimination (rhs -= K g), the Neumann load's normal (sign and
    measure) and the projection together."""
    prob = _bc_patch(d, p, 4, dir_sides, seed=5)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = prob.patches[0].nbasis - nfree
    neu = [s for s in range(1, 2 * d + 1) if s not in dir_sides]
    bcs = [(0, s, 1, 2, AFFINE) for s in dir_sides] + [(0, s, 2, 2, AFFINE) for s in neu]
    fixed = O.dirichlet_values(prob.patches, dofs, nfree, nfix, bcs)
    A, b = O.assemble(prob.patches, dofs, nfree, 0, fixed=fixed)
    neu_load = O.neumann_load(prob.patches, dofs, nfree, bcs)
    import scipy.sparse.linalg as spla
    x = spla.spsolve(A.tocsc(), b + neu_load)
    expect, _ = _affine_at_ctrl(prob, dofs, nfree, 0)
    assert np.abs(x - expect).max() <= 1e-10 * np.abs(expect).max()
    # the same with the normal flipped (inward) does not reproduce u
    x_bad = spla.spsolve(A.tocsc(), b - neu_load)
    assert np.abs(x_bad - expect).max() > 1e-3


def test_mixed_boundary_manufactured_convergence():
    """u = sin(pi x) sin(pi y) on [0.25, 1.25]^2 (non-zero on the boundary):
    projected Dirichlet data on west/south, flux grad u . nv on east/north;
    the L2 error falls at order p + 1 = 3."""
    errs = []
    for n in (4, 8, 16):
        prob = _bc_patch(2, 2, n, [1, 3], seed=0, lo=[0.25, 0.25], hi=[1.25, 1.25], fkind=1, curved=False)
        dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
        nfix = prob.patches[0].nbasis - nfree
        bcs = [(0, 1, 1, 1, ()), (0, 3, 1, 1, ()), (0, 2, 2, 1, ()), (0, 4, 2, 1, ())]
        fixed = O.dirichlet_values(prob.patches, dofs, nfree, nfix, bcs)
        A, b = O.assemble(prob.patches, dofs, nfree, 0, fixed=fixed)
        b = b + O.neumann_load(prob.patches, dofs, nfree, bcs)
        x, it, err = O.cg_solve(A, b, tol=1e-13, maxit=5000)
        errs.append(O.l2_error(prob.patches, dofs, nfree, x, fixed=fixed))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print(errs, rates)
    assert rates[-1] > 2.8, rates


# ---------------------------------------------------------------------------
# fast diagonalization (gsPatchPreconditionersCreator)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("d,p,n", [(2, 2, 6), (3, 3, 4), (3, 2, 5)])
def test_kron_sum_is_the_unit_box_stiffness(d, p, n):
    """On the identity map with kappa = 1 the stiffness matrix is the
    Kronecker sum of the univariate matrices: the oracle's d-dimensional
    assembly pins the univariate matrices and the Kronecker ordering."""
    P = problems.box_patch(d, p, n)
    P.fkind = 1
    sides = list(range(1, 2 * d + 1))
    dofs, nfree = O.mapper([P], [(0, s) for s in sides], [])
    A, _ = O.assemble([P], dofs, nfree, 0)
    K = O.kron_sum(P, sides)
    assert K.shape == A.shape
    assert abs(K - A).max() <= 1e-14 * abs(A).max()


@pytest.mark.parametrize("d,p,sides,alpha", [(2, 3, [1, 2, 3, 4], 0.0), (2, 2, [1, 4], 0.3),
                                             (3, 2, [1, 3, 6], 0.0), (3, 3, [], 1.5)])
def test_fd_is_the_inverse_of_the_kronecker_sum(d, p, sides, alpha):
    """fastDiagonalizationOp applies (alpha M + sum_k M x .. x K_k x .. x M)^-1
    exactly, with Neumann ends kept (and alpha > 0 where no end is fixed)."""
    import scipy.sparse.linalg as spla
    P = problems.box_patch(d, p, 5)
    A = O.kron_sum(P, sides, alpha=alpha, beta=1.0)
    fd = O.fd_operator(P, sides, alpha=alpha)
    r = np.random.default_rng(3).standard_normal(A.shape[0])
    want = spla.spsolve(A.tocsc(), r)
    assert np.abs(fd(r) - want).max() <= 1e-10 * np.abs(want).max()


# ---------------------------------------------------------------------------
# subspace corrected mass smoother (gsPatchPreconditionersCreator)
# ---------------------------------------------------------------------------

def _graded_knots(n, p, seed):
    """Open knot vector on [0, 1] with n spans of random lengths (ratio up
    to 3) and simple interior knots."""
    w = np.random.default_rng(seed).uniform(1.0, 3.0, n)
    inner = np.cumsum(w)[:-1] / w.sum()
    return np.concatenate([np.zeros(p + 1), inner, np.ones(p + 1)])


def _graded_box(d, p, n, seed):
    P = problems.box_patch(d, p, n)
    P.knots = [_graded_knots(n, p, seed + k) for k in range(d)]
    return P


@pytest.mark.parametrize("p", [2, 3, 4, 5])
@pytest.mark.parametrize("dw,de", [(False, False), (True, False), (False, True), (True, True)])
def test_tilde_space_is_the_space_of_vanishing_odd_derivatives(p, dw, de):
    """tildeSpaceBasis: every column of B_tilde, as a spline on the free
    functions, has vanishing derivatives of the odd orders below p at both
    ends (evaluated with all active functions, so a wrong column offset or
    a kept Dirichlet function fails); B_tilde and B_l2compl together form
    an orthonormal-by-blocks basis of R^n, the complement has one column
    per constraint (p // 2 per end) and is orthogonal to B_tilde."""
    knots = _graded_knots(9, p, 7 * p + dw + 2 * de)
    nb = len(knots) - p - 1
    n = nb - dw - de
    Bt, Bl = O.tilde_space(knots, p, dw, de)
    assert Bt.shape == (n, n - 2 * (p // 2)) and Bl.shape == (n, 2 * (p // 2))
    assert np.abs(Bt.T @ Bl).max() <= 1e-13
    assert np.linalg.matrix_rank(np.hstack([Bt, Bl])) == n
    for u in (knots[0], knots[-1]):
        span = O.find_span(knots, p, u)
        ders = O.bspline_ders(knots, p, span, u, p - 1)
        for c in range(Bt.shape[1]):
            full = np.zeros(nb)
            full[int(dw):int(dw) + n] = Bt[:, c]
            act = full[span - p:span + 1]
            for i in range(1, p, 2):
                scale = np.abs(ders[i]).max()
                assert abs(ders[i] @ act) <= 1e-11 * scale, (u, c, i)
    # the constraints restricted to the complement have full rank
    rows = []
    for u in (knots[0], knots[-1]):
        span = O.find_span(knots, p, u)
        ders = O.bspline_ders(knots, p, span, u, p - 1)
        for i in range(1, p, 2):
            r = np.zeros(nb)
            r[span - p:span + 1] = ders[i]
            rows.append(r[int(dw):int(dw) + n])
    Cmap = np.array(rows).reshape(-1, n)
    assert np.linalg.matrix_rank(Cmap @ Bl) == Bl.shape[1]


@pytest.mark.parametrize("d,sides,sigma,alpha,beta,graded", [(2, [1, 2, 3, 4], 0.12, 0.0, 1.0, False),
                                                             (2, [2, 3], 0.3, 0.5, 2.0, True),
                                                             (3, [1, 6], 0.12, 0.0, 1.0, True),
                                                             (3, [], 0.2, 1.0, 0.5, False)])
def test_scms_degree_one_is_the_scaled_mass_inverse(d, sides, sigma, alpha, beta, graded):
    """For p = 1 there are no odd-derivative constraints: the tilde space is
    everything, only the interior type survives and the smoother is
    (x)_k M_k^-1 / (alpha + beta d / (sigma h^2)) with h the smallest knot
    span of the basis."""
    P = _graded_box(d, 1, 6, 3) if graded else problems.box_patch(d, 1, 6)
    S = O.scms_matrix(P, sides, sigma=sigma, alpha=alpha, beta=beta)
    h = min(np.diff(np.unique(P.knots[k])).min() for k in range(d))
    Minv = np.ones((1, 1))
    for k in reversed(range(d)):
        lo, hi = O._free_range(P, k, sides)
        _, M = O.tensor_1d(P.knots[k], 1, lo, hi)
        Minv = np.kron(Minv, np.linalg.inv(M))
    want = Minv / (alpha + beta * d / (sigma * h * h))
    assert np.abs(S - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("d,p,sides", [(2, 2, [1, 2, 3, 4]), (2, 3, [1, 4]), (2, 4, []), (3, 3, [2, 3, 5]),
                                       (3, 2, [1, 2, 3, 4, 5, 6])])
def test_scms_mass_limit_is_the_mass_inverse(d, p, sides):
    """beta -> 0 leaves sum_J T_J (T_J^T M T_J)^-1 T_J^T / alpha, which is
    M^-1 / alpha exactly when the subspaces are M-orthogonal and together
    span the space (B_compl = M^-1 B_l2compl, all 2^d types kept)."""
    P = _graded_box(d, p, 7 if d == 2 else 5, 11 * p + d)
    S = O.scms_matrix(P, sides, sigma=0.12, alpha=1.0, beta=1e-13)
    Minv = np.ones((1, 1))
    for k in reversed(range(d)):
        lo, hi = O._free_range(P, k, sides)
        _, M = O.tensor_1d(P.knots[k], p, lo, hi)
        Minv = np.kron(Minv, np.linalg.inv(M))
    assert np.abs(S - Minv).max() <= 1e-8 * np.abs(Minv).max()


@pytest.mark.parametrize("d,p,sides,alpha", [(2, 3, [1, 2, 3, 4], 0.0), (2, 2, [3], 0.7), (3, 3, [1, 6], 0.0),
                                             (3, 4, [], 1.0)])
def test_scms_is_the_galerkin_subspace_correction(d, p, sides, alpha):
    """Each correction is the Galerkin restriction of the full tensor
    operator alpha M + beta sum_{k in J} K_k (x) M_rest + beta (d - |J|) /
    (sigma h^2) M (stiffness in the boundary directions, scaled mass in the
    interior ones) to the range of T_J: assembled here from the full
    Kronecker matrices instead of the univariate blocks, and the result is
    symmetric positive definite."""
    sigma, beta = 0.12, 1.3
    P = _graded_box(d, p, 7 if d == 2 else 5, 5 * p + d)
    S = O.scms_matrix(P, sides, sigma=sigma, alpha=alpha, beta=beta)
    h = min(np.diff(np.unique(P.knots[k])).min() for k in range(d))
    Ks, Ms, Bt, Bc = [], [], [], []
    for k in range(d):
        lo, hi = O._free_range(P, k, sides)
        nb = len(P.knots[k]) - p - 1
        K, M = O.tensor_1d(P.knots[k], p, lo, hi)
        bt, bl = O.tilde_space(P.knots[k], p, lo == 1, hi == nb - 2)
        Ks.append(K)
        Ms.append(M)
        Bt.append(bt)
        Bc.append(np.linalg.solve(M, bl))

    def kr(ms):
        out = np.ones((1, 1))
        for m in reversed(ms):
            out = np.kron(out, m)
        return out
    want = np.zeros_like(S)
    for J in range(1 << d):
        inJ = [(J >> k) & 1 for k in range(d)]
        T = kr([Bc[k] if inJ[k] else Bt[k] for k in range(d)])
        if T.shape[1] == 0:
            continue
        A = (alpha + beta * (d - sum(inJ)) / (sigma * h * h)) * kr(Ms)
        for j in range(d):
            if inJ[j]:
                A = A + beta * kr([Ks[k] if k == j else Ms[k] for k in range(d)])
        want += T @ np.linalg.inv(T.T @ A @ T) @ T.T
    assert np.abs(S - want).max() <= 1e-10 * np.abs(want).max()
    assert np.abs(S - S.T).max() <= 1e-12 * np.abs(S).max()
    assert np.linalg.eigvalsh(0.5 * (S + S.T)).min() > 0


def test_mg_vcycle_restates_the_c_oracle_cycle():
    """The numpy V-cycle with a Richardson smoother equals the C oracle's
    cycle (gsMultiGridOp with gsRichardsonOp) on a three-level cube."""
    levels = [problems.poisson_noisy_cube(n=n, p=2, seed=4) for n in (2, 4, 8)]
    maps = [O.mapper(L.patches, L.dir_sides, L.ifaces) for L in levels]
    Af, bf = O.assemble(levels[-1].patches, maps[-1][0], maps[-1][1], 0)
    Ps = [O.transfer_multipatch(levels[i].patches, maps[i][0], maps[i][1], maps[i + 1][0], maps[i + 1][1])
          for i in range(2)]
    tau = 0.3 / np.abs(Af.diagonal()).max()
    omg = O.Multigrid(Af, Ps, npre=2, npost=1, damping=tau, smoother="richardson")
    As = [omg.matrix(l) for l in range(3)]

    def rich(l):
        def f(b, x, post):
            x += tau * (b - As[l] @ x)
        return f
    cyc = O.mg_vcycle(As, Ps, [None, rich(1), rich(2)], npre=2, npost=1)
    z = cyc(bf)
    zo = omg.apply(bf)
    assert np.linalg.norm(z - zo) <= 1e-12 * np.linalg.norm(zo)


def test_scms_damping_is_the_inverse_spectral_radius():
    """Ten power iterations of S A reach its largest eigenvalue (S and A
    symmetric positive definite, so S A has a real spectrum) to within a few
    percent, from below; the exact value comes from the generalized
    eigenproblem A v = mu S^-1 v."""
    import scipy.linalg as sla
    prob = problems.poisson_square(n=8, p=3)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    A, _ = O.assemble(prob.patches, dofs, nfree, 0)
    S = O.scms_matrix(prob.patches[0], [1, 2, 3, 4])
    lam = 1.0 / O.scms_damping(A, S)
    mu = sla.eigh(A.toarray(), np.linalg.inv(S), eigvals_only=True).max()
    assert 0.9 * mu <= lam <= mu * (1 + 1e-12)


def test_pcg_restates_the_cg_recurrences():
    """The callable-preconditioner PCG gives the C oracle's iteration count
    and solution with the Jacobi preconditioner."""
    prob = problems.poisson_noisy_cube(n=6, p=2, seed=3)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    A, b = O.assemble(prob.patches, dofs, nfree, 0)
    x, it, err = O.pcg(A, b, lambda r: r / A.diagonal(), tol=1e-10)
    xo, ito, erro = O.cg_solve(A, b, tol=1e-10)
    assert it == ito
    assert np.abs(x - xo).max() <= 1e-9 * np.abs(xo).max()


def test_gauss_seidel_sweeps_and_two_grid_match_closed_forms():
    """A forward Gauss-Seidel sweep is x <- (D + L)^-1 (b - U x), a reverse
    sweep x <- (D + U)^-1 (b - L x); the two-grid cycle runs forward sweeps
    before and reverse sweeps after the coarse correction (gsMultiGridOp's
    SymmSmooth with stepT)."""
    Af, bf, P = _two_grid_setup()
    A, Pd = Af.toarray(), P.toarray()
    lower, upper = np.tril(A), np.triu(A)

    def fwd(x):
        return np.linalg.solve(lower, bf - (A - lower) @ x)

    def bwd(x):
        return np.linalg.solve(upper, bf - (A - upper) @ x)
    x0 = np.random.default_rng(4).standard_normal(len(bf))
    assert np.abs(O.gs_sweep(Af, bf, x0) - fwd(x0)).max() <= 1e-12 * np.abs(fwd(x0)).max()
    assert np.abs(O.gs_sweep(Af, bf, x0, reverse=True) - bwd(x0)).max() <= 1e-12 * np.abs(bwd(x0)).max()
    Ac = Pd.T @ A @ Pd
    for m in (1, 2):
        mg = O.Multigrid(Af, [P], npre=m, npost=m, smoother="gauss_seidel")
        x = np.zeros(len(bf))
        for _ in range(m):
            x = fwd(x)
        x = x - Pd @ np.linalg.solve(Ac, Pd.T @ (A @ x - bf))
        for _ in range(m):
            x = bwd(x)
        z = mg.apply(bf)
        assert np.linalg.norm(z - x) <= 1e-11 * np.linalg.norm(x)


@pytest.mark.parametrize("d,p", [(2, 3), (3, 2)])
def test_dirichlet_interpolation_pins(d, p):
    """gsAssembler::computeDirichletDofsIntpl: an affine g is reproduced
    exactly (its trace lies in the side's isoparametric space); the side
    trace of the interpolant matches a manufactured g at every anchor
    (scipy's B-splines evaluate the trace); a later side overwrites the
    functions it shares with an earlier one (a homogeneous side writes
    zeros)."""
    from scipy.interpolate import BSpline
    sides = list(range(1, 2 * d + 1))
    prob = _bc_patch(d, p, 4, sides, seed=21)
    P = prob.patches[0]
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    nfix = P.nbasis - nfree
    fx = O.dirichlet_interpolation([P], dofs, nfree, nfix, [(0, s, 1, 2, AFFINE) for s in sides])
    _, expect = _affine_at_ctrl(prob, dofs, nfree, 0)
    assert np.abs(fx - expect).max() <= 1e-13
    # interpolation property on side 1 (west, u_0 = 0) for g = prod sin(pi x)
    fx = O.dirichlet_interpolation([P], dofs, nfree, nfix, [(0, 1, 1, 1, ())])
    nb = P.nbasis_dir()
    coef = np.zeros(nb[1:][::-1])
    for idx in np.ndindex(*nb[1:][::-1]):
        m = 0 + nb[0] * (idx[-1] + (nb[1] * idx[0] if d == 3 else 0))
        coef[idx] = fx[dofs[m] - nfree]
    grs = [O._greville(P.knots[k], P.p[k]) for k in range(1, d)]
    for js in np.ndindex(*[len(g) for g in grs]):
        u = [0.0] + [grs[t][js[t]] for t in range(d - 1)]
        want = np.prod(np.sin(np.pi * O.geometry_at(P, u)))
        vals = [np.array([BSpline(P.knots[k], np.eye(nb[k])[j], P.p[k])(u[k]) for j in range(nb[k])])
                for k in range(1, d)]
        got = np.einsum('i,i->', vals[0], coef) if d == 2 else np.einsum('j,i,ij->', vals[0], vals[1], coef)
        assert abs(got - want) <= 1e-12, (js, got, want)
    # overwrite order at the corner of sides 1 and 3
    corner = dofs[0] - nfree
    a = O.dirichlet_interpolation([P], dofs, nfree, nfix, [(0, 1, 1, 2, AFFINE), (0, 3, 1, 0, ())])
    b = O.dirichlet_interpolation([P], dofs, nfree, nfix, [(0, 3, 1, 0, ()), (0, 1, 1, 2, AFFINE)])
    assert a[corner] == 0.0
    assert abs(b[corner] - expect[corner]) <= 1e-13
