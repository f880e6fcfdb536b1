"""Fast diagonalization (gsPatchPreconditionersCreator::fastDiagonalizationOp)
on the GPU through the C ABI: the apply against the oracle's restatement,
CG preconditioned by it against the oracle's CG (iterations and solution) in
FP64 and mixed precision, the one-iteration solve where it is the exact
inverse, and its error paths."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from gismo_amd import gsx, problems

pytestmark = pytest.mark.gpu


def _curved(d, p, n, seed, sides):
    P = problems.box_patch(d, p, n)
    rng = np.random.default_rng(seed)
    P.ctrl = P.ctrl + rng.uniform(-0.1 / n, 0.1 / n, P.ctrl.shape)
    P.fkind = 1
    return problems.Problem([P], [(0, s) for s in sides], [])


@pytest.mark.parametrize("d,p,n,sides,alpha", [(2, 3, 9, [1, 2, 3, 4], 0.0), (2, 2, 7, [1, 4], 0.25),
                                               (3, 3, 5, [1, 2, 3, 4, 5, 6], 0.0), (3, 2, 4, [2, 3, 5], 0.0),
                                               (3, 4, 3, [1, 2, 3, 4, 5, 6], 2.0),
                                               (2, 2, 120, [1, 2, 3, 4], 0.0)])
def test_fd_apply_matches_oracle(ctx, d, p, n, sides, alpha):
    print(f"[case] fd apply d={d} p={p} n={n} sides={sides} alpha={alpha}", flush=True)
    prob = _curved(d, p, n, 5, sides)
    gp = gsx.Problem(ctx, prob)
    fd = gsx.FastDiag(gp, alpha=alpha)
    r = np.random.default_rng(11).standard_normal(gp.nfree)
    want = O.fd_operator(prob.patches[0], sides, alpha=alpha)(r)
    rd = ctx.to_device(r)
    zd = ctx.array(gp.nfree)
    fd.apply(rd, zd)
    got = zd.download()
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()
    fd.apply(rd, rd)   # in place
    assert np.abs(rd.download() - got).max() <= 1e-14 * np.abs(got).max()
    for o in (rd, zd, fd, gp):
        o.close()


@pytest.mark.parametrize("n,p", [(8, 3), (10, 2)])
def test_fd_pcg_matches_oracle(ctx, n, p):
    """Noisy cube with lognormal kappa: FD-preconditioned CG takes the
    oracle's iterations (within one: rounding of the recurrences) and
    reaches its solution; far fewer iterations than Jacobi-CG."""
    print(f"[case] fd pcg cube n={n} p={p}", flush=True)
    prob = problems.poisson_noisy_cube(n=n, p=p, seed=20261202)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    Ao, bo = O.assemble(prob.patches, dofs, nfree, 0)
    sides = [s for _, s in prob.dir_sides]
    fdo = O.fd_operator(prob.patches[0], sides)
    xo, ito, _ = O.pcg(Ao, bo, fdo, tol=1e-10, maxit=500)
    _, itj, _ = O.cg_solve(Ao, bo, tol=1e-10, maxit=2000)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.MIXED)
    fd = gsx.FastDiag(gp)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.fd_cg_solve(A, fd, b, x, tol=1e-10, maxit=500)
    assert ok and rel < 1e-10
    assert abs(it - ito) <= 1, (it, ito)
    assert it < itj / 2, (it, itj)
    xs = spla.spsolve(Ao.tocsc(), bo)
    assert np.abs(x.download() - xs).max() <= 1e-8 * np.abs(xs).max()
    # mixed: FP32 iterations and factors, FP64 reliable updates
    xm = ctx.array(gp.nfree).zero()
    itm, relm, okm = gsx.fd_cg_solve(A, fd, b, xm, tol=1e-10, maxit=500, precision=gsx.MIXED)
    assert okm and relm < 1e-10
    assert itm <= ito + 4, (itm, ito)
    assert np.abs(xm.download() - xs).max() <= 1e-8 * np.abs(xs).max()
    for o in (x, xm, fd, A, b, gp):
        o.close()


@pytest.mark.parametrize("d,p", [(2, 3), (3, 2)])
def test_fd_pcg_exact_on_the_unit_box(ctx, d, p):
    """Identity map, kappa = 1: the stiffness is the Kronecker sum, so the
    preconditioned CG converges in one iteration."""
    P = problems.box_patch(d, p, 12 if d == 2 else 6)
    P.fkind = 1
    prob = problems.Problem([P], problems.all_sides(d), [])
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.FP64)
    fd = gsx.FastDiag(gp)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.fd_cg_solve(A, fd, b, x, tol=1e-10, maxit=50)
    assert ok and it == 1, (it, rel)
    for o in (x, fd, A, b, gp):
        o.close()


def test_fd_error_paths(ctx):
    # several patches: no tensor structure
    grid = problems.square_grid(2, 1, n=4, p=2)
    gp = gsx.Problem(ctx, grid)
    with pytest.raises(gsx.GsxError):
        gsx.FastDiag(gp)
    gp.close()
    # no Dirichlet end and alpha = 0: singular
    P = problems.box_patch(2, 2, 4)
    P.fkind = 1
    gp = gsx.Problem(ctx, problems.Problem([P], [], []))
    with pytest.raises(gsx.GsxError):
        gsx.FastDiag(gp)
    fd = gsx.FastDiag(gp, alpha=1.0)   # with a mass term it is regular
    fd.close()
    gp.close()


# ---------------------------------------------------------------------------
# subspace corrected mass smoother (subspaceCorrectedMassSmootherOp)
# ---------------------------------------------------------------------------

def _graded_curved(d, p, n, seed, sides):
    prob = _curved(d, p, n, seed, sides)
    P = prob.patches[0]
    rng = np.random.default_rng(seed + 1)
    knots = []
    for _ in range(d):
        w = rng.uniform(1.0, 3.0, n)
        knots.append(np.concatenate([np.zeros(p + 1), np.cumsum(w)[:-1] / w.sum(), np.ones(p + 1)]))
    P.knots = knots
    return prob


@pytest.mark.parametrize("d,p,n,sides,sigma,alpha,beta,graded", [
    (2, 1, 8, [1, 2, 3, 4], 0.12, 0.0, 1.0, False),
    (2, 2, 9, [1, 4], 0.12, 0.0, 1.0, True),
    (2, 3, 10, [1, 2, 3, 4], 0.3, 0.5, 2.0, True),
    (2, 4, 12, [], 0.12, 1.0, 1.0, False),
    (2, 3, 2, [1, 2, 3, 4], 0.12, 0.0, 1.0, False),   # no interior functions: B_tilde empty
    (2, 4, 40, [2, 3], 0.12, 0.0, 1.0, True),
    (3, 2, 5, [1, 2, 3, 4, 5, 6], 0.12, 0.0, 1.0, False),
    (3, 3, 6, [2, 3, 5], 0.2, 0.0, 1.5, True),
    (3, 4, 5, [1, 2, 3, 4, 5, 6], 0.12, 0.3, 1.0, False),
])
def test_scms_apply_matches_oracle(ctx, d, p, n, sides, sigma, alpha, beta, graded):
    """The device smoother (per type: reduce and expand with the
    generalized eigenvectors of the boundary complement, a diagonal
    scaling, the interior mass inverses) against the oracle's dense
    restatement that builds and solves each correction as it stands."""
    print(f"[case] scms apply d={d} p={p} n={n} sides={sides} graded={graded}", flush=True)
    prob = (_graded_curved if graded else _curved)(d, p, n, 17, sides)
    gp = gsx.Problem(ctx, prob)
    sm = gsx.SubspaceMassSmoother(gp, sigma=sigma, alpha=alpha, beta=beta)
    S = O.scms_matrix(prob.patches[0], sides, sigma=sigma, alpha=alpha, beta=beta)
    assert S.shape == (gp.nfree, gp.nfree)
    r = np.random.default_rng(12).standard_normal(gp.nfree)
    want = S @ r
    rd = ctx.to_device(r)
    zd = ctx.array(gp.nfree)
    sm.apply(rd, zd)
    got = zd.download()
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    sm.apply(rd, rd)   # in place
    assert np.abs(rd.download() - got).max() <= 1e-14 * np.abs(got).max()
    for o in (rd, zd, sm, gp):
        o.close()


def test_scms_large_is_symmetric_positive_and_linear(ctx):
    """At 400^2 free functions (p = 3) the dense oracle is out of reach:
    (u, S v) = (v, S u), (v, S v) > 0 and S(a u + v) = a S u + S v."""
    print("[case] scms large 2d p=3 n=400", flush=True)
    prob = _graded_curved(2, 3, 400, 23, [1, 2, 3, 4])
    gp = gsx.Problem(ctx, prob)
    sm = gsx.SubspaceMassSmoother(gp)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(gp.nfree), rng.standard_normal(gp.nfree)
    out = []
    for w in (u, v, 0.75 * u + v):
        wd, zd = ctx.to_device(w), ctx.array(gp.nfree)
        sm.apply(wd, zd)
        out.append(zd.download())
        wd.close()
        zd.close()
    Su, Sv, Sw = out
    assert abs(u @ Sv - v @ Su) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(Sv)
    assert v @ Sv > 0 and u @ Su > 0
    assert np.abs(Sw - (0.75 * Su + Sv)).max() <= 1e-13 * np.abs(Sw).max()
    sm.close()
    gp.close()


@pytest.mark.parametrize("d,p,n,sides", [(2, 3, 12, [4]), (3, 2, 6, [1, 2, 3, 4, 5, 6])])
def test_scms_pcg_matches_oracle(ctx, d, p, n, sides):
    """CG on the stiffness matrix with the smoother as preconditioner (the
    reference's preconditioner test uses it so, on a mass-shifted matrix):
    the oracle's iterations within three and its solution, FP64 and mixed."""
    print(f"[case] scms pcg d={d} p={p} n={n} sides={sides}", flush=True)
    prob = _curved(d, p, n, 29, sides)
    dofs, nfree = O.mapper(prob.patches, prob.dir_sides, prob.ifaces)
    Ao, bo = O.assemble(prob.patches, dofs, nfree, 0)
    S = O.scms_matrix(prob.patches[0], sides)
    xo, ito, _ = O.pcg(Ao, bo, lambda r: S @ r, tol=1e-10, maxit=1000)
    gp = gsx.Problem(ctx, prob)
    A, b = gp.assemble(gsx.PATTERN_EXPR, gsx.MIXED)
    sm = gsx.SubspaceMassSmoother(gp)
    x = ctx.array(gp.nfree).zero()
    it, rel, ok = gsx.fd_cg_solve(A, sm, b, x, tol=1e-10, maxit=1000)
    assert ok and rel < 1e-10
    assert abs(it - ito) <= 3, (it, ito)
    xs = spla.spsolve(Ao.tocsc(), bo)
    assert np.abs(x.download() - xs).max() <= 1e-8 * np.abs(xs).max()
    xm = ctx.array(gp.nfree).zero()
    itm, relm, okm = gsx.fd_cg_solve(A, sm, b, xm, tol=1e-10, maxit=1000, precision=gsx.MIXED)
    assert okm and relm < 1e-10
    assert itm <= ito * 3 // 2, (itm, ito)   # FP32 factors: slower once the recurrence nears its floor
    assert np.abs(xm.download() - xs).max() <= 1e-8 * np.abs(xs).max()
    for o in (x, xm, sm, A, b, gp):
        o.close()


def test_scms_error_paths(ctx):
    grid = problems.square_grid(2, 1, n=4, p=2)
    gp = gsx.Problem(ctx, grid)
    with pytest.raises(gsx.GsxError):
        gsx.SubspaceMassSmoother(gp)
    gp.close()
    P = problems.box_patch(2, 2, 4)
    P.fkind = 1
    gp = gsx.Problem(ctx, problems.Problem([P], problems.all_sides(2), []))
    for kw in (dict(beta=0.0), dict(sigma=0.0), dict(alpha=-1.0)):
        with pytest.raises(gsx.GsxError):
            gsx.SubspaceMassSmoother(gp, **kw)
    gp.close()
